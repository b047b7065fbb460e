"""ZSTD of device-resident streams (pipelined chunked copy + streaming frame,
and the per-field batch form) against the host sz_lossless_compress of the
same bytes, which equals the reference's call (tests/test_host.py): byte-equal
frames at sizes around the 8 MiB chunk, empty and tiny inputs, and on the
streams of a real batched compression."""
import numpy as np
import pytest
import torch

import cpsz_amd as cz
from cpsz_amd.synth import batch_2d_shapes, field_2d_device

pytestmark = pytest.mark.gpu


def _bytes(n, seed):
    rng = np.random.default_rng(seed)
    # compressible: small values with runs, some random bytes
    a = rng.integers(0, 4, n, dtype=np.uint8)
    a[rng.random(n) < 0.05] = rng.integers(0, 256, dtype=np.uint8)
    return a


@pytest.mark.parametrize("n", [0, 1, 1000, (8 << 20) - 1, 8 << 20, (8 << 20) + 1, 20_000_003])
def test_lossless_dev_equals_host(n):
    print(f"[lossless] n={n}", flush=True)
    h = _bytes(n, n % 97)
    d = torch.from_numpy(h).cuda() if n else torch.empty(1, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    for level in (3, 1):
        got = cz.sz_lossless_compress_dev(d.data_ptr(), n, level)
        assert np.array_equal(got, cz.sz_lossless_compress(h, level)), level


def test_lossless_batch_dev_equals_host():
    sizes = [0, 5, 70000, 3_000_001, 131072, 9_000_000, 17]
    offs, pos = [], 0
    for s in sizes:
        offs.append(pos)
        pos += s + 123  # gaps between the streams
    h = _bytes(pos, 5)
    d = torch.from_numpy(h).cuda()
    torch.cuda.synchronize()
    got = cz.sz_lossless_compress_batch_dev(d.data_ptr(), offs, sizes, 3)
    for k, (o, s) in enumerate(zip(offs, sizes)):
        assert np.array_equal(got[k], cz.sz_lossless_compress(h[o:o + s], 3)), k


def test_lossless_batch_on_compressed_streams():
    shapes = batch_2d_shapes(12, seed=4)
    dev = [field_2d_device(min(r1, 700), min(r2, 700), seed=30 + k) for k, (r1, r2) in enumerate(shapes)]
    fl = [(u.data_ptr(), v.data_ptr(), u.shape[0], u.shape[1]) for u, v in dev]
    torch.cuda.synchronize()  # the library runs on its own stream
    base, off, size = cz.compress_2d_batch_dev(fl, 0.1)
    torch.cuda.synchronize()
    frames = cz.sz_lossless_compress_batch_dev(base, off, size, 3)
    for k, (o, s) in enumerate(zip(off, size)):
        one = cz.sz_lossless_compress_dev(base + int(o), int(s), 3)
        assert np.array_equal(frames[k], one), k
        back = cz.sz_lossless_decompress(frames[k], int(s))
        assert back.size == s
