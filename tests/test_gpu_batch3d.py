"""Batched 3D codec on the GPU: every field's stream equals the single-field
codec's (and the oracle's), decoded fields equal the single-field decoder's,
for uneven shapes in one persistent wavefront launch, with and without the
autotuner."""
import ctypes as C

import numpy as np
import pytest
import torch

import cpsz_amd as cz
from _libs import load_oracle
from cpsz_amd.synth import field_3d

pytestmark = pytest.mark.gpu

SHAPES = [(3, 130, 5), (20, 65, 33), (9, 64, 64), (33, 33, 33), (4, 3, 200), (17, 129, 40), (6, 1, 70),
          (2, 70, 2)]


def _host(ptr, n):
    out = np.empty(n, np.uint8)
    torch.cuda.synchronize()
    C.CDLL("libamdhip64.so").hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(ptr), C.c_size_t(n), 2)
    return out


@pytest.fixture(scope="module")
def batch():
    fields = [field_3d(*shape, 40 + k) for k, shape in enumerate(SHAPES)]
    dev = [tuple(torch.from_numpy(x).cuda() for x in f) for f in fields]
    return fields, dev


def _fields(dev):
    return [(u.data_ptr(), v.data_ptr(), w.data_ptr(), *SHAPES[k]) for k, (u, v, w) in enumerate(dev)]


@pytest.mark.parametrize("tune", [False, True])
def test_batch3d_streams_match_single_field(batch, tune):
    print(f"[batch3d] {len(SHAPES)} fields tune={tune}", flush=True)
    fields, dev = batch
    base, off, size = cz.compress_3d_batch_dev(_fields(dev), 0.05, tune=tune)
    orc = load_oracle()
    allb = _host(base, int(off[-1] + size[-1]))
    for k, ((U, V, W), shape) in enumerate(zip(fields, SHAPES)):
        s = allb[off[k]:off[k] + size[k]]
        single = cz.sz_compress_cp_preserve_3d_offline_log(U, V, W, 0.05)
        assert np.array_equal(s, single), shape
        if k % 3 == 0:
            assert np.array_equal(s, orc.compress_3d(U, V, W, 0.05)), shape
    w, g = cz.batch_3d_config(SHAPES)
    assert w >= 1 and g >= 1


def test_batch3d_round_trip_matches_single_decoder(batch):
    fields, dev = batch
    base, off, size = cz.compress_3d_batch_dev(_fields(dev), 0.05)
    outs = [tuple(torch.empty(s, dtype=torch.float32, device="cuda") for _ in range(3)) for s in SHAPES]
    cz.decompress_3d_batch_dev([(base + int(o), int(n)) for o, n in zip(off, size)], SHAPES,
                               [tuple(t.data_ptr() for t in o) for o in outs])
    torch.cuda.synchronize()
    allb = _host(base, int(off[-1] + size[-1]))
    for k, shape in enumerate(SHAPES):
        d = cz.sz_decompress_cp_preserve_3d_online_log(allb[off[k]:off[k] + size[k]], *shape)
        for a, b in zip(outs[k], d):
            assert np.array_equal(a.cpu().numpy().view(np.uint32), b.view(np.uint32)), shape
