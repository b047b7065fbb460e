"""A caller written against the reference's C++ API (drop-in header + libcpsz_hip.so):
compress -> ZSTD -> unZSTD -> decompress, checked against the oracle."""
import os
import subprocess

import numpy as np
import pytest

from _libs import load_oracle
from cpsz_amd.synth import field_2d, field_3d

pytestmark = pytest.mark.gpu
DEMO = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "dropin_roundtrip")


def _run(tmp_path, kind, comps, dims, eb):
    inp = tmp_path / "in.bin"
    np.concatenate([c.ravel() for c in comps]).astype(np.float32).tofile(inp)
    out = str(tmp_path / "o")
    r = subprocess.run([DEMO, kind, str(inp), out] + [str(d) for d in dims] + [str(eb)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr
    stream = np.fromfile(out + ".stream", np.uint8)
    dec = np.fromfile(out + ".out", np.float32).reshape(len(comps), *dims)
    return stream, dec


def test_dropin_2d(tmp_path):
    orc = load_oracle()
    U, V = field_2d(130, 170, 4)
    stream, dec = _run(tmp_path, "2d", (U, V), (130, 170), 0.1)
    from _libs import mask_2d_padding
    ref = orc.compress_2d(U, V, 0.1)
    assert np.array_equal(mask_2d_padding(stream, 130, 170), mask_2d_padding(ref, 130, 170))
    for a, b in zip(dec, orc.decompress_2d(ref, 130, 170)):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_dropin_3d(tmp_path):
    orc = load_oracle()
    U, V, W = field_3d(20, 24, 28, 4)
    stream, dec = _run(tmp_path, "3d", (U, V, W), (20, 24, 28), 0.05)
    ref = orc.compress_3d(U, V, W, 0.05)
    assert np.array_equal(stream, ref)
    for a, b in zip(dec, orc.decompress_3d(ref, 20, 24, 28)):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_dropin_2d_offline_log(tmp_path):
    orc = load_oracle()
    U, V = field_2d(90, 140, 5)
    stream, dec = _run(tmp_path, "2dlog", (U, V), (90, 140), 0.1)
    from _libs import mask_2d_log_padding
    ref = orc.compress_2d_log(U, V, 0.1)
    assert np.array_equal(mask_2d_log_padding(stream, 90, 140), mask_2d_log_padding(ref, 90, 140))
    for a, b in zip(dec, orc.decompress_2d_log(ref, 90, 140)):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
