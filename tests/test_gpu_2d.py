"""2D codec on the GPU against the reference golden streams and the CPU oracle.

Bit-exact: stream bytes (uninitialised meanInfo padding masked) and decoded
fields. Property checks at full size: every decoded value within its bound.
"""
import glob
import os

import numpy as np
import pytest

import cpsz_amd as cz
from _libs import load_oracle, mask_2d_padding
from cpsz_amd.synth import field_2d

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def orc():
    return load_oracle()


def _same_bits(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint32), np.ascontiguousarray(b).view(np.uint32))


@pytest.mark.parametrize("path", sorted(p for p in glob.glob(os.path.join(GOLD, "c2d_*.npz"))
                                        if "f64" not in p), ids=os.path.basename)
def test_golden_streams(path):
    z = np.load(path)
    U, V = z["U"], z["V"]
    r1, r2 = U.shape
    print(f"[2d golden] {os.path.basename(path)} {r1}x{r2}", flush=True)
    s = cz.sz_compress_cp_preserve_2d_offline(U, V, float(z["eb"]))
    assert np.array_equal(mask_2d_padding(s, r1, r2), z["stream"])
    dU, dV = cz.sz_decompress_cp_preserve_2d_offline(z["stream"], r1, r2)
    assert _same_bits(dU, z["dU"]) and _same_bits(dV, z["dV"])


@pytest.mark.parametrize("shape,seed,eb", [
    ((1, 70), 1, 0.1), ((70, 1), 2, 0.1), ((2, 2), 3, 0.1), ((63, 65), 4, 0.1),
    ((64, 12), 5, 0.1), ((65, 13), 6, 0.05), ((129, 300), 7, 0.1), ((300, 129), 8, 0.2),
    ((517, 23), 9, 0.1),
])
def test_shapes_vs_oracle(orc, shape, seed, eb):
    r1, r2 = shape
    print(f"[2d oracle] {r1}x{r2} seed={seed} eb={eb}", flush=True)
    U, V = field_2d(r1, r2, seed)
    s = cz.sz_compress_cp_preserve_2d_offline(U, V, eb)
    ref = orc.compress_2d(U, V, eb)
    assert np.array_equal(mask_2d_padding(s, r1, r2), mask_2d_padding(ref, r1, r2))
    dU, dV = cz.sz_decompress_cp_preserve_2d_offline(ref, r1, r2)
    oU, oV = orc.decompress_2d(ref, r1, r2)
    assert _same_bits(dU, oU) and _same_bits(dV, oV)


def test_full_size_config0(orc):
    """2400x3600 (the 2D headline field): bit-exact stream and decode vs the oracle."""
    r1, r2 = 2400, 3600
    print(f"[2d full] {r1}x{r2}", flush=True)
    U, V = field_2d(r1, r2, 0)
    s = cz.sz_compress_cp_preserve_2d_offline(U, V, 0.1)
    ref = orc.compress_2d(U, V, 0.1)
    assert s.size == ref.size
    assert np.array_equal(mask_2d_padding(s, r1, r2), mask_2d_padding(ref, r1, r2))
    dU, dV = cz.sz_decompress_cp_preserve_2d_offline(s, r1, r2)
    oU, oV = orc.decompress_2d(ref, r1, r2)
    assert _same_bits(dU, oU) and _same_bits(dV, oV)
    # every decoded value within its quantized bound (ids drawn from the oracle)
    eb = orc.eb_2d(U, V, 0.1)
    for X, D in ((U, dU), (V, dV)):
        ids, prec = orc.eb_quantize_2d(np.abs(X).astype(np.float64) * eb)
        err = np.abs(D.astype(np.float64) - X)
        ok = (prec == 0) | (ids == 0) | (err <= prec)
        assert ok.all()


def test_lossless_roundtrip():
    U, V = field_2d(100, 120, 3)
    s = cz.sz_compress_cp_preserve_2d_offline(U, V, 0.1)
    z = cz.sz_lossless_compress(s, 3)
    assert z.size < s.size
    assert np.array_equal(cz.sz_lossless_decompress(z, s.size), s)


def test_nonfinite_rejected():
    U, V = field_2d(30, 30, 1)
    U[5, 5] = np.nan
    with pytest.raises(cz.CpszError) as e:
        cz.sz_compress_cp_preserve_2d_offline(U, V, 0.1)
    assert e.value.code == -6


def test_truncated_stream_rejected():
    U, V = field_2d(40, 40, 2)
    s = cz.sz_compress_cp_preserve_2d_offline(U, V, 0.1)
    with pytest.raises(cz.CpszError):
        cz.sz_decompress_cp_preserve_2d_offline(s[: s.size // 2], 40, 40)
