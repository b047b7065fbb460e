"""3D offline_log codec on the GPU against the reference decoder fixtures and
the CPU oracle: bit-exact streams and decoded fields, edge shapes, full size."""
import glob
import os

import numpy as np
import pytest

import cpsz_amd as cz
from _libs import load_oracle
from cpsz_amd.synth import field_3d

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def orc():
    return load_oracle()


def _same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint32), np.ascontiguousarray(b).view(np.uint32))


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "c3d_*.npz"))),
                         ids=os.path.basename)
def test_golden(path):
    z = np.load(path)
    U, V, W = z["U"], z["V"], z["W"]
    print(f"[3d golden] {os.path.basename(path)} {U.shape}", flush=True)
    s = cz.sz_compress_cp_preserve_3d_offline_log(U, V, W, float(z["eb"]))
    assert np.array_equal(s, z["stream"])
    d = cz.sz_decompress_cp_preserve_3d_online_log(z["stream"], *U.shape)
    for a, b in zip(d, (z["dU"], z["dV"], z["dW"])):
        assert _same(a, b)


@pytest.mark.parametrize("shape,seed,eb", [
    ((2, 2, 2), 1, 0.05), ((1, 5, 70), 2, 0.05), ((70, 1, 3), 3, 0.05), ((3, 130, 5), 4, 0.05),
    ((20, 65, 33), 5, 0.05), ((33, 64, 64), 6, 0.1), ((17, 129, 40), 7, 0.01),
])
def test_shapes_vs_oracle(orc, shape, seed, eb):
    print(f"[3d oracle] {shape} seed={seed} eb={eb}", flush=True)
    U, V, W = field_3d(*shape, seed)
    try:
        ref = orc.compress_3d(U, V, W, eb)
    except AssertionError:
        # no predictable vertex: the reference cannot write an empty code section
        with pytest.raises(cz.CpszError):
            cz.sz_compress_cp_preserve_3d_offline_log(U, V, W, eb)
        return
    s = cz.sz_compress_cp_preserve_3d_offline_log(U, V, W, eb)
    assert np.array_equal(s, ref)
    d = cz.sz_decompress_cp_preserve_3d_online_log(ref, *shape)
    o = orc.decompress_3d(ref, *shape)
    for a, b in zip(d, o):
        assert _same(a, b)


def test_full_size_256(orc):
    """256^3 (the 3D headline field): bit-exact stream and decode vs the oracle."""
    shape = (256, 256, 256)
    print("[3d full] 256^3", flush=True)
    U, V, W = field_3d(*shape, 0)
    s = cz.sz_compress_cp_preserve_3d_offline_log(U, V, W, 0.05)
    ref = orc.compress_3d(U, V, W, 0.05)
    assert s.size == ref.size and np.array_equal(s, ref)
    d = cz.sz_decompress_cp_preserve_3d_online_log(s, *shape)
    o = orc.decompress_3d(ref, *shape)
    for a, b in zip(d, o):
        assert _same(a, b)


def test_zero_and_negative_values(orc):
    U, V, W = field_3d(10, 12, 14, 9)
    U[2:5, 3:6, :] = 0
    V[:, :, 4] = -V[:, :, 4]
    W[7, 7, 7] = 0
    s = cz.sz_compress_cp_preserve_3d_offline_log(U, V, W, 0.05)
    assert np.array_equal(s, orc.compress_3d(U, V, W, 0.05))
    for a, b in zip(cz.sz_decompress_cp_preserve_3d_online_log(s, 10, 12, 14),
                    orc.decompress_3d(s, 10, 12, 14)):
        assert _same(a, b)


def test_nonfinite_rejected_3d():
    U, V, W = field_3d(8, 8, 8, 1)
    W[1, 2, 3] = np.inf
    with pytest.raises(cz.CpszError) as e:
        cz.sz_compress_cp_preserve_3d_offline_log(U, V, W, 0.05)
    assert e.value.code == -6


@pytest.mark.parametrize("kind", ["tiny", "small", "large", "integers"])
def test_bound_fast_path_extremes(orc, kind):
    """Fields whose vertex bounds sit near id thresholds, are exactly 0, or
    come from sums beyond 2^124: the approximate bound pass must leave those
    vertices to the exact recomputation, so the stream stays identical."""
    shape = (20, 33, 40)
    U, V, W = field_3d(*shape, 11)
    if kind == "integers":
        rng = np.random.default_rng(11)
        U, V, W = (rng.integers(-3, 4, size=shape).astype(np.float32) for _ in range(3))
    else:
        scale = {"tiny": 1e-13, "small": 1e-6, "large": 1e12}[kind]
        U, V, W = (np.float32(scale) * x for x in (U, V, W))
    print(f"[3d fast-path] {kind} {shape}", flush=True)
    try:
        ref = orc.compress_3d(U, V, W, 0.05)
    except AssertionError:
        with pytest.raises(cz.CpszError):
            cz.sz_compress_cp_preserve_3d_offline_log(U, V, W, 0.05)
        return
    s = cz.sz_compress_cp_preserve_3d_offline_log(U, V, W, 0.05)
    assert np.array_equal(s, ref)


def test_fast_bound_premises():
    """The approximate bound pass assumes the fast cube root is within 2 ulp of
    the reference's over [0, 2^126) (checked here exhaustively) and that the
    fast sign bound stays inside its error envelope (checked on 2.6e8 hashed
    argument pairs, from far apart to nearly equal)."""
    print("[3d fast-bound premises]", flush=True)
    ulp, ratio = cz.check_fast_bounds(0)
    assert ulp <= 2
    assert ratio <= 1.0
    print(f"max ulp {ulp}, max error/envelope {ratio:.3f}", flush=True)


def _id_section_nbytes_offset(s, n):
    """Offset of the eb-id Huffman section's bitstream byte count (layout:
    int base | int radius | 3 sign maps | size_t nunp | floats | size_t n |
    size_t nodes | u32 tree_size | tree | size_t nbytes | bits)."""
    smap = (n - 1) // 8 + 1
    off = 8 + 3 * smap
    nunp = int(np.frombuffer(s[off:off + 8].tobytes(), np.uint64)[0])
    off += 8 + 4 * nunp + 8
    ts = int(np.frombuffer(s[off + 8:off + 12].tobytes(), np.uint32)[0])
    return off + 12 + ts


@pytest.mark.parametrize("case", ["truncated", "huge_nbytes", "wrapping_nbytes"])
def test_malformed_3d_stream_rejected(case):
    """Truncated streams and bitstream byte counts that overflow the bounds
    check are rejected as stream errors before any decode kernel runs."""
    print(f"[3d] malformed stream: {case}", flush=True)
    U, V, W = field_3d(9, 10, 11, 5)
    s = cz.sz_compress_cp_preserve_3d_offline_log(U, V, W, 0.05).copy()
    if case == "truncated":
        s = s[: s.size // 2]
    else:
        at = _id_section_nbytes_offset(s, U.size)
        bad = (1 << 63) if case == "huge_nbytes" else (1 << 64) - 16
        s[at:at + 8] = np.frombuffer(np.uint64(bad).tobytes(), np.uint8)
    with pytest.raises(cz.CpszError) as e:
        cz.sz_decompress_cp_preserve_3d_online_log(s, 9, 10, 11)
    assert e.value.code == -4
