"""The 2D bounds pass for float fields approximates the last FP64 steps of the
reference's triangle bound (src/sz_compress_cp_preserve_2d.cpp:125-172) in
FP32 and falls back to the exact bound wherever the approximation does not
fix a vertex's eb codes. These tests hold it to the all-FP64 kernel (itself
checked against the oracle elsewhere): identical streams on fields chosen to
reach every fallback (products outside the FP32 range, ties and zeros,
cancelling sign sums, critical-point triangles) and on the 96-field batch."""
import ctypes as C

import numpy as np
import pytest
import torch

import cpsz_amd as cz
from _libs import load_oracle, mask_2d_padding
from cpsz_amd.synth import batch_2d_shapes, field_2d, field_2d_device

pytestmark = pytest.mark.gpu


def _fields():
    rng = np.random.default_rng(7)
    r1, r2 = 150, 230
    U, V = field_2d(r1, r2, 5)
    out = [("smooth", U, V)]
    # per-vertex scale 2^[-45, 45]: products leave [2^-60, 2^58] but stay finite
    s = np.exp2(rng.integers(-45, 46, size=(r1, r2))).astype(np.float32)
    out.append(("wide_magnitudes", (U * s).astype(np.float32), (V * s[::-1]).astype(np.float32)))
    # small integers: exact zeros, equal products, det == 0, many critical points
    out.append(("integers", rng.integers(-3, 4, size=(r1, r2)).astype(np.float32),
                rng.integers(-3, 4, size=(r1, r2)).astype(np.float32)))
    # near-constant: p and n of the sign bounds agree to ~7 digits
    out.append(("near_constant", (1 + 1e-7 * rng.standard_normal((r1, r2))).astype(np.float32),
                (-1 + 1e-7 * rng.standard_normal((r1, r2))).astype(np.float32)))
    # zero regions next to a smooth field
    Z, W = U.copy(), V.copy()
    Z[:, : r2 // 2] = 0
    W[r1 // 3:, :] = 0
    out.append(("zero_regions", Z, W))
    # white noise: a critical point in a large share of the triangles
    out.append(("noise", rng.standard_normal((r1, r2)).astype(np.float32),
                rng.standard_normal((r1, r2)).astype(np.float32)))
    out.append(("huge", (U * np.float32(3e18)).astype(np.float32), (V * np.float32(3e18)).astype(np.float32)))
    out.append(("tiny", (U * np.float32(2e-19)).astype(np.float32), (V * np.float32(2e-19)).astype(np.float32)))
    return out


@pytest.mark.parametrize("eb", [0.1, 0.01, 0.7])
def test_approximate_bounds_give_the_exact_streams(eb):
    orc = load_oracle()
    try:
        for name, U, V in _fields():
            print(f"[bounds2d] {name} eb={eb}", flush=True)
            cz.set_exact_2d_bounds(True)
            exact = cz.sz_compress_cp_preserve_2d_offline(U, V, eb)
            cz.set_exact_2d_bounds(False)
            fast = cz.sz_compress_cp_preserve_2d_offline(U, V, eb)
            assert np.array_equal(exact, fast), name
            if name in ("wide_magnitudes", "integers", "near_constant"):
                ref = orc.compress_2d(U, V, eb)
                assert np.array_equal(mask_2d_padding(fast, *U.shape), mask_2d_padding(ref, *U.shape)), name
    finally:
        cz.set_exact_2d_bounds(False)


def test_approximate_bounds_batch_96_fields():
    """Every stream of the 96-field batch (220M vertices) byte for byte."""
    shapes = batch_2d_shapes(96, seed=0)
    dev = [field_2d_device(r1, r2, seed=1000 + k) for k, (r1, r2) in enumerate(shapes)]
    fl = [(u.data_ptr(), v.data_ptr(), r1, r2) for (u, v), (r1, r2) in zip(dev, shapes)]
    torch.cuda.synchronize()  # the library runs on its own stream
    hip = C.CDLL("libamdhip64.so")
    got = []
    try:
        for exact in (True, False):
            print(f"[bounds2d] batch of 96, exact={exact}", flush=True)
            cz.set_exact_2d_bounds(exact)
            base, off, size = cz.compress_2d_batch_dev(fl, 0.1)
            n = int(off[-1] + size[-1])
            t = torch.empty(n, dtype=torch.uint8, device="cuda")
            torch.cuda.synchronize()
            assert hip.hipMemcpy(C.c_void_p(t.data_ptr()), C.c_void_p(base), C.c_size_t(n), 3) == 0
            got.append((t, off.copy(), size.copy()))
    finally:
        cz.set_exact_2d_bounds(False)
    (a, oa, sa), (b, ob, sb) = got
    assert np.array_equal(sa, sb)
    for k in range(len(shapes)):
        assert torch.equal(a[int(oa[k]):int(oa[k] + sa[k])], b[int(ob[k]):int(ob[k] + sb[k])]), (k, shapes[k])
