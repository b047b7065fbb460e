"""Host checks (no GPU) of the premises behind the 2D bounds pass's FP32
approximation: the cell-table entry of every triangle, computed by the
library's own header with square roots and reciprocals moved by up to 4 ulp,
encloses the oracle's FP64 bound (max_eb_to_keep_position_and_type,
src/sz_compress_cp_preserve_2d.cpp:125-172); exact entries are exact; and
the division-free ratio test agrees with IEEE division."""
import numpy as np
import pytest

import cpsz_amd as cz
from _libs import load_oracle

REL, ABS = 2.0 ** -16, 2.0 ** -50  # the enclosure the kernel assumes


def _triangles(n, seed):
    rng = np.random.default_rng(seed)
    sets = {
        "uniform": rng.uniform(-1, 1, (n, 6)),
        # magnitudes over 2^[-40, 40]: products leave the FP32 fast range
        "wide": rng.uniform(-1, 1, (n, 6)) * np.exp2(rng.integers(-40, 41, (n, 6))),
        # p and n of the sign bounds close: near-constant vertex values
        "near_equal": np.sign(rng.uniform(-1, 1, (n, 6))) * (1 + 1e-6 * rng.standard_normal((n, 6))),
        # small integers: zeros, ties, det == 0, critical points
        "integers": rng.integers(-3, 4, (n, 6)).astype(np.float64),
        # smooth-field-like: the three vertices of a small triangle differ slightly
        "smooth": np.concatenate([rng.uniform(-1, 1, (n, 1)) + 1e-3 * rng.standard_normal((n, 3)),
                                  rng.uniform(-1, 1, (n, 1)) + 1e-3 * rng.standard_normal((n, 3))], axis=1),
    }
    return {k: v.astype(np.float32) for k, v in sets.items()}


@pytest.mark.parametrize("ulps", [0, 1, 4])
def test_fast_entries_enclose_the_exact_bound(ulps):
    orc = load_oracle()
    for name, uv in _triangles(40000, 11 + ulps).items():
        e = cz.tri_bound_2d_fast_host(uv, ulps).astype(np.float64)
        # the entry does not depend on the triangle's shape; the exact bound
        # does only through the critical-point branch, which the entry marks
        for lower in (0, 1):
            x = orc.tri_bound_2d(uv, lower)
            need = e == -1
            ex0, ex1 = e == 0, e == 2
            approx = ~(need | ex0 | ex1)
            assert np.all(x[ex0] == 0), name
            assert np.all(x[ex1] == 1), name
            a = e[approx]
            assert np.all((a > 2.0 ** -56) & (a <= 1 + 2.0 ** -19)), name
            lo, hi = a * (1 - REL) - ABS, a * (1 + REL) + ABS
            xa = x[approx]
            bad = ~((lo <= xa) & (xa <= hi))
            assert not bad.any(), (name, lower, uv[approx][bad][:3], a[bad][:3], xa[bad][:3])
            # and the approximation is close (the envelope is not vacuous)
            assert np.all(np.abs(a - xa) <= 2.0 ** -19 * xa + 2.0 ** -49), name
        if name in ("uniform", "smooth", "near_equal"):
            assert approx.mean() > 0.5, (name, approx.mean())


def test_ratio_test_matches_division():
    rng = np.random.default_rng(5)
    n = 200000
    d = rng.uniform(-1, 1, n) * np.exp2(rng.integers(-100, 100, n))
    # quotients within a few ulp of 1 on both sides, and far from it
    k = rng.integers(-64, 65, n)
    det = np.where(rng.uniform(size=n) < 0.8, d * (1 + k * 2.0 ** -53), d * rng.uniform(-3, 3, n))
    det[:1000] = np.nextafter(d[:1000], np.inf)
    det[1000:2000] = np.nextafter(d[1000:2000], -np.inf)
    d[2000:2100] = 0.0
    d[2100:2200] = -0.0
    with np.errstate(divide="ignore", invalid="ignore"):
        want = det / d >= 1
    assert np.array_equal(cz.ratio_ge1_host(det, d), want)
