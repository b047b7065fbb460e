"""Pins the CPU oracle: reference golden vectors, live reference comparisons
(when oracle/_ref is built), closed forms and brute-force invariants."""
import glob
import math
import os

import numpy as np
import pytest

from _libs import Reference, have_reference, load_oracle, mask_2d_log_padding, mask_2d_padding

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


@pytest.fixture(scope="module")
def orc():
    return load_oracle()


def _bits(a):
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)


# ------------------------------------------------------------ golden vectors --
@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "c2d_*.npz"))),
                         ids=lambda p: os.path.basename(p))
def test_2d_golden(orc, path):
    z = np.load(path)
    U, V = z["U"], z["V"]
    r1, r2 = U.shape
    item = U.dtype.itemsize
    s = mask_2d_padding(orc.compress_2d(U, V, float(z["eb"])), r1, r2, item)
    assert np.array_equal(s, z["stream"])
    dU, dV = orc.decompress_2d(z["stream"], r1, r2, U.dtype)
    assert np.array_equal(_bits(dU), _bits(z["dU"]))
    assert np.array_equal(_bits(dV), _bits(z["dV"]))


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "c2l_*.npz"))),
                         ids=lambda p: os.path.basename(p))
def test_2d_log_golden(orc, path):
    """offline_log pair against the compiled reference's fixtures (exact zeros included)."""
    z = np.load(path)
    U, V = z["U"], z["V"]
    r1, r2 = U.shape
    s = mask_2d_log_padding(orc.compress_2d_log(U, V, float(z["eb"])), r1, r2)
    assert np.array_equal(s, z["stream"])
    dU, dV = orc.decompress_2d_log(z["stream"], r1, r2)
    assert np.array_equal(_bits(dU), _bits(z["dU"]))
    assert np.array_equal(_bits(dV), _bits(z["dV"]))


def test_huffman_golden(orc):
    z = np.load(os.path.join(GOLD, "huffman.npz"))
    for i, state in enumerate(z["states"]):
        syms, sec = z[f"syms{i}"], z[f"sec{i}"]
        assert np.array_equal(orc.huff_encode(int(state), syms), sec), i
        dec, used = orc.huff_decode(sec, syms.size)
        assert used == sec.size
        assert np.array_equal(dec, syms), i


def test_lorenzo_golden(orc):
    z = np.load(os.path.join(GOLD, "lorenzo2d_45x38.npz"))
    sec, codes, recon = orc.lorenzo_2d(z["data"], z["prec"])
    sec = sec.copy()
    sec[10:13] = 0
    assert np.array_equal(sec, z["sec"])
    # reconstruction honours every per-vertex bound
    err = np.abs(recon.astype(np.float64) - z["data"])
    assert np.all(err <= z["prec"] + (z["prec"] == 0) * 0)


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "c3d_*.npz"))),
                         ids=lambda p: os.path.basename(p))
def test_3d_golden(orc, path):
    z = np.load(path)
    U, V, W = z["U"], z["V"], z["W"]
    s = orc.compress_3d(U, V, W, float(z["eb"]))
    assert np.array_equal(s, z["stream"])
    d = orc.decompress_3d(z["stream"], *U.shape)
    for a, b in zip(d, (z["dU"], z["dV"], z["dW"])):
        assert np.array_equal(_bits(a), _bits(b))


# ------------------------------------------------------ live reference runs --
@pytest.mark.skipif(not have_reference(), reason="reference build absent (oracle/_ref)")
@pytest.mark.parametrize("seed", range(6))
def test_2d_live_reference(orc, seed):
    from cpsz_amd.synth import field_2d
    rng = np.random.default_rng(seed)
    r1, r2 = int(rng.integers(20, 90)), int(rng.integers(20, 90))
    eb = float(rng.choice([0.1, 0.05, 0.2]))
    U, V = field_2d(r1, r2, 100 + seed)
    ref = Reference()
    a = mask_2d_padding(orc.compress_2d(U, V, eb), r1, r2)
    b = mask_2d_padding(ref.compress_2d(U, V, eb), r1, r2)
    assert np.array_equal(a, b)


@pytest.mark.skipif(not have_reference(), reason="reference build absent (oracle/_ref)")
@pytest.mark.parametrize("seed", range(6))
def test_2d_log_live_reference(orc, seed):
    """offline_log pair against the live reference build on random shapes, bounds
    and zero patterns (streams byte-equal, decoded fields bit-equal)."""
    from cpsz_amd.synth import field_2d
    rng = np.random.default_rng(50 + seed)
    # The reference itself faults on some small inputs (fields two columns wide;
    # narrow fields with many zeros, whose unpredictable values outgrow its
    # output buffer estimate), so narrow shapes are pinned through the
    # fixtures and the GPU edge-shape tests.
    r1, r2 = int(rng.integers(8, 90)), int(rng.integers(8, 90))
    eb = float(rng.choice([0.1, 0.01, 0.3]))
    U, V = field_2d(r1, r2, 200 + seed)
    if seed % 2 and min(r1, r2) >= 32:
        U[rng.random(U.shape) < 0.05] = 0
        V[rng.random(V.shape) < 0.05] = 0
    ref = Reference()
    a = mask_2d_log_padding(orc.compress_2d_log(U, V, eb), r1, r2)
    b = mask_2d_log_padding(ref.compress_2d_log(U, V, eb), r1, r2)
    assert np.array_equal(a, b)
    for x, y in zip(orc.decompress_2d_log(b, r1, r2), ref.decompress_2d_log(b, r1, r2)):
        assert np.array_equal(_bits(x), _bits(y))


@pytest.mark.skipif(not have_reference(), reason="reference build absent (oracle/_ref)")
def test_3d_live_reference_decoder(orc):
    from cpsz_amd.synth import field_3d
    ref = Reference()
    U, V, W = field_3d(11, 15, 19, 77)
    s = orc.compress_3d(U, V, W, 0.05)
    for a, b in zip(orc.decompress_3d(s, 11, 15, 19), ref.decompress_3d(s, 11, 15, 19)):
        assert np.array_equal(_bits(a), _bits(b))


# ------------------------------------------------------- closed forms / math --
def test_acbrt1_matches_cbrt_within_documented_error(orc):
    # include/sz_cp_preserve_utils.hpp:167-171 documents relative error in [0, 1.16e-6]
    rng = np.random.default_rng(0)
    x = np.concatenate([np.float32(10.0) ** rng.uniform(-30, 30, 200000).astype(np.float32),
                        np.array([1, 8, 27, 0.125, 1e-20, 3.0e30], np.float32)])
    a = orc.acbrt1(x).astype(np.float64)
    c = np.cbrt(x.astype(np.float64))
    rel = (a - c) / c
    assert rel.min() > -1e-7 and rel.max() < 1.3e-6


def _tri_cp(u, v):
    # zero inside the closed triangle of the three (u, v) vectors
    d = [u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]]
    D = sum(d)
    return D != 0 and all(x * D >= 0 for x in d)


def test_tri_bound_preserves_cp_existence_bruteforce(orc):
    """Perturbing every component by a relative factor within the derived bound
    never creates or destroys a critical point in the triangle."""
    rng = np.random.default_rng(1)
    tri = rng.normal(size=(3000, 6)).astype(np.float32)
    for lower in (0, 1):
        eb = orc.tri_bound_2d(tri, lower)
        assert np.all((eb >= 0) & (eb <= 1))
        for t, e in zip(tri.astype(np.float64), eb):
            base = _tri_cp(t[:3], t[3:])
            for _ in range(6):
                s = rng.uniform(-1, 1, 6) * 0.999 * e
                p = t * (1 + s)
                assert _tri_cp(p[:3], p[3:]) == base


def test_tri_bound_special_cases(orc):
    # det == 0 -> 0 (src/sz_compress_cp_preserve_2d.cpp:135-136)
    z = np.array([[1, 2, 3, 2, 4, 6]], np.float32)  # parallel vectors
    assert orc.tri_bound_2d(z, 0)[0] == 0.0
    # vectors all in one quadrant -> no CP; one sign-flip fails: bound from sqrt formula
    t = np.array([[1, 2, 4, 1, 1, 3]], np.float32)
    e = orc.tri_bound_2d(t, 0)[0]
    assert 0 < e <= 1


def _tet_cp(uvw):
    v = uvw.reshape(3, 4).T  # 4 vectors
    def det3(a, b, c):
        return float(np.linalg.det(np.array([a, b, c])))
    d = [-det3(v[1], v[2], v[3]), det3(v[0], v[2], v[3]), -det3(v[0], v[1], v[3]), det3(v[0], v[1], v[2])]
    D = sum(d)
    return D != 0 and all(x * D >= 0 for x in d)


def test_tet_bound_preserves_cp_existence_bruteforce(orc):
    rng = np.random.default_rng(2)
    tets = rng.normal(size=(1500, 12)).astype(np.float32)
    eb = orc.tet_bound_3d(tets)
    assert np.all((eb >= 0) & (eb <= 1))
    for t, e in zip(tets.astype(np.float64), eb):
        if e == 1:
            continue  # CP with a same-signed component: any perturbation keeps it
        base = _tet_cp(t)
        for _ in range(5):
            p = t * (1 + rng.uniform(-1, 1, 12) * 0.99 * e)
            assert _tet_cp(p) == base


def test_eb_quantize_2d_invariants(orc):
    eps = np.finfo(np.float64).eps
    rng = np.random.default_rng(3)
    e = np.concatenate([eps * 4.0 ** rng.uniform(-1, 40, 100000), [0, eps, 2 * eps, 4 * eps, 16 * eps]])
    ids, prec = orc.eb_quantize_2d(e)
    ok = e > eps
    assert np.all(ids[~ok] == 0) and np.all(prec[~ok] == 0)
    # precision is exactly 4^id eps and brackets the bound (within the log2 rounding window)
    assert np.all(prec[ok] == eps * 4.0 ** ids[ok])
    r = e[ok] / prec[ok]
    assert np.all(r >= 1 - 1e-14) and np.all(r < 4)


def test_eb_id_3d_invariants(orc):
    feps = float(np.finfo(np.float32).eps)
    rng = np.random.default_rng(4)
    e = 10.0 ** rng.uniform(-9, 0, 100000)
    ids = orc.eb_id_3d(e)
    a = np.log2(1 + e)
    ok = a > feps
    lo = feps * 2.0 ** ids[ok]
    assert np.all(lo <= a[ok] * (1 + 1e-15)) and np.all(a[ok] < 2 * lo)
    assert np.all(ids[~ok] == 0)
    assert orc.eb_id_3d(np.array([0.0, -1.0]))[0] == 0


# ------------------------------------------------------------- CP judge ----
def _grid_2d(r1, r2, f):
    y, x = np.mgrid[0:r1, 0:r2].astype(np.float64)
    u, v = f(x, y)
    return u.astype(np.float32), v.astype(np.float32)


@pytest.mark.parametrize("fun,typ", [
    (lambda x, y: (x - 5.3, y - 4.6), 2),            # repelling node
    (lambda x, y: (-(x - 5.3), -(y - 4.6)), 1),      # attracting node
    (lambda x, y: (x - 5.3, -(y - 4.6)), 3),         # saddle
    (lambda x, y: (-(y - 4.6), x - 5.3), 6),         # center
    (lambda x, y: (0.3 * (x - 5.3) - (y - 4.6), (x - 5.3) + 0.3 * (y - 4.6)), 5),  # repelling focus
    (lambda x, y: (-0.3 * (x - 5.3) - (y - 4.6), (x - 5.3) - 0.3 * (y - 4.6)), 4),  # attracting focus
])
def test_cp_2d_linear_fields(orc, fun, typ):
    U, V = _grid_2d(10, 12, fun)
    cps = orc.cp_2d(U, V)
    assert list(cps.values()) == [typ]
    # the CP (x=5.3, y=4.6) lies in cell (i=4, j=5) at local (0.3, 0.6): y > x puts it in
    # the triangle (i,j),(i+1,j),(i+1,j+1), id 2*(i*(r2-1)+j)
    assert list(cps.keys()) == [2 * (4 * 11 + 5)]


def _grid_3d(n, f):
    z, y, x = np.mgrid[0:n, 0:n, 0:n].astype(np.float64)
    return [c.astype(np.float32) for c in f(x, y, z)]


@pytest.mark.parametrize("fun,typ", [
    (lambda x, y, z: (1 * (x - 3.2), 2 * (y - 2.55), 3 * (z - 3.9)), 1),
    (lambda x, y, z: (-1 * (x - 3.2), -2 * (y - 2.55), -3 * (z - 3.9)), 7),
    (lambda x, y, z: (1 * (x - 3.2), -2 * (y - 2.55), 3 * (z - 3.9)), 3),
    (lambda x, y, z: (-1 * (x - 3.2), -2 * (y - 2.55), 3 * (z - 3.9)), 5),
    (lambda x, y, z: (0.2 * (x - 3.2) - (y - 2.55), (x - 3.2) + 0.2 * (y - 2.55), -(z - 3.9)), 4),
])
def test_cp_3d_linear_fields(orc, fun, typ):
    U, V, W = _grid_3d(7, fun)
    cps = orc.cp_3d(U, V, W)
    assert len(cps) == 1 and list(cps.values()) == [typ]
    sid = list(cps.keys())[0]
    assert sid // 6 == 3 * 49 + 2 * 7 + 3  # cell (i=3, j=2, k=3)


def _type_from_eigs(ev, dim):
    """Type code from eigenvalues (test/cp_extraction_2d.cpp:32-39,
    test/cp_extraction_3d.cpp:30-38)."""
    cplx = bool(np.any(np.abs(ev.imag) > 1e-9))
    if dim == 2:
        if not cplx:
            e0, e1 = sorted(ev.real)[::-1]
            return 3 if e0 * e1 < 0 else (1 if e0 < 0 else 2)
        return 4 if ev.real[0] < 0 else 5
    neg = int(np.sum(ev.real < 0))
    return 2 * neg + (2 if cplx else 1)


def test_cp_types_match_eigenvalues_random_linear(orc):
    """Random linear fields A (x - c): one critical point whose type is fixed
    by the eigenvalues of A (numpy's eigensolver); cases near a structural
    boundary (a real part or imaginary part near 0) are skipped."""
    rng = np.random.default_rng(3)
    checked = {2: 0, 3: 0}
    for trial in range(120):
        dim = 2 if trial % 2 else 3
        A = rng.standard_normal((dim, dim))
        ev = np.linalg.eigvals(A)
        if np.min(np.abs(ev.real)) < 0.05 or np.any((np.abs(ev.imag) > 0) & (np.abs(ev.imag) < 0.05)):
            continue
        c = rng.uniform(2.2, 3.8, dim)
        if dim == 2:
            U, V = _grid_2d(7, 7, lambda x, y: tuple(A[r, 0] * (x - c[0]) + A[r, 1] * (y - c[1]) for r in range(2)))
            cps = orc.cp_2d(U, V)
        else:
            U, V, W = _grid_3d(7, lambda x, y, z: tuple(A[r, 0] * (x - c[0]) + A[r, 1] * (y - c[1]) +
                                                        A[r, 2] * (z - c[2]) for r in range(3)))
            cps = orc.cp_3d(U, V, W)
        assert len(cps) >= 1
        # float32 rounding of the field can perturb J slightly; the type is robust here
        assert set(cps.values()) == {_type_from_eigs(ev, dim)}, (trial, A, ev, cps)
        checked[dim] += 1
    assert checked[2] > 20 and checked[3] > 20


def test_cp_2d_matches_float_barycentric(orc):
    """On a generic random field the exact predicate agrees with a float64
    barycentric solve wherever the latter is far from a tie."""
    from cpsz_amd.synth import field_2d
    U, V = field_2d(60, 70, 5)
    cps = orc.cp_2d(U, V)
    r1, r2 = U.shape
    found = {}
    for i in range(1, r1 - 2):
        for j in range(1, r2 - 2):
            for t, idx in enumerate(([(i, j), (i + 1, j), (i + 1, j + 1)], [(i, j), (i, j + 1), (i + 1, j + 1)])):
                u = [float(U[a]) for a in idx]
                v = [float(V[a]) for a in idx]
                if _tri_cp(u, v):
                    found[2 * (i * (r2 - 1) + j) + t] = 1
    assert set(found) == set(cps)
