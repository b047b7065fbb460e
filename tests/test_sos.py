"""Simulation of Simplicity (Edelsbrunner & Muecke 1990) in the critical-point
judge: pins on the CPU.

* The determinant sign under the symbolic perturbation, against a third,
  independent evaluation: the perturbed determinant expanded into monomials in
  Python with exact rational arithmetic, its first nonzero coefficient in
  increasing exponent order taken. Both the oracle (generic expansion) and the
  library's host evaluation of its device predicate (a fixed term table) must
  agree with it on small-integer matrices full of ties, zeros and repeated rows.
  The test also shows that it has power: the same expansion with the vertex
  order reversed (the highest index perturbed most) disagrees on some of these
  matrices.
* What the mathematics fixes for a whole field: for a linear field whose zero
  sits exactly on a grid vertex, an edge or inside a cell, the SoS judge finds
  it in exactly one simplex, with the type the closed-simplex judge gives the
  same zero (a linear field has one Jacobian), while the closed test counts it
  in every simplex that touches it; on generic fields the two judges agree.
"""
import itertools
from fractions import Fraction

import numpy as np
import pytest

import cpsz_amd as cz
from _libs import load_oracle


@pytest.fixture(scope="module")
def orc():
    return load_oracle()


def _perm_sign(p):
    s = 1
    for i in range(len(p)):
        for j in range(i + 1, len(p)):
            if p[i] > p[j]:
                s = -s
    return s


def sos_sign_symbolic(rows, idx, order="rank"):
    """Sign of det(M + E) for infinitesimal E: E[r][c] = eps^(2^(d*rank + c)),
    rank = position of the row's vertex in increasing global index."""
    d = len(idx)
    M = [[Fraction(float(rows[r][c])) for c in range(d)] for r in range(d)]
    rank = {r: k for k, r in enumerate(sorted(range(d), key=lambda r: idx[r]))}
    coef = {}
    for perm in itertools.permutations(range(d)):
        sg = _perm_sign(perm)
        # each row takes either the entry or the perturbation of (r, perm[r])
        for pick in itertools.product((0, 1), repeat=d):
            e = 0
            val = Fraction(sg)
            for r in range(d):
                c = perm[r]
                if pick[r]:
                    k = rank[r] if order == "rank" else d - 1 - rank[r]
                    e += 1 << (d * k + c)
                else:
                    val *= M[r][c]
            coef[e] = coef.get(e, 0) + val
    for e in sorted(coef):
        if coef[e] != 0:
            return 1 if coef[e] > 0 else -1
    return 0


def _cases(d, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        m = rng.integers(-2, 3, size=(d, d)).astype(np.float32)
        kind = rng.integers(0, 4)
        if kind == 1:
            m[rng.integers(0, d)] = 0  # a zero vertex vector
        elif kind == 2:
            m[1] = m[0] * rng.integers(-1, 2)  # parallel rows
        elif kind == 3:
            m[:, rng.integers(0, d)] = 0  # a zero component
        idx = rng.permutation(50)[:d].astype(np.int64)
        out.append((m, idx))
    return out


@pytest.mark.parametrize("d", [2, 3])
def test_sos_det_matches_symbolic_expansion(orc, d):
    cases = _cases(d, 1500 if d == 2 else 700, seed=d)
    differs = 0
    for m, idx in cases:
        want = sos_sign_symbolic(m, idx)
        assert want in (-1, 1)
        assert orc.sos_det(m, idx) == want, (m, idx)
        assert cz.sos_det(m, idx) == want, (m, idx)
        differs += sos_sign_symbolic(m, idx, order="reversed") != want
    assert differs > 0  # the pin tells the vertex orders apart


def test_sos_det_equals_plain_sign_when_nonzero(orc):
    rng = np.random.default_rng(7)
    for _ in range(2000):
        d = int(rng.integers(2, 4))
        m = rng.standard_normal((d, d)).astype(np.float32)
        idx = rng.permutation(1000)[:d].astype(np.int64)
        plain = np.sign(np.linalg.det(m.astype(np.float64)))
        if plain != 0:
            assert orc.sos_det(m, idx) == plain and cz.sos_det(m, idx) == plain


def _linear_2d(r1, r2, x0, y0, J):
    """u, v = J @ (x - x0, y - y0) on the grid (x = column j, y = row i), exact
    in float for the small integers used here."""
    y, x = np.mgrid[0:r1, 0:r2].astype(np.float64)
    dx, dy = x - x0, y - y0
    U = J[0][0] * dx + J[0][1] * dy
    V = J[1][0] * dx + J[1][1] * dy
    return U.astype(np.float32), V.astype(np.float32)


@pytest.mark.parametrize("where", [(5.0, 6.0), (5.5, 6.0), (5.0, 6.5), (5.5, 6.5), (5.25, 6.5), (5.3, 6.7)],
                         ids=["vertex", "h-edge", "v-edge", "diagonal", "lower-cell-edge", "inside"])
@pytest.mark.parametrize("J", [((1, 0), (0, 1)), ((1, 2), (-3, 1)), ((2, 1), (1, -1))], ids=["id", "focus", "saddle"])
def test_sos_zero_counted_once_2d(orc, where, J):
    x0, y0 = where
    U, V = _linear_2d(14, 13, x0, y0, J)
    sos = orc.cp_2d_sos(U, V)
    closed = orc.cp_2d(U, V)
    assert len(sos) == 1, sos
    t = next(iter(sos.values()))
    # the closed test sees the same zero (where it does not skip a zero vertex
    # vector), in one or more simplices, each with the same type
    if where != (5.0, 6.0):
        assert len(closed) >= 1 and set(closed.values()) == {t}
    else:
        assert closed == {}  # the reference skips simplices with a zero vertex vector


@pytest.mark.parametrize("where", [(4, 5, 6), (4.5, 5, 6), (4.5, 5.5, 6), (4.5, 5.5, 6.5), (4.3, 5.6, 6.2)],
                         ids=["vertex", "edge", "face", "diagonal", "inside"])
def test_sos_zero_counted_once_3d(orc, where):
    r = 11
    z, y, x = np.mgrid[0:r, 0:r, 0:r].astype(np.float64)
    x0, y0, z0 = where
    J = np.array([[1, 1, 0], [0, 1, 1], [1, 0, 2]], np.float64)
    d = np.stack([x - x0, y - y0, z - z0])
    U, V, W = (np.tensordot(J[c], d, axes=1).astype(np.float32) for c in range(3))
    sos = orc.cp_3d_sos(U, V, W)
    assert len(sos) == 1, sos
    closed = orc.cp_3d(U, V, W)
    if where != (4, 5, 6):
        assert set(closed.values()) == set(sos.values())


def test_sos_agrees_with_closed_test_on_generic_fields(orc):
    from cpsz_amd.synth import field_2d
    U, V = field_2d(160, 170, seed=3)
    assert orc.cp_2d_sos(U, V) == orc.cp_2d(U, V)
    rng = np.random.default_rng(5)
    U, V, W = (rng.standard_normal((12, 13, 14)).astype(np.float32) for _ in range(3))
    assert orc.cp_3d_sos(U, V, W) == orc.cp_3d(U, V, W)
