"""CPU checks of the 3D learned-bound proposer: the library's host evaluation
of the exact tetrahedron predicate against the oracle's, the oracle's 3D
bisection labels against the oracle's whole-field critical-point judge, the
3D features' definition, and the numpy MLP restatement's gradient at the 3D
feature width."""
import numpy as np
import pytest

import _mlp_ref as M
import cpsz_amd as cz
from _libs import load_oracle
from cpsz_amd.synth import field_3d


@pytest.fixture(scope="module")
def orc():
    return load_oracle()


def test_tet_predicate_matches_oracle(orc):
    """Random tetrahedra (general position, tetrahedra around the origin,
    small-integer vectors with ties and zero vertices) of every judge
    tetrahedron t: existence and type agree with the oracle."""
    rng = np.random.default_rng(11)
    cases = [rng.standard_normal((4, 3)) for _ in range(300)]
    for _ in range(300):  # shifted so that the origin is often inside
        f = rng.standard_normal((4, 3))
        cases.append(f - f.mean(0) + rng.standard_normal(3) * 0.3)
    cases += [rng.integers(-2, 3, (4, 3)).astype(np.float64) for _ in range(400)]
    found = set()
    for f in cases:
        f = f.astype(np.float32)
        for t in range(6):
            a, b = cz.tet_cp_emulate(f, t), orc.tet_cp_3d(f, t)
            assert a == b, (f.tolist(), t, a, b)
            found.add(a)
    assert -1 in found and len(found) >= 5  # critical points of several types occur


def _corners(U, V, W, v, eb):
    i, j, k = v
    x = [float(A[i, j, k]) for A in (U, V, W)]
    up = 1.0 + eb
    for c in range(8):
        F = [A.copy() for A in (U, V, W)]
        for p in range(3):
            F[p][i, j, k] = np.float32(x[p] * up if (c >> p) & 1 else x[p] / up)
        yield F


def test_labels3d_match_the_critical_point_judge(orc):
    """At the label every corner move of the vertex leaves the field's critical
    points unchanged; one more doubling changes them (or the gain is capped).
    The judge is the oracle's whole-field extractor. (At eb 0.05 nearly every
    label of this small field is capped; a larger cap makes labels tight.)"""
    eb_max = 0.5
    shape = (10, 11, 12)
    U, V, W = field_3d(*shape, 3)
    E = orc.eb_3d(U, V, W, eb_max)
    base = orc.cp_3d(U, V, W)
    assert base
    near = set()
    for sid in base:
        org = sid // 6
        i, j, k = org // (11 * 12), org // 12 % 11, org % 12
        near.update((i + a, j + b, k + c) for a in (0, 1) for b in (0, 1) for c in (0, 1))
    rng = np.random.default_rng(0)
    rand = [(int(rng.integers(0, 10)), int(rng.integers(0, 11)), int(rng.integers(0, 12))) for _ in range(25)]
    pick = sorted(near)[:60] + rand
    idx = np.array([(i * 11 + j) * 12 + k for i, j, k in pick], np.uint32)
    L = orc.learn_labels_3d(U, V, W, E, eb_max, idx)
    tight = 0
    for v, g in zip(pick, L):
        e = E[v]
        if not e > 0:
            assert g == 0
            continue
        gmax = 0
        while gmax < 24 and np.ldexp(e, gmax) < eb_max:
            gmax += 1
        for F in _corners(U, V, W, v, min(eb_max, np.ldexp(e, int(g)))):
            assert orc.cp_3d(*F) == base, (v, g)
        if g < gmax:
            changed = any(orc.cp_3d(*F) != base for F in _corners(U, V, W, v, min(eb_max, np.ldexp(e, int(g) + 1))))
            assert changed, (v, g)
            tight += 1
    assert tight > 0


def test_features3d_definition(orc):
    shape = (7, 8, 9)
    U, V, W = field_3d(*shape, 4)
    E = orc.eb_3d(U, V, W, 0.05)
    X = M.bf16_to_f32(orc.learn_features_3d(U, V, W, E, 0.05)).reshape(*shape, 96)
    i, j, k = 3, 4, 5
    patch = [A[i - 1:i + 2, j - 1:j + 2, k - 1:k + 2] for A in (U, V, W)]
    s = max(np.abs(p).max() for p in patch)
    ref = M.bf16_to_f32(M.bf16_bits(np.concatenate([(p / s).ravel() for p in patch])))
    assert np.array_equal(X[i, j, k, :81], ref)
    assert np.abs(X[..., :81]).max() <= 1.0
    assert np.all(X[..., 82:] == 0)
    assert X[0, 0, 0, 0] == 0  # outside the field
    e = E[i, j, k]
    assert X[i, j, k, 81] == M.bf16_to_f32(M.bf16_bits(np.float32(max(np.float32(np.log2(e / 0.05)), -64) / 16)))


def test_mlp_reference_gradient_finite_differences_3d():
    rng = np.random.default_rng(2)
    n = M.layout(M.F3 * M.H + 3 * M.H + M.H * M.H + 1)[0][-1] + 1
    p = (rng.standard_normal(n) * 0.2).astype(np.float64)
    X = M.bf16_bits(rng.standard_normal((16, M.F3)).astype(np.float32))
    T = rng.uniform(0, 8, 16)
    g, _ = M.grad(p, X, T, bf=False)
    for k in list(rng.choice(n, 40, replace=False)) + [n - 1, 5]:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (M.grad(pp, X, T, bf=False)[1] - M.grad(pm, X, T, bf=False)[1]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-5 * max(1.0, abs(fd)), (k, fd, g[k])
