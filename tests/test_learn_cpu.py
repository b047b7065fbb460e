"""CPU checks of the learned-bound proposer's checkers: the oracle's bisection
labels against the oracle's whole-field critical-point judge, the features'
definition, and the numpy MLP restatement's gradient against finite
differences."""
import numpy as np
import pytest

import _mlp_ref as M
from _libs import load_oracle
from cpsz_amd.synth import field_2d


@pytest.fixture(scope="module")
def orc():
    return load_oracle()


def _corner_fields(U, V, i, j, eb):
    u, v = float(U[i, j]), float(V[i, j])
    du, dv = abs(u) * eb, abs(v) * eb
    for su in (-1, 1):
        for sv in (-1, 1):
            U2, V2 = U.copy(), V.copy()
            U2[i, j] = np.float32(u + su * du)
            V2[i, j] = np.float32(v + sv * dv)
            yield U2, V2


def test_labels_match_the_critical_point_judge(orc):
    """At the label's gain every corner move of the vertex leaves the field's
    critical points (existence and type) unchanged; one more doubling changes
    them (or the gain is capped). The judge is the oracle's whole-field
    extractor, independent of the labeler's local per-triangle test."""
    eb_max = 0.1
    U, V = field_2d(40, 44, 3)
    E = orc.eb_2d(U, V, eb_max)
    rng = np.random.default_rng(0)
    base = orc.cp_2d(U, V)
    inside = lambda i, j: 2 <= i <= 37 and 2 <= j <= 41  # noqa: E731  all 6 incident cells judged
    # vertices around critical-point cells (where labels are tight) plus random ones
    near = set()
    for sid in base:
        ci, cj = divmod(sid // 2, 43)
        near.update((ci + a, cj + b) for a in (-1, 0, 1, 2) for b in (-1, 0, 1, 2) if inside(ci + a, cj + b))
    verts = [(i, j) for i in range(2, 38) for j in range(2, 42)]
    pick = sorted(near) + [verts[k] for k in rng.choice(len(verts), 60, replace=False)]
    idx = np.array([i * 44 + j for i, j in pick], np.uint32)
    L = orc.learn_labels_2d(U, V, E, eb_max, idx)
    assert L.max() > 0  # the analytic bound is conservative somewhere
    tight = 0
    for (i, j), g in zip(pick, L):
        e = E[i, j]
        if not e > 0:
            assert g == 0
            continue
        gmax = 0
        while gmax < 24 and np.ldexp(e, gmax) < eb_max:
            gmax += 1
        for U2, V2 in _corner_fields(U, V, i, j, min(eb_max, np.ldexp(e, int(g)))):
            assert orc.cp_2d(U2, V2) == base, (i, j, g)
        if g < gmax:
            changed = any(orc.cp_2d(U2, V2) != base
                          for U2, V2 in _corner_fields(U, V, i, j, min(eb_max, np.ldexp(e, int(g) + 1))))
            assert changed, (i, j, g)
            tight += 1
    assert tight > 0


def test_features_definition(orc):
    U, V = field_2d(20, 23, 4)
    E = orc.eb_2d(U, V, 0.1)
    X = M.bf16_to_f32(orc.learn_features_2d(U, V, E, 0.1)).reshape(20, 23, 64)
    i, j = 7, 9
    pu, pv = U[i - 2:i + 3, j - 2:j + 3], V[i - 2:i + 3, j - 2:j + 3]
    s = max(np.abs(pu).max(), np.abs(pv).max())
    ref = M.bf16_to_f32(M.bf16_bits(np.concatenate([(pu / s).ravel(), (pv / s).ravel()])))
    assert np.array_equal(X[i, j, :50], ref)
    assert np.abs(X[:, :, :50]).max() <= 1.0
    assert np.all(X[:, :, 51:] == 0)
    assert X[0, 0, 0] == 0  # outside the field
    e = E[i, j]
    assert X[i, j, 50] == M.bf16_to_f32(M.bf16_bits(np.float32(max(np.float32(np.log2(e / 0.1)), -64) / 16)))


def test_mlp_reference_gradient_finite_differences():
    rng = np.random.default_rng(1)
    p = (rng.standard_normal(M.N_PARAMS) * 0.2).astype(np.float64)
    X = M.bf16_bits(rng.standard_normal((16, 64)).astype(np.float32))
    T = rng.uniform(0, 8, 16)
    g, _ = M.grad(p, X, T, bf=False)
    for k in list(rng.choice(M.N_PARAMS, 40, replace=False)) + [M.OFF_B3, M.OFF_W3 + 5, M.OFF_B1 + 3]:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (M.grad(pp, X, T, bf=False)[1] - M.grad(pm, X, T, bf=False)[1]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-5 * max(1.0, abs(fd)), (k, fd, g[k])
