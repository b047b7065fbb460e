"""Parity pins of the 3D offline_log compressor's oracle.

The reference compressor (src/sz_compress_cp_preserve_3d.cpp) does not build
here: it includes an FTK header. Its header-only helpers do
(include/sz_cp_preserve_utils.hpp), so they are compiled from the reference
header into oracle/_ref and compared with the oracle bit for bit. The rest is
pinned by what the mathematics fixes:

* the per-tetrahedron bound (:264-358) on integer-valued tetrahedra, where every
  triple product and cofactor sum is an exact integer: the cofactors are the
  signed 3x3 determinants of the other three vectors, their positive and
  negative Leibniz terms are summed exactly, and the expected bound follows from
  the reference's sign bound on the float-rounded sums;
* the per-vertex bound (:376-418) as the minimum over the Kuhn tetrahedra of
  every cell that holds the vertex, each taken in path order (first step,
  second step, far corner, origin) — generated from the axis permutations, not
  from a table;
* the driver (:423-547) through the compiled reference decoder: every vertex
  the stream marks predictable decodes within its quantized bound, strictly,
  and keeps its sign; every other vertex is stored exactly.

Each pin takes the oracle (tests/_libs.Oracle) and the compiled reference
(tests/_libs.Reference) and raises AssertionError on a mismatch. The oracle's
mutant builds (ORC_MUTANT in oracle/orc_3d.cpp) must each fail at least one.
"""
import ctypes as C
import itertools

import numpy as np

FEPS = float(np.finfo(np.float32).eps)
ID_MAX = 24  # (int)(log2(1 / FLT_EPSILON) / log2(2)) + 1, :458


def _bits(a):
    a = np.asarray(a)
    return a.view({4: np.uint32, 8: np.uint64}[a.dtype.itemsize])


def _same_bits(a, b):
    """Equal bit patterns, any NaN equal to any NaN."""
    a, b = np.asarray(a), np.asarray(b)
    nan = np.isnan(a) & np.isnan(b)
    return np.all((_bits(a) == _bits(b)) | nan)


# ---------------------------------------------------------------- helpers ----
def pin_acbrt1_exhaustive(orc, ref):
    """acbrt1 (:173-184) on every non-negative float bit pattern up to +inf."""
    fp = C.cast(orc.lib.orc_acbrt1_one, C.c_void_p).value
    bad, first = ref.acbrt1_compare(fp, 0, 0x7F800001)
    assert bad == 0, f"acbrt1 differs on {bad} inputs, first bit pattern {first:#x}"


def _sign_bound_inputs(rng, n=400000):
    e = rng.uniform(-44, 38, (2, n))
    p = (10.0 ** e[0]).astype(np.float32)
    q = (10.0 ** e[1]).astype(np.float32)
    near = p * (1 + rng.uniform(-1e-6, 1e-6, n)).astype(np.float32)  # P close to N
    tiny = np.array([0, 1e-45, 1e-40, 1e-38, 1.2e-38, 3e38, np.inf], np.float32)
    grid_p, grid_n = np.meshgrid(tiny, tiny)
    pos = np.concatenate([p, p, p, np.zeros(100, np.float32), grid_p.ravel()])
    neg = np.concatenate([q, near, np.zeros(n, np.float32), q[:100], grid_n.ravel()])
    return pos, neg


def pin_sign_bound(orc, ref):
    """max_eb_to_keep_sign_3d_offline (:186-191) on random, near-equal, zero and
    extreme (subnormal, huge, infinite) argument pairs."""
    pos, neg = _sign_bound_inputs(np.random.default_rng(11))
    a, b = orc.sign_bound_3d(pos, neg), ref.sign_bound_3d(pos, neg)
    assert _same_bits(a, b), "sign bound differs"


def pin_same_direction(orc, ref):
    """same_direction (:52-57) on every sign pattern of four values, signed
    zeros and NaN included."""
    vals = np.array([-2.5, -0.0, 0.0, 1.5, np.nan], np.float32)
    q4 = np.array(list(itertools.product(vals, repeat=4)), np.float32)
    assert np.array_equal(orc.same_direction_3d(q4), ref.same_direction(q4)), "same_direction differs"


def pin_eb_quantize(orc, ref):
    """eb_exponential_quantize (:14-22) with the 3D compressor's arguments (base
    2, log2(2), FLT_EPSILON) on log2(1 + eb) of random bounds, and on values
    within a few ulps of every id threshold 2^k * FLT_EPSILON."""
    rng = np.random.default_rng(12)
    rnd = np.log2(1 + 10.0 ** rng.uniform(-9, 0.5, 200000))
    thr = FEPS * 2.0 ** np.arange(0, 30)
    near = (thr[:, None] * (1 + np.arange(-6, 7)[None, :] * np.finfo(np.float64).eps)).ravel()
    eb = np.concatenate([rnd, near, [0.0, FEPS, -1.0, np.log2(2.0)]])
    i0, e0 = orc.eb_exp_quantize(eb, 2, 1.0, FEPS)
    i1, e1 = ref.eb_exp_quantize(eb, 2, 1.0, FEPS)
    assert np.array_equal(i0, i1) and _same_bits(e0, e1), "eb_exponential_quantize differs"


def pin_log_transform(orc, ref):
    """log_transform<float> (:205-221): log2f|x| and sign, -100 for zeros."""
    rng = np.random.default_rng(13)
    x = (rng.choice([-1, 1], 300000) * 10.0 ** rng.uniform(-44, 38, 300000)).astype(np.float32)
    x[::97] = 0
    x[1::97] = -0.0
    x = np.concatenate([x, np.array([np.inf, -np.inf, 1e-45, -1e-45, 1.0, -1.0], np.float32)])
    l0, s0 = orc.log_transform_3d(x)
    l1, s1 = ref.log_transform(x)
    assert _same_bits(l0, l1) and np.array_equal(s0, s1), "log_transform differs"


# ------------------------------------------------------- per-tet bound ------
def _leibniz_parts(a, b, c):
    """Sums of the positive and of the negative terms of det[a; b; c] (rows are
    vectors), exact for integer inputs."""
    pos = neg = 0
    for perm in itertools.permutations(range(3)):
        sgn = 1
        for i in range(3):
            for j in range(i + 1, 3):
                if perm[i] > perm[j]:
                    sgn = -sgn
        t = sgn * a[perm[0]] * b[perm[1]] * c[perm[2]]
        if t >= 0:
            pos += t
        else:
            neg -= t
    return pos, neg


def _expected_tet_bound(vec, ref):
    """Bound of one integer-valued tetrahedron (4 vectors (u, v, w))."""
    # M_k = (-1)^(k+1) det of the other three vectors in index order
    parts, M = [], []
    for k in range(4):
        rows = [vec[i] for i in range(4) if i != k]
        p, n = _leibniz_parts(*rows)
        if k % 2 == 0:
            p, n = n, p
        parts.append((p, n))
        M.append(p - n)
    Mt = sum(M)
    if Mt == 0:
        return 0.0
    # M / Mk > 1 in double is exact here: same sign and |M| > |Mk|
    flag = [m == 0 or (m * Mt > 0 and abs(Mt) > abs(m)) for m in M]
    if all(flag):
        def same(c):
            s = [np.sign(vec[i][c]) for i in range(4)]
            return s[0] != 0 and all(x == s[0] for x in s)
        return 1.0 if same(0) or same(1) or same(2) else 0.0
    P = [np.float32(p) for p, _ in parts]
    N = [np.float32(n) for _, n in parts]
    eb = np.float32(0)
    for k in range(4):
        if flag[k]:
            continue
        o = [i for i in range(4) if i != k]
        ps = np.float32(np.float32(P[o[0]] + P[o[1]]) + P[o[2]])
        ns = np.float32(np.float32(N[o[0]] + N[o[1]]) + N[o[2]])
        e = ref.sign_bound_3d(np.array([P[k], ps], np.float32), np.array([N[k], ns], np.float32))
        eb = max(eb, min(e[0], e[1]))
    return float(eb)


def _integer_tets(rng, n):
    mag = rng.integers(120, 256, (n, 12))
    sgn = rng.choice([-1, 1], (n, 12))
    t = mag * sgn
    t[: n // 4] = rng.integers(-9, 10, (n // 4, 12))  # ties and zero cofactors
    t[n // 4: n // 3, :4] = np.abs(t[n // 4: n // 3, :4])  # same-signed u
    return t


def pin_tet_bound_integer(orc, ref, n=3000):
    """max_eb_to_keep_position_and_type_3d_offline (:264-358) on integer-valued
    tetrahedra against the exact cofactor construction."""
    t = _integer_tets(np.random.default_rng(14), n)
    got = orc.tet_bound_3d(t.astype(np.float32))
    for row, g in zip(t, got):
        vec = [(int(row[i]), int(row[4 + i]), int(row[8 + i])) for i in range(4)]
        want = _expected_tet_bound(vec, ref)
        assert _bits(np.float64(g)) == _bits(np.float64(want)), (row.tolist(), g, want)


# --------------------------------------------------- per-vertex bounds ------
def kuhn_tets():
    """The six tetrahedra of a unit cell that share its main diagonal, one per
    axis order: vertices (first step, second step, far corner, origin), each an
    offset (k, j, i) with k fastest."""
    out = []
    for a, b, _ in itertools.permutations(range(3)):
        e1 = [0, 0, 0]
        e1[a] = 1
        e2 = list(e1)
        e2[b] = 1
        out.append((tuple(e1), tuple(e2), (1, 1, 1), (0, 0, 0)))
    return out


def expected_vertex_bounds(orc, U, V, W, max_pwr_eb):
    r1, r2, r3 = U.shape
    eb = np.full(U.shape, max_pwr_eb, np.float64)
    ci, cj, ck = np.meshgrid(np.arange(r1 - 1), np.arange(r2 - 1), np.arange(r3 - 1), indexing="ij")
    ci, cj, ck = ci.ravel(), cj.ravel(), ck.ravel()
    for tet in kuhn_tets():
        idx = [(ci + o[2], cj + o[1], ck + o[0]) for o in tet]
        uvw = np.stack([F[ix] for F in (U, V, W) for ix in idx], axis=1)
        b = orc.tet_bound_3d(uvw)
        for ix in idx:
            np.minimum.at(eb, ix, b)
    return eb


def _fields(rng):
    """Normal, small-integer (ties, zero cofactors) and wide-range fields; in the
    last, triple products span more than 2^29, so the double cofactor sums round
    and their order of accumulation shows in the bits."""
    out = []
    for shape in [(6, 7, 9), (5, 9, 4), (2, 3, 11)]:
        out.append([rng.normal(size=shape).astype(np.float32) for _ in range(3)])
        out.append([rng.integers(-3, 4, shape).astype(np.float32) for _ in range(3)])
        out.append([(rng.choice([-1, 1], shape) * 10.0 ** rng.uniform(-8, 8, shape)).astype(np.float32)
                    for _ in range(3)])
    return out


def pin_vertex_bounds(orc, ref):
    """Per-vertex bound (:376-418): minimum of max_pwr_eb and the bounds of the
    Kuhn tetrahedra holding the vertex."""
    rng = np.random.default_rng(15)
    for U, V, W in _fields(rng):
        for eb in (0.05, 1.0):
            got = orc.eb_3d(U, V, W, eb)
            want = expected_vertex_bounds(orc, U, V, W, eb)
            assert np.array_equal(_bits(got), _bits(want)), (U.shape, eb)


# --------------------------------------------------------------- driver ------
def stream_ids_3d(stream, n, ref):
    """Per-vertex eb ids of an offline_log stream (layout :556-580), decoded by
    the compiled reference Huffman decoder."""
    smap = (n - 1) // 8 + 1
    off = 8 + 3 * smap
    nunp = int(np.frombuffer(stream[off:off + 8].tobytes(), np.uint64)[0])
    off += 8 + 4 * nunp
    assert int(np.frombuffer(stream[off:off + 8].tobytes(), np.uint64)[0]) == n
    ids, _ = ref.huff_decode(512, stream[off + 8:], n)
    return ids


def quarter_power_field(rng, shape):
    """Components +-2^(a/4), integer a with |a/4| in [2, 15]: their log2f is
    exactly a/4, so Lorenzo differences in the log domain are multiples of 1/4
    and land exactly on the quantization boundaries of ids 21 and 22."""
    out = []
    for _ in range(3):
        a = rng.integers(8, 61, shape) * rng.choice([-1, 1], shape)
        v = (2.0 ** (a / 4.0)).astype(np.float32) * rng.choice([-1, 1], shape).astype(np.float32)
        out.append(v)
    return out


def _check_decode(orc, ref, U, V, W, eb, exact_logs):
    """exact_logs: the decoded log2 values are exactly recoverable (quarter
    powers), so the bound is checked strictly; otherwise log2f(exp2f(d)) may
    sit an ulp or two from d, and that much slack is allowed."""
    s = orc.compress_3d(U, V, W, eb)
    n = U.size
    ids = stream_ids_3d(s, n, ref).reshape(U.shape)
    dec = ref.decompress_3d(s, *U.shape)
    pred = (ids != 0) & (ids != ID_MAX)
    bound = np.where(pred, FEPS * 2.0 ** np.where(pred, ids, 0), 0)
    for x, y in zip((U, V, W), dec):
        assert np.array_equal(_bits(x[~pred]), _bits(y[~pred])), "unpredictable vertex not stored exactly"
        xs, ys = x[pred], y[pred]
        assert np.all(np.sign(xs) == np.sign(ys)), "sign lost"
        nz = xs != 0
        lx = ref.log_transform(xs[nz])[0].astype(np.float64)  # glibc log2f, as the codec
        ly = ref.log_transform(ys[nz])[0].astype(np.float64)
        slack = 0.0 if exact_logs else 4 * np.spacing(np.abs(lx).astype(np.float32)).astype(np.float64)
        assert np.all(np.abs(ly - lx) < bound[pred][nz] + slack), "decoded value on or beyond its bound"
    return int(pred.sum())


def pin_decode_within_bound(orc, ref):
    """Driver (:423-547) through the compiled reference decoder: predictable
    vertices decode strictly inside their quantized log-domain bound (the
    reference rejects an error equal to the bound, :507), with their sign;
    others are stored exactly. Quarter-power fields hit the equality case."""
    rng = np.random.default_rng(16)
    for shape in [(7, 9, 11), (4, 13, 6)]:
        for eb in (0.3, 0.9):
            assert _check_decode(orc, ref, *quarter_power_field(rng, shape), eb, True) > 0
    from cpsz_amd.synth import field_3d
    _check_decode(orc, ref, *field_3d(9, 10, 12, 3), 0.05, False)


def pin_stream_ids(orc, ref):
    """Stream eb ids (:466-547): the quantized vertex bound, or the
    unpredictable id (24) when a component does not quantize; id 0 exactly for
    a bound at or below FLT_EPSILON after log2(1 + eb)."""
    rng = np.random.default_rng(17)
    for U, V, W in _fields(rng)[:2]:
        s = orc.compress_3d(U, V, W, 0.05)
        ids = stream_ids_3d(s, U.size, ref)
        b = expected_vertex_bounds(orc, U, V, W, 0.05).ravel()
        log2 = C.CDLL("libm.so.6").log2  # the reference's double log2
        log2.argtypes, log2.restype = [C.c_double], C.c_double
        want = np.zeros(b.size, np.int32)
        pos = b > 0
        want[pos], _ = ref.eb_exp_quantize(np.array([log2(1 + x) for x in b[pos]]), 2, 1.0, FEPS)
        # bound 0: id 0; else the quantized id, or the unpredictable id when a
        # component does not quantize or the id is 0 (:470-520)
        ok = np.where(pos, ((ids == want) & (want > 0)) | (ids == ID_MAX), ids == 0)
        assert np.all(ok), "stream id is neither the quantized bound nor the unpredictable id"


PINS = [pin_acbrt1_exhaustive, pin_sign_bound, pin_same_direction, pin_eb_quantize, pin_log_transform,
        pin_tet_bound_integer, pin_vertex_bounds, pin_stream_ids, pin_decode_within_bound]
