"""Host-side parts of the HIP library, testable without a GPU: the C ABI's
exported symbols, the eb-id threshold tables and the Huffman tree builder."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import cpsz_amd as cz
from _libs import load_oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_exports_every_declared_symbol():
    decl = open(os.path.join(ROOT, "include", "cpsz.h")).read()
    names = sorted(set(re.findall(r"\b(cpsz_[a-z0-9_]+)\s*\(", decl)))
    assert len(names) >= 20
    L = cz.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_version_and_no_device_error():
    L = cz.lib()
    assert L.cpsz_version().startswith(b"cpsz")
    assert cz.device_count() >= 0


def test_dropin_header_caller_links_and_host_stage(tmp_path):
    """A caller of the reference's C++ API builds against the drop-in header; on a
    machine without a GPU the codec fails cleanly (no exit, no fallback) and says why."""
    import subprocess
    demo = os.path.join(ROOT, "tests", "cpp", "dropin_roundtrip")
    assert os.path.exists(demo), "build() compiles tests/cpp/dropin_roundtrip"
    if cz.device_count() > 0:
        pytest.skip("a GPU is present: covered by test_gpu_dropin")
    inp = tmp_path / "in.bin"
    np.zeros(2 * 30 * 40, np.float32).tofile(inp)
    r = subprocess.run([demo, "2d", str(inp), str(tmp_path / "o"), "30", "40", "0.1"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and r.stdout.startswith("fail compress") and "HIP" in r.stdout


def test_dropin_3d_double_fails_to_compile(tmp_path):
    """The double 3D pair is deleted in the drop-in: using it is a compile error,
    while the float 3D and double 2D forms compile."""
    import shutil
    import subprocess
    if not shutil.which("g++"):
        pytest.skip("no g++")
    inc = os.path.join(ROOT, "include")
    cases = {
        "bad": "double* a = nullptr; size_t n; return sz_compress_cp_preserve_3d_offline_log<double>"
               "(a, a, a, 2, 2, 2, n, false, 0.1) != nullptr;",
        "bad2": "double *u = nullptr, *v = nullptr, *w = nullptr; "
                "sz_decompress_cp_preserve_3d_online_log<double>(nullptr, 2, 2, 2, u, v, w); return 0;",
        "ok3f": "float* a = nullptr; size_t n; return sz_compress_cp_preserve_3d_offline_log<float>"
                "(a, a, a, 2, 2, 2, n, false, 0.1) != nullptr;",
        "ok2d": "double* a = nullptr; size_t n; return sz_compress_cp_preserve_2d_offline<double>"
                "(a, a, 2, 2, n, false, 0.1) != nullptr;",
    }
    for name, body in cases.items():
        src = tmp_path / f"{name}.cpp"
        src.write_text('#include "sz_cp_preserve_dropin.hpp"\nint main() { ' + body + " }\n")
        r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I" + inc, str(src)], capture_output=True,
                           text=True, timeout=120)
        if name.startswith("bad"):
            assert r.returncode != 0 and "deleted" in r.stderr, name
        else:
            assert r.returncode == 0, (name, r.stderr[:2000])


def test_dropin_lossless_matches_reference_zstd_level3():
    """The ZSTD stage is the reference's (level 3 frame, src/sz_lossless.cpp:11-45)."""
    rng = np.random.default_rng(1)
    data = rng.integers(0, 7, 100000).astype(np.uint8)
    z = cz.sz_lossless_compress(data, 3)
    assert np.array_equal(cz.sz_lossless_decompress(z, data.size), data)
    from _libs import Reference, have_reference
    if have_reference():
        assert np.array_equal(z, Reference().zstd_compress(data, 3))


def test_dropin_lossless_header_helpers():
    """is_lossless_compressed_data and the 64 KiB header read of the drop-in's C ABI
    (src/sz_lossless.cpp:3-9,47-62): frame detection, and the first 65536 bytes
    of a frame, zero-filled past a shorter frame's end."""
    L = cz.lib()
    L.cpsz_lossless_is_frame.argtypes = [C.c_void_p, C.c_size_t]
    L.cpsz_lossless_decompress_prefix.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t,
                                                  C.POINTER(C.c_size_t)]
    rng = np.random.default_rng(2)
    from _libs import Reference, have_reference
    ref = Reference() if have_reference() else None
    for n in (10, 1000, 65536, 200000):
        data = rng.integers(0, 5, n).astype(np.uint8)
        z = cz.sz_lossless_compress(data, 3)
        assert L.cpsz_lossless_is_frame(z.ctypes.data, z.size) == 1
        assert L.cpsz_lossless_is_frame(data.ctypes.data, min(n, 4)) == 0
        out = np.zeros(65536, np.uint8)
        got = C.c_size_t()
        assert L.cpsz_lossless_decompress_prefix(z.ctypes.data, z.size, out.ctypes.data, 65536,
                                                 C.byref(got)) == 0
        m = min(n, 65536)
        assert got.value == m and np.array_equal(out[:m], data[:m]) and not out[m:].any()
        if ref is not None:
            assert ref.is_lossless(z) == 1 and ref.is_lossless(data[:4]) == -1
            head = ref.zstd_decompress65536(z)
            if n <= 65536:  # a longer frame does not fit the reference's buffer at all
                assert head.size == 65536 and np.array_equal(head[:m], out[:m])


def _tables():
    L = cz.lib()
    L.cpsz_eb_tables.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.c_void_p, C.c_int,
                                 C.POINTER(C.c_int)]
    q2 = np.zeros(1024)
    q3 = np.zeros(128)
    n2, n3 = C.c_int(), C.c_int()
    assert L.cpsz_eb_tables(q2.ctypes.data, 1024, C.byref(n2), q3.ctypes.data, 128, C.byref(n3)) == 0
    return q2[:n2.value], q3[:n3.value]


def test_eb_thresholds_match_oracle_ids():
    orc = load_oracle()
    q2, q3 = _tables()
    assert np.isinf(q2[-1]) and np.isinf(q3[-1])
    t = q2[1:-1]
    ids_at, _ = orc.eb_quantize_2d(t)
    ids_below, _ = orc.eb_quantize_2d(np.nextafter(t, 0))
    k = np.arange(1, t.size + 1)
    assert np.array_equal(ids_at, k) and np.array_equal(ids_below, k - 1)
    t3 = q3[1:-1]
    k3 = np.arange(1, t3.size + 1)
    assert np.array_equal(orc.eb_id_3d(t3), k3)
    assert np.array_equal(orc.eb_id_3d(np.nextafter(t3, 0)), k3 - 1)


def _tree(freq, state_num):
    L = cz.lib()
    L.cpsz_huffman_tree.argtypes = [C.c_void_p, C.c_size_t, C.c_uint32, C.c_void_p, C.c_size_t,
                                    C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    buf = np.zeros(1 << 22, np.uint8)
    ts, nc = C.c_size_t(), C.c_size_t()
    f = np.ascontiguousarray(freq, np.uint64)
    rc = L.cpsz_huffman_tree(f.ctypes.data, f.size, state_num, buf.ctypes.data, buf.size, C.byref(ts),
                             C.byref(nc))
    assert rc == 0
    return buf[:ts.value].copy(), nc.value


@pytest.mark.parametrize("seed", range(8))
def test_huffman_tree_matches_oracle(seed):
    orc = load_oracle()
    rng = np.random.default_rng(seed)
    state = [2048, 131072, 512][seed % 3]
    k = int(rng.integers(1, 2000 if state > 600 else 400))
    n = int(rng.integers(1, 50000))
    syms = (rng.geometric(0.02, n) % k + state // 4).astype(np.int32)
    freq = np.bincount(syms, minlength=2 * state)
    tree, nodes = _tree(freq, state)
    sec = orc.huff_encode(state, syms)
    node_count = int(np.frombuffer(sec[:8].tobytes(), np.uint64)[0])
    tree_size = int(np.frombuffer(sec[8:12].tobytes(), np.uint32)[0])
    assert nodes == node_count
    assert np.array_equal(tree, sec[12:12 + tree_size])


def test_huffman_tree_builds_in_sequence_match_oracle():
    """Tree builds reuse per-thread working storage: alternating large and
    small alphabets on one thread must each still match the oracle's tree."""
    orc = load_oracle()
    rng = np.random.default_rng(77)
    for it in range(10):
        state = [131072, 512][it % 2]
        k = int(rng.integers(20000, 30000)) if state > 600 else int(rng.integers(1, 60))
        syms = (rng.geometric(0.001 if state > 600 else 0.1, 60000) % k + state // 4).astype(np.int32)
        freq = np.bincount(syms, minlength=2 * state)
        tree, nodes = _tree(freq, state)
        sec = orc.huff_encode(state, syms)
        tree_size = int(np.frombuffer(sec[8:12].tobytes(), np.uint32)[0])
        assert nodes == int(np.frombuffer(sec[:8].tobytes(), np.uint64)[0])
        assert np.array_equal(tree, sec[12:12 + tree_size])


def _decode_host(tree, node_count, bits, n):
    L = cz.lib()
    L.cpsz_huffman_decode_host.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_void_p, C.c_size_t,
                                           C.c_size_t, C.c_void_p]
    tree = np.ascontiguousarray(tree)
    bits = np.ascontiguousarray(bits)
    out = np.empty(max(n, 1), np.uint32)
    rc = L.cpsz_huffman_decode_host(tree.ctypes.data, tree.size, node_count, bits.ctypes.data, bits.size, n,
                                    out.ctypes.data)
    return rc, out[:n]


def test_huffman_decoder_rejects_malformed_trees():
    """Tree validation before the decode tables are built: children out of
    range, pointing backwards, shared by two parents or equal to each other are
    errors, never a crash; the untouched tree still decodes."""
    orc = load_oracle()
    rng = np.random.default_rng(5)
    state = 2048
    syms = (rng.geometric(0.002, 20000) % 1000 + 512).astype(np.int32)
    sec = orc.huff_encode(state, syms)
    nc = int(np.frombuffer(sec[:8].tobytes(), np.uint64)[0])
    ts = int(np.frombuffer(sec[8:12].tobytes(), np.uint32)[0])
    tree = sec[12:12 + ts].copy()
    nbytes = int(np.frombuffer(sec[12 + ts:20 + ts].tobytes(), np.uint64)[0])
    bits = sec[20 + ts:20 + ts + nbytes].copy()
    rc, out = _decode_host(tree, nc, bits, syms.size)
    assert rc == 0 and np.array_equal(out, syms.astype(np.uint32))
    assert nc > 256 and nc <= 65536  # u16 child indices
    left = tree[1:1 + 2 * nc].view(np.uint16)
    right = tree[1 + 2 * nc:1 + 4 * nc].view(np.uint16)
    internal = np.flatnonzero(left[1:] != 0) + 1
    p = int(internal[len(internal) // 2])
    for field, value in [("left", nc + 5), ("right", 0xFFFF), ("left", p), ("right", int(left[p])),
                         ("right", int(right[0]))]:
        bad = tree.copy()
        arr = (bad[1:1 + 2 * nc] if field == "left" else bad[1 + 2 * nc:1 + 4 * nc]).view(np.uint16)
        arr[p] = value
        rc, _ = _decode_host(bad, nc, bits, syms.size)
        assert rc != 0, (field, value)


def _emulate(x, c, wide):
    L = cz.lib()
    L.cpsz_lorenzo2d_emulate.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t, C.c_int,
                                         C.c_void_p, C.c_void_p]
    r1, r2 = x.shape
    q = np.empty((r1, r2), np.uint16)
    rec = np.empty((r1, r2), np.float32)
    x = np.ascontiguousarray(x, np.float32)
    c = np.ascontiguousarray(c, np.uint16)
    assert L.cpsz_lorenzo2d_emulate(x.ctypes.data, c.ctypes.data, r1, r2, int(wide), q.ctypes.data,
                                    rec.ctypes.data) == 0
    return q, rec


def _block_major(a, bs=12):
    r1, r2 = a.shape
    out = []
    for bi in range(0, r1, bs):
        for bj in range(0, r2, bs):
            out.append(a[bi:bi + bs, bj:bj + bs].ravel())
    return np.concatenate(out)


@pytest.mark.parametrize("shape,seed,eb", [((45, 38), 1, 0.1), ((97, 130), 2, 0.05),
                                           ((64, 64), 3, 0.3), ((30, 200), 4, 0.01)])
@pytest.mark.parametrize("wide", [False, True])
def test_lorenzo2d_arithmetic_matches_oracle(shape, seed, eb, wide):
    """The device step arithmetic (run on the host) reproduces the reference codes."""
    from cpsz_amd.synth import field_2d
    orc = load_oracle()
    U, V = field_2d(*shape, seed)
    ebv = orc.eb_2d(U, V, eb)
    for X in (U, V):
        ids, prec = orc.eb_quantize_2d(np.abs(X).astype(np.float64) * ebv)
        c = np.where(prec == 0, 0, ids + 1).astype(np.uint16)
        q, rec = _emulate(X, c, wide)
        _, codes, orec = orc.lorenzo_2d(X, prec)
        assert np.array_equal(_block_major(q.astype(np.int32)), codes)
        assert np.array_equal(rec.view(np.uint32), orec.view(np.uint32))


def test_lorenzo2d_arithmetic_extreme_bounds():
    """Tiny, zero and eps-window precisions, large values: FP32 and double paths agree."""
    rng = np.random.default_rng(7)
    x = (rng.normal(size=(40, 50)) * 10.0 ** rng.integers(-20, 20, size=(40, 50))).astype(np.float32)
    c = rng.integers(0, 70, size=(40, 50)).astype(np.uint16)
    qa, ra = _emulate(x, c, False)
    qb, rb = _emulate(x, c, True)
    assert np.array_equal(qa, qb) and np.array_equal(ra.view(np.uint32), rb.view(np.uint32))


def _libm_eval(name, x):
    L = cz.lib()
    f = getattr(L, name)
    f.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p]
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(x)
    assert f(x.ctypes.data, x.size, out.ctypes.data) == 0
    return out


def _glibc(name, x):
    m = C.CDLL("libm.so.6")
    f = getattr(m, name)
    f.argtypes = [C.c_float]
    f.restype = C.c_float
    return np.array([f(float(v)) for v in x], np.float32)


def test_log2f_matches_host_libm():
    rng = np.random.default_rng(11)
    # random bit patterns over every positive binade, plus a dense slice of
    # [0.98, 1.05] where the host log2f deviates most often
    bits = rng.integers(1, 0x7F800000, size=300000, dtype=np.uint32)
    dense = np.arange(np.float32(0.98).view(np.uint32), np.float32(1.05).view(np.uint32), 7,
                      dtype=np.uint32)
    x = np.concatenate([bits, dense[:400000]]).view(np.float32)
    got = _libm_eval("cpsz_eval_log2f", x)
    want = _glibc("log2f", x)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_exp2f_matches_host_libm():
    rng = np.random.default_rng(12)
    y = np.concatenate([rng.uniform(-99, 128, 300000), rng.uniform(-1, 1, 200000)]).astype(np.float32)
    got = _libm_eval("cpsz_eval_exp2f", y)
    want = _glibc("exp2f", y)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def _parse_3d(orc, s, n):
    """(ids, codes) of a 3D offline_log stream."""
    smap = (n - 1) // 8 + 1
    off = 8 + 3 * smap
    nunp = int(np.frombuffer(s[off:off + 8].tobytes(), np.uint64)[0])
    off += 8 + 4 * nunp
    off += 8  # eb_quant_num
    ids, used = orc.huff_decode(s[off:], n)
    off += used
    nq = int(np.frombuffer(s[off:off + 8].tobytes(), np.uint64)[0])
    codes, _ = orc.huff_decode(s[off + 8:], nq)
    return ids, codes


def _emulate_3d(U, V, W, eb):
    L = cz.lib()
    r1, r2, r3 = U.shape
    n = U.size
    L.cpsz_eb3d_emulate.argtypes = [C.c_void_p] * 3 + [C.c_size_t] * 3 + [C.c_double, C.c_void_p]
    L.cpsz_lorenzo3d_emulate.argtypes = [C.c_void_p, C.c_void_p] + [C.c_size_t] * 3 + [C.c_void_p] * 2
    L.cpsz_eb_code_3d.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p]
    e = np.empty(n, np.float64)
    assert L.cpsz_eb3d_emulate(U.ctypes.data, V.ctypes.data, W.ctypes.data, r1, r2, r3, eb, e.ctypes.data) == 0
    code = np.empty(n, np.uint8)
    assert L.cpsz_eb_code_3d(e.ctypes.data, n, code.ctypes.data) == 0
    logs = []
    for X in (U, V, W):
        a = np.abs(X.ravel())
        lg = _libm_eval("cpsz_eval_log2f", np.where(a == 0, 1, a))
        logs.append(np.where(a == 0, np.float32(-100), lg))
    Lx = np.ascontiguousarray(np.concatenate(logs), np.float32)
    vid = np.empty(n, np.uint8)
    q = np.empty(3 * n, np.uint16)
    assert L.cpsz_lorenzo3d_emulate(Lx.ctypes.data, code.ctypes.data, r1, r2, r3, vid.ctypes.data,
                                    q.ctypes.data) == 0
    return e, code, vid, q.reshape(3, n)


@pytest.mark.parametrize("shape,seed,eb", [((12, 13, 14), 1, 0.05), ((9, 17, 33), 2, 0.05),
                                           ((16, 16, 16), 3, 0.1), ((7, 40, 9), 4, 0.02)])
def test_3d_arithmetic_matches_oracle(shape, seed, eb):
    from cpsz_amd.synth import field_3d
    orc = load_oracle()
    U, V, W = field_3d(*shape, seed)
    e, code, vid, q = _emulate_3d(U, V, W, eb)
    eo = orc.eb_3d(U, V, W, eb).ravel()
    assert np.array_equal(e.view(np.uint64), eo.view(np.uint64))
    ido = orc.eb_id_3d(eo)
    pos = eo > 0
    assert np.array_equal(code[pos].astype(np.int32), ido[pos]) and np.all(code[~pos] == 255)
    s = orc.compress_3d(U, V, W, eb)
    ids, codes = _parse_3d(orc, s, U.size)
    assert np.array_equal(vid.astype(np.int32), ids)
    pred = (vid != 0) & (vid != 24)
    assert np.array_equal(q[:, pred].T.ravel().astype(np.int32), codes)


@pytest.mark.parametrize("seed", range(6))
def test_two_level_decode_tables_match_oracle_encoder(seed):
    """The GPU decoder's two-level tables (11-bit level 1, level-2 tables over
    up to 12 more bits, tree walk beyond), run on the host, decode the oracle's
    Huffman sections exactly, including codes longer than 24 bits."""
    orc = load_oracle()
    rng = np.random.default_rng(100 + seed)
    state = [131072, 2048, 512][seed % 3]
    if seed < 3:  # broad alphabets: most codes 10-20 bits
        k = min(int(rng.integers(2000, 30000)), state - state // 4 - 1)  # symbols stay below state
        syms = (rng.zipf(1.3, 200000) % k + state // 4).astype(np.int32)
    else:  # very skewed: code lengths up to ~40 bits
        syms = (np.minimum(rng.geometric(0.45, 100000), 60) + state // 4).astype(np.int32)
    sec = orc.huff_encode(state, syms)
    node_count = int(np.frombuffer(sec[:8].tobytes(), np.uint64)[0])
    tree_size = int(np.frombuffer(sec[8:12].tobytes(), np.uint32)[0])
    tree = sec[12:12 + tree_size]
    nbytes = int(np.frombuffer(sec[12 + tree_size:20 + tree_size].tobytes(), np.uint64)[0])
    bits = np.ascontiguousarray(sec[20 + tree_size:20 + tree_size + nbytes])
    out = np.empty(syms.size, np.uint32)
    L = cz.lib()
    L.cpsz_huffman_decode_host.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_void_p, C.c_size_t,
                                           C.c_size_t, C.c_void_p]
    tree = np.ascontiguousarray(tree)
    assert L.cpsz_huffman_decode_host(tree.ctypes.data, tree.size, node_count, bits.ctypes.data, nbytes,
                                      syms.size, out.ctypes.data) == 0
    assert np.array_equal(out, syms.astype(np.uint32))


def test_eb_table_2d_log_matches_reference_quantizer():
    """The device's offline_log code table against log2 of this host's libm and
    the reference's eb_exponential_quantize (include/sz_cp_preserve_utils.hpp:14-22,
    base 2 over DBL_EPSILON, src/sz_compress_cp_preserve_2d.cpp:327-331): at every
    threshold and its predecessor, and on random bounds over 40 decades."""
    L = cz.lib()
    q = np.zeros(128)
    n = C.c_int()
    assert L.cpsz_eb_table_2d_log(q.ctypes.data, 128, C.byref(n)) == 0
    q = q[:n.value]
    assert np.isinf(q[-1]) and np.all(np.diff(q[1:-1]) > 0)
    libm = C.CDLL("libm.so.6")
    libm.log2.restype = C.c_double
    libm.log2.argtypes = [C.c_double]
    rng = np.random.default_rng(3)
    e = list(10.0 ** rng.uniform(-17, 3, 4000))
    for t in q[1:-1]:
        e += [t, np.nextafter(t, 0)]
    e = np.array(e)
    a = np.array([libm.log2(1 + x) for x in e])
    from _libs import Reference, have_reference
    quant = Reference() if have_reference() else load_oracle()
    ids, _ = quant.eb_exp_quantize(a, 2, 1.0, np.finfo(np.float64).eps)
    want = np.where(a <= np.finfo(np.float64).eps, 0, ids + 1)
    got = np.searchsorted(q[1:-1], e, side="right")  # number of thresholds <= e
    assert np.array_equal(got, want)
