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
