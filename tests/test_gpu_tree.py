"""Huffman trees built on the GPU against the host builder (whose trees equal
the oracle's and the compiled reference's: tests/test_host.py): tree bytes and
node counts byte-equal on tie-heavy alphabets of 1 to 60000 symbols, and the
device codebook equal to the codes the tree bytes define (preorder walk:
left 0, right 1)."""
import ctypes as C

import numpy as np
import pytest

import cpsz_amd as cz

pytestmark = pytest.mark.gpu


def _host_tree(freq, state):
    L = cz.lib()
    L.cpsz_huffman_tree.argtypes = [C.c_void_p, C.c_size_t, C.c_uint32, C.c_void_p, C.c_size_t,
                                    C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    buf = np.zeros(1 << 22, np.uint8)
    ts, nc = C.c_size_t(), C.c_size_t()
    f = np.ascontiguousarray(freq, np.uint64)
    assert L.cpsz_huffman_tree(f.ctypes.data, f.size, state, buf.ctypes.data, buf.size, C.byref(ts),
                               C.byref(nc)) == 0
    return buf[:ts.value].copy(), nc.value


def _dev_tree(freq, state):
    buf = np.zeros(1 << 22, np.uint8)
    ts, nc = C.c_size_t(), C.c_size_t()
    f = np.ascontiguousarray(freq, np.uint64)
    codes = np.zeros(f.size, np.uint64)
    lens = np.zeros(f.size, np.uint8)
    rc = cz.lib().cpsz_huffman_tree_dev(0, f.ctypes.data, f.size, state, buf.ctypes.data, buf.size, C.byref(ts),
                                        C.byref(nc), codes.ctypes.data, lens.ctypes.data)
    assert rc == 0, cz.lib().cpsz_last_error()
    return buf[:ts.value].copy(), nc.value, codes, lens


def _codes_from_tree(tree, N):
    """symbol -> (code, length) by walking the serialized tree."""
    w = 1 if N <= 256 else (2 if N <= 65536 else 4)
    dt = {1: np.uint8, 2: np.uint16, 4: np.uint32}[w]
    L = np.frombuffer(tree[1:1 + N * w].tobytes(), dt)
    R = np.frombuffer(tree[1 + N * w:1 + 2 * N * w].tobytes(), dt)
    Cs = np.frombuffer(tree[1 + 2 * N * w:1 + 2 * N * w + 4 * N].tobytes(), np.uint32)
    T = tree[1 + 2 * N * w + 4 * N:1 + 2 * N * w + 5 * N]
    out, stack = {}, [(0, 0, 0)]
    while stack:
        n, code, ln = stack.pop()
        if T[n]:
            out[int(Cs[n])] = (code, ln)
        else:
            stack.append((int(L[n]), code << 1, ln + 1))
            stack.append((int(R[n]), code << 1 | 1, ln + 1))
    return out


def _freq(rng, state, k, n, p):
    syms = (rng.geometric(p, n) % k + state // 4).astype(np.int64)
    return np.bincount(syms, minlength=2 * state)


CASES = [(2048, 1, 10, 0.5), (2048, 2, 10, 0.5), (2048, 13, 4000, 0.3), (512, 300, 50000, 0.02),
         (131072, 3000, 200000, 0.002), (131072, 9000, 400000, 0.0005), (131072, 30000, 300000, 0.00005),
         (131072, 60000, 200000, 0.00001)]


@pytest.mark.parametrize("state,k,n,p", CASES)
def test_device_tree_equals_host(state, k, n, p):
    print(f"[tree] state={state} k={k} n={n}", flush=True)
    rng = np.random.default_rng(k)
    freq = _freq(rng, state, k, n, p)
    ht, hn = _host_tree(freq, state)
    dt, dn, codes, lens = _dev_tree(freq, state)
    assert dn == hn and np.array_equal(dt, ht)
    want = _codes_from_tree(ht, hn)
    for s, (code, ln) in want.items():
        assert lens[s] == ln
        assert int(codes[s]) == ((code << (64 - ln)) & ((1 << 64) - 1) if ln else 0)


def test_device_tree_equal_counts_everywhere():
    """All counts equal (every removal a tie) and a Fibonacci-like spread (the
    deepest codes): the heap's tie-breaks and 64-bit code paths."""
    for freq_vals in (np.full(1000, 7), np.array([1, 1, 2, 3, 5, 8, 13, 21, 34, 55, 89, 144, 233, 377, 610])):
        state = 2048
        freq = np.zeros(2 * state, np.uint64)
        freq[100:100 + freq_vals.size] = freq_vals
        ht, hn = _host_tree(freq, state)
        dt, dn, _, _ = _dev_tree(freq, state)
        assert dn == hn and np.array_equal(dt, ht)
