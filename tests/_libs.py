"""ctypes bindings of the CPU oracle and of the reference build (test-only)."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORC_PATH = os.path.join(ROOT, "oracle", "liborc.so")
REF_PATH = os.path.join(ROOT, "oracle", "_ref", "libcpsz_ref.so")

_P = C.c_void_p
_SZ = C.c_size_t


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def load_oracle(path=ORC_PATH):
    lib = C.CDLL(path)
    pp = C.POINTER(C.c_void_p)
    for name, args in {
        "orc_compress_2d_f32": [_P, _P, _SZ, _SZ, C.c_double, pp, C.POINTER(_SZ)],
        "orc_compress_2d_f64": [_P, _P, _SZ, _SZ, C.c_double, pp, C.POINTER(_SZ)],
        "orc_decompress_2d_f32": [_P, _SZ, _SZ, _SZ, _P, _P],
        "orc_decompress_2d_f64": [_P, _SZ, _SZ, _SZ, _P, _P],
        "orc_compress_2d_log_f32": [_P, _P, _SZ, _SZ, C.c_double, pp, C.POINTER(_SZ)],
        "orc_decompress_2d_log_f32": [_P, _SZ, _SZ, _SZ, _P, _P],
        "orc_eb_2d_f32": [_P, _P, _SZ, _SZ, C.c_double, _P],
        "orc_eb_quantize_2d": [_P, _SZ, _P, _P],
        "orc_tri_bound_2d": [_P, _SZ, C.c_int, _P],
        "orc_lorenzo_2d_f32": [_P, _P, _SZ, _SZ, _P, _P, pp, C.POINTER(_SZ)],
        "orc_huff_encode": [C.c_uint32, _P, _SZ, pp, C.POINTER(_SZ)],
        "orc_huff_decode": [_P, _SZ, _SZ, _P, C.POINTER(_SZ)],
        "orc_compress_3d_f32": [_P, _P, _P, _SZ, _SZ, _SZ, C.c_double, pp, C.POINTER(_SZ)],
        "orc_compress_3d_f64": [_P, _P, _P, _SZ, _SZ, _SZ, C.c_double, pp, C.POINTER(_SZ)],
        "orc_decompress_3d_f32": [_P, _SZ, _SZ, _SZ, _SZ, _P, _P, _P],
        "orc_decompress_3d_f64": [_P, _SZ, _SZ, _SZ, _SZ, _P, _P, _P],
        "orc_eb_3d_f32": [_P, _P, _P, _SZ, _SZ, _SZ, C.c_double, _P],
        "orc_eb_id_3d": [_P, _SZ, _P],
        "orc_acbrt1": [_P, _SZ, _P],
        "orc_tet_bound_3d": [_P, _SZ, _P],
        "orc_sign_bound_3d": [_P, _P, _SZ, _P],
        "orc_same_direction_3d": [_P, _SZ, _P],
        "orc_eb_exp_quantize": [_P, _SZ, C.c_int, C.c_double, C.c_double, _P, _P],
        "orc_log_transform_3d_f32": [_P, _SZ, _P, _P],
    }.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_int
    lib.orc_cp_2d.argtypes = [_P, _P, _SZ, _SZ, _P, _P, _SZ]
    lib.orc_learn_features_2d.argtypes = [_P, _P, _P, _SZ, _SZ, C.c_double, _P, _SZ, _P]
    lib.orc_learn_labels_2d.argtypes = [_P, _P, _P, _SZ, _SZ, C.c_double, _P, _SZ, _P]
    lib.orc_learn_features_3d.argtypes = [_P, _P, _P, _P, _SZ, _SZ, _SZ, C.c_double, _P, _SZ, _P]
    lib.orc_learn_labels_3d.argtypes = [_P, _P, _P, _P, _SZ, _SZ, _SZ, C.c_double, _P, _SZ, _P]
    lib.orc_tet_cp_3d.argtypes = [_P, C.c_int]
    lib.orc_tet_cp_3d.restype = C.c_int
    lib.orc_cp_2d.restype = C.c_long
    lib.orc_cp_3d.argtypes = [_P, _P, _P, _SZ, _SZ, _SZ, _P, _P, _SZ]
    lib.orc_cp_3d.restype = C.c_long
    lib.orc_cp_2d_sos.argtypes = [_P, _P, _SZ, _SZ, _P, _P, _SZ]
    lib.orc_cp_2d_sos.restype = C.c_long
    lib.orc_cp_3d_sos.argtypes = [_P, _P, _P, _SZ, _SZ, _SZ, _P, _P, _SZ]
    lib.orc_cp_3d_sos.restype = C.c_long
    lib.orc_sos_det.argtypes = [C.c_int, _P, _P]
    lib.orc_sos_det.restype = C.c_int
    lib.orc_free.argtypes = [_P]
    lib.orc_acbrt1_one.argtypes = [C.c_float]
    lib.orc_acbrt1_one.restype = C.c_float
    return Oracle(lib)


def _take(lib, ptr, size):
    out = np.frombuffer(C.string_at(ptr, size), dtype=np.uint8).copy()
    lib.orc_free(ptr)
    return out


class Oracle:
    def __init__(self, lib):
        self.lib = lib

    def compress_2d(self, U, V, eb=0.1):
        r1, r2 = U.shape
        f = self.lib.orc_compress_2d_f32 if U.dtype == np.float32 else self.lib.orc_compress_2d_f64
        p, s = C.c_void_p(), _SZ()
        assert f(_ptr(U), _ptr(V), r1, r2, eb, C.byref(p), C.byref(s)) == 0
        return _take(self.lib, p, s.value)

    def decompress_2d(self, stream, r1, r2, dtype=np.float32):
        U = np.empty((r1, r2), dtype)
        V = np.empty((r1, r2), dtype)
        f = self.lib.orc_decompress_2d_f32 if dtype == np.float32 else self.lib.orc_decompress_2d_f64
        assert f(_ptr(stream), stream.size, r1, r2, _ptr(U), _ptr(V)) == 0
        return U, V

    def compress_2d_log(self, U, V, eb=0.1):
        r1, r2 = U.shape
        p, n = C.c_void_p(), _SZ()
        assert self.lib.orc_compress_2d_log_f32(_ptr(U), _ptr(V), r1, r2, eb, C.byref(p), C.byref(n)) == 0
        return _take(self.lib, p, n.value)

    def decompress_2d_log(self, stream, r1, r2):
        U = np.empty((r1, r2), np.float32)
        V = np.empty((r1, r2), np.float32)
        assert self.lib.orc_decompress_2d_log_f32(_ptr(stream), stream.size, r1, r2, _ptr(U), _ptr(V)) == 0
        return U, V

    def eb_2d(self, U, V, eb=0.1):
        out = np.empty(U.shape, np.float64)
        self.lib.orc_eb_2d_f32(_ptr(U), _ptr(V), U.shape[0], U.shape[1], eb, _ptr(out))
        return out

    def eb_quantize_2d(self, eb_u):
        eb_u = np.ascontiguousarray(eb_u, np.float64)
        ids = np.empty(eb_u.shape, np.int32)
        prec = np.empty(eb_u.shape, np.float64)
        self.lib.orc_eb_quantize_2d(_ptr(eb_u), eb_u.size, _ptr(ids), _ptr(prec))
        return ids, prec

    def tri_bound_2d(self, uv6, lower):
        uv6 = np.ascontiguousarray(uv6, np.float32)
        out = np.empty(uv6.shape[0], np.float64)
        self.lib.orc_tri_bound_2d(_ptr(uv6), uv6.shape[0], int(lower), _ptr(out))
        return out

    def lorenzo_2d(self, data, prec):
        r1, r2 = data.shape
        codes = np.empty(r1 * r2, np.int32)
        recon = np.empty((r1, r2), np.float32)
        p, s = C.c_void_p(), _SZ()
        assert self.lib.orc_lorenzo_2d_f32(_ptr(data), _ptr(prec), r1, r2, _ptr(codes), _ptr(recon),
                                           C.byref(p), C.byref(s)) == 0
        return _take(self.lib, p, s.value), codes, recon

    def huff_encode(self, state_num, syms):
        syms = np.ascontiguousarray(syms, np.int32)
        p, s = C.c_void_p(), _SZ()
        assert self.lib.orc_huff_encode(state_num, _ptr(syms), syms.size, C.byref(p), C.byref(s)) == 0
        return _take(self.lib, p, s.value)

    def huff_decode(self, sec, n):
        out = np.empty(n, np.int32)
        used = _SZ()
        assert self.lib.orc_huff_decode(_ptr(sec), sec.size, n, _ptr(out), C.byref(used)) == 0
        return out, used.value

    def compress_3d(self, U, V, W, eb=0.05):
        r1, r2, r3 = U.shape
        f = self.lib.orc_compress_3d_f32 if U.dtype == np.float32 else self.lib.orc_compress_3d_f64
        p, s = C.c_void_p(), _SZ()
        assert f(_ptr(U), _ptr(V), _ptr(W), r1, r2, r3, eb, C.byref(p), C.byref(s)) == 0
        return _take(self.lib, p, s.value)

    def decompress_3d(self, stream, r1, r2, r3, dtype=np.float32):
        out = [np.empty((r1, r2, r3), dtype) for _ in range(3)]
        f = self.lib.orc_decompress_3d_f32 if dtype == np.float32 else self.lib.orc_decompress_3d_f64
        assert f(_ptr(stream), stream.size, r1, r2, r3, *[_ptr(o) for o in out]) == 0
        return out

    def eb_3d(self, U, V, W, eb=0.05):
        out = np.empty(U.shape, np.float64)
        self.lib.orc_eb_3d_f32(_ptr(U), _ptr(V), _ptr(W), *U.shape, eb, _ptr(out))
        return out

    def eb_id_3d(self, eb):
        eb = np.ascontiguousarray(eb, np.float64)
        ids = np.empty(eb.shape, np.int32)
        self.lib.orc_eb_id_3d(_ptr(eb), eb.size, _ptr(ids))
        return ids

    def acbrt1(self, x):
        x = np.ascontiguousarray(x, np.float32)
        out = np.empty_like(x)
        self.lib.orc_acbrt1(_ptr(x), x.size, _ptr(out))
        return out

    def sign_bound_3d(self, pos, neg):
        pos, neg = (np.ascontiguousarray(x, np.float32) for x in (pos, neg))
        out = np.empty(pos.shape, np.float32)
        self.lib.orc_sign_bound_3d(_ptr(pos), _ptr(neg), pos.size, _ptr(out))
        return out

    def same_direction_3d(self, q4):
        q4 = np.ascontiguousarray(q4, np.float32)
        out = np.empty(q4.shape[0], np.int32)
        self.lib.orc_same_direction_3d(_ptr(q4), q4.shape[0], _ptr(out))
        return out

    def eb_exp_quantize(self, eb, base, log_of_base, threshold):
        eb = np.ascontiguousarray(eb, np.float64)
        ids = np.empty(eb.shape, np.int32)
        out = np.empty(eb.shape, np.float64)
        self.lib.orc_eb_exp_quantize(_ptr(eb), eb.size, base, log_of_base, threshold, _ptr(ids), _ptr(out))
        return ids, out

    def log_transform_3d(self, x):
        x = np.ascontiguousarray(x, np.float32)
        logv = np.empty(x.shape, np.float32)
        sign = np.empty(x.shape, np.uint8)
        self.lib.orc_log_transform_3d_f32(_ptr(x), x.size, _ptr(logv), _ptr(sign))
        return logv, sign

    def tet_bound_3d(self, uvw12):
        uvw12 = np.ascontiguousarray(uvw12, np.float32)
        out = np.empty(uvw12.shape[0], np.float64)
        self.lib.orc_tet_bound_3d(_ptr(uvw12), uvw12.shape[0], _ptr(out))
        return out

    def learn_features_2d(self, U, V, E, eb, idx=None):
        U, V = (np.ascontiguousarray(x, np.float32) for x in (U, V))
        E = np.ascontiguousarray(E, np.float64)
        m = U.size if idx is None else len(idx)
        idx_a = None if idx is None else np.ascontiguousarray(idx, np.uint32)
        X = np.zeros((m, 64), np.uint16)
        self.lib.orc_learn_features_2d(_ptr(U), _ptr(V), _ptr(E), U.shape[0], U.shape[1], eb,
                                       None if idx_a is None else _ptr(idx_a), m, _ptr(X))
        return X

    def learn_labels_2d(self, U, V, E, eb, idx=None):
        U, V = (np.ascontiguousarray(x, np.float32) for x in (U, V))
        E = np.ascontiguousarray(E, np.float64)
        m = U.size if idx is None else len(idx)
        idx_a = None if idx is None else np.ascontiguousarray(idx, np.uint32)
        L = np.zeros(m, np.uint8)
        self.lib.orc_learn_labels_2d(_ptr(U), _ptr(V), _ptr(E), U.shape[0], U.shape[1], eb,
                                     None if idx_a is None else _ptr(idx_a), m, _ptr(L))
        return L

    def learn_features_3d(self, U, V, W, E, eb, idx=None):
        U, V, W = (np.ascontiguousarray(x, np.float32) for x in (U, V, W))
        E = np.ascontiguousarray(E, np.float64)
        m = U.size if idx is None else len(idx)
        idx_a = None if idx is None else np.ascontiguousarray(idx, np.uint32)
        X = np.zeros((m, 96), np.uint16)
        self.lib.orc_learn_features_3d(_ptr(U), _ptr(V), _ptr(W), _ptr(E), *U.shape, eb,
                                       None if idx_a is None else _ptr(idx_a), m, _ptr(X))
        return X

    def learn_labels_3d(self, U, V, W, E, eb, idx=None):
        U, V, W = (np.ascontiguousarray(x, np.float32) for x in (U, V, W))
        E = np.ascontiguousarray(E, np.float64)
        m = U.size if idx is None else len(idx)
        idx_a = None if idx is None else np.ascontiguousarray(idx, np.uint32)
        L = np.zeros(m, np.uint8)
        self.lib.orc_learn_labels_3d(_ptr(U), _ptr(V), _ptr(W), _ptr(E), *U.shape, eb,
                                     None if idx_a is None else _ptr(idx_a), m, _ptr(L))
        return L

    def tet_cp_3d(self, f, t):
        f = np.ascontiguousarray(f, np.float32).reshape(12)
        return self.lib.orc_tet_cp_3d(_ptr(f), t)

    def cp_2d(self, U, V):
        r1, r2 = U.shape
        cap = 2 * r1 * r2
        sid = np.empty(cap, np.int64)
        typ = np.empty(cap, np.int32)
        n = self.lib.orc_cp_2d(_ptr(U), _ptr(V), r1, r2, _ptr(sid), _ptr(typ), cap)
        return dict(zip(sid[:n].tolist(), typ[:n].tolist()))

    def cp_3d(self, U, V, W):
        r1, r2, r3 = U.shape
        cap = 6 * r1 * r2 * r3
        sid = np.empty(cap, np.int64)
        typ = np.empty(cap, np.int32)
        n = self.lib.orc_cp_3d(_ptr(U), _ptr(V), _ptr(W), r1, r2, r3, _ptr(sid), _ptr(typ), cap)
        return dict(zip(sid[:n].tolist(), typ[:n].tolist()))

    def cp_2d_sos(self, U, V):
        r1, r2 = U.shape
        cap = 2 * r1 * r2
        sid = np.empty(cap, np.int64)
        typ = np.empty(cap, np.int32)
        n = self.lib.orc_cp_2d_sos(_ptr(U), _ptr(V), r1, r2, _ptr(sid), _ptr(typ), cap)
        return dict(zip(sid[:n].tolist(), typ[:n].tolist()))

    def cp_3d_sos(self, U, V, W):
        r1, r2, r3 = U.shape
        cap = 6 * r1 * r2 * r3
        sid = np.empty(cap, np.int64)
        typ = np.empty(cap, np.int32)
        n = self.lib.orc_cp_3d_sos(_ptr(U), _ptr(V), _ptr(W), r1, r2, r3, _ptr(sid), _ptr(typ), cap)
        return dict(zip(sid[:n].tolist(), typ[:n].tolist()))

    def sos_det(self, rows, idx):
        rows = np.ascontiguousarray(rows, np.float32).ravel()
        idx = np.ascontiguousarray(idx, np.int64)
        return self.lib.orc_sos_det(len(idx), _ptr(rows), _ptr(idx))


def have_reference():
    return os.path.exists(REF_PATH)


class Reference:
    """The reference CPU library compiled from its own sources (oracle/_ref)."""

    def __init__(self):
        lib = C.CDLL(REF_PATH)
        lib.ref_compress_2d_f32.argtypes = [_P, _P, _SZ, _SZ, C.c_double, C.POINTER(_SZ)]
        lib.ref_compress_2d_f32.restype = C.c_void_p
        lib.ref_compress_2d_f64.argtypes = [_P, _P, _SZ, _SZ, C.c_double, C.POINTER(_SZ)]
        lib.ref_compress_2d_f64.restype = C.c_void_p
        lib.ref_decompress_2d_f32.argtypes = [_P, _SZ, _SZ, _P, _P]
        lib.ref_decompress_2d_f64.argtypes = [_P, _SZ, _SZ, _P, _P]
        lib.ref_decompress_3d_f32.argtypes = [_P, _SZ, _SZ, _SZ, _P, _P, _P]
        lib.ref_decompress_3d_f64.argtypes = [_P, _SZ, _SZ, _SZ, _P, _P, _P]
        lib.ref_huff_encode.argtypes = [_SZ, _P, _SZ, _P]
        lib.ref_huff_encode.restype = _SZ
        lib.ref_huff_decode.argtypes = [_SZ, _P, _SZ, _P]
        lib.ref_huff_decode.restype = _SZ
        lib.ref_lorenzo_2d_f32.argtypes = [_P, _P, _SZ, _SZ, C.POINTER(_SZ)]
        lib.ref_lorenzo_2d_f32.restype = C.c_void_p
        lib.ref_free.argtypes = [_P]
        lib.ref_zstd_compress.argtypes = [C.c_int, _P, _SZ, C.POINTER(C.c_void_p)]
        lib.ref_zstd_compress.restype = _SZ
        lib.ref_zstd_decompress.argtypes = [_P, _SZ, C.POINTER(C.c_void_p), _SZ]
        lib.ref_zstd_decompress.restype = _SZ
        lib.ref_compress_2d_log_f32.argtypes = [_P, _P, _SZ, _SZ, C.c_double, C.POINTER(_SZ)]
        lib.ref_compress_2d_log_f32.restype = C.c_void_p
        lib.ref_decompress_2d_log_f32.argtypes = [_P, _SZ, _SZ, _P, _P]
        lib.ref_is_lossless.argtypes = [_P, _SZ]
        lib.ref_is_lossless.restype = C.c_int
        lib.ref_zstd_decompress65536.argtypes = [_P, _SZ, C.POINTER(C.c_void_p)]
        lib.ref_zstd_decompress65536.restype = _SZ
        lib.ref_acbrt1_one.argtypes = [C.c_float]
        lib.ref_acbrt1_one.restype = C.c_float
        lib.ref_acbrt1_compare.argtypes = [_P, C.c_uint32, C.c_uint32, C.c_int, C.POINTER(C.c_uint32)]
        lib.ref_acbrt1_compare.restype = C.c_uint64
        lib.ref_sign_bound_3d.argtypes = [_P, _P, _SZ, _P]
        lib.ref_same_direction.argtypes = [_P, _SZ, _P]
        lib.ref_eb_exp_quantize.argtypes = [_P, _SZ, C.c_int, C.c_double, C.c_double, _P, _P]
        lib.ref_log_transform_f32.argtypes = [_P, _SZ, _P, _P]
        self.lib = lib

    # header-only helpers of the 3D compressor (include/sz_cp_preserve_utils.hpp)
    def acbrt1_compare(self, fn_ptr, lo, hi, threads=8):
        """Number of float bit patterns in [lo, hi) where the C function at
        fn_ptr differs from the reference acbrt1, and the first of them."""
        first = C.c_uint32()
        bad = self.lib.ref_acbrt1_compare(fn_ptr, lo, hi, threads, C.byref(first))
        return int(bad), first.value

    def sign_bound_3d(self, pos, neg):
        pos, neg = (np.ascontiguousarray(x, np.float32) for x in (pos, neg))
        out = np.empty(pos.shape, np.float32)
        self.lib.ref_sign_bound_3d(_ptr(pos), _ptr(neg), pos.size, _ptr(out))
        return out

    def same_direction(self, q4):
        q4 = np.ascontiguousarray(q4, np.float32)
        out = np.empty(q4.shape[0], np.int32)
        self.lib.ref_same_direction(_ptr(q4), q4.shape[0], _ptr(out))
        return out

    def eb_exp_quantize(self, eb, base, log_of_base, threshold):
        eb = np.ascontiguousarray(eb, np.float64)
        ids = np.empty(eb.shape, np.int32)
        out = np.empty(eb.shape, np.float64)
        self.lib.ref_eb_exp_quantize(_ptr(eb), eb.size, base, log_of_base, threshold, _ptr(ids), _ptr(out))
        return ids, out

    def log_transform(self, x):
        x = np.ascontiguousarray(x, np.float32)
        logv = np.empty(x.shape, np.float32)
        sign = np.empty(x.shape, np.uint8)
        self.lib.ref_log_transform_f32(_ptr(x), x.size, _ptr(logv), _ptr(sign))
        return logv, sign

    def _take(self, p, n):
        out = np.frombuffer(C.string_at(p, n), dtype=np.uint8).copy()
        self.lib.ref_free(p)
        return out

    def zstd_compress(self, data, level=3):
        data = np.ascontiguousarray(data, np.uint8)
        p = C.c_void_p()
        n = self.lib.ref_zstd_compress(level, _ptr(data), data.size, C.byref(p))
        return self._take(p, n)

    def compress_2d_log(self, U, V, eb=0.1):
        r1, r2 = U.shape
        s = _SZ()
        p = self.lib.ref_compress_2d_log_f32(_ptr(U), _ptr(V), r1, r2, eb, C.byref(s))
        return self._take(p, s.value)

    def decompress_2d_log(self, stream, r1, r2):
        U = np.empty((r1, r2), np.float32)
        V = np.empty((r1, r2), np.float32)
        self.lib.ref_decompress_2d_log_f32(_ptr(stream), r1, r2, _ptr(U), _ptr(V))
        return U, V

    def is_lossless(self, data):
        data = np.ascontiguousarray(data, np.uint8)
        return self.lib.ref_is_lossless(_ptr(data), data.size)

    def zstd_decompress65536(self, data):
        data = np.ascontiguousarray(data, np.uint8)
        p = C.c_void_p()
        n = self.lib.ref_zstd_decompress65536(_ptr(data), data.size, C.byref(p))
        return self._take(p, n)

    def compress_2d(self, U, V, eb=0.1):
        r1, r2 = U.shape
        s = _SZ()
        f = self.lib.ref_compress_2d_f32 if U.dtype == np.float32 else self.lib.ref_compress_2d_f64
        p = f(_ptr(U), _ptr(V), r1, r2, eb, C.byref(s))
        return self._take(p, s.value)

    def decompress_2d(self, stream, r1, r2, dtype=np.float32):
        U = np.empty((r1, r2), dtype)
        V = np.empty((r1, r2), dtype)
        f = self.lib.ref_decompress_2d_f32 if dtype == np.float32 else self.lib.ref_decompress_2d_f64
        f(_ptr(stream), r1, r2, _ptr(U), _ptr(V))
        return U, V

    def decompress_3d(self, stream, r1, r2, r3, dtype=np.float32):
        out = [np.empty((r1, r2, r3), dtype) for _ in range(3)]
        f = self.lib.ref_decompress_3d_f32 if dtype == np.float32 else self.lib.ref_decompress_3d_f64
        f(_ptr(stream), r1, r2, r3, *[_ptr(o) for o in out])
        return out

    def huff_encode(self, state_num, syms):
        syms = np.ascontiguousarray(syms, np.int32)
        buf = np.zeros(syms.size * 8 + (1 << 22), np.uint8)
        n = self.lib.ref_huff_encode(state_num, _ptr(syms), syms.size, _ptr(buf))
        return buf[:n].copy()

    def huff_decode(self, state_num, sec, n):
        out = np.empty(n, np.int32)
        buf = np.concatenate([sec, np.zeros(64, np.uint8)])
        used = self.lib.ref_huff_decode(state_num, _ptr(buf), n, _ptr(out))
        return out, used

    def lorenzo_2d(self, data, prec):
        r1, r2 = data.shape
        s = _SZ()
        p = self.lib.ref_lorenzo_2d_f32(_ptr(data), _ptr(prec), r1, r2, C.byref(s))
        return self._take(p, s.value)


def mask_2d_log_padding(stream, r1, r2, itemsize=4):
    """mask_2d_padding for the 2D offline_log layout: the two sign maps of
    ceil(n/8) bytes sit between the id section and the u section."""
    s = stream.copy()
    hdr = 4 + 8 + 8 * 3
    eb_size, u_size, v_size = np.frombuffer(s[12:36].tobytes(), dtype=np.uint64)
    smap = (r1 * r2 - 1) // 8 + 1
    for off in (hdr + int(eb_size) + 2 * smap, hdr + int(eb_size) + 2 * smap + int(u_size)):
        s[off + 10: off + 9 + itemsize] = 0
    return s


def mask_2d_padding(stream, r1, r2, itemsize=4):
    """Zero the uninitialised meanInfo padding bytes of a 2D offline stream.

    Layout: int base | double thr | size_t eb_size,u_size,v_size | eb | u | v;
    each component section starts int block | char | int radius | meanInfo
    (bool + padding + T), the padding lying at section offsets 10..(9+itemsize-1).
    """
    s = stream.copy()
    hdr = 4 + 8 + 8 * 3
    eb_size, u_size, v_size = np.frombuffer(s[12:36].tobytes(), dtype=np.uint64)
    for off in (hdr + int(eb_size), hdr + int(eb_size) + int(u_size)):
        s[off + 10: off + 9 + itemsize] = 0
    return s
