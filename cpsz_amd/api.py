"""ctypes binding of libcpsz_hip.so, mirroring the reference's C++ API names.

Host-buffer functions take/return numpy arrays; the ``*_dev`` functions take
raw device pointers (e.g. ``torch.Tensor.data_ptr()``) and an optional HIP
stream handle. Any failure raises :class:`CpszError` carrying the library's
status code and message. There is no CPU fallback: if the HIP library is not
built, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcpsz_hip.so")

_P = C.c_void_p
_SZ = C.c_size_t

# name -> argtypes (all return int)
_SIGS = {
    "cpsz_compress_2d_f32": [_P, _P, _SZ, _SZ, C.c_double, C.POINTER(_P), C.POINTER(_SZ)],
    "cpsz_decompress_2d_f32": [_P, _SZ, _SZ, _SZ, _P, _P],
    "cpsz_compress_3d_f32": [_P, _P, _P, _SZ, _SZ, _SZ, C.c_double, C.POINTER(_P), C.POINTER(_SZ)],
    "cpsz_decompress_3d_f32": [_P, _SZ, _SZ, _SZ, _SZ, _P, _P, _P],
    "cpsz_stream_size_2d": [_P, C.POINTER(_SZ)],
    "cpsz_stream_size_3d": [_P, _SZ, _SZ, _SZ, C.POINTER(_SZ)],
    "cpsz_compress_2d_f32_dev": [C.c_int, _P, _P, _P, _SZ, _SZ, C.c_double, C.POINTER(_P), C.POINTER(_SZ)],
    "cpsz_decompress_2d_f32_dev": [C.c_int, _P, _P, _SZ, _SZ, _SZ, _P, _P],
    "cpsz_compress_3d_f32_dev": [C.c_int, _P, _P, _P, _P, _SZ, _SZ, _SZ, C.c_double, C.POINTER(_P),
                                 C.POINTER(_SZ)],
    "cpsz_decompress_3d_f32_dev": [C.c_int, _P, _P, _SZ, _SZ, _SZ, _SZ, _P, _P, _P],
    "cpsz_lossless_compress": [C.c_int, _P, _SZ, C.POINTER(_P), C.POINTER(_SZ)],
    "cpsz_lossless_decompress": [_P, _SZ, _P, _SZ],
    "cpsz_lossless_compress_dev": [C.c_int, _P, C.c_int, _P, _SZ, C.POINTER(_P), C.POINTER(_SZ)],
    "cpsz_lossless_compress_batch_dev": [C.c_int, _P, C.c_int, _P, _P, _P, _SZ, _P, _P],
    "cpsz_critical_points_2d_dev": [C.c_int, _P, _P, _P, _SZ, _SZ, _P, _P, _SZ, C.POINTER(_SZ)],
    "cpsz_compress_2d_f64": [_P, _P, _SZ, _SZ, C.c_double, C.POINTER(_P), C.POINTER(_SZ)],
    "cpsz_decompress_2d_f64": [_P, _SZ, _SZ, _SZ, _P, _P],
    "cpsz_compress_2d_f64_dev": [C.c_int, _P, _P, _P, _SZ, _SZ, C.c_double, C.POINTER(_P), C.POINTER(_SZ)],
    "cpsz_decompress_2d_f64_dev": [C.c_int, _P, _P, _SZ, _SZ, _SZ, _P, _P],
    "cpsz_compress_2d_log_f32": [_P, _P, _SZ, _SZ, C.c_double, C.POINTER(_P), C.POINTER(_SZ)],
    "cpsz_decompress_2d_log_f32": [_P, _SZ, _SZ, _SZ, _P, _P],
    "cpsz_compress_2d_log_f32_dev": [C.c_int, _P, _P, _P, _SZ, _SZ, C.c_double, C.POINTER(_P), C.POINTER(_SZ)],
    "cpsz_decompress_2d_log_f32_dev": [C.c_int, _P, _P, _SZ, _SZ, _SZ, _P, _P],
    "cpsz_stream_size_2d_log": [_P, _SZ, _SZ, C.POINTER(_SZ)],
    "cpsz_eb_table_2d_log": [_P, C.c_int, C.POINTER(C.c_int)],
    "cpsz_critical_points_2d_noalloc_dev": [C.c_int, _P, _P, _P, _SZ, _SZ, _P, _P, _SZ, C.POINTER(_SZ)],
    "cpsz_critical_points_2d_sos_dev": [C.c_int, _P, _P, _P, _SZ, _SZ, _P, _P, _SZ, C.POINTER(_SZ)],
    "cpsz_critical_points_3d_sos_dev": [C.c_int, _P, _P, _P, _P, _SZ, _SZ, _SZ, _P, _P, _SZ, C.POINTER(_SZ)],
    "cpsz_sos_det": [C.c_int, _P, _P, C.POINTER(C.c_int)],
    "cpsz_huffman_tree_dev": [C.c_int, _P, _SZ, C.c_uint32, _P, _SZ, C.POINTER(_SZ), C.POINTER(_SZ), _P, _P],
    "cpsz_critical_points_3d_dev": [C.c_int, _P, _P, _P, _P, _SZ, _SZ, _SZ, _P, _P, _SZ,
                                    C.POINTER(_SZ)],
    "cpsz_set_timing": [C.c_int, C.c_int],
    "cpsz_last_stages": [C.c_int, C.c_char_p, _SZ, _P, C.c_int],
    "cpsz_device_count": [],
    "cpsz_compress_2d_batch_dev": [C.c_int, _P, _SZ, _P, _P, _P, _P, C.c_double, C.c_int, C.POINTER(_P), _P,
                                   _P],
    "cpsz_decompress_2d_batch_dev": [C.c_int, _P, _SZ, _P, _P, _P, _P, _P, _P],
    "cpsz_batch_2d_config": [C.c_int, _SZ, _P, _P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "cpsz_batch_2d_set_config": [C.c_int, _SZ, _P, _P, C.c_int, C.c_int, C.c_int],
    "cpsz_compress_3d_batch_dev": [C.c_int, _P, _SZ, _P, _P, _P, _P, _P, _P, C.c_double, C.c_int,
                                   C.POINTER(_P), _P, _P],
    "cpsz_decompress_3d_batch_dev": [C.c_int, _P, _SZ, _P, _P, _P, _P, _P, _P, _P, _P],
    "cpsz_batch_3d_config": [C.c_int, _SZ, _P, _P, _P, C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "cpsz_learn_init_params": [C.c_uint64, _P],
    "cpsz_learn_analytic_eb_2d_dev": [C.c_int, _P, _P, _P, _SZ, _SZ, C.c_double, _P],
    "cpsz_learn_features_2d_dev": [C.c_int, _P, _P, _P, _P, _SZ, _SZ, C.c_double, _P, _SZ, _P],
    "cpsz_learn_labels_2d_dev": [C.c_int, _P, _P, _P, _P, _SZ, _SZ, C.c_double, _P, _SZ, _P],
    "cpsz_gemm_bf16_dev": [C.c_int, _P, C.c_int, C.c_int, C.c_int, _P, C.c_int64, C.c_int64, _P,
                           C.c_int64, C.c_int64, _P, C.c_int, C.c_int],
    "cpsz_mlp_forward_dev": [C.c_int, _P, _P, _P, _SZ, _P],
    "cpsz_mlp_grad_dev": [C.c_int, _P, _P, _P, _P, _SZ, _P, _P],
    "cpsz_learn_train_2d_dev": [C.c_int, _P, _SZ, _P, _P, _P, _P, C.c_double, _SZ, C.c_int, C.c_int,
                                C.c_float, C.c_uint64, _P, _P],
    "cpsz_compress_2d_learned_f32_dev": [C.c_int, _P, _P, _P, _SZ, _SZ, C.c_double, _P, C.POINTER(_P),
                                         C.POINTER(_SZ), C.POINTER(C.c_int), C.POINTER(_SZ),
                                         C.POINTER(_SZ)],
    "cpsz_compress_2d_learned_f32": [_P, _P, _SZ, _SZ, C.c_double, _P, C.POINTER(_P), C.POINTER(_SZ)],
    "cpsz_learn_init_params_3d": [C.c_uint64, _P],
    "cpsz_learn_analytic_eb_3d_dev": [C.c_int, _P, _P, _P, _P, _SZ, _SZ, _SZ, C.c_double, _P],
    "cpsz_learn_features_3d_dev": [C.c_int, _P, _P, _P, _P, _P, _SZ, _SZ, _SZ, C.c_double, _P, _SZ, _P],
    "cpsz_learn_labels_3d_dev": [C.c_int, _P, _P, _P, _P, _P, _SZ, _SZ, _SZ, C.c_double, _P, _SZ, _P],
    "cpsz_mlp_forward_3d_dev": [C.c_int, _P, _P, _P, _SZ, _P],
    "cpsz_mlp_grad_3d_dev": [C.c_int, _P, _P, _P, _P, _SZ, _P, _P],
    "cpsz_learn_train_3d_dev": [C.c_int, _P, _SZ, _P, _P, _P, _P, _P, _P, C.c_double, _SZ, C.c_int, C.c_int,
                                C.c_float, C.c_uint64, _P, _P],
    "cpsz_learn_infer_gains_dev": [C.c_int, _P, C.c_int, _P, _P, _P, _P, _SZ, _SZ, _SZ, C.c_double, _P, _SZ, _SZ,
                                   _P],
    "cpsz_compress_2d_learned_batch_dev": [C.c_int, _P, _SZ, _P, _P, _P, _P, C.c_double, _P, C.c_int,
                                           C.POINTER(_P), _P, _P, C.POINTER(C.c_int), C.POINTER(_SZ),
                                           C.POINTER(_SZ)],
    "cpsz_compress_3d_learned_batch_dev": [C.c_int, _P, _SZ, _P, _P, _P, _P, _P, _P, C.c_double, _P, C.c_int,
                                           C.POINTER(_P), _P, _P, C.POINTER(C.c_int), C.POINTER(_SZ),
                                           C.POINTER(_SZ)],
    "cpsz_tet_cp_emulate": [_P, C.c_int, C.POINTER(C.c_int)],
    "cpsz_check_fast_bounds": [C.c_int, C.POINTER(C.c_uint32), C.POINTER(C.c_double)],
    "cpsz_set_exact_2d_bounds": [C.c_int],
    "cpsz_tri_bound_2d_fast_host": [_P, _SZ, C.c_int, _P],
    "cpsz_ratio_ge1_host": [_P, _P, _SZ, _P],
}


class CpszError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"cpsz error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    """Loads the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CpszError(-5, f"{LIB_PATH} is missing; build it with __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        L.cpsz_last_error.restype = C.c_char_p
        L.cpsz_last_error.argtypes = []
        L.cpsz_version.restype = C.c_char_p
        L.cpsz_free.argtypes = [_P]
        L.cpsz_free.restype = None
        L.cpsz_learn_param_count.restype = _SZ
        L.cpsz_learn_param_count.argtypes = []
        L.cpsz_learn_param_count_3d.restype = _SZ
        L.cpsz_learn_param_count_3d.argtypes = []
        _lib = L
    return _lib


def _check(rc: int):
    if rc < 0:
        raise CpszError(rc, lib().cpsz_last_error().decode())
    return rc


def _f32(a):
    a = np.ascontiguousarray(a)
    if a.dtype != np.float32:
        raise TypeError("cpsz works on float32 fields")
    return a


def _ptr(a):
    return a.ctypes.data_as(_P)


def _take(p, n):
    out = np.frombuffer(C.string_at(p, n), dtype=np.uint8).copy()
    lib().cpsz_free(p)
    return out


# ------------------------------------------------------------- host API ----
def sz_compress_cp_preserve_2d_offline(U, V, max_pwr_eb: float = 0.1) -> np.ndarray:
    """Compresses row-major r1 x r2 fields U, V (float32, or float64 for the
    reference's double instantiation); returns the stream."""
    if np.asarray(U).dtype == np.float64:
        U, V = np.ascontiguousarray(U, np.float64), np.ascontiguousarray(V, np.float64)
        if U.shape != V.shape or U.ndim != 2:
            raise ValueError("U and V must be 2D arrays of the same shape")
        p, s = _P(), _SZ()
        _check(lib().cpsz_compress_2d_f64(_ptr(U), _ptr(V), *U.shape, max_pwr_eb, C.byref(p), C.byref(s)))
        return _take(p, s.value)
    U, V = _f32(U), _f32(V)
    if U.shape != V.shape or U.ndim != 2:
        raise ValueError("U and V must be 2D arrays of the same shape")
    r1, r2 = U.shape
    p, s = _P(), _SZ()
    _check(lib().cpsz_compress_2d_f32(_ptr(U), _ptr(V), r1, r2, max_pwr_eb, C.byref(p), C.byref(s)))
    return _take(p, s.value)


def sz_decompress_cp_preserve_2d_offline(stream, r1: int, r2: int, dtype=np.float32):
    """dtype float64 decodes a stream of the double instantiation."""
    stream = np.ascontiguousarray(np.frombuffer(bytes(stream), np.uint8) if isinstance(stream, (bytes, bytearray)) else stream, np.uint8)
    U = np.empty((r1, r2), dtype)
    V = np.empty((r1, r2), dtype)
    f = lib().cpsz_decompress_2d_f64 if np.dtype(dtype) == np.float64 else lib().cpsz_decompress_2d_f32
    _check(f(_ptr(stream), stream.size, r1, r2, _ptr(U), _ptr(V)))
    return U, V


def sz_compress_cp_preserve_2d_offline_log(U, V, max_pwr_eb: float = 0.1) -> np.ndarray:
    """The 2D offline_log compressor (bounds in the log domain); returns the stream."""
    U, V = _f32(U), _f32(V)
    if U.shape != V.shape or U.ndim != 2:
        raise ValueError("U and V must be 2D arrays of the same shape")
    r1, r2 = U.shape
    p, s = _P(), _SZ()
    _check(lib().cpsz_compress_2d_log_f32(_ptr(U), _ptr(V), r1, r2, max_pwr_eb, C.byref(p), C.byref(s)))
    return _take(p, s.value)


def sz_decompress_cp_preserve_2d_offline_log(stream, r1: int, r2: int):
    stream = np.ascontiguousarray(np.frombuffer(bytes(stream), np.uint8) if isinstance(stream, (bytes, bytearray)) else stream, np.uint8)
    U = np.empty((r1, r2), np.float32)
    V = np.empty((r1, r2), np.float32)
    _check(lib().cpsz_decompress_2d_log_f32(_ptr(stream), stream.size, r1, r2, _ptr(U), _ptr(V)))
    return U, V


def sz_compress_cp_preserve_3d_offline_log(U, V, W, max_pwr_eb: float = 0.1) -> np.ndarray:
    U, V, W = _f32(U), _f32(V), _f32(W)
    if not (U.shape == V.shape == W.shape) or U.ndim != 3:
        raise ValueError("U, V, W must be 3D arrays of the same shape")
    r1, r2, r3 = U.shape
    p, s = _P(), _SZ()
    _check(lib().cpsz_compress_3d_f32(_ptr(U), _ptr(V), _ptr(W), r1, r2, r3, max_pwr_eb, C.byref(p),
                                      C.byref(s)))
    return _take(p, s.value)


def sz_decompress_cp_preserve_3d_online_log(stream, r1: int, r2: int, r3: int):
    stream = np.ascontiguousarray(stream, np.uint8)
    out = [np.empty((r1, r2, r3), np.float32) for _ in range(3)]
    _check(lib().cpsz_decompress_3d_f32(_ptr(stream), stream.size, r1, r2, r3,
                                        *[_ptr(o) for o in out]))
    return tuple(out)


def sz_lossless_compress(data, level: int = 3) -> np.ndarray:
    data = np.ascontiguousarray(data, np.uint8)
    p, s = _P(), _SZ()
    _check(lib().cpsz_lossless_compress(level, _ptr(data), data.size, C.byref(p), C.byref(s)))
    return _take(p, s.value)


def sz_lossless_compress_dev(d_in: int, n: int, level: int = 3, device: int = 0, stream: int = 0) -> np.ndarray:
    """sz_lossless_compress of n bytes at device address d_in (pipelined D2H)."""
    p, s = _P(), _SZ()
    _check(lib().cpsz_lossless_compress_dev(device, stream or None, level, d_in, n, C.byref(p), C.byref(s)))
    return _take(p, s.value)


def sz_lossless_compress_batch_dev(base: int, off, size, level: int = 3, device: int = 0, stream: int = 0):
    """One ZSTD frame per device stream base + off[k] (size[k] bytes)."""
    off = np.ascontiguousarray(off, np.uint64)
    size = np.ascontiguousarray(size, np.uint64)
    nf = size.size
    outs = (_P * nf)()
    sizes = np.zeros(nf, np.uint64)
    _check(lib().cpsz_lossless_compress_batch_dev(device, stream or None, level, base, _ptr(off), _ptr(size), nf,
                                                  C.cast(outs, _P), _ptr(sizes)))
    return [_take(outs[k], int(sizes[k])) for k in range(nf)]


def sz_lossless_decompress(data, size: int) -> np.ndarray:
    data = np.ascontiguousarray(data, np.uint8)
    out = np.empty(size, np.uint8)
    _check(lib().cpsz_lossless_decompress(_ptr(data), data.size, _ptr(out), size))
    return out


def stream_size_2d(stream) -> int:
    s = _SZ()
    _check(lib().cpsz_stream_size_2d(_ptr(np.ascontiguousarray(stream, np.uint8)), C.byref(s)))
    return s.value


# ----------------------------------------------------------- device API ----
def compress_2d_dev(dU: int, dV: int, r1: int, r2: int, max_pwr_eb: float = 0.1, device: int = 0,
                    stream: int = 0):
    """Returns (device pointer, size) of the stream in a library-owned buffer."""
    p, s = _P(), _SZ()
    _check(lib().cpsz_compress_2d_f32_dev(device, stream or None, dU, dV, r1, r2, max_pwr_eb,
                                          C.byref(p), C.byref(s)))
    return p.value, s.value


def decompress_2d_dev(d_in: int, in_size: int, r1: int, r2: int, dU: int, dV: int, device: int = 0,
                      stream: int = 0):
    _check(lib().cpsz_decompress_2d_f32_dev(device, stream or None, d_in, in_size, r1, r2, dU, dV))


def compress_3d_dev(dU: int, dV: int, dW: int, r1: int, r2: int, r3: int, max_pwr_eb: float = 0.1,
                    device: int = 0, stream: int = 0):
    p, s = _P(), _SZ()
    _check(lib().cpsz_compress_3d_f32_dev(device, stream or None, dU, dV, dW, r1, r2, r3, max_pwr_eb,
                                          C.byref(p), C.byref(s)))
    return p.value, s.value


def decompress_3d_dev(d_in: int, in_size: int, r1: int, r2: int, r3: int, dU: int, dV: int, dW: int,
                      device: int = 0, stream: int = 0):
    _check(lib().cpsz_decompress_3d_f32_dev(device, stream or None, d_in, in_size, r1, r2, r3, dU, dV,
                                            dW))


def critical_points_2d_dev(dU: int, dV: int, r1: int, r2: int, device: int = 0, stream: int = 0,
                           cap: int = 1 << 20):
    ids = np.empty(cap, np.int64)
    types = np.empty(cap, np.int32)
    cnt = _SZ()
    _check(lib().cpsz_critical_points_2d_dev(device, stream or None, dU, dV, r1, r2, _ptr(ids),
                                             _ptr(types), cap, C.byref(cnt)))
    n = min(cnt.value, cap)
    if cnt.value > cap:
        raise CpszError(-3, f"{cnt.value} critical points exceed cap {cap}")
    return dict(zip(ids[:n].tolist(), types[:n].tolist()))


def critical_points_3d_dev(dU: int, dV: int, dW: int, r1: int, r2: int, r3: int, device: int = 0,
                           stream: int = 0, cap: int = 1 << 22):
    ids = np.empty(cap, np.int64)
    types = np.empty(cap, np.int32)
    cnt = _SZ()
    _check(lib().cpsz_critical_points_3d_dev(device, stream or None, dU, dV, dW, r1, r2, r3,
                                             _ptr(ids), _ptr(types), cap, C.byref(cnt)))
    if cnt.value > cap:
        raise CpszError(-3, f"{cnt.value} critical points exceed cap {cap}")
    n = cnt.value
    return dict(zip(ids[:n].tolist(), types[:n].tolist()))


def critical_points_2d_sos_dev(dU: int, dV: int, r1: int, r2: int, device: int = 0, stream: int = 0,
                               cap: int = 1 << 20):
    """The 2D judge under Simulation of Simplicity: {simplex id: type}."""
    ids = np.empty(cap, np.int64)
    types = np.empty(cap, np.int32)
    cnt = _SZ()
    _check(lib().cpsz_critical_points_2d_sos_dev(device, stream or None, dU, dV, r1, r2, _ptr(ids), _ptr(types), cap,
                                                 C.byref(cnt)))
    if cnt.value > cap:
        raise CpszError(-3, f"{cnt.value} critical points exceed cap {cap}")
    n = cnt.value
    return dict(zip(ids[:n].tolist(), types[:n].tolist()))


def critical_points_3d_sos_dev(dU: int, dV: int, dW: int, r1: int, r2: int, r3: int, device: int = 0,
                               stream: int = 0, cap: int = 1 << 22):
    """The 3D judge under Simulation of Simplicity: {simplex id: type}."""
    ids = np.empty(cap, np.int64)
    types = np.empty(cap, np.int32)
    cnt = _SZ()
    _check(lib().cpsz_critical_points_3d_sos_dev(device, stream or None, dU, dV, dW, r1, r2, r3, _ptr(ids),
                                                 _ptr(types), cap, C.byref(cnt)))
    if cnt.value > cap:
        raise CpszError(-3, f"{cnt.value} critical points exceed cap {cap}")
    n = cnt.value
    return dict(zip(ids[:n].tolist(), types[:n].tolist()))


def sos_det(rows, idx) -> int:
    """Host evaluation of the device's SoS determinant sign (d = len(idx))."""
    rows = np.ascontiguousarray(rows, np.float32).ravel()
    idx = np.ascontiguousarray(idx, np.int64)
    s = C.c_int()
    _check(lib().cpsz_sos_det(len(idx), _ptr(rows), _ptr(idx), C.byref(s)))
    return s.value


def set_timing(on: bool = True, device: int = 0):
    _check(lib().cpsz_set_timing(device, int(on)))


def last_stages(device: int = 0):
    names = C.create_string_buffer(4096)
    ms = (C.c_float * 64)()
    n = _check(lib().cpsz_last_stages(device, names, 4096, C.cast(ms, _P), 64))
    keys = names.value.decode().split(";") if n else []
    return list(zip(keys, [ms[i] for i in range(n)]))


def device_count() -> int:
    return lib().cpsz_device_count()


# ---------------------------------------------------------------- learned eb --
def learn_param_count() -> int:
    return int(lib().cpsz_learn_param_count())


def learn_init_params(seed: int = 0) -> np.ndarray:
    p = np.zeros(learn_param_count(), np.float32)
    _check(lib().cpsz_learn_init_params(seed, _ptr(p)))
    return p


def learn_analytic_eb_2d_dev(dU: int, dV: int, r1: int, r2: int, eb: float, dE: int, device: int = 0,
                             stream: int = 0):
    _check(lib().cpsz_learn_analytic_eb_2d_dev(device, stream or None, dU, dV, r1, r2, eb, dE))


def learn_features_2d_dev(dU, dV, dE, r1, r2, eb, d_idx, m, dX, device=0, stream=0):
    _check(lib().cpsz_learn_features_2d_dev(device, stream or None, dU, dV, dE, r1, r2, eb, d_idx or None,
                                            m, dX))


def learn_labels_2d_dev(dU, dV, dE, r1, r2, eb, d_idx, m, dL, device=0, stream=0):
    _check(lib().cpsz_learn_labels_2d_dev(device, stream or None, dU, dV, dE, r1, r2, eb, d_idx or None,
                                          m, dL))


def gemm_bf16_dev(M, N, K, A, sam, sak, B, sbk, sbn, Cm, ldc, split=1, device=0, stream=0):
    _check(lib().cpsz_gemm_bf16_dev(device, stream or None, M, N, K, A, sam, sak, B, sbk, sbn, Cm, ldc,
                                    split))


def mlp_forward_dev(params: np.ndarray, dX: int, m: int, dY: int, device: int = 0, stream: int = 0):
    params = np.ascontiguousarray(params, np.float32)
    _check(lib().cpsz_mlp_forward_dev(device, stream or None, _ptr(params), dX, m, dY))


def mlp_grad_dev(params: np.ndarray, dX: int, dT: int, m: int, device: int = 0, stream: int = 0):
    params = np.ascontiguousarray(params, np.float32)
    g = np.zeros_like(params)
    loss = np.zeros(1, np.float32)
    _check(lib().cpsz_mlp_grad_dev(device, stream or None, _ptr(params), dX, dT, m, _ptr(g), _ptr(loss)))
    return g, float(loss[0])


def learn_train_2d_dev(fields, eb: float, params: np.ndarray, samples_per_field: int = 16384,
                       epochs: int = 4, batch: int = 4096, lr: float = 1e-3, seed: int = 0,
                       device: int = 0, stream: int = 0):
    """fields: list of (dU, dV, r1, r2) with device pointers. Returns (params, loss per epoch)."""
    n = len(fields)
    du = (C.c_void_p * n)(*[f[0] for f in fields])
    dv = (C.c_void_p * n)(*[f[1] for f in fields])
    r1 = (C.c_size_t * n)(*[f[2] for f in fields])
    r2 = (C.c_size_t * n)(*[f[3] for f in fields])
    p = np.ascontiguousarray(params, np.float32).copy()
    losses = np.zeros(epochs, np.float32)
    _check(lib().cpsz_learn_train_2d_dev(device, stream or None, n, du, dv, r1, r2, eb, samples_per_field,
                                         epochs, batch, lr, seed, _ptr(p), _ptr(losses)))
    return p, losses


def compress_2d_learned_dev(dU: int, dV: int, r1: int, r2: int, eb: float, params: np.ndarray,
                            device: int = 0, stream: int = 0):
    """Returns (device pointer, size, verification rounds, boosted vertices, fallback vertices)."""
    params = np.ascontiguousarray(params, np.float32)
    p, s, it, bo, fb = _P(), _SZ(), C.c_int(), _SZ(), _SZ()
    _check(lib().cpsz_compress_2d_learned_f32_dev(device, stream or None, dU, dV, r1, r2, eb, _ptr(params),
                                                  C.byref(p), C.byref(s), C.byref(it), C.byref(bo),
                                                  C.byref(fb)))
    return p.value, s.value, it.value, bo.value, fb.value


def sz_compress_cp_preserve_2d_learned(U, V, eb: float, params: np.ndarray) -> np.ndarray:
    """Learned-bound 2D compression of host arrays (reference stream layout)."""
    U = np.ascontiguousarray(U, np.float32)
    V = np.ascontiguousarray(V, np.float32)
    params = np.ascontiguousarray(params, np.float32)
    p, s = _P(), _SZ()
    _check(lib().cpsz_compress_2d_learned_f32(_ptr(U), _ptr(V), U.shape[0], U.shape[1], eb, _ptr(params),
                                              C.byref(p), C.byref(s)))
    return _take(p, s.value)


# --------------------------------------------------------------- 2D batches --
def _arr(ctype, vals):
    return (ctype * len(vals))(*vals)


def compress_2d_batch_dev(fields, eb: float, tune: bool = False, device: int = 0, stream: int = 0):
    """fields: list of (dU, dV, r1, r2) device pointers. Returns (base pointer of the
    library-owned stream buffer, offsets, sizes)."""
    n = len(fields)
    base = _P()
    off = np.zeros(n, np.uint64)
    size = np.zeros(n, np.uint64)
    _check(lib().cpsz_compress_2d_batch_dev(device, stream or None, n, _arr(C.c_void_p, [f[0] for f in fields]),
                                            _arr(C.c_void_p, [f[1] for f in fields]),
                                            _arr(C.c_size_t, [f[2] for f in fields]),
                                            _arr(C.c_size_t, [f[3] for f in fields]), eb, int(tune),
                                            C.byref(base), _ptr(off), _ptr(size)))
    return base.value, off.astype(np.int64), size.astype(np.int64)


def decompress_2d_batch_dev(streams, shapes, outs, device: int = 0, stream: int = 0):
    """streams: list of (device pointer, size); shapes: (r1, r2); outs: (dU, dV) pointers."""
    n = len(streams)
    _check(lib().cpsz_decompress_2d_batch_dev(device, stream or None, n, _arr(C.c_void_p, [s[0] for s in streams]),
                                              _arr(C.c_size_t, [s[1] for s in streams]),
                                              _arr(C.c_size_t, [s[0] for s in shapes]),
                                              _arr(C.c_size_t, [s[1] for s in shapes]),
                                              _arr(C.c_void_p, [o[0] for o in outs]),
                                              _arr(C.c_void_p, [o[1] for o in outs])))


def batch_2d_config(shapes, device: int = 0):
    """(workers, grid cap, trees on the GPU) the batch calls use for batches of
    these shapes."""
    n = len(shapes)
    w, g, t = C.c_int(), C.c_int(), C.c_int()
    _check(lib().cpsz_batch_2d_config(device, n, _arr(C.c_size_t, [s[0] for s in shapes]),
                                      _arr(C.c_size_t, [s[1] for s in shapes]), C.byref(w), C.byref(g),
                                      C.byref(t)))
    return w.value, g.value, bool(t.value)


def batch_2d_set_config(shapes, workers: int, grid_cap: int, device_trees: bool, device: int = 0):
    """Pin the configuration for batches of these shapes."""
    _check(lib().cpsz_batch_2d_set_config(device, len(shapes), _arr(C.c_size_t, [s[0] for s in shapes]),
                                          _arr(C.c_size_t, [s[1] for s in shapes]), workers, grid_cap,
                                          int(device_trees)))


# --------------------------------------------------------------- 3D batches --
def compress_3d_batch_dev(fields, eb: float, tune: bool = False, device: int = 0, stream: int = 0):
    """fields: list of (dU, dV, dW, r1, r2, r3) device pointers. Returns (base pointer of the
    library-owned stream buffer, offsets, sizes)."""
    n = len(fields)
    base = _P()
    off = np.zeros(n, np.uint64)
    size = np.zeros(n, np.uint64)
    _check(lib().cpsz_compress_3d_batch_dev(device, stream or None, n,
                                            *[_arr(C.c_void_p, [f[k] for f in fields]) for k in range(3)],
                                            *[_arr(C.c_size_t, [f[k] for f in fields]) for k in range(3, 6)],
                                            eb, int(tune), C.byref(base), _ptr(off), _ptr(size)))
    return base.value, off.astype(np.int64), size.astype(np.int64)


def decompress_3d_batch_dev(streams, shapes, outs, device: int = 0, stream: int = 0):
    """streams: list of (device pointer, size); shapes: (r1, r2, r3); outs: (dU, dV, dW) pointers."""
    n = len(streams)
    _check(lib().cpsz_decompress_3d_batch_dev(device, stream or None, n, _arr(C.c_void_p, [s[0] for s in streams]),
                                              _arr(C.c_size_t, [s[1] for s in streams]),
                                              *[_arr(C.c_size_t, [s[k] for s in shapes]) for k in range(3)],
                                              *[_arr(C.c_void_p, [o[k] for o in outs]) for k in range(3)]))


def batch_3d_config(shapes, device: int = 0):
    """(workers, grid cap) the 3D batch calls use for batches of these shapes."""
    n = len(shapes)
    w, g = C.c_int(), C.c_int()
    _check(lib().cpsz_batch_3d_config(device, n, *[_arr(C.c_size_t, [s[k] for s in shapes]) for k in range(3)],
                                      C.byref(w), C.byref(g)))
    return w.value, g.value


# ------------------------------------------------------------- learned eb 3D --
def learn_param_count_3d() -> int:
    return int(lib().cpsz_learn_param_count_3d())


def learn_init_params_3d(seed: int = 0) -> np.ndarray:
    p = np.zeros(learn_param_count_3d(), np.float32)
    _check(lib().cpsz_learn_init_params_3d(seed, _ptr(p)))
    return p


def learn_analytic_eb_3d_dev(dU, dV, dW, r1, r2, r3, eb, dE, device=0, stream=0):
    _check(lib().cpsz_learn_analytic_eb_3d_dev(device, stream or None, dU, dV, dW, r1, r2, r3, eb, dE))


def learn_features_3d_dev(dU, dV, dW, dE, r1, r2, r3, eb, d_idx, m, dX, device=0, stream=0):
    _check(lib().cpsz_learn_features_3d_dev(device, stream or None, dU, dV, dW, dE, r1, r2, r3, eb,
                                            d_idx or None, m, dX))


def learn_labels_3d_dev(dU, dV, dW, dE, r1, r2, r3, eb, d_idx, m, dL, device=0, stream=0):
    _check(lib().cpsz_learn_labels_3d_dev(device, stream or None, dU, dV, dW, dE, r1, r2, r3, eb,
                                          d_idx or None, m, dL))


def mlp_forward_3d_dev(params: np.ndarray, dX: int, m: int, dY: int, device: int = 0, stream: int = 0):
    params = np.ascontiguousarray(params, np.float32)
    _check(lib().cpsz_mlp_forward_3d_dev(device, stream or None, _ptr(params), dX, m, dY))


def mlp_grad_3d_dev(params: np.ndarray, dX: int, dT: int, m: int, device: int = 0, stream: int = 0):
    params = np.ascontiguousarray(params, np.float32)
    g = np.zeros_like(params)
    loss = np.zeros(1, np.float32)
    _check(lib().cpsz_mlp_grad_3d_dev(device, stream or None, _ptr(params), dX, dT, m, _ptr(g), _ptr(loss)))
    return g, float(loss[0])


def learn_train_3d_dev(fields, eb: float, params: np.ndarray, samples_per_field: int = 16384,
                       epochs: int = 4, batch: int = 4096, lr: float = 1e-3, seed: int = 0,
                       device: int = 0, stream: int = 0):
    """fields: list of (dU, dV, dW, r1, r2, r3) with device pointers. Returns (params, loss per epoch)."""
    n = len(fields)
    p = np.ascontiguousarray(params, np.float32).copy()
    losses = np.zeros(epochs, np.float32)
    _check(lib().cpsz_learn_train_3d_dev(device, stream or None, n,
                                         *[_arr(C.c_void_p, [f[k] for f in fields]) for k in range(3)],
                                         *[_arr(C.c_size_t, [f[k] for f in fields]) for k in range(3, 6)],
                                         eb, samples_per_field, epochs, batch, lr, seed, _ptr(p), _ptr(losses)))
    return p, losses


def compress_2d_learned_batch_dev(fields, eb: float, params: np.ndarray, tune: bool = False, device: int = 0,
                                  stream: int = 0):
    """fields: list of (dU, dV, r1, r2) device pointers. Returns (base pointer, offsets, sizes,
    verification rounds, boosted vertices, fallback vertices)."""
    n = len(fields)
    params = np.ascontiguousarray(params, np.float32)
    base = _P()
    off = np.zeros(n, np.uint64)
    size = np.zeros(n, np.uint64)
    it, bo, fb = C.c_int(), _SZ(), _SZ()
    _check(lib().cpsz_compress_2d_learned_batch_dev(device, stream or None, n,
                                                    *[_arr(C.c_void_p, [f[k] for f in fields]) for k in range(2)],
                                                    *[_arr(C.c_size_t, [f[k] for f in fields]) for k in range(2, 4)],
                                                    eb, _ptr(params), int(tune), C.byref(base), _ptr(off), _ptr(size),
                                                    C.byref(it), C.byref(bo), C.byref(fb)))
    return base.value, off.astype(np.int64), size.astype(np.int64), it.value, bo.value, fb.value


def compress_3d_learned_batch_dev(fields, eb: float, params: np.ndarray, tune: bool = False, device: int = 0,
                                  stream: int = 0):
    """fields: list of (dU, dV, dW, r1, r2, r3) device pointers. Returns (base pointer, offsets, sizes,
    verification rounds, boosted vertices, fallback vertices)."""
    n = len(fields)
    params = np.ascontiguousarray(params, np.float32)
    base = _P()
    off = np.zeros(n, np.uint64)
    size = np.zeros(n, np.uint64)
    it, bo, fb = C.c_int(), _SZ(), _SZ()
    _check(lib().cpsz_compress_3d_learned_batch_dev(device, stream or None, n,
                                                    *[_arr(C.c_void_p, [f[k] for f in fields]) for k in range(3)],
                                                    *[_arr(C.c_size_t, [f[k] for f in fields]) for k in range(3, 6)],
                                                    eb, _ptr(params), int(tune), C.byref(base), _ptr(off), _ptr(size),
                                                    C.byref(it), C.byref(bo), C.byref(fb)))
    return base.value, off.astype(np.int64), size.astype(np.int64), it.value, bo.value, fb.value


def check_fast_bounds(device: int = 0):
    """Device check of the approximate 3D bound pass's premises: (largest ulp
    distance of the fast cube root over [0, 2^126), largest error-to-envelope
    ratio of the fast sign bound over hashed argument pairs)."""
    u, r = C.c_uint32(), C.c_double()
    _check(lib().cpsz_check_fast_bounds(device, C.byref(u), C.byref(r)))
    return u.value, r.value


def set_exact_2d_bounds(on: bool) -> None:
    """Process-wide choice of the 2D bounds kernel for float fields: False
    (default) = FP32 approximations of the last FP64 steps with exact fix-ups,
    True = every step in FP64. Both give identical streams."""
    _check(lib().cpsz_set_exact_2d_bounds(int(bool(on))))


def tri_bound_2d_fast_host(uv6, ulps: int = 0):
    """Host evaluation of the approximate 2D bound's cell-table entries for
    triangles uv6 (n x 6: u0, u1, u2, v0, v1, v2), square roots and reciprocals
    moved by up to `ulps` ulp: 0 / 2.0 = exact 0 / 1, -1 = exact bound needed,
    else an approximation."""
    uv6 = np.ascontiguousarray(uv6, np.float32).reshape(-1, 6)
    out = np.empty(uv6.shape[0], np.float32)
    _check(lib().cpsz_tri_bound_2d_fast_host(_ptr(uv6), uv6.shape[0], int(ulps), _ptr(out)))
    return out


def ratio_ge1_host(det, d):
    """Host evaluation of the bounds pass's division-free (det / d >= 1)."""
    det = np.ascontiguousarray(det, np.float64)
    d = np.ascontiguousarray(d, np.float64)
    out = np.empty(det.size, np.uint8)
    _check(lib().cpsz_ratio_ge1_host(_ptr(det), _ptr(d), det.size, _ptr(out)))
    return out.astype(bool)


def tet_cp_emulate(f, t: int) -> int:
    """Host evaluation of the device judge's tetrahedron predicate (-1 or the type code)."""
    f = np.ascontiguousarray(f, np.float32).reshape(12)
    ty = C.c_int()
    _check(lib().cpsz_tet_cp_emulate(_ptr(f), t, C.byref(ty)))
    return ty.value
