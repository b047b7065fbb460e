// extern "C" boundary of libcpsz_hip.so (declared in include/cpsz.h).
#define ZSTD_STATIC_LINKING_ONLY  // the block-level API (ZSTD_compressBegin_advanced, ZSTD_compressContinue)
#include <zstd.h>

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cpsz.h"
#include "batch2d.hpp"
#include "batch3d.hpp"
#include "codec2d.hpp"
#include "codec3d.hpp"
#include "cp.hpp"
#include "cp_math.hpp"
#include "ctx.hpp"
#include "bounds2d_fast.hpp"
#include "eb3d_math.hpp"
#include "huff_gpu.hpp"
#include "huff_host.hpp"
#include "learn.hpp"

namespace cpsz {
namespace {

thread_local std::string g_err;
bool g_exiting = false;

struct ThreadCtxs {
  std::map<int, Ctx*> m;
  ~ThreadCtxs() {
    if (g_exiting) return;  // the HIP runtime may already be gone at process exit
    for (auto& kv : m) delete kv.second;
  }
};
thread_local ThreadCtxs g_ctxs;

Ctx& ctx_of(int device) {
  static bool reg = (std::atexit([] { g_exiting = true; }), true);
  (void)reg;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    fail(kErrHip, "no HIP device available");
  CPSZ_REQUIRE(device >= 0 && device < count, kErrArg, "device index out of range");
  Ctx*& c = g_ctxs.m[device];
  if (!c) c = new Ctx(device);
  CPSZ_HIP(hipSetDevice(device));
  return *c;
}

int current_device() {
  int d = 0;
  if (hipGetDevice(&d) != hipSuccess) d = 0;
  return d;
}

template <class F>
int guarded(F f) {
  try {
    f();
    g_err.clear();
    return kOk;
  } catch (const Error& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "host allocation failed";
    return kErrAlloc;
  } catch (const std::exception& e) {
    g_err = e.what();
    return kErrInternal;
  }
}

struct StreamScope {  // temporarily routes a context onto a caller stream
  Ctx& c;
  StreamScope(Ctx& ctx, void* s) : c(ctx) { c.set_stream((hipStream_t)s); }
  ~StreamScope() { c.set_stream(nullptr); }
};

// Device copies of host buffers, owned for one call.
struct DevMem {
  void* p = nullptr;
  explicit DevMem(size_t bytes) {
    if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
      p = nullptr;
      fail(kErrAlloc, "device allocation failed");
    }
  }
  ~DevMem() {
    if (p) hipFree(p);
  }
};

}  // namespace
}  // namespace cpsz

using namespace cpsz;

extern "C" {

const char* cpsz_last_error(void) { return g_err.c_str(); }
const char* cpsz_version(void) { return "cpsz-mi355x 0.1 (gfx950)"; }
int cpsz_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}
void cpsz_free(void* p) { std::free(p); }

int cpsz_compress_2d_f32(const float* U, const float* V, size_t r1, size_t r2, double max_pwr_eb,
                         unsigned char** out, size_t* out_size) {
  return guarded([&] {
    CPSZ_REQUIRE(U && V && out && out_size && r1 && r2, kErrArg, "null argument or empty field");
    Ctx& c = ctx_of(current_device());
    const size_t n = r1 * r2;
    float* d = c.buf<float>("api.in", 2 * n);
    CPSZ_HIP(hipMemcpyAsync(d, U, n * 4, hipMemcpyHostToDevice, c.stream()));
    CPSZ_HIP(hipMemcpyAsync(d + n, V, n * 4, hipMemcpyHostToDevice, c.stream()));
    DeviceStream s;
    compress_2d_device(c, d, d + n, r1, r2, max_pwr_eb, s);
    unsigned char* h = (unsigned char*)std::malloc(s.size ? s.size : 1);
    CPSZ_REQUIRE(h, kErrAlloc, "host allocation failed");
    CPSZ_HIP(hipMemcpyAsync(h, s.data, s.size, hipMemcpyDeviceToHost, c.stream()));
    CPSZ_HIP(hipStreamSynchronize(c.stream()));
    *out = h;
    *out_size = s.size;
  });
}

int cpsz_decompress_2d_f32(const unsigned char* in, size_t in_size, size_t r1, size_t r2, float* U,
                           float* V) {
  return guarded([&] {
    CPSZ_REQUIRE(in && U && V && r1 && r2 && in_size >= 36, kErrArg, "bad arguments");
    Ctx& c = ctx_of(current_device());
    const size_t n = r1 * r2;
    uint8_t* din = c.buf<uint8_t>("api.stream", in_size + 16);
    CPSZ_HIP(hipMemcpyAsync(din, in, in_size, hipMemcpyHostToDevice, c.stream()));
    float* d = c.buf<float>("api.out", 2 * n);
    decompress_2d_device(c, din, in, in_size, r1, r2, d, d + n);
    CPSZ_HIP(hipMemcpyAsync(U, d, n * 4, hipMemcpyDeviceToHost, c.stream()));
    CPSZ_HIP(hipMemcpyAsync(V, d + n, n * 4, hipMemcpyDeviceToHost, c.stream()));
    CPSZ_HIP(hipStreamSynchronize(c.stream()));
  });
}

int cpsz_compress_2d_f64(const double* U, const double* V, size_t r1, size_t r2, double max_pwr_eb,
                         unsigned char** out, size_t* out_size) {
  return guarded([&] {
    CPSZ_REQUIRE(U && V && out && out_size && r1 && r2, kErrArg, "null argument or empty field");
    Ctx& c = ctx_of(current_device());
    const size_t n = r1 * r2;
    double* d = c.buf<double>("api.in64", 2 * n);
    CPSZ_HIP(hipMemcpyAsync(d, U, n * 8, hipMemcpyHostToDevice, c.stream()));
    CPSZ_HIP(hipMemcpyAsync(d + n, V, n * 8, hipMemcpyHostToDevice, c.stream()));
    DeviceStream s;
    compress_2d_f64_device(c, d, d + n, r1, r2, max_pwr_eb, s);
    unsigned char* h = (unsigned char*)std::malloc(s.size ? s.size : 1);
    CPSZ_REQUIRE(h, kErrAlloc, "host allocation failed");
    CPSZ_HIP(hipMemcpyAsync(h, s.data, s.size, hipMemcpyDeviceToHost, c.stream()));
    CPSZ_HIP(hipStreamSynchronize(c.stream()));
    *out = h;
    *out_size = s.size;
  });
}

int cpsz_decompress_2d_f64(const unsigned char* in, size_t in_size, size_t r1, size_t r2, double* U, double* V) {
  return guarded([&] {
    CPSZ_REQUIRE(in && U && V && r1 && r2 && in_size >= 36, kErrArg, "bad arguments");
    Ctx& c = ctx_of(current_device());
    const size_t n = r1 * r2;
    uint8_t* din = c.buf<uint8_t>("api.stream", in_size + 16);
    CPSZ_HIP(hipMemcpyAsync(din, in, in_size, hipMemcpyHostToDevice, c.stream()));
    double* d = c.buf<double>("api.out64", 2 * n);
    decompress_2d_f64_device(c, din, in, in_size, r1, r2, d, d + n);
    CPSZ_HIP(hipMemcpyAsync(U, d, n * 8, hipMemcpyDeviceToHost, c.stream()));
    CPSZ_HIP(hipMemcpyAsync(V, d + n, n * 8, hipMemcpyDeviceToHost, c.stream()));
    CPSZ_HIP(hipStreamSynchronize(c.stream()));
  });
}

int cpsz_compress_2d_f64_dev(int device, void* stream, const double* dU, const double* dV, size_t r1, size_t r2,
                             double max_pwr_eb, const unsigned char** d_out, size_t* out_size) {
  return guarded([&] {
    CPSZ_REQUIRE(d_out && out_size, kErrArg, "null output argument");
    Ctx& c = ctx_of(device);
    StreamScope ss(c, stream);
    DeviceStream s;
    compress_2d_f64_device(c, dU, dV, r1, r2, max_pwr_eb, s);
    *d_out = s.data;
    *out_size = s.size;
  });
}

int cpsz_decompress_2d_f64_dev(int device, void* stream, const unsigned char* d_in, size_t in_size, size_t r1,
                               size_t r2, double* dU, double* dV) {
  return guarded([&] {
    Ctx& c = ctx_of(device);
    StreamScope ss(c, stream);
    decompress_2d_f64_device(c, d_in, nullptr, in_size, r1, r2, dU, dV);
  });
}

int cpsz_compress_2d_log_f32(const float* U, const float* V, size_t r1, size_t r2, double max_pwr_eb,
                             unsigned char** out, size_t* out_size) {
  return guarded([&] {
    CPSZ_REQUIRE(U && V && out && out_size && r1 && r2, kErrArg, "null argument or empty field");
    Ctx& c = ctx_of(current_device());
    const size_t n = r1 * r2;
    float* d = c.buf<float>("api.in", 2 * n);
    CPSZ_HIP(hipMemcpyAsync(d, U, n * 4, hipMemcpyHostToDevice, c.stream()));
    CPSZ_HIP(hipMemcpyAsync(d + n, V, n * 4, hipMemcpyHostToDevice, c.stream()));
    DeviceStream s;
    compress_2d_log_device(c, d, d + n, r1, r2, max_pwr_eb, s);
    unsigned char* h = (unsigned char*)std::malloc(s.size ? s.size : 1);
    CPSZ_REQUIRE(h, kErrAlloc, "host allocation failed");
    CPSZ_HIP(hipMemcpyAsync(h, s.data, s.size, hipMemcpyDeviceToHost, c.stream()));
    CPSZ_HIP(hipStreamSynchronize(c.stream()));
    *out = h;
    *out_size = s.size;
  });
}

int cpsz_decompress_2d_log_f32(const unsigned char* in, size_t in_size, size_t r1, size_t r2, float* U,
                               float* V) {
  return guarded([&] {
    CPSZ_REQUIRE(in && U && V && r1 && r2 && in_size >= 36, kErrArg, "bad arguments");
    Ctx& c = ctx_of(current_device());
    const size_t n = r1 * r2;
    uint8_t* din = c.buf<uint8_t>("api.stream", in_size + 16);
    CPSZ_HIP(hipMemcpyAsync(din, in, in_size, hipMemcpyHostToDevice, c.stream()));
    float* d = c.buf<float>("api.out", 2 * n);
    decompress_2d_log_device(c, din, in, in_size, r1, r2, d, d + n);
    CPSZ_HIP(hipMemcpyAsync(U, d, n * 4, hipMemcpyDeviceToHost, c.stream()));
    CPSZ_HIP(hipMemcpyAsync(V, d + n, n * 4, hipMemcpyDeviceToHost, c.stream()));
    CPSZ_HIP(hipStreamSynchronize(c.stream()));
  });
}

int cpsz_compress_2d_log_f32_dev(int device, void* stream, const float* dU, const float* dV, size_t r1,
                                 size_t r2, double max_pwr_eb, const unsigned char** d_out, size_t* out_size) {
  return guarded([&] {
    CPSZ_REQUIRE(d_out && out_size, kErrArg, "null output argument");
    Ctx& c = ctx_of(device);
    StreamScope ss(c, stream);
    DeviceStream s;
    compress_2d_log_device(c, dU, dV, r1, r2, max_pwr_eb, s);
    *d_out = s.data;
    *out_size = s.size;
  });
}

int cpsz_decompress_2d_log_f32_dev(int device, void* stream, const unsigned char* d_in, size_t in_size,
                                   size_t r1, size_t r2, float* dU, float* dV) {
  return guarded([&] {
    Ctx& c = ctx_of(device);
    StreamScope ss(c, stream);
    decompress_2d_log_device(c, d_in, nullptr, in_size, r1, r2, dU, dV);
  });
}

int cpsz_stream_size_2d_log(const unsigned char* in, size_t r1, size_t r2, size_t* size) {
  return guarded([&] {
    CPSZ_REQUIRE(in && size && r1 && r2, kErrArg, "bad arguments");
    size_t eb, u, v;
    std::memcpy(&eb, in + 12, 8);
    std::memcpy(&u, in + 20, 8);
    std::memcpy(&v, in + 28, 8);
    *size = 36 + eb + 2 * ((r1 * r2 - 1) / 8 + 1) + u + v;
  });
}

int cpsz_compress_3d_f32(const float* U, const float* V, const float* W, size_t r1, size_t r2,
                         size_t r3, double max_pwr_eb, unsigned char** out, size_t* out_size) {
  return guarded([&] {
    CPSZ_REQUIRE(U && V && W && out && out_size && r1 && r2 && r3, kErrArg, "bad arguments");
    Ctx& c = ctx_of(current_device());
    const size_t n = r1 * r2 * r3;
    float* d = c.buf<float>("api.in", 3 * n);
    CPSZ_HIP(hipMemcpyAsync(d, U, n * 4, hipMemcpyHostToDevice, c.stream()));
    CPSZ_HIP(hipMemcpyAsync(d + n, V, n * 4, hipMemcpyHostToDevice, c.stream()));
    CPSZ_HIP(hipMemcpyAsync(d + 2 * n, W, n * 4, hipMemcpyHostToDevice, c.stream()));
    DeviceStream s;
    compress_3d_device(c, d, d + n, d + 2 * n, r1, r2, r3, max_pwr_eb, s);
    unsigned char* h = (unsigned char*)std::malloc(s.size ? s.size : 1);
    CPSZ_REQUIRE(h, kErrAlloc, "host allocation failed");
    CPSZ_HIP(hipMemcpyAsync(h, s.data, s.size, hipMemcpyDeviceToHost, c.stream()));
    CPSZ_HIP(hipStreamSynchronize(c.stream()));
    *out = h;
    *out_size = s.size;
  });
}

int cpsz_decompress_3d_f32(const unsigned char* in, size_t in_size, size_t r1, size_t r2,
                           size_t r3, float* U, float* V, float* W) {
  return guarded([&] {
    CPSZ_REQUIRE(in && U && V && W && r1 && r2 && r3 && in_size >= 8, kErrArg, "bad arguments");
    Ctx& c = ctx_of(current_device());
    const size_t n = r1 * r2 * r3;
    uint8_t* din = c.buf<uint8_t>("api.stream", in_size + 16);
    CPSZ_HIP(hipMemcpyAsync(din, in, in_size, hipMemcpyHostToDevice, c.stream()));
    float* d = c.buf<float>("api.out", 3 * n);
    decompress_3d_device(c, din, in, in_size, r1, r2, r3, d, d + n, d + 2 * n);
    CPSZ_HIP(hipMemcpyAsync(U, d, n * 4, hipMemcpyDeviceToHost, c.stream()));
    CPSZ_HIP(hipMemcpyAsync(V, d + n, n * 4, hipMemcpyDeviceToHost, c.stream()));
    CPSZ_HIP(hipMemcpyAsync(W, d + 2 * n, n * 4, hipMemcpyDeviceToHost, c.stream()));
    CPSZ_HIP(hipStreamSynchronize(c.stream()));
  });
}

int cpsz_stream_size_2d(const unsigned char* in, size_t* size) {
  return guarded([&] {
    CPSZ_REQUIRE(in && size, kErrArg, "null argument");
    size_t eb, u, v;
    std::memcpy(&eb, in + 12, 8);
    std::memcpy(&u, in + 20, 8);
    std::memcpy(&v, in + 28, 8);
    *size = 36 + eb + u + v;
  });
}

int cpsz_stream_size_3d(const unsigned char* in, size_t r1, size_t r2, size_t r3, size_t* size) {
  return guarded([&] {
    CPSZ_REQUIRE(in && size && r1 && r2 && r3, kErrArg, "bad arguments");
    const size_t n = r1 * r2 * r3;
    size_t off = 8 + 3 * ((n - 1) / 8 + 1);
    size_t unp;
    std::memcpy(&unp, in + off, 8);
    off += 8 + 4 * unp + 8;  // unpredictable values, eb_quant_num
    for (int s = 0; s < 2; s++) {
      size_t nodes, nbytes;
      uint32_t tree;
      std::memcpy(&nodes, in + off, 8);
      std::memcpy(&tree, in + off + 8, 4);
      std::memcpy(&nbytes, in + off + 12 + tree, 8);
      off += 20 + tree + nbytes;
      if (s == 0) off += 8;  // data_quant_num
      (void)nodes;
    }
    *size = off;
  });
}

int cpsz_compress_2d_f32_dev(int device, void* stream, const float* dU, const float* dV, size_t r1,
                             size_t r2, double max_pwr_eb, const unsigned char** d_out,
                             size_t* out_size) {
  return guarded([&] {
    CPSZ_REQUIRE(d_out && out_size, kErrArg, "null output argument");
    Ctx& c = ctx_of(device);
    StreamScope ss(c, stream);
    DeviceStream s;
    compress_2d_device(c, dU, dV, r1, r2, max_pwr_eb, s);
    *d_out = s.data;
    *out_size = s.size;
  });
}

int cpsz_decompress_2d_f32_dev(int device, void* stream, const unsigned char* d_in, size_t in_size,
                               size_t r1, size_t r2, float* dU, float* dV) {
  return guarded([&] {
    Ctx& c = ctx_of(device);
    StreamScope ss(c, stream);
    decompress_2d_device(c, d_in, nullptr, in_size, r1, r2, dU, dV);
  });
}

int cpsz_compress_3d_f32_dev(int device, void* stream, const float* dU, const float* dV,
                             const float* dW, size_t r1, size_t r2, size_t r3, double max_pwr_eb,
                             const unsigned char** d_out, size_t* out_size) {
  return guarded([&] {
    CPSZ_REQUIRE(d_out && out_size, kErrArg, "null output argument");
    Ctx& c = ctx_of(device);
    StreamScope ss(c, stream);
    DeviceStream s;
    compress_3d_device(c, dU, dV, dW, r1, r2, r3, max_pwr_eb, s);
    *d_out = s.data;
    *out_size = s.size;
  });
}

int cpsz_decompress_3d_f32_dev(int device, void* stream, const unsigned char* d_in, size_t in_size,
                               size_t r1, size_t r2, size_t r3, float* dU, float* dV, float* dW) {
  return guarded([&] {
    Ctx& c = ctx_of(device);
    StreamScope ss(c, stream);
    decompress_3d_device(c, d_in, nullptr, in_size, r1, r2, r3, dU, dV, dW);
  });
}

int cpsz_lossless_compress(int level, const unsigned char* in, size_t n, unsigned char** out,
                           size_t* out_size) {
  return guarded([&] {
    CPSZ_REQUIRE(in && out && out_size, kErrArg, "null argument");
    // The reference sizes its buffer as dataLength*1.2 (200 below 100 bytes),
    // src/sz_lossless.cpp:18-23; ZSTD_compressBound is never smaller.
    const size_t cap = ZSTD_compressBound(n);
    unsigned char* o = (unsigned char*)std::malloc(cap);
    CPSZ_REQUIRE(o, kErrAlloc, "host allocation failed");
    const size_t r = ZSTD_compress(o, cap, in, n, level);
    if (ZSTD_isError(r)) {
      std::free(o);
      fail(kErrInternal, std::string("ZSTD_compress: ") + ZSTD_getErrorName(r));
    }
    *out = o;
    *out_size = r;
  });
}

// ZSTD of a device-resident stream: the bytes come to one pinned host buffer
// in 8 MiB chunks on the caller's stream while the host compresses the chunks
// that have arrived, with zstd's block-level API (ZSTD_compressBegin_advanced
// with the parameters ZSTD_compress derives from the level and the size, then
// ZSTD_compressContinue per chunk and ZSTD_compressEnd). The chunks are
// contiguous and multiples of the 128 KiB block, so the compressor sees the
// same window and the same block boundaries as the one-shot ZSTD_compress of
// src/sz_lossless.cpp:11-29: the frame is byte-identical to the reference's.
int cpsz_lossless_compress_dev(int device, void* stream, int level, const unsigned char* d_in, size_t n,
                               unsigned char** out, size_t* out_size) {
  return guarded([&] {
    CPSZ_REQUIRE(d_in && out && out_size, kErrArg, "null argument");
    Ctx& c = ctx_of(device);
    StreamScope ss(c, stream);
    hipStream_t st = c.stream();
    const size_t cap = ZSTD_compressBound(n);
    unsigned char* o = (unsigned char*)std::malloc(cap ? cap : 1);
    CPSZ_REQUIRE(o, kErrAlloc, "host allocation failed");
    std::unique_ptr<unsigned char, void (*)(void*)> keep(o, &std::free);
    constexpr size_t kChunk = 8u << 20;  // a multiple of ZSTD_BLOCKSIZE_MAX
    static_assert(kChunk % ZSTD_BLOCKSIZE_MAX == 0, "chunks must hold whole blocks");
    uint8_t* h = (uint8_t*)c.host_buf("zs.all", n + 1);
    const size_t nch = n ? div_up(n, kChunk) : 1;
    std::vector<hipEvent_t> ev(nch, nullptr);
    ZSTD_CCtx* z = ZSTD_createCCtx();
    auto cleanup = [&] {
      ZSTD_freeCCtx(z);
      for (auto e : ev)
        if (e) hipEventDestroy(e);
    };
    try {
      CPSZ_REQUIRE(z, kErrAlloc, "ZSTD_createCCtx failed");
      for (size_t k = 0; k < nch; k++) {
        const size_t at = k * kChunk, len = std::min(kChunk, n - at);
        CPSZ_HIP(hipEventCreateWithFlags(&ev[k], hipEventDisableTiming));
        if (len) CPSZ_HIP(hipMemcpyAsync(h + at, d_in + at, len, hipMemcpyDeviceToHost, st));
        CPSZ_HIP(hipEventRecord(ev[k], st));
      }
      const ZSTD_parameters prm = ZSTD_getParams(level, n, 0);
      size_t r = ZSTD_compressBegin_advanced(z, nullptr, 0, prm, n);
      CPSZ_REQUIRE(!ZSTD_isError(r), kErrInternal, std::string("ZSTD: ") + ZSTD_getErrorName(r));
      size_t produced = 0;
      for (size_t k = 0; k < nch; k++) {
        CPSZ_HIP(hipEventSynchronize(ev[k]));
        const size_t at = k * kChunk, len = std::min(kChunk, n - at);
        r = k + 1 < nch ? ZSTD_compressContinue(z, o + produced, cap - produced, h + at, len)
                        : ZSTD_compressEnd(z, o + produced, cap - produced, h + at, len);
        CPSZ_REQUIRE(!ZSTD_isError(r), kErrInternal, std::string("ZSTD: ") + ZSTD_getErrorName(r));
        produced += r;
      }
      cleanup();
      *out = keep.release();
      *out_size = produced;
    } catch (...) {
      hipStreamSynchronize(st);
      cleanup();
      throw;
    }
  });
}

// Per-field ZSTD frames of a batch of device-resident streams: every field's
// bytes are queued to pinned host memory in order (one event each), and host
// threads compress fields as their bytes arrive, each into its own frame with
// the reference's one-shot call. outs[k] is malloc'd (free with cpsz_free).
int cpsz_lossless_compress_batch_dev(int device, void* stream, int level, const unsigned char* d_base,
                                     const size_t* off, const size_t* size, size_t nf, unsigned char** outs,
                                     size_t* out_sizes) {
  return guarded([&] {
    CPSZ_REQUIRE(d_base && off && size && outs && out_sizes && nf > 0, kErrArg, "null argument or empty batch");
    Ctx& c = ctx_of(device);
    StreamScope ss(c, stream);
    hipStream_t st = c.stream();
    std::vector<size_t> po(nf + 1, 0);
    for (size_t k = 0; k < nf; k++) po[k + 1] = po[k] + size[k];
    uint8_t* h = (uint8_t*)c.host_buf("zsb.all", po[nf] + 1);
    std::vector<hipEvent_t> ev(nf, nullptr);
    for (size_t k = 0; k < nf; k++) outs[k] = nullptr;
    std::atomic<size_t> next{0};
    std::atomic<int> failed{0};
    std::string first_error;
    std::mutex em;
    auto release = [&] {
      for (auto e : ev)
        if (e) hipEventDestroy(e);
    };
    try {
      for (size_t k = 0; k < nf; k++) {
        CPSZ_HIP(hipEventCreateWithFlags(&ev[k], hipEventDisableTiming));
        if (size[k]) CPSZ_HIP(hipMemcpyAsync(h + po[k], d_base + off[k], size[k], hipMemcpyDeviceToHost, st));
        CPSZ_HIP(hipEventRecord(ev[k], st));
      }
      const int nt = (int)std::min<size_t>(nf, 16);
      std::vector<std::thread> th;
      for (int t = 0; t < nt; t++)
        th.emplace_back([&] {
          ZSTD_CCtx* z = ZSTD_createCCtx();
          for (size_t k; !failed && (k = next++) < nf;) {
            const char* err = nullptr;
            if (!z) err = "ZSTD_createCCtx failed";
            else if (hipEventSynchronize(ev[k]) != hipSuccess) err = "hipEventSynchronize failed";
            unsigned char* o = nullptr;
            size_t r = 0;
            if (!err) {
              const size_t cap = ZSTD_compressBound(size[k]);
              o = (unsigned char*)std::malloc(cap ? cap : 1);
              if (!o) err = "host allocation failed";
              else {
                r = ZSTD_compressCCtx(z, o, cap, h + po[k], size[k], level);
                if (ZSTD_isError(r)) err = ZSTD_getErrorName(r);
              }
            }
            if (err) {
              std::free(o);
              std::lock_guard<std::mutex> g(em);
              if (!failed.exchange(1)) first_error = err;
              break;
            }
            outs[k] = o;
            out_sizes[k] = r;
          }
          ZSTD_freeCCtx(z);
        });
      for (auto& x : th) x.join();
      release();
    } catch (...) {
      CPSZ_HIP(hipStreamSynchronize(st));
      release();
      throw;
    }
    if (failed) {
      for (size_t k = 0; k < nf; k++) {
        std::free(outs[k]);
        outs[k] = nullptr;
      }
      fail(kErrInternal, "batched ZSTD: " + first_error);
    }
  });
}

int cpsz_lossless_decompress(const unsigned char* in, size_t n, unsigned char* out, size_t out_size) {
  return guarded([&] {
    CPSZ_REQUIRE(in && out, kErrArg, "null argument");
    const size_t r = ZSTD_decompress(out, out_size, in, n);
    CPSZ_REQUIRE(!ZSTD_isError(r), kErrStream, std::string("ZSTD_decompress: ") + ZSTD_getErrorName(r));
    CPSZ_REQUIRE(r == out_size, kErrStream, "ZSTD frame size differs from the expected size");
  });
}

int cpsz_lossless_is_frame(const unsigned char* in, size_t n) {
  if (!in) return 0;
  return ZSTD_getFrameContentSize(in, n) != ZSTD_CONTENTSIZE_ERROR ? 1 : 0;
}

int cpsz_lossless_decompress_prefix(const unsigned char* in, size_t n, unsigned char* out, size_t out_size,
                                    size_t* produced) {
  return guarded([&] {
    CPSZ_REQUIRE(in && out && produced, kErrArg, "null argument");
    ZSTD_DStream* ds = ZSTD_createDStream();
    CPSZ_REQUIRE(ds, kErrAlloc, "ZSTD_createDStream failed");
    ZSTD_inBuffer ib{in, n, 0};
    ZSTD_outBuffer ob{out, out_size, 0};
    size_t r = ZSTD_initDStream(ds);
    // stops when the prefix is full, the frame ends or the input runs out
    while (!ZSTD_isError(r) && ob.pos < ob.size) {
      const size_t before_in = ib.pos, before_out = ob.pos;
      r = ZSTD_decompressStream(ds, &ob, &ib);
      if (r == 0 || (ib.pos == before_in && ob.pos == before_out)) break;
    }
    ZSTD_freeDStream(ds);
    CPSZ_REQUIRE(!ZSTD_isError(r), kErrStream, std::string("ZSTD_decompressStream: ") + ZSTD_getErrorName(r));
    // ZSTD may use the space past the decoded bytes as scratch: clear it
    std::memset(out + ob.pos, 0, out_size - ob.pos);
    *produced = ob.pos;
  });
}

int cpsz_critical_points_2d_dev(int device, void* stream, const float* dU, const float* dV,
                                size_t r1, size_t r2, int64_t* ids, int32_t* types, size_t cap,
                                size_t* count) {
  return guarded([&] {
    Ctx& c = ctx_of(device);
    StreamScope ss(c, stream);
    critical_points_2d(c, dU, dV, r1, r2, ids, types, cap, count);
  });
}

int cpsz_critical_points_2d_noalloc_dev(int device, void* stream, const float* dU, const float* dV, size_t r1,
                                        size_t r2, int64_t* ids, int32_t* types, size_t cap, size_t* count) {
  return guarded([&] {
    critical_points_2d_noalloc(device, (hipStream_t)stream, dU, dV, r1, r2, ids, types, cap, count);
  });
}

int cpsz_critical_points_3d_dev(int device, void* stream, const float* dU, const float* dV,
                                const float* dW, size_t r1, size_t r2, size_t r3, int64_t* ids,
                                int32_t* types, size_t cap, size_t* count) {
  return guarded([&] {
    Ctx& c = ctx_of(device);
    StreamScope ss(c, stream);
    critical_points_3d(c, dU, dV, dW, r1, r2, r3, ids, types, cap, count);
  });
}

int cpsz_critical_points_2d_sos_dev(int device, void* stream, const float* dU, const float* dV, size_t r1,
                                    size_t r2, int64_t* ids, int32_t* types, size_t cap, size_t* count) {
  return guarded([&] {
    Ctx& c = ctx_of(device);
    StreamScope ss(c, stream);
    critical_points_2d(c, dU, dV, r1, r2, ids, types, cap, count, true);
  });
}

int cpsz_critical_points_3d_sos_dev(int device, void* stream, const float* dU, const float* dV, const float* dW,
                                    size_t r1, size_t r2, size_t r3, int64_t* ids, int32_t* types, size_t cap,
                                    size_t* count) {
  return guarded([&] {
    Ctx& c = ctx_of(device);
    StreamScope ss(c, stream);
    critical_points_3d(c, dU, dV, dW, r1, r2, r3, ids, types, cap, count, true);
  });
}

int cpsz_sos_det(int d, const float* rows, const int64_t* idx, int* sign) {
  return guarded([&] {
    CPSZ_REQUIRE(rows && idx && sign && (d == 2 || d == 3), kErrArg, "bad arguments");
    if (d == 2) *sign = sos_cross(rows, rows + 2, idx[0], idx[1]);
    else *sign = sos_det3(rows, rows + 3, rows + 6, idx[0], idx[1], idx[2]);
  });
}

int cpsz_eb_tables(double* q2d, int cap2d, int* n2d, double* q3d, int cap3d, int* n3d) {
  return guarded([&] {
    const EbTables& t = eb_tables();
    for (int i = 0; i < (int)t.q2d.size() && i < cap2d; i++) q2d[i] = t.q2d[i];
    for (int i = 0; i < (int)t.q3d.size() && i < cap3d; i++) q3d[i] = t.q3d[i];
    *n2d = (int)t.q2d.size();
    *n3d = (int)t.q3d.size();
  });
}

int cpsz_eb_table_2d_log(double* q, int cap, int* n) {
  return guarded([&] {
    CPSZ_REQUIRE(q && n, kErrArg, "null argument");
    const EbTables& t = eb_tables();
    for (int i = 0; i < (int)t.q2l.size() && i < cap; i++) q[i] = t.q2l[i];
    *n = (int)t.q2l.size();
  });
}

int cpsz_huffman_tree(const uint64_t* freq, size_t nbins, uint32_t state_num, uint8_t* tree,
                      size_t cap, size_t* tree_size, size_t* node_count) {
  return guarded([&] {
    HuffEncodeTable t;
    huff_build(freq, nbins, state_num, t);
    CPSZ_REQUIRE(t.tree.size() <= cap, kErrAlloc, "tree buffer too small");
    std::memcpy(tree, t.tree.data(), t.tree.size());
    *tree_size = t.tree.size();
    *node_count = t.node_count;
  });
}

int cpsz_huffman_tree_dev(int device, const uint64_t* freq, size_t nbins, uint32_t state_num, uint8_t* tree,
                          size_t cap, size_t* tree_size, size_t* node_count, uint64_t* codes, uint8_t* lens) {
  return guarded([&] {
    CPSZ_REQUIRE(freq && tree && tree_size && node_count && nbins < (1u << 24), kErrArg, "bad arguments");
    Ctx& c = ctx_of(device);
    hipStream_t st = c.stream();
    std::vector<uint32_t> sf;
    for (size_t s = 0; s < nbins; s++)
      if (freq[s]) {
        CPSZ_REQUIRE(freq[s] < (1ull << 32), kErrUnsupported, "counts must stay below 2^32");
        sf.push_back((uint32_t)s);
      }
    const uint32_t k = (uint32_t)sf.size();
    CPSZ_REQUIRE(k > 0, kErrArg, "empty histogram");
    for (uint32_t q = 0; q < k; q++) sf.push_back((uint32_t)freq[sf[q]]);
    const size_t N = 2 * (size_t)k - 1, tb = 13 * N + 1;
    uint32_t* d_sf = c.buf<uint32_t>("ht.sf", 2 * (size_t)k + 4);
    uint64_t* d_meta = c.buf<uint64_t>("ht.meta", 8);  // base (0), count, info[4]
    uint64_t* d_heap = c.buf<uint64_t>("ht.heap", k + 2);
    uint8_t* d_tree = c.buf<uint8_t>("ht.tree", tb);
    uint64_t* d_code = c.buf<uint64_t>("ht.code", nbins);
    uint8_t* d_len = c.buf<uint8_t>("ht.len", nbins);
    uint64_t meta[2] = {0, k};
    CPSZ_HIP(hipMemcpyAsync(d_sf, sf.data(), 8 * (size_t)k, hipMemcpyHostToDevice, st));
    CPSZ_HIP(hipMemcpyAsync(d_meta, meta, 16, hipMemcpyHostToDevice, st));
    TreeJob j{};
    j.syms = d_sf;
    j.freqs = d_sf + k;
    j.count = (const uint32_t*)(d_meta + 1);
    j.base = d_meta;
    j.state_num = state_num;
    j.max_leaves = k;
    j.heap = d_heap;
    j.tree = d_tree;
    j.code = d_code;
    j.len = d_len;
    j.info = d_meta + 2;
    TreeJob* d_job = c.buf<TreeJob>("ht.job", 1);
    CPSZ_HIP(hipMemcpyAsync(d_job, &j, sizeof(j), hipMemcpyHostToDevice, st));
    build_trees_device(c, d_job, 1);
    uint64_t info[4];
    CPSZ_HIP(hipMemcpyAsync(info, d_meta + 2, 32, hipMemcpyDeviceToHost, st));
    CPSZ_HIP(hipStreamSynchronize(st));
    CPSZ_REQUIRE(info[3] == 0, info[3] == 3 ? kErrUnsupported : kErrInternal,
                 "device huffman tree failed (code " + std::to_string(info[3]) + ")");
    CPSZ_REQUIRE(info[1] <= cap, kErrAlloc, "tree buffer too small");
    CPSZ_HIP(hipMemcpy(tree, d_tree, info[1], hipMemcpyDeviceToHost));
    if (codes) CPSZ_HIP(hipMemcpy(codes, d_code, 8 * nbins, hipMemcpyDeviceToHost));
    if (lens) CPSZ_HIP(hipMemcpy(lens, d_len, nbins, hipMemcpyDeviceToHost));
    *tree_size = info[1];
    *node_count = info[0];
  });
}

int cpsz_huffman_decode_host(const uint8_t* tree, size_t tree_size, size_t node_count, const uint8_t* bits,
                             size_t nbytes, size_t n, uint32_t* out) {
  return guarded([&] {
    CPSZ_REQUIRE(tree && out && (bits || !nbytes), kErrArg, "null argument");
    decode_host(tree, tree_size, node_count, bits, nbytes, n, out);
  });
}

int cpsz_prepare_libm_tables(const char* path) {
  return guarded([&] { libm_prepare(path ? std::string(path) : libm_cache_path()); });
}

int cpsz_eval_log2f(const float* x, size_t n, float* out) {
  return guarded([&] {
    const LibmView v = libm_tables().log2f.view();
    for (size_t i = 0; i < n; i++) out[i] = host_log2f(v, x[i]);
  });
}

int cpsz_eval_exp2f(const float* y, size_t n, float* out) {
  return guarded([&] {
    const LibmView v = libm_tables().exp2f.view();
    for (size_t i = 0; i < n; i++) out[i] = host_exp2f(v, y[i]);
  });
}

int cpsz_eb3d_emulate(const float* U, const float* V, const float* W, size_t r1, size_t r2,
                      size_t r3, double max_pwr_eb, double* eb) {
  return guarded([&] {
    CPSZ_REQUIRE(U && V && W && eb && r1 && r2 && r3, kErrArg, "bad arguments");
    eb3d_emulate(U, V, W, r1, r2, r3, max_pwr_eb, eb);
  });
}

int cpsz_lorenzo3d_emulate(const float* L, const uint8_t* code, size_t r1, size_t r2, size_t r3,
                           uint8_t* vid, uint16_t* q) {
  return guarded([&] {
    CPSZ_REQUIRE(L && code && vid && q && r1 && r2 && r3, kErrArg, "bad arguments");
    lorenzo3d_emulate(L, code, r1, r2, r3, vid, q);
  });
}

int cpsz_eb_code_3d(const double* eb, size_t n, uint8_t* code) {
  return guarded([&] {
    const auto& q = eb_tables().q3d;
    for (size_t i = 0; i < n; i++) code[i] = (uint8_t)eb_code_3d(eb[i], q.data(), (int)q.size());
  });
}

int cpsz_lorenzo2d_emulate(const float* x, const uint16_t* c, size_t r1, size_t r2, int wide,
                           uint16_t* q, float* recon) {
  return guarded([&] {
    CPSZ_REQUIRE(x && c && q && r1 && r2, kErrArg, "bad arguments");
    lorenzo2d_emulate(x, c, r1, r2, wide != 0, q, recon);
  });
}

int cpsz_set_timing(int device, int on) {
  return guarded([&] { ctx_of(device).timing = on != 0; });
}

int cpsz_last_stages(int device, char* names, size_t names_cap, float* ms, int cap) {
  int count = 0;
  int rc = guarded([&] {
    Ctx& c = ctx_of(device);
    std::string joined;
    for (const auto& s : c.stages) {
      if (count < cap) ms[count] = s.second;
      if (!joined.empty()) joined += ";";
      joined += s.first;
      count++;
    }
    if (names && names_cap) {
      std::strncpy(names, joined.c_str(), names_cap - 1);
      names[names_cap - 1] = 0;
    }
  });
  return rc ? rc : count;
}

int cpsz_compress_2d_batch_dev(int device, void* stream, size_t nfields, const float* const* dU,
                               const float* const* dV, const size_t* r1, const size_t* r2,
                               double max_pwr_eb, int tune, const unsigned char** d_base,
                               size_t* offsets, size_t* sizes) {
  return guarded([&] {
    CPSZ_REQUIRE(nfields && dU && dV && r1 && r2 && d_base && offsets && sizes, kErrArg, "null argument");
    Ctx& c = ctx_of(device);
    StreamScope ss(c, stream);
    std::vector<Field2d> f(nfields);
    for (size_t k = 0; k < nfields; k++) f[k] = Field2d{dU[k], dV[k], r1[k], r2[k]};
    const Batch2dConfig cfg = batch_2d_config(c, f, max_pwr_eb, tune != 0);
    BatchOut o;
    compress_2d_batch(c, f, max_pwr_eb, cfg, o);
    *d_base = o.base;
    for (size_t k = 0; k < nfields; k++) {
      offsets[k] = o.off[k];
      sizes[k] = o.size[k];
    }
  });
}

int cpsz_decompress_2d_batch_dev(int device, void* stream, size_t nfields,
                                 const unsigned char* const* d_in, const size_t* in_size,
                                 const size_t* r1, const size_t* r2, float* const* dU,
                                 float* const* dV) {
  return guarded([&] {
    CPSZ_REQUIRE(nfields && d_in && in_size && r1 && r2 && dU && dV, kErrArg, "null argument");
    Ctx& c = ctx_of(device);
    StreamScope ss(c, stream);
    std::vector<Field2d> sig(nfields);
    std::vector<Field2dOut> f(nfields);
    std::vector<const uint8_t*> in(nfields);
    std::vector<size_t> sz(nfields);
    for (size_t k = 0; k < nfields; k++) {
      sig[k] = Field2d{nullptr, nullptr, r1[k], r2[k]};
      f[k] = Field2dOut{dU[k], dV[k], r1[k], r2[k]};
      in[k] = d_in[k];
      sz[k] = in_size[k];
    }
    decompress_2d_batch(c, in, sz, f, batch_2d_config(c, sig, 0.0, false));
  });
}

int cpsz_batch_2d_config(int device, size_t nfields, const size_t* r1, const size_t* r2, int* workers,
                         int* grid_cap, int* device_trees) {
  return guarded([&] {
    CPSZ_REQUIRE(nfields && r1 && r2 && workers && grid_cap, kErrArg, "null argument");
    Ctx& c = ctx_of(device);
    std::vector<Field2d> sig(nfields);
    for (size_t k = 0; k < nfields; k++) sig[k] = Field2d{nullptr, nullptr, r1[k], r2[k]};
    const Batch2dConfig cfg = batch_2d_config(c, sig, 0.0, false);
    *workers = cfg.workers;
    *grid_cap = cfg.grid_cap;
    if (device_trees) *device_trees = cfg.device_trees ? 1 : 0;
  });
}

int cpsz_batch_2d_set_config(int device, size_t nfields, const size_t* r1, const size_t* r2, int workers,
                             int grid_cap, int device_trees) {
  return guarded([&] {
    CPSZ_REQUIRE(nfields && r1 && r2 && workers >= 1 && workers <= 64 && grid_cap >= 1, kErrArg, "bad arguments");
    int cur = 0;
    CPSZ_HIP(hipGetDevice(&cur));
    CPSZ_HIP(hipSetDevice(device));
    std::vector<Field2d> sig(nfields);
    for (size_t k = 0; k < nfields; k++) sig[k] = Field2d{nullptr, nullptr, r1[k], r2[k]};
    Batch2dConfig cfg;
    cfg.workers = workers;
    cfg.grid_cap = grid_cap;
    cfg.device_trees = device_trees != 0;
    batch_2d_set_config(sig, cfg);
    CPSZ_HIP(hipSetDevice(cur));
  });
}

int cpsz_compress_3d_batch_dev(int device, void* stream, size_t nfields, const float* const* dU,
                               const float* const* dV, const float* const* dW, const size_t* r1,
                               const size_t* r2, const size_t* r3, double max_pwr_eb, int tune,
                               const unsigned char** d_base, size_t* offsets, size_t* sizes) {
  return guarded([&] {
    CPSZ_REQUIRE(nfields && dU && dV && dW && r1 && r2 && r3 && d_base && offsets && sizes, kErrArg,
                 "null argument");
    Ctx& c = ctx_of(device);
    StreamScope ss(c, stream);
    std::vector<Field3d> f(nfields);
    for (size_t k = 0; k < nfields; k++) f[k] = Field3d{dU[k], dV[k], dW[k], r1[k], r2[k], r3[k]};
    const Batch2dConfig cfg = batch_3d_config(c, f, max_pwr_eb, tune != 0);
    BatchOut o;
    compress_3d_batch(c, f, max_pwr_eb, cfg, o);
    *d_base = o.base;
    for (size_t k = 0; k < nfields; k++) {
      offsets[k] = o.off[k];
      sizes[k] = o.size[k];
    }
  });
}

int cpsz_decompress_3d_batch_dev(int device, void* stream, size_t nfields,
                                 const unsigned char* const* d_in, const size_t* in_size,
                                 const size_t* r1, const size_t* r2, const size_t* r3, float* const* dU,
                                 float* const* dV, float* const* dW) {
  return guarded([&] {
    CPSZ_REQUIRE(nfields && d_in && in_size && r1 && r2 && r3 && dU && dV && dW, kErrArg, "null argument");
    Ctx& c = ctx_of(device);
    StreamScope ss(c, stream);
    std::vector<Field3d> sig(nfields);
    std::vector<Field3dOut> f(nfields);
    std::vector<const uint8_t*> in(nfields);
    std::vector<size_t> sz(nfields);
    for (size_t k = 0; k < nfields; k++) {
      sig[k] = Field3d{nullptr, nullptr, nullptr, r1[k], r2[k], r3[k]};
      f[k] = Field3dOut{dU[k], dV[k], dW[k], r1[k], r2[k], r3[k]};
      in[k] = d_in[k];
      sz[k] = in_size[k];
    }
    decompress_3d_batch(c, in, sz, f, batch_3d_config(c, sig, 0.0, false));
  });
}

int cpsz_batch_3d_config(int device, size_t nfields, const size_t* r1, const size_t* r2, const size_t* r3,
                         int* workers, int* grid_cap) {
  return guarded([&] {
    CPSZ_REQUIRE(nfields && r1 && r2 && r3 && workers && grid_cap, kErrArg, "null argument");
    Ctx& c = ctx_of(device);
    std::vector<Field3d> sig(nfields);
    for (size_t k = 0; k < nfields; k++) sig[k] = Field3d{nullptr, nullptr, nullptr, r1[k], r2[k], r3[k]};
    const Batch2dConfig cfg = batch_3d_config(c, sig, 0.0, false);
    *workers = cfg.workers;
    *grid_cap = cfg.grid_cap;
  });
}

size_t cpsz_learn_param_count(void) { return kNumParams; }

int cpsz_learn_init_params(uint64_t seed, float* params) {
  return guarded([&] {
    CPSZ_REQUIRE(params, kErrArg, "null params");
    mlp_init_params(seed, params);
  });
}

int cpsz_learn_analytic_eb_2d_dev(int device, void* stream, const float* dU, const float* dV, size_t r1,
                                  size_t r2, double max_pwr_eb, double* dE) {
  return guarded([&] {
    Ctx& c = ctx_of(device);
    StreamScope ss(c, stream);
    analytic_eb_2d(c, dU, dV, r1, r2, max_pwr_eb, dE);
  });
}

int cpsz_learn_features_2d_dev(int device, void* stream, const float* dU, const float* dV,
                               const double* dE, size_t r1, size_t r2, double max_pwr_eb,
                               const uint32_t* d_idx, size_t m, uint16_t* dX) {
  return guarded([&] {
    Ctx& c = ctx_of(device);
    StreamScope ss(c, stream);
    features_2d(c, dU, dV, dE, r1, r2, max_pwr_eb, d_idx, m, dX);
  });
}

int cpsz_learn_labels_2d_dev(int device, void* stream, const float* dU, const float* dV,
                             const double* dE, size_t r1, size_t r2, double max_pwr_eb,
                             const uint32_t* d_idx, size_t m, uint8_t* dL) {
  return guarded([&] {
    Ctx& c = ctx_of(device);
    StreamScope ss(c, stream);
    labels_2d(c, dU, dV, dE, r1, r2, max_pwr_eb, d_idx, m, dL);
  });
}

int cpsz_gemm_bf16_dev(int device, void* stream, int M, int N, int K, const uint16_t* A, int64_t sam,
                       int64_t sak, const uint16_t* B, int64_t sbk, int64_t sbn, float* C, int ldc,
                       int split) {
  return guarded([&] {
    CPSZ_REQUIRE(A && B && C, kErrArg, "null operand");
    Ctx& c = ctx_of(device);
    StreamScope ss(c, stream);
    if (split > 1) CPSZ_HIP(hipMemsetAsync(C, 0, (size_t)M * ldc * sizeof(float), c.stream()));
    gemm_bf16(c, M, N, K, A, (long)sam, (long)sak, B, (long)sbk, (long)sbn, C, ldc, split);
  });
}

int cpsz_mlp_forward_dev(int device, void* stream, const float* params, const uint16_t* dX, size_t m,
                         float* dY) {
  return guarded([&] {
    CPSZ_REQUIRE(params && dX && dY, kErrArg, "null argument");
    Ctx& c = ctx_of(device);
    StreamScope ss(c, stream);
    float* P = c.buf<float>("api.p", kNumParams);
    CPSZ_HIP(hipMemcpyAsync(P, params, kNumParams * 4, hipMemcpyHostToDevice, c.stream()));
    mlp_forward(c, P, dX, m, dY);
    CPSZ_HIP(hipStreamSynchronize(c.stream()));
  });
}

int cpsz_mlp_grad_dev(int device, void* stream, const float* params, const uint16_t* dX,
                      const float* dT, size_t m, float* grad, float* loss) {
  return guarded([&] {
    CPSZ_REQUIRE(params && dX && dT && grad && loss, kErrArg, "null argument");
    Ctx& c = ctx_of(device);
    StreamScope ss(c, stream);
    float* P = c.buf<float>("api.p", kNumParams);
    float* G = c.buf<float>("api.g", kNumParams + 1);
    CPSZ_HIP(hipMemcpyAsync(P, params, kNumParams * 4, hipMemcpyHostToDevice, c.stream()));
    mlp_grad(c, P, dX, dT, m, G, G + kNumParams);
    CPSZ_HIP(hipMemcpyAsync(grad, G, kNumParams * 4, hipMemcpyDeviceToHost, c.stream()));
    CPSZ_HIP(hipMemcpyAsync(loss, G + kNumParams, 4, hipMemcpyDeviceToHost, c.stream()));
    CPSZ_HIP(hipStreamSynchronize(c.stream()));
  });
}

int cpsz_learn_train_2d_dev(int device, void* stream, size_t nfields, const float* const* dU,
                            const float* const* dV, const size_t* r1, const size_t* r2,
                            double max_pwr_eb, size_t samples_per_field, int epochs, int batch,
                            float lr, uint64_t seed, float* params, float* loss_per_epoch) {
  return guarded([&] {
    CPSZ_REQUIRE(nfields && dU && dV && r1 && r2 && params, kErrArg, "null argument");
    Ctx& c = ctx_of(device);
    StreamScope ss(c, stream);
    std::vector<TrainField> f(nfields);
    for (size_t k = 0; k < nfields; k++) f[k] = TrainField{dU[k], dV[k], r1[k], r2[k]};
    train_2d(c, f, max_pwr_eb, samples_per_field, epochs, batch, lr, seed, params, loss_per_epoch);
  });
}

int cpsz_compress_2d_learned_f32_dev(int device, void* stream, const float* dU, const float* dV,
                                     size_t r1, size_t r2, double max_pwr_eb, const float* params,
                                     const unsigned char** d_out, size_t* out_size, int* iterations,
                                     size_t* boosted, size_t* fallback) {
  return guarded([&] {
    CPSZ_REQUIRE(d_out && out_size, kErrArg, "null output argument");
    Ctx& c = ctx_of(device);
    StreamScope ss(c, stream);
    DeviceStream s;
    LearnedStats st;
    compress_2d_learned(c, dU, dV, r1, r2, max_pwr_eb, params, s, st);
    *d_out = s.data;
    *out_size = s.size;
    if (iterations) *iterations = st.iterations;
    if (boosted) *boosted = st.boosted_vertices;
    if (fallback) *fallback = st.fallback_vertices;
  });
}

int cpsz_compress_2d_learned_f32(const float* U, const float* V, size_t r1, size_t r2,
                                 double max_pwr_eb, const float* params, unsigned char** out,
                                 size_t* out_size) {
  return guarded([&] {
    CPSZ_REQUIRE(U && V && params && out && out_size && r1 && r2, kErrArg, "null argument or empty field");
    Ctx& c = ctx_of(current_device());
    const size_t n = r1 * r2;
    float* d = c.buf<float>("api.lin", 2 * n);
    CPSZ_HIP(hipMemcpyAsync(d, U, n * 4, hipMemcpyHostToDevice, c.stream()));
    CPSZ_HIP(hipMemcpyAsync(d + n, V, n * 4, hipMemcpyHostToDevice, c.stream()));
    DeviceStream s;
    LearnedStats st;
    compress_2d_learned(c, d, d + n, r1, r2, max_pwr_eb, params, s, st);
    unsigned char* h = (unsigned char*)std::malloc(s.size ? s.size : 1);
    CPSZ_REQUIRE(h, kErrAlloc, "host allocation failed");
    CPSZ_HIP(hipMemcpyAsync(h, s.data, s.size, hipMemcpyDeviceToHost, c.stream()));
    CPSZ_HIP(hipStreamSynchronize(c.stream()));
    *out = h;
    *out_size = s.size;
  });
}

int cpsz_compress_2d_learned_batch_dev(int device, void* stream, size_t nfields, const float* const* dU,
                                       const float* const* dV, const size_t* r1, const size_t* r2,
                                       double max_pwr_eb, const float* params, int tune,
                                       const unsigned char** d_base, size_t* offsets, size_t* sizes,
                                       int* iterations, size_t* boosted, size_t* fallback) {
  return guarded([&] {
    CPSZ_REQUIRE(nfields && dU && dV && r1 && r2 && params && d_base && offsets && sizes, kErrArg,
                 "null argument");
    Ctx& c = ctx_of(device);
    StreamScope ss(c, stream);
    std::vector<TrainField> f(nfields);
    std::vector<Field2d> sig(nfields);
    for (size_t k = 0; k < nfields; k++) {
      f[k] = TrainField{dU[k], dV[k], r1[k], r2[k]};
      sig[k] = Field2d{dU[k], dV[k], r1[k], r2[k]};
    }
    const Batch2dConfig cfg = batch_2d_config(c, sig, max_pwr_eb, tune != 0);
    const uint8_t* base = nullptr;
    std::vector<size_t> off, size;
    LearnedStats st;
    compress_2d_learned_batch(c, f, max_pwr_eb, params, cfg.workers, cfg.grid_cap, base, off, size, st);
    *d_base = base;
    for (size_t k = 0; k < nfields; k++) {
      offsets[k] = off[k];
      sizes[k] = size[k];
    }
    if (iterations) *iterations = st.iterations;
    if (boosted) *boosted = st.boosted_vertices;
    if (fallback) *fallback = st.fallback_vertices;
  });
}

int cpsz_learn_infer_gains_dev(int device, void* stream, int dim, const float* dU, const float* dV,
                               const float* dW, const double* dE, size_t r1, size_t r2, size_t r3,
                               double max_pwr_eb, const float* params, size_t v0, size_t m, uint8_t* d_gain) {
  return guarded([&] {
    CPSZ_REQUIRE(params && (dim == 2 || dim == 3), kErrArg, "bad arguments");
    const size_t n = r1 * r2 * (dim == 3 ? r3 : 1);
    CPSZ_REQUIRE(v0 <= n && m <= n - v0, kErrArg, "vertex range outside the field");
    Ctx& c = ctx_of(device);
    StreamScope ss(c, stream);
    const size_t np = dim == 2 ? kNumParams : kNumParams3;
    float* P = c.buf<float>("api.infer.p", np);
    CPSZ_HIP(hipMemcpyAsync(P, params, np * 4, hipMemcpyHostToDevice, c.stream()));
    mlp_infer_gains(c, dim, dU, dV, dW, dE, r1, r2, dim == 3 ? r3 : 1, max_pwr_eb, P, v0, m, d_gain);
    CPSZ_HIP(hipStreamSynchronize(c.stream()));  // the parameter upload reads the caller's host buffer
  });
}

// ---- learned proposer, 3D
size_t cpsz_learn_param_count_3d(void) { return kNumParams3; }

int cpsz_learn_init_params_3d(uint64_t seed, float* params) {
  return guarded([&] {
    CPSZ_REQUIRE(params, kErrArg, "null params");
    mlp_init_params(seed, params, kFeatDim3, kFeatUsed3);
  });
}

int cpsz_check_fast_bounds(int device, uint32_t* max_ulp, double* max_ratio) {
  return guarded([&] {
    CPSZ_REQUIRE(max_ulp && max_ratio, kErrArg, "null argument");
    Ctx& c = ctx_of(device);
    check_fast_bounds(c, max_ulp, max_ratio);
  });
}

namespace {
// correctly rounded square root and reciprocal moved by up to g_ulps units in
// the last place, from a hash of a running count
thread_local uint64_t g_perturb_n = 0;
thread_local int g_ulps = 0;
float perturb(float x) {
  if (!g_ulps || !(x > 0) || !std::isfinite(x)) return x;
  uint64_t h = (g_perturb_n++ + 1) * 0x9E3779B97F4A7C15ull;
  h ^= h >> 31;
  const int d = (int)(h % (uint64_t)(2 * g_ulps + 1)) - g_ulps;
  uint32_t u;
  std::memcpy(&u, &x, 4);
  u += (uint32_t)d;
  std::memcpy(&x, &u, 4);
  return x;
}
struct PerturbedRoots {
  static float sqrt(float x) { return perturb(std::sqrt(x)); }
  static float rcp(float x) { return perturb(1.0f / x); }
};
}  // namespace

int cpsz_tri_bound_2d_fast_host(const float* uv6, size_t n, int ulps, float* entry) {
  return guarded([&] {
    CPSZ_REQUIRE(uv6 && entry && ulps >= 0 && ulps < 64, kErrArg, "bad argument");
    g_ulps = ulps;
    g_perturb_n = 0;
    for (size_t i = 0; i < n; i++) {
      const float* t = uv6 + 6 * i;
      entry[i] = tri_fast<PerturbedRoots>(t[0], t[1], t[2], t[3], t[4], t[5], false);
    }
    g_ulps = 0;
  });
}

int cpsz_ratio_ge1_host(const double* det, const double* d, size_t n, unsigned char* out) {
  return guarded([&] {
    CPSZ_REQUIRE(det && d && out, kErrArg, "null argument");
    for (size_t i = 0; i < n; i++) out[i] = ratio_ge1(det[i], d[i]) ? 1 : 0;
  });
}

int cpsz_set_exact_2d_bounds(int on) {
  return guarded([&] { set_exact_2d_bounds(on != 0); });
}

int cpsz_learn_analytic_eb_3d_dev(int device, void* stream, const float* dU, const float* dV, const float* dW,
                                  size_t r1, size_t r2, size_t r3, double max_pwr_eb, double* dE) {
  return guarded([&] {
    Ctx& c = ctx_of(device);
    StreamScope ss(c, stream);
    analytic_eb_3d(c, dU, dV, dW, r1, r2, r3, max_pwr_eb, dE);
  });
}

int cpsz_learn_features_3d_dev(int device, void* stream, const float* dU, const float* dV, const float* dW,
                               const double* dE, size_t r1, size_t r2, size_t r3, double max_pwr_eb,
                               const uint32_t* d_idx, size_t m, uint16_t* dX) {
  return guarded([&] {
    Ctx& c = ctx_of(device);
    StreamScope ss(c, stream);
    features_3d(c, dU, dV, dW, dE, r1, r2, r3, max_pwr_eb, d_idx, m, dX);
  });
}

int cpsz_learn_labels_3d_dev(int device, void* stream, const float* dU, const float* dV, const float* dW,
                             const double* dE, size_t r1, size_t r2, size_t r3, double max_pwr_eb,
                             const uint32_t* d_idx, size_t m, uint8_t* dL) {
  return guarded([&] {
    Ctx& c = ctx_of(device);
    StreamScope ss(c, stream);
    labels_3d(c, dU, dV, dW, dE, r1, r2, r3, max_pwr_eb, d_idx, m, dL);
  });
}

int cpsz_mlp_forward_3d_dev(int device, void* stream, const float* params, const uint16_t* dX, size_t m,
                            float* dY) {
  return guarded([&] {
    CPSZ_REQUIRE(params && dX && dY, kErrArg, "null argument");
    Ctx& c = ctx_of(device);
    StreamScope ss(c, stream);
    float* P = c.buf<float>("api.p3", kNumParams3);
    CPSZ_HIP(hipMemcpyAsync(P, params, kNumParams3 * 4, hipMemcpyHostToDevice, c.stream()));
    mlp_forward(c, P, dX, m, dY, kFeatDim3);
    CPSZ_HIP(hipStreamSynchronize(c.stream()));
  });
}

int cpsz_mlp_grad_3d_dev(int device, void* stream, const float* params, const uint16_t* dX, const float* dT,
                         size_t m, float* grad, float* loss) {
  return guarded([&] {
    CPSZ_REQUIRE(params && dX && dT && grad && loss, kErrArg, "null argument");
    Ctx& c = ctx_of(device);
    StreamScope ss(c, stream);
    float* P = c.buf<float>("api.p3", kNumParams3);
    float* G = c.buf<float>("api.g3", kNumParams3 + 1);
    CPSZ_HIP(hipMemcpyAsync(P, params, kNumParams3 * 4, hipMemcpyHostToDevice, c.stream()));
    mlp_grad(c, P, dX, dT, m, G, G + kNumParams3, kFeatDim3);
    CPSZ_HIP(hipMemcpyAsync(grad, G, kNumParams3 * 4, hipMemcpyDeviceToHost, c.stream()));
    CPSZ_HIP(hipMemcpyAsync(loss, G + kNumParams3, 4, hipMemcpyDeviceToHost, c.stream()));
    CPSZ_HIP(hipStreamSynchronize(c.stream()));
  });
}

int cpsz_learn_train_3d_dev(int device, void* stream, size_t nfields, const float* const* dU,
                            const float* const* dV, const float* const* dW, const size_t* r1, const size_t* r2,
                            const size_t* r3, double max_pwr_eb, size_t samples_per_field, int epochs, int batch,
                            float lr, uint64_t seed, float* params, float* loss_per_epoch) {
  return guarded([&] {
    CPSZ_REQUIRE(nfields && dU && dV && dW && r1 && r2 && r3 && params, kErrArg, "null argument");
    Ctx& c = ctx_of(device);
    StreamScope ss(c, stream);
    std::vector<TrainField3> f(nfields);
    for (size_t k = 0; k < nfields; k++) f[k] = TrainField3{dU[k], dV[k], dW[k], r1[k], r2[k], r3[k]};
    train_3d(c, f, max_pwr_eb, samples_per_field, epochs, batch, lr, seed, params, loss_per_epoch);
  });
}

int cpsz_compress_3d_learned_batch_dev(int device, void* stream, size_t nfields, const float* const* dU,
                                       const float* const* dV, const float* const* dW, const size_t* r1,
                                       const size_t* r2, const size_t* r3, double max_pwr_eb,
                                       const float* params, int tune, const unsigned char** d_base,
                                       size_t* offsets, size_t* sizes, int* iterations, size_t* boosted,
                                       size_t* fallback) {
  return guarded([&] {
    CPSZ_REQUIRE(nfields && dU && dV && dW && r1 && r2 && r3 && params && d_base && offsets && sizes, kErrArg,
                 "null argument");
    Ctx& c = ctx_of(device);
    StreamScope ss(c, stream);
    std::vector<TrainField3> f(nfields);
    std::vector<Field3d> sig(nfields);
    for (size_t k = 0; k < nfields; k++) {
      f[k] = TrainField3{dU[k], dV[k], dW[k], r1[k], r2[k], r3[k]};
      sig[k] = Field3d{dU[k], dV[k], dW[k], r1[k], r2[k], r3[k]};
    }
    const Batch2dConfig cfg = batch_3d_config(c, sig, max_pwr_eb, tune != 0);
    const uint8_t* base = nullptr;
    std::vector<size_t> off, size;
    LearnedStats st;
    compress_3d_learned_batch(c, f, max_pwr_eb, params, cfg.workers, cfg.grid_cap, base, off, size, st);
    *d_base = base;
    for (size_t k = 0; k < nfields; k++) {
      offsets[k] = off[k];
      sizes[k] = size[k];
    }
    if (iterations) *iterations = st.iterations;
    if (boosted) *boosted = st.boosted_vertices;
    if (fallback) *fallback = st.fallback_vertices;
  });
}

int cpsz_tet_cp_emulate(const float* f, int t, int* type) {
  return guarded([&] {
    CPSZ_REQUIRE(f && type && t >= 0 && t < 6, kErrArg, "bad argument");
    float g[4][3];
    for (int p = 0; p < 4; p++)
      for (int q = 0; q < 3; q++) g[p][q] = f[3 * p + q];
    *type = tet_cp(g, judge_tet_a(t), judge_tet_b(t));
  });
}

}  // extern "C"
