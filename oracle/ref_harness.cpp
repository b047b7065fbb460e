// extern "C" wrappers around the reference library's public API, compiled
// together with the reference sources by oracle/Makefile into oracle/_ref/.
// TEST INFRASTRUCTURE: used only to produce golden vectors and as the timed
// CPU baseline. The reference prints progress chatter to stdout; every call
// here silences stdout for its duration so callers see clean output.
#include <fcntl.h>
#include <unistd.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "sz_compress_3d.hpp"
#include "sz_compress_cp_preserve_2d.hpp"
#include "sz_compression_utils.hpp"
#include "sz_decompress_3d.hpp"
#include "sz_decompress_cp_preserve_2d.hpp"
#include "sz_decompress_cp_preserve_3d.hpp"
#include "sz_decompression_utils.hpp"
#include "sz_lossless.hpp"
#include "sz_cp_preserve_utils.hpp"

namespace {
struct Quiet {
  int saved = -1;
  Quiet() {
    std::fflush(stdout);
    saved = dup(1);
    int dn = open("/dev/null", O_WRONLY);
    if (dn >= 0) {
      dup2(dn, 1);
      close(dn);
    }
  }
  ~Quiet() {
    std::fflush(stdout);
    if (saved >= 0) {
      dup2(saved, 1);
      close(saved);
    }
  }
};
}  // namespace

extern "C" {

void ref_free(void* p) { std::free(p); }

// sz_compress_cp_preserve_2d_offline<float> (include/sz_compress_cp_preserve_2d.hpp:8-10)
unsigned char* ref_compress_2d_f32(const float* U, const float* V, size_t r1, size_t r2, double eb,
                                   size_t* size) {
  Quiet q;
  size_t s = 0;
  unsigned char* out = sz_compress_cp_preserve_2d_offline<float>(U, V, r1, r2, s, false, eb);
  *size = s;
  return out;
}
unsigned char* ref_compress_2d_f64(const double* U, const double* V, size_t r1, size_t r2,
                                   double eb, size_t* size) {
  Quiet q;
  size_t s = 0;
  unsigned char* out = sz_compress_cp_preserve_2d_offline<double>(U, V, r1, r2, s, false, eb);
  *size = s;
  return out;
}
// sz_decompress_cp_preserve_2d_offline (include/sz_decompress_cp_preserve_2d.hpp:8-11)
int ref_decompress_2d_f32(const unsigned char* s, size_t r1, size_t r2, float* U, float* V) {
  Quiet q;
  float* u = nullptr;
  float* v = nullptr;
  sz_decompress_cp_preserve_2d_offline<float>(s, r1, r2, u, v);
  std::memcpy(U, u, r1 * r2 * sizeof(float));
  std::memcpy(V, v, r1 * r2 * sizeof(float));
  std::free(u);
  std::free(v);
  return 0;
}
int ref_decompress_2d_f64(const unsigned char* s, size_t r1, size_t r2, double* U, double* V) {
  Quiet q;
  double* u = nullptr;
  double* v = nullptr;
  sz_decompress_cp_preserve_2d_offline<double>(s, r1, r2, u, v);
  std::memcpy(U, u, r1 * r2 * sizeof(double));
  std::memcpy(V, v, r1 * r2 * sizeof(double));
  std::free(u);
  std::free(v);
  return 0;
}
// sz_decompress_cp_preserve_3d_online_log (include/sz_decompress_cp_preserve_3d.hpp:8-11)
int ref_decompress_3d_f32(const unsigned char* s, size_t r1, size_t r2, size_t r3, float* U,
                          float* V, float* W) {
  Quiet q;
  float *u = nullptr, *v = nullptr, *w = nullptr;
  sz_decompress_cp_preserve_3d_online_log<float>(s, r1, r2, r3, u, v, w);
  const size_t n = r1 * r2 * r3;
  std::memcpy(U, u, n * 4);
  std::memcpy(V, v, n * 4);
  std::memcpy(W, w, n * 4);
  std::free(u);
  std::free(v);
  std::free(w);
  return 0;
}
int ref_decompress_3d_f64(const unsigned char* s, size_t r1, size_t r2, size_t r3, double* U,
                          double* V, double* W) {
  Quiet q;
  double *u = nullptr, *v = nullptr, *w = nullptr;
  sz_decompress_cp_preserve_3d_online_log<double>(s, r1, r2, r3, u, v, w);
  const size_t n = r1 * r2 * r3;
  std::memcpy(U, u, n * 8);
  std::memcpy(V, v, n * 8);
  std::memcpy(W, w, n * 8);
  std::free(u);
  std::free(v);
  std::free(w);
  return 0;
}
// Huffman_encode_tree_and_data (src/sz_compression_utils.cpp:61-82); buf must
// hold the section plus the encoder's 8-byte overrun.
size_t ref_huff_encode(size_t state_num, const int* syms, size_t n, unsigned char* buf) {
  Quiet q;
  unsigned char* p = buf;
  Huffman_encode_tree_and_data(state_num, syms, n, p);
  return (size_t)(p - buf);
}
// Huffman_decode_tree_and_data (src/sz_decompression_utils.cpp:92-108)
size_t ref_huff_decode(size_t state_num, const unsigned char* buf, size_t n, int* out) {
  Quiet q;
  const unsigned char* p = buf;
  int* t = Huffman_decode_tree_and_data(state_num, n, p);
  std::memcpy(out, t, n * sizeof(int));
  std::free(t);
  return (size_t)(p - buf);
}
// sz_compress_2d_with_eb<float> (src/sz_compress_3d.cpp:196-236)
unsigned char* ref_lorenzo_2d_f32(const float* data, const double* prec, size_t r1, size_t r2,
                                  size_t* size) {
  Quiet q;
  size_t s = 0;
  unsigned char* out = sz_compress_2d_with_eb<float>(data, prec, r1, r2, s);
  *size = s;
  return out;
}
// sz_compress/decompress_cp_preserve_2d_offline_log<float>
// (src/sz_compress_cp_preserve_2d.cpp:285-374, src/sz_decompress_cp_preserve_2d.cpp:54-99)
unsigned char* ref_compress_2d_log_f32(const float* U, const float* V, size_t r1, size_t r2, double eb,
                                       size_t* size) {
  size_t s = 0;
  unsigned char* out = sz_compress_cp_preserve_2d_offline_log<float>(U, V, r1, r2, s, false, eb);
  *size = s;
  return out;
}
int ref_decompress_2d_log_f32(const unsigned char* s, size_t r1, size_t r2, float* U, float* V) {
  float *u = nullptr, *v = nullptr;
  sz_decompress_cp_preserve_2d_offline_log<float>(s, r1, r2, u, v);
  if (!u || !v) return -1;
  std::memcpy(U, u, 4 * r1 * r2);
  std::memcpy(V, v, 4 * r1 * r2);
  std::free(u);
  std::free(v);
  return 0;
}

// sz_lossless_compress / _decompress (src/sz_lossless.cpp:11-45), ZSTD.
size_t ref_zstd_compress(int level, unsigned char* data, size_t len, unsigned char** out) {
  return sz_lossless_compress(ZSTD_COMPRESSOR, level, data, len, out);
}
size_t ref_zstd_decompress(unsigned char* data, size_t len, unsigned char** out, size_t target) {
  return sz_lossless_decompress(ZSTD_COMPRESSOR, data, len, out, target);
}
// is_lossless_compressed_data / sz_lossless_decompress65536bytes (src/sz_lossless.cpp:3-9,47-62)
int ref_is_lossless(unsigned char* data, size_t len) { return is_lossless_compressed_data(data, len); }
size_t ref_zstd_decompress65536(unsigned char* data, size_t len, unsigned char** out) {
  return sz_lossless_decompress65536bytes(ZSTD_COMPRESSOR, data, len, out);
}

// ---- header-only helpers of the 3D offline_log compressor -------------------
// (include/sz_cp_preserve_utils.hpp). The compressor's own source does not build
// here (src/sz_compress_cp_preserve_3d.cpp includes FTK), but these helpers it
// calls compile from the reference header alone.

// acbrt1 (:173-184) for one value.
float ref_acbrt1_one(float x) { return acbrt1(x); }

// Compares acbrt1 (:173-184) with `f` on every float bit pattern in [lo, hi)
// using `threads` threads; returns the number of inputs whose results differ
// bit for bit and the smallest such pattern in *first (hi when none).
uint64_t ref_acbrt1_compare(float (*f)(float), uint32_t lo, uint32_t hi, int threads, uint32_t* first) {
  if (threads < 1) threads = 1;
  std::atomic<uint64_t> bad{0};
  std::atomic<uint32_t> lowest{hi};
  std::vector<std::thread> pool;
  const uint64_t span = (uint64_t)hi - lo;
  for (int t = 0; t < threads; t++)
    pool.emplace_back([&, t] {
      const uint32_t a = (uint32_t)(lo + span * t / threads), b = (uint32_t)(lo + span * (t + 1) / threads);
      uint64_t nb = 0;
      uint32_t lw = hi;
      for (uint32_t bits = a; bits < b; bits++) {
        float x, r0, r1;
        std::memcpy(&x, &bits, 4);
        r0 = acbrt1(x);
        r1 = f(x);
        if (std::memcmp(&r0, &r1, 4) != 0) {
          if (!nb) lw = bits;
          nb++;
        }
      }
      bad += nb;
      uint32_t cur = lowest.load();
      while (lw < cur && !lowest.compare_exchange_weak(cur, lw)) {
      }
    });
  for (auto& th : pool) th.join();
  *first = lowest.load();
  return bad.load();
}

// max_eb_to_keep_sign_3d_offline (:186-191) elementwise.
void ref_sign_bound_3d(const float* pos, const float* neg, size_t n, float* out) {
  for (size_t i = 0; i < n; i++) out[i] = max_eb_to_keep_sign_3d_offline(pos[i], neg[i]);
}

// same_direction (:52-57) over groups of four values.
void ref_same_direction(const float* q4, size_t n, int* out) {
  for (size_t i = 0; i < n; i++) out[i] = same_direction(q4[4 * i], q4[4 * i + 1], q4[4 * i + 2], q4[4 * i + 3]);
}

// eb_exponential_quantize (:14-22) elementwise.
void ref_eb_exp_quantize(const double* eb, size_t n, int base, double log_of_base, double threshold, int* ids,
                         double* eb_out) {
  for (size_t i = 0; i < n; i++) {
    double e = eb[i];
    ids[i] = eb_exponential_quantize(e, base, log_of_base, threshold);
    eb_out[i] = e;
  }
}

// log_transform<float> (:205-221).
void ref_log_transform_f32(const float* x, size_t n, float* logv, unsigned char* sign) {
  float* l = log_transform(x, sign, n);
  std::memcpy(logv, l, n * sizeof(float));
  std::free(l);
}

}  // extern "C"
