// extern "C" wrappers around the reference library's public API, compiled
// together with the reference sources by oracle/Makefile into oracle/_ref/.
// TEST INFRASTRUCTURE: used only to produce golden vectors and as the timed
// CPU baseline. The reference prints progress chatter to stdout; every call
// here silences stdout for its duration so callers see clean output.
#include <fcntl.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "sz_compress_3d.hpp"
#include "sz_compress_cp_preserve_2d.hpp"
#include "sz_compression_utils.hpp"
#include "sz_decompress_3d.hpp"
#include "sz_decompress_cp_preserve_2d.hpp"
#include "sz_decompress_cp_preserve_3d.hpp"
#include "sz_decompression_utils.hpp"
#include "sz_lossless.hpp"

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
// sz_lossless_compress / _decompress (src/sz_lossless.cpp:11-45), ZSTD.
size_t ref_zstd_compress(int level, unsigned char* data, size_t len, unsigned char** out) {
  return sz_lossless_compress(ZSTD_COMPRESSOR, level, data, len, out);
}
size_t ref_zstd_decompress(unsigned char* data, size_t len, unsigned char** out, size_t target) {
  return sz_lossless_decompress(ZSTD_COMPRESSOR, data, len, out, target);
}

}  // extern "C"
