// Drop-in C++ entry points for the cpSZ plain codec on MI355X.
//
// Same names, template signatures, default arguments and buffer ownership as
// the reference headers, so a caller of the reference compiles unchanged
// against this header and links libcpsz_hip.so instead of the reference
// library:
//   sz_compress_cp_preserve_2d_offline     include/sz_compress_cp_preserve_2d.hpp:8-10
//   sz_decompress_cp_preserve_2d_offline   include/sz_decompress_cp_preserve_2d.hpp:8-11
//   sz_compress_cp_preserve_2d_offline_log include/sz_compress_cp_preserve_2d.hpp:12-14
//   sz_decompress_cp_preserve_2d_offline_log include/sz_decompress_cp_preserve_2d.hpp:13-15
//   sz_compress_cp_preserve_3d_offline_log include/sz_compress_cp_preserve_3d.hpp:6-8
//   sz_decompress_cp_preserve_3d_online_log include/sz_decompress_cp_preserve_3d.hpp:8-11
//   ZSTD_COMPRESSOR, is_lossless_compressed_data, sz_lossless_compress,
//   sz_lossless_decompress, sz_lossless_decompress65536bytes
//                                          include/sz_lossless.hpp:9-15
// Streams are byte-compatible with the reference in both directions.
//
// Differences a caller can observe:
//  * the 2D pair has both instantiations (float and double); the 3D pair is
//    float only: its double instantiations are deleted (a compile error where
//    used), because the reference's double 3D path applies the C library's
//    double log2 and exp2, whose rounding the device does not reproduce;
//  * failures (no GPU, non-finite input, corrupt stream) are reported as
//    nullptr / compressed_size = 0 / null outputs instead of the reference's
//    exit(0) or silent garbage; cpsz_last_error() gives the message;
//  * `transpose` is ignored, as in the reference.
#ifndef SZ_CP_PRESERVE_DROPIN_HPP_
#define SZ_CP_PRESERVE_DROPIN_HPP_

#include <cstddef>
#include <cstdlib>

#include "cpsz.h"

template <typename T>
unsigned char* sz_compress_cp_preserve_2d_offline(const T* U, const T* V, size_t r1, size_t r2,
                                                  size_t& compressed_size, bool transpose = false,
                                                  double max_pwr_eb = 0.1);
template <typename T>
void sz_decompress_cp_preserve_2d_offline(const unsigned char* compressed, size_t r1, size_t r2,
                                          T*& U, T*& V);
template <typename T>
unsigned char* sz_compress_cp_preserve_2d_offline_log(const T* U, const T* V, size_t r1, size_t r2,
                                                      size_t& compressed_size, bool transpose = false,
                                                      double max_pwr_eb = 0.1);
template <typename T>
void sz_decompress_cp_preserve_2d_offline_log(const unsigned char* compressed, size_t r1, size_t r2, T*& U,
                                              T*& V);
template <typename T>
unsigned char* sz_compress_cp_preserve_3d_offline_log(const T* U, const T* V, const T* W, size_t r1,
                                                      size_t r2, size_t r3, size_t& compressed_size,
                                                      bool transpose = false, double max_pwr_eb = 0.1);
template <typename T>
void sz_decompress_cp_preserve_3d_online_log(const unsigned char* compressed, size_t r1, size_t r2,
                                             size_t r3, T*& U, T*& V, T*& W);

namespace cpsz_dropin {
template <typename T>
inline void release(T*& p) {
  std::free(p);
  p = nullptr;
}
template <typename T>
inline T* fresh(size_t n) {
  return static_cast<T*>(std::malloc((n ? n : 1) * sizeof(T)));
}
}  // namespace cpsz_dropin

template <>
inline unsigned char* sz_compress_cp_preserve_2d_offline<float>(const float* U, const float* V, size_t r1,
                                                                size_t r2, size_t& compressed_size, bool,
                                                                double max_pwr_eb) {
  unsigned char* out = nullptr;
  size_t size = 0;
  if (cpsz_compress_2d_f32(U, V, r1, r2, max_pwr_eb, &out, &size) != CPSZ_OK) {
    compressed_size = 0;
    return nullptr;
  }
  compressed_size = size;
  return out;
}

template <>
inline unsigned char* sz_compress_cp_preserve_2d_offline<double>(const double* U, const double* V, size_t r1,
                                                                 size_t r2, size_t& compressed_size, bool,
                                                                 double max_pwr_eb) {
  unsigned char* out = nullptr;
  size_t size = 0;
  if (cpsz_compress_2d_f64(U, V, r1, r2, max_pwr_eb, &out, &size) != CPSZ_OK) {
    compressed_size = 0;
    return nullptr;
  }
  compressed_size = size;
  return out;
}

// The reference frees incoming non-null outputs and returns malloc'd ones
// (src/sz_decompress_cp_preserve_2d.cpp:11-12).
template <>
inline void sz_decompress_cp_preserve_2d_offline<float>(const unsigned char* compressed, size_t r1,
                                                        size_t r2, float*& U, float*& V) {
  cpsz_dropin::release(U);
  cpsz_dropin::release(V);
  size_t size = 0;
  if (cpsz_stream_size_2d(compressed, &size) != CPSZ_OK) return;
  float* u = cpsz_dropin::fresh<float>(r1 * r2);
  float* v = cpsz_dropin::fresh<float>(r1 * r2);
  if (!u || !v || cpsz_decompress_2d_f32(compressed, size, r1, r2, u, v) != CPSZ_OK) {
    std::free(u);
    std::free(v);
    return;
  }
  U = u;
  V = v;
}

template <>
inline void sz_decompress_cp_preserve_2d_offline<double>(const unsigned char* compressed, size_t r1, size_t r2,
                                                         double*& U, double*& V) {
  cpsz_dropin::release(U);
  cpsz_dropin::release(V);
  size_t size = 0;
  if (cpsz_stream_size_2d(compressed, &size) != CPSZ_OK) return;
  double* u = cpsz_dropin::fresh<double>(r1 * r2);
  double* v = cpsz_dropin::fresh<double>(r1 * r2);
  if (!u || !v || cpsz_decompress_2d_f64(compressed, size, r1, r2, u, v) != CPSZ_OK) {
    std::free(u);
    std::free(v);
    return;
  }
  U = u;
  V = v;
}

template <>
inline unsigned char* sz_compress_cp_preserve_2d_offline_log<float>(const float* U, const float* V, size_t r1,
                                                                    size_t r2, size_t& compressed_size, bool,
                                                                    double max_pwr_eb) {
  unsigned char* out = nullptr;
  size_t size = 0;
  if (cpsz_compress_2d_log_f32(U, V, r1, r2, max_pwr_eb, &out, &size) != CPSZ_OK) {
    compressed_size = 0;
    return nullptr;
  }
  compressed_size = size;
  return out;
}

// src/sz_decompress_cp_preserve_2d.cpp:57-58 frees incoming outputs; results are malloc'd.
template <>
inline void sz_decompress_cp_preserve_2d_offline_log<float>(const unsigned char* compressed, size_t r1,
                                                            size_t r2, float*& U, float*& V) {
  cpsz_dropin::release(U);
  cpsz_dropin::release(V);
  size_t size = 0;
  if (cpsz_stream_size_2d_log(compressed, r1, r2, &size) != CPSZ_OK) return;
  float* u = cpsz_dropin::fresh<float>(r1 * r2);
  float* v = cpsz_dropin::fresh<float>(r1 * r2);
  if (!u || !v || cpsz_decompress_2d_log_f32(compressed, size, r1, r2, u, v) != CPSZ_OK) {
    std::free(u);
    std::free(v);
    return;
  }
  U = u;
  V = v;
}

template <>
inline unsigned char* sz_compress_cp_preserve_3d_offline_log<float>(const float* U, const float* V,
                                                                    const float* W, size_t r1, size_t r2,
                                                                    size_t r3, size_t& compressed_size,
                                                                    bool, double max_pwr_eb) {
  unsigned char* out = nullptr;
  size_t size = 0;
  if (cpsz_compress_3d_f32(U, V, W, r1, r2, r3, max_pwr_eb, &out, &size) != CPSZ_OK) {
    compressed_size = 0;
    return nullptr;
  }
  compressed_size = size;
  return out;
}

// The double 3D pair is deleted: a caller that instantiates it fails to
// compile instead of receiving streams whose bits differ from the reference's.
template <>
unsigned char* sz_compress_cp_preserve_3d_offline_log<double>(const double*, const double*, const double*, size_t,
                                                              size_t, size_t, size_t& compressed_size, bool,
                                                              double) = delete;

// src/sz_decompress_cp_preserve_3d.cpp:10-12 frees and reallocates U, V, W.
template <>
inline void sz_decompress_cp_preserve_3d_online_log<float>(const unsigned char* compressed, size_t r1,
                                                           size_t r2, size_t r3, float*& U, float*& V,
                                                           float*& W) {
  cpsz_dropin::release(U);
  cpsz_dropin::release(V);
  cpsz_dropin::release(W);
  size_t size = 0;
  if (cpsz_stream_size_3d(compressed, r1, r2, r3, &size) != CPSZ_OK) return;
  const size_t n = r1 * r2 * r3;
  float* u = cpsz_dropin::fresh<float>(n);
  float* v = cpsz_dropin::fresh<float>(n);
  float* w = cpsz_dropin::fresh<float>(n);
  if (!u || !v || !w || cpsz_decompress_3d_f32(compressed, size, r1, r2, r3, u, v, w) != CPSZ_OK) {
    std::free(u);
    std::free(v);
    std::free(w);
    return;
  }
  U = u;
  V = v;
  W = w;
}

template <>
void sz_decompress_cp_preserve_3d_online_log<double>(const unsigned char*, size_t, size_t, size_t, double*& U,
                                                     double*& V, double*& W) = delete;

// Host ZSTD stage with the reference's names and signatures
// (include/sz_lossless.hpp:9-15, src/sz_lossless.cpp:3-62). The outputs are
// malloc'd, as in the reference. An unknown compressor id returns 0 and leaves
// the output pointer untouched, as the reference's default branch does; a
// ZSTD failure returns 0 with a null output instead of an unchecked buffer.
#ifndef ZSTD_COMPRESSOR
#define ZSTD_COMPRESSOR 1
#endif

inline int is_lossless_compressed_data(unsigned char* compressedBytes, size_t cmpSize) {
  return cpsz_lossless_is_frame(compressedBytes, cmpSize) ? ZSTD_COMPRESSOR : -1;
}

inline unsigned long sz_lossless_compress(int losslessCompressor, int level, unsigned char* data,
                                          unsigned long dataLength, unsigned char** compressBytes) {
  if (losslessCompressor != ZSTD_COMPRESSOR) return 0;
  size_t out = 0;
  *compressBytes = nullptr;
  if (cpsz_lossless_compress(level, data, dataLength, compressBytes, &out) != CPSZ_OK) return 0;
  return (unsigned long)out;
}

inline unsigned long sz_lossless_decompress(int losslessCompressor, unsigned char* compressBytes,
                                            unsigned long cmpSize, unsigned char** oriData,
                                            unsigned long targetOriSize) {
  if (losslessCompressor != ZSTD_COMPRESSOR) return 0;
  *oriData = static_cast<unsigned char*>(std::malloc(targetOriSize ? targetOriSize : 1));
  if (!*oriData) return 0;
  if (cpsz_lossless_decompress(compressBytes, cmpSize, *oriData, targetOriSize) != CPSZ_OK) {
    std::free(*oriData);
    *oriData = nullptr;
    return 0;
  }
  return targetOriSize;
}

// The first 65536 bytes of the frame, zero-filled past its end (the reference
// reads stream headers this way).
inline unsigned long sz_lossless_decompress65536bytes(int losslessCompressor, unsigned char* compressBytes,
                                                      unsigned long cmpSize, unsigned char** oriData) {
  if (losslessCompressor != ZSTD_COMPRESSOR) return 0;
  *oriData = static_cast<unsigned char*>(std::calloc(65536, 1));
  if (!*oriData) return 0;
  size_t got = 0;
  if (cpsz_lossless_decompress_prefix(compressBytes, cmpSize, *oriData, 65536, &got) != CPSZ_OK) {
    std::free(*oriData);
    *oriData = nullptr;
    return 0;
  }
  return 65536;
}

#endif  // SZ_CP_PRESERVE_DROPIN_HPP_
