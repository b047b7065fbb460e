// One element of the 2D Lorenzo codec, shared by the wavefront kernels and
// the host emulation used to test the arithmetic on the CPU.
//
// Compressor step (include/sz_compression_utils.hpp:50-65 with
// pred = (L + U) - UL in float, include/sz_prediction.hpp:10-14):
//   diff = x - pred (float, widened); quant_diff = |diff|/E + 1 (double);
//   unpredictable unless quant_diff < 65536 and |x' - x| <= E, where
//   x' = (float)(pred + 2 q E) and q = +-(int)(quant_diff / 2).
// With E = 2^(2 id - 52) a power of two the FP32 form below is exact:
//   |diff|/E + 1 < 65536   <=>  a = |diff| * 2^(52 - 2 id) < 65535
//   (int)(quant_diff / 2)  ==   (floor(a) + 1) >> 1
//   (float)(pred + 2qE)    ==   pred + q * 2^(2 id - 51) in FP32, because
//       2qE is an exact float (|2q| <= 2^16, id <= 81) and a double sum of two
//       floats rounded to float equals the single FP32 rounding (the double
//       sum is exact when the exponents differ by <= 28; otherwise the smaller
//       term is below 1/32 ulp of the larger and both roundings return it).
// Eb code c: 0 = zero precision (always unpredictable), c >= 1: id = c - 1.
// `shift` is log2 of the id base: 2 for the offline codec (E = 4^id eps), 1
// for offline_log (E = 2^id eps, ids <= 62, never the wide path); the
// argument above holds for both, since E stays a power of two.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace cpsz {

template <bool WIDE>
__host__ __device__ inline float l2d_compress_step(float pred, float x, uint32_t c, uint16_t& code, int shift = 2) {
  if (!WIDE) {
    // branch-free: code 0 is computed with id 0 and then rejected
    const int id = c ? (int)c - 1 : 0;
    const int pe = shift * id - 52;  // log2 E
    const float diff = x - pred;
    const float a = ldexpf(fabsf(diff), -pe);
    const float qf = floorf(fmaf(floorf(a), 0.5f, 0.5f));
    const float qs = copysignf(qf, diff);
    const float rec = pred + ldexpf(qs, pe + 1);
    const bool ok = (c != 0) & (a < 65535.0f) & (fabsf(rec - x) <= ldexpf(1.0f, pe));
    code = ok ? (uint16_t)((int)qs + 32768) : 0;
    return ok ? rec : x;
  } else {
    if (c == 0) {
      code = 0;
      return x;
    }
    const int id = (int)c - 1;
    const double prec = ldexp(1.0, 2 * id - 52);
    const double diff = x - pred;
    double qd = fabs(diff) / prec + 1;
    if (qd < 65536) {
      qd = (diff > 0) ? qd : -qd;
      const int qi = (int)(qd / 2) + 32768;
      const float dec = (float)((double)pred + 2 * (qi - 32768) * prec);
      if (fabs((double)(dec - x)) <= prec) {
        code = (uint16_t)qi;
        return dec;
      }
    }
    code = 0;
    return x;
  }
}

// Decompressor step (include/sz_decompression_utils.hpp:63-73): q == 0 takes
// the next unpredictable value; otherwise (float)(pred + 2(q - 32768) eb) with
// eb = 0 for stream id 0 and 2^(2 sid - 52) otherwise
// (src/sz_decompress_cp_preserve_2d.cpp:30-33).
template <bool WIDE>
__host__ __device__ inline float l2d_decompress_step(float pred, uint32_t q, uint32_t sid, float raw,
                                                     int shift = 2) {
  if (!WIDE) {
    const float kE = sid ? ldexpf((float)(2 * ((int)q - 32768)), shift * (int)sid - 52) : 0.f;
    return q == 0 ? raw : pred + kE;
  }
  if (q == 0) return raw;
  const double prec = sid ? ldexp(1.0, 2 * (int)sid - 52) : 0.0;
  return (float)((double)pred + 2 * ((int)q - 32768) * prec);
}

}  // namespace cpsz
