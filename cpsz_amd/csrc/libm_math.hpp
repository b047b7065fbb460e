// Host-libm-exact log2f / exp2f for the device.
//
// The 3D codec applies the host C library's log2f to every value
// (include/sz_cp_preserve_utils.hpp:205-221) and exp2f to every decoded log
// value (src/sz_decompress_cp_preserve_3d.cpp:97-102). Those functions are not
// correctly rounded, so the device reproduces them as:
//   1. y = log2 / exp2 evaluated in double with a fixed sequence of IEEE
//      operations (identical on host and device, no libm);
//   2. f = (float)y;
//   3. if y lies within the input binade's "band" of a float rounding
//      midpoint, consult that binade's exception table (built once on the host
//      by an exhaustive sweep comparing f with the host libm): a bitmap whose
//      hits take the float on the other side of y, or a sorted (x, result) list.
// Outside the band the host libm result equals f for every input of the binade
// (established by the exhaustive sweep), so the lookup is skipped.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>

namespace cpsz {

__host__ __device__ inline uint64_t dbits(double x) {
  uint64_t b;
  __builtin_memcpy(&b, &x, 8);
  return b;
}
__host__ __device__ inline double dfrom(uint64_t b) {
  double x;
  __builtin_memcpy(&x, &b, 8);
  return x;
}
__host__ __device__ inline uint32_t fbits(float x) {
  uint32_t b;
  __builtin_memcpy(&b, &x, 4);
  return b;
}
__host__ __device__ inline float ffrom(uint32_t b) {
  float x;
  __builtin_memcpy(&x, &b, 4);
  return x;
}

// log2(x) for finite x > 0: x = m 2^e, m in [sqrt(1/2), sqrt(2)),
// ln m = 2 atanh(s) with s = (m-1)/(m+1), series to s^23.
__host__ __device__ inline double log2_dd(float xf) {
  const double x = (double)xf;  // exact; float subnormals become normal doubles
  const uint64_t b = dbits(x);
  int e = (int)((b >> 52) & 0x7FF) - 1023;
  double m = dfrom((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
  if (m > 1.4142135623730951) {
    m = m * 0.5;
    e += 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double z = s * s;
  double p = 1.0 / 23.0;
  p = 1.0 / 21.0 + z * p;
  p = 1.0 / 19.0 + z * p;
  p = 1.0 / 17.0 + z * p;
  p = 1.0 / 15.0 + z * p;
  p = 1.0 / 13.0 + z * p;
  p = 1.0 / 11.0 + z * p;
  p = 1.0 / 9.0 + z * p;
  p = 1.0 / 7.0 + z * p;
  p = 1.0 / 5.0 + z * p;
  p = 1.0 / 3.0 + z * p;
  const double s2 = s + s;
  const double lnm = s2 + s2 * (z * p);
  return lnm * 1.4426950408889634 + (double)e;
}

// exp2(y) for finite y: y = k + f with k = floor(y + 1/2), f in [-1/2, 1/2],
// exp(f ln 2) by Taylor series to degree 15, scaled by 2^k.
__host__ __device__ inline double exp2_dd(float yf) {
  const double y = (double)yf;
  const double kd = floor(y + 0.5);
  const double f = y - kd;
  const double r = f * 0.6931471805599453;
  double p = 1.0 / 1307674368000.0;  // 1/15!
  p = 1.0 / 87178291200.0 + r * p;
  p = 1.0 / 6227020800.0 + r * p;
  p = 1.0 / 479001600.0 + r * p;
  p = 1.0 / 39916800.0 + r * p;
  p = 1.0 / 3628800.0 + r * p;
  p = 1.0 / 362880.0 + r * p;
  p = 1.0 / 40320.0 + r * p;
  p = 1.0 / 5040.0 + r * p;
  p = 1.0 / 720.0 + r * p;
  p = 1.0 / 120.0 + r * p;
  p = 1.0 / 24.0 + r * p;
  p = 1.0 / 6.0 + r * p;
  p = 0.5 + r * p;
  p = 1.0 + r * p;
  p = 1.0 + r * p;
  int k = (int)kd;
  if (k > 1100) k = 1100;
  if (k < -1100) k = -1100;
  return ldexp(p, k);
}

// Float neighbour of f on the side of y (toward +inf when y == f).
__host__ __device__ inline float side_neighbour(double y, float f) {
  const uint32_t b = fbits(f);
  const bool up = y >= (double)f;
  const bool neg = (b >> 31) != 0;
  if ((b & 0x7FFFFFFFu) == 0) return up ? ffrom(1u) : ffrom(0x80000001u);
  return ffrom((up != neg) ? b + 1 : b - 1);
}

// Distance of y from the rounding midpoint between f = (float)y and its
// neighbour on y's side, in units of that float spacing: 0 at the midpoint,
// 0.5 when y is exactly f.
__host__ __device__ inline float midpoint_distance(double y, float f) {
  const float nb = side_neighbour(y, f);
  const double fd = (double)f, nd = (double)nb;
  const double mid = (fd + nd) * 0.5;
  const double ulp = fabs(nd - fd);
  if (!(ulp > 0 && ulp <= 1.0e308)) return 0.5f;
  // ulp is the spacing of adjacent floats, a power of two in [2^-149, 2^104]:
  // multiplying by its reciprocal (built from the exponent) equals the division
  // exactly, both being exact scalings by a power of two
  const double inv = dfrom((uint64_t)(2046 - (int)(dbits(ulp) >> 52)) << 52);
  return (float)(fabs(y - mid) * inv);
}

// Exception tables for one function, laid out per input binade ("slot").
struct LibmView {
  const float* band;       // per slot: -1 = no exceptions, else max midpoint distance
  const uint8_t* kind;     // per slot: 0 none, 1 sorted list, 2 bitmap (other-side rule)
  const uint32_t* off;     // per slot: offset into list arrays / bitmap words
  const uint32_t* cnt;     // per slot: list length
  const uint32_t* list_x;  // sorted input bit patterns
  const uint32_t* list_y;  // host-libm results
  const uint32_t* bitmap;  // 2^23 bits per bitmap slot (indexed by mantissa)
};

__host__ __device__ inline float libm_fix(const LibmView& v, uint32_t slot, uint32_t xb, double y) {
  const float f = (float)y;
  const float band = v.band[slot];
  if (band < 0.f || midpoint_distance(y, f) > band) return f;
  const uint8_t kind = v.kind[slot];
  if (kind == 2) {
    const uint32_t m = xb & 0x7FFFFFu;
    const uint32_t w = v.bitmap[v.off[slot] + (m >> 5)];
    return ((w >> (m & 31)) & 1u) ? side_neighbour(y, f) : f;
  }
  if (kind == 1) {
    uint32_t lo = v.off[slot], hi = lo + v.cnt[slot];
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      const uint32_t a = v.list_x[mid];
      if (a == xb) return ffrom(v.list_y[mid]);
      if (a < xb) lo = mid + 1;
      else hi = mid;
    }
  }
  return f;
}

// log2f as the host libm computes it, for finite x > 0.
__host__ __device__ inline float host_log2f(const LibmView& v, float x) {
  const uint32_t xb = fbits(x);
  return libm_fix(v, xb >> 23, xb, log2_dd(x));
}

// exp2f as the host libm computes it, for finite y in [-99, 128).
__host__ __device__ inline float host_exp2f(const LibmView& v, float y) {
  const uint32_t yb = fbits(y);
  return libm_fix(v, yb >> 23, yb, exp2_dd(y));  // slot = sign and exponent (9 bits)
}

}  // namespace cpsz
