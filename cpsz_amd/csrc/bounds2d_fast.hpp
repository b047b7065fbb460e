// The FP32-approximated 2D triangle bound for float fields, shared by the
// bounds kernel and its host check.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>

namespace cpsz {

__host__ __device__ inline uint32_t f32_bits(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  return u;
}

// The reference's bound is a function of the six float products: it sums
// them in double (det, the p/n sums of the sign bounds), divides det by three
// of those sums, and takes |sqrt p - sqrt n| / (sqrt p + sqrt n) with
// correctly rounded double square roots and divisions. Here the sums and the
// branch decisions stay exact, and only the last step is approximated, in
// FP32, as t ~ |p - n| / (p + n + 2 sqrt(p n)) from the exact p, n (relative
// error below 2^-20). The reference's rounded value lies within 5 * 2^-53 of
// the real t (absolute; t <= 1), so the true bound is enclosed in
// [a (1 - 2^-16) - 2^-50, a (1 + 2^-16) + 2^-50] for an approximation a; a
// vertex whose eb codes that enclosure does not fix takes the exact bound.
//
// A triangle's entry in the cell table (FP32):
//   0            the exact bound 0
//   kExact1      the exact bound 1
//   +inf         no triangle (cell outside the field)
//   kNeedExact   only the FP64 bound will do (critical-point triangles,
//                products outside [2^-60, 2^58])
//   otherwise    an approximation a in (2^-56, 1 + 2^-19]
constexpr float kExact1 = 2.0f;
constexpr float kNeedExact = -1.0f;

// products in [2^-60, 2^58] (or 0) keep every FP32 step below in the normal range
__host__ __device__ inline bool prod_ok(float f) {
  const uint32_t e = f32_bits(f) & 0x7FFFFFFFu;
  return e == 0 || (e >= 0x21800000u && e <= 0x5C800000u);
}
// a vertex value in [2^-30, 2^29] (or 0): products of two such values pass prod_ok
__host__ __device__ inline bool value_ok(float f) {
  const uint32_t e = f32_bits(f) & 0x7FFFFFFFu;
  return e == 0 || (e >= 0x30800000u && e <= 0x4E000000u);
}

// |sqrt p - sqrt n| / (sqrt p + sqrt n) = |p - n| / (p + n + 2 sqrt(p n)); num = |p - n|.
// R supplies the square root and reciprocal (on the device the 1-ulp
// hardware instructions; the enclosure below leaves room for several ulp).
template <class R>
__host__ __device__ inline float approx_sign_bound(float p, float n, float num) {
  const float s = R::sqrt(p * n);
  return num * R::rcp((p + n) + 2.0f * s);
}

// max_eb_to_keep_sign_2d_offline_2(a, b): 1 unless the terms have opposite
// signs, then p, n are the terms' magnitudes and p - n = a + b
template <class R>
__host__ __device__ inline float fbound2(float a, float b) {
  const bool mixed = (a > 0 && b < 0) || (a < 0 && b > 0);
  const float r = approx_sign_bound<R>(fabsf(a), fabsf(b), fabsf(a + b));
  return mixed ? r : kExact1;
}

// max_eb_to_keep_sign_2d_offline_4 from its p and n (or both doubled), summed
// in the reference's order by the caller (a zero part adds nothing to a
// non-negative sum); products stay below 2^58, so doubled sums below 2^61
template <class R>
__host__ __device__ inline float fbound4(double p, double n) {
  const float r = approx_sign_bound<R>((float)p, (float)n, (float)fabs(p - n));
  return (p == 0 || n == 0) ? kExact1 : r;
}

// (det / d >= 1) in double: the rounded quotient reaches 1 exactly when
// det / d >= 1 - 2^-54, and det - d (exact in sign) tells which side of 1 the
// quotient is on; only quotients within 2^-50 below 1 (or d == 0) divide
__host__ __device__ inline bool ratio_ge1(double det, double d) {
  const double t = det - d;
  const bool ge = t == 0 || ((t > 0) == (d > 0));
  if (d == 0 || (!ge && fabs(t) <= 0x1p-50 * fabs(d))) return det / d >= 1;
  return ge;
}

// The cell table entry of one triangle (vertex order as dtri's). checked: the
// tile's values all pass value_ok, so the products need no check.
template <class R>
__host__ __device__ inline float tri_fast(float u0, float u1, float u2, float v0, float v1, float v2, bool checked) {
  const float f01 = u0 * v1, f10 = u1 * v0, f02 = u0 * v2, f20 = u2 * v0, f12 = u1 * v2, f21 = u2 * v1;
  const double u0v1 = f01, u1v0 = f10, u0v2 = f02, u2v0 = f20, u1v2 = f12, u2v1 = f21;
  const double det = u0v1 - u1v0 + u1v2 - u2v1 + u2v0 - u0v2;
  if (!(det != 0)) return 0.0f;
  if (!checked && !(prod_ok(f01) && prod_ok(f10) && prod_ok(f02) && prod_ok(f20) && prod_ok(f12) && prod_ok(f21)))
    return kNeedExact;
  const bool f1 = ratio_ge1(det, u2v0 - u0v2);
  const bool f2 = ratio_ge1(det, u1v2 - u2v1);
  const bool f3 = ratio_ge1(det, u0v1 - u1v0);
  if (f1 && f2 && f3) return kNeedExact;
  // twice the positive and negative parts of the signed terms u0v1, -u1v0 |
  // u1v2, -u2v1 | u2v0, -u0v2: t + |t| and |t| - t are exact, and sums of
  // doubled terms are the doubled sums bit for bit; fbound4 is scale-free
  const double p0 = u0v1 + fabs(u0v1), n0 = fabs(u0v1) - u0v1, p1 = fabs(u1v0) - u1v0, n1 = u1v0 + fabs(u1v0);
  const double p2 = u1v2 + fabs(u1v2), n2 = fabs(u1v2) - u1v2, p3 = fabs(u2v1) - u2v1, n3 = u2v1 + fabs(u2v1);
  const double p4 = u2v0 + fabs(u2v0), n4 = fabs(u2v0) - u2v0, p5 = fabs(u0v2) - u0v2, n5 = u0v2 + fabs(u0v2);
  const double pa = p0 + p1, na = n0 + n1, pb = p2 + p3, nb = n2 + n3;
  const float m1 = fminf(fbound2<R>(f20, -f02), fbound4<R>((pa + p2) + p3, (na + n2) + n3));
  const float m2 = fminf(fbound2<R>(f12, -f21), fbound4<R>((pa + p4) + p5, (na + n4) + n5));
  const float m3 = fminf(fbound2<R>(f01, -f10), fbound4<R>((pb + p4) + p5, (nb + n4) + n5));
  return fmaxf(fmaxf(f1 ? 0.0f : m1, f2 ? 0.0f : m2), f3 ? 0.0f : m3);
}

// The enclosure of the reference's bound around an approximation a > 0.
__host__ __device__ inline double enclosure_lo(double a) { return a * (1 - 0x1p-16) - 0x1p-50; }
__host__ __device__ inline double enclosure_hi(double a) { return a * (1 + 0x1p-16) + 0x1p-50; }

}  // namespace cpsz
