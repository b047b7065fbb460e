// 3D offline_log arithmetic shared by the device kernels and host emulations.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>

namespace cpsz {

// acbrt1 (include/sz_cp_preserve_utils.hpp:173-184): integer first guess and
// two Newton steps, all FP32 with a correctly rounded division.
__host__ __device__ inline float acbrt1_f(float x0) {
  int32_t ix;
  __builtin_memcpy(&ix, &x0, 4);
  ix = ix / 4 + ix / 16;
  ix = ix + ix / 16;
  ix = ix + ix / 256;
  ix = 0x2a5137a0 + ix;
  float x;
  __builtin_memcpy(&x, &ix, 4);
  x = 0.33333333f * (2.0f * x + x0 / (x * x));
  x = 0.33333333f * (2.0f * x + x0 / (x * x));
  return x;
}

// acbrt1_f for arguments with a clear sign bit (sums of magnitudes): the
// integer steps on a non-negative bit pattern are exact as unsigned shifts,
// which cost one instruction where signed division by 4, 16, 256 costs three.
__host__ __device__ inline float acbrt1_nonneg_f(float x0) {
  uint32_t ix;
  __builtin_memcpy(&ix, &x0, 4);
  ix = (ix >> 2) + (ix >> 4);
  ix = ix + (ix >> 4);
  ix = ix + (ix >> 8);
  ix = 0x2a5137a0u + ix;
  float x;
  __builtin_memcpy(&x, &ix, 4);
  x = 0.33333333f * (2.0f * x + x0 / (x * x));
  x = 0.33333333f * (2.0f * x + x0 / (x * x));
  return x;
}

// max_eb_to_keep_sign_3d_offline (:186-191); p and n are sums of magnitudes
// (never negative, never -0: the sums start from +0)
__host__ __device__ inline float cube_sign_bound(float p, float n) {
  if (p == 0 && n == 0) return 0;
  const float P = acbrt1_nonneg_f(p), N = acbrt1_nonneg_f(n);
  return fabsf(P - N) / (P + N);
}

// ---- approximate bounds with a proven error envelope
//
// The five correctly rounded divisions of cube_sign_bound cost about ten
// instructions each; the approximate form replaces every a / b with
// a * rcp(b) (hardware reciprocal, 1 ulp). For arguments in [0, 2^126):
//  * acbrt1_fast is within 2 ulp of acbrt1_nonneg_f (checked exhaustively
//    over every float of that range, subnormals included, on the device), so
//    its relative error is at most eps = 2^-22;
//  * with P' = P(1+a), N' = N(1+b), |a|, |b| <= eps = 4u, u = 2^-24, every
//    float operation rounded to nearest (relative error <= u) and the
//    reciprocal within 2u: writing E = |P-N|/(P+N), the exact bound lies in
//    [E(1-u)^2/(1+u), E(1+u)^2/(1-u)] and the fast one in
//    [(E-eps)(1-u)^2(1-2u)/((1+eps)(1+u)), (E+eps)(1+u)^2(1+2u)/((1-eps)(1-u))],
//    so |fast - exact| <= 12.01u e + 4.0001u, below e * kEbRel + kEbAbs.
// Minima and maxima of values that each obey such an envelope obey it too
// (the maps e -> e (1 -/+ R) -/+ A are increasing), so a vertex bound F'
// assembled from approximate tetrahedron bounds satisfies
// F in [(F' - A)(1 - R), (F' + A)(1 + R)]. Vertices whose id is not constant
// over that interval are recomputed exactly.
constexpr double kEbAbs = 0x1.2p-22;  // A = 4.5u
constexpr double kEbRel = 0x1p-20;    // R = 16u
// Sign-bound arguments are the eight per-cofactor sums or sums of three of
// them: the eight below 2^124 keeps every argument below 2^126.
constexpr uint32_t kSafeBits = 0x7D800000u;  // bit pattern of 2^124

__host__ __device__ inline float rcp_approx(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}

__host__ __device__ inline float acbrt1_fast(float x0) {
  uint32_t ix;
  __builtin_memcpy(&ix, &x0, 4);
  ix = (ix >> 2) + (ix >> 4);
  ix = ix + (ix >> 4);
  ix = ix + (ix >> 8);
  ix = 0x2a5137a0u + ix;
  float x;
  __builtin_memcpy(&x, &ix, 4);
  x = 0.33333333f * (2.0f * x + x0 * rcp_approx(x * x));
  x = 0.33333333f * (2.0f * x + x0 * rcp_approx(x * x));
  return x;
}

// cube_sign_bound within the envelope above, for p, n in [0, 2^126).
__host__ __device__ inline float cube_sign_bound_fast(float p, float n) {
  if (p == 0 && n == 0) return 0;
  const float P = acbrt1_fast(p), N = acbrt1_fast(n);
  return fabsf(P - N) * rcp_approx(P + N);
}

__host__ __device__ inline void acc3(double v, double& p, double& n) {
  if (v >= 0) p += v;
  else n += -v;
}

__host__ __device__ inline int sgn3(float v) { return (0.f < v) - (v < 0.f); }
__host__ __device__ inline bool same_dir(float a, float b, float c, float d) {
  const int s = sgn3(a);
  return s != 0 && s == sgn3(b) && s == sgn3(c) && s == sgn3(d);
}

// max_eb_to_keep_position_and_type_3d_offline
// (src/sz_compress_cp_preserve_3d.cpp:264-358); vertex 3 is the cell origin.
// FAST evaluates the sign bounds with cube_sign_bound_fast and returns 0 (which
// marks every vertex of the tetrahedron for exact recomputation) when an
// argument falls outside that function's range.
template <bool FAST>
__host__ __device__ inline double tet_bound_t(float u0, float u1, float u2, float u3, float v0, float v1,
                                            float v2, float v3, float w0, float w1, float w2,
                                            float w3) {
  const double u3v1w2 = u3 * v1 * w2, u3v2w1 = u3 * v2 * w1, u3v2w0 = u3 * v2 * w0,
               u3v0w2 = u3 * v0 * w2, u3v0w1 = u3 * v0 * w1, u3v1w0 = u3 * v1 * w0;
  const double u1v3w2 = u1 * v3 * w2, u2v3w1 = u2 * v3 * w1, u2v3w0 = u2 * v3 * w0,
               u0v3w2 = u0 * v3 * w2, u0v3w1 = u0 * v3 * w1, u1v3w0 = u1 * v3 * w0;
  const double u1v2w3 = u1 * v2 * w3, u2v1w3 = u2 * v1 * w3, u0v2w3 = u0 * v2 * w3,
               u2v0w3 = u2 * v0 * w3, u0v1w3 = u0 * v1 * w3, u1v0w3 = u1 * v0 * w3;
  const double u0v1w2 = u0 * v1 * w2, u0v2w1 = u0 * v2 * w1, u1v2w0 = u1 * v2 * w0,
               u1v0w2 = u1 * v0 * w2, u2v0w1 = u2 * v0 * w1, u2v1w0 = u2 * v1 * w0;
  const double u3_0 = -u3v1w2 + u3v2w1, u3_1 = -u3v2w0 + u3v0w2, u3_2 = -u3v0w1 + u3v1w0;
  const double v3_0 = u1v3w2 - u2v3w1, v3_1 = u2v3w0 - u0v3w2, v3_2 = u0v3w1 - u1v3w0;
  const double w3_0 = -u1v2w3 + u2v1w3, w3_1 = u0v2w3 - u2v0w3, w3_2 = -u0v1w3 + u1v0w3;
  const double c_4 = u0v1w2 - u0v2w1 + u1v2w0 - u1v0w2 + u2v0w1 - u2v1w0;
  const double M0 = u3_0 + v3_0 + w3_0;
  const double M1 = u3_1 + v3_1 + w3_1;
  const double M2 = u3_2 + v3_2 + w3_2;
  const double M3 = c_4;
  const double M = M0 + M1 + M2 + M3;
  if (M == 0) return 0;
  // The reference tests M / Mk > 1 with a correctly rounded division; for
  // Mk != 0 that equals M > Mk (Mk > 0) or M < Mk (Mk < 0): the smallest
  // quotient above 1 that two doubles can form exceeds 1 + 2^-53, so rounding
  // never moves it to 1, and NaN / infinity compare the same way in both forms.
  auto above = [M](double mk) { return mk == 0 || (mk > 0 ? M > mk : M < mk); };
  const bool f0 = above(M0);
  const bool f1 = above(M1);
  const bool f2 = above(M2);
  const bool f3 = above(M3);
  if (f0 && f1 && f2 && f3) {
    if (same_dir(u0, u1, u2, u3) || same_dir(v0, v1, v2, v3) || same_dir(w0, w1, w2, w3)) return 1;
    return 0;
  }
  // FAST: a NaN product (an overflowed float product times 0) makes M NaN;
  // such tetrahedra go to the exact recomputation. The sign split then runs
  // without selects: p accumulates v + |v| (2v or +0) and n accumulates
  // |v| - v (0 or -2v), both exact, and the sums are halved at the end. Adding
  // +0 changes nothing, and doubling every term doubles every rounded partial
  // sum exactly (the terms are float products, far from the double range's
  // ends), so the halved sums are the reference's bit for bit.
  if (FAST && M != M) return 0;
  double p0 = 0, n0 = 0, p1 = 0, n1 = 0, p2 = 0, n2 = 0, p3 = 0, n3 = 0;
  auto acc = [](double v, double& p, double& n) {
    if constexpr (FAST) {
      p += v + fabs(v);
      n += fabs(v) - v;
    } else {
      acc3(v, p, n);
    }
  };
  acc(-u3v1w2, p0, n0); acc(u3v2w1, p0, n0); acc(u1v3w2, p0, n0);
  acc(-u2v3w1, p0, n0); acc(-u1v2w3, p0, n0); acc(u2v1w3, p0, n0);
  acc(-u3v2w0, p1, n1); acc(u3v0w2, p1, n1); acc(u2v3w0, p1, n1);
  acc(-u0v3w2, p1, n1); acc(u0v2w3, p1, n1); acc(-u2v0w3, p1, n1);
  acc(-u3v0w1, p2, n2); acc(u3v1w0, p2, n2); acc(u0v3w1, p2, n2);
  acc(-u1v3w0, p2, n2); acc(-u0v1w3, p2, n2); acc(u1v0w3, p2, n2);
  acc(u0v1w2, p3, n3); acc(-u0v2w1, p3, n3); acc(u1v2w0, p3, n3);
  acc(-u1v0w2, p3, n3); acc(u2v0w1, p3, n3); acc(-u2v1w0, p3, n3);
  if constexpr (FAST) {
    p0 *= 0.5;
    p1 *= 0.5;
    p2 *= 0.5;
    p3 *= 0.5;
    n0 *= 0.5;
    n1 *= 0.5;
    n2 *= 0.5;
    n3 *= 0.5;
  }
  const float pm0 = (float)p0, pm1 = (float)p1, pm2 = (float)p2, pm3 = (float)p3;
  const float nm0 = (float)n0, nm1 = (float)n1, nm2 = (float)n2, nm3 = (float)n3;
  float eb = 0;
  auto csb = [](float p, float n) {
    if constexpr (FAST) return cube_sign_bound_fast(p, n);
    else return cube_sign_bound(p, n);
  };
  if (!f0) {
    const float e1 = csb(pm0, nm0), e2 = csb(pm1 + pm2 + pm3, nm1 + nm2 + nm3);
    const float m = e1 < e2 ? e1 : e2;
    eb = eb > m ? eb : m;
  }
  if (!f1) {
    const float e1 = csb(pm1, nm1), e2 = csb(pm0 + pm2 + pm3, nm0 + nm2 + nm3);
    const float m = e1 < e2 ? e1 : e2;
    eb = eb > m ? eb : m;
  }
  if (!f2) {
    const float e1 = csb(pm2, nm2), e2 = csb(pm0 + pm1 + pm3, nm0 + nm1 + nm3);
    const float m = e1 < e2 ? e1 : e2;
    eb = eb > m ? eb : m;
  }
  if (!f3) {
    const float e1 = csb(pm3, nm3), e2 = csb(pm0 + pm1 + pm2, nm0 + nm1 + nm2);
    const float m = e1 < e2 ? e1 : e2;
    eb = eb > m ? eb : m;
  }
  if constexpr (FAST) {
    // NaN and infinity fail this test too (as unsigned bit patterns)
    uint32_t b = 0;
    for (float x : {pm0, pm1, pm2, pm3, nm0, nm1, nm2, nm3}) {
      uint32_t xb;
      __builtin_memcpy(&xb, &x, 4);
      b = b > xb ? b : xb;
    }
    if (b >= kSafeBits) return 0;
  }
  return eb;
}

__host__ __device__ inline double tet_bound(float u0, float u1, float u2, float u3, float v0, float v1, float v2,
                                            float v3, float w0, float w1, float w2, float w3) {
  return tet_bound_t<false>(u0, u1, u2, u3, v0, v1, v2, v3, w0, w1, w2, w3);
}

// Tetrahedra of a cell as (k, j, i) offsets of vertices 0..2; vertex 3 is the
// cell origin (src/sz_compress_cp_preserve_3d.cpp:9-46, first six entries).
__host__ __device__ inline void tet_offsets(int t, int v, int& di, int& dj, int& dk) {
  // rows: t; columns: vertex 0..2; packed as dk | dj << 1 | di << 2
  constexpr unsigned char kT[6][3] = {{4, 6, 7}, {2, 6, 7}, {4, 5, 7}, {1, 5, 7}, {2, 3, 7}, {1, 3, 7}};
  const unsigned char c = kT[t][v];
  dk = c & 1;
  dj = (c >> 1) & 1;
  di = (c >> 2) & 1;
}

// Vertex eb code for the 3D id mapping: 255 when the bound is not positive
// (stream id 0), else the id (>= 0) of eb_exponential_quantize(log2(1+eb), 2,
// 1, FLT_EPSILON) through the host-libm threshold table q3d.
__host__ __device__ inline uint32_t eb_code_3d(double eb, const double* q3d, int nq) {
  if (!(eb > 0)) return 255;
  uint32_t k = 0;
  while ((int)k + 1 < nq && eb >= q3d[k + 1]) k++;
  return k;
}

// One component of the 3D log-domain Lorenzo step (:498-516). E = 2^(id-23).
// FP32 form of the reference's double arithmetic, exact for the same reasons
// as the 2D step (E a power of two, |2qE| an exact float for id <= 33); note
// the reference rejects when |x' - x| >= E.
__host__ __device__ inline float l3d_step(float pred, float cur, int id, bool& ok, uint16_t& code) {
  const float diff = cur - pred;
  const float a = ldexpf(fabsf(diff), 23 - id);
  const float qf = floorf(fmaf(floorf(a), 0.5f, 0.5f));
  const float qs = copysignf(qf, diff);
  const float dec = pred + ldexpf(qs, id - 22);
  ok = (a < 65535.0f) & (fabsf(dec - cur) < ldexpf(1.0f, id - 23));
  code = (uint16_t)((int)qs + 32768);
  return dec;
}

}  // namespace cpsz
