// 2D critical-point-preserving codec on gfx950 (the drop-in for
// sz_compress_cp_preserve_2d_offline / sz_decompress_cp_preserve_2d_offline).
//
// Device pipeline (compress), all inputs resident in HBM:
//   k_eb2d        per-cell triangle bounds in LDS (8x64 vertex tiles + 1-cell
//                 halo), ordered per-vertex min, |U|,|V| scaling and the base-4
//                 eb id -> eb code per vertex and component, id histogram
//   k_lorenzo2d_c strip-pipelined wavefront Lorenzo + linear quantization:
//                 one wave = 64 rows (lane = row), the row above arrives through
//                 DPP wave_shr:1; strips hand their last row to the next strip
//                 as tagged agent-coherent values the consumer spins on
//   Huffman       histogram -> host tree (reference tie-break) -> encode
//   unpredictable values compacted in 12x12 block-major order
// Decompress mirrors it: parallel Huffman decode of ids and codes, block-major
// unpredictable scatter, wavefront reconstruction.
//
// Bit-exactness: every float/double operation follows the reference's types
// and order (src/sz_compress_cp_preserve_2d.cpp:17-172, quantize in
// include/sz_compression_utils.hpp:50-65); the Lorenzo step runs in FP32, which
// reproduces the reference's double arithmetic exactly (the argument is in the
// comments of lorenzo2d_step.hpp).
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "codec2d.hpp"
#include "pool.hpp"
#include "stream_reader.hpp"
#include "huff_gpu.hpp"
#include "libm_math.hpp"
#include "lorenzo2d_step.hpp"
#include "bounds2d_fast.hpp"

namespace cpsz {

// ----------------------------------------------------------------- eb bound --
namespace {

__device__ __forceinline__ double dminf(double a, double b) { return (a < b) ? a : b; }
__device__ __forceinline__ double dmaxm(double a, double b) { return (a > b) ? a : b; }

template <typename V>
__device__ __forceinline__ void dacc(V v, double& p, double& n) {
  if (v >= 0) p += v;
  else n += -v;
}

// max_eb_to_keep_sign (include/sz_cp_preserve_utils.hpp:59-90), degrees 1, 2
__device__ __forceinline__ double dsign_bound(double p, double n, int deg) {
  if (n == 0 || p == 0) return 1;
  double P = p, N = n;
  if (deg == 2) {
    P = sqrt(p);
    N = sqrt(n);
  }
  return fabs(P - N) / (P + N);
}

__device__ __forceinline__ double dbound2(double a, double b) {
  double p = 0, n = 0;
  dacc(a, p, n);
  dacc(b, p, n);
  return dsign_bound(p, n, 2);
}
__device__ __forceinline__ double dbound4(double a, double b, double c, double d) {
  double p = 0, n = 0;
  dacc(a, p, n);
  dacc(b, p, n);
  dacc(c, p, n);
  dacc(d, p, n);
  return dsign_bound(p, n, 2);
}

// max_eb_to_keep_sign_eigen_delta_2 (:75-123); rarely taken (CP triangles).
// T is the field's value type: the reference computes c[], m, C and the
// coefficients in T.
template <typename T>
__device__ __noinline__ double deigen(T u0, T u1, T u2, T v0, T v1, T v2, T x0, T x1, T x2, T y0, T y1, T y2) {
  double eb = 1;
  T c[4];
  {
    const T det = (x0 - x2) * (y1 - y2) - (x1 - x2) * (y0 - y2);
    c[0] = (y1 - y2) / det;
    c[1] = -(y0 - y2) / det;
    c[2] = -(x1 - x2) / det;
    c[3] = (x0 - x2) / det;
    double p = 0, n = 0;
    dacc(c[0] * u0, p, n);
    dacc(c[1] * u1, p, n);
    dacc(-(c[0] + c[1]) * u2, p, n);
    dacc(c[2] * v0, p, n);
    dacc(c[3] * v1, p, n);
    dacc(-(c[2] + c[3]) * v2, p, n);
    eb = dsign_bound(p, n, 1);
    double p2 = 0, n2 = 0;
    dacc(u0 * v1, p2, n2);
    dacc(-u0 * v2, p2, n2);
    dacc(u1 * v2, p2, n2);
    dacc(-u1 * v0, p2, n2);
    dacc(u2 * v0, p2, n2);
    dacc(-u2 * v1, p2, n2);
    eb = dminf(eb, dsign_bound(p2, n2, 2));
  }
  const T m = c[1] * c[2] - c[0] * c[3];
  const T C = (-m) * (u0 * v1 - u0 * v2 + u1 * v2 - u1 * v0 + u2 * v0 - u2 * v1);
  if (C == 0) return 0;
  if (C < 0) return eb;
  const T co[6] = {c[0] * u0, c[1] * u1, -(c[1] + c[0]) * u2, c[2] * v0, c[3] * v1, -(c[3] + c[2]) * v2};
  double p = 0, n = 0;
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 6; j++) dacc(co[i] * co[j], p, n);
  dacc(-4 * m * u1 * v0, p, n);
  dacc(4 * m * u2 * v0, p, n);
  dacc(4 * m * u0 * v1, p, n);
  dacc(-4 * m * u2 * v1, p, n);
  dacc(-4 * m * u0 * v2, p, n);
  dacc(4 * m * u1 * v2, p, n);
  return dminf(eb, dsign_bound(p, n, 2));
}

// max_eb_to_keep_position_and_type (src/sz_compress_cp_preserve_2d.cpp:125-172)
// for the two triangle shapes of a cell (coordinates passed "reversed", :187-206).
template <typename T>
__device__ double dtri(T u0, T u1, T u2, T v0, T v1, T v2, int lower) {
  const double u0v1 = u0 * v1, u1v0 = u1 * v0, u0v2 = u0 * v2, u2v0 = u2 * v0, u1v2 = u1 * v2,
               u2v1 = u2 * v1;
  const double det = u0v1 - u1v0 + u1v2 - u2v1 + u2v0 - u0v2;
  if (!(det != 0)) return 0;
  const bool f1 = (det / (u2v0 - u0v2) >= 1);
  const bool f2 = (det / (u1v2 - u2v1) >= 1);
  const bool f3 = (det / (u0v1 - u1v0) >= 1);
  if (f1 && f2 && f3) {
    const double e1 = dminf(dbound2(u2v0, -u0v2), dbound4(u0v1, -u1v0, u1v2, -u2v1));
    const double e2 = dminf(dbound2(u1v2, -u2v1), dbound4(u0v1, -u1v0, u2v0, -u0v2));
    const double e3 = dminf(dbound2(u0v1, -u1v0), dbound4(u1v2, -u2v1, u2v0, -u0v2));
    // X_upper = {{0,0},{1,0},{1,1}}, X_lower = {{0,0},{0,1},{1,1}}; x = X[.][1], y = X[.][0]
    const T o = 0, l = 1;
    const double e4 = lower ? dminf(e3, deigen<T>(u0, u1, u2, v0, v1, v2, o, l, l, o, o, l))
                            : dminf(e3, deigen<T>(u0, u1, u2, v0, v1, v2, o, o, l, o, l, l));
    double eb = dminf(e1, e2);
    return dminf(eb, e4);
  }
  double eb = 0;
  if (!f1) eb = dmaxm(eb, dminf(dbound2(u2v0, -u0v2), dbound4(u0v1, -u1v0, u1v2, -u2v1)));
  if (!f2) eb = dmaxm(eb, dminf(dbound2(u1v2, -u2v1), dbound4(u0v1, -u1v0, u2v0, -u0v2)));
  if (!f3) eb = dmaxm(eb, dminf(dbound2(u0v1, -u1v0), dbound4(u1v2, -u2v1, u2v0, -u0v2)));
  return eb;
}

// The FP32-approximated triangle bound (tri_fast) and its constants are in
// bounds2d_fast.hpp, shared with the host check.
struct DeviceRoots {
  __device__ static float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
  __device__ static float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
};

// eb code: 0 = zero precision; c >= 1: reference id c-1 with precision
// 4^(c-1) * DBL_EPSILON (id 0 keeps precision eps in the compressor).
__device__ __forceinline__ uint32_t eb_code_2d(double x, const double* __restrict__ q2d, int nq) {
  if (!(x > 2.220446049250313080847263336181640625e-16)) return 0;
  const int e = (int)((__double_as_longlong(x) >> 52) & 0x7FF) - 1023 + 52;  // exponent of x/eps
  int k = e >> 1;
  if (k + 1 < nq && x >= q2d[k + 1]) k++;
  return (uint32_t)k + 1;
}

// Eb code from a threshold table q[1..nq-2] (q[nq-1] = +inf): the largest k
// with x >= q[k], 0 below q[1].
__device__ __forceinline__ uint32_t code_from_table(double x, const double* __restrict__ q, int nq) {
  if (!(x >= q[1])) return 0;
  int lo = 1, hi = nq - 2;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (x >= q[mid]) lo = mid;
    else hi = mid - 1;
  }
  return (uint32_t)lo;
}

constexpr int kTH = 16, kTW = 32;  // vertex tile (17x33 cells: 9 wave rounds of triangles per 512 vertices)

template <typename T>
__device__ __forceinline__ void eb2d_tile(const T* __restrict__ U, const T* __restrict__ V, int r1, int r2,
                                          double max_pwr_eb, const double* __restrict__ q2d, int nq,
                                          uint16_t* __restrict__ cu, uint16_t* __restrict__ cv,
                                          uint32_t* __restrict__ hist, int* __restrict__ flags, const EbOverride& ov,
                                          const double* __restrict__ qlog, int nqlog, int bx, int by) {
  __shared__ T su[kTH + 2][kTW + 2];
  __shared__ T sv[kTH + 2][kTW + 2];
  __shared__ double tb[kTH + 1][kTW + 1][2];
  __shared__ uint32_t sh_hist[kMaxEbCode];
  const int tid = threadIdx.x;
  const int i0 = by * kTH, j0 = bx * kTW;
  for (int k = tid; k < kMaxEbCode; k += 256) sh_hist[k] = 0;
  for (int k = tid; k < (kTH + 2) * (kTW + 2); k += 256) {
    const int r = k / (kTW + 2), c = k % (kTW + 2);
    const int gi = i0 - 1 + r, gj = j0 - 1 + c;
    T a = 0, b = 0;
    if (gi >= 0 && gi < r1 && gj >= 0 && gj < r2) {
      a = U[(size_t)gi * r2 + gj];
      b = V[(size_t)gi * r2 + gj];
    }
    su[r][c] = a;
    sv[r][c] = b;
  }
  __syncthreads();
  for (int k = tid; k < (kTH + 1) * (kTW + 1); k += 256) {
    const int cr = k / (kTW + 1), cc = k % (kTW + 1);
    const int ci = i0 - 1 + cr, cj = j0 - 1 + cc;
    double up = INFINITY, lo = INFINITY;
    if (ci >= 0 && ci < r1 - 1 && cj >= 0 && cj < r2 - 1) {
      // upper: (i,j),(i+1,j),(i+1,j+1); lower: (i,j),(i,j+1),(i+1,j+1)
      up = dtri<T>(su[cr][cc], su[cr + 1][cc], su[cr + 1][cc + 1], sv[cr][cc], sv[cr + 1][cc],
                   sv[cr + 1][cc + 1], 0);
      lo = dtri<T>(su[cr][cc], su[cr][cc + 1], su[cr + 1][cc + 1], sv[cr][cc], sv[cr][cc + 1],
                   sv[cr + 1][cc + 1], 1);
    }
    tb[cr][cc][0] = up;
    tb[cr][cc][1] = lo;
  }
  __syncthreads();
  int bad = 0;
  for (int k = tid; k < kTH * kTW; k += 256) {
    const int vr = k / kTW, vc = k % kTW;
    const int i = i0 + vr, j = j0 + vc;
    if (i >= r1 || j >= r2) continue;
    const T u = su[vr + 1][vc + 1], v = sv[vr + 1][vc + 1];
    // incident triangles in the reference's visiting order (cells row-major, upper first)
    double eb = max_pwr_eb;
    eb = dminf(eb, tb[vr][vc][0]);
    eb = dminf(eb, tb[vr][vc][1]);
    eb = dminf(eb, tb[vr][vc + 1][0]);
    eb = dminf(eb, tb[vr + 1][vc][1]);
    eb = dminf(eb, tb[vr + 1][vc + 1][0]);
    eb = dminf(eb, tb[vr + 1][vc + 1][1]);
    const size_t vix = (size_t)i * r2 + j;
    if (ov.eb_out) ov.eb_out[vix] = eb;
    if (ov.gain && !(ov.fallback && ov.fallback[vix]) && ov.gain[vix])
      eb = dminf(max_pwr_eb, ldexp(eb, (int)ov.gain[vix]));
    if (ov.fallback && ov.fallback[vix] == 2) eb = 0;  // stored exactly
    if (!isfinite(u) || !isfinite(v) || !(eb == eb)) bad = 1;
    if (qlog) {
      // offline_log: one code per vertex for both components, from the bound
      // itself (src/sz_compress_cp_preserve_2d.cpp:327-331)
      const uint32_t a = code_from_table(eb, qlog, nqlog);
      cu[(size_t)i * r2 + j] = (uint16_t)a;
      atomicAdd(&sh_hist[a ? a - 1 : 0], 1u);
      continue;
    }
    // the reference's eb/DBL_EPSILON overflows from 2^972 on (undefined id)
    // eb_u = fabs(U[i]) * eb[i] (src/sz_compress_cp_preserve_2d.cpp:231)
    const double au = (double)fabs(u) * eb, av = (double)fabs(v) * eb;
    if (au >= 0x1p972 || av >= 0x1p972) bad = 1;
    const uint32_t a = eb_code_2d(au, q2d, nq);
    const uint32_t b = eb_code_2d(av, q2d, nq);
    cu[(size_t)i * r2 + j] = (uint16_t)a;
    cv[(size_t)i * r2 + j] = (uint16_t)b;
    atomicAdd(&sh_hist[a ? a - 1 : 0], 1u);
    atomicAdd(&sh_hist[b ? b - 1 : 0], 1u);
    if (a > 82 || b > 82) atomicOr(&flags[1], 1);  // beyond the FP32 Lorenzo path
  }
  if (bad) atomicOr(&flags[0], 1);
  __syncthreads();
  for (int k = tid; k < kMaxEbCode; k += 256)
    if (sh_hist[k]) atomicAdd(&hist[k], sh_hist[k]);
}

// code(x) >= c for the eb code function above (c = code(y) for some y >= x):
// the codes step at eps and at the thresholds q2d[1..]
__device__ __forceinline__ bool code_reaches(uint32_t c, double x, const double* __restrict__ q2d) {
  if (c == 0) return true;
  if (c == 1) return x > 2.220446049250313080847263336181640625e-16;
  return x >= q2d[c - 1];
}

// Marks, in cu, a vertex whose bound the fix-up pass computes in FP64 (eb
// codes stay far below).
constexpr uint16_t kFixCode = 0xFFFF;

// Every field of a batch in one launch (float fields): a job per field, its
// tiles numbered from blk0 row by row.
struct Eb2dJob {
  const float* U;
  const float* V;
  int r1, r2, gx;  // gx: tiles per tile row
  uint16_t* cu;
  uint16_t* cv;
  uint32_t* hist;
  int* flags;
  EbOverride ov;
  uint32_t blk0;
};

__device__ __forceinline__ const Eb2dJob& eb_job_of(const Eb2dJob* __restrict__ jobs, int nj, uint32_t tile) {
  int lo = 0, hi = nj - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].blk0 <= tile) lo = mid;
    else hi = mid - 1;
  }
  return jobs[lo];
}

// eb2d_tile for float fields without overrides or log bounds: the cell table
// holds tri_fast entries and each vertex takes its eb codes from the
// enclosure of its bound. Vertices that need the FP64 bound get kFixCode, and
// the tile goes on the fix-up list (fix_tiles, counted by fix_n) for
// k_eb2d_fix, which also counts them in the histogram.
__device__ __forceinline__ void eb2d_tile_fast(const float* __restrict__ U, const float* __restrict__ V, int r1,
                                               int r2, double max_pwr_eb, const double* __restrict__ q2d, int nq,
                                               uint16_t* __restrict__ cu, uint16_t* __restrict__ cv,
                                               uint32_t* __restrict__ hist, int* __restrict__ flags, int bx, int by,
                                               uint32_t tile, uint32_t* __restrict__ fix_tiles,
                                               uint32_t* __restrict__ fix_n) {
  __shared__ float su[kTH + 2][kTW + 2];
  __shared__ float sv[kTH + 2][kTW + 2];
  __shared__ float tb[kTH + 1][kTW + 1][2];
  __shared__ uint32_t sh_hist[kMaxEbCode];
  __shared__ int sh_fix;
  const int tid = threadIdx.x;
  const int i0 = by * kTH, j0 = bx * kTW;
  for (int k = tid; k < kMaxEbCode; k += 256) sh_hist[k] = 0;
  if (tid == 0) sh_fix = 0;
  int in_range = 1;
  for (int k = tid; k < (kTH + 2) * (kTW + 2); k += 256) {
    const int r = k / (kTW + 2), c = k % (kTW + 2);
    const int gi = i0 - 1 + r, gj = j0 - 1 + c;
    float a = 0, b = 0;
    if (gi >= 0 && gi < r1 && gj >= 0 && gj < r2) {
      a = U[(size_t)gi * r2 + gj];
      b = V[(size_t)gi * r2 + gj];
    }
    su[r][c] = a;
    sv[r][c] = b;
    in_range &= (int)(value_ok(a) && value_ok(b));
  }
  const bool checked = __syncthreads_and(in_range) != 0;
  // one triangle per item: upper (i,j),(i+1,j),(i+1,j+1); lower (i,j),(i,j+1),(i+1,j+1)
  for (int k = tid; k < 2 * (kTH + 1) * (kTW + 1); k += 256) {
    const int cell = k >> 1, lower = k & 1;
    const int cr = cell / (kTW + 1), cc = cell % (kTW + 1);
    const int ci = i0 - 1 + cr, cj = j0 - 1 + cc;
    float t = INFINITY;
    if (ci >= 0 && ci < r1 - 1 && cj >= 0 && cj < r2 - 1) {
      const int mr = lower ? cr : cr + 1, mc = lower ? cc + 1 : cc;
      t = tri_fast<DeviceRoots>(su[cr][cc], su[mr][mc], su[cr + 1][cc + 1], sv[cr][cc], sv[mr][mc], sv[cr + 1][cc + 1],
                                   checked);
    }
    tb[cr][cc][lower] = t;
  }
  __syncthreads();
  int bad = 0, fix = 0;
  for (int k = tid; k < kTH * kTW; k += 256) {
    const int vr = k / kTW, vc = k % kTW;
    const int i = i0 + vr, j = j0 + vc;
    if (i >= r1 || j >= r2) continue;
    const float u = su[vr + 1][vc + 1], v = sv[vr + 1][vc + 1];
    // the six incident triangles (for the reference, in the order cells
    // row-major, upper first; the order matters only through NaN bounds,
    // which go to the fix-up pass)
    const float m = fminf(fminf(fminf(tb[vr][vc][0], tb[vr][vc][1]), fminf(tb[vr][vc + 1][0], tb[vr + 1][vc][1])),
                          fminf(tb[vr + 1][vc + 1][0], tb[vr + 1][vc + 1][1]));
    uint32_t a, b;
    if (m > 0 && m < kExact1) {
      // the smallest bound is an approximation: exact 1s and missing cells do not matter
      const double am = m;
      const double lo = dminf(max_pwr_eb, enclosure_lo(am));
      const double hi = dminf(max_pwr_eb, enclosure_hi(am));
      const double uh = (double)fabsf(u) * hi, vh = (double)fabsf(v) * hi;
      a = eb_code_2d(uh, q2d, nq);
      b = eb_code_2d(vh, q2d, nq);
      if (!(uh < 0x1p972 && vh < 0x1p972 && code_reaches(a, (double)fabsf(u) * lo, q2d) &&
            code_reaches(b, (double)fabsf(v) * lo, q2d))) {
        cu[(size_t)i * r2 + j] = kFixCode;
        fix = 1;
        continue;
      }
    } else if (m >= 0) {
      // exact: 0, 1 or no triangle
      const double eb = m == 0 ? 0.0 : (m == kExact1 ? dminf(max_pwr_eb, 1.0) : max_pwr_eb);
      // eb_u = fabs(U[i]) * eb[i] (src/sz_compress_cp_preserve_2d.cpp:231); the
      // reference's eb/DBL_EPSILON overflows from 2^972 on (undefined id)
      const double au = (double)fabsf(u) * eb, av = (double)fabsf(v) * eb;
      if (au >= 0x1p972 || av >= 0x1p972 || !isfinite(u) || !isfinite(v)) bad = 1;
      a = eb_code_2d(au, q2d, nq);
      b = eb_code_2d(av, q2d, nq);
    } else {
      cu[(size_t)i * r2 + j] = kFixCode;
      fix = 1;
      continue;
    }
    cu[(size_t)i * r2 + j] = (uint16_t)a;
    cv[(size_t)i * r2 + j] = (uint16_t)b;
    atomicAdd(&sh_hist[a ? a - 1 : 0], 1u);
    atomicAdd(&sh_hist[b ? b - 1 : 0], 1u);
    if (a > 82 || b > 82) atomicOr(&flags[1], 1);  // beyond the FP32 Lorenzo path
  }
  if (bad) atomicOr(&flags[0], 1);
  if (fix) sh_fix = 1;
  __syncthreads();
  for (int k = tid; k < kMaxEbCode; k += 256)
    if (sh_hist[k]) atomicAdd(&hist[k], sh_hist[k]);
  if (tid == 0 && sh_fix) fix_tiles[atomicAdd(fix_n, 1u)] = tile;
}

// The FP64 bounds of the vertices eb2d_tile_fast marked, tile by tile from the
// fix-up list: the six incident triangles in the reference's order, from
// global memory (missing cells contribute +inf, as in eb2d_tile).
__global__ void __launch_bounds__(256) k_eb2d_fix(const Eb2dJob* __restrict__ jobs, int nj,
                                                  const uint32_t* __restrict__ fix_tiles,
                                                  const uint32_t* __restrict__ fix_n, double max_pwr_eb,
                                                  const double* __restrict__ q2d, int nq) {
  const uint32_t n = *fix_n;
  for (uint32_t e = blockIdx.x; e < n; e += gridDim.x) {
    const uint32_t tile = fix_tiles[e];
    const Eb2dJob& J = eb_job_of(jobs, nj, tile);
    const uint32_t t = tile - J.blk0;
    const int i0 = (int)(t / (uint32_t)J.gx) * kTH, j0 = (int)(t % (uint32_t)J.gx) * kTW;
    const int r1 = J.r1, r2 = J.r2;
    int bad = 0;
    for (int k = threadIdx.x; k < kTH * kTW; k += blockDim.x) {
      const int i = i0 + k / kTW, j = j0 + k % kTW;
      if (i >= r1 || j >= r2) continue;
      const size_t vix = (size_t)i * r2 + j;
      if (J.cu[vix] != kFixCode) continue;
      auto tri = [&](int ci, int cj, int lower) -> double {
        if (ci < 0 || ci >= r1 - 1 || cj < 0 || cj >= r2 - 1) return INFINITY;
        const size_t c0 = (size_t)ci * r2 + cj, c2 = c0 + r2 + 1, c1 = lower ? c0 + 1 : c0 + r2;
        return dtri<float>(J.U[c0], J.U[c1], J.U[c2], J.V[c0], J.V[c1], J.V[c2], lower);
      };
      double eb = max_pwr_eb;
      eb = dminf(eb, tri(i - 1, j - 1, 0));
      eb = dminf(eb, tri(i - 1, j - 1, 1));
      eb = dminf(eb, tri(i - 1, j, 0));
      eb = dminf(eb, tri(i, j - 1, 1));
      eb = dminf(eb, tri(i, j, 0));
      eb = dminf(eb, tri(i, j, 1));
      const float u = J.U[vix], v = J.V[vix];
      const double au = (double)fabsf(u) * eb, av = (double)fabsf(v) * eb;
      if (au >= 0x1p972 || av >= 0x1p972 || !(eb == eb) || !isfinite(u) || !isfinite(v)) bad = 1;
      const uint32_t a = eb_code_2d(au, q2d, nq), b = eb_code_2d(av, q2d, nq);
      J.cu[vix] = (uint16_t)a;
      J.cv[vix] = (uint16_t)b;
      atomicAdd(&J.hist[a ? a - 1 : 0], 1u);
      atomicAdd(&J.hist[b ? b - 1 : 0], 1u);
      if (a > 82 || b > 82) atomicOr(&J.flags[1], 1);
    }
    if (bad) atomicOr(&J.flags[0], 1);
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_eb2d(const T* __restrict__ U, const T* __restrict__ V,
                                              int r1, int r2, double max_pwr_eb,
                                              const double* __restrict__ q2d, int nq,
                                              uint16_t* __restrict__ cu, uint16_t* __restrict__ cv,
                                              uint32_t* __restrict__ hist, int* __restrict__ flags,
                                              EbOverride ov, const double* __restrict__ qlog, int nqlog) {
  eb2d_tile<T>(U, V, r1, r2, max_pwr_eb, q2d, nq, cu, cv, hist, flags, ov, qlog, nqlog, blockIdx.x, blockIdx.y);
}

// APPROX: eb2d_tile_fast (jobs without overrides), with its fix-up list
template <bool APPROX>
__global__ void __launch_bounds__(256) k_eb2d_multi(const Eb2dJob* __restrict__ jobs, int nj, double max_pwr_eb,
                                                    const double* __restrict__ q2d, int nq,
                                                    uint32_t* __restrict__ fix_tiles, uint32_t* __restrict__ fix_n) {
  const Eb2dJob& J = eb_job_of(jobs, nj, blockIdx.x);
  const uint32_t t = blockIdx.x - J.blk0;
  const int bx = (int)(t % (uint32_t)J.gx), by = (int)(t / (uint32_t)J.gx);
  if constexpr (APPROX)
    eb2d_tile_fast(J.U, J.V, J.r1, J.r2, max_pwr_eb, q2d, nq, J.cu, J.cv, J.hist, J.flags, bx, by, blockIdx.x,
                   fix_tiles, fix_n);
  else
    eb2d_tile<float>(J.U, J.V, J.r1, J.r2, max_pwr_eb, q2d, nq, J.cu, J.cv, J.hist, J.flags, J.ov, nullptr, 0, bx,
                     by);
}

// ------------------------------------------------------------ wavefront ----
constexpr int kCH = 32;  // steps per chunk (boundary publish / prefetch granularity)

__device__ __forceinline__ float shr1(float src, float lane0) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(lane0), __float_as_int(src),
                                                    0x138 /* wave_shr:1 */, 0xF, 0xF, false));
}

struct Strip2d {
  const float* x[2];
  const uint16_t* c[2];  // compress: eb codes; decompress: stream ids
  uint16_t* q[2];        // quant codes (written by compress, read by decompress)
  const float* raw[2];   // decompress: unpredictable value per element
  float* out[2];         // decompress output
  uint64_t* bnd;         // [2][nstrips][r2] boundary rows (kBndTag | float bits, 0 = not yet)
  int* prog;             // [2][nstrips] unused; [2 * nstrips] the item ticket
  int* ticket;
  int r1, r2, nstrips;
  int wide;   // ids above 81: double-precision step
  int shift;  // log2 of the eb base: 2 (offline, 4^id eps) or 1 (offline_log, 2^id eps)
  // compress of a batch: the eb pass's flags, read in place of `wide`
  // ([1]); a field with [0] set (non-finite input) is skipped as a whole
  const int* flags;
  // batch: issue priority (0-3) of the field's waves on their SIMD, higher
  // for longer wavefronts
  int prio;
};

// One item (component, 64-row strip) of a field's wavefront. The strip's last
// row is handed to the next strip value by value: lane 63 stores each value
// with a tag in the high word (kBndTag | float bits; the rows are zeroed before
// the launch), and the consumer spins until the values it needs carry the tag.
// No progress counter and no wait for store completion sit on the producer's
// path. The strip's own rows for chunk c+1 are loaded while chunk c computes.
constexpr uint64_t kBndTag = 1ull << 32;

__device__ __forceinline__ uint64_t ld_coh64(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_coh64(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct StripChunk {
  float xs[kCH];         // compress: field values
  uint32_t cs[kCH / 2];  // eb codes (compress) / stream ids (decompress), two u16 per word
  uint32_t qs[kCH / 2];  // decompress: quantization codes, two u16 per word
  float raws[kCH];       // decompress: unpredictable value where the code is 0
};

__device__ __forceinline__ uint32_t lo16(uint32_t w, int k) { return (k & 1) ? w >> 16 : w & 0xFFFFu; }

// Interior chunks move a lane's kCH consecutive elements of its row with
// 16-byte loads from the 16-byte-aligned block holding the first element (a
// 16-byte-aligned load never leaves the pages of the row it starts in), then
// shift them into place with selects. The rows of a strip are 64 different
// lines per instruction, so wide loads cut the lines the L1 and L2 serve per
// element by 4 (8 for the u16 arrays) against element-wise loads.
__device__ __forceinline__ uint32_t sel4(uint32_t a, uint32_t b, uint32_t c, uint32_t d, uint32_t q) {
  const uint32_t lo = (q & 1) ? b : a, hi = (q & 1) ? d : c;
  return (q & 2) ? hi : lo;
}

// kCH floats from p (4-byte aligned) into v
__device__ __forceinline__ void ld_row_f32(const float* p, float* v) {
  const uintptr_t a = (uintptr_t)p;
  const uint32_t s = (uint32_t)(a >> 2) & 3;
  const uint4* w = (const uint4*)(a & ~(uintptr_t)15);
  uint32_t d[kCH + 4];
#pragma unroll
  for (int i = 0; i < kCH / 4; i++) {
    const uint4 x = w[i];
    d[4 * i] = x.x;
    d[4 * i + 1] = x.y;
    d[4 * i + 2] = x.z;
    d[4 * i + 3] = x.w;
  }
  uint4 t = make_uint4(0, 0, 0, 0);
  if (s) t = w[kCH / 4];
  d[kCH] = t.x;
  d[kCH + 1] = t.y;
  d[kCH + 2] = t.z;
  d[kCH + 3] = t.w;
#pragma unroll
  for (int k = 0; k < kCH; k++) v[k] = __uint_as_float(sel4(d[k], d[k + 1], d[k + 2], d[k + 3], s));
}

// kCH u16 from p (2-byte aligned) into kCH/2 words, element k in word k/2
// (low half for even k)
__device__ __forceinline__ void ld_row_u16(const uint16_t* p, uint32_t* e) {
  const uintptr_t a = (uintptr_t)p;
  const uint32_t s = (uint32_t)(a >> 1) & 7;  // u16 offset in the 16-byte block
  const uint4* w = (const uint4*)(a & ~(uintptr_t)15);
  uint32_t d[kCH / 2 + 4];
#pragma unroll
  for (int i = 0; i < kCH / 8; i++) {
    const uint4 x = w[i];
    d[4 * i] = x.x;
    d[4 * i + 1] = x.y;
    d[4 * i + 2] = x.z;
    d[4 * i + 3] = x.w;
  }
  uint4 t = make_uint4(0, 0, 0, 0);
  if (s) t = w[kCH / 8];
  d[kCH / 2] = t.x;
  d[kCH / 2 + 1] = t.y;
  d[kCH / 2 + 2] = t.z;
  d[kCH / 2 + 3] = t.w;
  const uint32_t q = s >> 1, b = (s & 1) * 2;
  uint32_t g[kCH / 2 + 1];
#pragma unroll
  for (int m = 0; m <= kCH / 2; m++)
    g[m] = m + 3 < kCH / 2 + 4 ? sel4(d[m], d[m + 1], d[m + 2], d[m + 3], q) : d[m];
#pragma unroll
  for (int m = 0; m < kCH / 2; m++) e[m] = __builtin_amdgcn_alignbyte(g[m + 1], g[m], b);
}

// kCH u16 (two per word, as ld_row_u16) to p (2-byte aligned): 4-byte
// stores, with a 2-byte store at each end when p is not 4-byte aligned
__device__ __forceinline__ void st_row_u16(uint16_t* p, const uint32_t* e) {
  const uintptr_t a = (uintptr_t)p;
  const bool odd = (a & 2) != 0;
  uint32_t* w = (uint32_t*)((a + 2) & ~(uintptr_t)3);  // first 4-byte boundary at or after p
  if (odd) {
    p[0] = (uint16_t)e[0];
    p[kCH - 1] = (uint16_t)(e[kCH / 2 - 1] >> 16);
  }
#pragma unroll
  for (int m = 0; m < kCH / 2; m++) {
    const uint32_t v = odd ? (m + 1 < kCH / 2 ? __builtin_amdgcn_alignbyte(e[m + 1], e[m], 2) : 0u) : e[m];
    if (!odd || m + 1 < kCH / 2) w[m] = v;
  }
}

// kCH floats to p (4-byte aligned): 8-byte stores, with a 4-byte store at
// each end when p is not 8-byte aligned
__device__ __forceinline__ void st_row_f32(float* p, const float* v) {
  const uintptr_t a = (uintptr_t)p;
  const bool odd = (a & 4) != 0;
  float2* w = (float2*)((a + 4) & ~(uintptr_t)7);
  if (odd) {
    p[0] = v[0];
    p[kCH - 1] = v[kCH - 1];
  }
#pragma unroll
  for (int m = 0; m < kCH / 2; m++) {
    const float x = odd ? v[2 * m + 1] : v[2 * m];
    const float y = odd ? (2 * m + 2 < kCH ? v[2 * m + 2] : 0.f) : v[2 * m + 1];
    if (!odd || m + 1 < kCH / 2) w[m] = make_float2(x, y);
  }
}

// SHIFT (a.shift as a compile-time constant) keeps the quantum's exponent a
// shift of the id: a runtime factor costs an integer multiply per element.
template <bool COMPRESS, bool WIDE, int SHIFT>
__device__ __forceinline__ void strip_item(const Strip2d& a, int item, int lane) {
  const int r1 = a.r1, r2 = a.r2;
  const int comp = item & 1, s = item >> 1;
  const int row = s * 64 + lane;
  const bool rvalid = row < r1;
  const size_t roff = (size_t)(rvalid ? row : 0) * r2;
  // (selects rather than a.x[comp]: an indexed member would put a copied
  // descriptor in scratch memory)
  const float* X = COMPRESS ? (comp ? a.x[1] : a.x[0]) + roff : nullptr;
  const uint16_t* C = (comp ? a.c[1] : a.c[0]) + roff;
  uint16_t* Q = (comp ? a.q[1] : a.q[0]) + roff;
  const float* RAW = COMPRESS ? nullptr : (comp ? a.raw[1] : a.raw[0]) + roff;
  float* OUT = COMPRESS ? nullptr : (comp ? a.out[1] : a.out[0]) + roff;
  const bool has_up = s > 0;
  const bool publish = s + 1 < a.nstrips;
  const uint64_t* bnd_in = a.bnd + ((size_t)comp * a.nstrips + (s > 0 ? s - 1 : 0)) * r2;
  uint64_t* bnd_out = a.bnd + ((size_t)comp * a.nstrips + s) * r2;
  float Rprev = 0.f, Uprev = 0.f;
  const int steps = r2 + 63;
  // An interior chunk has every lane's kCH columns inside the row and all 64
  // rows present, so none of its steps needs a bounds mask; most chunks are
  // interior and run the unmasked body, with wide loads and buffered stores.
  const bool full_rows = s * 64 + 64 <= r1;
  auto interior = [&](int t0) { return full_rows && t0 >= 63 && t0 + kCH <= r2; };
  auto load = [&](StripChunk& B, int t0, auto inner) {
    constexpr bool IN = decltype(inner)::value;
    if constexpr (IN) {
      const int j0 = t0 - lane;
      ld_row_u16(C + j0, B.cs);
      if (COMPRESS) ld_row_f32(X + j0, B.xs);
      else ld_row_u16(Q + j0, B.qs);
      return;
    }
#pragma unroll
    for (int k = 0; k < kCH; k += 2) {
      uint32_t c2[2], q2[2];
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int j = t0 + k + h - lane;
        const bool act = rvalid && j >= 0 && j < r2;
        c2[h] = act ? C[j] : 0u;
        if (COMPRESS) B.xs[k + h] = act ? X[j] : 0.f;
        else q2[h] = act ? Q[j] : 0u;
      }
      B.cs[k >> 1] = c2[0] | c2[1] << 16;
      if (!COMPRESS) B.qs[k >> 1] = q2[0] | q2[1] << 16;
    }
  };
  // decompress: the unpredictable values are read only where the code is 0
  // (a few percent of elements), once the chunk's codes have arrived
  auto load_raw = [&](StripChunk& B, int t0, auto inner) {
    constexpr bool IN = decltype(inner)::value;
#pragma unroll
    for (int k = 0; k < kCH; k++) {
      const int j = t0 + k - lane;
      const bool act = IN || (rvalid && j >= 0 && j < r2);
      B.raws[k] = (act && lo16(B.qs[k >> 1], k) == 0) ? RAW[j] : 0.f;
    }
  };
  auto chunk = [&](const StripChunk& B, StripChunk& N, int t0, auto inner) {
    constexpr bool IN = decltype(inner)::value;
    float inc = 0.f;
    if (has_up && t0 < r2) {
      const bool want = lane < kCH && t0 + lane < r2;
      uint64_t v = want ? ld_coh64(bnd_in + t0 + lane) : kBndTag;
      while (!__all((v >> 32) != 0)) {
        __builtin_amdgcn_s_sleep(1);
        if ((v >> 32) == 0) v = ld_coh64(bnd_in + t0 + lane);
      }
      inc = __uint_as_float((uint32_t)v);
    }
    const bool more = t0 + kCH < steps, next_in = interior(t0 + kCH);
    if (more) {
      if (next_in) load(N, t0 + kCH, std::true_type{});
      else load(N, t0 + kCH, std::false_type{});
    }
    uint32_t qo[IN && COMPRESS ? kCH / 2 : 1];
    float oo[IN && !COMPRESS ? kCH : 1];
    // the boundary row's chunk (lane 63's values) gathers in lane k and
    // leaves in one store after the chunk: the next strip waits for whole
    // chunks anyway (one step later than per-step stores, one store in place
    // of kCH)
    uint32_t pubv = 0;
#pragma unroll
    for (int k = 0; k < kCH; k++) {
      if (!COMPRESS && k == kCH / 2 && more) {
        if (next_in) load_raw(N, t0 + kCH, std::true_type{});
        else load_raw(N, t0 + kCH, std::false_type{});
      }
      const int t = t0 + k;
      const int j = t - lane;
      const bool act = IN || (rvalid && j >= 0 && j < r2);
      const float lane0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(inc), k));
      const float Uv = shr1(Rprev, lane0);
      const float pred = (Rprev + Uv) - Uprev;
      const uint32_t ck = lo16(B.cs[k >> 1], k);
      float R;
      if (COMPRESS) {
        uint16_t code;
        R = l2d_compress_step<WIDE>(pred, B.xs[k], ck, code, SHIFT);
        if constexpr (IN) {
          if (k & 1) qo[k >> 1] |= (uint32_t)code << 16;
          else qo[k >> 1] = code;
        } else if (act) {
          Q[j] = code;
        }
      } else {
        R = l2d_decompress_step<WIDE>(pred, lo16(B.qs[k >> 1], k), ck, B.raws[k], SHIFT);
        if constexpr (IN) oo[k] = R;
        else if (act) OUT[j] = R;
      }
      if (publish) {
        const uint32_t r63 = (uint32_t)__builtin_amdgcn_readlane(__float_as_int(R), 63);
        pubv = lane == k ? r63 : pubv;
      }
      Uprev = Uv;
      Rprev = (!IN && j < 0) ? 0.f : R;
    }
    if (publish) {
      const int jp = t0 - 63 + lane;
      if (lane < kCH && jp >= 0 && jp < r2) st_coh64(bnd_out + jp, kBndTag | pubv);
    }
    if constexpr (IN) {
      if constexpr (COMPRESS) st_row_u16(Q + (t0 - lane), qo);
      else st_row_f32(OUT + (t0 - lane), oo);
    }
  };
  auto run = [&](const StripChunk& B, StripChunk& N, int t0) {
    if (interior(t0)) chunk(B, N, t0, std::true_type{});
    else chunk(B, N, t0, std::false_type{});
  };
  StripChunk A, B;
  load(A, 0, std::false_type{});
  if (!COMPRESS) load_raw(A, 0, std::false_type{});
  for (int t0 = 0; t0 < steps; t0 += 2 * kCH) {
    run(A, B, t0);
    if (t0 + kCH < steps) run(B, A, t0 + kCH);
  }
}

// One strip variant per kernel (the host picks it): the unrolled strip loop
// is large, and one variant keeps the kernel's hot code and registers small.
template <bool COMPRESS, bool WIDE, int SHIFT>
__global__ void __launch_bounds__(64) k_lorenzo2d(Strip2d a) {
  const int lane = threadIdx.x;
  const int items = 2 * a.nstrips;
  for (;;) {
    int item = 0;
    if (lane == 0) item = atomicAdd(a.ticket, 1);
    item = __shfl(item, 0);
    if (item >= items) return;
    strip_item<COMPRESS, WIDE, SHIFT>(a, item, lane);
  }
}

// All fields of a batch in one persistent launch: items are numbered field by
// field (start[f] = first item of field f) and taken from one ticket counter,
// so every wait (strip s on strip s-1 of the same field) targets an earlier
// ticket. Each field keeps its own boundary rows and progress counters.
struct StripBatch {
  const Strip2d* f;
  const int* start;
  int nf, items;
  int* ticket;
};

// Clears every field's boundary rows and counters (block row y = field) and
// the batch ticket, from the field table already on the device.
__global__ void __launch_bounds__(256) k_zero_fields(const Strip2d* __restrict__ f, int* __restrict__ ticket) {
  const Strip2d& a = f[blockIdx.y];
  const size_t nb = (size_t)2 * a.nstrips * a.r2 * 2, nz = 2 * (size_t)a.nstrips + 1;
  const size_t step = (size_t)gridDim.x * blockDim.x;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nb; i += step) ((int*)a.bnd)[i] = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nz; i += step) a.prog[i] = 0;
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *ticket = 0;
}

template <bool COMPRESS>
__global__ void __launch_bounds__(64) k_lorenzo2d_batch(StripBatch b) {
  const int lane = threadIdx.x;
  for (;;) {
    int item = 0;
    if (lane == 0) item = atomicAdd(b.ticket, 1);
    item = __shfl(item, 0);
    if (item >= b.items) return;
    int lo = 0, hi = b.nf - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (b.start[mid] <= item) lo = mid;
      else hi = mid - 1;
    }
    const Strip2d a = b.f[lo];
    // The batch ends when its longest wavefront does, and the waves of a
    // SIMD share its issue slots: the longer a field's dependency chain, the
    // higher its waves' priority, so the critical chains step at nearly their
    // solo speed while the short fields fill the remaining slots.
    switch (a.prio) {
      case 3: __builtin_amdgcn_s_setprio(3); break;
      case 2: __builtin_amdgcn_s_setprio(2); break;
      case 1: __builtin_amdgcn_s_setprio(1); break;
      default: __builtin_amdgcn_s_setprio(0); break;
    }
    bool wide = a.wide;
    if (a.flags) {
      if (a.flags[0]) continue;  // every item of the field skips: no strip waits on another
      wide = a.flags[1] != 0;
    }
    if (wide) strip_item<COMPRESS, true, 2>(a, item - b.start[lo], lane);
    else strip_item<COMPRESS, false, 2>(a, item - b.start[lo], lane);  // batches are plain offline fields
  }
}

// ---- T = double wavefront ----------------------------------------------------
// The reference's double instantiation predicts and reconstructs in double
// (pred = (float)((L + U) - UL), include/sz_compress_block_processing.hpp:127;
// quantize / recover in double, include/sz_compression_utils.hpp:50-65,
// include/sz_decompression_utils.hpp:63-73), so this strip wavefront carries
// doubles: the same one-wave-per-64-row-strip structure as the float kernel,
// the row above arriving through two 32-bit DPP shifts, the strip's last row
// handed to the next strip through 64-bit values that replace a NaN pattern
// no reconstruction takes (every value is finite: non-finite input is rejected).
constexpr uint64_t kBndEmpty = 0x7FF4DEADBEEF0000ull;

struct Strip2d64 {
  const double* x[2];
  const uint16_t* c[2];  // compress: eb codes; decompress: stream ids
  uint16_t* q[2];
  const double* raw[2];
  double* out[2];
  uint64_t* bnd;  // [2][nstrips][r2], kBndEmpty until published
  int* ticket;
  int r1, r2, nstrips;
};

__device__ __forceinline__ double shr1_d(double src, double lane0) {
  const uint64_t b = (uint64_t)__double_as_longlong(src), z = (uint64_t)__double_as_longlong(lane0);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp((int)(uint32_t)z, (int)(uint32_t)b, 0x138, 0xF, 0xF, false);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp((int)(uint32_t)(z >> 32), (int)(uint32_t)(b >> 32), 0x138,
                                                            0xF, 0xF, false);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

__device__ __forceinline__ double readlane_d(double v, int k) {
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, k);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), k);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// quantize<double> (include/sz_compression_utils.hpp:50-65) for eb code c (0:
// zero precision, c >= 1: precision 4^(c-1) eps)
__device__ __forceinline__ double l2d_compress_step64(float pred, double x, uint32_t c, uint16_t& code) {
  code = 0;
  if (c == 0) return x;
  const double prec = ldexp(1.0, 2 * ((int)c - 1) - 52);
  const double diff = x - (double)pred;
  double qd = fabs(diff) / prec + 1;
  if (qd < 65536) {
    qd = (diff > 0) ? qd : -qd;
    const int qi = (int)(qd / 2) + 32768;
    const double dec = (double)pred + 2 * (qi - 32768) * prec;
    if (fabs(dec - x) <= prec) {
      code = (uint16_t)qi;
      return dec;
    }
  }
  return x;
}

template <bool COMPRESS>
__device__ void strip_item64(const Strip2d64& a, int item, int lane) {
  const int r1 = a.r1, r2 = a.r2;
  const int comp = item & 1, s = item >> 1;
  const int row = s * 64 + lane;
  const bool rvalid = row < r1;
  const size_t roff = (size_t)(rvalid ? row : 0) * r2;
  const bool has_up = s > 0, publish = s + 1 < a.nstrips;
  const uint64_t* bnd_in = a.bnd + ((size_t)comp * a.nstrips + (s > 0 ? s - 1 : 0)) * r2;
  uint64_t* bnd_out = a.bnd + ((size_t)comp * a.nstrips + s) * r2;
  double Rprev = 0, Uprev = 0;
  const int steps = r2 + 63;
  for (int t0 = 0; t0 < steps; t0 += kCH) {
    double inc = 0;
    if (has_up && t0 < r2) {
      const bool want = lane < kCH && t0 + lane < r2;
      uint64_t v = want ? ld_coh64(bnd_in + t0 + lane) : 0;
      while (!__all(!want || v != kBndEmpty)) {
        __builtin_amdgcn_s_sleep(1);
        if (want && v == kBndEmpty) v = ld_coh64(bnd_in + t0 + lane);
      }
      inc = __longlong_as_double((long long)v);
    }
    for (int k = 0; k < kCH && t0 + k < steps; k++) {
      const int j = t0 + k - lane;
      const bool act = rvalid && j >= 0 && j < r2;
      const double lane0 = readlane_d(inc, k);
      const double Uv = shr1_d(Rprev, lane0);
      const float pred = (float)((Rprev + Uv) - Uprev);
      double R;
      if (COMPRESS) {
        uint16_t code;
        R = l2d_compress_step64(pred, act ? a.x[comp][roff + j] : 0.0, act ? a.c[comp][roff + j] : 0u, code);
        if (act) a.q[comp][roff + j] = code;
      } else {
        const uint32_t q = act ? a.q[comp][roff + j] : 0u, sid = act ? a.c[comp][roff + j] : 0u;
        const double prec = sid ? ldexp(1.0, 2 * (int)sid - 52) : 0.0;  // pow(4, id) * DBL_EPSILON
        R = q == 0 ? (act ? a.raw[comp][roff + j] : 0.0) : (double)pred + 2 * ((int)q - 32768) * prec;
        if (act) a.out[comp][roff + j] = R;
      }
      if (publish && lane == 63 && act) st_coh64(bnd_out + j, (uint64_t)__double_as_longlong(R));
      Uprev = Uv;
      Rprev = j < 0 ? 0.0 : R;
    }
  }
}

template <bool COMPRESS>
__global__ void __launch_bounds__(64) k_lorenzo2d_f64(Strip2d64 a) {
  const int lane = threadIdx.x;
  const int items = 2 * a.nstrips;
  for (;;) {
    int item = 0;
    if (lane == 0) item = atomicAdd(a.ticket, 1);
    item = __shfl(item, 0);
    if (item >= items) return;
    strip_item64<COMPRESS>(a, item, lane);
  }
}

__global__ void k_fill_u64(uint64_t* p, size_t n, uint64_t v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_zero_i32(int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// ------------------------------------------------------- huffman sources ----
// m ids: the u codes then the v codes (offline, m = 2n), or the one code per
// vertex of offline_log (m = n)
struct SrcIds2d {
  const uint16_t* cu;  // 256-byte aligned
  const uint16_t* cv;
  size_t n, m;
  static constexpr int kChunk = 512;
  // four consecutive ids per item: one 8-byte load where the four lie in one
  // array at an 8-byte boundary, element loads otherwise
  static constexpr int kMaxItems = kChunk / 4, kPer = 4;
  __device__ int get(size_t c, int idx, uint32_t* sym) const {
    const size_t p0 = c * kChunk + 4 * (size_t)idx;
    uint32_t x[4] = {0, 0, 0, 0};
    int k = 0;
    const uint16_t* src = p0 + 4 <= n ? cu + p0 : (p0 >= n && p0 + 4 <= m ? cv + (p0 - n) : nullptr);
    if (src && ((uintptr_t)src & 7) == 0) {
      const uint2 w = *reinterpret_cast<const uint2*>(src);
      x[0] = w.x & 0xFFFFu;
      x[1] = w.x >> 16;
      x[2] = w.y & 0xFFFFu;
      x[3] = w.y >> 16;
      k = 4;
    } else {
      for (; k < 4 && p0 + k < m; k++) x[k] = p0 + k < n ? cu[p0 + k] : cv[p0 + k - n];
    }
#pragma unroll
    for (int q = 0; q < 4; q++) sym[q] = x[q] ? x[q] - 1 : 0;
    return k;
  }
};

struct SrcCodes2d {
  const uint16_t* q;
  int r1, r2, nby;
  // optional (batched compress): the values the codec ran on, the global
  // exclusive scan of the blocks' zero counts at this component's first
  // block, and where the encoder gathers the unpredictable values
  const float* x = nullptr;
  const uint64_t* uoff = nullptr;
  float* unp = nullptr;
  static constexpr bool kUnpred = true;
  static constexpr int kMaxItems = kBlock2d * kBlock2d, kPer = 1;
  // block row and column (32-bit division when the block index allows it)
  __device__ void block(size_t b, int& bi, int& bj) const {
    if (b < (1ull << 32)) {
      bi = (int)((uint32_t)b / (uint32_t)nby);
      bj = (int)((uint32_t)b - (uint32_t)bi * (uint32_t)nby);
    } else {
      bi = (int)(b / nby);
      bj = (int)(b - (size_t)bi * nby);
    }
  }
  __device__ size_t elem(size_t b, int idx) const {
    int bi, bj;
    block(b, bi, bj);
    const int sy = min(kBlock2d, r2 - bj * kBlock2d);
    const int ii = (int)((uint32_t)idx / (uint32_t)sy), jj = idx - ii * sy;
    return (size_t)(bi * kBlock2d + ii) * r2 + bj * kBlock2d + jj;
  }
  __device__ int get(size_t b, int idx, uint32_t* sym) const {
    int bi, bj;
    block(b, bi, bj);
    const int sx = min(kBlock2d, r1 - bi * kBlock2d), sy = min(kBlock2d, r2 - bj * kBlock2d);
    const int k = min(idx, sx * sy - 1), ii = (int)((uint32_t)k / (uint32_t)sy), jj = k - ii * sy;
    sym[0] = q[(size_t)(bi * kBlock2d + ii) * r2 + bj * kBlock2d + jj];
    return idx < sx * sy;
  }
  template <class F>
  __device__ void each(size_t b, F f) const {
    const int bi = (int)(b / nby), bj = (int)(b - (size_t)bi * nby);
    const int i0 = bi * kBlock2d, j0 = bj * kBlock2d;
    const int sx = min(kBlock2d, r1 - i0), sy = min(kBlock2d, r2 - j0);
    for (int ii = 0; ii < sx; ii++) {
      const uint16_t* row = q + (size_t)(i0 + ii) * r2 + j0;
      for (int jj = 0; jj < sy; jj++) f((uint32_t)row[jj]);
    }
  }
};

struct SinkIds2d {
  static constexpr int kBytes = 2;
  uint16_t* su;
  uint16_t* sv;
  size_t n;
  __device__ void put(uint64_t p, uint32_t s) const {
    if (p < n) su[p] = (uint16_t)s;
    else sv[p - n] = (uint16_t)s;
  }
};

// The id streams when every id symbol is below 256 (the float codec's ids
// stay below 90): staged as bytes, so a block's whole output fits its LDS
// stage and leaves in coalesced stores, and decoded up to 3 codes per step.
struct SinkIds2dB {
  static constexpr int kBytes = 1;
  uint16_t* su;
  uint16_t* sv;
  size_t n;
  __device__ void put(uint64_t p, uint32_t s) const {
    if (p < n) su[p] = (uint16_t)s;
    else sv[p - n] = (uint16_t)s;
  }
};

// block-major position -> row-major element
struct SinkCodes2d {
  static constexpr int kBytes = 2;
  uint16_t* q;
  int r1, r2, nby;
  // optional: zero-filled per-block counts of code 0 (the unpredictable
  // marker), added up while the codes are written
  uint32_t* zc = nullptr;
  // optional: zero-filled per-block masks of the code-0 positions (three
  // words per block, bit k for the block's k-th element in row-major order)
  uint64_t* zm = nullptr;
  __device__ void put(uint64_t p, uint32_t s) const {
    const uint64_t rowblk = (uint64_t)kBlock2d * r2;
    int bi, bj, ii, jj;
    uint32_t k;  // position within the block
    if (p < (1ull << 32) && rowblk < (1ull << 32)) {  // 32-bit divisions (a 64-bit one is a long sequence)
      const uint32_t p32 = (uint32_t)p, rb = (uint32_t)rowblk;
      const uint32_t b = p32 / rb, rem = p32 - b * rb;
      bi = (int)b;
      const uint32_t sx = (uint32_t)min(kBlock2d, r1 - bi * kBlock2d);
      bj = min((int)(rem / (sx * kBlock2d)), nby - 1);
      const uint32_t rem2 = rem - (uint32_t)bj * sx * kBlock2d;
      const uint32_t sy = (uint32_t)min(kBlock2d, r2 - bj * kBlock2d);
      ii = (int)(rem2 / sy);
      jj = (int)(rem2 - (uint32_t)ii * sy);
      k = rem2;
    } else {
      bi = (int)(p / rowblk);
      const uint64_t rem = p - (uint64_t)bi * rowblk;
      const int sx = min(kBlock2d, r1 - bi * kBlock2d);
      bj = (int)(rem / ((uint64_t)sx * kBlock2d));
      if (bj > nby - 1) bj = nby - 1;
      const uint32_t rem2 = (uint32_t)(rem - (uint64_t)bj * sx * kBlock2d);
      const int sy = min(kBlock2d, r2 - bj * kBlock2d);
      ii = rem2 / sy;
      jj = rem2 - ii * sy;
      k = rem2;
    }
    q[(size_t)(bi * kBlock2d + ii) * r2 + bj * kBlock2d + jj] = (uint16_t)s;
    if (zm && s == 0)
      atomicOr((unsigned long long*)&zm[3 * ((size_t)bi * nby + bj) + (k >> 6)], 1ull << (k & 63));
    if (zc) {
      // one atomic per block among the active lanes' zeros: consecutive
      // positions (the staged writes) fall in one or two blocks
      const uint32_t b = (uint32_t)bi * (uint32_t)nby + (uint32_t)bj;
      const bool z = s == 0;
      uint64_t m = __ballot(z);
      while (m) {
        const int leader = __ffsll((long long)m) - 1;
        const uint32_t bl = (uint32_t)__shfl((int)b, leader);
        const uint64_t same = __ballot(z && b == bl);
        if ((int)(threadIdx.x & 63) == leader) atomicAdd(&zc[bl], (uint32_t)__popcll(same));
        m &= ~same;
      }
    }
  }
};

// Unpredictable values in block-major order. One wave per 12x12 block (4
// blocks per 256-thread workgroup); its lanes take the block's elements in
// row-major order, 64 at a time, so a row's 12 codes are read by consecutive
// lanes, and the ranks of the zero codes come from a ballot.
constexpr int kUnpWaves = 4;
inline int unp_grid(size_t nb) { return (int)div_up(nb, (size_t)kUnpWaves); }

struct BlockView {
  int i0, j0, sy, m;
  __device__ BlockView(size_t b, int r1, int r2, int nby) {
    const int bi = (int)(b / nby), bj = (int)(b - (size_t)bi * nby);
    i0 = bi * kBlock2d;
    j0 = bj * kBlock2d;
    const int sx = min(kBlock2d, r1 - i0);
    sy = min(kBlock2d, r2 - j0);
    m = sx * sy;
  }
  __device__ size_t elem(int k, int r2) const {
    const int ii = k / sy, jj = k - ii * sy;
    return (size_t)(i0 + ii) * r2 + j0 + jj;
  }
};

__global__ void __launch_bounds__(256) k_block_zeros(const uint16_t* __restrict__ q, int r1, int r2, int nby,
                                                     size_t nb, uint32_t* __restrict__ cnt) {
  const size_t b = (size_t)blockIdx.x * kUnpWaves + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (b >= nb) return;  // uniform across the wave
  const BlockView v(b, r1, r2, nby);
  uint32_t z = 0;
  for (int k0 = 0; k0 < v.m; k0 += 64) {
    const int k = k0 + lane;
    const bool zero = k < v.m && q[v.elem(k, r2)] == 0;
    z += (uint32_t)__popcll(__ballot(zero));
  }
  if (lane == 0) cnt[b] = z;
}

template <typename T>
__global__ void __launch_bounds__(256) k_gather_unpred(const uint16_t* __restrict__ q, const T* __restrict__ x,
                                                       int r1, int r2, int nby, size_t nb,
                                                       const uint64_t* __restrict__ off, T* __restrict__ out) {
  const size_t b = (size_t)blockIdx.x * kUnpWaves + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (b >= nb) return;
  const BlockView v(b, r1, r2, nby);
  uint64_t o = off[b];
  const uint64_t below = (1ull << lane) - 1;
  for (int k0 = 0; k0 < v.m; k0 += 64) {
    const int k = k0 + lane;
    const size_t e = k < v.m ? v.elem(k, r2) : 0;
    const bool zero = k < v.m && q[e] == 0;
    const uint64_t mask = __ballot(zero);
    if (zero) out[o + __popcll(mask & below)] = x[e];
    o += __popcll(mask);
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_scatter_unpred(const uint16_t* __restrict__ q, const T* __restrict__ vals,
                                                        int r1, int r2, int nby, size_t nb,
                                                        const uint64_t* __restrict__ off, uint64_t count,
                                                        T* __restrict__ raw, int* __restrict__ err) {
  const size_t b = (size_t)blockIdx.x * kUnpWaves + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (b >= nb) return;
  const BlockView v(b, r1, r2, nby);
  uint64_t o = off[b];
  const uint64_t below = (1ull << lane) - 1;
  for (int k0 = 0; k0 < v.m; k0 += 64) {
    const int k = k0 + lane;
    const size_t e = k < v.m ? v.elem(k, r2) : 0;
    const bool zero = k < v.m && q[e] == 0;
    const uint64_t mask = __ballot(zero);
    if (zero) {
      const uint64_t r = o + __popcll(mask & below);
      if (r < count) raw[e] = vals[r];
      else *err = 1;
    }
    o += __popcll(mask);
  }
}

// offline_log transform (include/sz_cp_preserve_utils.hpp:205-221): host-libm
// log2f of |x| (-100 for 0) and the MSB-first sign bitmaps (1 = positive):
// one wave per 64 values; reversing the ballot's 64 bits and then its byte
// order puts value 8k+j at bit 7-j of byte k. Block y = component.
__global__ void __launch_bounds__(256) k_log2d(const float* __restrict__ U, const float* __restrict__ V, size_t n,
                                               size_t nq, LibmView lv, float* __restrict__ LU,
                                               float* __restrict__ LV, uint64_t* __restrict__ smap) {
  const int c = blockIdx.y, lane = threadIdx.x & 63;
  const float* X = c ? V : U;
  float* L = c ? LV : LU;
  const size_t q = (size_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (q >= nq) return;  // uniform across the wave
  const size_t e = q * 64 + lane;
  const float x = e < n ? X[e] : 0.f;
  if (e < n) L[e] = x != 0 ? host_log2f(lv, fabsf(x)) : -100.f;
  const uint64_t m = __ballot(x > 0);
  if (lane == 0) smap[c * nq + q] = __builtin_bswap64(__builtin_bitreverse64(m));
}

// offline_log decode, in place (src/sz_decompress_cp_preserve_2d.cpp:80-92):
// 0 below -99, else +-exp2f (host libm) by the sign bitmap.
__global__ void __launch_bounds__(256) k_finish2d(float* __restrict__ U, float* __restrict__ V, size_t n,
                                                  const uint8_t* __restrict__ signs, size_t smap, LibmView ev) {
  const int c = blockIdx.y;
  float* X = c ? V : U;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float lx = X[e];
  const bool pos = (signs[c * smap + e / 8] >> (7 - (e % 8))) & 1;
  float y = 0.f;
  if (!(lx < -99)) {
    const float t = host_exp2f(ev, lx);
    y = pos ? t : -t;
  }
  X[e] = y;
}

struct U16Hist {
  const uint16_t* p;  // 8-byte aligned (the code arrays are 256-byte aligned)
  size_t n;
  static constexpr int kChunk = 1024;
  // four codes per item, one 8-byte load
  static constexpr int kMaxItems = kChunk / 4, kPer = 4;
  __device__ int get(size_t c, int idx, uint32_t* sym) const {
    const size_t i = c * kChunk + 4 * (size_t)idx;
    if (i + 4 <= n) {
      const uint2 w = *reinterpret_cast<const uint2*>(p + i);
      sym[0] = w.x & 0xFFFFu;
      sym[1] = w.x >> 16;
      sym[2] = w.y & 0xFFFFu;
      sym[3] = w.y >> 16;
      return 4;
    }
    int k = 0;
    for (; k < 4 && i + k < n; k++) sym[k] = p[i + k];
    for (int q = k; q < 4; q++) sym[q] = 0;
    return k;
  }
};

// ---- batched entropy stage: many fields' streams per launch -----------------
// One component of one field for the per-block unpredictable passes
// (kUnpWaves 12x12 blocks per workgroup, from workgroup blk0).
struct UnpJob {
  const uint16_t* q;
  const float* x;        // gather: values the codec ran on; scatter: the stream's unpredictable values
  int r1, r2, nby;
  size_t nb;
  uint32_t* cnt;         // per-block zero counts
  const uint64_t* off;   // global exclusive scan of the counts, at this component's first block
  float* out;            // gather: this component's unpredictable values; scatter: per-element values
  uint64_t count;        // scatter: values in the stream
  int* err;              // scatter: set when the codes ask for more values than the stream holds
  uint32_t blk0;
  const uint64_t* zm;    // scatter: per-block code-0 masks (SinkCodes2d::zm), else the codes are read
};

__global__ void __launch_bounds__(256) k_block_zeros_multi(const UnpJob* __restrict__ jobs, int nj) {
  const UnpJob& J = jobs[job_of(jobs, nj, blockIdx.x)];
  const size_t b = (size_t)(blockIdx.x - J.blk0) * kUnpWaves + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (b >= J.nb) return;  // uniform across the wave
  const BlockView v(b, J.r1, J.r2, J.nby);
  uint32_t z = 0;
  for (int k0 = 0; k0 < v.m; k0 += 64) {
    const int k = k0 + lane;
    const bool zero = k < v.m && J.q[v.elem(k, J.r2)] == 0;
    z += (uint32_t)__popcll(__ballot(zero));
  }
  if (lane == 0) J.cnt[b] = z;
}

__global__ void __launch_bounds__(256) k_scatter_unpred_multi(const UnpJob* __restrict__ jobs, int nj) {
  const UnpJob& J = jobs[job_of(jobs, nj, blockIdx.x)];
  const size_t b = (size_t)(blockIdx.x - J.blk0) * kUnpWaves + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (b >= J.nb) return;
  const BlockView v(b, J.r1, J.r2, J.nby);
  uint64_t o = J.off[b] - J.off[0];
  const uint64_t below = (1ull << lane) - 1;
  for (int k0 = 0; k0 < v.m; k0 += 64) {
    const int k = k0 + lane;
    const size_t e = k < v.m ? v.elem(k, J.r2) : 0;
    const bool zero = J.zm ? ((J.zm[3 * b + (k0 >> 6)] >> lane) & 1) != 0 : (k < v.m && J.q[e] == 0);
    const uint64_t mask = __ballot(zero);
    if (zero) {
      const uint64_t r = o + __popcll(mask & below);
      if (r < J.count) J.out[e] = J.x[r];
      else *J.err = 1;
    }
    o += __popcll(mask);
  }
}

// The present symbols of each histogram (one 1024-thread workgroup per
// histogram), in symbol order, packed into one list: histogram h gets
// count[h] entries from base[h].
struct CompactJob {
  const uint32_t* hist;
  uint32_t nbins;
};

__global__ void __launch_bounds__(1024) k_compact_hist(const CompactJob* __restrict__ jobs, uint32_t* __restrict__ count,
                                                       uint64_t* __restrict__ base,
                                                       unsigned long long* __restrict__ total,
                                                       uint32_t* __restrict__ syms, uint32_t* __restrict__ freqs) {
  __shared__ uint32_t wsum[16];
  __shared__ uint64_t b0;
  const CompactJob J = jobs[blockIdx.x];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint32_t c = 0;
  for (uint32_t i = threadIdx.x; i < J.nbins; i += 1024) c += J.hist[i] != 0;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if (lane == 0) wsum[wv] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t k = 0;
    for (int x = 0; x < 16; x++) k += wsum[x];
    const uint64_t at = atomicAdd(total, (unsigned long long)k);
    count[blockIdx.x] = k;
    base[blockIdx.x] = at;
    b0 = at;
  }
  __syncthreads();
  uint64_t pos = b0;
  const uint64_t below = (1ull << lane) - 1;
  for (uint32_t i0 = 0; i0 < J.nbins; i0 += 1024) {
    const uint32_t i = i0 + threadIdx.x;
    const uint32_t f = i < J.nbins ? J.hist[i] : 0u;
    const uint64_t m = __ballot(f != 0);
    if (lane == 0) wsum[wv] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (int x = 0; x < 16; x++) {
      const uint32_t y = wsum[x];
      before += x < wv ? y : 0u;
      all += y;
    }
    if (f) {
      const uint64_t at = pos + before + (uint32_t)__popcll(m & below);
      syms[at] = i;
      freqs[at] = f;
    }
    pos += all;
    __syncthreads();
  }
}

// Codebooks of many trees into their dense tables (one workgroup per tree).
struct CbJob {
  const uint32_t* syms;
  const uint64_t* codes;
  const uint8_t* lens;
  uint32_t k;
  uint64_t* code;
  uint8_t* len;
};

__global__ void __launch_bounds__(256) k_scatter_codebooks(const CbJob* __restrict__ jobs) {
  const CbJob& J = jobs[blockIdx.x];
  for (uint32_t i = threadIdx.x; i < J.k; i += blockDim.x) {
    J.code[J.syms[i]] = J.codes[i];
    J.len[J.syms[i]] = J.lens[i];
  }
}

// A host-built tree moved to where the device tree builder puts its results:
// codebook scattered into the dense tables, serialized bytes, info words.
struct HostTreeJob {
  CbJob cb;
  const uint8_t* tree_src;
  uint8_t* tree_dst;
  uint64_t tree_n;
  uint64_t* info_dst;
  uint64_t info[4];
};

__global__ void __launch_bounds__(256) k_host_trees_in(const HostTreeJob* __restrict__ jobs) {
  const HostTreeJob& J = jobs[blockIdx.x];
  for (uint32_t i = threadIdx.x; i < J.cb.k; i += blockDim.x) {
    J.cb.code[J.cb.syms[i]] = J.cb.codes[i];
    J.cb.len[J.cb.syms[i]] = J.cb.lens[i];
  }
  for (uint64_t i = threadIdx.x; i < J.tree_n; i += blockDim.x) J.tree_dst[i] = J.tree_src[i];
  if (threadIdx.x < 4) J.info_dst[threadIdx.x] = J.info[threadIdx.x];
}

// out[j] = a[j].hi - a[j].lo: per-stream sums read off global exclusive scans.
struct ScanDiff {
  const uint64_t* lo;
  const uint64_t* hi;
};

__global__ void k_scan_diffs(const ScanDiff* __restrict__ d, int n, uint64_t* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) out[j] = *d[j].hi - *d[j].lo;
}

// Stream assembly: one workgroup per piece (at most 64 KiB) copies len bytes
// to out + dst. The destination's aligned words are stored whole, each built
// from the two source words that cover it (alignbyte); the at most 3 bytes
// before and after them go byte by byte, so neighbouring pieces never share a
// word store.
struct CopyPiece {
  uint64_t dst;
  const uint8_t* src;
  uint32_t len;
  uint32_t pad;
};

__device__ __forceinline__ void copy_piece(uint8_t* __restrict__ d, const uint8_t* __restrict__ s, uint32_t n) {
  const uint32_t head = min(n, (uint32_t)((4u - ((uint32_t)(uintptr_t)d & 3u)) & 3u));
  const uint32_t nw = (n - head) >> 2, tail0 = head + 4 * nw;
  if (threadIdx.x < head) d[threadIdx.x] = s[threadIdx.x];
  if (threadIdx.x < n - tail0) d[tail0 + threadIdx.x] = s[tail0 + threadIdx.x];
  uint32_t* dw = (uint32_t*)(d + head);
  const uint8_t* sp = s + head;
  const uint32_t mis = (uint32_t)(uintptr_t)sp & 3u;
  const uint32_t* sw = (const uint32_t*)(sp - mis);
  if (mis == 0) {
    for (uint32_t k = threadIdx.x; k < nw; k += blockDim.x) dw[k] = sw[k];
  } else {
    for (uint32_t k = threadIdx.x; k < nw; k += blockDim.x)
      dw[k] = __builtin_amdgcn_alignbyte(sw[k + 1], sw[k], mis);
  }
}

__global__ void __launch_bounds__(256) k_assemble(const CopyPiece* __restrict__ pieces, uint8_t* __restrict__ out) {
  const CopyPiece P = pieces[blockIdx.x];
  copy_piece(out + P.dst, P.src, P.len);
}

// ---- device-side layout of a batch's streams: every size and offset follows
// from the per-stream totals and the device-built trees, so no host round trip
// sits between the scans and the assembled streams.
struct Layout2d {
  // set by the host
  uint64_t off;              // stream offset in the batch output
  uint64_t cap;              // bytes available there
  uint64_t nb;               // 12x12 blocks
  const uint64_t* tot;       // bits of the eb/u/v streams, unpredictable u/v counts
  const uint64_t* tinfo[3];  // device tree info of the eb/u/v trees (node count, tree bytes, ...)
  const uint8_t* tree[3];
  EncJob<SrcIds2d>* ej_ids;  // jobs whose word arrays k_layout2d fills in
  EncJob<SrcCodes2d>* ej_codes;  // [2], also their unpredictable-value arrays
  // set by k_layout2d
  uint32_t* words;  // eb, u, v word arrays, consecutive
  float* unp;       // u then v unpredictable values
  uint64_t bytes[3], wn[3], unp_n[2], size;
  uint64_t at[11];  // hdr, h_eb, eb bits, hu1, unp u, hu2, u bits, hv1, unp v, hv2, v bits
};

// One thread: sizes, offsets, the word and value arrays of every field (in
// the preallocated pools: *err set when they would overflow), sizes[k].
__global__ void k_layout2d(Layout2d* __restrict__ L, int nf, uint32_t* words, uint64_t words_cap, float* unp,
                           uint64_t unp_cap, uint64_t* __restrict__ sizes, uint64_t* __restrict__ words_used,
                           int* __restrict__ err) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint64_t wpos = 0, upos = 0;
  for (int k = 0; k < nf; k++) {
    Layout2d& F = L[k];
    uint64_t tb[3];
    for (int t = 0; t < 3; t++) {
      F.bytes[t] = (F.tot[t] + 7) / 8;
      F.wn[t] = (F.bytes[t] + 3) / 4 + 2;
      tb[t] = F.tinfo[t][1];
    }
    F.unp_n[0] = F.tot[3];
    F.unp_n[1] = F.tot[4];
    const uint64_t ind = (F.nb - 1) / 8 + 1;
    const uint64_t h_eb = 8 + 4 + tb[0] + 8, h1 = 4 + 1 + 4 + 8 + 8 + 8;
    const uint64_t hu2 = ind + 8 + 4 + tb[1] + 8, hv2 = ind + 8 + 4 + tb[2] + 8;
    uint64_t p = F.off;
    const uint64_t sec[11] = {36, h_eb, F.bytes[0], h1, 4 * F.unp_n[0], hu2, F.bytes[1], h1, 4 * F.unp_n[1], hv2,
                              F.bytes[2]};
    for (int q = 0; q < 11; q++) {
      F.at[q] = p;
      p += sec[q];
    }
    F.size = p - F.off;
    sizes[k] = F.size;
    const uint64_t wneed = (F.wn[0] + F.wn[1] + F.wn[2] + 63) & ~63ull;
    const uint64_t uneed = (F.unp_n[0] + F.unp_n[1] + 63) & ~63ull;
    if (F.size > F.cap || wpos + wneed > words_cap || upos + uneed > unp_cap) {
      // this and the remaining fields keep null arrays: the passes that
      // follow skip them, and the host reports the error
      *err = 1;
      break;
    }
    F.words = words + wpos;
    F.unp = unp + upos;
    wpos += wneed;
    upos += uneed;
    F.ej_ids->words = F.words;
    F.ej_codes[0].words = F.words + F.wn[0];
    F.ej_codes[1].words = F.words + F.wn[0] + F.wn[1];
    F.ej_codes[0].src.unp = F.unp;
    F.ej_codes[1].src.unp = F.unp + F.unp_n[0];
  }
  *words_used = wpos;
}

__global__ void k_fill_words(uint32_t* __restrict__ w, const uint64_t* __restrict__ n) {
  const uint64_t m = *n;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x)
    w[i] = 0;
}

__device__ __forceinline__ void put_bytes(uint8_t* p, uint64_t v, int nbytes) {
  for (int b = 0; b < nbytes; b++) p[b] = (uint8_t)(v >> (8 * b));
}

// One workgroup per field: the stream header, the id tree head, the two
// compressor heads and the two code tree heads (all-zero block indicators
// first), trees copied from the device builds (the host layout of
// src/sz_compress_cp_preserve_2d.cpp:258-269, src/sz_compress_3d.cpp:219-228).
__global__ void __launch_bounds__(256) k_write_heads2d(const Layout2d* __restrict__ L, uint8_t* __restrict__ out) {
  const Layout2d& F = L[blockIdx.x];
  if (!F.words) return;  // not laid out (an overflow the host reports)
  const uint64_t ind = (F.nb - 1) / 8 + 1;
  if (threadIdx.x == 0) {
    uint8_t* h = out + F.at[0];
    put_bytes(h, 4, 4);
    put_bytes(h + 4, 0x3CB0000000000000ull, 8);  // DBL_EPSILON
    put_bytes(h + 12, F.bytes[0] + 8 + 4 + F.tinfo[0][1] + 8, 8);
    put_bytes(h + 20, F.at[7] - F.at[3], 8);
    put_bytes(h + 28, F.off + F.size - F.at[7], 8);
    for (int c = 0; c < 2; c++) {
      uint8_t* q = out + F.at[c ? 7 : 3];
      put_bytes(q, kBlock2d, 4);
      q[4] = 0;
      put_bytes(q + 5, kRadius, 4);
      put_bytes(q + 9, 0, 8);   // meanInfo<float>{false, 0}
      put_bytes(q + 17, 0, 8);  // reg_count
      put_bytes(q + 25, F.unp_n[c], 8);
    }
  }
  for (int t = 0; t < 3; t++) {
    uint8_t* q = out + (t == 0 ? F.at[1] : F.at[t == 1 ? 5 : 9] + ind);
    const uint64_t tb = F.tinfo[t][1];
    if (threadIdx.x == 0) {
      put_bytes(q, F.tinfo[t][0], 8);
      put_bytes(q + 8, tb, 4);
      put_bytes(q + 12 + tb, F.bytes[t], 8);
    }
    for (uint64_t b = threadIdx.x; b < tb; b += blockDim.x) q[12 + b] = F.tree[t][b];
  }
  for (uint64_t b = threadIdx.x; b < ind; b += blockDim.x) {
    out[F.at[5] + b] = 0;
    out[F.at[9] + b] = 0;
  }
}

// The five device sections of every field (bitstreams and unpredictable
// values) into place, 64 KiB pieces spread over the workgroups of the
// field's grid row.
__global__ void __launch_bounds__(256) k_assemble2d(const Layout2d* __restrict__ L, uint8_t* __restrict__ out) {
  const Layout2d& F = L[blockIdx.y];
  if (!F.words) return;
  constexpr uint64_t kP = 65536;
  const uint8_t* src[5] = {(const uint8_t*)F.words, (const uint8_t*)F.unp, (const uint8_t*)(F.words + F.wn[0]),
                           (const uint8_t*)(F.unp + F.unp_n[0]), (const uint8_t*)(F.words + F.wn[0] + F.wn[1])};
  const uint64_t len[5] = {F.bytes[0], 4 * F.unp_n[0], F.bytes[1], 4 * F.unp_n[1], F.bytes[2]};
  const uint64_t dst[5] = {F.at[2], F.at[4], F.at[6], F.at[8], F.at[10]};
  uint64_t p = blockIdx.x;
  for (int q = 0; q < 5; q++) {
    const uint64_t np = (len[q] + kP - 1) / kP;
    for (; p < np; p += gridDim.x) {
      const uint64_t o = p * kP;
      copy_piece(out + dst[q] + o, src[q] + o, (uint32_t)min(kP, len[q] - o));
    }
    p -= np;
  }
}

}  // namespace

// ---------------------------------------------------------------- host ----
namespace {

template <typename V>
void put(std::vector<uint8_t>& o, const V& v) {
  const size_t at = o.size();
  o.resize(at + sizeof(V));
  std::memcpy(o.data() + at, &v, sizeof(V));
}

struct Piece {  // a contiguous part of the output stream
  size_t off;
  const void* src;  // host (staging) or device pointer
  size_t len;
  bool device;
};

// Host-written parts of a 2D stream (src/sz_compress_cp_preserve_2d.cpp:258-269,
// src/sz_compress_3d.cpp:219-228): the id section's tree head, the two
// components' compressor heads (block size, radius, meanInfo<T>, unpredictable
// count) and tree heads (the all-zero block indicator bytes first), and the
// stream header with the three section sizes. vb: bytes of T.
// One pinned staging area mirrored on the device: tables are filled on the
// host and sent with one copy; take() returns the device address of a slot and
// its host address through `host`.
struct Arena {
  uint8_t* h = nullptr;
  uint8_t* d = nullptr;
  size_t used = 0, cap = 0;
  template <class T>
  T* take(size_t count, T** host) {
    used = (used + 15) & ~(size_t)15;
    CPSZ_REQUIRE(used + count * sizeof(T) <= cap, kErrInternal, "staging table overflow");
    *host = (T*)(h + used);
    T* dp = (T*)(d + used);
    used += count * sizeof(T);
    return dp;
  }
};

struct Heads2d {
  std::vector<uint8_t> hdr, h_eb, hu1, hu2, hv1, hv2;
  size_t eb_size = 0, u_size = 0, v_size = 0;
};

void heads_2d(const HuffEncodeTable& t_eb, const HuffEncodeTable& t_u, const HuffEncodeTable& t_v, size_t bytes_eb,
              size_t bytes_u, size_t bytes_v, uint64_t unp_u, uint64_t unp_v, size_t nb, bool log, size_t vb,
              Heads2d& H) {
  const size_t ind_bytes = (nb - 1) / 8 + 1;
  auto huff_head = [](std::vector<uint8_t>& h, const HuffEncodeTable& t, size_t nbytes) {
    put(h, (size_t)t.node_count);
    put(h, (uint32_t)t.tree.size());
    h.insert(h.end(), t.tree.begin(), t.tree.end());
    put(h, (size_t)nbytes);
  };
  auto comp_head = [&](std::vector<uint8_t>& h, uint64_t unp_count) {
    put(h, (int)kBlock2d);
    put(h, (char)0);
    put(h, (int)kRadius);
    put(h, (uint64_t)0);  // meanInfo<T>{false, 0}: flag byte, padding bytes, T 0 (8 or 16 bytes)
    if (vb == 8) put(h, (uint64_t)0);
    put(h, (size_t)0);    // reg_count
    put(h, (size_t)unp_count);
  };
  H.hdr.clear();
  H.h_eb.clear();
  H.hu1.clear();
  H.hv1.clear();
  huff_head(H.h_eb, t_eb, bytes_eb);
  comp_head(H.hu1, unp_u);
  comp_head(H.hv1, unp_v);
  H.hu2.assign(ind_bytes, 0);
  H.hv2.assign(ind_bytes, 0);
  huff_head(H.hu2, t_u, bytes_u);
  huff_head(H.hv2, t_v, bytes_v);
  H.eb_size = H.h_eb.size() + bytes_eb;
  H.u_size = H.hu1.size() + vb * unp_u + H.hu2.size() + bytes_u;
  H.v_size = H.hv1.size() + vb * unp_v + H.hv2.size() + bytes_v;
  put(H.hdr, (int)(log ? 2 : 4));
  put(H.hdr, (double)DBL_EPSILON);
  put(H.hdr, H.eb_size);
  put(H.hdr, H.u_size);
  put(H.hdr, H.v_size);
}

// Lorenzo wavefront for fields f[0..nf): one launch; a single field keeps the
// plain kernel, several share one persistent launch.
template <bool COMPRESS>
void launch_wavefront(Ctx& ctx, std::vector<Strip2d>& f, int grid_cap) {
  hipStream_t st = ctx.stream();
  if (f.size() == 1 && !f[0].flags) {
    Strip2d& sa = f[0];
    const int nz = 2 * sa.nstrips + 1;
    k_zero_i32<<<(int)div_up(nz, 256), 256, 0, st>>>(sa.prog, nz);
    const size_t nb = (size_t)2 * sa.nstrips * sa.r2 * 2;  // tagged boundary rows as 32-bit words
    k_zero_i32<<<(int)div_up(nb, 256), 256, 0, st>>>((int*)sa.bnd, (int)nb);
    const int items = 2 * sa.nstrips;
    const int grid = std::min(items, grid_cap);
    if (sa.wide) k_lorenzo2d<COMPRESS, true, 2><<<grid, 64, 0, st>>>(sa);
    else if (sa.shift == 1) k_lorenzo2d<COMPRESS, false, 1><<<grid, 64, 0, st>>>(sa);
    else k_lorenzo2d<COMPRESS, false, 2><<<grid, 64, 0, st>>>(sa);
    CPSZ_HIP(hipGetLastError());
    return;
  }
  // Tickets go out field by field, the fields with the longest wavefront
  // (one strip's hand-off lag per strip plus the row length) first, so the
  // tallest fields are not left waiting for a free wave at the end.
  const size_t nf = f.size();
  CPSZ_REQUIRE(nf <= 65535, kErrUnsupported, "2D batch: at most 65535 fields per call");
  for (const Strip2d& x : f) CPSZ_REQUIRE(x.shift == 2, kErrInternal, "2D batch wavefront: plain offline fields only");
  std::vector<Strip2d> g(f);
  std::stable_sort(g.begin(), g.end(), [](const Strip2d& a, const Strip2d& b) {
    return (int64_t)a.nstrips * 96 + a.r2 > (int64_t)b.nstrips * 96 + b.r2;
  });
  std::vector<int> start(nf + 1, 0);
  size_t max_words = 0;
  const double lmax = (double)g[0].nstrips * 96 + g[0].r2;
  for (size_t k = 0; k < nf; k++) {
    const double l = ((double)g[k].nstrips * 96 + g[k].r2) / lmax;
    g[k].prio = l >= 0.75 ? 3 : (l >= 0.5 ? 2 : (l >= 0.25 ? 1 : 0));
    start[k + 1] = start[k] + 2 * g[k].nstrips;
    max_words = std::max(max_words, (size_t)2 * g[k].nstrips * g[k].r2 * 2);
  }
  const size_t tbytes = nf * sizeof(Strip2d), sbytes = (nf + 1) * sizeof(int);
  uint8_t* d = ctx.buf<uint8_t>("w2b.tab", tbytes + sbytes + 64);
  int* ticket = (int*)(d + tbytes + sbytes + 16 - ((tbytes + sbytes) & 15));
  uint8_t* h = (uint8_t*)ctx.host_table("w2b.tab", tbytes + sbytes);
  std::memcpy(h, g.data(), tbytes);
  std::memcpy(h + tbytes, start.data(), sbytes);
  ctx.upload_table("w2b.tab", d, h, tbytes + sbytes);
  k_zero_fields<<<dim3((unsigned)std::min<size_t>(div_up(max_words, 256 * 8), 64), (unsigned)nf), 256, 0, st>>>(
      (const Strip2d*)d, ticket);
  StripBatch b{(const Strip2d*)d, (const int*)(d + tbytes), (int)nf, start[nf], ticket};
  const int grid = std::min(start[nf], grid_cap);
  k_lorenzo2d_batch<COMPRESS><<<grid, 64, 0, st>>>(b);
  CPSZ_HIP(hipGetLastError());
}

size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

// ------------------------------------------------------------- staged form --
size_t c2_scratch_bytes(size_t r1, size_t r2) {
  const size_t n = r1 * r2, ns = div_up(r1, 64);
  return 4 * al256(2 * n) + al256((kMaxEbCode + 2 * 65536) * 4) + al256(16) + al256(2 * ns * r2 * 8) +
         al256((2 * ns + 1) * 4);
}

void c2_bind(C2Work& w, uint8_t* p) {
  const size_t n = w.r1 * w.r2, ns = div_up(w.r1, 64);
  auto take = [&](size_t bytes) {
    uint8_t* q = p;
    p += al256(bytes);
    return q;
  };
  w.cu = (uint16_t*)take(2 * n);
  w.cv = (uint16_t*)take(2 * n);
  w.qu = (uint16_t*)take(2 * n);
  w.qv = (uint16_t*)take(2 * n);
  w.hist = (uint32_t*)take((kMaxEbCode + 2 * 65536) * 4);
  w.flags = (int*)take(16);
  w.bnd = (uint64_t*)take(2 * ns * w.r2 * 8);
  w.prog = (int*)take((2 * ns + 1) * 4);
}

namespace {
std::atomic<bool> g_exact_bounds{false};
}
void set_exact_2d_bounds(bool on) { g_exact_bounds.store(on); }
bool exact_2d_bounds() { return g_exact_bounds.load(); }

namespace {
// The fast bounds pass over a job table already on its way to the device
// (blk tiles), then the FP64 fix-up of the vertices it marked. Graph-capture
// safe: the list counter is cleared by a kernel and nothing waits.
void launch_eb_fast(Ctx& ctx, const Eb2dJob* dj, int nj, uint32_t blk, double max_pwr_eb) {
  hipStream_t st = ctx.stream();
  uint32_t* fx = ctx.buf<uint32_t>("eb.fix", (size_t)blk + 1);  // [0] count, [1..] tiles
  k_fill_u32<<<1, 64, 0, st>>>(fx, 1, 0);
  k_eb2d_multi<true><<<blk, 256, 0, st>>>(dj, nj, max_pwr_eb, ctx.d_q2d(), ctx.q2d_size(), fx + 1, fx);
  // a listed tile costs a few dependent global loads per marked vertex: enough
  // blocks to keep every CU's slots filled (70 VGPRs: 7 blocks per CU)
  k_eb2d_fix<<<std::min<uint32_t>(blk, 2048), 256, 0, st>>>(dj, nj, fx + 1, fx, max_pwr_eb, ctx.d_q2d(),
                                                            ctx.q2d_size());
}
}  // namespace

void c2_eb(Ctx& ctx, C2Work& w, double max_pwr_eb, const EbOverride& ov) {
  CPSZ_REQUIRE((w.f64 ? w.dU64 && w.dV64 : w.dU && w.dV) && w.r1 >= 1 && w.r2 >= 1, kErrArg,
               "compress_2d: bad arguments");
  CPSZ_REQUIRE(w.r1 < (1u << 30) && w.r2 < (1u << 30) && w.r1 * w.r2 < (1ull << 40), kErrUnsupported,
               "compress_2d: field too large");
  hipStream_t st = ctx.stream();
  CPSZ_HIP(hipMemsetAsync(w.hist, 0, (kMaxEbCode + 2 * 65536) * sizeof(uint32_t), st));
  CPSZ_HIP(hipMemsetAsync(w.flags, 0, 4 * sizeof(int), st));
  dim3 grid((unsigned)div_up(w.r2, kTW), (unsigned)div_up(w.r1, kTH));
  if (w.f64) {
    k_eb2d<double><<<grid, 256, 0, st>>>(w.dU64, w.dV64, (int)w.r1, (int)w.r2, max_pwr_eb, ctx.d_q2d(),
                                         ctx.q2d_size(), w.cu, w.cv, w.hist, w.flags, ov, nullptr, 0);
  } else if (!w.log && !has_override(ov) && !exact_2d_bounds()) {
    Eb2dJob J{};
    J.U = w.dU;
    J.V = w.dV;
    J.r1 = (int)w.r1;
    J.r2 = (int)w.r2;
    J.gx = (int)grid.x;
    J.cu = w.cu;
    J.cv = w.cv;
    J.hist = w.hist;
    J.flags = w.flags;
    J.blk0 = 0;
    const size_t tiles = (size_t)grid.x * grid.y;
    CPSZ_REQUIRE(tiles < (1ull << 31), kErrUnsupported, "compress_2d: field too large");
    Eb2dJob* dj = ctx.buf<Eb2dJob>("c2eb.job", 1);
    Eb2dJob* hj = (Eb2dJob*)ctx.host_table("c2eb.job", sizeof(Eb2dJob));
    *hj = J;
    ctx.upload_table("c2eb.job", dj, hj, sizeof(Eb2dJob));
    launch_eb_fast(ctx, dj, 1, (uint32_t)tiles, max_pwr_eb);
  } else {
    k_eb2d<float><<<grid, 256, 0, st>>>(w.dU, w.dV, (int)w.r1, (int)w.r2, max_pwr_eb, ctx.d_q2d(), ctx.q2d_size(),
                                        w.cu, w.cv, w.hist, w.flags, ov, w.log ? ctx.d_q2l() : nullptr,
                                        ctx.q2l_size());
  }
  CPSZ_HIP(hipGetLastError());
}

// Bounds of every field (plain offline, float) in one launch: the fields'
// histograms and flags move to two contiguous arrays, cleared at once. The
// flags stay on the device (the wavefront reads them there).
void c2_eb_batch(Ctx& ctx, C2Work* const* w, size_t nf, double max_pwr_eb, const std::vector<EbOverride>* ov) {
  hipStream_t st = ctx.stream();
  constexpr size_t HN = kMaxEbCode + 2 * 65536;
  uint32_t* hist = ctx.buf<uint32_t>("b2eb.hist", nf * HN);
  int* flags = ctx.buf<int>("b2eb.flags", 4 * nf);
  std::vector<Eb2dJob> jobs(nf);
  uint32_t blk = 0;
  bool approx = !exact_2d_bounds();
  for (size_t k = 0; k < nf; k++) {
    C2Work& W = *w[k];
    CPSZ_REQUIRE(W.dU && W.dV && W.r1 >= 1 && W.r2 >= 1 && !W.log && !W.f64, kErrArg, "compress_2d: bad arguments");
    CPSZ_REQUIRE(W.r1 < (1u << 30) && W.r2 < (1u << 30) && W.r1 * W.r2 < (1ull << 40), kErrUnsupported,
                 "compress_2d: field too large");
    W.hist = hist + k * HN;
    W.flags = flags + 4 * k;
    W.flags_on_device = true;
    Eb2dJob& J = jobs[k];
    J.U = W.dU;
    J.V = W.dV;
    J.r1 = (int)W.r1;
    J.r2 = (int)W.r2;
    J.gx = (int)div_up(W.r2, kTW);
    J.cu = W.cu;
    J.cv = W.cv;
    J.hist = W.hist;
    J.flags = W.flags;
    J.ov = ov ? (*ov)[k] : EbOverride();
    approx = approx && !has_override(J.ov);
    J.blk0 = blk;
    const size_t tiles = div_up(W.r2, kTW) * div_up(W.r1, kTH);
    CPSZ_REQUIRE(blk + tiles < (1ull << 31), kErrUnsupported, "2D batch: too many tiles in one call");
    blk += (uint32_t)tiles;
  }
  Eb2dJob* dj = ctx.buf<Eb2dJob>("b2eb.jobs", nf);
  const double* q2d = ctx.d_q2d();
  Eb2dJob* hj = (Eb2dJob*)ctx.host_table("b2eb.jobs", nf * sizeof(Eb2dJob));
  std::memcpy(hj, jobs.data(), nf * sizeof(Eb2dJob));
  ctx.upload_table("b2eb.jobs", dj, hj, nf * sizeof(Eb2dJob));
  k_fill_u32<<<1024, 256, 0, st>>>(hist, nf * HN, 0);
  k_fill_u32<<<(unsigned)div_up(4 * nf, 256), 256, 0, st>>>((uint32_t*)flags, 4 * nf, 0);
  if (approx)
    launch_eb_fast(ctx, dj, (int)nf, blk, max_pwr_eb);
  else
    k_eb2d_multi<false><<<blk, 256, 0, st>>>(dj, (int)nf, max_pwr_eb, q2d, ctx.q2d_size(), nullptr, nullptr);
  CPSZ_HIP(hipGetLastError());
}

void c2_check(Ctx& ctx, C2Work* const* w, size_t nf) {
  hipStream_t st = ctx.stream();
  int* hf = (int*)ctx.host_buf("c2.hflags", std::max<size_t>(nf, 1) * 16);
  for (size_t k = 0; k < nf; k++)
    CPSZ_HIP(hipMemcpyAsync(hf + 4 * k, w[k]->flags, 16, hipMemcpyDeviceToHost, st));
  CPSZ_HIP(hipStreamSynchronize(st));
  for (size_t k = 0; k < nf; k++) {
    CPSZ_REQUIRE(hf[4 * k] == 0, kErrNonFinite,
                 "compress_2d: non-finite input or bound (the reference cannot encode it)");
    w[k]->wide = hf[4 * k + 1] != 0;
  }
}

namespace {
void launch_wavefront64(Ctx& ctx, Strip2d64& sa, int* prog, int grid_cap) {
  hipStream_t st = ctx.stream();
  sa.ticket = prog;
  k_zero_i32<<<1, 64, 0, st>>>(prog, 1);
  const size_t nb = (size_t)2 * sa.nstrips * sa.r2;
  k_fill_u64<<<(int)std::min<size_t>(div_up(nb, 256), 4096), 256, 0, st>>>(sa.bnd, nb, kBndEmpty);
  const int grid = std::min(2 * sa.nstrips, grid_cap);
  k_lorenzo2d_f64<true><<<grid, 64, 0, st>>>(sa);
  CPSZ_HIP(hipGetLastError());
}
}  // namespace

void c2_wavefront(Ctx& ctx, C2Work* const* w, size_t nf, int grid_cap) {
  if (nf == 1 && w[0]->f64) {
    Strip2d64 sa{};
    sa.x[0] = w[0]->dU64;
    sa.x[1] = w[0]->dV64;
    sa.c[0] = w[0]->cu;
    sa.c[1] = w[0]->cv;
    sa.q[0] = w[0]->qu;
    sa.q[1] = w[0]->qv;
    sa.bnd = w[0]->bnd;
    sa.r1 = (int)w[0]->r1;
    sa.r2 = (int)w[0]->r2;
    sa.nstrips = (int)div_up(w[0]->r1, 64);
    launch_wavefront64(ctx, sa, w[0]->prog, grid_cap);
    return;
  }
  for (size_t k = 0; k < nf; k++)
    CPSZ_REQUIRE(!w[k]->f64, kErrUnsupported, "compress_2d: double fields go one at a time");
  std::vector<Strip2d> f(nf);
  for (size_t k = 0; k < nf; k++) {
    Strip2d& sa = f[k];
    sa = Strip2d{};
    const bool lg = w[k]->log;
    sa.x[0] = lg ? w[k]->lu : w[k]->dU;
    sa.x[1] = lg ? w[k]->lv : w[k]->dV;
    sa.c[0] = w[k]->cu;
    sa.c[1] = lg ? w[k]->cu : w[k]->cv;
    sa.shift = lg ? 1 : 2;
    sa.q[0] = w[k]->qu;
    sa.q[1] = w[k]->qv;
    sa.nstrips = (int)div_up(w[k]->r1, 64);
    sa.bnd = w[k]->bnd;
    sa.prog = w[k]->prog;
    sa.ticket = w[k]->prog + 2 * sa.nstrips;
    sa.r1 = (int)w[k]->r1;
    sa.r2 = (int)w[k]->r2;
    sa.wide = w[k]->wide;
    sa.flags = w[k]->flags_on_device ? w[k]->flags : nullptr;
    CPSZ_REQUIRE(!(lg && (sa.wide || sa.flags)), kErrInternal, "compress_2d_log: eb id beyond the base-2 range");
  }
  launch_wavefront<true>(ctx, f, grid_cap);
}

void c2_entropy(Ctx& ctx, const C2Work& w, DeviceStream& out) {
  hipStream_t st = ctx.stream();
  const size_t r1 = w.r1, r2 = w.r2, n = r1 * r2;
  // the values the codec ran on (unpredictable ones are stored as they are)
  const float *dU = w.log ? w.lu : w.dU, *dV = w.log ? w.lv : w.dV;
  const uint16_t *cu = w.cu, *cv = w.cv, *qu = w.qu, *qv = w.qv;
  uint32_t* hist = w.hist;
  // ---- histograms of the quant codes
  {
    const uint32_t win_lo = 32768 - 8192, win_n = 16384;
    U16Hist hu{qu, n}, hv{qv, n};
    const size_t nch = div_up(n, U16Hist::kChunk);
    const int grid = (int)std::min<size_t>(encode_blocks(nch), 512);
    k_hist<<<grid, 256, win_n * 4, st>>>(hu, nch, hist + kMaxEbCode, win_lo, win_n);
    k_hist<<<grid, 256, win_n * 4, st>>>(hv, nch, hist + kMaxEbCode + 65536, win_lo, win_n);
    CPSZ_HIP(hipGetLastError());
  }
  ctx.mark("hist");
  // ---- host trees
  const size_t hist_n = kMaxEbCode + 2 * 65536;
  uint32_t* hh = (uint32_t*)ctx.host_buf("c2.hist", hist_n * 4);
  CPSZ_HIP(hipMemcpyAsync(hh, hist, hist_n * 4, hipMemcpyDeviceToHost, st));
  const hipEvent_t hist_done = ctx.record_point();
  // per-block unpredictable counts and their scans need no tree: they run
  // while the host builds the trees
  const int nby = (int)div_up(r2, kBlock2d);
  const size_t nb = div_up(r1, kBlock2d) * nby;
  const size_t m_ids = w.log ? n : 2 * n;
  const size_t nch_ids = div_up(m_ids, SrcIds2d::kChunk);
  SrcIds2d s_ids{cu, cv, n, m_ids};
  SrcCodes2d s_u{qu, (int)r1, (int)r2, nby}, s_v{qv, (int)r1, (int)r2, nby};
  uint64_t* bits = ctx.buf<uint64_t>("c2.bits", nch_ids + 2 * nb);
  uint64_t* offs = ctx.buf<uint64_t>("c2.offs", nch_ids + 2 * nb + 8);
  uint32_t* zc = ctx.buf<uint32_t>("c2.zc", 2 * nb);
  uint64_t* zoff = ctx.buf<uint64_t>("c2.zoff", 2 * nb + 4);
  uint64_t* totals = ctx.buf<uint64_t>("c2.totals", 8);
  k_block_zeros<<<unp_grid(nb), 256, 0, st>>>(qu, (int)r1, (int)r2, nby, nb, zc);
  k_block_zeros<<<unp_grid(nb), 256, 0, st>>>(qv, (int)r1, (int)r2, nby, nb, zc + nb);
  CPSZ_HIP(hipGetLastError());
  scan_u32_to_u64(ctx, zc, zoff, nb, totals + 3);
  scan_u32_to_u64(ctx, zc + nb, zoff + nb, nb, totals + 4);
  CPSZ_HIP(hipEventSynchronize(hist_done));
  // the v tree is built on the helper thread beside the eb and u trees, and
  // the id stream's bit counts run on the device meanwhile; the tables are
  // kept per thread so their storage is reused from call to call (the helper
  // reaches the calling thread's objects through references: a thread_local
  // named in the job would be the helper thread's own)
  thread_local HuffEncodeTable t_eb, t_u, t_v_keep;
  thread_local std::vector<uint64_t> f, fv_keep;
  HuffEncodeTable& t_v = t_v_keep;
  std::vector<uint64_t>& fv = fv_keep;
  f.resize(65536);
  fv.resize(65536);
  uint64_t* cb_code = ctx.buf<uint64_t>("c2.cbcode", 3 * 65536);
  uint8_t* cb_len = ctx.buf<uint8_t>("c2.cblen", 3 * 65536);
  for (int k = 0; k < 65536; k++) fv[k] = hh[kMaxEbCode + 65536 + k];
  {
    HelperJob v_tree([&] { huff_build(fv.data(), 65536, 2 * kCapacity, t_v); }, !ctx.batch_worker);
    for (int k = 0; k < kMaxEbCode; k++) f[k] = hh[k];
    huff_build(f.data(), kMaxEbCode, 2 * 1024, t_eb);
    upload_codebook(ctx, t_eb, cb_code, cb_len, "c2.eb");
    k_chunk_bits<<<(int)encode_blocks(nch_ids), kEncThreads, 0, st>>>(s_ids, nch_ids, cb_len, bits);
    CPSZ_HIP(hipGetLastError());
    scan_u64(ctx, bits, offs, nch_ids, totals + 0);
    for (int k = 0; k < 65536; k++) f[k] = hh[kMaxEbCode + k];
    huff_build(f.data(), 65536, 2 * kCapacity, t_u);
    v_tree.join();
  }
  ctx.mark("tree");
  // ---- codebooks and encode pass A (bits per chunk)
  upload_codebook(ctx, t_u, cb_code + 65536, cb_len + 65536, "c2.u");
  upload_codebook(ctx, t_v, cb_code + 2 * 65536, cb_len + 2 * 65536, "c2.v");
  k_chunk_bits<<<(int)encode_blocks(nb), kEncThreads, 0, st>>>(s_u, nb, cb_len + 65536, bits + nch_ids);
  k_chunk_bits<<<(int)encode_blocks(nb), kEncThreads, 0, st>>>(s_v, nb, cb_len + 2 * 65536, bits + nch_ids + nb);
  CPSZ_HIP(hipGetLastError());
  scan_u64(ctx, bits + nch_ids, offs + nch_ids, nb, totals + 1);
  scan_u64(ctx, bits + nch_ids + nb, offs + nch_ids + nb, nb, totals + 2);
  uint64_t* ht = (uint64_t*)ctx.host_buf("c2.totals", 8 * 8);
  CPSZ_HIP(hipMemcpyAsync(ht, totals, 5 * 8, hipMemcpyDeviceToHost, st));
  CPSZ_HIP(hipStreamSynchronize(st));
  const uint64_t bits_eb = ht[0], bits_u = ht[1], bits_v = ht[2], unp_u = ht[3], unp_v = ht[4];
  ctx.mark("scan");
  // ---- encode pass B into aligned scratch
  const size_t bytes_eb = div_up(bits_eb, 8), bytes_u = div_up(bits_u, 8), bytes_v = div_up(bits_v, 8);
  const size_t w_eb = div_up(bytes_eb, 4) + 2, w_u = div_up(bytes_u, 4) + 2, w_v = div_up(bytes_v, 4) + 2;
  uint32_t* words = ctx.buf<uint32_t>("c2.words", w_eb + w_u + w_v);
  k_fill_u32<<<1024, 256, 0, st>>>(words, w_eb + w_u + w_v, 0);
  k_encode<<<(int)encode_blocks(nch_ids), kEncThreads, 0, st>>>(s_ids, nch_ids, cb_code, cb_len, offs, bits, words);
  k_encode<<<(int)encode_blocks(nb), kEncThreads, 0, st>>>(s_u, nb, cb_code + 65536, cb_len + 65536, offs + nch_ids,
                                                   bits + nch_ids, words + w_eb);
  k_encode<<<(int)encode_blocks(nb), kEncThreads, 0, st>>>(s_v, nb, cb_code + 2 * 65536, cb_len + 2 * 65536,
                                                   offs + nch_ids + nb, bits + nch_ids + nb, words + w_eb + w_u);
  // unpredictable values in the field's own type (T = double: 8 bytes each)
  const size_t vb = w.f64 ? 8 : 4;
  uint8_t* unp = ctx.buf<uint8_t>("c2.unp", vb * (unp_u + unp_v + 1));
  if (w.f64) {
    k_gather_unpred<double><<<unp_grid(nb), 256, 0, st>>>(qu, w.dU64, (int)r1, (int)r2, nby, nb, zoff, (double*)unp);
    k_gather_unpred<double><<<unp_grid(nb), 256, 0, st>>>(qv, w.dV64, (int)r1, (int)r2, nby, nb, zoff + nb,
                                                          (double*)unp + unp_u);
  } else {
    k_gather_unpred<float><<<unp_grid(nb), 256, 0, st>>>(qu, dU, (int)r1, (int)r2, nby, nb, zoff, (float*)unp);
    k_gather_unpred<float><<<unp_grid(nb), 256, 0, st>>>(qv, dV, (int)r1, (int)r2, nby, nb, zoff + nb,
                                                         (float*)unp + unp_u);
  }
  CPSZ_HIP(hipGetLastError());
  ctx.mark("encode");
  // ---- stream layout
  Heads2d H;
  heads_2d(t_eb, t_u, t_v, bytes_eb, bytes_u, bytes_v, unp_u, unp_v, nb, w.log, vb, H);
  const std::vector<uint8_t>&hdr = H.hdr, &h_eb = H.h_eb, &hu1 = H.hu1, &hu2 = H.hu2, &hv1 = H.hv1, &hv2 = H.hv2;
  const size_t eb_size = H.eb_size, u_size = H.u_size, v_size = H.v_size;
  // offline_log: the two sign bitmaps follow the id section
  // (src/sz_compress_cp_preserve_2d.cpp:358-366)
  const size_t smap = w.log ? (n - 1) / 8 + 1 : 0;
  const size_t total = hdr.size() + eb_size + 2 * smap + u_size + v_size;
  uint8_t* dout = ctx.buf<uint8_t>("c2.out", total + 16);
  // host pieces go through one pinned staging buffer
  std::vector<const std::vector<uint8_t>*> hp = {&hdr, &h_eb, &hu1, &hu2, &hv1, &hv2};
  size_t hbytes = 0;
  for (auto* p : hp) hbytes += p->size();
  uint8_t* stage = (uint8_t*)ctx.host_buf("c2.stage", hbytes);
  std::vector<Piece> pieces;
  size_t pos = 0, spos = 0;
  auto host_piece = [&](const std::vector<uint8_t>& v) {
    std::memcpy(stage + spos, v.data(), v.size());
    pieces.push_back({pos, stage + spos, v.size(), false});
    pos += v.size();
    spos += v.size();
  };
  auto dev_piece = [&](const void* src, size_t len) {
    pieces.push_back({pos, src, len, true});
    pos += len;
  };
  host_piece(hdr);
  host_piece(h_eb);
  dev_piece(words, bytes_eb);
  if (w.log) {
    const size_t nq = div_up(n, 64);
    if (nq * 8 == smap) dev_piece(w.smap, 2 * smap);  // the two bitmaps are already contiguous
    else
      for (int c = 0; c < 2; c++) dev_piece(w.smap + c * nq, smap);
  }
  host_piece(hu1);
  dev_piece(unp, vb * unp_u);
  host_piece(hu2);
  dev_piece(words + w_eb, bytes_u);
  host_piece(hv1);
  dev_piece(unp + vb * unp_u, vb * unp_v);
  host_piece(hv2);
  dev_piece(words + w_eb + w_u, bytes_v);
  CPSZ_REQUIRE(pos == total, kErrInternal, "stream layout mismatch");
  for (const Piece& p : pieces) {
    if (!p.len) continue;
    CPSZ_HIP(hipMemcpyAsync(dout + p.off, p.src, p.len,
                            p.device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
  }
  ctx.mark("assemble");
  out.data = dout;
  out.size = total;
  CPSZ_HIP(hipStreamSynchronize(st));
}

// Batched entropy stage (plain offline, float), on the device from the codes
// to the finished streams: every field's histograms, block zero counts and
// their scan, the present symbols of every histogram, the Huffman trees (one
// wave per tree, huff_tree.hip), chunk bit counts and one scan for every
// stream, then the layout (sizes, offsets, word arrays) computed by a kernel
// from the per-stream totals, the encodes, the unpredictable gathers, the
// headers and the assembly. The host only fills the job tables up front and
// reads the sizes at the end. Each field's stream is byte-identical to
// c2_entropy's.
void c2_entropy_batch(Ctx& ctx, C2Work* const* w, size_t nf, uint8_t* dout, const std::vector<size_t>& off,
                      const std::vector<size_t>& cap, std::vector<size_t>& size, int workers, bool device_trees) {
  hipStream_t st = ctx.stream();
  constexpr size_t HN = kMaxEbCode + 2 * 65536;  // dense codebook entries per field
  constexpr uint32_t kMaxCodeLeaves = 65536, kMaxIdLeaves = kMaxEbCode;
  struct G {
    size_t n, nb, nby, nch_ids, zb, cb;
  };
  std::vector<G> g(nf);
  size_t Z = 0, CB = 0, words_cap = 0, unp_cap = 0;
  for (size_t k = 0; k < nf; k++) {
    const C2Work& W = *w[k];
    CPSZ_REQUIRE(!W.log && !W.f64, kErrInternal, "batched entropy: plain float fields only");
    G& x = g[k];
    x.n = W.r1 * W.r2;
    x.nby = div_up(W.r2, kBlock2d);
    x.nb = div_up(W.r1, kBlock2d) * x.nby;
    x.nch_ids = div_up(2 * x.n, (size_t)SrcIds2d::kChunk);
    x.zb = Z;
    Z += 2 * x.nb;
    x.cb = CB;
    CB += x.nch_ids + 2 * x.nb;
    words_cap += cap[k] / 4 + 6 + 64;
    unp_cap += 2 * x.n + 64;
  }
  const size_t ntree = 3 * nf;
  // ---- device pools (fixed per batch shape)
  uint32_t* zc = ctx.buf<uint32_t>("b2e.zc", Z + 1);
  uint64_t* zoff = ctx.buf<uint64_t>("b2e.zoff", Z + 1);
  uint64_t* bits = ctx.buf<uint64_t>("b2e.bits", CB + 1);
  uint64_t* offs = ctx.buf<uint64_t>("b2e.offs", CB + 1);
  uint64_t* cb_code = ctx.buf<uint64_t>("b2e.cbcode", nf * HN);
  uint8_t* cb_len = ctx.buf<uint8_t>("b2e.cblen", nf * HN);
  uint32_t* d_syms = ctx.buf<uint32_t>("b2e.syms", nf * (kMaxIdLeaves + 2 * (size_t)kMaxCodeLeaves));
  uint32_t* d_freqs = ctx.buf<uint32_t>("b2e.freqs", nf * (kMaxIdLeaves + 2 * (size_t)kMaxCodeLeaves));
  // per tree: heap (code trees only: id trees fit LDS), tree bytes, info
  const size_t heap_n = kMaxCodeLeaves + 2;
  const size_t tb_code = 13 * (2 * (size_t)kMaxCodeLeaves - 1) + 1, tb_id = 13 * (2 * (size_t)kMaxIdLeaves - 1) + 1;
  const size_t tb_field = (tb_id + 2 * tb_code + 3 * 255) & ~(size_t)255;
  uint64_t* t_heap = ctx.buf<uint64_t>("b2e.theap", 2 * nf * heap_n);
  uint8_t* t_tree = ctx.buf<uint8_t>("b2e.ttree", nf * tb_field);
  uint8_t* meta = ctx.buf<uint8_t>("b2e.meta", ntree * (12 + 32) + nf * 8 + 5 * nf * 8 + 64);
  unsigned long long* d_total = (unsigned long long*)meta;        // present symbols, all trees
  uint64_t* d_words_used = (uint64_t*)(meta + 8);
  int* d_err = (int*)(meta + 16);                                  // layout pool overflow
  uint64_t* d_base = (uint64_t*)(meta + 64);
  uint64_t* d_info = d_base + ntree;                               // 4 per tree
  uint64_t* d_sizes = d_info + 4 * ntree;
  uint64_t* d_tot = d_sizes + nf;
  uint32_t* d_count = (uint32_t*)(d_tot + 5 * nf);
  uint32_t* words = ctx.buf<uint32_t>("b2e.words", words_cap);
  float* unp = ctx.buf<float>("b2e.unp", unp_cap);
  // ---- job tables, filled on the host and sent in one copy
  for (size_t k = 0; k < nf; k++)
    CPSZ_REQUIRE(w[k]->flags == w[0]->flags + 4 * k, kErrInternal, "batched entropy: eb flags not contiguous");
  Arena A;
  const size_t abytes = 2 * nf * sizeof(HistJob<U16Hist>) + ntree * sizeof(CompactJob) +
                        ntree * sizeof(TreeJob) + nf * sizeof(EncJob<SrcIds2d>) + 2 * nf * sizeof(EncJob<SrcCodes2d>) +
                        5 * nf * sizeof(ScanDiff) + nf * sizeof(Layout2d) + 16 * 64;
  A.d = ctx.buf<uint8_t>("b2e.tab", abytes);
  A.h = (uint8_t*)ctx.host_table("b2e.tab", abytes);
  A.cap = abytes;
  HistJob<U16Hist>* hj;
  HistJob<U16Hist>* d_hj = A.take(2 * nf, &hj);
  CompactJob* cj;
  CompactJob* d_cj = A.take(ntree, &cj);
  TreeJob* tj;
  TreeJob* d_tj = A.take(ntree, &tj);
  EncJob<SrcIds2d>* ej_i;
  EncJob<SrcIds2d>* d_ej_i = A.take(nf, &ej_i);
  EncJob<SrcCodes2d>* ej_c;
  EncJob<SrcCodes2d>* d_ej_c = A.take(2 * nf, &ej_c);
  ScanDiff* sd;
  ScanDiff* d_sd = A.take(5 * nf, &sd);
  Layout2d* lj;
  Layout2d* d_lj = A.take(nf, &lj);
  uint32_t hblk = 0, iblk = 0, cblk = 0;
  for (size_t k = 0; k < nf; k++) {
    const C2Work& W = *w[k];
    const G& x = g[k];
    const size_t nch = div_up(x.n, (size_t)U16Hist::kChunk);
    for (int c = 0; c < 2; c++) {
      HistJob<U16Hist>& h = hj[2 * k + c];
      h.src = U16Hist{c ? W.qv : W.qu, x.n};
      h.nchunks = nch;
      h.hist = W.hist + kMaxEbCode + c * 65536;
      // 128 chunks per block: a block clears and flushes its 16384-bin window once per 131072 codes
      h.nblk = (uint32_t)std::min<size_t>(div_up(nch, 128), 512);
      h.blk0 = hblk;
      hblk += h.nblk;
    }
    for (int t = 0; t < 3; t++) {
      const size_t tt = 3 * k + t;
      cj[tt] = CompactJob{W.hist + (t == 0 ? 0 : kMaxEbCode + (t - 1) * 65536), t == 0 ? (uint32_t)kMaxEbCode : 65536u};
      TreeJob& T = tj[tt];
      T = TreeJob{};
      T.syms = d_syms;
      T.freqs = d_freqs;
      T.count = d_count + tt;
      T.base = d_base + tt;
      T.state_num = t == 0 ? 2 * 1024 : 2 * kCapacity;
      T.max_leaves = t == 0 ? kMaxIdLeaves : kMaxCodeLeaves;
      T.heap = t == 0 ? nullptr : t_heap + (2 * k + t - 1) * heap_n;
      T.tree = t_tree + k * tb_field + (t == 0 ? 0 : ((tb_id + 255) & ~(size_t)255) + (t - 1) * ((tb_code + 255) & ~(size_t)255));
      const size_t at = k * HN + (t == 0 ? 0 : kMaxEbCode + (t - 1) * 65536);
      T.code = cb_code + at;
      T.len = cb_len + at;
      T.info = d_info + 4 * tt;
    }
    EncJob<SrcIds2d>& e = ej_i[k];
    e.src = SrcIds2d{W.cu, W.cv, x.n, 2 * x.n};
    e.nchunks = x.nch_ids;
    e.code = cb_code + k * HN;
    e.len = cb_len + k * HN;
    e.bits = bits + x.cb;
    e.off = offs + x.cb;
    e.words = nullptr;
    e.blk0 = iblk;
    e.zc = nullptr;
    iblk += (uint32_t)encode_blocks(x.nch_ids);
    for (int c = 0; c < 2; c++) {
      EncJob<SrcCodes2d>& f = ej_c[2 * k + c];
      // the blocks' zero counts come with their bit counts; the encoder
      // gathers the unpredictable values (its array comes from k_layout2d)
      f.src = SrcCodes2d{c ? W.qv : W.qu, (int)W.r1, (int)W.r2, (int)x.nby, c ? W.dV : W.dU,
                         zoff + x.zb + c * x.nb, nullptr};
      f.zc = zc + x.zb + c * x.nb;
      f.nchunks = x.nb;
      f.code = cb_code + k * HN + kMaxEbCode + c * 65536;
      f.len = cb_len + k * HN + kMaxEbCode + c * 65536;
      f.bits = bits + x.cb + x.nch_ids + c * x.nb;
      f.off = offs + x.cb + x.nch_ids + c * x.nb;
      f.words = nullptr;
      f.blk0 = cblk;
      cblk += (uint32_t)encode_blocks(x.nb);
    }
    const size_t c0 = x.cb, c1 = c0 + x.nch_ids, c2 = c1 + x.nb, c3 = c2 + x.nb;
    sd[5 * k + 0] = ScanDiff{offs + c0, offs + c1};
    sd[5 * k + 1] = ScanDiff{offs + c1, offs + c2};
    sd[5 * k + 2] = ScanDiff{offs + c2, offs + c3};
    sd[5 * k + 3] = ScanDiff{zoff + x.zb, zoff + x.zb + x.nb};
    sd[5 * k + 4] = ScanDiff{zoff + x.zb + x.nb, zoff + x.zb + 2 * x.nb};
    Layout2d& l = lj[k];
    l = Layout2d{};
    l.off = off[k];
    l.cap = cap[k];
    l.nb = x.nb;
    l.tot = d_tot + 5 * k;
    for (int t = 0; t < 3; t++) {
      l.tinfo[t] = d_info + 4 * (3 * k + t);
      l.tree[t] = tj[3 * k + t].tree;
    }
    l.ej_ids = d_ej_i + k;
    l.ej_codes = d_ej_c + 2 * k;
  }
  ctx.upload_table("b2e.tab", A.d, A.h, A.used);
  // ---- histograms, present symbols, trees
  const uint32_t win_lo = 32768 - 4096, win_n = 8192;
  k_hist_multi<U16Hist><<<hblk, 256, win_n * 4, st>>>(d_hj, (int)(2 * nf), win_lo, win_n);
  CPSZ_HIP(hipGetLastError());
  k_zero_i32<<<1, 64, 0, st>>>((int*)meta, 16);  // total, words used, error flag
  k_compact_hist<<<(unsigned)ntree, 1024, 0, st>>>(d_cj, d_count, d_base, d_total, d_syms, d_freqs);
  CPSZ_HIP(hipGetLastError());
  if (device_trees) {
    ctx.mark("hist");
    build_trees_device(ctx, d_tj, (int)ntree);
  } else {
    // present symbols to the host, trees on the worker pool, results moved to
    // where the device builder would have put them
    CPSZ_REQUIRE(!ctx.capturing, kErrInternal, "batched entropy: host trees cannot be captured");
    const size_t cm = 12 * ntree;
    uint8_t* hc = (uint8_t*)ctx.host_buf("b2e.hcnt", cm + 16 * nf + 64);
    CPSZ_HIP(hipMemcpyAsync(hc, d_total, 8, hipMemcpyDeviceToHost, st));
    CPSZ_HIP(hipMemcpyAsync(hc + 64, d_base, 8 * ntree, hipMemcpyDeviceToHost, st));
    CPSZ_HIP(hipMemcpyAsync(hc + 64 + 8 * ntree, d_count, 4 * ntree, hipMemcpyDeviceToHost, st));
    CPSZ_HIP(hipMemcpyAsync(hc + 64 + cm, w[0]->flags, 16 * nf, hipMemcpyDeviceToHost, st));
    CPSZ_HIP(hipStreamSynchronize(st));
    for (size_t k = 0; k < nf; k++)  // a skipped field's codes are not the codec's: report it before its trees
      CPSZ_REQUIRE(((const int*)(hc + 64 + cm))[4 * k] == 0, kErrNonFinite,
                   "compress_2d: non-finite input or bound (the reference cannot encode it)");
    const uint64_t present = *(const uint64_t*)hc;
    const uint64_t* h_base = (const uint64_t*)(hc + 64);
    const uint32_t* h_count = (const uint32_t*)(hc + 64 + 8 * ntree);
    uint32_t* h_sf = (uint32_t*)ctx.host_buf("b2e.sf", 8 * present + 8);
    CPSZ_HIP(hipMemcpyAsync(h_sf, d_syms, 4 * present, hipMemcpyDeviceToHost, st));
    CPSZ_HIP(hipMemcpyAsync(h_sf + present, d_freqs, 4 * present, hipMemcpyDeviceToHost, st));
    CPSZ_HIP(hipStreamSynchronize(st));
    ctx.mark("hist");
    // kept per calling thread so their storage is reused; the pool threads
    // reach them through a reference (a thread_local named inside a job would
    // be the worker's own)
    thread_local std::vector<HuffEncodeTable> tabs_keep;
    std::vector<HuffEncodeTable>& tabs = tabs_keep;
    if (tabs.size() < ntree) tabs.resize(ntree);
    // the upload table's layout follows from the leaf counts alone, so each
    // worker writes its tree's results straight into it
    auto tree_bytes_of = [](size_t k) {
      const size_t N = 2 * k - 1, w = N <= 256 ? 1 : (N <= 65536 ? 2 : 4);
      return 1 + 2 * N * w + 5 * N;
    };
    size_t tbytes = 0;
    for (size_t t = 0; t < ntree; t++) tbytes += h_count[t] * 13 + tree_bytes_of(h_count[t] ? h_count[t] : 1) + 64;
    Arena B;
    const size_t bbytes = tbytes + ntree * sizeof(HostTreeJob) + 64;
    B.d = ctx.buf<uint8_t>("b2e.tabB", bbytes);
    B.h = (uint8_t*)ctx.host_table("b2e.tabB", bbytes);
    B.cap = bbytes;
    HostTreeJob* hj2;
    HostTreeJob* d_hj2 = B.take(ntree, &hj2);
    std::vector<uint8_t*> hdst(4 * ntree);
    for (size_t t = 0; t < ntree; t++) {
      const size_t kk = h_count[t];
      HostTreeJob& J = hj2[t];
      J.cb.codes = B.take(kk, (uint64_t**)&hdst[4 * t]);
      J.cb.syms = B.take(kk, (uint32_t**)&hdst[4 * t + 1]);
      J.cb.lens = B.take(kk, &hdst[4 * t + 2]);
      J.tree_src = B.take(tree_bytes_of(kk ? kk : 1), &hdst[4 * t + 3]);
      J.cb.k = (uint32_t)kk;
      J.cb.code = tj[t].code;
      J.cb.len = tj[t].len;
      J.tree_dst = tj[t].tree;
      J.info_dst = tj[t].info;
    }
    {
      std::vector<std::function<void(Ctx&)>> jobs(ntree);
      for (size_t t = 0; t < ntree; t++)
        jobs[t] = [&, t](Ctx&) {
          CPSZ_REQUIRE(h_count[t] > 0, kErrInternal, "batched entropy: empty histogram");
          HuffEncodeTable& T = tabs[t];
          huff_build_sparse(h_sf + h_base[t], h_sf + present + h_base[t], h_count[t], tj[t].state_num, T);
          const size_t kk = T.syms.size();
          CPSZ_REQUIRE(kk == h_count[t] && T.tree.size() == tree_bytes_of(kk), kErrInternal,
                       "batched entropy: tree size mismatch");
          std::memcpy(hdst[4 * t], T.codes.data(), 8 * kk);
          std::memcpy(hdst[4 * t + 1], T.syms.data(), 4 * kk);
          std::memcpy(hdst[4 * t + 2], T.lens.data(), kk);
          std::memcpy(hdst[4 * t + 3], T.tree.data(), T.tree.size());
          HostTreeJob& J = hj2[t];
          J.tree_n = T.tree.size();
          J.info[0] = T.node_count;
          J.info[1] = T.tree.size();
          J.info[2] = (uint64_t)T.max_len;
          J.info[3] = 0;
        };
      pool_run(pool_for(ctx.device()), jobs, std::max(workers, 1));
    }
    ctx.upload_table("b2e.tabB", B.d, B.h, B.used);
    k_host_trees_in<<<(unsigned)ntree, 256, 0, st>>>(d_hj2);
    CPSZ_HIP(hipGetLastError());
  }
  ctx.mark("tree");
  // ---- chunk bit counts (and the code blocks' zero counts), one scan for
  // every stream, per-stream totals
  k_chunk_bits_multi<SrcIds2d><<<iblk, kEncThreads, 0, st>>>(d_ej_i, (int)nf);
  k_chunk_bits_multi<SrcCodes2d><<<cblk, kEncThreads, 0, st>>>(d_ej_c, (int)(2 * nf));
  CPSZ_HIP(hipGetLastError());
  scan_u64(ctx, bits, offs, CB, offs + CB);
  scan_u32_to_u64(ctx, zc, zoff, Z, zoff + Z);
  k_scan_diffs<<<(unsigned)div_up(5 * nf, 256), 256, 0, st>>>(d_sd, (int)(5 * nf), d_tot);
  CPSZ_HIP(hipGetLastError());
  ctx.mark("scan");
  // ---- layout, encode (with the unpredictable gather), headers, assembly
  k_layout2d<<<1, 1, 0, st>>>(d_lj, (int)nf, words, words_cap, unp, unp_cap, d_sizes, d_words_used, d_err);
  k_fill_words<<<2048, 256, 0, st>>>(words, d_words_used);
  k_encode_multi<SrcIds2d><<<iblk, kEncThreads, 0, st>>>(d_ej_i, (int)nf);
  k_encode_multi<SrcCodes2d><<<cblk, kEncThreads, 0, st>>>(d_ej_c, (int)(2 * nf));
  CPSZ_HIP(hipGetLastError());
  ctx.mark("encode");
  k_write_heads2d<<<(unsigned)nf, 256, 0, st>>>(d_lj, dout);
  k_assemble2d<<<dim3(64, (unsigned)nf), 256, 0, st>>>(d_lj, dout);
  CPSZ_HIP(hipGetLastError());
  ctx.mark("assemble");
  // ---- tree infos, sizes and error flags back to the host
  const size_t mb = ntree * 32 + nf * 8;
  uint8_t* hm = (uint8_t*)ctx.host_buf("b2e.hmeta", mb + 16 * nf + 64);
  CPSZ_HIP(hipMemcpyAsync(hm, d_info, mb, hipMemcpyDeviceToHost, st));
  CPSZ_HIP(hipMemcpyAsync(hm + mb, d_err, 4, hipMemcpyDeviceToHost, st));
  CPSZ_HIP(hipMemcpyAsync(hm + mb + 64, w[0]->flags, 16 * nf, hipMemcpyDeviceToHost, st));
  if (!ctx.capturing) c2_entropy_batch_finish(ctx, nf, cap, size);
}

void c2_entropy_batch_finish(Ctx& ctx, size_t nf, const std::vector<size_t>& cap, std::vector<size_t>& size) {
  const size_t ntree = 3 * nf, mb = ntree * 32 + nf * 8;
  const uint8_t* hm = (const uint8_t*)ctx.host_buf("b2e.hmeta", mb + 16 * nf + 64);
  CPSZ_HIP(hipStreamSynchronize(ctx.stream()));
  const int* hflags = (const int*)(hm + mb + 64);
  for (size_t k = 0; k < nf; k++)
    CPSZ_REQUIRE(hflags[4 * k] == 0, kErrNonFinite,
                 "compress_2d: non-finite input or bound (the reference cannot encode it)");
  const uint64_t* hinfo = (const uint64_t*)hm;
  const uint64_t* hsizes = hinfo + 4 * ntree;
  for (size_t t = 0; t < ntree; t++) {
    const uint64_t e = hinfo[4 * t + 3];
    CPSZ_REQUIRE(e != 3, kErrUnsupported, "symbol beyond the coder's state count");
    CPSZ_REQUIRE(e == 0, kErrInternal, "device huffman tree failed (code " + std::to_string(e) + ")");
  }
  CPSZ_REQUIRE(*(const int*)(hm + mb) == 0, kErrInternal, "batched entropy: layout pools overflowed");
  for (size_t k = 0; k < nf; k++) {
    CPSZ_REQUIRE(hsizes[k] <= cap[k], kErrInternal, "batch stream bound exceeded");
    size[k] = hsizes[k];
  }
}

void eb_2d_device(Ctx& ctx, const float* dU, const float* dV, size_t r1, size_t r2, double max_pwr_eb,
                  double* dE) {
  CPSZ_REQUIRE(dU && dV && dE && r1 >= 1 && r2 >= 1, kErrArg, "eb_2d: bad arguments");
  C2Work w;
  w.dU = dU;
  w.dV = dV;
  w.r1 = r1;
  w.r2 = r2;
  c2_bind(w, ctx.buf<uint8_t>("c2.scratch", c2_scratch_bytes(r1, r2)));
  EbOverride ov;
  ov.eb_out = dE;
  c2_eb(ctx, w, max_pwr_eb, ov);
}

void compress_2d_device(Ctx& ctx, const float* dU, const float* dV, size_t r1, size_t r2,
                        double max_pwr_eb, DeviceStream& out, const EbOverride& ov) {
  ctx.stages.clear();
  ctx.mark("start");
  C2Work w;
  w.dU = dU;
  w.dV = dV;
  w.r1 = r1;
  w.r2 = r2;
  CPSZ_REQUIRE(r1 >= 1 && r2 >= 1 && r1 < (1u << 30) && r2 < (1u << 30), kErrArg, "compress_2d: bad shape");
  c2_bind(w, ctx.buf<uint8_t>("c2.scratch", c2_scratch_bytes(r1, r2)));
  c2_eb(ctx, w, max_pwr_eb, ov);
  ctx.mark("eb");
  C2Work* wp = &w;
  c2_check(ctx, &wp, 1);
  c2_wavefront(ctx, &wp, 1, 4096);
  ctx.mark("lorenzo");
  c2_entropy(ctx, w, out);
  ctx.finish_marks();
}

void compress_2d_log_device(Ctx& ctx, const float* dU, const float* dV, size_t r1, size_t r2,
                            double max_pwr_eb, DeviceStream& out) {
  ctx.stages.clear();
  ctx.mark("start");
  CPSZ_REQUIRE(dU && dV && r1 >= 1 && r2 >= 1 && r1 < (1u << 30) && r2 < (1u << 30), kErrArg,
               "compress_2d_log: bad arguments");
  hipStream_t st = ctx.stream();
  C2Work w;
  w.dU = dU;
  w.dV = dV;
  w.r1 = r1;
  w.r2 = r2;
  w.log = true;
  const size_t n = r1 * r2, nq = div_up(n, 64);
  c2_bind(w, ctx.buf<uint8_t>("c2.scratch", c2_scratch_bytes(r1, r2)));
  float* L = ctx.buf<float>("c2l.log", 2 * n);
  uint64_t* sm = ctx.buf<uint64_t>("c2l.smap", 2 * nq);
  w.lu = L;
  w.lv = L + n;
  w.smap = sm;
  c2_eb(ctx, w, max_pwr_eb, EbOverride());
  k_log2d<<<dim3((unsigned)div_up(nq, 4), 2), 256, 0, st>>>(dU, dV, n, nq, ctx.libm_dev().log2f, L, L + n, sm);
  CPSZ_HIP(hipGetLastError());
  ctx.mark("eb");
  C2Work* wp = &w;
  c2_check(ctx, &wp, 1);
  c2_wavefront(ctx, &wp, 1, 4096);
  ctx.mark("lorenzo");
  c2_entropy(ctx, w, out);
  ctx.finish_marks();
}

namespace {
// Reads stream bytes either from a host copy or from device memory.
struct HuffSection {
  size_t node_count;
  uint32_t tree_size;
  std::vector<uint8_t> tree;
  size_t nbytes;
  size_t bits_off;  // absolute offset of the bitstream
  size_t end;
};

HuffSection parse_huff(StreamReader& r, size_t off) {
  HuffSection h;
  h.node_count = r.get<size_t>(off);
  h.tree_size = r.get<uint32_t>(off + 8);
  CPSZ_REQUIRE(h.node_count >= 1 && h.node_count < (1u << 24), kErrStream, "bad huffman node count");
  const uint8_t* t = r.read(off + 12, h.tree_size);
  h.tree.assign(t, t + h.tree_size);
  h.nbytes = r.get<size_t>(off + 12 + h.tree_size);
  h.bits_off = off + 20 + h.tree_size;
  CPSZ_REQUIRE(h.nbytes <= r.size && h.bits_off + h.nbytes <= r.size, kErrStream,
               "huffman bitstream truncated");
  h.end = h.bits_off + h.nbytes;
  return h;
}

// The header fields and section positions of a 2D stream.
struct D2Plan {
  HuffSection h_eb;
  struct Comp {
    size_t off = 0, unp_count = 0, unp_off = 0;
    HuffSection h;
  } comp[2];
  int nby = 0;
  size_t nb = 0;
  size_t sign_off = 0;  // offline_log: the two sign bitmaps
};

// log: the offline_log layout (base 2, the sign bitmaps between the id
// section and the u section, src/sz_decompress_cp_preserve_2d.cpp:60-79).
void d2_parse(StreamReader& rd, size_t r1, size_t r2, D2Plan& p, bool log = false, size_t vb = 4) {
  const size_t n = r1 * r2;
  const int base = rd.get<int>(0);
  const double thr = rd.get<double>(4);
  const size_t u_size = rd.get<size_t>(20);
  CPSZ_REQUIRE(base == (log ? 2 : 4) && thr == DBL_EPSILON, kErrUnsupported,
               log ? "decompress_2d_log: stream base/threshold differ from the offline_log 2D format"
                   : "decompress_2d: stream base/threshold differ from the offline 2D format");
  p.h_eb = parse_huff(rd, 36);
  // components follow the id section (and the sign bitmaps) directly; the u
  // section spans u_size bytes
  p.nby = (int)div_up(r2, kBlock2d);
  p.nb = div_up(r1, kBlock2d) * p.nby;
  size_t off = p.h_eb.end;
  if (log) {
    const size_t smap = (n - 1) / 8 + 1;
    p.sign_off = off;
    rd.read(off, 2 * smap);  // bounds check
    off += 2 * smap;
  }
  for (int c = 0; c < 2; c++) {
    D2Plan::Comp& k = p.comp[c];
    k.off = off;
    const int bs = rd.get<int>(off);
    const int radius = rd.get<int>(off + 5);
    // meanInfo<T> is 2 * sizeof(T) bytes: bool, padding, T mean
    const uint8_t use_mean = rd.get<uint8_t>(off + 9);
    const size_t reg = rd.get<size_t>(off + 9 + 2 * vb);
    k.unp_count = rd.get<size_t>(off + 17 + 2 * vb);
    CPSZ_REQUIRE(bs == kBlock2d && radius == kRadius && use_mean == 0 && reg == 0, kErrUnsupported,
                 "decompress_2d: stream uses block settings the offline compressor never writes");
    CPSZ_REQUIRE(k.unp_count <= n, kErrStream, "bad unpredictable count");
    k.unp_off = off + 25 + 2 * vb;
    const size_t ind = k.unp_off + vb * k.unp_count;
    const size_t ind_bytes = (p.nb - 1) / 8 + 1;
    const uint8_t* ib = rd.read(ind, ind_bytes);
    for (size_t b = 0; b < ind_bytes; b++)
      CPSZ_REQUIRE(ib[b] == 0, kErrUnsupported, "regression blocks are not produced by the 2D codec");
    k.h = parse_huff(rd, ind + ind_bytes);
    off = c == 0 ? off + u_size : k.h.end;
  }
}

// Largest symbol of a parsed tree (ids above 81 need the wide wavefront path).
uint32_t max_symbol(const HuffDecodeTree& t) {
  uint32_t m = 0;
  for (size_t i = 0; i < t.node_count; i++)
    if ((t.leaf[i] || (i == 0 && t.root_leaf)) && t.sym[i] > m) m = t.sym[i];
  return m;
}

// Unpredictable values to per-element raw arrays (block-major order).
void d2_unpred(Ctx& ctx, const uint8_t* d_in, const D2Plan& p, D2Work& w) {
  hipStream_t st = ctx.stream();
  const size_t r1 = w.r1, r2 = w.r2, n = r1 * r2, nb = p.nb;
  uint32_t* zc = ctx.buf<uint32_t>("d2.zc", 2 * nb);
  uint64_t* zoff = ctx.buf<uint64_t>("d2.zoff", 2 * nb + 4);
  CPSZ_HIP(hipMemsetAsync(w.err, 0, 8, st));
  uint16_t* qs[2] = {w.qu, w.qv};
  const size_t vb = w.f64 ? 8 : 4;
  for (int c = 0; c < 2; c++) {
    uint8_t* vals = ctx.buf<uint8_t>(c ? "d2.unpv" : "d2.unpu", vb * (p.comp[c].unp_count + 1));
    if (p.comp[c].unp_count)
      CPSZ_HIP(hipMemcpyAsync(vals, d_in + p.comp[c].unp_off, vb * p.comp[c].unp_count, hipMemcpyDeviceToDevice,
                              st));
    k_block_zeros<<<unp_grid(nb), 256, 0, st>>>(qs[c], (int)r1, (int)r2, p.nby, nb, zc + c * nb);
    scan_u32_to_u64(ctx, zc + c * nb, zoff + c * (nb + 1), nb, zoff + c * (nb + 1) + nb);
    if (w.f64)
      k_scatter_unpred<double><<<unp_grid(nb), 256, 0, st>>>(qs[c], (const double*)vals, (int)r1, (int)r2, p.nby,
                                                             nb, zoff + c * (nb + 1), p.comp[c].unp_count,
                                                             (double*)w.raw + c * n, w.err);
    else
      k_scatter_unpred<float><<<unp_grid(nb), 256, 0, st>>>(qs[c], (const float*)vals, (int)r1, (int)r2, p.nby, nb,
                                                            zoff + c * (nb + 1), p.comp[c].unp_count,
                                                            w.raw + c * n, w.err);
    CPSZ_HIP(hipGetLastError());
  }
}

}  // namespace

size_t d2_scratch_bytes(size_t r1, size_t r2, bool f64) {
  const size_t n = r1 * r2, ns = div_up(r1, 64);
  return 4 * al256(2 * n) + al256((f64 ? 16 : 8) * n) + al256(2 * ns * r2 * 8) + al256((2 * ns + 1) * 4) +
         al256(16);
}

void d2_bind(D2Work& w, uint8_t* p) {
  const size_t n = w.r1 * w.r2, ns = div_up(w.r1, 64);
  auto take = [&](size_t bytes) {
    uint8_t* q = p;
    p += al256(bytes);
    return q;
  };
  w.sid_u = (uint16_t*)take(2 * n);
  w.sid_v = (uint16_t*)take(2 * n);
  w.qu = (uint16_t*)take(2 * n);
  w.qv = (uint16_t*)take(2 * n);
  w.raw = (float*)take((w.f64 ? 16 : 8) * n);  // T = double: 2n doubles
  w.bnd = (uint64_t*)take(2 * ns * w.r2 * 8);
  w.prog = (int*)take((2 * ns + 1) * 4);
  w.err = (int*)take(16);
}

void d2_entropy(Ctx& ctx, const uint8_t* d_in, const uint8_t* h_in, size_t in_size, D2Work& w) {
  CPSZ_REQUIRE(d_in && w.r1 >= 1 && w.r2 >= 1, kErrArg, "decompress_2d: bad arguments");
  hipStream_t st = ctx.stream();
  const size_t r1 = w.r1, r2 = w.r2, n = r1 * r2;
  uint16_t *sid_u = w.sid_u, *sid_v = w.sid_v, *qu = w.qu, *qv = w.qv;
  const bool log = w.log;
  StreamReader rd(ctx, h_in, d_in, in_size);
  D2Plan p;
  d2_parse(rd, r1, r2, p, w.log, w.f64 ? 8 : 4);
  if (w.log) w.signs = d_in + p.sign_off;
  const HuffSection& h_eb = p.h_eb;
  const int nby = p.nby;
  ctx.mark("parse");
  // ---- bitstreams into aligned scratch
  auto stage_bits = [&](const HuffSection& h, const char* name) {
    // slack words: a speculative decode near the end peeks up to ~110 bits past it
    uint32_t* w = ctx.buf<uint32_t>(name, div_up(h.nbytes, 4) + 8);
    CPSZ_HIP(hipMemsetAsync(w, 0, (div_up(h.nbytes, 4) + 8) * 4, st));
    if (h.nbytes)
      CPSZ_HIP(hipMemcpyAsync(w, d_in + h.bits_off, h.nbytes, hipMemcpyDeviceToDevice, st));
    return w;
  };
  // The code trees are parsed and tabulated while the device decodes the id
  // stream (the u tree on the helper thread, the v tree on this one between
  // queuing the id decode and waiting for it); then the u and v streams decode
  // in lockstep. The tables are kept per calling thread (their storage is
  // reused); the helper fills them through the pointer (a thread_local named
  // in the job would be the helper thread's own).
  thread_local HostDecoder hc_keep[2];
  HostDecoder* hc = hc_keep;
  auto prep = [&](int c) {
    prepare_decoder(p.comp[c].h.tree.data(), p.comp[c].h.tree.size(), p.comp[c].h.node_count, hc[c]);
  };
  uint32_t id_max = 0;
  {
    HuffDecodeTree t;
    CPSZ_REQUIRE(huff_parse_tree(h_eb.tree.data(), h_eb.tree.size(), h_eb.node_count, t), kErrStream,
                 "malformed id tree");
    id_max = max_symbol(t);
  }
  {
    HelperJob u_ready([&] { prep(0); }, !ctx.batch_worker);
    uint32_t* w = stage_bits(h_eb, "d2.w.eb");
    DeviceDecoder dd = upload_decoder(ctx, h_eb.tree.data(), h_eb.tree.size(), h_eb.node_count, "d2.eb");
    if (id_max < 256)
      decode_stream(ctx, dd, w, h_eb.nbytes, log ? n : 2 * n, SinkIds2dB{sid_u, sid_v, n}, "d2.eb",
                    [&] { prep(1); });
    else
      decode_stream(ctx, dd, w, h_eb.nbytes, log ? n : 2 * n, SinkIds2d{sid_u, sid_v, n}, "d2.eb",
                    [&] { prep(1); });
    u_ready.join();  // rethrows a malformed code tree
  }
  {
    uint32_t* wu = stage_bits(p.comp[0].h, "d2.w.u");
    uint32_t* wv = stage_bits(p.comp[1].h, "d2.w.v");
    const DeviceDecoder du = upload_decoder(ctx, hc[0], "d2.u");
    const DeviceDecoder dv = upload_decoder(ctx, hc[1], "d2.v");
    DecJob<SinkCodes2d> jobs[2] = {
        {&du, wu, p.comp[0].h.nbytes, n, SinkCodes2d{qu, (int)r1, (int)r2, nby}, "d2.u"},
        {&dv, wv, p.comp[1].h.nbytes, n, SinkCodes2d{qv, (int)r1, (int)r2, nby}, "d2.v"}};
    decode_streams(ctx, jobs, 2);
  }
  ctx.mark("huffman");
  d2_unpred(ctx, d_in, p, w);
  w.wide = id_max > 81;
  ctx.mark("unpred");
}

// Unpredictable values of every field of a batch (float): the stream's
// values copied to aligned arrays in one assembly launch, the per-block zero
// counts of every component, one global scan, one scatter launch.
// zeros_counted: the code decoders already filled b2u.zc (SinkCodes2d::zc)
void d2_unpred_batch(Ctx& ctx, const std::vector<const uint8_t*>& d_in, const std::vector<D2Plan>& plan,
                     D2Work* const* w, size_t nf, bool zeros_counted, const uint64_t* zmask) {
  hipStream_t st = ctx.stream();
  size_t Z = 0, V = 0, npieces = 0;
  constexpr size_t kPiece = 65536;
  std::vector<size_t> zb(nf), vb(2 * nf);
  for (size_t k = 0; k < nf; k++) {
    CPSZ_REQUIRE(!w[k]->f64 && !w[k]->log, kErrInternal, "batched unpredictable scatter: plain float fields only");
    zb[k] = Z;
    Z += 2 * plan[k].nb;
    for (int c = 0; c < 2; c++) {
      vb[2 * k + c] = V;
      V += (plan[k].comp[c].unp_count + 63) & ~(size_t)63;
      npieces += div_up(4 * plan[k].comp[c].unp_count, kPiece);
    }
  }
  uint32_t* zc = ctx.buf<uint32_t>("b2u.zc", Z + 1);
  uint64_t* zoff = ctx.buf<uint64_t>("b2u.zoff", Z + 1);
  float* vals = ctx.buf<float>("b2u.vals", V + 1);
  int* errs = ctx.buf<int>("b2u.err", 4 * nf);
  Arena A;
  const size_t bytes = 2 * nf * sizeof(UnpJob) + npieces * sizeof(CopyPiece) + 64;
  A.h = (uint8_t*)ctx.host_buf("b2u.tab", bytes);
  A.d = ctx.buf<uint8_t>("b2u.tab", bytes);
  A.cap = bytes;
  UnpJob* uj;
  UnpJob* d_uj = A.take(2 * nf, &uj);
  CopyPiece* cp;
  CopyPiece* d_cp = A.take(npieces, &cp);
  size_t np = 0;
  uint32_t blk = 0;
  for (size_t k = 0; k < nf; k++) {
    const D2Plan& p = plan[k];
    D2Work& W = *w[k];
    W.err = errs + 4 * k;
    for (int c = 0; c < 2; c++) {
      UnpJob& u = uj[2 * k + c];
      u = UnpJob{};
      u.q = c ? W.qv : W.qu;
      u.x = vals + vb[2 * k + c];
      u.r1 = (int)W.r1;
      u.r2 = (int)W.r2;
      u.nby = p.nby;
      u.nb = p.nb;
      u.cnt = zc + zb[k] + c * p.nb;
      u.off = zoff + zb[k] + c * p.nb;
      u.out = W.raw + c * W.r1 * W.r2;
      u.count = p.comp[c].unp_count;
      u.err = W.err;
      u.blk0 = blk;
      u.zm = zmask ? zmask + 3 * (zb[k] + c * p.nb) : nullptr;
      blk += (uint32_t)unp_grid(p.nb);
      const size_t len = 4 * p.comp[c].unp_count;
      for (size_t o = 0; o < len; o += kPiece)
        cp[np++] = CopyPiece{4 * vb[2 * k + c] + o, d_in[k] + p.comp[c].unp_off + o,
                             (uint32_t)std::min(kPiece, len - o), 0};
    }
  }
  CPSZ_HIP(hipMemcpyAsync(A.d, A.h, A.used, hipMemcpyHostToDevice, st));
  k_fill_u32<<<1, 256, 0, st>>>((uint32_t*)errs, 4 * nf, 0);
  if (np) k_assemble<<<(unsigned)np, 256, 0, st>>>(d_cp, (uint8_t*)vals);
  if (!zeros_counted) k_block_zeros_multi<<<blk, 256, 0, st>>>(d_uj, (int)(2 * nf));
  CPSZ_HIP(hipGetLastError());
  scan_u32_to_u64(ctx, zc, zoff, Z, zoff + Z);
  k_scatter_unpred_multi<<<blk, 256, 0, st>>>(d_uj, (int)(2 * nf));
  CPSZ_HIP(hipGetLastError());
  // the pinned table is reused by the next call only after the stream passes here
  CPSZ_HIP(hipStreamSynchronize(st));
}

constexpr uint32_t kIdWarm = 2048;  // speculative warm-up of the id streams, bits

// Batched form: every field's headers parsed and trees tabulated on the
// workers; all decoders uploaded in one copy and all bitstreams staged in one
// launch; the id streams of every field decoded together, then the u and v
// streams of every field together (one grid per pass, one host check per
// round); the unpredictable scatter of each field on the workers.
void d2_entropy_batch(Ctx& ctx, const std::vector<const uint8_t*>& d_in, const std::vector<size_t>& in_size,
                      D2Work* const* w, size_t nf, int workers) {
  hipStream_t st = ctx.stream();
  std::vector<D2Plan> plan(nf);
  std::vector<HostDecoder> hd(3 * nf);
  // Every stream's header pieces come to the host in four batched rounds (a
  // gather of the pieces on the device, one copy, one wait each): the start of
  // the stream (header, id tree), the u and v section heads, their indicator
  // bytes and tree heads, then their trees. The offsets follow the layout
  // d2_parse reads; a piece outside what was fetched is read from the device
  // by the parser itself.
  for (size_t k = 0; k < nf; k++)
    CPSZ_REQUIRE(d_in[k] && w[k]->r1 >= 1 && w[k]->r2 >= 1, kErrArg, "decompress_2d: bad arguments");
  std::vector<std::vector<StreamReader::Range>> pre(nf);
  {
    struct Want {
      size_t k, off, len;
    };
    // one round: gathers the wanted pieces into a device staging buffer, copies
    // it to pinned host memory and records the ranges; returns their host copies
    int round = 0;
    auto fetch = [&](const std::vector<Want>& want) {
      std::vector<const uint8_t*> where(want.size(), nullptr);
      size_t total = 0, npc = 0;
      for (const Want& x : want) {
        total += (x.len + 15) & ~(size_t)15;
        npc += div_up(x.len, (size_t)65536);
      }
      if (!total) return where;
      const std::string tag = "b2d.pre" + std::to_string(round++);
      uint8_t* dst = ctx.buf<uint8_t>(tag.c_str(), total);
      uint8_t* hst = (uint8_t*)ctx.host_buf(tag.c_str(), total);
      CopyPiece* hp = (CopyPiece*)ctx.host_table((tag + "t").c_str(), npc * sizeof(CopyPiece));
      CopyPiece* dp = ctx.buf<CopyPiece>((tag + "t").c_str(), npc);
      size_t at = 0, np = 0;
      for (size_t i = 0; i < want.size(); i++) {
        const Want& x = want[i];
        for (size_t o = 0; o < x.len; o += 65536)
          hp[np++] = CopyPiece{at + o, d_in[x.k] + x.off + o, (uint32_t)std::min<size_t>(65536, x.len - o), 0};
        where[i] = hst + at;
        pre[x.k].push_back({x.off, x.len, hst + at});
        at += (x.len + 15) & ~(size_t)15;
      }
      ctx.upload_table((tag + "t").c_str(), dp, hp, np * sizeof(CopyPiece));
      k_assemble<<<(unsigned)np, 256, 0, st>>>(dp, dst);
      CPSZ_HIP(hipGetLastError());
      CPSZ_HIP(hipMemcpyAsync(hst, dst, at, hipMemcpyDeviceToHost, st));
      CPSZ_HIP(hipStreamSynchronize(st));
      return where;
    };
    constexpr size_t kStart = 32u << 10;  // header + id tree (at most 1024 leaves: < 27 KiB)
    std::vector<Want> w1(nf);
    for (size_t k = 0; k < nf; k++) w1[k] = {k, 0, std::min(kStart, in_size[k])};
    const std::vector<const uint8_t*> h1 = fetch(w1);
    // section starts: u after the id section, v u_size bytes later
    std::vector<Want> w2;
    for (size_t k = 0; k < nf; k++) {
      const uint8_t* h = h1[k];
      if (!h || w1[k].len < 48) continue;
      uint64_t u_size, nbytes;
      uint32_t tsz;
      std::memcpy(&u_size, h + 20, 8);
      std::memcpy(&tsz, h + 44, 4);
      if ((uint64_t)48 + tsz + 8 > w1[k].len) continue;
      std::memcpy(&nbytes, h + 48 + tsz, 8);
      const uint64_t u0 = 56 + (uint64_t)tsz + nbytes;
      if (nbytes > in_size[k] || u0 > in_size[k]) continue;
      // section head: block size, radius, meanInfo, regression count, then
      // the unpredictable count (at +25) and values (from +33)
      if (in_size[k] - u0 >= 33) {
        w2.push_back({k, u0, 33});
      }
      if (u_size <= in_size[k] - u0 && in_size[k] - u0 - u_size >= 33) {
        w2.push_back({k, u0 + u_size, 33});
      }
    }
    const std::vector<const uint8_t*> h2 = fetch(w2);
    std::vector<Want> w3;
    for (size_t i = 0; i < w2.size(); i++) {
      const size_t k = w2[i].k;
      uint64_t unp;
      std::memcpy(&unp, h2[i] + 25, 8);
      const size_t nb = div_up(w[k]->r1, kBlock2d) * div_up(w[k]->r2, kBlock2d);
      const uint64_t ind = w2[i].off + 33 + 4 * (uint64_t)std::min<uint64_t>(unp, w[k]->r1 * w[k]->r2);
      const uint64_t len = (nb - 1) / 8 + 1 + 12;  // indicator bytes, tree head
      if (ind > in_size[k] || len > in_size[k] - ind) continue;
      w3.push_back({k, ind, len});
    }
    const std::vector<const uint8_t*> h3 = fetch(w3);
    std::vector<Want> w4;
    for (size_t i = 0; i < w3.size(); i++) {
      const size_t k = w3[i].k;
      uint32_t tsz;
      std::memcpy(&tsz, h3[i] + w3[i].len - 4, 4);
      const uint64_t t0 = w3[i].off + w3[i].len;
      if ((uint64_t)tsz + 8 > in_size[k] - t0) continue;
      w4.push_back({k, t0, (size_t)tsz + 8});  // tree and bitstream size
    }
    fetch(w4);
  }
  {
    std::vector<std::function<void(Ctx&)>> jobs(nf);
    for (size_t k = 0; k < nf; k++)
      jobs[k] = [&, k](Ctx& wc) {
        StreamReader rd(wc, nullptr, d_in[k], in_size[k]);
        rd.pre = pre[k];
        d2_parse(rd, w[k]->r1, w[k]->r2, plan[k]);
        const D2Plan& p = plan[k];
        prepare_decoder(p.h_eb.tree.data(), p.h_eb.tree.size(), p.h_eb.node_count, hd[3 * k]);
        for (int c = 0; c < 2; c++)
          prepare_decoder(p.comp[c].h.tree.data(), p.comp[c].h.tree.size(), p.comp[c].h.node_count,
                          hd[3 * k + 1 + c]);
        w[k]->wide = max_symbol(hd[3 * k].d) > 81;
      };
    pool_run(pool_for(ctx.device()), jobs, workers);
  }
  ctx.mark("parse");
  std::vector<const HostDecoder*> hp(3 * nf);
  for (size_t k = 0; k < 3 * nf; k++) hp[k] = &hd[k];
  std::vector<DeviceDecoder> dd(3 * nf);
  upload_decoders(ctx, hp.data(), (int)(3 * nf), dd.data(), "b2d.dec", &pool_for(ctx.device()), workers);
  ctx.mark("upload");
  std::vector<BitsCopy> bc(3 * nf);
  for (size_t k = 0; k < nf; k++) {
    const D2Plan& p = plan[k];
    bc[3 * k] = BitsCopy{d_in[k] + p.h_eb.bits_off, p.h_eb.nbytes, nullptr, 0};
    for (int c = 0; c < 2; c++)
      bc[3 * k + 1 + c] = BitsCopy{d_in[k] + p.comp[c].h.bits_off, p.comp[c].h.nbytes, nullptr, 0};
  }
  stage_bits_multi(ctx, bc.data(), (int)(3 * nf), "b2d.bits");
  ctx.mark("stage");
  std::vector<DecJob<SinkIds2d>> ji(nf);
  std::vector<DecJob<SinkIds2dB>> ji8(nf);
  std::vector<DecJob<SinkCodes2d>> jc(2 * nf);
  // the code decoders count each 12x12 block's zero codes as they write them
  // (the unpredictable scatter's offsets), in the array d2_unpred_batch scans
  size_t Z = 0;
  for (size_t k = 0; k < nf; k++) Z += 2 * plan[k].nb;
  uint32_t* zc = ctx.buf<uint32_t>("b2u.zc", Z + 1);
  k_fill_u32<<<(unsigned)std::min<size_t>(div_up(Z, 256), 2048), 256, 0, st>>>(zc, Z, 0);
  // and record where the zeros are, so that the scatter reads 24 bytes of masks per
  // block instead of the block's codes
  uint64_t* zm = ctx.buf<uint64_t>("b2u.zm", 3 * Z + 1);
  k_fill_u32<<<(unsigned)std::min<size_t>(div_up(6 * Z, 256), 2048), 256, 0, st>>>((uint32_t*)zm, 6 * Z, 0);
  CPSZ_HIP(hipGetLastError());
  size_t zb = 0;
  bool ids8 = true;
  for (size_t k = 0; k < nf; k++) ids8 = ids8 && max_symbol(hd[3 * k].d) < 256;
  for (size_t k = 0; k < nf; k++) {
    const size_t r1 = w[k]->r1, r2 = w[k]->r2, n = r1 * r2;
    const D2Plan& p = plan[k];
    // eb ids: codes of 1-5 bits over long runs of equal ids resynchronize
    // slowly from an arbitrary bit, so their speculative warm-up is longer
    ji[k] = {&dd[3 * k], bc[3 * k].dst, p.h_eb.nbytes, 2 * n, SinkIds2d{w[k]->sid_u, w[k]->sid_v, n}, "b2d",
             kIdWarm};
    ji8[k] = {&dd[3 * k], bc[3 * k].dst, p.h_eb.nbytes, 2 * n, SinkIds2dB{w[k]->sid_u, w[k]->sid_v, n}, "b2d",
              kIdWarm};
    uint16_t* qs[2] = {w[k]->qu, w[k]->qv};
    for (int c = 0; c < 2; c++)
      jc[2 * k + c] = {&dd[3 * k + 1 + c], bc[3 * k + 1 + c].dst, p.comp[c].h.nbytes, n,
                       SinkCodes2d{qs[c], (int)r1, (int)r2, p.nby, zc + zb + c * p.nb, zm + 3 * (zb + c * p.nb)},
                       "b2d"};
    zb += 2 * p.nb;
  }
  if (ids8) decode_streams_multi(ctx, ji8.data(), (int)nf, "b2d.ids");
  else decode_streams_multi(ctx, ji.data(), (int)nf, "b2d.ids");
  decode_streams_multi(ctx, jc.data(), (int)(2 * nf), "b2d.codes");
  ctx.mark("huffman");
  d2_unpred_batch(ctx, d_in, plan, w, nf, true, zm);
  ctx.mark("unpred");
}

void d2_wavefront(Ctx& ctx, D2Work* const* w, size_t nf, int grid_cap) {
  if (nf == 1 && w[0]->f64) {
    hipStream_t st = ctx.stream();
    const size_t n = w[0]->r1 * w[0]->r2;
    Strip2d64 sa{};
    sa.c[0] = w[0]->sid_u;
    sa.c[1] = w[0]->sid_v;
    sa.q[0] = w[0]->qu;
    sa.q[1] = w[0]->qv;
    sa.raw[0] = (const double*)w[0]->raw;
    sa.raw[1] = (const double*)w[0]->raw + n;
    sa.out[0] = w[0]->dU64;
    sa.out[1] = w[0]->dV64;
    sa.bnd = w[0]->bnd;
    sa.r1 = (int)w[0]->r1;
    sa.r2 = (int)w[0]->r2;
    sa.nstrips = (int)div_up(w[0]->r1, 64);
    sa.ticket = w[0]->prog;
    k_zero_i32<<<1, 64, 0, st>>>(w[0]->prog, 1);
    const size_t nb = (size_t)2 * sa.nstrips * sa.r2;
    k_fill_u64<<<(int)std::min<size_t>(div_up(nb, 256), 4096), 256, 0, st>>>(sa.bnd, nb, kBndEmpty);
    k_lorenzo2d_f64<false><<<std::min(2 * sa.nstrips, grid_cap), 64, 0, st>>>(sa);
    CPSZ_HIP(hipGetLastError());
    return;
  }
  for (size_t k = 0; k < nf; k++)
    CPSZ_REQUIRE(!w[k]->f64, kErrUnsupported, "decompress_2d: double fields go one at a time");
  std::vector<Strip2d> f(nf);
  for (size_t k = 0; k < nf; k++) {
    Strip2d& sa = f[k];
    sa = Strip2d{};
    const size_t n = w[k]->r1 * w[k]->r2;
    sa.c[0] = w[k]->sid_u;
    sa.c[1] = w[k]->log ? w[k]->sid_u : w[k]->sid_v;
    sa.shift = w[k]->log ? 1 : 2;
    sa.q[0] = w[k]->qu;
    sa.q[1] = w[k]->qv;
    sa.raw[0] = w[k]->raw;
    sa.raw[1] = w[k]->raw + n;
    sa.out[0] = w[k]->dU;
    sa.out[1] = w[k]->dV;
    sa.nstrips = (int)div_up(w[k]->r1, 64);
    sa.bnd = w[k]->bnd;
    sa.prog = w[k]->prog;
    sa.ticket = w[k]->prog + 2 * sa.nstrips;
    sa.r1 = (int)w[k]->r1;
    sa.r2 = (int)w[k]->r2;
    sa.wide = w[k]->wide;
    CPSZ_REQUIRE(!(w[k]->log && sa.wide), kErrStream, "decompress_2d_log: eb id beyond the base-2 range");
  }
  launch_wavefront<false>(ctx, f, grid_cap);
}

void d2_check(Ctx& ctx, D2Work* const* w, size_t nf) {
  hipStream_t st = ctx.stream();
  int* he = (int*)ctx.host_buf("d2.herr", std::max<size_t>(nf, 1) * 16);
  for (size_t k = 0; k < nf; k++) CPSZ_HIP(hipMemcpyAsync(he + 4 * k, w[k]->err, 8, hipMemcpyDeviceToHost, st));
  CPSZ_HIP(hipStreamSynchronize(st));
  for (size_t k = 0; k < nf; k++)
    CPSZ_REQUIRE(he[4 * k] == 0, kErrStream, "decompress_2d: more unpredictable codes than values");
}

void decompress_2d_device(Ctx& ctx, const uint8_t* d_in, const uint8_t* h_in, size_t in_size,
                          size_t r1, size_t r2, float* dU, float* dV) {
  CPSZ_REQUIRE(d_in && dU && dV && r1 >= 1 && r2 >= 1, kErrArg, "decompress_2d: bad arguments");
  ctx.stages.clear();
  ctx.mark("start");
  D2Work w;
  w.r1 = r1;
  w.r2 = r2;
  w.dU = dU;
  w.dV = dV;
  d2_bind(w, ctx.buf<uint8_t>("d2.scratch", d2_scratch_bytes(r1, r2)));
  CPSZ_HIP(hipMemsetAsync(w.err, 0, 16, ctx.stream()));
  d2_entropy(ctx, d_in, h_in, in_size, w);
  D2Work* wp = &w;
  d2_wavefront(ctx, &wp, 1, 4096);
  ctx.mark("lorenzo");
  d2_check(ctx, &wp, 1);
  ctx.finish_marks();
}

void decompress_2d_log_device(Ctx& ctx, const uint8_t* d_in, const uint8_t* h_in, size_t in_size, size_t r1,
                              size_t r2, float* dU, float* dV) {
  CPSZ_REQUIRE(d_in && dU && dV && r1 >= 1 && r2 >= 1, kErrArg, "decompress_2d_log: bad arguments");
  ctx.stages.clear();
  ctx.mark("start");
  D2Work w;
  w.r1 = r1;
  w.r2 = r2;
  w.dU = dU;
  w.dV = dV;
  w.log = true;
  d2_bind(w, ctx.buf<uint8_t>("d2.scratch", d2_scratch_bytes(r1, r2)));
  CPSZ_HIP(hipMemsetAsync(w.err, 0, 16, ctx.stream()));
  d2_entropy(ctx, d_in, h_in, in_size, w);
  D2Work* wp = &w;
  d2_wavefront(ctx, &wp, 1, 4096);
  ctx.mark("lorenzo");
  const size_t n = r1 * r2;
  k_finish2d<<<dim3((unsigned)div_up(n, 256), 2), 256, 0, ctx.stream()>>>(dU, dV, n, w.signs, (n - 1) / 8 + 1,
                                                                         ctx.libm_dev().exp2f);
  CPSZ_HIP(hipGetLastError());
  ctx.mark("finish");
  d2_check(ctx, &wp, 1);
  ctx.finish_marks();
}

void compress_2d_f64_device(Ctx& ctx, const double* dU, const double* dV, size_t r1, size_t r2,
                            double max_pwr_eb, DeviceStream& out) {
  ctx.stages.clear();
  ctx.mark("start");
  CPSZ_REQUIRE(dU && dV && r1 >= 1 && r2 >= 1 && r1 < (1u << 30) && r2 < (1u << 30), kErrArg,
               "compress_2d (double): bad arguments");
  C2Work w;
  w.f64 = true;
  w.dU64 = dU;
  w.dV64 = dV;
  w.r1 = r1;
  w.r2 = r2;
  c2_bind(w, ctx.buf<uint8_t>("c2.scratch", c2_scratch_bytes(r1, r2)));
  c2_eb(ctx, w, max_pwr_eb, EbOverride());
  ctx.mark("eb");
  C2Work* wp = &w;
  c2_check(ctx, &wp, 1);
  c2_wavefront(ctx, &wp, 1, 4096);
  ctx.mark("lorenzo");
  c2_entropy(ctx, w, out);
  ctx.finish_marks();
}

void decompress_2d_f64_device(Ctx& ctx, const uint8_t* d_in, const uint8_t* h_in, size_t in_size, size_t r1,
                              size_t r2, double* dU, double* dV) {
  CPSZ_REQUIRE(d_in && dU && dV && r1 >= 1 && r2 >= 1, kErrArg, "decompress_2d (double): bad arguments");
  ctx.stages.clear();
  ctx.mark("start");
  D2Work w;
  w.r1 = r1;
  w.r2 = r2;
  w.f64 = true;
  w.dU64 = dU;
  w.dV64 = dV;
  d2_bind(w, ctx.buf<uint8_t>("d2.scratch", d2_scratch_bytes(r1, r2, true)));
  CPSZ_HIP(hipMemsetAsync(w.err, 0, 16, ctx.stream()));
  d2_entropy(ctx, d_in, h_in, in_size, w);
  D2Work* wp = &w;
  d2_wavefront(ctx, &wp, 1, 4096);
  ctx.mark("lorenzo");
  d2_check(ctx, &wp, 1);
  ctx.finish_marks();
}

// Host emulation of the wavefront kernels' arithmetic (row-major order; any
// order that respects the Lorenzo dependencies yields the same values).
void lorenzo2d_emulate(const float* x, const uint16_t* c, size_t r1, size_t r2, bool wide,
                       uint16_t* q, float* recon) {
  std::vector<float> R((r1 + 1) * (r2 + 1), 0.f);
  const size_t W = r2 + 1;
  for (size_t i = 0; i < r1; i++)
    for (size_t j = 0; j < r2; j++) {
      float* rp = &R[(i + 1) * W + j + 1];
      const float pred = (rp[-1] + rp[-(ptrdiff_t)W]) - rp[-(ptrdiff_t)W - 1];
      uint16_t code;
      const size_t e = i * r2 + j;
      *rp = wide ? l2d_compress_step<true>(pred, x[e], c[e], code)
                 : l2d_compress_step<false>(pred, x[e], c[e], code);
      q[e] = code;
      if (recon) recon[e] = *rp;
    }
}

void lorenzo2d_emulate_decode(const uint16_t* q, const uint16_t* sid, const float* raw, size_t r1,
                              size_t r2, bool wide, float* out) {
  std::vector<float> R((r1 + 1) * (r2 + 1), 0.f);
  const size_t W = r2 + 1;
  for (size_t i = 0; i < r1; i++)
    for (size_t j = 0; j < r2; j++) {
      float* rp = &R[(i + 1) * W + j + 1];
      const float pred = (rp[-1] + rp[-(ptrdiff_t)W]) - rp[-(ptrdiff_t)W - 1];
      const size_t e = i * r2 + j;
      *rp = wide ? l2d_decompress_step<true>(pred, q[e], sid[e], raw[e])
                 : l2d_decompress_step<false>(pred, q[e], sid[e], raw[e]);
      out[e] = *rp;
    }
}

}  // namespace cpsz
