// 3D offline_log codec on gfx950: the drop-in for
// sz_compress_cp_preserve_3d_offline_log / sz_decompress_cp_preserve_3d_online_log.
//
// Compress (inputs in HBM):
//   k_eb3d       per-cell 6-tet bounds, marching along i over 16-plane strips
//                of a 15x63 vertex tile, folded into per-vertex minima with
//                LDS atomic-min on float bits, then the id code from the
//                host-libm threshold table
//   k_signmap    three MSB-first sign bitmaps
//   k_log        host-libm-exact log2f of |x| (-100 for zeros)
//   k_wave3d     wavefront: a workgroup owns 4 consecutive planes of a 64-row
//                j block (one wave per plane, lane = j); waves sweep k with a
//                one-step skew per lane; j-1 values arrive through DPP
//                wave_shr:1, plane i-1 values through LDS inside the workgroup
//                and agent-coherent loads behind progress counters across them
//   Huffman      ids (512 states) and the codes of predictable vertices
// Decompress mirrors it and finishes with host-libm-exact exp2f.
//
// Wave layout: every per-vertex array the wavefront reads or writes (log
// values, eb codes, stream ids, quantization codes, reconstruction) is stored
// in the wavefront's own order, [plane i][j block][step t][lane], with vertex
// (i, j, k) at step t = k + (j mod 64). At each step a wave then touches one
// contiguous 64-element row per array instead of 64 rows r3 elements apart.
// The producers (eb, log, scatter) and consumers (Huffman sources, gather,
// finish) translate from and to row-major vertex order.
#include <cfloat>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "codec3d.hpp"
#include "eb3d_math.hpp"
#include "huff_gpu.hpp"
#include "libm_math.hpp"
#include "pool.hpp"
#include "stream_reader.hpp"

namespace cpsz {

namespace {


constexpr int kP3 = 8;  // planes (waves) per wavefront workgroup
constexpr int kC3 = 4;  // steps per wavefront chunk (a multiple of 4)

// Wave layout: item (plane i, 64-row block b) holds T steps (r3 + 63 rounded
// up to whole chunks) of 64 lanes; lane l's values for steps 4u..4u+3 are
// contiguous, so one 16-byte load gives a lane four steps of a float array.
// Vertex (i, j, k) sits at step t = k + (j mod 64) of lane j mod 64.
struct Skew3 {
  int r1, r2, r3, nb, T;
  size_t per;  // elements per component
  __host__ __device__ size_t blk(int i, int b) const { return ((size_t)i * nb + b) * (size_t)T * 64; }
  __host__ __device__ size_t at(int i, int j, int k) const {
    const int l = j & 63, t = k + l;
    return blk(i, j >> 6) + (size_t)(t >> 2) * 256 + l * 4 + (t & 3);
  }
  // vertex index (row-major, < 2^32) -> wave-layout index
  __device__ size_t at_v(size_t v) const {
    const uint32_t vv = (uint32_t)v, row = vv / (uint32_t)r3, k = vv - row * (uint32_t)r3;
    const uint32_t i = row / (uint32_t)r2, j = row - i * (uint32_t)r2;
    return at((int)i, (int)j, (int)k);
  }
};

Skew3 make_skew(size_t r1, size_t r2, size_t r3) {
  Skew3 s;
  s.r1 = (int)r1;
  s.r2 = (int)r2;
  s.r3 = (int)r3;
  s.nb = (int)div_up(r2, 64);
  s.T = (int)(div_up(r3 + 63, kC3) * kC3);
  s.per = r1 * (size_t)s.nb * s.T * 64;
  CPSZ_REQUIRE(s.per < (1ull << 31), kErrUnsupported,
               "3D field too thin for the wavefront layout (r1 * ceil(r2/64) * (r3+63) * 64 >= 2^31)");
  return s;
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Per-vertex bounds, marching along i. A block owns an output tile of kMJ
// rows x kMK columns over kMS consecutive planes and walks the cell layers
// between them; each layer's 16 x 64 cells (origin rows j0-1 .. j0+14,
// columns k0-1 .. k0+62) are computed once, so only a one-cell halo in j and k
// (and one extra layer per plane strip) is recomputed: 1.16 cell evaluations
// per vertex instead of the 1.43 of a cubic tile. Two vertex planes sit in LDS;
// a thread takes one cell, evaluates its six tetrahedra from registers, folds
// them into the eight corner minima and applies those to the two planes'
// partial minima with LDS atomic-min on the float bits (the bounds are floats,
// >= 0; non-negative floats order like their bit patterns). A plane is final
// after the layers on both sides of it: then its ids are written (wave layout)
// and its slot is reset for plane + 2.
// FAST evaluates the tetrahedra with the approximate sign bounds (tet_bound_t):
// a vertex whose id is constant over the error envelope of its approximate
// bound gets that id; the others are appended to `fix` (count in *nfix) and
// recomputed exactly by k_eb3d_fix. FAST never writes eb_out.
constexpr int kMJ = 15, kMK = 63, kMS = 16;
constexpr int kVJ = kMJ + 2, kVK = kMK + 2;  // vertex halo box of one plane
constexpr uint32_t kInfBits = 0x7f800000u;

template <bool FAST>
__global__ void __launch_bounds__(256, 4) k_eb3d(const float* __restrict__ U, const float* __restrict__ V,
                                              const float* __restrict__ W, int r1, int r2, int r3,
                                              double max_pwr_eb, const double* __restrict__ q3d, int nq,
                                              Skew3 sk, uint8_t* __restrict__ code, int* __restrict__ flags,
                                              const uint8_t* __restrict__ gain, const uint8_t* __restrict__ fb,
                                              double* __restrict__ eb_out, uint32_t* __restrict__ fix,
                                              uint32_t* __restrict__ nfix) {
  __shared__ float sv[2][3][kVJ][kVK];
  __shared__ uint32_t seb[2][kMJ][kMK];
  const int tid = threadIdx.x;
  const int k0 = blockIdx.x * kMK, j0 = blockIdx.y * kMJ, i0 = blockIdx.z * kMS;
  const int i1 = min(i0 + kMS, r1);
  const size_t d0 = (size_t)r2 * r3, d1 = r3;
  const float* src[3] = {U, V, W};
  int bad = 0;
  for (int e = tid; e < 2 * kMJ * kMK; e += 256) (&seb[0][0][0])[e] = kInfBits;
  auto load_plane = [&](int pl) {
    const int slot = pl & 1;
    for (int e = tid; e < kVJ * kVK; e += 256) {
      const int b = e / kVK, c = e - b * kVK;
      const int gj = j0 - 1 + b, gk = k0 - 1 + c;
      const bool in = pl >= 0 && pl < r1 && gj >= 0 && gj < r2 && gk >= 0 && gk < r3;
      const size_t g = in ? (size_t)pl * d0 + (size_t)gj * d1 + gk : 0;
#pragma unroll
      for (int p = 0; p < 3; p++) {
        const float x = in ? src[p][g] : 0.f;
        if (!isfinite(x)) bad = 1;
        sv[slot][p][b][c] = x;
      }
    }
  };
  // cell origins of a layer inside the grid and this block's halo box
  const int cj_lo = max(j0 - 1, 0), cj_hi = min(j0 + kMJ - 1, r2 - 2);
  const int ck_lo = max(k0 - 1, 0), ck_hi = min(k0 + kMK - 1, r3 - 2);
  const int nj = cj_hi - cj_lo + 1, nk = ck_hi - ck_lo + 1;
  const int ncell = (nj > 0 && nk > 0) ? nj * nk : 0;
  load_plane(i0 - 1);
  for (int L = i0 - 1; L < i1; L++) {
    load_plane(L + 1);
    __syncthreads();
    if (L >= 0 && L <= r1 - 2) {
      for (int e = tid; e < ncell; e += 256) {
        const int cj = cj_lo + e / nk, ck = ck_lo + e % nk;
        const int b = cj - (j0 - 1), c = ck - (k0 - 1);
        float x[3][8];  // corner values, corner = di << 2 | dj << 1 | dk
#pragma unroll
        for (int cr = 0; cr < 8; cr++)
#pragma unroll
          for (int p = 0; p < 3; p++) x[p][cr] = sv[(L + (cr >> 2)) & 1][p][b + ((cr >> 1) & 1)][c + (cr & 1)];
        float cm[8];
#pragma unroll
        for (int cr = 0; cr < 8; cr++) cm[cr] = __uint_as_float(kInfBits);
#pragma unroll
        for (int t = 0; t < 6; t++) {
          int cv[3];
#pragma unroll
          for (int v = 0; v < 3; v++) {
            int di, dj, dk;
            tet_offsets(t, v, di, dj, dk);
            cv[v] = di << 2 | dj << 1 | dk;
          }
          const double eb = tet_bound_t<FAST>(x[0][cv[0]], x[0][cv[1]], x[0][cv[2]], x[0][0], x[1][cv[0]],
                                              x[1][cv[1]], x[1][cv[2]], x[1][0], x[2][cv[0]], x[2][cv[1]],
                                              x[2][cv[2]], x[2][0]);
          if (!(eb == eb)) bad = 1;
          const float ef = (float)eb;  // exact: the bound is a float, 0 or 1
#pragma unroll
          for (int v = 0; v < 3; v++) cm[cv[v]] = fminf(cm[cv[v]], ef);
          cm[0] = fminf(cm[0], ef);
        }
#pragma unroll
        for (int cr = 0; cr < 8; cr++) {
          const int pl = L + (cr >> 2), tj = cj + ((cr >> 1) & 1) - j0, tk = ck + (cr & 1) - k0;
          if (pl >= i0 && pl < i1 && tj >= 0 && tj < kMJ && tk >= 0 && tk < kMK)
            atomicMin(&seb[pl & 1][tj][tk], __float_as_uint(cm[cr]));
        }
      }
    }
    __syncthreads();
    if (L >= i0) {  // plane L has seen layers L-1 and L
      for (int e = tid; e < kMJ * kMK; e += 256) {
        const int tj = e / kMK, tk = e - tj * kMK;
        const int gj = j0 + tj, gk = k0 + tk;
        if (gj >= r2 || gk >= r3) continue;
        const uint32_t bits = seb[L & 1][tj][tk];
        seb[L & 1][tj][tk] = kInfBits;
        const size_t v = (size_t)L * d0 + (size_t)gj * d1 + gk;
        const int g = gain && !(fb && fb[v]) ? gain[v] : 0;
        if constexpr (FAST) {
          const double F = (double)__uint_as_float(bits);
          double lo = fmin(max_pwr_eb, (F - kEbAbs) * (1 - kEbRel));
          double hi = fmin(max_pwr_eb, (F + kEbAbs) * (1 + kEbRel));
          if (g) {
            lo = fmin(max_pwr_eb, ldexp(lo, g));
            hi = fmin(max_pwr_eb, ldexp(hi, g));
          }
          const uint32_t c = eb_code_3d(lo, q3d, nq);
          const bool sure = lo > 0 && c == eb_code_3d(hi, q3d, nq);
          code[sk.at(L, gj, gk)] = (uint8_t)c;
          const uint64_t m = __ballot(!sure);
          if (!sure) {
            const uint32_t lead = (uint32_t)__ffsll((unsigned long long)m) - 1;
            uint32_t base = 0;
            if (__lane_id() == lead) base = atomicAdd(nfix, (uint32_t)__popcll(m));
            base = __shfl(base, (int)lead);
            fix[base + lanes_below(m)] = (uint32_t)v;
          }
        } else {
          double eb = fmin(max_pwr_eb, (double)__uint_as_float(bits));
          if (eb_out) eb_out[v] = eb;
          if (g) eb = fmin(max_pwr_eb, ldexp(eb, g));
          code[sk.at(L, gj, gk)] = (uint8_t)eb_code_3d(eb, q3d, nq);
        }
      }
    }
  }
  if (bad) atomicOr(flags, 1);
}

// Exact bounds for the vertices k_eb3d<true> could not decide: the minimum of
// the exact bounds of the 24 tetrahedra around the vertex (those of the eight
// cells that hold it). The set of tetrahedra and their bounds equal the
// marching kernel's, so the minimum does too. Grid-stride over the list.
__global__ void __launch_bounds__(256) k_eb3d_fix(const float* __restrict__ U, const float* __restrict__ V,
                                                  const float* __restrict__ W, int r1, int r2, int r3,
                                                  double max_pwr_eb, const double* __restrict__ q3d, int nq,
                                                  Skew3 sk, uint8_t* __restrict__ code, int* __restrict__ flags,
                                                  const uint8_t* __restrict__ gain,
                                                  const uint8_t* __restrict__ fb,
                                                  const uint32_t* __restrict__ fix,
                                                  const uint32_t* __restrict__ nfix) {
  const uint32_t count = *nfix;
  const size_t d0 = (size_t)r2 * r3, d1 = r3;
  int bad = 0;
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < count; e += gridDim.x * blockDim.x) {
    const uint32_t v = fix[e];
    const int i = (int)(v / d0), rem = (int)(v - (size_t)i * d0), j = rem / r3, k = rem - j * r3;
    // the 3 x 3 x 3 vertex neighbourhood once (clamped at the borders, where
    // the cells that would use the clamped values are skipped)
    float nb[3][27];
#pragma unroll
    for (int o = 0; o < 27; o++) {
      const int gi = min(max(i + o / 9 - 1, 0), r1 - 1), gj = min(max(j + (o / 3) % 3 - 1, 0), r2 - 1),
                gk = min(max(k + o % 3 - 1, 0), r3 - 1);
      const size_t g = (size_t)gi * d0 + (size_t)gj * d1 + gk;
      nb[0][o] = U[g];
      nb[1][o] = V[g];
      nb[2][o] = W[g];
    }
    // cell c has its origin at the vertex minus c's offsets; the vertex is
    // that cell's corner c
    float m = __uint_as_float(kInfBits);
#pragma unroll
    for (int c = 0; c < 8; c++) {
      const int ci = i - (c >> 2), cj = j - ((c >> 1) & 1), ck = k - (c & 1);
      if (ci < 0 || cj < 0 || ck < 0 || ci > r1 - 2 || cj > r2 - 2 || ck > r3 - 2) continue;
      float x[3][8];
#pragma unroll
      for (int cr = 0; cr < 8; cr++) {
        const int o = (1 - (c >> 2) + (cr >> 2)) * 9 + (1 - ((c >> 1) & 1) + ((cr >> 1) & 1)) * 3 +
                      (1 - (c & 1) + (cr & 1));
        x[0][cr] = nb[0][o];
        x[1][cr] = nb[1][o];
        x[2][cr] = nb[2][o];
      }
#pragma unroll
      for (int t = 0; t < 6; t++) {
        int cv[3];
        bool holds = c == 0;
#pragma unroll
        for (int q = 0; q < 3; q++) {
          int di, dj, dk;
          tet_offsets(t, q, di, dj, dk);
          cv[q] = di << 2 | dj << 1 | dk;
          holds = holds || cv[q] == c;
        }
        if (!holds) continue;  // c and t are the same in every lane: no divergence
        const double eb = tet_bound(x[0][cv[0]], x[0][cv[1]], x[0][cv[2]], x[0][0], x[1][cv[0]], x[1][cv[1]],
                                    x[1][cv[2]], x[1][0], x[2][cv[0]], x[2][cv[1]], x[2][cv[2]], x[2][0]);
        if (!(eb == eb)) bad = 1;
        m = fminf(m, (float)eb);
      }
    }
    double eb = fmin(max_pwr_eb, (double)m);
    if (gain && !(fb && fb[v]) && gain[v]) eb = fmin(max_pwr_eb, ldexp(eb, (int)gain[v]));
    code[sk.at(i, j, k)] = (uint8_t)eb_code_3d(eb, q3d, nq);
  }
  if (bad) atomicOr(flags, 1);
}

// Self-check of the premises of the approximate bounds (cube_sign_bound_fast):
// the largest ulp distance between acbrt1_fast and acbrt1_nonneg_f over the
// bit patterns [lo, hi), and, over hashed argument pairs in [0, 2^126) with
// ratios from far apart to nearly equal, the largest |fast - exact| / (e R + A).
__global__ void __launch_bounds__(256) k_check_cbrt(uint32_t lo, uint32_t hi, unsigned* __restrict__ out) {
  unsigned worst = 0;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint64_t b = (uint64_t)lo + blockIdx.x * blockDim.x + threadIdx.x; b < hi; b += stride) {
    const float x = __uint_as_float((uint32_t)b);
    const uint32_t a = __float_as_uint(acbrt1_nonneg_f(x)), f = __float_as_uint(acbrt1_fast(x));
    const unsigned d = a > f ? a - f : f - a;
    worst = d > worst ? d : worst;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned t = __shfl_xor(worst, o);
    worst = t > worst ? t : worst;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(out, worst);
}

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

__global__ void __launch_bounds__(256) k_check_csb(uint32_t iters, uint32_t seed, unsigned* __restrict__ out) {
  float worst = 0.f;
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  for (uint32_t it = 0; it < iters; it++) {
    const uint32_t h1 = mix32(g * 2654435761u + it * 40503u + seed), h2 = mix32(h1 ^ 0x9e3779b9u),
                   h3 = mix32(h2 + 12345u);
    const float p = ldexpf(1.0f + (h1 & 0xFFFFFF) * 0x1p-24f, (int)(h2 % 250) - 125);
    const float t = ldexpf((h3 & 0xFFFF) * 0x1p-16f, -(int)((h3 >> 16) % 30));
    float n = (h2 & 0x100) ? p * (1.0f + t) : p / (1.0f + t);
    if ((h3 & 0x7) == 0) n = 0;
    if (!(n < 0x1p125f)) continue;
    const float e = cube_sign_bound(p, n), f = cube_sign_bound_fast(p, n);
    const float r = (float)(fabs((double)f - (double)e) / ((double)e * kEbRel + kEbAbs));
    worst = fmaxf(worst, r);
  }
  for (int o = 32; o > 0; o >>= 1) worst = fmaxf(worst, __shfl_xor(worst, o));
  if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(worst));  // non-negative: bits order as values
}

// Three MSB-first sign bitmaps (1 = positive): a wave ballots 64 consecutive
// values per group; reversing the 64 ballot bits and then the byte order puts
// value 8k+j at bit 7-j of byte k. Block y = component; each wave takes 16
// consecutive groups, loading all of them before its ballots.
constexpr int kSignGroups = 16;
__global__ void __launch_bounds__(256) k_signmap(const float* __restrict__ U, const float* __restrict__ V,
                                                 const float* __restrict__ W, size_t n, size_t nq,
                                                 uint64_t* __restrict__ out) {
  const int p = blockIdx.y, lane = threadIdx.x & 63;
  const float* X = p == 0 ? U : (p == 1 ? V : W);
  const size_t q0 = ((size_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * kSignGroups;
  float v[kSignGroups];
#pragma unroll
  for (int r = 0; r < kSignGroups; r++) {
    const size_t e = (q0 + r) * 64 + lane;
    v[r] = e < n ? X[e] : 0.f;
  }
#pragma unroll
  for (int r = 0; r < kSignGroups; r++) {
    const uint64_t m = __ballot(v[r] > 0);
    if (lane == 0 && q0 + r < nq) out[p * nq + q0 + r] = __builtin_bswap64(__builtin_bitreverse64(m));
  }
}

// log_transform (include/sz_cp_preserve_utils.hpp:205-221): log2f(|x|), -100
// for 0, written in wave layout. Block x = one (i, j) row, y = k range.
__global__ void k_log(const float* __restrict__ U, const float* __restrict__ V,
                      const float* __restrict__ W, Skew3 s, LibmView lv, float* __restrict__ L) {
  const uint32_t l = xcd_block(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
  const uint32_t row = l % gridDim.x;
  const int k = (int)(l / gridDim.x) * blockDim.x + threadIdx.x;
  if (k >= s.r3) return;
  const int i = (int)(row / (uint32_t)s.r2), j = (int)(row - (uint32_t)i * s.r2);
  const size_t v = (size_t)row * s.r3 + k, o = s.at(i, j, k);
  const float x0 = U[v], x1 = V[v], x2 = W[v];
  L[o] = x0 != 0 ? host_log2f(lv, fabsf(x0)) : -100.f;
  L[s.per + o] = x1 != 0 ? host_log2f(lv, fabsf(x1)) : -100.f;
  L[2 * s.per + o] = x2 != 0 ? host_log2f(lv, fabsf(x2)) : -100.f;
}

// ------------------------------------------------------------- wavefront --
// A workgroup of kP3 waves owns kP3 consecutive planes of one 64-row j block.
// Wave w sweeps plane g*kP3 + w in chunks of kC3 steps and computes chunk c in
// iteration c + w: a systolic pipeline with one barrier per iteration. Plane
// i-1 values reach wave w > 0 through LDS from wave w-1 (written one iteration
// earlier); wave 0 reads them from the previous group's last plane in HBM, and
// lane 0 reads the j0-1 row from the block to the left, both behind
// per-(plane, block) progress counters. Everything chunk c+1 needs is loaded
// while chunk c computes; a chunk's progress is published one iteration after
// its stores were issued.

__device__ __forceinline__ int ld_coh(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_coh(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float ld_cohf(const float* p) { return __int_as_float(ld_coh((const int*)p)); }
// 16-byte agent-coherent load / store as two 8-byte atomics (p 16-byte aligned)
__device__ __forceinline__ float4 ld_coh4(const float* p) {
  const unsigned long long* q = (const unsigned long long*)p;
  const unsigned long long lo = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const unsigned long long hi = __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return make_float4(__uint_as_float((uint32_t)lo), __uint_as_float((uint32_t)(lo >> 32)),
                     __uint_as_float((uint32_t)hi), __uint_as_float((uint32_t)(hi >> 32)));
}
__device__ __forceinline__ void st_coh4(float* p, float4 v) {
  unsigned long long* q = (unsigned long long*)p;
  __hip_atomic_store(q, ((unsigned long long)__float_as_uint(v.y) << 32) | __float_as_uint(v.x),
                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(q + 1, ((unsigned long long)__float_as_uint(v.w) << 32) | __float_as_uint(v.z),
                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float shr1(float src, float lane0) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(lane0), __float_as_int(src),
                                                    0x138 /* wave_shr:1 */, 0xF, 0xF, false));
}
__device__ __forceinline__ float comp4(const float4& v, int s) {
  return s == 0 ? v.x : (s == 1 ? v.y : (s == 2 ? v.z : v.w));
}

// global address space: pointers read from memory (the batch table) are
// generic; accessed through these the loads and stores stay global (not flat)
// accesses, which LDS waits do not wait for
#define GLB __attribute__((address_space(1)))
// constant address space: the batch table is not written while the kernel
// runs, so uniform reads of it become scalar loads
#define CST __attribute__((address_space(4)))

struct Wave3d {
  const float* lx;       // compress: log values [3][per]; decompress: initial values of unpredictables
  const uint8_t* code;   // compress: eb codes; decompress: stream ids
  const uint16_t* qin;   // decompress: quant codes [3][per]
  float* R;              // reconstruction [3][per] (log domain)
  uint16_t* qout;        // compress: quant codes [3][per]
  uint8_t* vid;          // compress: stream id per vertex
  int* prog;             // [r1][nb] completed steps
  Skew3 s;               // all arrays above are in wave layout
};

constexpr int kG3 = kC3 / 4;  // 4-step groups per chunk

// what one chunk of one wave reads
struct Chunk3 {
  float4 x[3][kG3];   // log values (compress) / initial values (decompress)
  uint32_t cd[kG3];   // 4 eb codes (compress) / stream ids (decompress) per group
  uint2 q[3][kG3];    // decompress: 4 quantization codes per group
  float4 up[3][kG3];  // wave 0: plane i-1 from HBM
  float bl[3], bu[3]; // lane s: row j0-1 at step t0 + s (own plane, plane above)
};

// Fields of one launch: a single field travels in the kernel arguments (one),
// several through a device table f[0..nf) with start[k] = first item of field
// k. Items are numbered field by field and taken from one ticket counter, so
// every wait (on an earlier plane group or block of the same field) targets an
// earlier ticket, whose workgroup is already resident.
struct WaveBatch3d {
  Wave3d one;
  const Wave3d* f;  // null: the single field `one`
  const int* start;
  int nf, items;
  int* ticket;
};

// items of one field: (plane group, j block) pairs
__host__ __device__ inline int wave3d_items(const Skew3& s) { return (s.r1 + kP3 - 1) / kP3 * s.nb; }

template <bool COMPRESS>
__global__ void __launch_bounds__(64 * kP3) k_wave3d(WaveBatch3d bt) {
  __shared__ float4 hand[kP3 - 1][2][kG3][3][64];
  __shared__ int s_item;
  // the wave index is uniform: tests on it become scalar branches
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(bt.ticket, 1);
    __syncthreads();
    const int gitem = __builtin_amdgcn_readfirstlane(s_item);  // uniform: the fields' pointers stay scalar
    __syncthreads();
    if (gitem >= bt.items) return;
    const CST Wave3d* tab = (const CST Wave3d*)bt.f;
    const CST int* start = (const CST int*)bt.start;
    int fi = 0;
    if (bt.f) {
      int hi = bt.nf - 1;
      while (fi < hi) {
        const int mid = (fi + hi + 1) >> 1;
        if (start[mid] <= gitem) fi = mid;
        else hi = mid - 1;
      }
    }
    Wave3d a = bt.one;
    if (bt.f) {
      const CST Wave3d* e = tab + fi;
      a.lx = e->lx;
      a.code = e->code;
      a.qin = e->qin;
      a.R = e->R;
      a.qout = e->qout;
      a.vid = e->vid;
      a.prog = e->prog;
      a.s.r1 = e->s.r1;
      a.s.r2 = e->s.r2;
      a.s.r3 = e->s.r3;
      a.s.nb = e->s.nb;
      a.s.T = e->s.T;
      a.s.per = e->s.per;
    }
    const GLB float* lx = (const GLB float*)a.lx;
    const GLB uint8_t* code = (const GLB uint8_t*)a.code;
    const GLB uint16_t* qin = (const GLB uint16_t*)a.qin;
    GLB float* R = (GLB float*)a.R;
    GLB uint16_t* qout = (GLB uint16_t*)a.qout;
    GLB uint8_t* vid = (GLB uint8_t*)a.vid;
    GLB int* prog = (GLB int*)a.prog;
    const int item = bt.f ? gitem - start[fi] : gitem;
    const Skew3 sk = a.s;
    const int r1 = sk.r1, r3 = sk.r3, nb = sk.nb, T = sk.T;
    const size_t per = sk.per;
    const int nch = T / kC3;
    const int g = item / nb, b = item - g * nb;
    const int i = g * kP3 + w;
    const bool pv = i < r1;
    const size_t base = pv ? sk.blk(i, b) : 0;
    const size_t ubase = (pv && i > 0) ? sk.blk(i - 1, b) : 0;
    const size_t lbase = (pv && b > 0) ? sk.blk(i, b - 1) : 0;
    const size_t lubase = (pv && b > 0 && i > 0) ? sk.blk(i - 1, b - 1) : 0;
    const int* prog_up = (const int*)(prog + (pv && i > 0 ? (size_t)(i - 1) * nb + b : 0));
    const int* prog_left = (const int*)(prog + (pv && b > 0 ? (size_t)i * nb + b - 1 : 0));
    int* prog_me = (int*)(prog + (pv ? (size_t)i * nb + b : 0));
    int seen_up = 0, seen_left = 0;
    // counters read one iteration ahead (VGPR values, converted when used, so
    // that their loads overlap the chunk's compute instead of stalling it)
    int ahead_up = 0, ahead_left = 0;
    float Rp[3] = {0, 0, 0}, Bp[3] = {0, 0, 0}, Dp[3] = {0, 0, 0}, Fp[3] = {0, 0, 0};

    auto fetch = [&](Chunk3& X, int c) {
      const int t0 = c * kC3;
      if (w == 0 && i > 0) {
        const int need = min(t0 + kC3, T);
        seen_up = max(seen_up, __builtin_amdgcn_readfirstlane(ahead_up));
        while (seen_up < need) {
          seen_up = __builtin_amdgcn_readfirstlane(ld_coh(prog_up));
          if (seen_up < need) __builtin_amdgcn_s_sleep(1);
        }
      }
      if (b > 0) {
        const int need = min(t0 + kC3 + 63, T);
        seen_left = max(seen_left, __builtin_amdgcn_readfirstlane(ahead_left));
        while (seen_left < need) {
          seen_left = __builtin_amdgcn_readfirstlane(ld_coh(prog_left));
          if (seen_left < need) __builtin_amdgcn_s_sleep(1);
        }
      }
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
#pragma unroll
      for (int u = 0; u < kG3; u++) {
        const size_t o = (size_t)(t0 / 4 + u) * 256 + lane * 4;
        X.cd[u] = *(const GLB uint32_t*)(code + base + o);
#pragma unroll
        for (int p = 0; p < 3; p++) {
          X.x[p][u] = *(const float4*)(lx + p * per + base + o);
          if (!COMPRESS) X.q[p][u] = *(const uint2*)(qin + p * per + base + o);
          if (w == 0)
            X.up[p][u] = i > 0 ? ld_coh4((const float*)(R + p * per + ubase + o)) : make_float4(0.f, 0.f, 0.f, 0.f);
        }
      }
#pragma unroll
      for (int p = 0; p < 3; p++) X.bl[p] = X.bu[p] = 0.f;
      if (b > 0 && lane < kC3 && t0 + lane < r3) {
        const int t = t0 + lane + 63;
        const size_t o = (size_t)(t >> 2) * 256 + 63 * 4 + (t & 3);
#pragma unroll
        for (int p = 0; p < 3; p++) {
          X.bl[p] = ld_cohf((const float*)(R + p * per + lbase + o));
          if (i > 0) X.bu[p] = ld_cohf((const float*)(R + p * per + lubase + o));
        }
      }
      if (w == 0 && i > 0 && seen_up < T) ahead_up = ld_coh(prog_up);
      if (b > 0 && seen_left < T) ahead_left = ld_coh(prog_left);
    };

    auto compute = [&](const Chunk3& X, int c) {
      const int t0 = c * kC3;
      float4 up[3][kG3];
#pragma unroll
      for (int u = 0; u < kG3; u++)
#pragma unroll
        for (int p = 0; p < 3; p++) {
          const float4 h = hand[w > 0 ? w - 1 : 0][c & 1][u][p][lane];  // wave 0 reads and discards
          up[p][u] = w == 0 ? X.up[p][u] : h;
        }
#pragma unroll
      for (int u = 0; u < kG3; u++) {
        float rr[3][4];
        uint32_t vv = 0, qq[3][2] = {{0, 0}, {0, 0}, {0, 0}};
#pragma unroll
        for (int s4 = 0; s4 < 4; s4++) {
          const int s = u * 4 + s4;
          const int k = t0 + s - lane;
          const uint32_t code = (X.cd[u] >> (8 * s4)) & 0xFFu;
          float res[3], dec[3];
          bool okall = true;
          uint16_t qc[3];
#pragma unroll
          for (int p = 0; p < 3; p++) {
            const float l0b = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(X.bl[p]), s));
            const float l0u = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(X.bu[p]), s));
            const float Bv = shr1(Rp[p], l0b);   // L(i, j-1, k)
            const float Fv = shr1(Dp[p], l0u);   // L(i-1, j-1, k)
            const float Dv = comp4(up[p][u], s4);  // L(i-1, j, k)
            // pred = d0 + d3 + d5 + d6 - d1 - d2 - d4 (src/sz_compress_cp_preserve_3d.cpp:498)
            const float pred = Fp[p] + Dv + Bv + Rp[p] - Fv - Dp[p] - Bp[p];
            const float xv = comp4(X.x[p][u], s4);
            if (COMPRESS) {
              bool ok;
              dec[p] = l3d_step(pred, xv, (int)code, ok, qc[p]);
              okall = okall & ok;
            } else {
              const uint32_t qw = (s4 < 2) ? X.q[p][u].x : X.q[p][u].y;
              const int qv = (int)((qw >> (16 * (s4 & 1))) & 0xFFFFu);
              const float rec = pred + ldexpf((float)(2 * (qv - 32768)), (int)code - 23);
              res[p] = ((code == 0u) | (code == 24u)) ? xv : rec;
            }
            Bp[p] = Bv;
            Fp[p] = Fv;
            Dp[p] = Dv;
          }
          if (COMPRESS) {
            // codes 1..254 are predictable candidates (255: zero bound, 0: lossless)
            const bool pred_ok = ((code - 1u) < 254u) & okall;
#pragma unroll
            for (int p = 0; p < 3; p++) {
              res[p] = pred_ok ? dec[p] : comp4(X.x[p][u], s4);
              qq[p][s4 >> 1] |= (uint32_t)qc[p] << (16 * (s4 & 1));
            }
            const uint32_t id = pred_ok ? code : 24u;
            vv |= (code == 255u ? 0u : id) << (8 * s4);
          }
#pragma unroll
          for (int p = 0; p < 3; p++) {
            Rp[p] = k < 0 ? 0.f : res[p];
            rr[p][s4] = Rp[p];
          }
        }
        const size_t o = (size_t)(t0 / 4 + u) * 256 + lane * 4;
#pragma unroll
        for (int p = 0; p < 3; p++) {
          const float4 v = make_float4(rr[p][0], rr[p][1], rr[p][2], rr[p][3]);
          if (w < kP3 - 1) hand[w][c & 1][u][p][lane] = v;
          if (!COMPRESS || w == kP3 - 1 || lane == 63) st_coh4((float*)(R + p * per + base + o), v);
          if (COMPRESS) *(uint2*)(qout + p * per + base + o) = make_uint2(qq[p][0], qq[p][1]);
        }
        if (COMPRESS) *(GLB uint32_t*)(vid + base + o) = vv;
      }
    };

    auto iter = [&](Chunk3& cur, Chunk3& nxt, int it) {
      const int c = it - w;
      if (pv) {
        if (c >= 1 && c <= nch) {  // chunk c-1's stores were issued last iteration
          __atomic_signal_fence(__ATOMIC_SEQ_CST);
          __builtin_amdgcn_s_waitcnt(0);
          __atomic_signal_fence(__ATOMIC_SEQ_CST);
          if (lane == 0) st_coh(prog_me, min(c * kC3, T));
        }
        if (c + 1 >= 0 && c + 1 < nch) fetch(nxt, c + 1);
        if (c >= 0 && c < nch) compute(cur, c);
      }
      __syncthreads();
    };

    Chunk3 A, B;
    for (int it = -1; it < nch + kP3 - 1; it += 2) {
      iter(A, B, it);
      iter(B, A, it + 1);
    }
    if (pv) {
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      __builtin_amdgcn_s_waitcnt(0);
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      if (lane == 0) st_coh(prog_me, T);
    }
  }
}

__global__ void k_zero_i32_3d(int* p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// ---------------------------------------------------------- huffman ------
constexpr int kVChunk = 256;  // vertices per chunk

__device__ __forceinline__ bool is_pred3(uint32_t id) { return id != 0 && id != 24; }

// Huffman sources: stream (row-major) order over wave-layout arrays.
struct SrcIds3d {
  const uint8_t* vid;
  Skew3 s;
  size_t n;
  static constexpr int kMaxItems = kVChunk, kPer = 1;
  __device__ int get(size_t c, int idx, uint32_t* sym) const {
    const size_t v = c * kVChunk + idx;
    sym[0] = vid[s.at_v(v < n ? v : n - 1)];
    return v < n;
  }
};

struct SrcCodes3d {
  const uint8_t* vid;
  const uint16_t* q;
  Skew3 s;
  size_t n;
  static constexpr int kMaxItems = kVChunk, kPer = 3;
  __device__ int get(size_t c, int idx, uint32_t* sym) const {
    const size_t v = c * kVChunk + idx, x = s.at_v(v < n ? v : n - 1);
    sym[0] = q[x];  // loads are unconditional so that a wave issues them all at once
    sym[1] = q[s.per + x];
    sym[2] = q[2 * s.per + x];
    return v < n && is_pred3(vid[x]) ? 3 : 0;
  }
};

// The chunk kernels below run one wave per 256-vertex chunk (4 waves per
// block); lane l takes vertices c*256 + 64*r + l, r = 0..3, and ranks the
// predictable / unpredictable ones with ballots.
#define CHUNK_WAVE_PROLOGUE                                                 \
  const size_t c = (size_t)xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6); \
  const int lane = threadIdx.x & 63;                                         \
  if (c >= nch) return;

// per-chunk predictable / unpredictable vertex counts; vid in wave layout
// (SKEW) or row-major
template <bool SKEW>
__global__ void __launch_bounds__(256) k_chunk_counts3d(const uint8_t* __restrict__ vid, Skew3 s, size_t n,
                                                        size_t nch, uint32_t* __restrict__ npred,
                                                        uint32_t* __restrict__ nunp) {
  CHUNK_WAVE_PROLOGUE
  uint32_t p = 0, u = 0;
  for (int r = 0; r < 4; r++) {
    const size_t v = c * kVChunk + r * 64 + lane;
    const bool in = v < n;
    const bool pr = in && is_pred3(vid[SKEW ? s.at_v(in ? v : 0) : (in ? v : 0)]);
    const uint64_t mp = __ballot(pr), mu = __ballot(in && !pr);
    p += __popcll(mp);
    u += __popcll(mu);
  }
  if (lane == 0) {
    npred[c] = p;
    nunp[c] = u;
  }
}

// compress: original values of unpredictable vertices, 3 floats each, in order
__global__ void __launch_bounds__(256) k_gather_unpred3d(const uint8_t* __restrict__ vid, Skew3 s,
                                                         const float* __restrict__ U,
                                                         const float* __restrict__ V,
                                                         const float* __restrict__ W, size_t n,
                                                         size_t nch, const uint64_t* __restrict__ uoff,
                                                         float* __restrict__ out) {
  CHUNK_WAVE_PROLOGUE
  uint64_t o = uoff[c];
  for (int r = 0; r < 4; r++) {
    const size_t v = c * kVChunk + r * 64 + lane;
    const bool un = v < n && !is_pred3(vid[s.at_v(v < n ? v : 0)]);
    const uint64_t m = __ballot(un);
    if (un) {
      const uint64_t at = 3 * (o + lanes_below(m));
      out[at] = U[v];
      out[at + 1] = V[v];
      out[at + 2] = W[v];
    }
    o += __popcll(m);
  }
}

struct SinkU8 {
  static constexpr int kBytes = 1;
  uint8_t* p;
  __device__ void put(uint64_t i, uint32_t s) const { p[i] = (uint8_t)s; }
};
struct SinkU16 {
  static constexpr int kBytes = 2;
  uint16_t* p;
  __device__ void put(uint64_t i, uint32_t s) const { p[i] = (uint16_t)s; }
};

// decompress: stream ids, codes and initial (log) values of unpredictable
// vertices, from stream order into wave layout
__global__ void __launch_bounds__(256) k_scatter3d(const uint8_t* __restrict__ vid, Skew3 s,
                                                   const uint16_t* __restrict__ qc,
                                                   const float* __restrict__ unp, size_t n, size_t nch,
                                                   const uint64_t* __restrict__ poff,
                                                   const uint64_t* __restrict__ uoff, uint64_t nq,
                                                   uint64_t nu, LibmView lv, uint8_t* __restrict__ vs,
                                                   uint16_t* __restrict__ q, float* __restrict__ init,
                                                   int* __restrict__ err) {
  CHUNK_WAVE_PROLOGUE
  uint64_t po = poff[c], uo = uoff[c];
  for (int r = 0; r < 4; r++) {
    const size_t v = c * kVChunk + r * 64 + lane;
    const bool in = v < n;
    const uint32_t id = in ? vid[v] : 0;
    const bool pr = in && is_pred3(id), un = in && !pr;
    const uint64_t mp = __ballot(pr), mu = __ballot(un);
    if (in) {
      const size_t x = s.at_v(v);
      vs[x] = (uint8_t)id;
      if (pr) {
        const uint64_t a = 3 * (po + lanes_below(mp));
        if (a + 3 > nq) {
          *err = 1;
        } else {
          q[x] = qc[a];
          q[s.per + x] = qc[a + 1];
          q[2 * s.per + x] = qc[a + 2];
        }
      } else {
        const uint64_t a = 3 * (uo + lanes_below(mu));
        if (a + 3 > nu) {
          *err = 1;
        } else {
          for (int p = 0; p < 3; p++) {
            const float xv = unp[a + p];
            init[p * s.per + x] = (xv == 0) ? -100.f : host_log2f(lv, fabsf(xv));
          }
        }
      }
    }
    po += __popcll(mp);
    uo += __popcll(mu);
  }
}

// final values (src/sz_decompress_cp_preserve_3d.cpp:89-104)
__global__ void __launch_bounds__(256) k_finish3d(const uint8_t* __restrict__ vid, Skew3 s,
                                                  const float* __restrict__ R,
                                                  const uint8_t* __restrict__ smap, size_t smap_bytes,
                                                  const float* __restrict__ unp, size_t n, size_t nch,
                                                  const uint64_t* __restrict__ uoff, uint64_t nu,
                                                  LibmView ev, float* __restrict__ U,
                                                  float* __restrict__ V, float* __restrict__ W) {
  CHUNK_WAVE_PROLOGUE
  uint64_t uo = uoff[c];
  float* out[3] = {U, V, W};
  for (int r = 0; r < 4; r++) {
    const size_t v = c * kVChunk + r * 64 + lane;
    const bool in = v < n;
    const bool un = in && !is_pred3(vid[in ? v : 0]);
    const uint64_t mu = __ballot(un);
    if (un) {
      const uint64_t a = 3 * (uo + lanes_below(mu));
      if (a + 3 <= nu)
        for (int p = 0; p < 3; p++) out[p][v] = unp[a + p];
    } else if (in) {
      const size_t x = s.at_v(v);
      for (int p = 0; p < 3; p++) {
        const float lx = R[p * s.per + x];
        const bool pos = (smap[p * smap_bytes + v / 8] >> (7 - (v % 8))) & 1;
        float y;
        if (lx < -99) {
          y = 0;
        } else {
          const float t = host_exp2f(ev, lx);
          y = pos ? t : -t;
        }
        out[p][v] = y;
      }
    }
    uo += __popcll(mu);
  }
}
#undef CHUNK_WAVE_PROLOGUE

template <typename V>
void put(std::vector<uint8_t>& o, const V& v) {
  const size_t at = o.size();
  o.resize(at + sizeof(V));
  std::memcpy(o.data() + at, &v, sizeof(V));
}

size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

// Lorenzo wavefront over fields f[0..nf) in one launch (see WaveBatch3d).
template <bool COMPRESS>
void launch_wave3d(Ctx& ctx, const std::vector<Wave3d>& f, int grid_cap) {
  hipStream_t st = ctx.stream();
  const size_t nf = f.size();
  CPSZ_REQUIRE(nf > 0 && grid_cap > 0, kErrArg, "3D wavefront: empty launch");
  std::vector<int> start(nf + 1, 0);
  for (size_t k = 0; k < nf; k++) {
    const size_t nz = (size_t)f[k].s.r1 * f[k].s.nb;
    k_zero_i32_3d<<<(int)std::min<size_t>(div_up(nz, 256), 1024), 256, 0, st>>>(f[k].prog, nz);
    CPSZ_REQUIRE((size_t)start[k] + wave3d_items(f[k].s) < (1u << 30), kErrUnsupported,
                 "3D wavefront: too many items in one launch");
    start[k + 1] = start[k] + wave3d_items(f[k].s);
  }
  WaveBatch3d bt{};
  bt.nf = (int)nf;
  bt.items = start[nf];
  if (nf == 1) {
    bt.one = f[0];
    bt.ticket = ctx.buf<int>("w3.ticket", 1);
  } else {
    const size_t tbytes = nf * sizeof(Wave3d), sbytes = (nf + 1) * sizeof(int);
    uint8_t* h = (uint8_t*)ctx.host_buf("w3.tab", tbytes + sbytes);
    std::memcpy(h, f.data(), tbytes);
    std::memcpy(h + tbytes, start.data(), sbytes);
    uint8_t* d = ctx.buf<uint8_t>("w3.tab", tbytes + sbytes + 64);
    CPSZ_HIP(hipMemcpyAsync(d, h, tbytes + sbytes, hipMemcpyHostToDevice, st));
    bt.f = (const Wave3d*)d;
    bt.start = (const int*)(d + tbytes);
    bt.ticket = (int*)(d + ((tbytes + sbytes + 15) & ~(size_t)15));
  }
  k_zero_i32_3d<<<1, 64, 0, st>>>(bt.ticket, 1);
  const int grid = std::min(bt.items, grid_cap);
  k_wave3d<COMPRESS><<<grid, 64 * kP3, 0, st>>>(bt);
  CPSZ_HIP(hipGetLastError());
  // the pinned host table is rewritten by the next batch launch only after
  // this copy has run
  if (nf > 1) CPSZ_HIP(hipStreamSynchronize(st));
}

void check_dims_3d(size_t r1, size_t r2, size_t r3, const char* who) {
  CPSZ_REQUIRE(r1 && r2 && r3, kErrArg, std::string(who) + ": bad arguments");
  CPSZ_REQUIRE(r1 < (1u << 20) && r2 < (1u << 20) && r3 < (1u << 20) && r1 * r2 * r3 < (1ull << 32),
               kErrUnsupported, std::string(who) + ": field too large");
}

template <typename V>
V* carve(uint8_t*& p, size_t count) {
  V* q = (V*)p;
  p += al256(count * sizeof(V));
  return q;
}

}  // namespace

// ------------------------------------------------------------- compress ----
size_t c3_scratch_bytes(size_t r1, size_t r2, size_t r3) {
  const Skew3 sk = make_skew(r1, r2, r3);
  const size_t n = r1 * r2 * r3, nq = div_up(n, 64);
  return al256(sk.per) + 2 * al256(3 * sk.per * 4) + al256(3 * sk.per * 2) + al256(sk.per) + al256(3 * nq * 8) +
         al256(r1 * sk.nb * 4) + al256(16);
}

void c3_bind(C3Work& w, uint8_t* p) {
  const Skew3 sk = make_skew(w.r1, w.r2, w.r3);
  const size_t n = w.r1 * w.r2 * w.r3, nq = div_up(n, 64);
  w.ebc = carve<uint8_t>(p, sk.per);
  w.L = carve<float>(p, 3 * sk.per);
  w.R = carve<float>(p, 3 * sk.per);
  w.q = carve<uint16_t>(p, 3 * sk.per);
  w.vid = carve<uint8_t>(p, sk.per);
  w.smap = carve<uint64_t>(p, 3 * nq);
  w.prog = carve<int>(p, w.r1 * sk.nb);
  w.flags = carve<int>(p, 4);
}

void c3_eb(Ctx& ctx, C3Work& w, double max_pwr_eb, const EbOverride& ov) {
  CPSZ_REQUIRE(w.dU && w.dV && w.dW, kErrArg, "compress_3d: bad arguments");
  check_dims_3d(w.r1, w.r2, w.r3, "compress_3d");
  hipStream_t st = ctx.stream();
  const size_t r1 = w.r1, r2 = w.r2, r3 = w.r3, n = r1 * r2 * r3;
  const Skew3 sk = make_skew(r1, r2, r3);
  const Ctx::LibmDev& lm = ctx.libm_dev();
  CPSZ_HIP(hipMemsetAsync(w.flags, 0, 16, st));
  {
    dim3 grid((unsigned)div_up(r3, kMK), (unsigned)div_up(r2, kMJ), (unsigned)div_up(r1, kMS));
    if (ov.eb_out) {
      k_eb3d<false><<<grid, 256, 0, st>>>(w.dU, w.dV, w.dW, (int)r1, (int)r2, (int)r3, max_pwr_eb, ctx.d_q3d(),
                                          ctx.q3d_size(), sk, w.ebc, w.flags, ov.gain, ov.fallback, ov.eb_out,
                                          nullptr, nullptr);
    } else {
      // the undecided-vertex list (at most n entries) lives in R, which the
      // wavefront fills only later
      uint32_t* fix = (uint32_t*)w.R;
      uint32_t* nfix = (uint32_t*)(w.flags + 1);
      k_eb3d<true><<<grid, 256, 0, st>>>(w.dU, w.dV, w.dW, (int)r1, (int)r2, (int)r3, max_pwr_eb, ctx.d_q3d(),
                                         ctx.q3d_size(), sk, w.ebc, w.flags, ov.gain, ov.fallback, nullptr, fix,
                                         nfix);
      k_eb3d_fix<<<1024, 256, 0, st>>>(w.dU, w.dV, w.dW, (int)r1, (int)r2, (int)r3, max_pwr_eb, ctx.d_q3d(),
                                       ctx.q3d_size(), sk, w.ebc, w.flags, ov.gain, ov.fallback, fix, nfix);
    }
    CPSZ_HIP(hipGetLastError());
  }
  ctx.mark("eb");
  const size_t nq = div_up(n, 64);
  k_signmap<<<dim3((unsigned)div_up(nq, 4 * kSignGroups), 3), 256, 0, st>>>(w.dU, w.dV, w.dW, n, nq, w.smap);
  {
    const int bx = r3 >= 256 ? 256 : (int)div_up(r3, 64) * 64;
    dim3 grid((unsigned)(r1 * r2), (unsigned)div_up(r3, bx));
    k_log<<<grid, bx, 0, st>>>(w.dU, w.dV, w.dW, sk, lm.log2f, w.L);
  }
  CPSZ_HIP(hipGetLastError());
  ctx.mark("log");
}

void c3_check(Ctx& ctx, C3Work* const* w, size_t nf) {
  int* h = (int*)ctx.host_buf("c3.hflags", 16 * nf);
  for (size_t k = 0; k < nf; k++)
    CPSZ_HIP(hipMemcpyAsync(h + 4 * k, w[k]->flags, 16, hipMemcpyDeviceToHost, ctx.stream()));
  CPSZ_HIP(hipStreamSynchronize(ctx.stream()));
  for (size_t k = 0; k < nf; k++)
    CPSZ_REQUIRE(h[4 * k] == 0, kErrNonFinite,
                 "compress_3d: non-finite input or bound (values too large for the reference's "
                 "float triple products)");
}

void c3_wavefront(Ctx& ctx, C3Work* const* w, size_t nf, int grid_cap) {
  std::vector<Wave3d> f(nf);
  for (size_t k = 0; k < nf; k++) {
    f[k] = Wave3d{};
    f[k].lx = w[k]->L;
    f[k].code = w[k]->ebc;
    f[k].R = w[k]->R;
    f[k].qout = w[k]->q;
    f[k].vid = w[k]->vid;
    f[k].prog = w[k]->prog;
    f[k].s = make_skew(w[k]->r1, w[k]->r2, w[k]->r3);
  }
  launch_wave3d<true>(ctx, f, grid_cap);
}

void c3_entropy(Ctx& ctx, const C3Work& w, DeviceStream& out) {
  hipStream_t st = ctx.stream();
  const size_t n = w.r1 * w.r2 * w.r3;
  const Skew3 sk = make_skew(w.r1, w.r2, w.r3);
  // ---- histograms
  uint32_t* hist = ctx.buf<uint32_t>("c3.hist", kMaxEbCode + 65536);
  CPSZ_HIP(hipMemsetAsync(hist, 0, (kMaxEbCode + 65536) * 4, st));
  const size_t nch = div_up(n, kVChunk);
  SrcIds3d s_ids{w.vid, sk, n};
  SrcCodes3d s_codes{w.vid, w.q, sk, n};
  {
    const int grid = (int)std::min<size_t>(encode_blocks(nch), 512);
    k_hist<<<grid, 256, kMaxEbCode * 4, st>>>(s_ids, nch, hist, 0, kMaxEbCode);
    k_hist<<<grid, 256, 16384 * 4, st>>>(s_codes, nch, hist + kMaxEbCode, 32768 - 8192, 16384);
    CPSZ_HIP(hipGetLastError());
  }
  uint32_t* hh = (uint32_t*)ctx.host_buf("c3.hist", (kMaxEbCode + 65536) * 4);
  CPSZ_HIP(hipMemcpyAsync(hh, hist, (kMaxEbCode + 65536) * 4, hipMemcpyDeviceToHost, st));
  const hipEvent_t hist_done = ctx.record_point();
  // per-chunk vertex counts and their scans need no tree: they run while the
  // host builds the trees
  uint64_t* bits = ctx.buf<uint64_t>("c3.bits", 2 * nch);
  uint64_t* offs = ctx.buf<uint64_t>("c3.offs", 2 * nch + 4);
  uint32_t* cnt = ctx.buf<uint32_t>("c3.cnt", 2 * nch);
  uint64_t* coff = ctx.buf<uint64_t>("c3.coff", 2 * nch + 4);
  uint64_t* totals = ctx.buf<uint64_t>("c3.totals", 8);
  k_chunk_counts3d<true><<<(int)div_up(nch, 4), 256, 0, st>>>(w.vid, sk, n, nch, cnt, cnt + nch);
  CPSZ_HIP(hipGetLastError());
  scan_u32_to_u64(ctx, cnt, coff, nch, totals + 2);
  scan_u32_to_u64(ctx, cnt + nch, coff + nch + 1, nch, totals + 3);
  CPSZ_HIP(hipEventSynchronize(hist_done));
  ctx.mark("hist");
  thread_local std::vector<uint64_t> f;
  f.resize(65536);
  thread_local HuffEncodeTable t_id, t_q;  // kept per thread: their storage is reused
  uint64_t* cb_code = ctx.buf<uint64_t>("c3.cbcode", 2 * 65536);
  uint8_t* cb_len = ctx.buf<uint8_t>("c3.cblen", 2 * 65536);
  for (int k = 0; k < kMaxEbCode; k++) f[k] = hh[k];
  huff_build(f.data(), kMaxEbCode, 2 * 256, t_id);
  // the id stream's bit counts run on the device while the host builds the
  // code tree
  upload_codebook(ctx, t_id, cb_code, cb_len, "c3.id");
  k_chunk_bits<<<(int)encode_blocks(nch), kEncThreads, 0, st>>>(s_ids, nch, cb_len, bits);
  CPSZ_HIP(hipGetLastError());
  scan_u64(ctx, bits, offs, nch, totals + 0);
  uint64_t npred_syms = 0;
  for (int k = 0; k < 65536; k++) npred_syms += (f[k] = hh[kMaxEbCode + k]);
  CPSZ_REQUIRE(npred_syms > 0, kErrUnsupported,
               "compress_3d: no predictable vertex (the reference cannot encode an empty code "
               "section)");
  huff_build(f.data(), 65536, 2 * kCapacity, t_q);
  ctx.mark("tree");
  upload_codebook(ctx, t_q, cb_code + 65536, cb_len + 65536, "c3.q");
  k_chunk_bits<<<(int)encode_blocks(nch), kEncThreads, 0, st>>>(s_codes, nch, cb_len + 65536, bits + nch);
  CPSZ_HIP(hipGetLastError());
  scan_u64(ctx, bits + nch, offs + nch, nch, totals + 1);
  uint64_t* ht = (uint64_t*)ctx.host_buf("c3.totals", 64);
  CPSZ_HIP(hipMemcpyAsync(ht, totals, 32, hipMemcpyDeviceToHost, st));
  CPSZ_HIP(hipStreamSynchronize(st));
  const uint64_t bits_id = ht[0], bits_q = ht[1], npred = ht[2], nunp = ht[3];
  CPSZ_REQUIRE(npred + nunp == n, kErrInternal, "vertex count mismatch");
  ctx.mark("scan");
  const size_t by_id = div_up(bits_id, 8), by_q = div_up(bits_q, 8);
  const size_t w_id = div_up(by_id, 4) + 2, w_q = div_up(by_q, 4) + 2;
  uint32_t* words = ctx.buf<uint32_t>("c3.words", w_id + w_q);
  k_fill_u32<<<1024, 256, 0, st>>>(words, w_id + w_q, 0);
  k_encode<<<(int)encode_blocks(nch), kEncThreads, 0, st>>>(s_ids, nch, cb_code, cb_len, offs, bits, words);
  k_encode<<<(int)encode_blocks(nch), kEncThreads, 0, st>>>(s_codes, nch, cb_code + 65536, cb_len + 65536, offs + nch,
                                                    bits + nch, words + w_id);
  float* unp = ctx.buf<float>("c3.unp", 3 * nunp + 1);
  k_gather_unpred3d<<<(int)div_up(nch, 4), 256, 0, st>>>(w.vid, sk, w.dU, w.dV, w.dW, n, nch, coff + nch + 1,
                                                         unp);
  CPSZ_HIP(hipGetLastError());
  ctx.mark("encode");
  // ---- stream (src/sz_compress_cp_preserve_3d.cpp:556-579)
  auto huff_head = [](const HuffEncodeTable& t, size_t nbytes) {
    std::vector<uint8_t> h;
    put(h, (size_t)t.node_count);
    put(h, (uint32_t)t.tree.size());
    h.insert(h.end(), t.tree.begin(), t.tree.end());
    put(h, (size_t)nbytes);
    return h;
  };
  const size_t smap = (n - 1) / 8 + 1, nq = div_up(n, 64);
  std::vector<uint8_t> h0, h1, h2;
  put(h0, (int)2);
  put(h0, (int)kRadius);
  put(h1, (size_t)(3 * nunp));
  put(h2, (size_t)n);
  {
    std::vector<uint8_t> a = huff_head(t_id, by_id);
    h2.insert(h2.end(), a.begin(), a.end());
  }
  std::vector<uint8_t> h3;
  put(h3, (size_t)(3 * npred));
  {
    std::vector<uint8_t> a = huff_head(t_q, by_q);
    h3.insert(h3.end(), a.begin(), a.end());
  }
  const size_t total = h0.size() + 3 * smap + h1.size() + 12 * nunp + h2.size() + by_id + h3.size() + by_q;
  uint8_t* dout = ctx.buf<uint8_t>("c3.out", total + 16);
  uint8_t* stage = (uint8_t*)ctx.host_buf("c3.stage", h0.size() + h1.size() + h2.size() + h3.size());
  size_t pos = 0, spos = 0;
  auto host_piece = [&](const std::vector<uint8_t>& v) {
    std::memcpy(stage + spos, v.data(), v.size());
    CPSZ_HIP(hipMemcpyAsync(dout + pos, stage + spos, v.size(), hipMemcpyHostToDevice, st));
    pos += v.size();
    spos += v.size();
  };
  auto dev_piece = [&](const void* src, size_t len) {
    if (len) CPSZ_HIP(hipMemcpyAsync(dout + pos, src, len, hipMemcpyDeviceToDevice, st));
    pos += len;
  };
  host_piece(h0);
  if (nq * 8 == smap) dev_piece(w.smap, 3 * smap);  // the three bitmaps are already contiguous
  else
    for (int p = 0; p < 3; p++) dev_piece(w.smap + p * nq, smap);
  host_piece(h1);
  dev_piece(unp, 12 * nunp);
  host_piece(h2);
  dev_piece(words, by_id);
  host_piece(h3);
  dev_piece(words + w_id, by_q);
  CPSZ_REQUIRE(pos == total, kErrInternal, "3D stream layout mismatch");
  ctx.mark("assemble");
  out.data = dout;
  out.size = total;
  CPSZ_HIP(hipStreamSynchronize(st));  // the pinned header staging is reused by the next call
}

void compress_3d_device(Ctx& ctx, const float* dU, const float* dV, const float* dW, size_t r1,
                        size_t r2, size_t r3, double max_pwr_eb, DeviceStream& out, const EbOverride& ov) {
  CPSZ_REQUIRE(dU && dV && dW, kErrArg, "compress_3d: bad arguments");
  check_dims_3d(r1, r2, r3, "compress_3d");
  ctx.stages.clear();
  ctx.mark("start");
  C3Work w;
  w.dU = dU;
  w.dV = dV;
  w.dW = dW;
  w.r1 = r1;
  w.r2 = r2;
  w.r3 = r3;
  c3_bind(w, ctx.buf<uint8_t>("c3.scratch", c3_scratch_bytes(r1, r2, r3)));
  c3_eb(ctx, w, max_pwr_eb, ov);
  C3Work* wp = &w;
  c3_check(ctx, &wp, 1);
  c3_wavefront(ctx, &wp, 1, 1024);
  ctx.mark("lorenzo");
  c3_entropy(ctx, w, out);
  ctx.finish_marks();
}

void eb_3d_device(Ctx& ctx, const float* dU, const float* dV, const float* dW, size_t r1, size_t r2, size_t r3,
                  double max_pwr_eb, double* dE) {
  CPSZ_REQUIRE(dU && dV && dW && dE, kErrArg, "eb_3d: bad arguments");
  check_dims_3d(r1, r2, r3, "eb_3d");
  hipStream_t st = ctx.stream();
  const Skew3 sk = make_skew(r1, r2, r3);
  uint8_t* code = ctx.buf<uint8_t>("e3.code", sk.per);
  int* flags = ctx.buf<int>("e3.flags", 4);
  CPSZ_HIP(hipMemsetAsync(flags, 0, 16, st));
  dim3 grid((unsigned)div_up(r3, kMK), (unsigned)div_up(r2, kMJ), (unsigned)div_up(r1, kMS));
  k_eb3d<false><<<grid, 256, 0, st>>>(dU, dV, dW, (int)r1, (int)r2, (int)r3, max_pwr_eb, ctx.d_q3d(),
                                      ctx.q3d_size(), sk, code, flags, nullptr, nullptr, dE, nullptr, nullptr);
  CPSZ_HIP(hipGetLastError());
  int* h = (int*)ctx.host_buf("e3.hflags", 16);
  CPSZ_HIP(hipMemcpyAsync(h, flags, 16, hipMemcpyDeviceToHost, st));
  CPSZ_HIP(hipStreamSynchronize(st));
  CPSZ_REQUIRE(h[0] == 0, kErrNonFinite, "eb_3d: non-finite input or bound");
}

void check_fast_bounds(Ctx& ctx, uint32_t* max_ulp, double* max_ratio) {
  hipStream_t st = ctx.stream();
  unsigned* d = ctx.buf<unsigned>("chk.fast", 2);
  CPSZ_HIP(hipMemsetAsync(d, 0, 8, st));
  k_check_cbrt<<<4096, 256, 0, st>>>(0u, 0x7E800000u, d);  // [0, 2^126)
  k_check_csb<<<1024, 256, 0, st>>>(256u, 0x5eedu, d + 1);
  CPSZ_HIP(hipGetLastError());
  unsigned h[2];
  CPSZ_HIP(hipMemcpyAsync(h, d, 8, hipMemcpyDeviceToHost, st));
  CPSZ_HIP(hipStreamSynchronize(st));
  *max_ulp = h[0];
  float r;
  std::memcpy(&r, &h[1], 4);
  *max_ratio = r;
}

// ----------------------------------------------------------- decompress ----

size_t d3_scratch_bytes(size_t r1, size_t r2, size_t r3) {
  const Skew3 sk = make_skew(r1, r2, r3);
  const size_t n = r1 * r2 * r3, nch = div_up(n, kVChunk), smap = (n - 1) / 8 + 1;
  return al256(n) + al256((3 * n + 1) * 2) + al256((3 * n + 1) * 4) + al256(3 * sk.per * 2) +
         2 * al256(3 * sk.per * 4) + al256(sk.per) + al256(r1 * sk.nb * 4) + al256(16) + al256(3 * smap + 8) +
         al256(2 * nch * 4) + al256((2 * nch + 4) * 8);
}

void d3_bind(D3Work& w, uint8_t* p) {
  const Skew3 sk = make_skew(w.r1, w.r2, w.r3);
  const size_t n = w.r1 * w.r2 * w.r3, nch = div_up(n, kVChunk), smap = (n - 1) / 8 + 1;
  w.vid = carve<uint8_t>(p, n);
  w.qc = carve<uint16_t>(p, 3 * n + 1);
  w.unp = carve<float>(p, 3 * n + 1);
  w.q = carve<uint16_t>(p, 3 * sk.per);
  w.init = carve<float>(p, 3 * sk.per);
  w.R = carve<float>(p, 3 * sk.per);
  w.vs = carve<uint8_t>(p, sk.per);
  w.prog = carve<int>(p, w.r1 * sk.nb);
  w.err = carve<int>(p, 4);
  w.sm = carve<uint8_t>(p, 3 * smap + 8);
  w.cnt = carve<uint32_t>(p, 2 * nch);
  w.coff = carve<uint64_t>(p, 2 * nch + 4);
}

void d3_entropy(Ctx& ctx, const uint8_t* d_in, const uint8_t* h_in, size_t in_size, D3Work& w) {
  CPSZ_REQUIRE(d_in && w.dU && w.dV && w.dW, kErrArg, "decompress_3d: bad arguments");
  check_dims_3d(w.r1, w.r2, w.r3, "decompress_3d");
  hipStream_t st = ctx.stream();
  const size_t n = w.r1 * w.r2 * w.r3;
  const Skew3 sk = make_skew(w.r1, w.r2, w.r3);
  const Ctx::LibmDev& lm = ctx.libm_dev();
  StreamReader rd(ctx, h_in, d_in, in_size);
  const int base = rd.get<int>(0);
  const int radius = rd.get<int>(4);
  CPSZ_REQUIRE(base == 2 && radius == kRadius, kErrUnsupported,
               "decompress_3d: base/radius differ from the offline_log format");
  const size_t smap = (n - 1) / 8 + 1;
  size_t off = 8 + 3 * smap;
  const size_t nunp_f = rd.get<size_t>(off);
  CPSZ_REQUIRE(nunp_f % 3 == 0 && nunp_f <= 3 * n, kErrStream, "bad unpredictable count");
  const size_t unp_off = off + 8;
  off = unp_off + 4 * nunp_f;
  const size_t eb_num = rd.get<size_t>(off);
  CPSZ_REQUIRE(eb_num == n, kErrStream, "eb id count differs from the field size");
  auto parse = [&](size_t at, size_t& node_count, std::vector<uint8_t>& tree, size_t& nbytes,
                   size_t& bits_off) {
    node_count = rd.get<size_t>(at);
    const uint32_t ts = rd.get<uint32_t>(at + 8);
    CPSZ_REQUIRE(node_count >= 1 && node_count < (1u << 24), kErrStream, "bad huffman node count");
    const uint8_t* t = rd.read(at + 12, ts);
    tree.assign(t, t + ts);
    nbytes = rd.get<size_t>(at + 12 + ts);
    bits_off = at + 20 + ts;
    CPSZ_REQUIRE(bits_off <= in_size && nbytes <= in_size - bits_off, kErrStream, "huffman bitstream truncated");
    return bits_off + nbytes;
  };
  size_t nc_id, nb_id, bo_id, nc_q, nb_q, bo_q;
  std::vector<uint8_t> tr_id, tr_q;
  off = parse(off + 8, nc_id, tr_id, nb_id, bo_id);
  const size_t data_num = rd.get<size_t>(off);
  CPSZ_REQUIRE(data_num <= 3 * n, kErrStream, "bad code count");
  parse(off + 8, nc_q, tr_q, nb_q, bo_q);
  ctx.mark("parse");
  // the code tree (the large one) is parsed and tabulated on another host
  // thread while the device decodes the id stream
  // kept per calling thread (its storage is reused); the helper fills it
  // through the reference (a thread_local named in the job would be the
  // helper thread's own)
  thread_local HostDecoder hq_keep;
  HostDecoder& hq = hq_keep;
  HelperJob q_ready([&] { prepare_decoder(tr_q.data(), tr_q.size(), nc_q, hq); }, !ctx.batch_worker);
  auto stage_bits = [&](size_t bo, size_t nbytes, const char* name) {
    uint32_t* wb = ctx.buf<uint32_t>(name, div_up(nbytes, 4) + 8);
    CPSZ_HIP(hipMemsetAsync(wb, 0, (div_up(nbytes, 4) + 8) * 4, st));
    if (nbytes) CPSZ_HIP(hipMemcpyAsync(wb, d_in + bo, nbytes, hipMemcpyDeviceToDevice, st));
    return wb;
  };
  {
    uint32_t* wb = stage_bits(bo_id, nb_id, "d3.w.id");
    DeviceDecoder dd = upload_decoder(ctx, tr_id.data(), tr_id.size(), nc_id, "d3.id");
    decode_stream(ctx, dd, wb, nb_id, n, SinkU8{w.vid}, "d3.id");
  }
  // what needs the ids but not the codes is queued while the helper may still
  // be tabulating the code tree: the code bitstream's staging, per-chunk
  // vertex counts and their scans, the unpredictable values and sign bitmaps
  uint32_t* wq = stage_bits(bo_q, nb_q, "d3.w.q");
  const size_t nch = div_up(n, kVChunk);
  k_chunk_counts3d<false><<<(int)div_up(nch, 4), 256, 0, st>>>(w.vid, sk, n, nch, w.cnt, w.cnt + nch);
  scan_u32_to_u64(ctx, w.cnt, w.coff, nch, w.coff + 2 * nch + 2);
  scan_u32_to_u64(ctx, w.cnt + nch, w.coff + nch + 1, nch, w.coff + 2 * nch + 3);
  if (nunp_f) CPSZ_HIP(hipMemcpyAsync(w.unp, d_in + unp_off, 4 * nunp_f, hipMemcpyDeviceToDevice, st));
  CPSZ_HIP(hipMemsetAsync(w.err, 0, 16, st));
  CPSZ_HIP(hipMemcpyAsync(w.sm, d_in + 8, 3 * smap, hipMemcpyDeviceToDevice, st));
  q_ready.join();  // rethrows a malformed code tree
  {
    DeviceDecoder dd = upload_decoder(ctx, hq, "d3.q");
    decode_stream(ctx, dd, wq, nb_q, data_num, SinkU16{w.qc}, "d3.q");
  }
  ctx.mark("huffman");
  k_scatter3d<<<(int)div_up(nch, 4), 256, 0, st>>>(w.vid, sk, w.qc, w.unp, n, nch, w.coff, w.coff + nch + 1,
                                                   data_num, nunp_f, lm.log2f, w.vs, w.q, w.init, w.err);
  CPSZ_HIP(hipGetLastError());
  w.nunp = nunp_f;
  ctx.mark("scatter");
}

void d3_wavefront(Ctx& ctx, D3Work* const* w, size_t nf, int grid_cap) {
  std::vector<Wave3d> f(nf);
  for (size_t k = 0; k < nf; k++) {
    f[k] = Wave3d{};
    f[k].lx = w[k]->init;
    f[k].code = w[k]->vs;
    f[k].qin = w[k]->q;
    f[k].R = w[k]->R;
    f[k].prog = w[k]->prog;
    f[k].s = make_skew(w[k]->r1, w[k]->r2, w[k]->r3);
  }
  launch_wave3d<false>(ctx, f, grid_cap);
}

void d3_finish(Ctx& ctx, const D3Work& w) {
  const size_t n = w.r1 * w.r2 * w.r3, nch = div_up(n, kVChunk), smap = (n - 1) / 8 + 1;
  const Skew3 sk = make_skew(w.r1, w.r2, w.r3);
  k_finish3d<<<(int)div_up(nch, 4), 256, 0, ctx.stream()>>>(w.vid, sk, w.R, w.sm, smap, w.unp, n, nch,
                                                            w.coff + nch + 1, w.nunp, ctx.libm_dev().exp2f,
                                                            w.dU, w.dV, w.dW);
  CPSZ_HIP(hipGetLastError());
}

void d3_check(Ctx& ctx, D3Work* const* w, size_t nf) {
  int* h = (int*)ctx.host_buf("d3.herr", 16 * nf);
  for (size_t k = 0; k < nf; k++)
    CPSZ_HIP(hipMemcpyAsync(h + 4 * k, w[k]->err, 16, hipMemcpyDeviceToHost, ctx.stream()));
  CPSZ_HIP(hipStreamSynchronize(ctx.stream()));
  for (size_t k = 0; k < nf; k++)
    CPSZ_REQUIRE(h[4 * k] == 0, kErrStream, "decompress_3d: code/unpredictable counts inconsistent");
}

void decompress_3d_device(Ctx& ctx, const uint8_t* d_in, const uint8_t* h_in, size_t in_size,
                          size_t r1, size_t r2, size_t r3, float* dU, float* dV, float* dW) {
  CPSZ_REQUIRE(d_in && dU && dV && dW, kErrArg, "decompress_3d: bad arguments");
  check_dims_3d(r1, r2, r3, "decompress_3d");
  ctx.stages.clear();
  ctx.mark("start");
  D3Work w;
  w.r1 = r1;
  w.r2 = r2;
  w.r3 = r3;
  w.dU = dU;
  w.dV = dV;
  w.dW = dW;
  d3_bind(w, ctx.buf<uint8_t>("d3.scratch", d3_scratch_bytes(r1, r2, r3)));
  d3_entropy(ctx, d_in, h_in, in_size, w);
  D3Work* wp = &w;
  d3_wavefront(ctx, &wp, 1, 1024);
  ctx.mark("lorenzo");
  d3_finish(ctx, w);
  ctx.mark("finish");
  d3_check(ctx, &wp, 1);
  ctx.finish_marks();
}

// ------------------------------------------------------ host emulations ----
void eb3d_emulate(const float* U, const float* V, const float* W, size_t r1, size_t r2, size_t r3,
                  double max_pwr_eb, double* eb) {
  const size_t n = r1 * r2 * r3, d0 = r2 * r3, d1 = r3;
  for (size_t v = 0; v < n; v++) eb[v] = max_pwr_eb;
  for (size_t i = 0; i + 1 < r1; i++)
    for (size_t j = 0; j + 1 < r2; j++)
      for (size_t k = 0; k + 1 < r3; k++) {
        const size_t c = i * d0 + j * d1 + k;
        for (int t = 0; t < 6; t++) {
          size_t idx[4];
          for (int v = 0; v < 3; v++) {
            int di, dj, dk;
            tet_offsets(t, v, di, dj, dk);
            idx[v] = c + di * d0 + dj * d1 + dk;
          }
          idx[3] = c;
          const double e = tet_bound(U[idx[0]], U[idx[1]], U[idx[2]], U[idx[3]], V[idx[0]], V[idx[1]],
                                     V[idx[2]], V[idx[3]], W[idx[0]], W[idx[1]], W[idx[2]], W[idx[3]]);
          for (int v = 0; v < 4; v++) eb[idx[v]] = eb[idx[v]] < e ? eb[idx[v]] : e;
        }
      }
}

void lorenzo3d_emulate(const float* L0, const uint8_t* code, size_t r1, size_t r2, size_t r3,
                       uint8_t* vid, uint16_t* q) {
  const size_t n = r1 * r2 * r3, d0 = r2 * r3, d1 = r3;
  std::vector<float> R(L0, L0 + 3 * n);
  for (size_t i = 0; i < r1; i++)
    for (size_t j = 0; j < r2; j++)
      for (size_t k = 0; k < r3; k++) {
        const size_t v = i * d0 + j * d1 + k;
        const uint32_t c = code[v];
        float dec[3];
        uint16_t qc[3];
        bool okall = true;
        for (int p = 0; p < 3; p++) {
          const float* X = R.data() + p * n;
          auto at = [&](bool in, size_t e) { return in ? X[e] : 0.f; };
          const float e0 = at(i && j && k, v - d0 - d1 - 1), e1 = at(i && j, v - d0 - d1),
                      e2 = at(i && k, v - d0 - 1), e3 = at(i, v - d0), e4 = at(j && k, v - d1 - 1),
                      e5 = at(j, v - d1), e6 = at(k, v - 1);
          const float pred = e0 + e3 + e5 + e6 - e1 - e2 - e4;
          bool ok = false;
          if (c != 255 && c != 0) dec[p] = l3d_step(pred, X[v], (int)c, ok, qc[p]);
          okall = okall && ok;
        }
        const bool pr = c != 255 && c != 0 && okall;
        vid[v] = (uint8_t)(c == 255 ? 0 : (pr ? c : 24));
        for (int p = 0; p < 3; p++) {
          if (pr) {
            R[p * n + v] = dec[p];
            q[p * n + v] = qc[p];
          } else {
            q[p * n + v] = 0;
          }
        }
      }
}

}  // namespace cpsz
