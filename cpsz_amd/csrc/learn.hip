// Learned error-bound proposer for the 2D and 3D codecs: features, bisection
// labels, BF16 MFMA MLP (forward, backward, Adam) and the verified learned
// compressors (single 2D field; batches of 3D fields).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "batch3d.hpp"
#include "codec3d.hpp"
#include "cp.hpp"
#include "cp_math.hpp"
#include "learn.hpp"

namespace cpsz {

namespace {

__host__ __device__ inline uint16_t f2bf(float x) {
  uint32_t b;
  __builtin_memcpy(&b, &x, 4);
  b += 0x7FFFu + ((b >> 16) & 1u);  // round to nearest even
  return (uint16_t)(b >> 16);
}
__device__ inline float bf2f(uint16_t h) {
  const uint32_t b = (uint32_t)h << 16;
  float x;
  __builtin_memcpy(&x, &b, 4);
  return x;
}

// ------------------------------------------------------------- features --
// 5x5 patch of u then v (row-major, zero outside the field) divided by the
// patch's largest magnitude, then log2(analytic / max_pwr_eb) / 16 (-4 for a
// zero bound), then zeros up to kFeatDim.
__global__ void k_feat2d(const float* __restrict__ U, const float* __restrict__ V,
                         const double* __restrict__ E, int r1, int r2, double max_pwr_eb,
                         const uint32_t* __restrict__ idx, size_t v0, size_t m, uint16_t* __restrict__ X) {
  const size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (q >= m) return;
  const size_t v = idx ? idx[q] : v0 + q;
  const int i = (int)(v / (size_t)r2), j = (int)(v - (size_t)i * r2);
  float pu[25], pv[25];
  float s = 0.f;
  for (int a = 0; a < 5; a++)
    for (int b = 0; b < 5; b++) {
      const int ii = i + a - 2, jj = j + b - 2;
      const bool in = ii >= 0 && ii < r1 && jj >= 0 && jj < r2;
      const size_t o = in ? (size_t)ii * r2 + jj : 0;
      pu[a * 5 + b] = in ? U[o] : 0.f;
      pv[a * 5 + b] = in ? V[o] : 0.f;
      s = fmaxf(s, fmaxf(fabsf(pu[a * 5 + b]), fabsf(pv[a * 5 + b])));
    }
  if (!(s > 0.f)) s = 1.f;
  uint16_t* x = X + q * kFeatDim;
  for (int k = 0; k < 25; k++) {
    x[k] = f2bf(pu[k] / s);
    x[25 + k] = f2bf(pv[k] / s);
  }
  const double e = E[v];
  const float le = e > 0 ? fmaxf((float)log2(e / max_pwr_eb), -64.f) / 16.f : -4.f;
  x[50] = f2bf(le);
  for (int k = kFeatUsed; k < kFeatDim; k++) x[k] = 0;
}

// --------------------------------------------------------------- labels --
// Incident triangles of vertex (i, j): cell (di, dj) offset, kind t (0 upper
// (i,j),(i+1,j),(i+1,j+1); 1 lower (i,j),(i,j+1),(i+1,j+1)) and the vertex's
// position in the triangle.
__constant__ int kInc[6][4] = {{-1, -1, 0, 2}, {-1, -1, 1, 2}, {-1, 0, 0, 1},
                               {0, -1, 1, 1},  {0, 0, 0, 0},   {0, 0, 1, 0}};

__device__ inline void tri_vertex(int t, int p, int& di, int& dj) {
  // upper: (0,0),(1,0),(1,1); lower: (0,0),(0,1),(1,1)
  di = p == 0 ? 0 : (t == 0 ? 1 : (p == 1 ? 0 : 1));
  dj = p == 0 ? 0 : (t == 0 ? (p == 1 ? 0 : 1) : 1);
}

__global__ void k_labels2d(const float* __restrict__ U, const float* __restrict__ V,
                           const double* __restrict__ E, int r1, int r2, double max_pwr_eb,
                           const uint32_t* __restrict__ idx, size_t m, uint8_t* __restrict__ L) {
  const size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (q >= m) return;
  const size_t v = idx ? idx[q] : q;
  const int i = (int)(v / (size_t)r2), j = (int)(v - (size_t)i * r2);
  const double e = E[v];
  if (!(e > 0)) {
    L[q] = 0;
    return;
  }
  float f[6][3][2];
  int base[6], kind[6], pos[6];
  bool valid[6];
  for (int k = 0; k < 6; k++) {
    const int ci = i + kInc[k][0], cj = j + kInc[k][1];
    kind[k] = kInc[k][2];
    pos[k] = kInc[k][3];
    valid[k] = ci >= 0 && ci < r1 - 1 && cj >= 0 && cj < r2 - 1;
    base[k] = -1;
    for (int p = 0; p < 3; p++) {
      int di, dj;
      tri_vertex(kind[k], p, di, dj);
      const size_t o = valid[k] ? (size_t)(ci + di) * r2 + (cj + dj) : 0;
      f[k][p][0] = valid[k] ? U[o] : 0.f;
      f[k][p][1] = valid[k] ? V[o] : 0.f;
    }
    if (valid[k]) base[k] = tri_cp(f[k], kind[k]);
  }
  const float u = U[v], w = V[v];
  int gmax = 0;
  while (gmax < kMaxGain && ldexp(e, gmax) < max_pwr_eb) gmax++;
  auto ok = [&](int g) {
    const double eb = fmin(max_pwr_eb, ldexp(e, g));
    const double du = fabs((double)u) * eb, dv = fabs((double)w) * eb;
    for (int c = 0; c < 4; c++) {
      const float cu = (float)((double)u + ((c & 1) ? du : -du));
      const float cv = (float)((double)w + ((c & 2) ? dv : -dv));
      for (int k = 0; k < 6; k++) {
        if (!valid[k]) continue;
        float g3[3][2];
        for (int p = 0; p < 3; p++) {
          g3[p][0] = p == pos[k] ? cu : f[k][p][0];
          g3[p][1] = p == pos[k] ? cv : f[k][p][1];
        }
        if (tri_cp(g3, kind[k]) != base[k]) return false;
      }
    }
    return true;
  };
  int lo = 0, hi = gmax;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (ok(mid)) lo = mid;
    else hi = mid - 1;
  }
  L[q] = (uint8_t)lo;
}

// ----------------------------------------------------------------- GEMM --
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

// One wave per 32x32 tile of C, four waves per block (64x64). Operand layout
// of v_mfma_f32_32x32x8_bf16: lane t supplies A(t%32, 4*(t/32)+q) and
// B(4*(t/32)+q, t%32), q = 0..3; it receives D(8*(r/4) + 4*(t/32) + r%4, t%32)
// in accumulator r.
__global__ void __launch_bounds__(256) k_gemm_bf16(int M, int N, int K, const uint16_t* __restrict__ A,
                                                   long sam, long sak, const uint16_t* __restrict__ B,
                                                   long sbk, long sbn, float* __restrict__ C, int ldc,
                                                   int kchunk, int atomic) {
  const int w = threadIdx.x >> 6, t = threadIdx.x & 63;
  const int m0 = blockIdx.y * 64 + (w >> 1) * 32, n0 = blockIdx.x * 64 + (w & 1) * 32;
  if (m0 >= M || n0 >= N) return;
  const int kb = blockIdx.z * kchunk, ke = min(K, kb + kchunk);
  const int row = m0 + (t & 31), col = n0 + (t & 31), kq = (t >> 5) * 4;
  const bool rv = row < M, cvld = col < N;
  floatx16 acc;
  for (int r = 0; r < 16; r++) acc[r] = 0.f;
  for (int k0 = kb; k0 < ke; k0 += 8) {
    bf16x4 a, b;
    for (int q = 0; q < 4; q++) {
      const int k = k0 + kq + q;
      const bool kv = k < ke;
      a[q] = (short)((rv && kv) ? A[(long)row * sam + (long)k * sak] : 0);
      b[q] = (short)((cvld && kv) ? B[(long)k * sbk + (long)col * sbn] : 0);
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
  }
  for (int r = 0; r < 16; r++) {
    const int i = m0 + 8 * (r >> 2) + 4 * (t >> 5) + (r & 3);
    if (i < M && cvld) {
      float* c = C + (long)i * ldc + col;
      if (atomic) atomicAdd(c, acc[r]);
      else *c = acc[r];
    }
  }
}

// ------------------------------------------------------ fused inference --
// Feature build + both hidden layers + output in one kernel, for compression
// (training keeps the GEMM path, whose activations the backward pass needs).
// A wave takes 32 vertices at a time: its lanes write the BF16 features of the
// 5x5 (2D) / 3x3x3 (3D) patch into the wave's LDS tile (lane t: vertex t%32,
// features of half t/32), then layer 1 and layer 2 run on
// v_mfma_f32_32x32x16_bf16 with the transposed BF16 weights resident in LDS,
// bias + ReLU between them in FP32, and the output layer reduces the 32 lanes
// that share a row. The hidden rows overwrite the feature tile in place (a
// wave's LDS reads complete before its later writes). Operand layout: lane t
// supplies A(t%32, 8*(t/32)+q) and B(8*(t/32)+q, t%32), q = 0..7, and
// receives D(8*(r/4) + 4*(t/32) + r%4, t%32) in accumulator r.
typedef __bf16 bf16x8m __attribute__((ext_vector_type(8)));

template <int DIM>
struct InferCfg;
template <>
struct InferCfg<2> {
  static constexpr int F = kFeatDim, P = kFeatUsed - 1, WPB = 8;  // 50 patch values, then the bound
};
template <>
struct InferCfg<3> {
  static constexpr int F = kFeatDim3, P = kFeatUsed3 - 1, WPB = 8;  // 81 patch values, then the bound
};

__device__ __forceinline__ bf16x8m lds_bf16x8(const uint16_t* p) {
  bf16x8m v;
  __builtin_memcpy(&v, p, 16);
  return v;
}

// W1t[n][k] = bf16(W1[k][n]), W2t[n][k] = bf16(W2[k][n])
__global__ void k_weights_t(const float* __restrict__ P, int fdim, size_t w1, size_t w2, uint16_t* __restrict__ W1t,
                            uint16_t* __restrict__ W2t) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t n1 = (size_t)kHidden * fdim;
  if (i < n1) {
    const size_t n = i / fdim, k = i - n * fdim;
    W1t[i] = f2bf(P[w1 + k * kHidden + n]);
  } else if (i < n1 + (size_t)kHidden * kHidden) {
    const size_t j = i - n1, n = j / kHidden, k = j - n * kHidden;
    W2t[j] = f2bf(P[w2 + k * kHidden + n]);
  }
}

template <int DIM>
__global__ void __launch_bounds__(64 * InferCfg<DIM>::WPB) k_mlp_infer(
    const float* __restrict__ U, const float* __restrict__ V, const float* __restrict__ W,
    const double* __restrict__ E, int r1, int r2, int r3, double max_pwr_eb, const uint16_t* __restrict__ W1t,
    const uint16_t* __restrict__ W2t, const float* __restrict__ P, NetLayout L, size_t v0, size_t m,
    uint8_t* __restrict__ gain) {
  constexpr int F = InferCfg<DIM>::F, NP = InferCfg<DIM>::P, WPB = InferCfg<DIM>::WPB, HALF = F / 2;
  __shared__ __align__(16) uint16_t sW1[kHidden * F];
  __shared__ __align__(16) uint16_t sW2[kHidden * kHidden];
  __shared__ float sb1[kHidden], sb2[kHidden], sw3[kHidden];
  __shared__ __align__(16) uint16_t sbuf[WPB][32 * kHidden];  // features, then the hidden rows
  for (int i = threadIdx.x; i < kHidden * F; i += blockDim.x) sW1[i] = W1t[i];
  for (int i = threadIdx.x; i < kHidden * kHidden; i += blockDim.x) sW2[i] = W2t[i];
  for (int i = threadIdx.x; i < kHidden; i += blockDim.x) {
    sb1[i] = P[L.b1 + i];
    sb2[i] = P[L.b2 + i];
    sw3[i] = P[L.w3 + i];
  }
  __syncthreads();
  const float b3 = P[L.b3];
  const int wave = threadIdx.x >> 6, t = threadIdx.x & 63, col = t & 31, hf = t >> 5;
  uint16_t* buf = sbuf[wave];
  const float* src[3] = {U, V, W};
  for (size_t tile = (size_t)blockIdx.x * WPB + wave; tile * 32 < m; tile += (size_t)gridDim.x * WPB) {
    // ---- features (as k_feat2d / k_feat3d write them)
    const size_t q = tile * 32 + col;
    const bool valid = q < m;
    const size_t v = v0 + (valid ? q : 0);
    int i, j, k = 0;
    if (DIM == 2) {
      i = (int)(v / (size_t)r2);
      j = (int)(v - (size_t)i * r2);
    } else {
      const size_t d0 = (size_t)r2 * r3;
      i = (int)(v / d0);
      const size_t rem = v - (size_t)i * d0;
      j = (int)(rem / (size_t)r3);
      k = (int)(rem - (size_t)j * r3);
    }
    // patch value of feature f (0 outside the field); the largest magnitude
    // first, then the values again (cached) for the normalized BF16 features
    auto patch = [&](int f) -> float {
      int c;
      bool in;
      size_t o;
      if (DIM == 2) {
        c = f / 25;
        const int a = (f % 25) / 5, b = f % 5, ii = i + a - 2, jj = j + b - 2;
        in = ii >= 0 && ii < r1 && jj >= 0 && jj < r2;
        o = in ? (size_t)ii * r2 + jj : 0;
      } else {
        c = f / 27;
        const int oo = f % 27, ii = i + oo / 9 - 1, jj = j + (oo / 3) % 3 - 1, kk = k + oo % 3 - 1;
        in = ii >= 0 && ii < r1 && jj >= 0 && jj < r2 && kk >= 0 && kk < r3;
        o = in ? ((size_t)ii * r2 + jj) * r3 + kk : 0;
      }
      return in ? src[c][o] : 0.f;
    };
    float smax = 0.f;
    for (int ff = 0; ff < HALF; ff++) {
      const int f = hf * HALF + ff;
      if (f < NP) smax = fmaxf(smax, fabsf(patch(f)));
    }
    smax = fmaxf(smax, __shfl_xor(smax, 32));
    if (!(smax > 0.f)) smax = 1.f;
    const double e = E[v];
    const float le = e > 0 ? fmaxf((float)log2(e / max_pwr_eb), -64.f) / 16.f : -4.f;
    for (int ff = 0; ff < HALF; ff++) {
      const int f = hf * HALF + ff;
      const uint16_t h = f < NP ? f2bf(patch(f) / smax) : (f == NP ? f2bf(le) : (uint16_t)0);
      buf[col * F + f] = valid ? h : (uint16_t)0;
    }
    // ---- layer 1: Z1 = X W1 (K = F)
    floatx16 acc[4];
#pragma unroll
    for (int nt = 0; nt < 4; nt++)
      for (int r = 0; r < 16; r++) acc[nt][r] = 0.f;
#pragma unroll
    for (int ks = 0; ks < F / 16; ks++) {
      const bf16x8m a = lds_bf16x8(buf + col * F + ks * 16 + hf * 8);
#pragma unroll
      for (int nt = 0; nt < 4; nt++) {
        const bf16x8m b = lds_bf16x8(sW1 + (nt * 32 + col) * F + ks * 16 + hf * 8);
        acc[nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[nt], 0, 0, 0);
      }
    }
    // ---- H1 = relu(Z1 + b1) as BF16, over the feature tile
#pragma unroll
    for (int nt = 0; nt < 4; nt++)
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int row = 8 * (r >> 2) + 4 * hf + (r & 3), c = nt * 32 + col;
        const float z = acc[nt][r] + sb1[c];
        buf[row * kHidden + c] = f2bf(z > 0.f ? z : 0.f);
      }
    // ---- layer 2: Z2 = H1 W2 (K = 128)
#pragma unroll
    for (int nt = 0; nt < 4; nt++)
      for (int r = 0; r < 16; r++) acc[nt][r] = 0.f;
#pragma unroll
    for (int ks = 0; ks < kHidden / 16; ks++) {
      const bf16x8m a = lds_bf16x8(buf + col * kHidden + ks * 16 + hf * 8);
#pragma unroll
      for (int nt = 0; nt < 4; nt++) {
        const bf16x8m b = lds_bf16x8(sW2 + (nt * 32 + col) * kHidden + ks * 16 + hf * 8);
        acc[nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[nt], 0, 0, 0);
      }
    }
    // ---- y = b3 + sum_c relu(Z2 + b2)[c] w3[c]; the 32 lanes of a half share
    // rows, so each row's partial sums are reduced across them; lane col = r
    // writes row 8*(r/4) + 4*hf + r%4 of its half
#pragma unroll
    for (int r = 0; r < 16; r++) {
      float sacc = 0.f;
#pragma unroll
      for (int nt = 0; nt < 4; nt++) {
        const int c = nt * 32 + col;
        const float z = acc[nt][r] + sb2[c];
        sacc = fmaf(z > 0.f ? z : 0.f, sw3[c], sacc);
      }
      for (int o = 1; o < 32; o <<= 1) sacc += __shfl_xor(sacc, o);
      const size_t row = tile * 32 + 8 * (r >> 2) + 4 * hf + (r & 3);
      if (col == r && row < m) {  // the proposal gain = clamp(floor(y), 0, kMaxGain), as k_gain
        const float f = floorf(sacc + b3);
        gain[row] = (uint8_t)(f > 0.f ? (f < (float)kMaxGain ? f : (float)kMaxGain) : 0.f);
      }
    }
  }
}

// ----------------------------------------------------------- elementwise --
__global__ void k_to_bf16(const float* __restrict__ x, size_t n, uint16_t* __restrict__ y) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

// Z[r][c] += b[c]; H = relu(Z) as BF16 (and FP32 when Hf is set)
__global__ void k_bias_relu(float* __restrict__ Z, const float* __restrict__ b, size_t m,
                            uint16_t* __restrict__ Hb, float* __restrict__ Hf) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= m * kHidden) return;
  const float z = Z[i] + b[i % kHidden];
  Z[i] = z;
  const float h = z > 0.f ? z : 0.f;
  Hb[i] = f2bf(h);
  if (Hf) Hf[i] = h;
}

// y[r] = b3 + sum_c H2[r][c] w3[c] (sequential order)
__global__ void k_out(const float* __restrict__ H2, const float* __restrict__ w3, const float* __restrict__ b3,
                      size_t m, float* __restrict__ y) {
  const size_t r = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (r >= m) return;
  float s = 0.f;
  const float* h = H2 + r * kHidden;
  for (int c = 0; c < kHidden; c++) s = fmaf(h[c], w3[c], s);
  y[r] = s + b3[0];
}

__global__ void k_loss(const float* __restrict__ y, const float* __restrict__ t, size_t m,
                       float* __restrict__ dy, float* __restrict__ loss) {
  const size_t r = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  float l = 0.f;
  if (r < m) {
    const float d = y[r] - t[r];
    dy[r] = 2.f * d / (float)m;
    l = d * d / (float)m;
  }
  for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
  if ((threadIdx.x & 63) == 0 && l != 0.f) atomicAdd(loss, l);
}

constexpr int kRowsPerBlock = 64;

// Output layer backward: gW3 += H2^T dy, gb3 += sum dy, dZ2 = dy w3 * (Z2 > 0)
// (BF16), gb2 += column sums of dZ2. Block: 128 threads (one per column) x 64 rows.
__global__ void k_back_out(const float* __restrict__ H2, const float* __restrict__ Z2,
                           const float* __restrict__ dy, const float* __restrict__ w3, size_t m,
                           float* __restrict__ gW3, float* __restrict__ gb3, float* __restrict__ gb2,
                           uint16_t* __restrict__ dZ2) {
  const int c = threadIdx.x;
  const size_t r0 = blockIdx.x * (size_t)kRowsPerBlock;
  float sw = 0.f, sb = 0.f, sy = 0.f;
  for (int k = 0; k < kRowsPerBlock; k++) {
    const size_t r = r0 + k;
    if (r >= m) break;
    const size_t o = r * kHidden + c;
    const float g = dy[r];
    sw += H2[o] * g;
    const float dz = Z2[o] > 0.f ? g * w3[c] : 0.f;
    dZ2[o] = f2bf(dz);
    sb += dz;
    sy += g;
  }
  atomicAdd(&gW3[c], sw);
  atomicAdd(&gb2[c], sb);
  if (c == 0) atomicAdd(gb3, sy);
}

// dZ1 = dH1 * (Z1 > 0) (BF16), gb1 += column sums
__global__ void k_back_relu(const float* __restrict__ dH, const float* __restrict__ Z, size_t m,
                            float* __restrict__ gb, uint16_t* __restrict__ dZ) {
  const int c = threadIdx.x;
  const size_t r0 = blockIdx.x * (size_t)kRowsPerBlock;
  float sb = 0.f;
  for (int k = 0; k < kRowsPerBlock; k++) {
    const size_t r = r0 + k;
    if (r >= m) break;
    const size_t o = r * kHidden + c;
    const float dz = Z[o] > 0.f ? dH[o] : 0.f;
    dZ[o] = f2bf(dz);
    sb += dz;
  }
  atomicAdd(&gb[c], sb);
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, size_t n, float lr, float b1, float b2, float eps, float c1,
                       float c2) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  p[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

__global__ void k_gather_rows(const uint16_t* __restrict__ X, const float* __restrict__ T,
                              const uint32_t* __restrict__ perm, size_t m, int fdim, uint16_t* __restrict__ Xb,
                              float* __restrict__ Tb) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= m * fdim) return;
  const size_t r = i / fdim, c = i - r * fdim;
  const size_t s = perm[r];
  Xb[i] = X[s * fdim + c];
  if (c == 0) Tb[r] = T[s];
}

__global__ void k_u8_to_f32(const uint8_t* __restrict__ a, size_t n, float* __restrict__ b) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

// proposal: gain = clamp(floor(y), 0, kMaxGain)
__global__ void k_gain(const float* __restrict__ y, size_t n, uint8_t* __restrict__ g) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float f = floorf(y[i]);
  g[i] = (uint8_t)(f > 0.f ? (f < (float)kMaxGain ? f : (float)kMaxGain) : 0.f);
}

int blocks_for(size_t n, int bs = 256) { return (int)div_up(n, (size_t)bs); }

uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Network activations for up to `cap` rows.
struct Acts {
  float *Z1, *Z2, *H2, *y;
  uint16_t *H1b, *W1b, *W2b;
};

Acts acts(Ctx& ctx, size_t cap, int fdim) {
  Acts a;
  a.Z1 = ctx.buf<float>("nn.z1", cap * kHidden);
  a.Z2 = ctx.buf<float>("nn.z2", cap * kHidden);
  a.H2 = ctx.buf<float>("nn.h2", cap * kHidden);
  a.y = ctx.buf<float>("nn.y", cap);
  a.H1b = ctx.buf<uint16_t>("nn.h1b", cap * kHidden);
  a.W1b = ctx.buf<uint16_t>("nn.w1b", (size_t)fdim * kHidden);
  a.W2b = ctx.buf<uint16_t>("nn.w2b", (size_t)kHidden * kHidden);
  return a;
}

void weights_bf16(Ctx& ctx, const float* P, Acts& a, const NetLayout& L) {
  hipStream_t st = ctx.stream();
  const size_t n1 = (size_t)L.fdim * kHidden, n2 = (size_t)kHidden * kHidden;
  k_to_bf16<<<blocks_for(n1), 256, 0, st>>>(P + L.w1, n1, a.W1b);
  k_to_bf16<<<blocks_for(n2), 256, 0, st>>>(P + L.w2, n2, a.W2b);
}

// forward of rows [0, m) of X (m <= cap)
void forward_rows(Ctx& ctx, const float* P, Acts& a, const uint16_t* X, size_t m, const NetLayout& L) {
  hipStream_t st = ctx.stream();
  gemm_bf16(ctx, (int)m, kHidden, L.fdim, X, L.fdim, 1, a.W1b, kHidden, 1, a.Z1, kHidden, 1);
  k_bias_relu<<<blocks_for(m * kHidden), 256, 0, st>>>(a.Z1, P + L.b1, m, a.H1b, nullptr);
  gemm_bf16(ctx, (int)m, kHidden, kHidden, a.H1b, kHidden, 1, a.W2b, kHidden, 1, a.Z2, kHidden, 1);
  uint16_t* h2b = ctx.buf<uint16_t>("nn.h2b", m * kHidden);
  k_bias_relu<<<blocks_for(m * kHidden), 256, 0, st>>>(a.Z2, P + L.b2, m, h2b, a.H2);
  k_out<<<blocks_for(m), 256, 0, st>>>(a.H2, P + L.w3, P + L.b3, m, a.y);
  CPSZ_HIP(hipGetLastError());
}

void check_fdim(int fdim) {
  CPSZ_REQUIRE(fdim == kFeatDim || fdim == kFeatDim3, kErrArg, "MLP: unsupported feature width");
}

}  // namespace

void gemm_bf16(Ctx& ctx, int M, int N, int K, const uint16_t* A, long sam, long sak, const uint16_t* B,
               long sbk, long sbn, float* C, int ldc, int split) {
  CPSZ_REQUIRE(M >= 0 && N >= 0 && K >= 0 && split >= 1, kErrArg, "gemm_bf16: bad sizes");
  if (M == 0 || N == 0) return;
  int kchunk = (int)div_up((size_t)K, (size_t)split);
  kchunk = (int)div_up((size_t)kchunk, 8) * 8;
  if (kchunk == 0) kchunk = 8;
  const int nz = (int)div_up((size_t)(K ? K : 1), (size_t)kchunk);
  dim3 grid((unsigned)div_up((size_t)N, 64), (unsigned)div_up((size_t)M, 64), (unsigned)nz);
  k_gemm_bf16<<<grid, 256, 0, ctx.stream()>>>(M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, kchunk, nz > 1);
  CPSZ_HIP(hipGetLastError());
}

void mlp_init_params(uint64_t seed, float* p, int fdim, int fused) {
  check_fdim(fdim);
  CPSZ_REQUIRE(fused > 0 && fused <= fdim, kErrArg, "mlp_init_params: bad feature count");
  const NetLayout L = net_layout(fdim);
  uint64_t s = seed ^ 0x6c6561726e656462ull;
  auto uni = [&]() { return (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); };
  auto fill = [&](size_t off, size_t n, double scale) {
    for (size_t i = 0; i < n; i++) p[off + i] = (float)((2 * uni() - 1) * scale);
  };
  std::memset(p, 0, L.count * sizeof(float));
  fill(L.w1, (size_t)fdim * kHidden, std::sqrt(6.0 / fused));
  fill(L.w2, (size_t)kHidden * kHidden, std::sqrt(6.0 / kHidden));
  fill(L.w3, kHidden, std::sqrt(3.0 / kHidden));
  for (size_t r = fused; r < (size_t)fdim; r++)
    for (int c = 0; c < kHidden; c++) p[L.w1 + r * kHidden + c] = 0.f;  // padding rows
}

void analytic_eb_2d(Ctx& ctx, const float* dU, const float* dV, size_t r1, size_t r2, double max_pwr_eb,
                    double* dE) {
  eb_2d_device(ctx, dU, dV, r1, r2, max_pwr_eb, dE);
}

void features_2d(Ctx& ctx, const float* dU, const float* dV, const double* dE, size_t r1, size_t r2,
                 double max_pwr_eb, const uint32_t* d_idx, size_t m, uint16_t* dX) {
  CPSZ_REQUIRE(dU && dV && dE && dX && r1 * r2 < (1ull << 32), kErrArg, "features_2d: bad arguments");
  if (!m) return;
  k_feat2d<<<blocks_for(m), 256, 0, ctx.stream()>>>(dU, dV, dE, (int)r1, (int)r2, max_pwr_eb, d_idx, 0, m, dX);
  CPSZ_HIP(hipGetLastError());
}

void mlp_infer_gains(Ctx& ctx, int dim, const float* dU, const float* dV, const float* dW, const double* dE,
                     size_t r1, size_t r2, size_t r3, double max_pwr_eb, const float* d_params, size_t v0, size_t m,
                     uint8_t* d_gain) {
  CPSZ_REQUIRE(dim == 2 || dim == 3, kErrArg, "mlp_infer_gains: dim must be 2 or 3");
  CPSZ_REQUIRE(dU && dV && (dim == 2 || dW) && dE && d_params && d_gain, kErrArg, "mlp_infer_gains: bad arguments");
  if (!m) return;
  hipStream_t st = ctx.stream();
  const int fdim = dim == 2 ? kFeatDim : kFeatDim3;
  const NetLayout L = net_layout(fdim);
  uint16_t* W1t = ctx.buf<uint16_t>("inf.w1t", (size_t)kHidden * fdim);
  uint16_t* W2t = ctx.buf<uint16_t>("inf.w2t", (size_t)kHidden * kHidden);
  const size_t nw = (size_t)kHidden * fdim + (size_t)kHidden * kHidden;
  k_weights_t<<<blocks_for(nw), 256, 0, st>>>(d_params, fdim, L.w1, L.w2, W1t, W2t);
  const size_t tiles = div_up(m, (size_t)32);
  if (dim == 2) {
    const int grid = (int)std::min<size_t>(div_up(tiles, (size_t)InferCfg<2>::WPB), 2048);
    k_mlp_infer<2><<<grid, 64 * InferCfg<2>::WPB, 0, st>>>(dU, dV, nullptr, dE, (int)r1, (int)r2, 1, max_pwr_eb,
                                                          W1t, W2t, d_params, L, v0, m, d_gain);
  } else {
    const int grid = (int)std::min<size_t>(div_up(tiles, (size_t)InferCfg<3>::WPB), 2048);
    k_mlp_infer<3><<<grid, 64 * InferCfg<3>::WPB, 0, st>>>(dU, dV, dW, dE, (int)r1, (int)r2, (int)r3, max_pwr_eb,
                                                          W1t, W2t, d_params, L, v0, m, d_gain);
  }
  CPSZ_HIP(hipGetLastError());
}

void labels_2d(Ctx& ctx, const float* dU, const float* dV, const double* dE, size_t r1, size_t r2,
               double max_pwr_eb, const uint32_t* d_idx, size_t m, uint8_t* dL) {
  CPSZ_REQUIRE(dU && dV && dE && dL && r1 * r2 < (1ull << 32), kErrArg, "labels_2d: bad arguments");
  if (!m) return;
  k_labels2d<<<blocks_for(m, 128), 128, 0, ctx.stream()>>>(dU, dV, dE, (int)r1, (int)r2, max_pwr_eb, d_idx,
                                                           m, dL);
  CPSZ_HIP(hipGetLastError());
}

void mlp_forward(Ctx& ctx, const float* P, const uint16_t* dX, size_t m, float* dY, int fdim) {
  check_fdim(fdim);
  if (!m) return;
  const NetLayout L = net_layout(fdim);
  const size_t cap = std::min<size_t>(m, 1u << 18);
  Acts a = acts(ctx, cap, fdim);
  weights_bf16(ctx, P, a, L);
  for (size_t r0 = 0; r0 < m; r0 += cap) {
    const size_t mm = std::min(cap, m - r0);
    forward_rows(ctx, P, a, dX + r0 * fdim, mm, L);
    CPSZ_HIP(hipMemcpyAsync(dY + r0, a.y, mm * sizeof(float), hipMemcpyDeviceToDevice, ctx.stream()));
  }
}

void mlp_grad(Ctx& ctx, const float* P, const uint16_t* X, const float* T, size_t m, float* G, float* loss,
              int fdim) {
  CPSZ_REQUIRE(m > 0 && m <= (1u << 20), kErrArg, "mlp_grad: batch size out of range");
  check_fdim(fdim);
  const NetLayout L = net_layout(fdim);
  hipStream_t st = ctx.stream();
  Acts a = acts(ctx, m, fdim);
  weights_bf16(ctx, P, a, L);
  forward_rows(ctx, P, a, X, m, L);
  CPSZ_HIP(hipMemsetAsync(G, 0, L.count * sizeof(float), st));
  CPSZ_HIP(hipMemsetAsync(loss, 0, sizeof(float), st));
  float* dy = ctx.buf<float>("nn.dy", m);
  k_loss<<<blocks_for(m), 256, 0, st>>>(a.y, T, m, dy, loss);
  uint16_t* dZ2 = ctx.buf<uint16_t>("nn.dz2", m * kHidden);
  k_back_out<<<blocks_for(m, kRowsPerBlock), kHidden, 0, st>>>(a.H2, a.Z2, dy, P + L.w3, m, G + L.w3,
                                                               G + L.b3, G + L.b2, dZ2);
  // gW2 = H1^T dZ2 (split over the batch)
  const int split = (int)std::max<size_t>(1, std::min<size_t>(64, m / 512));
  gemm_bf16(ctx, kHidden, kHidden, (int)m, a.H1b, 1, kHidden, dZ2, kHidden, 1, G + L.w2, kHidden, split);
  // dH1 = dZ2 W2^T
  float* dH1 = ctx.buf<float>("nn.dh1", m * kHidden);
  gemm_bf16(ctx, (int)m, kHidden, kHidden, dZ2, kHidden, 1, a.W2b, 1, kHidden, dH1, kHidden, 1);
  uint16_t* dZ1 = ctx.buf<uint16_t>("nn.dz1", m * kHidden);
  k_back_relu<<<blocks_for(m, kRowsPerBlock), kHidden, 0, st>>>(dH1, a.Z1, m, G + L.b1, dZ1);
  // gW1 = X^T dZ1
  gemm_bf16(ctx, fdim, kHidden, (int)m, X, 1, fdim, dZ1, kHidden, 1, G + L.w1, kHidden, split);
  CPSZ_HIP(hipGetLastError());
}

void train_samples(Ctx& ctx, const uint16_t* Xall, const float* Tall, size_t S, int fdim, int epochs, int batch,
                   float lr, uint64_t& rs, float* params, float* loss_per_epoch) {
  check_fdim(fdim);
  CPSZ_REQUIRE(epochs > 0 && batch >= 32 && batch <= (1 << 20), kErrArg, "train: bad arguments");
  const size_t steps = S / batch;
  CPSZ_REQUIRE(steps > 0, kErrArg, "train: fewer samples than one batch");
  const NetLayout L = net_layout(fdim);
  const size_t np = L.count;
  hipStream_t st = ctx.stream();
  float* P = ctx.buf<float>("tr.p", np);
  float* G = ctx.buf<float>("tr.g", np);
  float* M1 = ctx.buf<float>("tr.m1", np);
  float* M2 = ctx.buf<float>("tr.m2", np);
  float* loss = ctx.buf<float>("tr.loss", 1);
  CPSZ_HIP(hipMemcpyAsync(P, params, np * 4, hipMemcpyHostToDevice, st));
  CPSZ_HIP(hipMemsetAsync(M1, 0, np * 4, st));
  CPSZ_HIP(hipMemsetAsync(M2, 0, np * 4, st));
  uint16_t* Xb = ctx.buf<uint16_t>("tr.xb", (size_t)batch * fdim);
  float* Tb = ctx.buf<float>("tr.tb", batch);
  uint32_t* perm = ctx.buf<uint32_t>("tr.perm", S);
  std::vector<uint32_t> hp(S);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  int t = 0;
  for (int ep = 0; ep < epochs; ep++) {
    for (size_t k = 0; k < S; k++) hp[k] = (uint32_t)k;
    for (size_t k = S - 1; k > 0; k--) std::swap(hp[k], hp[splitmix(rs) % (k + 1)]);
    CPSZ_HIP(hipMemcpyAsync(perm, hp.data(), S * 4, hipMemcpyHostToDevice, st));
    for (size_t b = 0; b < steps; b++) {
      k_gather_rows<<<blocks_for((size_t)batch * fdim), 256, 0, st>>>(Xall, Tall, perm + b * batch, batch, fdim,
                                                                       Xb, Tb);
      mlp_grad(ctx, P, Xb, Tb, batch, G, loss, fdim);
      ++t;
      const float c1 = 1.f - std::pow(b1, (float)t), c2 = 1.f - std::pow(b2, (float)t);
      k_adam<<<blocks_for(np), 256, 0, st>>>(P, G, M1, M2, np, lr, b1, b2, eps, c1, c2);
      CPSZ_HIP(hipGetLastError());
      float hl = 0;
      if (b + 1 == steps) {  // the epoch's last batch loss
        CPSZ_HIP(hipMemcpyAsync(&hl, loss, 4, hipMemcpyDeviceToHost, st));
        CPSZ_HIP(hipStreamSynchronize(st));
        if (loss_per_epoch) loss_per_epoch[ep] = hl;
      }
    }
    CPSZ_HIP(hipStreamSynchronize(st));  // the permutation buffer is rewritten next epoch
  }
  CPSZ_HIP(hipMemcpyAsync(params, P, np * 4, hipMemcpyDeviceToHost, st));
  CPSZ_HIP(hipStreamSynchronize(st));
}

void train_2d(Ctx& ctx, const std::vector<TrainField>& fields, double max_pwr_eb, size_t spf, int epochs,
              int batch, float lr, uint64_t seed, float* params, float* loss_per_epoch) {
  CPSZ_REQUIRE(!fields.empty() && spf > 0 && epochs > 0 && batch >= 32 && batch <= (1 << 20), kErrArg,
               "train_2d: bad arguments");
  hipStream_t st = ctx.stream();
  const size_t S = fields.size() * spf;
  uint16_t* Xall = ctx.buf<uint16_t>("tr.x", S * kFeatDim);
  float* Tall = ctx.buf<float>("tr.t", S);
  uint8_t* L8 = ctx.buf<uint8_t>("tr.l8", spf);
  uint32_t* idx = ctx.buf<uint32_t>("tr.idx", spf);
  uint32_t* hidx = (uint32_t*)ctx.host_buf("tr.idx", spf * 4);
  uint64_t rs = seed;
  for (size_t f = 0; f < fields.size(); f++) {
    const TrainField& F = fields[f];
    const size_t n = F.r1 * F.r2;
    CPSZ_REQUIRE(F.dU && F.dV && n > 0 && n < (1ull << 32), kErrArg, "train_2d: bad field");
    double* E = ctx.buf<double>("tr.e", n);
    analytic_eb_2d(ctx, F.dU, F.dV, F.r1, F.r2, max_pwr_eb, E);
    CPSZ_HIP(hipStreamSynchronize(st));  // host index buffer reuse
    for (size_t k = 0; k < spf; k++) hidx[k] = (uint32_t)(splitmix(rs) % n);
    CPSZ_HIP(hipMemcpyAsync(idx, hidx, spf * 4, hipMemcpyHostToDevice, st));
    features_2d(ctx, F.dU, F.dV, E, F.r1, F.r2, max_pwr_eb, idx, spf, Xall + f * spf * kFeatDim);
    labels_2d(ctx, F.dU, F.dV, E, F.r1, F.r2, max_pwr_eb, idx, spf, L8);
    k_u8_to_f32<<<blocks_for(spf), 256, 0, st>>>(L8, spf, Tall + f * spf);
    CPSZ_HIP(hipGetLastError());
  }
  train_samples(ctx, Xall, Tall, S, kFeatDim, epochs, batch, lr, rs, params, loss_per_epoch);
}

namespace {
void cp_list(Ctx& ctx, const float* dU, const float* dV, size_t r1, size_t r2, std::vector<int64_t>& ids,
             std::vector<int32_t>& types, bool sos = false) {
  size_t cap = 1 << 16, count = 0;
  for (;;) {
    ids.resize(cap);
    types.resize(cap);
    critical_points_2d(ctx, dU, dV, r1, r2, ids.data(), types.data(), cap, &count, sos);
    if (count <= cap) break;
    cap = count;
  }
  ids.resize(count);
  types.resize(count);
}

// simplices whose critical point appeared, vanished or changed type
std::vector<int64_t> cp_changes(const std::vector<int64_t>& id0, const std::vector<int32_t>& ty0,
                                const std::vector<int64_t>& id1, const std::vector<int32_t>& ty1) {
  std::vector<int64_t> bad;
  size_t a = 0, b = 0;
  while (a < id0.size() || b < id1.size()) {
    if (b == id1.size() || (a < id0.size() && id0[a] < id1[b])) bad.push_back(id0[a++]);
    else if (a == id0.size() || id1[b] < id0[a]) bad.push_back(id1[b++]);
    else {
      if (ty0[a] != ty1[b]) bad.push_back(id0[a]);
      a++;
      b++;
    }
  }
  return bad;
}

// Per-vertex fallback of the learned bound: 0 learned, 1 the analytic bound
// (which keeps every critical point of the reference's closed-simplex test),
// 2 stored exactly (eb 0). A triangle the closed test sees change sends its
// vertices to the analytic bound; one only the Simulation-of-Simplicity test
// sees change (a degenerate zero the analytic bound need not keep) sends
// them one step further.
void mark_triangles(const std::vector<int64_t>& bad, size_t r2, uint8_t* hf, bool sos) {
  for (int64_t sid : bad) {
    const int64_t cell = sid / 2;
    const int t = (int)(sid % 2);
    const size_t ci = (size_t)(cell / (int64_t)(r2 - 1)), cj = (size_t)(cell % (int64_t)(r2 - 1));
    const size_t v[3] = {ci * r2 + cj, t == 0 ? (ci + 1) * r2 + cj : ci * r2 + cj + 1, (ci + 1) * r2 + cj + 1};
    for (size_t x : v) hf[x] = sos ? (hf[x] ? 2 : 1) : (hf[x] ? hf[x] : 1);
  }
}

// The learned verification rounds: kLearnRounds with proposals, then the
// analytic bound for every vertex not yet fallen back, then up to two more
// rounds escalating to exact storage, then everything exact.
constexpr int kLearnRounds = 16;
constexpr int kLearnLast = kLearnRounds + 3;

// Falls back as the round requires; returns false when the field is done.
bool verify_round_2d(int it, const std::vector<int64_t>& id0, const std::vector<int32_t>& ty0,
                     const std::vector<int64_t>& is0, const std::vector<int32_t>& ts0,
                     const std::vector<int64_t>& id1, const std::vector<int32_t>& ty1,
                     const std::vector<int64_t>& is1, const std::vector<int32_t>& ts1, size_t r2, uint8_t* hf,
                     size_t n) {
  const std::vector<int64_t> bad = cp_changes(id0, ty0, id1, ty1);
  const std::vector<int64_t> bad_s = cp_changes(is0, ts0, is1, ts1);
  if (bad.empty() && bad_s.empty()) return false;
  CPSZ_REQUIRE(bad.empty() || it < kLearnRounds, kErrInternal,
               "learned compression: the plain bound changed a critical point");
  CPSZ_REQUIRE(it < kLearnLast, kErrInternal, "learned compression did not converge");
  mark_triangles(bad, r2, hf, false);
  mark_triangles(bad_s, r2, hf, true);
  if (it + 1 == kLearnRounds)
    for (size_t v = 0; v < n; v++) hf[v] = hf[v] ? hf[v] : 1;
  if (it + 1 == kLearnLast) std::fill(hf, hf + n, 2);
  return true;
}
}  // namespace

void compress_2d_learned(Ctx& ctx, const float* dU, const float* dV, size_t r1, size_t r2,
                         double max_pwr_eb, const float* params, DeviceStream& out, LearnedStats& stats) {
  CPSZ_REQUIRE(dU && dV && params && r1 * r2 < (1ull << 32), kErrArg, "compress_2d_learned: bad arguments");
  hipStream_t st = ctx.stream();
  const size_t n = r1 * r2;
  stats = LearnedStats();
  double* E = ctx.buf<double>("ln.e", n);
  analytic_eb_2d(ctx, dU, dV, r1, r2, max_pwr_eb, E);
  float* P = ctx.buf<float>("ln.p", kNumParams);
  CPSZ_HIP(hipMemcpyAsync(P, params, kNumParams * 4, hipMemcpyHostToDevice, st));
  uint8_t* gain = ctx.buf<uint8_t>("ln.gain", n);
  mlp_infer_gains(ctx, 2, dU, dV, nullptr, E, r1, r2, 1, max_pwr_eb, P, 0, n, gain);
  std::vector<uint8_t> hg(n), hf(n, 0);
  CPSZ_HIP(hipMemcpyAsync(hg.data(), gain, n, hipMemcpyDeviceToHost, st));
  std::vector<int64_t> id0, id1, is0, is1;
  std::vector<int32_t> ty0, ty1, ts0, ts1;
  cp_list(ctx, dU, dV, r1, r2, id0, ty0);  // synchronizes
  cp_list(ctx, dU, dV, r1, r2, is0, ts0, true);
  for (size_t v = 0; v < n; v++) stats.boosted_vertices += hg[v] != 0;
  uint8_t* fb = ctx.buf<uint8_t>("ln.fb", n);
  float* dU2 = ctx.buf<float>("ln.u2", n);
  float* dV2 = ctx.buf<float>("ln.v2", n);
  EbOverride ov;
  ov.gain = gain;
  ov.fallback = fb;
  for (int it = 0;; it++) {
    CPSZ_HIP(hipMemcpyAsync(fb, hf.data(), n, hipMemcpyHostToDevice, st));
    compress_2d_device(ctx, dU, dV, r1, r2, max_pwr_eb, out, ov);
    decompress_2d_device(ctx, out.data, nullptr, out.size, r1, r2, dU2, dV2);
    cp_list(ctx, dU2, dV2, r1, r2, id1, ty1);
    cp_list(ctx, dU2, dV2, r1, r2, is1, ts1, true);
    stats.iterations = it + 1;
    if (!verify_round_2d(it, id0, ty0, is0, ts0, id1, ty1, is1, ts1, r2, hf.data(), n)) break;
  }
  for (size_t v = 0; v < n; v++) stats.fallback_vertices += hf[v] && hg[v];
}

// =================================================================== 3D ====
namespace {

// 3x3x3 patch of u, then v, then w (di slowest, dk fastest; zero outside)
// divided by the patch's largest magnitude, then log2(analytic / max) / 16
// (-4 for a zero bound), zeros up to kFeatDim3. Vertex idx[q], or v0 + q.
__global__ void k_feat3d(const float* __restrict__ U, const float* __restrict__ V,
                         const float* __restrict__ W, const double* __restrict__ E, int r1, int r2, int r3,
                         double max_pwr_eb, const uint32_t* __restrict__ idx, size_t v0, size_t m,
                         uint16_t* __restrict__ X) {
  const size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (q >= m) return;
  const size_t v = idx ? idx[q] : v0 + q;
  const size_t d0 = (size_t)r2 * r3;
  const int i = (int)(v / d0);
  const size_t rem = v - (size_t)i * d0;
  const int j = (int)(rem / (size_t)r3), k = (int)(rem - (size_t)j * r3);
  const float* src[3] = {U, V, W};
  float pt[81];
  float s = 0.f;
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++)
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const int ii = i + a - 1, jj = j + b - 1, kk = k + c - 1;
        const bool in = ii >= 0 && ii < r1 && jj >= 0 && jj < r2 && kk >= 0 && kk < r3;
        const size_t o = in ? (size_t)ii * d0 + (size_t)jj * r3 + kk : 0;
#pragma unroll
        for (int p = 0; p < 3; p++) {
          const float x = in ? src[p][o] : 0.f;
          pt[p * 27 + a * 9 + b * 3 + c] = x;
          s = fmaxf(s, fabsf(x));
        }
      }
  if (!(s > 0.f)) s = 1.f;
  uint16_t* x = X + q * kFeatDim3;
#pragma unroll
  for (int t = 0; t < 81; t++) x[t] = f2bf(pt[t] / s);
  const double e = E[v];
  const float le = e > 0 ? fmaxf((float)log2(e / max_pwr_eb), -64.f) / 16.f : -4.f;
  x[81] = f2bf(le);
  for (int t = kFeatUsed3; t < kFeatDim3; t++) x[t] = 0;
}

// Bisection labels over the incident tetrahedra. Slot s = 6 c + t: tetrahedron
// t (judge order) of the cell whose corner c (di << 2 | dj << 1 | dk) is the
// vertex; incident when c is one of its corners {0, a(t), b(t), 7} and the
// cell lies inside the grid. Base types are kept as type + 1 in 4-bit fields.
__global__ void __launch_bounds__(128) k_labels3d(const float* __restrict__ U, const float* __restrict__ V,
                                                  const float* __restrict__ W, const double* __restrict__ E,
                                                  int r1, int r2, int r3, double max_pwr_eb,
                                                  const uint32_t* __restrict__ idx, size_t m,
                                                  uint8_t* __restrict__ L) {
  const size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (q >= m) return;
  const size_t v = idx ? idx[q] : q;
  const double e = E[v];
  if (!(e > 0)) {
    L[q] = 0;
    return;
  }
  const size_t d0 = (size_t)r2 * r3, d1 = r3;
  const int i = (int)(v / d0);
  const size_t rem = v - (size_t)i * d0;
  const int j = (int)(rem / d1), k = (int)(rem - (size_t)j * d1);
  // value of tetrahedron slot s, with the vertex itself replaced by xv when set
  auto eval = [&](int sl, const float* xv) {
    const int c = sl / 6, t = sl - 6 * (sl / 6);
    const int ca = judge_tet_a(t), cb = judge_tet_b(t);
    const size_t org = v - (size_t)(c >> 2) * d0 - (size_t)((c >> 1) & 1) * d1 - (size_t)(c & 1);
    const int cs[4] = {0, ca, cb, 7};
    float f[4][3];
    for (int p = 0; p < 4; p++) {
      const int cp = cs[p];
      if (xv && cp == c) {
        f[p][0] = xv[0];
        f[p][1] = xv[1];
        f[p][2] = xv[2];
      } else {
        const size_t o = org + (size_t)(cp >> 2) * d0 + (size_t)((cp >> 1) & 1) * d1 + (size_t)(cp & 1);
        f[p][0] = U[o];
        f[p][1] = V[o];
        f[p][2] = W[o];
      }
    }
    return tet_cp(f, ca, cb);
  };
  uint64_t inc = 0, bt0 = 0, bt1 = 0, bt2 = 0;
  for (int sl = 0; sl < 48; sl++) {
    const int c = sl / 6, t = sl - 6 * (sl / 6);
    const int ci = i - (c >> 2), cj = j - ((c >> 1) & 1), ck = k - (c & 1);
    if (ci < 0 || ci >= r1 - 1 || cj < 0 || cj >= r2 - 1 || ck < 0 || ck >= r3 - 1) continue;
    if (!(c == 0 || c == 7 || c == judge_tet_a(t) || c == judge_tet_b(t))) continue;
    inc |= 1ull << sl;
    const uint64_t ty = (uint64_t)(eval(sl, nullptr) + 1) << (4 * (sl & 15));
    if (sl < 16) bt0 |= ty;
    else if (sl < 32) bt1 |= ty;
    else bt2 |= ty;
  }
  const float x0[3] = {U[v], V[v], W[v]};
  int gmax = 0;
  while (gmax < kMaxGain && ldexp(e, gmax) < max_pwr_eb) gmax++;
  auto ok = [&](int g) {
    const double up = 1.0 + fmin(max_pwr_eb, ldexp(e, g));
    for (int corner = 0; corner < 8; corner++) {
      float xv[3];
      for (int p = 0; p < 3; p++)
        xv[p] = ((corner >> p) & 1) ? (float)((double)x0[p] * up) : (float)((double)x0[p] / up);
      for (uint64_t mm = inc; mm; mm &= mm - 1) {
        const int sl = __builtin_ctzll(mm);
        const uint64_t wd = sl < 16 ? bt0 : (sl < 32 ? bt1 : bt2);
        const int base = (int)((wd >> (4 * (sl & 15))) & 15u);
        if (eval(sl, xv) + 1 != base) return false;
      }
    }
    return true;
  };
  int lo = 0, hi = gmax;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (ok(mid)) lo = mid;
    else hi = mid - 1;
  }
  L[q] = (uint8_t)lo;
}

void check_field_3d(const float* dU, const float* dV, const float* dW, size_t r1, size_t r2, size_t r3,
                    const char* who) {
  CPSZ_REQUIRE(dU && dV && dW && r1 && r2 && r3 && r1 * r2 * r3 < (1ull << 32) && r1 < (1u << 20) &&
                   r2 < (1u << 20) && r3 < (1u << 20),
               kErrArg, std::string(who) + ": bad arguments");
}

void features_3d_range(Ctx& ctx, const float* dU, const float* dV, const float* dW, const double* dE, size_t r1,
                       size_t r2, size_t r3, double max_pwr_eb, const uint32_t* d_idx, size_t v0, size_t m,
                       uint16_t* dX) {
  if (!m) return;
  k_feat3d<<<blocks_for(m), 256, 0, ctx.stream()>>>(dU, dV, dW, dE, (int)r1, (int)r2, (int)r3, max_pwr_eb, d_idx,
                                                    v0, m, dX);
  CPSZ_HIP(hipGetLastError());
}

void cp_list_3d(Ctx& ctx, const float* dU, const float* dV, const float* dW, size_t r1, size_t r2, size_t r3,
                std::vector<int64_t>& ids, std::vector<int32_t>& types) {
  size_t cap = 1 << 16, count = 0;
  for (;;) {
    ids.resize(cap);
    types.resize(cap);
    critical_points_3d(ctx, dU, dV, dW, r1, r2, r3, ids.data(), types.data(), cap, &count);
    if (count <= cap) break;
    cap = count;
  }
  ids.resize(count);
  types.resize(count);
}

// simplices whose critical point appeared, vanished or changed type
std::vector<int64_t> cp_diff(const std::vector<int64_t>& id0, const std::vector<int32_t>& ty0,
                             const std::vector<int64_t>& id1, const std::vector<int32_t>& ty1) {
  std::vector<int64_t> bad;
  size_t a = 0, b = 0;
  while (a < id0.size() || b < id1.size()) {
    if (b == id1.size() || (a < id0.size() && id0[a] < id1[b])) bad.push_back(id0[a++]);
    else if (a == id0.size() || id1[b] < id0[a]) bad.push_back(id1[b++]);
    else {
      if (ty0[a] != ty1[b]) bad.push_back(id0[a]);
      a++;
      b++;
    }
  }
  return bad;
}

}  // namespace

void analytic_eb_3d(Ctx& ctx, const float* dU, const float* dV, const float* dW, size_t r1, size_t r2,
                    size_t r3, double max_pwr_eb, double* dE) {
  eb_3d_device(ctx, dU, dV, dW, r1, r2, r3, max_pwr_eb, dE);
}

void features_3d(Ctx& ctx, const float* dU, const float* dV, const float* dW, const double* dE, size_t r1,
                 size_t r2, size_t r3, double max_pwr_eb, const uint32_t* d_idx, size_t m, uint16_t* dX) {
  check_field_3d(dU, dV, dW, r1, r2, r3, "features_3d");
  CPSZ_REQUIRE(dE && dX, kErrArg, "features_3d: bad arguments");
  CPSZ_REQUIRE(d_idx || m <= r1 * r2 * r3, kErrArg, "features_3d: more rows than vertices");
  features_3d_range(ctx, dU, dV, dW, dE, r1, r2, r3, max_pwr_eb, d_idx, 0, m, dX);
}

void labels_3d(Ctx& ctx, const float* dU, const float* dV, const float* dW, const double* dE, size_t r1,
               size_t r2, size_t r3, double max_pwr_eb, const uint32_t* d_idx, size_t m, uint8_t* dL) {
  check_field_3d(dU, dV, dW, r1, r2, r3, "labels_3d");
  CPSZ_REQUIRE(dE && dL, kErrArg, "labels_3d: bad arguments");
  CPSZ_REQUIRE(d_idx || m <= r1 * r2 * r3, kErrArg, "labels_3d: more rows than vertices");
  if (!m) return;
  k_labels3d<<<blocks_for(m, 128), 128, 0, ctx.stream()>>>(dU, dV, dW, dE, (int)r1, (int)r2, (int)r3, max_pwr_eb,
                                                           d_idx, m, dL);
  CPSZ_HIP(hipGetLastError());
}

void train_3d(Ctx& ctx, const std::vector<TrainField3>& fields, double max_pwr_eb, size_t spf, int epochs,
              int batch, float lr, uint64_t seed, float* params, float* loss_per_epoch) {
  CPSZ_REQUIRE(!fields.empty() && spf > 0 && epochs > 0 && batch >= 32 && batch <= (1 << 20), kErrArg,
               "train_3d: bad arguments");
  hipStream_t st = ctx.stream();
  const size_t S = fields.size() * spf;
  uint16_t* Xall = ctx.buf<uint16_t>("tr3.x", S * kFeatDim3);
  float* Tall = ctx.buf<float>("tr3.t", S);
  uint8_t* L8 = ctx.buf<uint8_t>("tr3.l8", spf);
  uint32_t* idx = ctx.buf<uint32_t>("tr3.idx", spf);
  uint32_t* hidx = (uint32_t*)ctx.host_buf("tr3.idx", spf * 4);
  uint64_t rs = seed;
  for (size_t f = 0; f < fields.size(); f++) {
    const TrainField3& F = fields[f];
    check_field_3d(F.dU, F.dV, F.dW, F.r1, F.r2, F.r3, "train_3d");
    const size_t n = F.r1 * F.r2 * F.r3;
    double* E = ctx.buf<double>("tr3.e", n);
    analytic_eb_3d(ctx, F.dU, F.dV, F.dW, F.r1, F.r2, F.r3, max_pwr_eb, E);  // synchronizes
    for (size_t k = 0; k < spf; k++) hidx[k] = (uint32_t)(splitmix(rs) % n);
    CPSZ_HIP(hipMemcpyAsync(idx, hidx, spf * 4, hipMemcpyHostToDevice, st));
    features_3d_range(ctx, F.dU, F.dV, F.dW, E, F.r1, F.r2, F.r3, max_pwr_eb, idx, 0, spf,
                      Xall + f * spf * kFeatDim3);
    labels_3d(ctx, F.dU, F.dV, F.dW, E, F.r1, F.r2, F.r3, max_pwr_eb, idx, spf, L8);
    k_u8_to_f32<<<blocks_for(spf), 256, 0, st>>>(L8, spf, Tall + f * spf);
    CPSZ_HIP(hipGetLastError());
    CPSZ_HIP(hipStreamSynchronize(st));  // host index buffer reuse
  }
  train_samples(ctx, Xall, Tall, S, kFeatDim3, epochs, batch, lr, rs, params, loss_per_epoch);
}

void compress_3d_learned_batch(Ctx& ctx, const std::vector<TrainField3>& fields, double max_pwr_eb,
                               const float* params, int workers, int grid_cap, const uint8_t*& base,
                               std::vector<size_t>& off, std::vector<size_t>& size, LearnedStats& stats) {
  const size_t nf = fields.size();
  CPSZ_REQUIRE(nf > 0 && params, kErrArg, "compress_3d_learned_batch: bad arguments");
  hipStream_t st = ctx.stream();
  stats = LearnedStats();
  std::vector<size_t> nofs(nf + 1, 0);
  size_t maxn = 0;
  for (size_t k = 0; k < nf; k++) {
    const TrainField3& F = fields[k];
    check_field_3d(F.dU, F.dV, F.dW, F.r1, F.r2, F.r3, "compress_3d_learned_batch");
    const size_t n = F.r1 * F.r2 * F.r3;
    nofs[k + 1] = nofs[k] + n;
    maxn = std::max(maxn, n);
  }
  const size_t ntot = nofs[nf];
  // ---- proposals: gain = clamp(floor(network output), 0, 24) per vertex
  float* P = ctx.buf<float>("l3.p", kNumParams3);
  CPSZ_HIP(hipMemcpyAsync(P, params, kNumParams3 * 4, hipMemcpyHostToDevice, st));
  uint8_t* gain = ctx.buf<uint8_t>("l3.gain", ntot);
  uint8_t* fb = ctx.buf<uint8_t>("l3.fb", ntot);
  double* E = ctx.buf<double>("l3.e", maxn);
  for (size_t k = 0; k < nf; k++) {
    const TrainField3& F = fields[k];
    const size_t n = nofs[k + 1] - nofs[k];
    analytic_eb_3d(ctx, F.dU, F.dV, F.dW, F.r1, F.r2, F.r3, max_pwr_eb, E);
    mlp_infer_gains(ctx, 3, F.dU, F.dV, F.dW, E, F.r1, F.r2, F.r3, max_pwr_eb, P, 0, n, gain + nofs[k]);
  }
  std::vector<uint8_t> hg(ntot), hf(ntot, 0);
  CPSZ_HIP(hipMemcpyAsync(hg.data(), gain, ntot, hipMemcpyDeviceToHost, st));
  CPSZ_HIP(hipStreamSynchronize(st));
  for (size_t v = 0; v < ntot; v++) stats.boosted_vertices += hg[v] != 0;
  // ---- critical points of the originals
  std::vector<std::vector<int64_t>> id0(nf);
  std::vector<std::vector<int32_t>> ty0(nf);
  for (size_t k = 0; k < nf; k++) {
    const TrainField3& F = fields[k];
    cp_list_3d(ctx, F.dU, F.dV, F.dW, F.r1, F.r2, F.r3, id0[k], ty0[k]);
  }
  // ---- verify and fall back until every field keeps its critical points
  float* dec = ctx.buf<float>("l3.dec", 3 * ntot);
  std::vector<size_t> soff(nf + 1, 0);
  for (size_t k = 0; k < nf; k++) soff[k + 1] = soff[k] + (((nofs[k + 1] - nofs[k]) * 32 + (4u << 20) + 255) & ~(size_t)255);
  uint8_t* out = ctx.buf<uint8_t>("l3.out", soff[nf]);
  off.assign(soff.begin(), soff.end() - 1);
  size.assign(nf, 0);
  std::vector<size_t> pending(nf);
  for (size_t k = 0; k < nf; k++) pending[k] = k;
  Batch2dConfig cfg;
  cfg.workers = workers;
  cfg.grid_cap = grid_cap;
  const int kMaxIter = 16;
  for (int it = 0; it <= kMaxIter && !pending.empty(); it++) {
    if (it == kMaxIter)  // the plain method for whatever is left
      for (size_t k : pending) std::fill(hf.begin() + nofs[k], hf.begin() + nofs[k + 1], 1);
    std::vector<Field3d> sub;
    std::vector<EbOverride> ov;
    for (size_t k : pending) {
      const TrainField3& F = fields[k];
      sub.push_back(Field3d{F.dU, F.dV, F.dW, F.r1, F.r2, F.r3});
      EbOverride o;
      o.gain = gain + nofs[k];
      o.fallback = fb + nofs[k];
      ov.push_back(o);
      CPSZ_HIP(hipMemcpyAsync(fb + nofs[k], hf.data() + nofs[k], nofs[k + 1] - nofs[k], hipMemcpyHostToDevice, st));
    }
    BatchOut bo;
    compress_3d_batch(ctx, sub, max_pwr_eb, cfg, bo, &ov);
    std::vector<const uint8_t*> ins;
    std::vector<size_t> insz;
    std::vector<Field3dOut> outs;
    for (size_t q = 0; q < pending.size(); q++) {
      const size_t k = pending[q];
      const TrainField3& F = fields[k];
      ins.push_back(bo.base + bo.off[q]);
      insz.push_back(bo.size[q]);
      float* d = dec + 3 * nofs[k];
      const size_t n = nofs[k + 1] - nofs[k];
      outs.push_back(Field3dOut{d, d + n, d + 2 * n, F.r1, F.r2, F.r3});
    }
    decompress_3d_batch(ctx, ins, insz, outs, cfg);
    stats.iterations = it + 1;
    std::vector<size_t> still;
    for (size_t q = 0; q < pending.size(); q++) {
      const size_t k = pending[q];
      const TrainField3& F = fields[k];
      std::vector<int64_t> id1;
      std::vector<int32_t> ty1;
      cp_list_3d(ctx, outs[q].dU, outs[q].dV, outs[q].dW, F.r1, F.r2, F.r3, id1, ty1);
      const std::vector<int64_t> bad = cp_diff(id0[k], ty0[k], id1, ty1);
      if (bad.empty()) {
        CPSZ_REQUIRE(bo.size[q] <= soff[k + 1] - soff[k], kErrInternal, "learned stream bound exceeded");
        CPSZ_HIP(hipMemcpyAsync(out + soff[k], bo.base + bo.off[q], bo.size[q], hipMemcpyDeviceToDevice, st));
        size[k] = bo.size[q];
        continue;
      }
      CPSZ_REQUIRE(it < kMaxIter, kErrInternal, "learned compression: the plain bound changed a critical point");
      const size_t d0 = F.r2 * F.r3, d1 = F.r3;
      for (int64_t sid : bad) {
        const size_t org = (size_t)(sid / 6);
        const int which = (int)(sid % 6);
        int t = 0;
        while (judge_tet_sid(t) != which) t++;
        const int cs[4] = {0, judge_tet_a(t), judge_tet_b(t), 7};
        for (int c : cs) hf[nofs[k] + org + (size_t)(c >> 2) * d0 + (size_t)((c >> 1) & 1) * d1 + (c & 1)] = 1;
      }
      still.push_back(k);
    }
    // the accepted streams leave the batch buffer before the next round rewrites it
    CPSZ_HIP(hipStreamSynchronize(st));
    pending.swap(still);
  }
  CPSZ_REQUIRE(pending.empty(), kErrInternal, "learned compression did not converge");
  for (size_t v = 0; v < ntot; v++) stats.fallback_vertices += hf[v] && hg[v];
  base = out;
}

// ------------------------------------------------------- 2D learned batch --
// The single-field learned compressor over a batch: proposals per field in
// row chunks, then rounds of one batched compress (one persistent wavefront
// launch for all pending fields), one batched decompress and the exact judge
// per field; the vertices of every changed triangle fall back to the analytic
// bound until each field keeps its critical points.
void compress_2d_learned_batch(Ctx& ctx, const std::vector<TrainField>& fields, double max_pwr_eb,
                               const float* params, int workers, int grid_cap, const uint8_t*& base,
                               std::vector<size_t>& off, std::vector<size_t>& size, LearnedStats& stats) {
  const size_t nf = fields.size();
  CPSZ_REQUIRE(nf > 0 && params, kErrArg, "compress_2d_learned_batch: bad arguments");
  hipStream_t st = ctx.stream();
  stats = LearnedStats();
  std::vector<size_t> nofs(nf + 1, 0);
  size_t maxn = 0;
  for (size_t k = 0; k < nf; k++) {
    const TrainField& F = fields[k];
    const size_t n = F.r1 * F.r2;
    CPSZ_REQUIRE(F.dU && F.dV && n > 0 && n < (1ull << 32), kErrArg, "compress_2d_learned_batch: bad field");
    nofs[k + 1] = nofs[k] + n;
    maxn = std::max(maxn, n);
  }
  const size_t ntot = nofs[nf];
  // ---- proposals: gain = clamp(floor(network output), 0, 24) per vertex
  float* P = ctx.buf<float>("l2b.p", kNumParams);
  CPSZ_HIP(hipMemcpyAsync(P, params, kNumParams * 4, hipMemcpyHostToDevice, st));
  uint8_t* gain = ctx.buf<uint8_t>("l2b.gain", ntot);
  uint8_t* fb = ctx.buf<uint8_t>("l2b.fb", ntot);
  double* E = ctx.buf<double>("l2b.e", maxn);
  for (size_t k = 0; k < nf; k++) {
    const TrainField& F = fields[k];
    const size_t n = nofs[k + 1] - nofs[k];
    analytic_eb_2d(ctx, F.dU, F.dV, F.r1, F.r2, max_pwr_eb, E);
    mlp_infer_gains(ctx, 2, F.dU, F.dV, nullptr, E, F.r1, F.r2, 1, max_pwr_eb, P, 0, n, gain + nofs[k]);
  }
  std::vector<uint8_t> hg(ntot), hf(ntot, 0);
  CPSZ_HIP(hipMemcpyAsync(hg.data(), gain, ntot, hipMemcpyDeviceToHost, st));
  CPSZ_HIP(hipStreamSynchronize(st));
  for (size_t v = 0; v < ntot; v++) stats.boosted_vertices += hg[v] != 0;
  // ---- critical points of the originals (closed-simplex and SoS tests)
  std::vector<std::vector<int64_t>> id0(nf), is0(nf);
  std::vector<std::vector<int32_t>> ty0(nf), ts0(nf);
  for (size_t k = 0; k < nf; k++) {
    cp_list(ctx, fields[k].dU, fields[k].dV, fields[k].r1, fields[k].r2, id0[k], ty0[k]);
    cp_list(ctx, fields[k].dU, fields[k].dV, fields[k].r1, fields[k].r2, is0[k], ts0[k], true);
  }
  // ---- verify and fall back until every field keeps its critical points
  float* dec = ctx.buf<float>("l2b.dec", 2 * ntot);
  std::vector<size_t> soff(nf + 1, 0);
  for (size_t k = 0; k < nf; k++) soff[k + 1] = soff[k] + ((12 * (nofs[k + 1] - nofs[k]) + (1u << 20) + 255) & ~(size_t)255);
  uint8_t* out = ctx.buf<uint8_t>("l2b.out", soff[nf]);
  off.assign(soff.begin(), soff.end() - 1);
  size.assign(nf, 0);
  std::vector<size_t> pending(nf);
  for (size_t k = 0; k < nf; k++) pending[k] = k;
  Batch2dConfig cfg;
  cfg.workers = workers;
  cfg.grid_cap = grid_cap;
  for (int it = 0; !pending.empty(); it++) {
    std::vector<Field2d> sub;
    std::vector<EbOverride> ov;
    for (size_t k : pending) {
      const TrainField& F = fields[k];
      sub.push_back(Field2d{F.dU, F.dV, F.r1, F.r2});
      EbOverride o;
      o.gain = gain + nofs[k];
      o.fallback = fb + nofs[k];
      ov.push_back(o);
      CPSZ_HIP(hipMemcpyAsync(fb + nofs[k], hf.data() + nofs[k], nofs[k + 1] - nofs[k], hipMemcpyHostToDevice, st));
    }
    BatchOut bo;
    compress_2d_batch(ctx, sub, max_pwr_eb, cfg, bo, &ov);
    std::vector<const uint8_t*> ins;
    std::vector<size_t> insz;
    std::vector<Field2dOut> outs;
    for (size_t q = 0; q < pending.size(); q++) {
      const size_t k = pending[q];
      const TrainField& F = fields[k];
      ins.push_back(bo.base + bo.off[q]);
      insz.push_back(bo.size[q]);
      float* d = dec + 2 * nofs[k];
      outs.push_back(Field2dOut{d, d + (nofs[k + 1] - nofs[k]), F.r1, F.r2});
    }
    decompress_2d_batch(ctx, ins, insz, outs, cfg);
    stats.iterations = it + 1;
    std::vector<size_t> still;
    for (size_t q = 0; q < pending.size(); q++) {
      const size_t k = pending[q];
      const TrainField& F = fields[k];
      std::vector<int64_t> id1, is1;
      std::vector<int32_t> ty1, ts1;
      cp_list(ctx, outs[q].dU, outs[q].dV, F.r1, F.r2, id1, ty1);
      cp_list(ctx, outs[q].dU, outs[q].dV, F.r1, F.r2, is1, ts1, true);
      if (!verify_round_2d(it, id0[k], ty0[k], is0[k], ts0[k], id1, ty1, is1, ts1, F.r2, hf.data() + nofs[k],
                           nofs[k + 1] - nofs[k])) {
        CPSZ_REQUIRE(bo.size[q] <= soff[k + 1] - soff[k], kErrInternal, "learned stream bound exceeded");
        CPSZ_HIP(hipMemcpyAsync(out + soff[k], bo.base + bo.off[q], bo.size[q], hipMemcpyDeviceToDevice, st));
        size[k] = bo.size[q];
        continue;
      }
      still.push_back(k);
    }
    // the accepted streams leave the batch buffer before the next round rewrites it
    CPSZ_HIP(hipStreamSynchronize(st));
    pending.swap(still);
  }
  CPSZ_REQUIRE(pending.empty(), kErrInternal, "learned compression did not converge");
  for (size_t v = 0; v < ntot; v++) stats.fallback_vertices += hf[v] && hg[v];
  base = out;
}

}  // namespace cpsz
