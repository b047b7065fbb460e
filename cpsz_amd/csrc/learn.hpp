// Learned error-bound proposer for the 2D codec (no counterpart in the
// reference; its nearest precedent is the verify-and-halve loop of
// sz_compress_cp_preserve_2d_online_abs_relax_FN, src/sz_compress_cp_preserve_2d.cpp:1439-1502).
//
// A small MLP (51 -> 128 -> 128 -> 1, BF16 weights and activations on MFMA,
// FP32 accumulation) reads a vertex's 5x5 neighbourhood of (u, v) plus its
// analytic bound and proposes an integer gain g: bound = min(max, analytic *
// 2^g). Training labels come from bisection under an exact critical-point
// test. Every proposal is verified on the decompressed field with the exact
// judge; the vertices of any triangle whose critical point changed fall back
// to the analytic bound and the field is recompressed, so the result has zero
// critical-point errors and the plain method is both the fallback and the
// judge. Streams keep the reference layout (only the eb ids differ).
#pragma once
#include <cstdint>
#include <vector>

#include "codec2d.hpp"
#include "ctx.hpp"

namespace cpsz {

constexpr int kFeatDim = 64;   // 51 used, zero padded
constexpr int kFeatUsed = 51;  // 25 u + 25 v + log2(analytic / max)
constexpr int kHidden = 128;
constexpr int kMaxGain = 24;

// fp32 parameter block: W1[64][128], b1[128], W2[128][128], b2[128], w3[128], b3
constexpr size_t kOffW1 = 0;
constexpr size_t kOffB1 = kOffW1 + (size_t)kFeatDim * kHidden;
constexpr size_t kOffW2 = kOffB1 + kHidden;
constexpr size_t kOffB2 = kOffW2 + (size_t)kHidden * kHidden;
constexpr size_t kOffW3 = kOffB2 + kHidden;
constexpr size_t kOffB3 = kOffW3 + kHidden;
constexpr size_t kNumParams = kOffB3 + 1;

void mlp_init_params(uint64_t seed, float* params);

// Analytic per-vertex bound (the plain method's, before quantization).
void analytic_eb_2d(Ctx& ctx, const float* dU, const float* dV, size_t r1, size_t r2, double max_pwr_eb,
                    double* dE);
// Features of vertices idx[0..m) (all vertices in row-major order when idx is
// null), BF16 bits, kFeatDim per vertex.
void features_2d(Ctx& ctx, const float* dU, const float* dV, const double* dE, size_t r1, size_t r2,
                 double max_pwr_eb, const uint32_t* d_idx, size_t m, uint16_t* dX);
// Largest gain that keeps every incident triangle's critical point (existence
// and type) when the vertex alone moves to any corner of its bound box.
void labels_2d(Ctx& ctx, const float* dU, const float* dV, const double* dE, size_t r1, size_t r2,
               double max_pwr_eb, const uint32_t* d_idx, size_t m, uint8_t* dL);

// C[M][N] (fp32, row-major ldc) = sum_k A(m,k) B(k,n) with BF16 operands
// A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn]; M, N multiples of 32,
// K a multiple of 8. split > 1 sums K slices with atomics into a zeroed C.
void gemm_bf16(Ctx& ctx, int M, int N, int K, const uint16_t* A, long sam, long sak, const uint16_t* B,
               long sbk, long sbn, float* C, int ldc, int split);

void mlp_forward(Ctx& ctx, const float* d_params, const uint16_t* dX, size_t m, float* dY);
// Mean squared error of the network output against dT and its gradient.
void mlp_grad(Ctx& ctx, const float* d_params, const uint16_t* dX, const float* dT, size_t m, float* d_grad,
              float* d_loss);

struct TrainField {
  const float* dU;
  const float* dV;
  size_t r1, r2;
};
void train_2d(Ctx& ctx, const std::vector<TrainField>& fields, double max_pwr_eb, size_t samples_per_field,
              int epochs, int batch, float lr, uint64_t seed, float* params, float* loss_per_epoch);

struct LearnedStats {
  int iterations = 0;
  size_t fallback_vertices = 0;
  size_t boosted_vertices = 0;
};
void compress_2d_learned(Ctx& ctx, const float* dU, const float* dV, size_t r1, size_t r2,
                         double max_pwr_eb, const float* params, DeviceStream& out, LearnedStats& st);

}  // namespace cpsz
