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

constexpr int kFeatDim = 64;   // 2D: 51 used, zero padded
constexpr int kFeatUsed = 51;  // 25 u + 25 v + log2(analytic / max)
constexpr int kFeatDim3 = 96;  // 3D: 82 used, zero padded
constexpr int kFeatUsed3 = 82; // 27 u + 27 v + 27 w + log2(analytic / max)
constexpr int kHidden = 128;
constexpr int kMaxGain = 24;

// fp32 parameter block for feature width F: W1[F][128], b1[128],
// W2[128][128], b2[128], w3[128], b3
struct NetLayout {
  int fdim;
  size_t w1, b1, w2, b2, w3, b3, count;
};
constexpr NetLayout net_layout(int fdim) {
  return NetLayout{fdim,
                   0,
                   (size_t)fdim * kHidden,
                   (size_t)fdim * kHidden + kHidden,
                   (size_t)fdim * kHidden + kHidden + (size_t)kHidden * kHidden,
                   (size_t)fdim * kHidden + 2 * kHidden + (size_t)kHidden * kHidden,
                   (size_t)fdim * kHidden + 3 * kHidden + (size_t)kHidden * kHidden,
                   (size_t)fdim * kHidden + 3 * kHidden + (size_t)kHidden * kHidden + 1};
}
constexpr size_t kOffW1 = net_layout(kFeatDim).w1;
constexpr size_t kOffB1 = net_layout(kFeatDim).b1;
constexpr size_t kOffW2 = net_layout(kFeatDim).w2;
constexpr size_t kOffB2 = net_layout(kFeatDim).b2;
constexpr size_t kOffW3 = net_layout(kFeatDim).w3;
constexpr size_t kOffB3 = net_layout(kFeatDim).b3;
constexpr size_t kNumParams = net_layout(kFeatDim).count;
constexpr size_t kNumParams3 = net_layout(kFeatDim3).count;

// He-uniform weights (fused inputs of fdim), zero biases and padding rows.
void mlp_init_params(uint64_t seed, float* params, int fdim = kFeatDim, int fused = kFeatUsed);

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

// Proposal gains of vertices v0 .. v0+m (row-major) in one fused kernel:
// features from the field and the analytic bound dE, both hidden layers on
// MFMA with the weights in LDS, gain = clamp(floor(network output), 0, 24)
// into d_gain[0..m). dim 2 uses (dU, dV) and the 51-feature net, dim 3
// (dU, dV, dW) and the 82-feature net.
void mlp_infer_gains(Ctx& ctx, int dim, const float* dU, const float* dV, const float* dW, const double* dE,
                     size_t r1, size_t r2, size_t r3, double max_pwr_eb, const float* d_params, size_t v0, size_t m,
                     uint8_t* d_gain);

// C[M][N] (fp32, row-major ldc) = sum_k A(m,k) B(k,n) with BF16 operands
// A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn]; M, N multiples of 32,
// K a multiple of 8. split > 1 sums K slices with atomics into a zeroed C.
void gemm_bf16(Ctx& ctx, int M, int N, int K, const uint16_t* A, long sam, long sak, const uint16_t* B,
               long sbk, long sbn, float* C, int ldc, int split);

// dX: m rows of fdim BF16 features; d_params in the layout of net_layout(fdim).
void mlp_forward(Ctx& ctx, const float* d_params, const uint16_t* dX, size_t m, float* dY, int fdim = kFeatDim);
// Mean squared error of the network output against dT and its gradient.
void mlp_grad(Ctx& ctx, const float* d_params, const uint16_t* dX, const float* dT, size_t m, float* d_grad,
              float* d_loss, int fdim = kFeatDim);
// Adam on S samples (features Xall, targets Tall) in shuffled batches;
// params (host) is the start point and receives the result.
void train_samples(Ctx& ctx, const uint16_t* Xall, const float* Tall, size_t S, int fdim, int epochs, int batch,
                   float lr, uint64_t& rng, float* params, float* loss_per_epoch);

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
// Learned-bound compression of a batch of 2D fields (one persistent wavefront
// launch per verification round); streams in a library-owned device buffer.
void compress_2d_learned_batch(Ctx& ctx, const std::vector<TrainField>& fields, double max_pwr_eb,
                               const float* params, int workers, int grid_cap, const uint8_t*& base,
                               std::vector<size_t>& off, std::vector<size_t>& size, LearnedStats& st);

// ---- 3D (learn3d.hip): 3x3x3 patch of (u, v, w) plus the analytic bound;
// labels keep every incident tetrahedron's critical point (24 per interior
// vertex) when the vertex alone moves to any corner of its box
// x * (1 + eb)^(+-1) per component (the offline_log codec keeps the sign and
// |log2 x' - log2 x| < log2(1 + eb)).
void analytic_eb_3d(Ctx& ctx, const float* dU, const float* dV, const float* dW, size_t r1, size_t r2,
                    size_t r3, double max_pwr_eb, double* dE);
void features_3d(Ctx& ctx, const float* dU, const float* dV, const float* dW, const double* dE, size_t r1,
                 size_t r2, size_t r3, double max_pwr_eb, const uint32_t* d_idx, size_t m, uint16_t* dX);
void labels_3d(Ctx& ctx, const float* dU, const float* dV, const float* dW, const double* dE, size_t r1,
               size_t r2, size_t r3, double max_pwr_eb, const uint32_t* d_idx, size_t m, uint8_t* dL);
struct TrainField3 {
  const float *dU, *dV, *dW;
  size_t r1, r2, r3;
};
void train_3d(Ctx& ctx, const std::vector<TrainField3>& fields, double max_pwr_eb, size_t samples_per_field,
              int epochs, int batch, float lr, uint64_t seed, float* params, float* loss_per_epoch);
// Learned-bound compression of a batch of 3D fields (one persistent
// wavefront launch per verification round); the streams are returned in a
// library-owned device buffer like the batch codec's.
void compress_3d_learned_batch(Ctx& ctx, const std::vector<TrainField3>& fields, double max_pwr_eb,
                               const float* params, int workers, int grid_cap, const uint8_t*& base,
                               std::vector<size_t>& off, std::vector<size_t>& size, LearnedStats& st);

}  // namespace cpsz
