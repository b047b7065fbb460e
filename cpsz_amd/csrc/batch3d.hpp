// Batched 3D codec over many fields or timesteps (see batch3d.cpp).
#pragma once
#include <vector>

#include "batch2d.hpp"
#include "codec3d.hpp"

namespace cpsz {

struct Field3d {
  const float *dU, *dV, *dW;
  size_t r1, r2, r3;
};
struct Field3dOut {
  float *dU, *dV, *dW;
  size_t r1, r2, r3;
};

// ov: null, or one EbOverride per field (the learned proposer's hook).
void compress_3d_batch(Ctx& ctx, const std::vector<Field3d>& fields, double max_pwr_eb,
                       const Batch2dConfig& cfg, BatchOut& out, const std::vector<EbOverride>* ov = nullptr);
void decompress_3d_batch(Ctx& ctx, const std::vector<const uint8_t*>& d_in, const std::vector<size_t>& in_size,
                         const std::vector<Field3dOut>& fields, const Batch2dConfig& cfg);
// Cached configuration for this batch signature; with tune, times the
// candidates on this batch first (results are identical for every candidate).
Batch2dConfig batch_3d_config(Ctx& ctx, const std::vector<Field3d>& fields, double max_pwr_eb, bool tune);

}  // namespace cpsz
