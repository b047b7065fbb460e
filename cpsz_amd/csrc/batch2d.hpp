// Batched 2D codec over many fields of uneven sizes (see batch2d.cpp).
#pragma once
#include <vector>

#include "codec2d.hpp"

namespace cpsz {

struct Field2d {
  const float* dU;
  const float* dV;
  size_t r1, r2;
};
struct Field2dOut {
  float* dU;
  float* dV;
  size_t r1, r2;
};
struct Batch2dConfig {
  int workers = 8;     // concurrent entropy-stage workers (own stream each)
  int grid_cap = 4096; // workgroups of the persistent wavefront launch
  bool device_trees = false;  // Huffman trees on the GPU instead of the host pool
};
struct BatchOut {
  const uint8_t* base = nullptr;  // library-owned device buffer
  std::vector<size_t> off, size;  // per field
};

// ov: null, or one EbOverride per field (the learned proposer's hook).
void compress_2d_batch(Ctx& ctx, const std::vector<Field2d>& fields, double max_pwr_eb,
                       const Batch2dConfig& cfg, BatchOut& out, const std::vector<EbOverride>* ov = nullptr);
void decompress_2d_batch(Ctx& ctx, const std::vector<const uint8_t*>& d_in, const std::vector<size_t>& in_size,
                         const std::vector<Field2dOut>& fields, const Batch2dConfig& cfg);
// Cached configuration for this batch signature; with tune, times the
// candidates on this batch first (results are identical for every candidate).
Batch2dConfig batch_2d_config(Ctx& ctx, const std::vector<Field2d>& fields, double max_pwr_eb, bool tune);
// Pin the configuration used for this batch signature (replaces a tuned one).
void batch_2d_set_config(const std::vector<Field2d>& fields, const Batch2dConfig& cfg);

}  // namespace cpsz
