// Batched compression / decompression of many 2D fields of uneven sizes.
//
// Compress: ONE bounds launch over every field's tiles, ONE persistent
// wavefront launch over every field's strips (a single field cannot fill the
// GPU: its wavefront has at most 2 * ceil(r1/64) waves), then the entropy
// stage of all fields as a few launches over the whole batch (histograms,
// encodes, unpredictable gathers, assembly), with the host Huffman trees
// built in parallel on the worker pool. Decompress: per-field parsing and
// decoder tables on the workers, then every field's bitstreams decoded
// together (one grid per decode pass for all fields), the unpredictable
// scatter on the workers and one persistent wavefront launch. Every field's
// stream is byte-identical to the single-field codec's.
//
// The autotuner times candidate configurations (worker count, persistent grid
// size) on the caller's batch itself (outputs do not depend on them) and
// caches the fastest per device and batch signature.
#include <algorithm>
#include <chrono>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "batch2d.hpp"
#include "pool.hpp"

namespace cpsz {

namespace {

std::string signature(const std::vector<Field2d>& f) {
  size_t n = 0, rmax = 0;
  for (const auto& x : f) {
    n += x.r1 * x.r2;
    rmax = std::max(rmax, x.r1);
  }
  int dev = 0;
  (void)hipGetDevice(&dev);
  return std::to_string(dev) + ":" + std::to_string(f.size()) + ":" + std::to_string(n >> 20) + ":" +
         std::to_string(rmax >> 6);
}

std::mutex g_tune_m;
std::map<std::string, Batch2dConfig> g_tuned;

size_t stream_bound(size_t n) { return 12 * n + (1u << 20); }

}  // namespace

void compress_2d_batch(Ctx& ctx, const std::vector<Field2d>& fields, double max_pwr_eb,
                       const Batch2dConfig& cfg, BatchOut& out, const std::vector<EbOverride>* ov) {
  const size_t nf = fields.size();
  CPSZ_REQUIRE(nf > 0, kErrArg, "compress_2d_batch: empty batch");
  CPSZ_REQUIRE(!ov || ov->size() == nf, kErrArg, "compress_2d_batch: one override per field");
  ctx.stages.clear();
  ctx.mark("start");
  std::vector<size_t> soff(nf + 1, 0);
  out.off.assign(nf, 0);
  out.size.assign(nf, 0);
  size_t obytes = 0;
  for (size_t k = 0; k < nf; k++) {
    CPSZ_REQUIRE(fields[k].dU && fields[k].dV && fields[k].r1 && fields[k].r2, kErrArg,
                 "compress_2d_batch: bad field");
    soff[k + 1] = soff[k] + c2_scratch_bytes(fields[k].r1, fields[k].r2);
    out.off[k] = obytes;
    obytes += (stream_bound(fields[k].r1 * fields[k].r2) + 255) & ~(size_t)255;
  }
  uint8_t* scratch = ctx.buf<uint8_t>("b2c.scratch", soff[nf]);
  uint8_t* dout = ctx.buf<uint8_t>("b2c.out", obytes);
  std::vector<C2Work> w(nf);
  std::vector<C2Work*> wp(nf);
  for (size_t k = 0; k < nf; k++) {
    w[k].dU = fields[k].dU;
    w[k].dV = fields[k].dV;
    w[k].r1 = fields[k].r1;
    w[k].r2 = fields[k].r2;
    c2_bind(w[k], scratch + soff[k]);
    wp[k] = &w[k];
  }
  // every field's bounds in one launch on the caller's stream (which also
  // orders them after the caller's writes of the inputs)
  c2_eb_batch(ctx, wp.data(), nf, max_pwr_eb, ov);
  ctx.mark("eb");
  c2_wavefront(ctx, wp.data(), nf, cfg.grid_cap);
  ctx.mark("lorenzo");
  std::vector<size_t> cap(nf);
  for (size_t k = 0; k < nf; k++) cap[k] = stream_bound(w[k].r1 * w[k].r2);
  c2_entropy_batch(ctx, wp.data(), nf, dout, out.off, cap, out.size, cfg.workers, cfg.device_trees);
  ctx.mark("entropy");
  out.base = dout;
  ctx.finish_marks();
}

void decompress_2d_batch(Ctx& ctx, const std::vector<const uint8_t*>& d_in, const std::vector<size_t>& in_size,
                         const std::vector<Field2dOut>& fields, const Batch2dConfig& cfg) {
  const size_t nf = fields.size();
  CPSZ_REQUIRE(nf > 0 && d_in.size() == nf && in_size.size() == nf, kErrArg, "decompress_2d_batch: bad batch");
  hipStream_t st = ctx.stream();
  ctx.stages.clear();
  ctx.mark("start");
  std::vector<size_t> soff(nf + 1, 0);
  for (size_t k = 0; k < nf; k++) {
    CPSZ_REQUIRE(d_in[k] && fields[k].dU && fields[k].dV && fields[k].r1 && fields[k].r2, kErrArg,
                 "decompress_2d_batch: bad field");
    soff[k + 1] = soff[k] + d2_scratch_bytes(fields[k].r1, fields[k].r2);
  }
  uint8_t* scratch = ctx.buf<uint8_t>("b2d.scratch", soff[nf]);
  std::vector<D2Work> w(nf);
  std::vector<D2Work*> wp(nf);
  for (size_t k = 0; k < nf; k++) {
    w[k].r1 = fields[k].r1;
    w[k].r2 = fields[k].r2;
    w[k].dU = fields[k].dU;
    w[k].dV = fields[k].dV;
    d2_bind(w[k], scratch + soff[k]);
    wp[k] = &w[k];
  }
  CPSZ_HIP(hipStreamSynchronize(st));  // inputs written on the caller's stream
  d2_entropy_batch(ctx, d_in, in_size, wp.data(), nf, cfg.workers);
  ctx.mark("entropy");
  d2_wavefront(ctx, wp.data(), nf, cfg.grid_cap);
  d2_check(ctx, wp.data(), nf);
  ctx.mark("lorenzo");
  ctx.finish_marks();
}

void batch_2d_set_config(const std::vector<Field2d>& fields, const Batch2dConfig& cfg) {
  const std::string key = signature(fields);
  std::lock_guard<std::mutex> g(g_tune_m);
  g_tuned[key] = cfg;
}

Batch2dConfig batch_2d_config(Ctx& ctx, const std::vector<Field2d>& fields, double max_pwr_eb, bool tune) {
  const std::string key = signature(fields);
  {
    std::lock_guard<std::mutex> g(g_tune_m);
    auto it = g_tuned.find(key);
    if (it != g_tuned.end() || !tune) return it != g_tuned.end() ? it->second : Batch2dConfig();
  }
  const int workers[] = {4, 8, 16};
  const int grids[] = {1024, 4096};
  Batch2dConfig best;
  double best_t = 1e30;
  BatchOut o;
  compress_2d_batch(ctx, fields, max_pwr_eb, best, o);  // warm-up: buffers, pools, tables
  for (int wk : workers)
    for (int gr : grids) {
      Batch2dConfig c;
      c.workers = wk;
      c.grid_cap = gr;
      CPSZ_HIP(hipStreamSynchronize(ctx.stream()));
      const auto t0 = std::chrono::steady_clock::now();
      compress_2d_batch(ctx, fields, max_pwr_eb, c, o);
      CPSZ_HIP(hipStreamSynchronize(ctx.stream()));
      const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      if (t < best_t) {
        best_t = t;
        best = c;
      }
    }
  {  // trees on the GPU instead, at the best worker count and grid
    Batch2dConfig c = best;
    c.device_trees = true;
    CPSZ_HIP(hipStreamSynchronize(ctx.stream()));
    const auto t0 = std::chrono::steady_clock::now();
    compress_2d_batch(ctx, fields, max_pwr_eb, c, o);
    CPSZ_HIP(hipStreamSynchronize(ctx.stream()));
    const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (t < best_t) {
      best_t = t;
      best = c;
    }
  }
  std::lock_guard<std::mutex> g(g_tune_m);
  g_tuned[key] = best;
  return best;
}

}  // namespace cpsz
