// Batched compression / decompression of many 3D fields (timesteps of a
// simulation, ensemble members), possibly of different sizes.
//
// Compress: the eb, sign and log kernels run per field on the caller's
// stream, then ONE persistent wavefront launch covers every field's (plane
// group, row block) items (a 192^3 field alone offers 144 workgroups of work
// to 256 CUs), then the entropy stages (histograms, host Huffman trees,
// encode, stream assembly) of different fields run concurrently on the worker
// pool. Decompress: parsing, Huffman decoding and scatter per field on the
// workers, one persistent wavefront launch, then the exp2 finish per field.
// Every field's stream is byte-identical to the single-field codec's.
#include <algorithm>
#include <chrono>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "batch3d.hpp"
#include "pool.hpp"

namespace cpsz {

namespace {

std::string signature(const std::vector<Field3d>& f) {
  size_t n = 0, rmax = 0;
  for (const auto& x : f) {
    n += x.r1 * x.r2 * x.r3;
    rmax = std::max(rmax, x.r1);
  }
  int dev = 0;
  (void)hipGetDevice(&dev);
  return std::to_string(dev) + ":" + std::to_string(f.size()) + ":" + std::to_string(n >> 20) + ":" +
         std::to_string(rmax >> 2);
}

std::mutex g_tune_m;
std::map<std::string, Batch2dConfig> g_tuned;

// a 3D stream never exceeds its header, 3 sign bitmaps, 12 B per vertex of
// unpredictable values and the two Huffman sections (<= 4 B per symbol each
// plus their trees)
size_t stream_bound(size_t n) { return 32 * n + (4u << 20); }

}  // namespace

void compress_3d_batch(Ctx& ctx, const std::vector<Field3d>& fields, double max_pwr_eb,
                       const Batch2dConfig& cfg, BatchOut& out, const std::vector<EbOverride>* ov) {
  const size_t nf = fields.size();
  CPSZ_REQUIRE(nf > 0, kErrArg, "compress_3d_batch: empty batch");
  CPSZ_REQUIRE(!ov || ov->size() == nf, kErrArg, "compress_3d_batch: one override per field expected");
  ctx.stages.clear();
  ctx.mark("start");
  std::vector<size_t> soff(nf + 1, 0);
  out.off.assign(nf, 0);
  out.size.assign(nf, 0);
  size_t obytes = 0;
  for (size_t k = 0; k < nf; k++) {
    const Field3d& F = fields[k];
    CPSZ_REQUIRE(F.dU && F.dV && F.dW && F.r1 && F.r2 && F.r3, kErrArg, "compress_3d_batch: bad field");
    soff[k + 1] = soff[k] + c3_scratch_bytes(F.r1, F.r2, F.r3);
    out.off[k] = obytes;
    obytes += (stream_bound(F.r1 * F.r2 * F.r3) + 255) & ~(size_t)255;
  }
  uint8_t* scratch = ctx.buf<uint8_t>("b3c.scratch", soff[nf]);
  uint8_t* dout = ctx.buf<uint8_t>("b3c.out", obytes);
  std::vector<C3Work> w(nf);
  std::vector<C3Work*> wp(nf);
  for (size_t k = 0; k < nf; k++) {
    w[k].dU = fields[k].dU;
    w[k].dV = fields[k].dV;
    w[k].dW = fields[k].dW;
    w[k].r1 = fields[k].r1;
    w[k].r2 = fields[k].r2;
    w[k].r3 = fields[k].r3;
    c3_bind(w[k], scratch + soff[k]);
    wp[k] = &w[k];
  }
  // inputs written on the caller's stream must be complete before the workers'
  // streams read them
  CPSZ_HIP(hipStreamSynchronize(ctx.stream()));
  // bounds passes on the workers' streams: one field's grid fills only part of
  // the device, so several fields run at once
  {
    std::vector<std::function<void(Ctx&)>> eb_jobs(nf);
    for (size_t k = 0; k < nf; k++)
      eb_jobs[k] = [&, k](Ctx& wc) {
        c3_eb(wc, w[k], max_pwr_eb, ov ? (*ov)[k] : EbOverride());
        CPSZ_HIP(hipStreamSynchronize(wc.stream()));
      };
    pool_run(pool_for(ctx.device()), eb_jobs, cfg.workers);
  }
  c3_check(ctx, wp.data(), nf);
  ctx.mark("eb");
  c3_wavefront(ctx, wp.data(), nf, cfg.grid_cap);
  CPSZ_HIP(hipStreamSynchronize(ctx.stream()));
  ctx.mark("lorenzo");
  std::vector<std::function<void(Ctx&)>> jobs(nf);
  for (size_t k = 0; k < nf; k++) {
    jobs[k] = [&, k](Ctx& wc) {
      DeviceStream s;
      c3_entropy(wc, w[k], s);
      CPSZ_REQUIRE(s.size <= stream_bound(w[k].r1 * w[k].r2 * w[k].r3), kErrInternal,
                   "batch stream bound exceeded");
      CPSZ_HIP(hipMemcpyAsync(dout + out.off[k], s.data, s.size, hipMemcpyDeviceToDevice, wc.stream()));
      CPSZ_HIP(hipStreamSynchronize(wc.stream()));
      out.size[k] = s.size;
    };
  }
  pool_run(pool_for(ctx.device()), jobs, cfg.workers);
  ctx.mark("entropy");
  out.base = dout;
  ctx.finish_marks();
}

void decompress_3d_batch(Ctx& ctx, const std::vector<const uint8_t*>& d_in, const std::vector<size_t>& in_size,
                         const std::vector<Field3dOut>& fields, const Batch2dConfig& cfg) {
  const size_t nf = fields.size();
  CPSZ_REQUIRE(nf > 0 && d_in.size() == nf && in_size.size() == nf, kErrArg, "decompress_3d_batch: bad batch");
  hipStream_t st = ctx.stream();
  ctx.stages.clear();
  ctx.mark("start");
  std::vector<size_t> soff(nf + 1, 0);
  for (size_t k = 0; k < nf; k++) {
    const Field3dOut& F = fields[k];
    CPSZ_REQUIRE(d_in[k] && F.dU && F.dV && F.dW && F.r1 && F.r2 && F.r3, kErrArg,
                 "decompress_3d_batch: bad field");
    soff[k + 1] = soff[k] + d3_scratch_bytes(F.r1, F.r2, F.r3);
  }
  uint8_t* scratch = ctx.buf<uint8_t>("b3d.scratch", soff[nf]);
  std::vector<D3Work> w(nf);
  std::vector<D3Work*> wp(nf);
  for (size_t k = 0; k < nf; k++) {
    w[k].r1 = fields[k].r1;
    w[k].r2 = fields[k].r2;
    w[k].r3 = fields[k].r3;
    w[k].dU = fields[k].dU;
    w[k].dV = fields[k].dV;
    w[k].dW = fields[k].dW;
    d3_bind(w[k], scratch + soff[k]);
    wp[k] = &w[k];
  }
  CPSZ_HIP(hipStreamSynchronize(st));  // inputs written on the caller's stream
  std::vector<std::function<void(Ctx&)>> jobs(nf);
  for (size_t k = 0; k < nf; k++) {
    jobs[k] = [&, k](Ctx& wc) {
      d3_entropy(wc, d_in[k], nullptr, in_size[k], w[k]);
      CPSZ_HIP(hipStreamSynchronize(wc.stream()));
    };
  }
  pool_run(pool_for(ctx.device()), jobs, cfg.workers);
  ctx.mark("entropy");
  d3_wavefront(ctx, wp.data(), nf, cfg.grid_cap);
  ctx.mark("lorenzo");
  for (size_t k = 0; k < nf; k++) d3_finish(ctx, w[k]);
  ctx.mark("finish");
  d3_check(ctx, wp.data(), nf);
  ctx.finish_marks();
}

Batch2dConfig batch_3d_config(Ctx& ctx, const std::vector<Field3d>& fields, double max_pwr_eb, bool tune) {
  const std::string key = signature(fields);
  {
    std::lock_guard<std::mutex> g(g_tune_m);
    auto it = g_tuned.find(key);
    if (it != g_tuned.end() || !tune) return it != g_tuned.end() ? it->second : Batch2dConfig();
  }
  const int workers[] = {4, 8, 16};
  const int grids[] = {512, 1024, 2048};
  Batch2dConfig best;
  double best_t = 1e30;
  BatchOut o;
  compress_3d_batch(ctx, fields, max_pwr_eb, best, o);  // warm-up: buffers, pools, tables
  for (int wk : workers)
    for (int gr : grids) {
      Batch2dConfig c;
      c.workers = wk;
      c.grid_cap = gr;
      CPSZ_HIP(hipStreamSynchronize(ctx.stream()));
      const auto t0 = std::chrono::steady_clock::now();
      compress_3d_batch(ctx, fields, max_pwr_eb, c, o);
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
