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
#include <cstdio>
#include <cstdlib>
#include <cstring>
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

// A compress of one exact batch (device, input pointers, shapes, bound, grid)
// captured into a HIP graph. The job tables it reads are uploaded from its
// own pinned copy before every launch (other calls reuse the device tables
// in between); the device addresses baked into it stay valid while the
// context's buffer generation is unchanged.
struct Graph2d {
  hipGraphExec_t exec = nullptr;
  uint64_t generation = 0;
  std::vector<Ctx::Upload> uploads;
  uint8_t* staging = nullptr;  // pinned copy of the tables
  const uint8_t* dout = nullptr;
  std::vector<size_t> off, cap;
  uint64_t last_use = 0;
};
constexpr size_t kMaxGraphs = 8;
std::mutex g_graph_m;
std::map<std::string, Graph2d> g_graphs;
uint64_t g_graph_clock = 0;

void drop_graph(Graph2d& g) {
  if (g.exec) (void)hipGraphExecDestroy(g.exec);
  if (g.staging) (void)hipHostFree(g.staging);
  g.exec = nullptr;
  g.staging = nullptr;
}

std::string graph_key(int device, const std::vector<Field2d>& f, double eb, int grid_cap) {
  std::string k = std::to_string(device) + "|" + std::to_string(grid_cap) + "|" + (exact_2d_bounds() ? "x" : "a");
  uint64_t bits;
  std::memcpy(&bits, &eb, 8);
  k += "|" + std::to_string(bits);
  for (const auto& x : f)
    k += "|" + std::to_string((uintptr_t)x.dU) + "," + std::to_string((uintptr_t)x.dV) + "," +
         std::to_string(x.r1) + "," + std::to_string(x.r2);
  return k;
}

}  // namespace

namespace {

// The batched compress as stream-ordered work: bounds, wavefront, entropy
// (the last one waits for the stream and reports unless a graph is being
// captured).
void enqueue_2d_batch(Ctx& ctx, const std::vector<Field2d>& fields, double max_pwr_eb, const Batch2dConfig& cfg,
                      BatchOut& out, std::vector<size_t>& cap, const std::vector<EbOverride>* ov) {
  const size_t nf = fields.size();
  std::vector<size_t> soff(nf + 1, 0);
  out.off.assign(nf, 0);
  out.size.assign(nf, 0);
  size_t obytes = 0;
  for (size_t k = 0; k < nf; k++) {
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
  cap.resize(nf);
  for (size_t k = 0; k < nf; k++) cap[k] = stream_bound(w[k].r1 * w[k].r2);
  c2_entropy_batch(ctx, wp.data(), nf, dout, out.off, cap, out.size, cfg.workers, cfg.device_trees);
  ctx.mark("entropy");
  out.base = dout;
}

// Replays the captured graph of this exact batch if there is a valid one:
// kReplayed; kNone: capture one after an eager run; kEager: this batch could
// not be captured (run it eagerly, do not try again).
enum Replay { kReplayed, kNone, kEager };
Replay replay_2d_batch(Ctx& ctx, const std::string& key, size_t nf, BatchOut& out) {
  hipStream_t st = ctx.stream();
  std::lock_guard<std::mutex> g(g_graph_m);
  auto it = g_graphs.find(key);
  if (it == g_graphs.end()) return kNone;
  Graph2d& G = it->second;
  if (G.generation != ctx.generation()) {  // a buffer moved since the capture
    drop_graph(G);
    g_graphs.erase(it);
    return kNone;
  }
  if (!G.exec) return kEager;
  for (const auto& u : G.uploads)
    CPSZ_HIP(hipMemcpyAsync(u.dst, G.staging + u.off, u.bytes, hipMemcpyHostToDevice, st));
  CPSZ_HIP(hipGraphLaunch(G.exec, st));
  G.last_use = ++g_graph_clock;
  out.base = G.dout;
  out.off = G.off;
  out.size.assign(nf, 0);
  c2_entropy_batch_finish(ctx, nf, G.cap, out.size);
  return kReplayed;
}

// Captures the batch (just run eagerly, so every buffer has its size) into a
// graph for the next calls. A batch that cannot be captured is run eagerly.
void capture_2d_batch(Ctx& ctx, const std::string& key, const std::vector<Field2d>& fields, double max_pwr_eb,
                      const Batch2dConfig& cfg) {
  hipStream_t st = ctx.stream();
  const uint64_t gen = ctx.generation();
  ctx.uploads.clear();
  ctx.upload_bytes.clear();
  BatchOut o;
  std::vector<size_t> cap;
  hipGraph_t graph = nullptr;
  std::string why;
  hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed);
  bool ok = e == hipSuccess;
  if (!ok) why = std::string("begin capture: ") + hipGetErrorString(e);
  if (ok) {
    ctx.capturing = true;
    try {
      enqueue_2d_batch(ctx, fields, max_pwr_eb, cfg, o, cap, nullptr);
    } catch (const std::exception& x) {
      ok = false;
      why = x.what();
    }
    ctx.capturing = false;
    e = hipStreamEndCapture(st, &graph);
    if (e != hipSuccess && why.empty()) why = std::string("end capture: ") + hipGetErrorString(e);
    ok = ok && e == hipSuccess && graph;
    if (ok && ctx.generation() != gen) {
      ok = false;
      why = "a buffer moved during the capture";
    }
  }
  (void)hipGetLastError();
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  CPSZ_HIP(hipStreamIsCapturing(st, &cs));
  CPSZ_REQUIRE(cs == hipStreamCaptureStatusNone, kErrInternal, "batch graph: the stream is still capturing (" + why + ")");
  if (!ok && std::getenv("CPSZ_GRAPH_DEBUG")) std::fprintf(stderr, "cpsz: batch not captured: %s\n", why.c_str());
  Graph2d G;
  G.generation = gen;
  if (ok) ok = hipGraphInstantiate(&G.exec, graph, nullptr, nullptr, 0) == hipSuccess;
  if (graph) (void)hipGraphDestroy(graph);
  (void)hipGetLastError();
  if (ok && !ctx.upload_bytes.empty())
    ok = hipHostMalloc(&G.staging, ctx.upload_bytes.size(), hipHostMallocDefault) == hipSuccess;
  if (!ok) drop_graph(G);  // kept without a graph: the batch runs eagerly from now on
  if (ok && !ctx.upload_bytes.empty()) std::memcpy(G.staging, ctx.upload_bytes.data(), ctx.upload_bytes.size());
  G.uploads = ctx.uploads;
  G.dout = o.base;
  G.off = o.off;
  G.cap = cap;
  std::lock_guard<std::mutex> g(g_graph_m);
  if (g_graphs.size() >= kMaxGraphs) {  // forget the least recently used one
    auto old = g_graphs.begin();
    for (auto jt = g_graphs.begin(); jt != g_graphs.end(); ++jt)
      if (jt->second.last_use < old->second.last_use) old = jt;
    drop_graph(old->second);
    g_graphs.erase(old);
  }
  G.last_use = ++g_graph_clock;
  g_graphs[key] = G;
}

}  // namespace

void compress_2d_batch(Ctx& ctx, const std::vector<Field2d>& fields, double max_pwr_eb,
                       const Batch2dConfig& cfg, BatchOut& out, const std::vector<EbOverride>* ov) {
  const size_t nf = fields.size();
  CPSZ_REQUIRE(nf > 0, kErrArg, "compress_2d_batch: empty batch");
  CPSZ_REQUIRE(!ov || ov->size() == nf, kErrArg, "compress_2d_batch: one override per field");
  for (size_t k = 0; k < nf; k++)
    CPSZ_REQUIRE(fields[k].dU && fields[k].dV && fields[k].r1 && fields[k].r2, kErrArg,
                 "compress_2d_batch: bad field");
  ctx.stages.clear();
  ctx.mark("start");
  // with the trees on the GPU the compress has no host work between its
  // launches: the same batch is then replayed from a captured graph
  const bool graph = cfg.device_trees && !ov;
  const std::string key = graph ? graph_key(ctx.device(), fields, max_pwr_eb, cfg.grid_cap) : std::string();
  const Replay r = graph ? replay_2d_batch(ctx, key, nf, out) : kEager;
  if (r == kReplayed) {
    ctx.mark("graph");
    ctx.finish_marks();
    return;
  }
  std::vector<size_t> cap;
  enqueue_2d_batch(ctx, fields, max_pwr_eb, cfg, out, cap, ov);
  ctx.finish_marks();
  if (r == kNone) capture_2d_batch(ctx, key, fields, max_pwr_eb, cfg);
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
  {  // trees on the GPU instead, at the best worker count and grid: the
     // first call captures the batch, the second replays it
    Batch2dConfig c = best;
    c.device_trees = true;
    compress_2d_batch(ctx, fields, max_pwr_eb, c, o);
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
