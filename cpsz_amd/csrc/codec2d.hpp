// 2D offline codec entry points on device buffers (see codec2d.hip).
#pragma once
#include <vector>

#include "ctx.hpp"

namespace cpsz {

struct DeviceStream {
  const uint8_t* data = nullptr;  // library-owned device buffer (valid until the next call)
  size_t size = 0;
};

// Optional per-vertex change of the analytic bound (the learned proposer's
// hook; all null = the reference's plain method):
//   gain[v] > 0 and !fallback[v]: bound = min(max_pwr_eb, analytic * 2^gain[v])
//   eb_out[v] receives the analytic bound (before any gain).
struct EbOverride {
  const uint8_t* gain = nullptr;
  const uint8_t* fallback = nullptr;
  double* eb_out = nullptr;
};
inline bool has_override(const EbOverride& o) { return o.gain || o.fallback || o.eb_out; }

// Bounds of float fields without overrides are computed with the last FP64
// steps approximated in FP32 and exact fix-ups where the approximation does
// not fix a vertex's codes; the switch selects the all-FP64 kernel instead
// (process-wide; both give the same codes).
void set_exact_2d_bounds(bool on);
bool exact_2d_bounds();

// U, V: device pointers to r1 x r2 row-major float fields.
void compress_2d_device(Ctx& ctx, const float* dU, const float* dV, size_t r1, size_t r2,
                        double max_pwr_eb, DeviceStream& out, const EbOverride& ov = EbOverride());
// ---- staged form: the single-field calls run the stages in order; the batch
// driver runs eb per field, one wavefront launch for all fields, and the
// entropy stage of different fields concurrently on worker streams.
struct C2Work {
  const float* dU = nullptr;
  const float* dV = nullptr;
  size_t r1 = 0, r2 = 0;
  uint16_t *cu = nullptr, *cv = nullptr;  // eb codes (0 = zero precision, else id + 1)
  uint16_t *qu = nullptr, *qv = nullptr;  // quantization codes, row-major
  uint32_t* hist = nullptr;               // kMaxEbCode + 2 * 65536 bins
  int* flags = nullptr;                   // [0] non-finite, [1] ids above 81
  uint64_t* bnd = nullptr;  // wavefront boundary rows (tagged values)
  int* prog = nullptr;      // wavefront item ticket
  bool wide = false;
  // the flags stay on the device: the wavefront reads `wide` from flags[1]
  // and skips a field with flags[0] set (the caller reports it afterwards)
  bool flags_on_device = false;
  // offline_log: one eb code per vertex (cu) for both components, the codec
  // runs on the log values lu/lv, and smap holds the two sign bitmaps as
  // 64-bit words (ceil(n/64) per component)
  bool log = false;
  const float* lu = nullptr;
  const float* lv = nullptr;
  const uint64_t* smap = nullptr;
  // T = double (the reference's double instantiation): the fields are dU64/dV64
  bool f64 = false;
  const double* dU64 = nullptr;
  const double* dV64 = nullptr;
};
size_t c2_scratch_bytes(size_t r1, size_t r2);
void c2_bind(C2Work& w, uint8_t* scratch);  // carves c2_scratch_bytes of device memory
void c2_eb(Ctx& ctx, C2Work& w, double max_pwr_eb, const EbOverride& ov);
void c2_check(Ctx& ctx, C2Work* const* w, size_t nf);  // synchronizes; sets wide
// Every field's bounds in one launch (plain offline float fields; ov: one
// override per field or null); re-points hist/flags to contiguous arrays
// (w[k]->flags = w[0]->flags + 4 k) and leaves the flags on the device, with
// no synchronization: the batched entropy stage reports them.
void c2_eb_batch(Ctx& ctx, C2Work* const* w, size_t nf, double max_pwr_eb, const std::vector<EbOverride>* ov);
void c2_wavefront(Ctx& ctx, C2Work* const* w, size_t nf, int grid_cap);
void c2_entropy(Ctx& ctx, const C2Work& w, DeviceStream& out);
// Every field's entropy stage at once (plain offline float fields): streams
// written to dout + off[k] (at most cap[k] bytes each), sizes into size[k].
// Huffman trees are built on the GPU (device_trees: one wave per tree, no
// host round trip) or on up to `workers` host pool threads; the streams are
// identical either way.
// The eb flags of the batch (c2_eb_batch) are checked here too. While the
// context captures a graph, the stage only enqueues its work (device trees
// only); c2_entropy_batch_finish then waits for the stream and reports the
// sizes and errors.
void c2_entropy_batch(Ctx& ctx, C2Work* const* w, size_t nf, uint8_t* dout, const std::vector<size_t>& off,
                      const std::vector<size_t>& cap, std::vector<size_t>& size, int workers, bool device_trees);
void c2_entropy_batch_finish(Ctx& ctx, size_t nf, const std::vector<size_t>& cap, std::vector<size_t>& size);

struct D2Work {
  size_t r1 = 0, r2 = 0;
  float *dU = nullptr, *dV = nullptr;         // outputs
  uint16_t *sid_u = nullptr, *sid_v = nullptr;  // stream ids
  uint16_t *qu = nullptr, *qv = nullptr;      // quantization codes
  float* raw = nullptr;                       // unpredictable values per element (2n)
  uint64_t* bnd = nullptr;  // wavefront boundary rows (tagged values)
  int *prog = nullptr, *err = nullptr;
  bool wide = false;
  bool log = false;               // offline_log stream: ids shared by u and v, log values
  const uint8_t* signs = nullptr;  // offline_log: the stream's two sign bitmaps (device)
  // T = double: outputs dU64/dV64, raw holds 2n doubles (d2_scratch_bytes(.., true))
  bool f64 = false;
  double *dU64 = nullptr, *dV64 = nullptr;
};
size_t d2_scratch_bytes(size_t r1, size_t r2, bool f64 = false);
void d2_bind(D2Work& w, uint8_t* scratch);  // carves d2_scratch_bytes(r1, r2, w.f64)
void d2_entropy(Ctx& ctx, const uint8_t* d_in, const uint8_t* h_in, size_t in_size, D2Work& w);
// The entropy stage of many fields at once (headers and trees on `workers`
// pool threads, every field's bitstreams decoded together on ctx's stream).
void d2_entropy_batch(Ctx& ctx, const std::vector<const uint8_t*>& d_in, const std::vector<size_t>& in_size,
                      D2Work* const* w, size_t nf, int workers);
void d2_wavefront(Ctx& ctx, D2Work* const* w, size_t nf, int grid_cap);
void d2_check(Ctx& ctx, D2Work* const* w, size_t nf);

// 2D offline_log pair (sz_compress_cp_preserve_2d_offline_log /
// sz_decompress_cp_preserve_2d_offline_log): the offline bounds taken in the
// log domain, the codec run on log2f|x|, sign bitmaps in the stream.
void compress_2d_log_device(Ctx& ctx, const float* dU, const float* dV, size_t r1, size_t r2,
                            double max_pwr_eb, DeviceStream& out);
void decompress_2d_log_device(Ctx& ctx, const uint8_t* d_in, const uint8_t* h_in, size_t in_size, size_t r1,
                              size_t r2, float* dU, float* dV);

// The T = double instantiation of the offline pair (the reference's
// src/sz_compress_cp_preserve_2d.cpp:280-282, src/sz_decompress_cp_preserve_2d.cpp:46-52):
// products, predictions and reconstructions in double as the reference's
// double code computes them, double unpredictable values in the stream.
void compress_2d_f64_device(Ctx& ctx, const double* dU, const double* dV, size_t r1, size_t r2,
                            double max_pwr_eb, DeviceStream& out);
void decompress_2d_f64_device(Ctx& ctx, const uint8_t* d_in, const uint8_t* h_in, size_t in_size, size_t r1,
                              size_t r2, double* dU, double* dV);

// Analytic per-vertex bound of the plain method (before quantization).
void eb_2d_device(Ctx& ctx, const float* dU, const float* dV, size_t r1, size_t r2, double max_pwr_eb,
                  double* dE);
// d_in: device copy of the stream; h_in: optional host copy (saves readbacks).
void decompress_2d_device(Ctx& ctx, const uint8_t* d_in, const uint8_t* h_in, size_t in_size,
                          size_t r1, size_t r2, float* dU, float* dV);

// Host emulations of the 2D wavefront arithmetic (for CPU tests): eb codes c
// (0 = zero precision, else id + 1) -> quant codes and reconstruction; and the
// decoder's reconstruction from codes, stream ids and unpredictable values.
void lorenzo2d_emulate(const float* x, const uint16_t* c, size_t r1, size_t r2, bool wide,
                       uint16_t* q, float* recon);
void lorenzo2d_emulate_decode(const uint16_t* q, const uint16_t* sid, const float* raw, size_t r1,
                              size_t r2, bool wide, float* out);

}  // namespace cpsz
