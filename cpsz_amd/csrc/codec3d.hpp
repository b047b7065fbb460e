// 3D offline_log codec entry points on device buffers (see codec3d.hip).
#pragma once
#include "codec2d.hpp"

namespace cpsz {

// ov: optional per-vertex change of the analytic bound (row-major vertex
// index), as for the 2D codec.
void compress_3d_device(Ctx& ctx, const float* dU, const float* dV, const float* dW, size_t r1,
                        size_t r2, size_t r3, double max_pwr_eb, DeviceStream& out,
                        const EbOverride& ov = EbOverride());
void decompress_3d_device(Ctx& ctx, const uint8_t* d_in, const uint8_t* h_in, size_t in_size,
                          size_t r1, size_t r2, size_t r3, float* dU, float* dV, float* dW);

// Device check of the approximate bound's premises: the largest ulp distance
// of the fast cube root over [0, 2^126), and the largest ratio of the fast sign
// bound's error to its envelope over hashed argument pairs (<= 1 when sound).
void check_fast_bounds(Ctx& ctx, uint32_t* max_ulp, double* max_ratio);

// Analytic per-vertex bound of the plain method (before quantization), row-major.
void eb_3d_device(Ctx& ctx, const float* dU, const float* dV, const float* dW, size_t r1, size_t r2, size_t r3,
                  double max_pwr_eb, double* dE);

// ---- staged form: the single-field calls run the stages in order; the batch
// driver runs eb per field, one wavefront launch for all fields, and the
// entropy stage of different fields concurrently on worker streams. Arrays
// marked "wave layout" follow the wavefront's own order.
struct C3Work {
  const float *dU = nullptr, *dV = nullptr, *dW = nullptr;
  size_t r1 = 0, r2 = 0, r3 = 0;
  uint8_t* ebc = nullptr;    // eb codes (wave layout)
  float* L = nullptr;        // log2 |x| [3] (wave layout)
  float* R = nullptr;        // log-domain reconstruction [3] (wave layout)
  uint16_t* q = nullptr;     // quantization codes [3] (wave layout)
  uint8_t* vid = nullptr;    // stream ids (wave layout)
  uint64_t* smap = nullptr;  // sign bitmaps [3][ceil(n / 64)]
  int* prog = nullptr;       // wavefront progress counters
  int* flags = nullptr;      // [0] non-finite input or bound
};
size_t c3_scratch_bytes(size_t r1, size_t r2, size_t r3);
void c3_bind(C3Work& w, uint8_t* scratch);  // carves c3_scratch_bytes of device memory
void c3_eb(Ctx& ctx, C3Work& w, double max_pwr_eb, const EbOverride& ov);
void c3_check(Ctx& ctx, C3Work* const* w, size_t nf);  // synchronizes
void c3_wavefront(Ctx& ctx, C3Work* const* w, size_t nf, int grid_cap);
void c3_entropy(Ctx& ctx, const C3Work& w, DeviceStream& out);  // synchronizes

struct D3Work {
  size_t r1 = 0, r2 = 0, r3 = 0;
  float *dU = nullptr, *dV = nullptr, *dW = nullptr;  // outputs
  uint8_t* vid = nullptr;     // stream ids, row-major
  uint16_t* qc = nullptr;     // decoded quantization codes (stream order)
  float* unp = nullptr;       // unpredictable values (stream order)
  uint16_t* q = nullptr;      // quantization codes [3] (wave layout)
  float* init = nullptr;      // log values of unpredictables [3] (wave layout)
  float* R = nullptr;         // log-domain reconstruction [3] (wave layout)
  uint8_t* vs = nullptr;      // stream ids (wave layout)
  int *prog = nullptr, *err = nullptr;
  uint8_t* sm = nullptr;      // sign bitmaps copied from the stream
  uint32_t* cnt = nullptr;    // per-chunk predictable / unpredictable counts
  uint64_t* coff = nullptr;   // their exclusive scans
  size_t nunp = 0;            // unpredictable floats in the stream
};
size_t d3_scratch_bytes(size_t r1, size_t r2, size_t r3);
void d3_bind(D3Work& w, uint8_t* scratch);
void d3_entropy(Ctx& ctx, const uint8_t* d_in, const uint8_t* h_in, size_t in_size, D3Work& w);
void d3_wavefront(Ctx& ctx, D3Work* const* w, size_t nf, int grid_cap);
void d3_finish(Ctx& ctx, const D3Work& w);
void d3_check(Ctx& ctx, D3Work* const* w, size_t nf);  // synchronizes

// Host emulations of the 3D device arithmetic (CPU tests): per-vertex bounds,
// and the Lorenzo pass from log values and eb codes to stream ids and codes.
void eb3d_emulate(const float* U, const float* V, const float* W, size_t r1, size_t r2, size_t r3,
                  double max_pwr_eb, double* eb);
void lorenzo3d_emulate(const float* L, const uint8_t* code, size_t r1, size_t r2, size_t r3,
                       uint8_t* vid, uint16_t* q);

}  // namespace cpsz
