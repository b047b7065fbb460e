// 2D offline codec entry points on device buffers (see codec2d.hip).
#pragma once
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

// U, V: device pointers to r1 x r2 row-major float fields.
void compress_2d_device(Ctx& ctx, const float* dU, const float* dV, size_t r1, size_t r2,
                        double max_pwr_eb, DeviceStream& out, const EbOverride& ov = EbOverride());
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
