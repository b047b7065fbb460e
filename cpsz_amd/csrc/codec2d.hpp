// 2D offline codec entry points on device buffers (see codec2d.hip).
#pragma once
#include "ctx.hpp"

namespace cpsz {

struct DeviceStream {
  const uint8_t* data = nullptr;  // library-owned device buffer (valid until the next call)
  size_t size = 0;
};

// U, V: device pointers to r1 x r2 row-major float fields.
void compress_2d_device(Ctx& ctx, const float* dU, const float* dV, size_t r1, size_t r2,
                        double max_pwr_eb, DeviceStream& out);
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
