// 3D offline_log codec entry points on device buffers (see codec3d.hip).
#pragma once
#include "codec2d.hpp"

namespace cpsz {

void compress_3d_device(Ctx& ctx, const float* dU, const float* dV, const float* dW, size_t r1,
                        size_t r2, size_t r3, double max_pwr_eb, DeviceStream& out);
void decompress_3d_device(Ctx& ctx, const uint8_t* d_in, const uint8_t* h_in, size_t in_size,
                          size_t r1, size_t r2, size_t r3, float* dU, float* dV, float* dW);

// Host emulations of the 3D device arithmetic (CPU tests): per-vertex bounds,
// and the Lorenzo pass from log values and eb codes to stream ids and codes.
void eb3d_emulate(const float* U, const float* V, const float* W, size_t r1, size_t r2, size_t r3,
                  double max_pwr_eb, double* eb);
void lorenzo3d_emulate(const float* L, const uint8_t* code, size_t r1, size_t r2, size_t r3,
                       uint8_t* vid, uint16_t* q);

}  // namespace cpsz
