// 3D offline_log codec entry points on device buffers (see codec3d.hip).
#pragma once
#include "codec2d.hpp"

namespace cpsz {

void compress_3d_device(Ctx& ctx, const float* dU, const float* dV, const float* dW, size_t r1,
                        size_t r2, size_t r3, double max_pwr_eb, DeviceStream& out);
void decompress_3d_device(Ctx& ctx, const uint8_t* d_in, const uint8_t* h_in, size_t in_size,
                          size_t r1, size_t r2, size_t r3, float* dU, float* dV, float* dW);

}  // namespace cpsz
