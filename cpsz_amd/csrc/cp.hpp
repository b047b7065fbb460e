// Critical-point extraction on the GPU (see cp.hip).
#pragma once
#include "ctx.hpp"

namespace cpsz {

void critical_points_2d(Ctx& ctx, const float* dU, const float* dV, size_t r1, size_t r2,
                        int64_t* ids, int32_t* types, size_t cap, size_t* count);
void critical_points_3d(Ctx& ctx, const float* dU, const float* dV, const float* dW, size_t r1,
                        size_t r2, size_t r3, int64_t* ids, int32_t* types, size_t cap,
                        size_t* count);

}  // namespace cpsz
