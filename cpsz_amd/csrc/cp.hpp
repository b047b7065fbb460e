// Critical-point extraction on the GPU (see cp.hip).
#pragma once
#include "ctx.hpp"

namespace cpsz {

// sos: decide with Simulation of Simplicity (every zero in exactly one
// simplex) instead of the reference's closed-simplex test
void critical_points_2d(Ctx& ctx, const float* dU, const float* dV, size_t r1, size_t r2,
                        int64_t* ids, int32_t* types, size_t cap, size_t* count, bool sos = false);
// The 2D judge without any allocation: its scratch is a file-scope device
// array (per GPU, shared by every call on it; overlapping calls on one GPU
// race). At most kNoAllocMaxCp critical points; ids/types are caller host
// arrays of cap >= the count, filled sorted by id.
constexpr size_t kNoAllocMaxCp = 1u << 20;
void critical_points_2d_noalloc(int device, hipStream_t stream, const float* dU, const float* dV, size_t r1,
                                size_t r2, int64_t* ids, int32_t* types, size_t cap, size_t* count);
void critical_points_3d(Ctx& ctx, const float* dU, const float* dV, const float* dW, size_t r1,
                        size_t r2, size_t r3, int64_t* ids, int32_t* types, size_t cap,
                        size_t* count, bool sos = false);

}  // namespace cpsz
