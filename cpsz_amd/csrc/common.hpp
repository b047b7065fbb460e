// Shared definitions of the cpsz MI355X library: status codes, error
// propagation, HIP checks and small host/device helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

namespace cpsz {

enum Status : int {
  kOk = 0,
  kErrArg = -1,          // bad argument (null pointer, zero size, ...)
  kErrHip = -2,          // HIP runtime error
  kErrAlloc = -3,        // allocation failed / capacity exceeded
  kErrStream = -4,       // malformed or truncated compressed stream
  kErrUnsupported = -5,  // valid input outside what this library supports
  kErrNonFinite = -6,    // NaN/Inf input (the reference crashes on these)
  kErrInternal = -7,
};

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const std::string& msg) { throw Error(code, msg); }

#define CPSZ_HIP(x)                                                                         \
  do {                                                                                      \
    hipError_t e_ = (x);                                                                    \
    if (e_ != hipSuccess)                                                                   \
      ::cpsz::fail(::cpsz::kErrHip, std::string(#x) + " failed: " + hipGetErrorString(e_)); \
  } while (0)

#define CPSZ_REQUIRE(cond, code, msg) \
  do {                                \
    if (!(cond)) ::cpsz::fail((code), (msg)); \
  } while (0)

constexpr int kRadius = 32768;        // intv_radius (src/sz_compress_3d.cpp:204)
constexpr int kCapacity = 65536;      // capacity (src/sz_compress_3d.cpp:202)
constexpr int kBlock2d = 12;          // SZ block size of the 2D codec (include/sz_compress_3d.hpp:12)
constexpr int kMaxEbCode = 1024;      // histogram bins reserved for eb id codes

inline size_t div_up(size_t a, size_t b) { return (a + b - 1) / b; }

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2: block b runs on XCD b % 8. xcd_block renumbers the blocks so that each
// XCD receives a contiguous range of logical blocks; kernels whose
// neighbouring blocks read or write the same cache lines (stream-order passes
// over the wavefront layout, block rows of 2D tiles) then find those lines in
// one L2 instead of fetching them once per XCD.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
  const uint32_t per = nb / kXcds, rem = nb % kXcds, x = b % kXcds, k = b / kXcds;
  return (x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per) + k;
}

}  // namespace cpsz
