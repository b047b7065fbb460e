// Reader of a compressed stream held in device memory (or in host memory when
// the caller has a host copy). Header fields sit close together, so reads go
// through a cached 64 KiB window in pinned host memory: one device-to-host
// copy serves every field inside it instead of one copy (and one stream
// synchronization) per field. A returned pointer is valid until the next read.
#pragma once
#include <algorithm>
#include <cstring>
#include <vector>

#include "ctx.hpp"

namespace cpsz {

struct StreamReader {
  static constexpr size_t kWindow = 64u << 10;
  Ctx& ctx;
  const uint8_t* host;  // may be null
  const uint8_t* dev;
  size_t size;
  uint8_t* win = nullptr;
  size_t win_off = 0, win_len = 0;
  // host copies of stream ranges fetched beforehand (for many streams in a
  // few batched copies); reads inside one of them need no copy
  struct Range {
    size_t off, len;
    const uint8_t* p;
  };
  std::vector<Range> pre;

  StreamReader(Ctx& c, const uint8_t* h, const uint8_t* d, size_t n) : ctx(c), host(h), dev(d), size(n) {}

  const uint8_t* read(size_t off, size_t len) {
    CPSZ_REQUIRE(off <= size && len <= size - off, kErrStream, "stream truncated");
    if (host) return host + off;
    for (const Range& r : pre)
      if (off >= r.off && off + len <= r.off + r.len) return r.p + (off - r.off);
    if (win && off >= win_off && off + len <= win_off + win_len) return win + (off - win_off);
    if (len <= kWindow) {
      if (!win) win = (uint8_t*)ctx.host_buf("rd.window", kWindow);
      win_off = off;
      win_len = std::min(kWindow, size - off);
      if (win_len) {
        CPSZ_HIP(hipMemcpyAsync(win, dev + off, win_len, hipMemcpyDeviceToHost, ctx.stream()));
        CPSZ_HIP(hipStreamSynchronize(ctx.stream()));
      }
      return win;
    }
    // longer reads (a large Huffman tree) go through their own pinned buffer:
    // a pageable destination would be staged by the runtime in pieces
    uint8_t* big = (uint8_t*)ctx.host_buf("rd.big", len);
    CPSZ_HIP(hipMemcpyAsync(big, dev + off, len, hipMemcpyDeviceToHost, ctx.stream()));
    CPSZ_HIP(hipStreamSynchronize(ctx.stream()));
    return big;
  }
  template <typename V>
  V get(size_t off) {
    V v;
    std::memcpy(&v, read(off, sizeof(V)), sizeof(V));
    return v;
  }
};

}  // namespace cpsz
