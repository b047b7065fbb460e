// Per-device context of the library: stream, grow-only device buffers, pinned
// host staging, and the eb-id threshold tables derived from the host libm.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "common.hpp"
#include "libm_tables.hpp"

namespace cpsz {

// Smallest inputs at which the reference's eb id computation reaches each id,
// computed once per process with the host libm the reference links against:
//  2D (include/sz_cp_preserve_utils.hpp:14-22, base 4 over DBL_EPSILON):
//    q2d[k] = min { x > eps : (int)(log2(x/eps)/log2(4)) >= k }, k >= 1
//  3D (src/sz_compress_cp_preserve_3d.cpp:470-476, base 2 over FLT_EPSILON on
//    log2(1+eb)): q3d[k] = min { e > 0 : id3(e) >= k }, k >= 1
struct EbTables {
  std::vector<double> q2d;  // index k; q2d[0] unused; +inf past the range
  std::vector<double> q3d;
  // 2D offline_log (src/sz_compress_cp_preserve_2d.cpp:327-331, eb
  // log2(1+e) base 2 over DBL_EPSILON): q2l[k] = min { e : code(e) >= k },
  // code 0 = zero precision, code c >= 1 = id c - 1
  std::vector<double> q2l;
};
const EbTables& eb_tables();

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
};

class Ctx {
 public:
  explicit Ctx(int device);
  ~Ctx();
  Ctx(const Ctx&) = delete;
  Ctx& operator=(const Ctx&) = delete;

  int device() const { return device_; }
  hipStream_t stream() const { return stream_; }
  void set_stream(hipStream_t s) { stream_ = s ? s : own_stream_; }
  // Set on the contexts of batch workers: many fields already keep the host
  // cores busy, so a codec runs its host work inline instead of on a helper.
  bool batch_worker = false;

  // Named grow-only device scratch buffer (contents undefined).
  template <typename T>
  T* buf(const char* name, size_t count) {
    return static_cast<T*>(raw_buf(name, count * sizeof(T)));
  }
  void* raw_buf(const char* name, size_t bytes);
  // Pinned host staging buffer (grow-only).
  void* host_buf(const char* name, size_t bytes);

  // Device copies of the eb threshold tables.
  const double* d_q2d();
  const double* d_q3d();
  const double* d_q2l();
  int q2l_size() const { return (int)eb_tables().q2l.size(); }
  int q2d_size() const { return (int)eb_tables().q2d.size(); }
  int q3d_size() const { return (int)eb_tables().q3d.size(); }
  // Device copies of the libm exception tables (uploaded on first use).
  struct LibmDev {
    LibmView log2f, exp2f;
  };
  const LibmDev& libm_dev();

  // Records an event on the stream and returns it: a host wait on it lets
  // later work queued on the stream overlap the host's own work.
  hipEvent_t record_point();

  // Host-side job tables sent to the device. host_table() returns the named
  // pinned staging buffer once the previous upload from it has left (so it
  // can be refilled without a stream synchronization); upload_table() sends
  // it. While a graph is being captured (capturing), uploads are not put on
  // the stream but kept in `uploads` with a copy of their bytes, to be
  // replayed before each launch of the graph.
  void* host_table(const char* name, size_t bytes);
  void upload_table(const char* name, void* dst, const void* src, size_t bytes);
  struct Upload {
    void* dst;
    size_t off, bytes;  // into upload_bytes
  };
  bool capturing = false;
  std::vector<Upload> uploads;
  std::vector<uint8_t> upload_bytes;
  // Incremented whenever a named buffer moves (is freed to grow): device
  // addresses baked into a captured graph stay valid while it is unchanged.
  uint64_t generation() const { return generation_; }

  // Stage timings of the last call (ms), filled when timing is enabled.
  bool timing = false;
  std::vector<std::pair<std::string, float>> stages;
  void mark(const char* name);
  void finish_marks();

 private:
  int device_;
  hipStream_t stream_ = nullptr;
  hipStream_t own_stream_ = nullptr;
  std::map<std::string, DevBuf> bufs_;
  std::map<std::string, DevBuf> hbufs_;
  double* d_q2d_ = nullptr;
  double* d_q3d_ = nullptr;
  double* d_q2l_ = nullptr;
  LibmDev libm_{};
  std::vector<void*> libm_allocs_;
  bool libm_ready_ = false;
  std::vector<std::pair<std::string, hipEvent_t>> marks_;
  std::vector<hipEvent_t> event_pool_;
  hipEvent_t point_ = nullptr;
  std::map<std::string, hipEvent_t> table_sent_;
  uint64_t generation_ = 0;
};

// Process-wide context for a device (created on first use).
Ctx& ctx_for_device(int device);

}  // namespace cpsz
