#include "ctx.hpp"

#include <cfloat>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>

namespace cpsz {

namespace {

// The reference's 2D id: eb_exponential_quantize(eb, 4, log2(4), DBL_EPSILON).
int id2d_ref(double x) {
  const double thr = std::numeric_limits<double>::epsilon();
  if (x <= thr) return 0;
  return (int)(std::log2(x / thr) / std::log2(4.0));
}

// The reference's 3D id for a vertex bound e > 0 (before the unpredictable
// remapping): eb_exponential_quantize(log2(1+e), 2, log2(2), FLT_EPSILON).
int id3d_ref(double e) {
  const double thr = std::numeric_limits<float>::epsilon();
  double a = std::log2(1 + e);
  if (a <= thr) return 0;
  return (int)(std::log2(a / thr) / std::log2(2.0));
}

// The reference's 2D offline_log eb code for a vertex bound e >= 0
// (src/sz_compress_cp_preserve_2d.cpp:327-331): a = log2(1+e) quantized by
// eb_exponential_quantize(a, 2, log2(2), DBL_EPSILON). Code 0: a <= eps (zero
// precision); code c >= 1: id c - 1 with precision 2^(c-1) eps.
int code2l_ref(double e) {
  const double thr = std::numeric_limits<double>::epsilon();
  const double a = std::log2(1 + e);
  if (a <= thr) return 0;
  return (int)(std::log2(a / thr) / std::log2(2.0)) + 1;
}

double next_down(double x) { return std::nextafter(x, 0.0); }
double next_up(double x) { return std::nextafter(x, INFINITY); }

// Smallest positive double x in (lo, hi] with f(x) >= k, assuming f is
// nondecreasing; then checks monotonicity in a window around the threshold.
template <typename F>
double threshold(F f, int k, double lo, double hi) {
  uint64_t a, b;
  std::memcpy(&a, &lo, 8);
  std::memcpy(&b, &hi, 8);
  // invariant: f(bits a) < k, f(bits b) >= k
  while (b - a > 1) {
    uint64_t m = a + (b - a) / 2;
    double x;
    std::memcpy(&x, &m, 8);
    if (f(x) >= k) b = m;
    else a = m;
  }
  double t;
  std::memcpy(&t, &b, 8);
  double x = t;
  for (int i = 0; i < 2048; i++, x = next_up(x))
    CPSZ_REQUIRE(f(x) >= k, kErrInternal, "libm log2 is not monotone near an eb threshold");
  x = next_down(t);
  for (int i = 0; i < 2048 && x > lo; i++, x = next_down(x))
    CPSZ_REQUIRE(f(x) < k, kErrInternal, "libm log2 is not monotone near an eb threshold");
  return t;
}

EbTables build_eb_tables() {
  EbTables t;
  const double eps = std::numeric_limits<double>::epsilon();
  t.q2d.push_back(0);
  // Up to k = 511: beyond 4^511 eps = 2^970 the reference's x/eps overflows
  // for x >= 2^972 (the device path rejects such bounds).
  for (int k = 1; k <= 511; k++) {
    const double c = std::ldexp(eps, 2 * k);  // 4^k eps
    t.q2d.push_back(threshold(id2d_ref, k, c / 2, c));
  }
  t.q2d.push_back(INFINITY);
  t.q3d.push_back(0);
  for (int k = 1; k < 64; k++) {
    // id3 >= k once log2(1+e) >= 2^(k-23): e near 2^(2^(k-23)) - 1
    const double a = std::ldexp(1.0, k - 23);
    if (a > 1000) break;
    const double guess = std::exp2(a) - 1;
    double lo = guess * 0.5, hi = guess * 2;
    if (!std::isfinite(hi)) break;
    if (id3d_ref(lo) >= k || id3d_ref(hi) < k) fail(kErrInternal, "3D eb threshold bracket failed");
    t.q3d.push_back(threshold(id3d_ref, k, lo, hi));
  }
  t.q3d.push_back(INFINITY);
  t.q2l.push_back(0);
  // code >= 1 from the first e with 1 + e > 1 (e just above 2^-53)
  t.q2l.push_back(threshold(code2l_ref, 1, 0x1p-60, 0x1p-50));
  for (int k = 2; k < 80; k++) {
    // code >= k once log2(1+e) >= 2^(k-1) eps: e near 2^(2^(k-1) eps) - 1
    const double a = std::ldexp(std::numeric_limits<double>::epsilon(), k - 1);
    if (a > 1000) break;
    const double guess = a < 1e-3 ? a * 0.6931471805599453 : std::exp2(a) - 1;
    const double lo = guess * 0.5, hi = guess * 2;
    if (!std::isfinite(hi)) break;
    if (code2l_ref(lo) >= k || code2l_ref(hi) < k) fail(kErrInternal, "2D log eb threshold bracket failed");
    t.q2l.push_back(threshold(code2l_ref, k, lo, hi));
  }
  t.q2l.push_back(INFINITY);
  return t;
}

}  // namespace

const EbTables& eb_tables() {
  static EbTables t = build_eb_tables();
  return t;
}

Ctx::Ctx(int device) : device_(device) {
  CPSZ_HIP(hipSetDevice(device));
  CPSZ_HIP(hipStreamCreateWithFlags(&own_stream_, hipStreamNonBlocking));
  stream_ = own_stream_;
}

Ctx::~Ctx() {
  hipSetDevice(device_);
  for (auto& kv : bufs_) hipFree(kv.second.p);
  for (auto& kv : hbufs_) hipHostFree(kv.second.p);
  if (d_q2d_) hipFree(d_q2d_);
  if (d_q3d_) hipFree(d_q3d_);
  if (d_q2l_) hipFree(d_q2l_);
  for (void* p : libm_allocs_) hipFree(p);
  for (auto e : event_pool_) hipEventDestroy(e);
  for (auto& kv : table_sent_) hipEventDestroy(kv.second);
  if (point_) hipEventDestroy(point_);
  if (own_stream_) hipStreamDestroy(own_stream_);
}

void* Ctx::host_table(const char* name, size_t bytes) {
  auto it = table_sent_.find(name);
  // (a capture copies the table's bytes instead of sending them, and the
  // stream was idle when it began)
  if (it != table_sent_.end() && !capturing) CPSZ_HIP(hipEventSynchronize(it->second));
  return host_buf(name, bytes);
}

void Ctx::upload_table(const char* name, void* dst, const void* src, size_t bytes) {
  if (!bytes) return;
  if (capturing) {
    const size_t off = upload_bytes.size();
    upload_bytes.insert(upload_bytes.end(), (const uint8_t*)src, (const uint8_t*)src + bytes);
    uploads.push_back(Upload{dst, off, bytes});
    return;
  }
  CPSZ_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream_));
  hipEvent_t& e = table_sent_[name];
  if (!e) CPSZ_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  CPSZ_HIP(hipEventRecord(e, stream_));
}

hipEvent_t Ctx::record_point() {
  if (!point_) CPSZ_HIP(hipEventCreateWithFlags(&point_, hipEventDisableTiming));
  CPSZ_HIP(hipEventRecord(point_, stream_));
  return point_;
}

void* Ctx::raw_buf(const char* name, size_t bytes) {
  DevBuf& b = bufs_[name];
  if (bytes == 0) bytes = 1;
  if (b.cap < bytes) {
    if (b.p) {
      CPSZ_HIP(hipStreamSynchronize(stream_));
      CPSZ_HIP(hipFree(b.p));
      b.p = nullptr;
      b.cap = 0;
      generation_++;
    }
    size_t cap = bytes + bytes / 8 + 4096;
    if (hipMalloc(&b.p, cap) != hipSuccess) {
      b.p = nullptr;
      fail(kErrAlloc, std::string("device allocation failed for ") + name);
    }
    b.cap = cap;
  }
  return b.p;
}

void* Ctx::host_buf(const char* name, size_t bytes) {
  DevBuf& b = hbufs_[name];
  if (bytes == 0) bytes = 1;
  if (b.cap < bytes) {
    if (b.p) {
      CPSZ_HIP(hipStreamSynchronize(stream_));
      CPSZ_HIP(hipHostFree(b.p));
      b.p = nullptr;
      b.cap = 0;
      generation_++;
    }
    size_t cap = bytes + bytes / 8 + 4096;
    if (hipHostMalloc(&b.p, cap, hipHostMallocDefault) != hipSuccess) {
      b.p = nullptr;
      fail(kErrAlloc, std::string("pinned host allocation failed for ") + name);
    }
    b.cap = cap;
  }
  return b.p;
}

const double* Ctx::d_q2d() {
  if (!d_q2d_) {
    const auto& q = eb_tables().q2d;
    CPSZ_HIP(hipMalloc(&d_q2d_, q.size() * sizeof(double)));
    CPSZ_HIP(hipMemcpy(d_q2d_, q.data(), q.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  return d_q2d_;
}

const double* Ctx::d_q2l() {
  if (!d_q2l_) {
    const auto& q = eb_tables().q2l;
    CPSZ_HIP(hipMalloc(&d_q2l_, q.size() * sizeof(double)));
    CPSZ_HIP(hipMemcpy(d_q2l_, q.data(), q.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  return d_q2l_;
}

const double* Ctx::d_q3d() {
  if (!d_q3d_) {
    const auto& q = eb_tables().q3d;
    CPSZ_HIP(hipMalloc(&d_q3d_, q.size() * sizeof(double)));
    CPSZ_HIP(hipMemcpy(d_q3d_, q.data(), q.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  return d_q3d_;
}

const Ctx::LibmDev& Ctx::libm_dev() {
  if (!libm_ready_) {
    const LibmTables& t = libm_tables();
    auto up = [&](const void* src, size_t bytes) {
      void* p = nullptr;
      CPSZ_HIP(hipMalloc(&p, bytes ? bytes : 4));
      libm_allocs_.push_back(p);
      if (bytes) CPSZ_HIP(hipMemcpy(p, src, bytes, hipMemcpyHostToDevice));
      return p;
    };
    auto view = [&](const LibmTable& x) {
      LibmView v;
      v.band = (const float*)up(x.band.data(), x.band.size() * 4);
      v.kind = (const uint8_t*)up(x.kind.data(), x.kind.size());
      v.off = (const uint32_t*)up(x.off.data(), x.off.size() * 4);
      v.cnt = (const uint32_t*)up(x.cnt.data(), x.cnt.size() * 4);
      v.list_x = (const uint32_t*)up(x.list_x.data(), x.list_x.size() * 4);
      v.list_y = (const uint32_t*)up(x.list_y.data(), x.list_y.size() * 4);
      v.bitmap = (const uint32_t*)up(x.bitmap.data(), x.bitmap.size() * 4);
      return v;
    };
    libm_.log2f = view(t.log2f);
    libm_.exp2f = view(t.exp2f);
    libm_ready_ = true;
  }
  return libm_;
}

void Ctx::mark(const char* name) {
  if (!timing || capturing) return;
  hipEvent_t e;
  if (!event_pool_.empty()) {
    e = event_pool_.back();
    event_pool_.pop_back();
  } else {
    CPSZ_HIP(hipEventCreate(&e));
  }
  CPSZ_HIP(hipEventRecord(e, stream_));
  marks_.push_back({name, e});
}

void Ctx::finish_marks() {
  stages.clear();
  if (!timing || marks_.empty()) {
    for (auto& m : marks_) event_pool_.push_back(m.second);
    marks_.clear();
    return;
  }
  CPSZ_HIP(hipEventSynchronize(marks_.back().second));
  for (size_t i = 1; i < marks_.size(); i++) {
    float ms = 0;
    CPSZ_HIP(hipEventElapsedTime(&ms, marks_[i - 1].second, marks_[i].second));
    stages.push_back({marks_[i].first, ms});
  }
  for (auto& m : marks_) event_pool_.push_back(m.second);
  marks_.clear();
}

Ctx& ctx_for_device(int device) {
  static std::mutex mu;
  static std::map<int, std::unique_ptr<Ctx>> ctxs;
  std::lock_guard<std::mutex> g(mu);
  auto& p = ctxs[device];
  if (!p) p.reset(new Ctx(device));
  return *p;
}

}  // namespace cpsz
