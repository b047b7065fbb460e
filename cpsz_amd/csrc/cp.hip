// Critical-point extraction on the GPU: the "0 CP errors" judge.
//
// Simplices, ids and type codes follow test/cp_extraction_2d.cpp:62-147 and
// test/cp_extraction_3d.cpp:79-209. The reference decides whether the zero
// vector lies in a simplex with FTK's floating-point inverse interpolation;
// here it is decided exactly: the barycentric numerators (2x2 / 3x3
// determinants of float data) and their sum are evaluated in double with a
// forward error bound, and with error-free expansions (TwoSum / FMA-based
// TwoProduct, Shewchuk 1997) whenever the bound does not settle the sign.
// A simplex holds a critical point when every numerator has the sign of
// their sum (closed simplex) and the sum is non-zero; simplices with a zero
// vertex vector are skipped as in the reference. The SoS form
// (Simulation of Simplicity, cp_math.hpp) perturbs the vertex vectors
// symbolically instead, so each zero lies in exactly one simplex. Types come from the signs of
// the eigenvalues of the simplex Jacobian, read off the characteristic
// polynomial's coefficients with +, -, *, / only (no sqrt or transcendental
// functions), so host and device classify identically: 3D uses Cardano's
// discriminant for the root structure, then p(0) and b c - d for one real
// root plus a complex pair, Descartes' rule of signs for three real roots.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <numeric>
#include <vector>

#include "cp.hpp"
#include "cp_math.hpp"

namespace cpsz {

namespace {

struct CpOut {
  unsigned long long* count;
  long long* ids;
  int* types;
  unsigned long long cap;
  __device__ void push(long long id, int type) const {
    const unsigned long long k = atomicAdd(count, 1ull);
    if (k < cap) {
      ids[k] = id;
      types[k] = type;
    }
  }
};

__global__ void k_cp2d(const float* __restrict__ U, const float* __restrict__ V, int r1, int r2, CpOut out,
                       int sos) {
  const long long cell = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long nci = r1 - 3, ncj = r2 - 3;  // i in [1, r1-3], j in [1, r2-3]
  if (nci <= 0 || ncj <= 0 || cell >= nci * ncj) return;
  const int i = 1 + (int)(cell / ncj), j = 1 + (int)(cell % ncj);
  const size_t idx[2][3] = {{(size_t)i * r2 + j, (size_t)(i + 1) * r2 + j, (size_t)(i + 1) * r2 + j + 1},
                            {(size_t)i * r2 + j, (size_t)i * r2 + j + 1, (size_t)(i + 1) * r2 + j + 1}};
  const long long sid = 2ll * ((long long)i * (r2 - 1) + j);
  for (int t = 0; t < 2; t++) {
    float f[3][2];
    for (int p = 0; p < 3; p++) {
      f[p][0] = U[idx[t][p]];
      f[p][1] = V[idx[t][p]];
    }
    const long long gi[3] = {(long long)idx[t][0], (long long)idx[t][1], (long long)idx[t][2]};
    const int type = sos ? tri_cp_sos(f, gi, t) : tri_cp(f, t);
    if (type >= 0) out.push(sid + t, type);
  }
}

__global__ void k_cp3d(const float* __restrict__ U, const float* __restrict__ V,
                       const float* __restrict__ W, int r1, int r2, int r3, CpOut out, int sos) {
  const long long cell = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long nk = r3 - 1, nj = r2 - 1, ni = r1 - 1;
  if (ni <= 0 || nj <= 0 || nk <= 0 || cell >= ni * nj * nk) return;
  const int k = (int)(cell % nk), j = (int)((cell / nk) % nj), i = (int)(cell / (nk * nj));
  const size_t d0 = (size_t)r2 * r3, d1 = r3;
  const size_t base = i * d0 + j * d1 + k;
  auto at = [&](int c) { return base + (c >> 2) * d0 + ((c >> 1) & 1) * d1 + (c & 1); };
  for (int t = 0; t < 6; t++) {
    const int ca = judge_tet_a(t), cb = judge_tet_b(t);
    const size_t vi[4] = {base, at(ca), at(cb), base + d0 + d1 + 1};
    float f[4][3];
    for (int p = 0; p < 4; p++) {
      f[p][0] = U[vi[p]];
      f[p][1] = V[vi[p]];
      f[p][2] = W[vi[p]];
    }
    const long long gi[4] = {(long long)vi[0], (long long)vi[1], (long long)vi[2], (long long)vi[3]};
    const int type = sos ? tet_cp_sos(f, gi, ca, cb) : tet_cp(f, ca, cb);
    if (type >= 0) out.push(6ll * (long long)base + judge_tet_sid(t), type);
  }
}

void collect(Ctx& ctx, CpOut& o, unsigned long long* dcount, int64_t* ids, int32_t* types, size_t cap,
             size_t* count) {
  unsigned long long n = 0;
  CPSZ_HIP(hipMemcpyAsync(&n, dcount, 8, hipMemcpyDeviceToHost, ctx.stream()));
  CPSZ_HIP(hipStreamSynchronize(ctx.stream()));
  *count = (size_t)n;
  const size_t m = std::min<size_t>((size_t)n, (size_t)o.cap);
  std::vector<long long> hid(m);
  std::vector<int> ht(m);
  if (m) {
    CPSZ_HIP(hipMemcpyAsync(hid.data(), o.ids, m * 8, hipMemcpyDeviceToHost, ctx.stream()));
    CPSZ_HIP(hipMemcpyAsync(ht.data(), o.types, m * 4, hipMemcpyDeviceToHost, ctx.stream()));
    CPSZ_HIP(hipStreamSynchronize(ctx.stream()));
  }
  CPSZ_REQUIRE(n <= o.cap, kErrAlloc, "critical point buffer overflow");
  std::vector<size_t> order(m);
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return hid[a] < hid[b]; });
  for (size_t k = 0; k < m && k < cap; k++) {
    ids[k] = hid[order[k]];
    types[k] = ht[order[k]];
  }
}

CpOut make_out(Ctx& ctx, size_t cap_dev, unsigned long long*& dcount) {
  CpOut o;
  dcount = ctx.buf<unsigned long long>("cp.count", 1);
  CPSZ_HIP(hipMemsetAsync(dcount, 0, 8, ctx.stream()));
  o.count = dcount;
  o.ids = ctx.buf<long long>("cp.ids", cap_dev);
  o.types = ctx.buf<int>("cp.types", cap_dev);
  o.cap = cap_dev;
  return o;
}

// ---- the no-allocation form of the 2D judge ---------------------------------
// Its scratch (the counter and the unsorted (id, type) list) is this
// file-scope device array: every GPU the library runs on holds its own copy
// (12 MiB + 256 B) until the process ends. Host code reaches a device's copy
// only through the address hipGetSymbolAddress returns on that device.
__device__ __align__(256) uint8_t cp_scratch[256 + 12 * (size_t)kNoAllocMaxCp];

__global__ void k_zero_count(unsigned long long* c) { *c = 0; }

constexpr int kMaxDevices = 64;
std::mutex g_scratch_m;
uint8_t* g_scratch[kMaxDevices] = {};

// Address of this device's scratch (looked up once per device, with it current).
uint8_t* scratch_of(int device) {
  CPSZ_REQUIRE(device >= 0 && device < kMaxDevices, kErrArg, "device index out of range");
  std::lock_guard<std::mutex> g(g_scratch_m);
  if (!g_scratch[device]) {
    void* p = nullptr;
    CPSZ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(cp_scratch)));
    CPSZ_REQUIRE(p, kErrHip, "no device address for the critical-point scratch");
    g_scratch[device] = (uint8_t*)p;
  }
  return g_scratch[device];
}

}  // namespace

void critical_points_2d_noalloc(int device, hipStream_t st, const float* dU, const float* dV, size_t r1, size_t r2,
                                int64_t* ids, int32_t* types, size_t cap, size_t* count) {
  CPSZ_REQUIRE(dU && dV && ids && types && count, kErrArg, "critical_points_2d_noalloc: bad arguments");
  CPSZ_REQUIRE(r1 < (1u << 28) && r2 < (1u << 28), kErrUnsupported,
               "critical_points_2d_noalloc: field dimensions must stay below 2^28");
  int ndev = 0;
  CPSZ_HIP(hipGetDeviceCount(&ndev));
  CPSZ_REQUIRE(device >= 0 && device < ndev, kErrArg, "device index out of range");
  CPSZ_HIP(hipSetDevice(device));
  uint8_t* base = scratch_of(device);
  CpOut o;
  o.count = (unsigned long long*)base;
  o.ids = (long long*)(base + 256);
  o.types = (int*)(base + 256 + 8 * (size_t)kNoAllocMaxCp);
  o.cap = kNoAllocMaxCp;
  k_zero_count<<<1, 1, 0, st>>>(o.count);
  const size_t cells = (r1 > 3 && r2 > 3) ? (r1 - 3) * (r2 - 3) : 0;
  if (cells) k_cp2d<<<(int)div_up(cells, 256), 256, 0, st>>>(dU, dV, (int)r1, (int)r2, o, 0);
  CPSZ_HIP(hipGetLastError());
  unsigned long long n = 0;
  CPSZ_HIP(hipMemcpyAsync(&n, o.count, 8, hipMemcpyDeviceToHost, st));
  CPSZ_HIP(hipStreamSynchronize(st));
  *count = (size_t)n;
  CPSZ_REQUIRE(n <= (unsigned long long)kNoAllocMaxCp, kErrUnsupported,
               "critical_points_2d_noalloc: more critical points than its scratch holds (2^20)");
  CPSZ_REQUIRE(n <= cap, kErrArg, "critical_points_2d_noalloc: output arrays hold fewer entries than *count");
  if (!n) return;
  // into the caller's arrays, then sorted there by id: the type rides in the
  // low 4 bits of the id word, so one in-place sort of one array suffices
  CPSZ_HIP(hipMemcpyAsync(ids, o.ids, 8 * n, hipMemcpyDeviceToHost, st));
  CPSZ_HIP(hipMemcpyAsync(types, o.types, 4 * n, hipMemcpyDeviceToHost, st));
  CPSZ_HIP(hipStreamSynchronize(st));
  for (size_t k = 0; k < n; k++) ids[k] = ids[k] << 4 | (int64_t)(types[k] & 15);
  std::sort(ids, ids + n);
  for (size_t k = 0; k < n; k++) {
    types[k] = (int32_t)(ids[k] & 15);
    ids[k] >>= 4;
  }
}

void critical_points_2d(Ctx& ctx, const float* dU, const float* dV, size_t r1, size_t r2,
                        int64_t* ids, int32_t* types, size_t cap, size_t* count, bool sos) {
  CPSZ_REQUIRE(dU && dV && ids && types && count, kErrArg, "critical_points_2d: bad arguments");
  const size_t cells = (r1 > 3 && r2 > 3) ? (r1 - 3) * (r2 - 3) : 0;
  unsigned long long* dc;
  CpOut o = make_out(ctx, std::max<size_t>(std::min<size_t>(2 * cells, 1u << 24), 1), dc);
  if (cells) {
    k_cp2d<<<(int)div_up(cells, 256), 256, 0, ctx.stream()>>>(dU, dV, (int)r1, (int)r2, o, sos ? 1 : 0);
    CPSZ_HIP(hipGetLastError());
  }
  collect(ctx, o, dc, ids, types, cap, count);
}

void critical_points_3d(Ctx& ctx, const float* dU, const float* dV, const float* dW, size_t r1,
                        size_t r2, size_t r3, int64_t* ids, int32_t* types, size_t cap,
                        size_t* count, bool sos) {
  CPSZ_REQUIRE(dU && dV && dW && ids && types && count, kErrArg, "critical_points_3d: bad arguments");
  const size_t cells = (r1 > 1 && r2 > 1 && r3 > 1) ? (r1 - 1) * (r2 - 1) * (r3 - 1) : 0;
  unsigned long long* dc;
  CpOut o = make_out(ctx, std::max<size_t>(std::min<size_t>(6 * cells, 1u << 24), 1), dc);
  if (cells) {
    k_cp3d<<<(int)div_up(cells, 256), 256, 0, ctx.stream()>>>(dU, dV, dW, (int)r1, (int)r2, (int)r3, o,
                                                              sos ? 1 : 0);
    CPSZ_HIP(hipGetLastError());
  }
  collect(ctx, o, dc, ids, types, cap, count);
}

}  // namespace cpsz
