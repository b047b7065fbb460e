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
// vertex vector are skipped as in the reference. Types come from the signs of
// the eigenvalues of the simplex Jacobian, read off the characteristic
// polynomial's coefficients with +, -, *, / only (no sqrt or transcendental
// functions), so host and device classify identically: 3D uses Cardano's
// discriminant for the root structure, then p(0) and b c - d for one real
// root plus a complex pair, Descartes' rule of signs for three real roots.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "cp.hpp"
#include "cp_math.hpp"

namespace cpsz {

namespace {

__device__ __forceinline__ void add_triple(Expansion& e, double s, float a, float b, float c) {
  const double ab = (double)a * b;
  const double p = ab * (double)c;
  const double q = fma(ab, (double)c, -p);
  e.add(s * p);
  if (q != 0) e.add(s * q);
}

// sign of a sum of signed triple products: filtered, exact fallback
struct Triple {
  float a, b, c;
  double s;
};

__device__ int sign_triples(const Triple* t, int m) {
  double sum = 0, mag = 0;
  for (int i = 0; i < m; i++) {
    const double p = t[i].s * ((double)t[i].a * t[i].b) * t[i].c;
    sum += p;
    mag += fabs(p);
  }
  // each product carries <= 1 rounding (2 ulp), the sum <= m roundings
  const double bound = (double)(m + 4) * 2.3e-16 * mag;
  if (sum > bound) return 1;
  if (sum < -bound) return -1;
  Expansion e;
  for (int i = 0; i < m; i++) add_triple(e, t[i].s, t[i].a, t[i].b, t[i].c);
  return e.sign();
}

__device__ int type_3d(const double v[4][3], const int X[4][3]) {
  double Xm[3][3], Um[3][3];
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) {
      Xm[r][c] = X[c][r] - X[3][r];
      Um[r][c] = v[c][r] - v[3][r];
    }
  const double d = Xm[0][0] * (Xm[1][1] * Xm[2][2] - Xm[1][2] * Xm[2][1]) -
                   Xm[0][1] * (Xm[1][0] * Xm[2][2] - Xm[1][2] * Xm[2][0]) +
                   Xm[0][2] * (Xm[1][0] * Xm[2][1] - Xm[1][1] * Xm[2][0]);
  double inv[3][3];
  inv[0][0] = (Xm[1][1] * Xm[2][2] - Xm[1][2] * Xm[2][1]) / d;
  inv[0][1] = (Xm[0][2] * Xm[2][1] - Xm[0][1] * Xm[2][2]) / d;
  inv[0][2] = (Xm[0][1] * Xm[1][2] - Xm[0][2] * Xm[1][1]) / d;
  inv[1][0] = (Xm[1][2] * Xm[2][0] - Xm[1][0] * Xm[2][2]) / d;
  inv[1][1] = (Xm[0][0] * Xm[2][2] - Xm[0][2] * Xm[2][0]) / d;
  inv[1][2] = (Xm[0][2] * Xm[1][0] - Xm[0][0] * Xm[1][2]) / d;
  inv[2][0] = (Xm[1][0] * Xm[2][1] - Xm[1][1] * Xm[2][0]) / d;
  inv[2][1] = (Xm[0][1] * Xm[2][0] - Xm[0][0] * Xm[2][1]) / d;
  inv[2][2] = (Xm[0][0] * Xm[1][1] - Xm[0][1] * Xm[1][0]) / d;
  double J[3][3];
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) J[r][c] = inv[r][0] * Um[0][c] + inv[r][1] * Um[1][c] + inv[r][2] * Um[2][c];
  const double tr = J[0][0] + J[1][1] + J[2][2];
  const double m2 = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) + (J[0][0] * J[2][2] - J[0][2] * J[2][0]) +
                    (J[1][1] * J[2][2] - J[1][2] * J[2][1]);
  const double dj = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                    J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                    J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  const double b = -tr, c = m2, dd = -dj;
  const double q = (3 * c - b * b) / 9;
  const double r = (-27 * dd + b * (9 * c - 2 * b * b)) / 54;
  const double disc = q * q * q + r * r;
  if (fabs(disc) < 2.220446049250313e-16) return 0;
  if (disc > 0) {
    // one real root r0 (< 0 iff p(0) = dd > 0) and a pair s +- iw with
    // b c - dd = -2 s ((r0 + s)^2 + w^2)
    const int neg = (dd > 0) + 2 * (b * c - dd > 0);
    return 2 * neg + 2;
  }
  // three real roots: sign changes of p(-x) = -x^3 + b x^2 - c x + dd
  const double co[4] = {-1.0, b, -c, dd};
  int neg = 0;
  double last = -1.0;
  for (int k = 1; k < 4; k++) {
    if (co[k] == 0) continue;
    if ((co[k] > 0) != (last > 0)) neg++;
    last = co[k];
  }
  return 2 * neg + 1;
}

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

__global__ void k_cp2d(const float* __restrict__ U, const float* __restrict__ V, int r1, int r2, CpOut out) {
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
    const int type = tri_cp(f, t);
    if (type >= 0) out.push(sid + t, type);
  }
}

__global__ void k_cp3d(const float* __restrict__ U, const float* __restrict__ V,
                       const float* __restrict__ W, int r1, int r2, int r3, CpOut out) {
  const long long cell = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long nk = r3 - 1, nj = r2 - 1, ni = r1 - 1;
  if (ni <= 0 || nj <= 0 || nk <= 0 || cell >= ni * nj * nk) return;
  const int k = (int)(cell % nk), j = (int)((cell / nk) % nj), i = (int)(cell / (nk * nj));
  const size_t d0 = (size_t)r2 * r3, d1 = r3;
  const size_t base = i * d0 + j * d1 + k;
  // (id offset, vertex 1 and 2 as (di, dj, dk)); vertex 0 = (i,j,k), vertex 3 = (i+1,j+1,k+1)
  const int defs[6][7] = {{0, 1, 0, 0, 1, 1, 0}, {2, 0, 1, 0, 1, 1, 0}, {1, 1, 0, 0, 1, 0, 1},
                          {4, 0, 0, 1, 1, 0, 1}, {3, 0, 1, 0, 0, 1, 1}, {5, 0, 0, 1, 0, 1, 1}};
  for (int t = 0; t < 6; t++) {
    const int* D = defs[t];
    const size_t vi[4] = {base, base + D[1] * d0 + D[2] * d1 + D[3], base + D[4] * d0 + D[5] * d1 + D[6],
                          base + d0 + d1 + 1};
    float f[4][3];
    bool zero = false;
    for (int p = 0; p < 4; p++) {
      f[p][0] = U[vi[p]];
      f[p][1] = V[vi[p]];
      f[p][2] = W[vi[p]];
      zero = zero || (f[p][0] == 0 && f[p][1] == 0 && f[p][2] == 0);
    }
    if (zero) continue;
    // numerators: D_p = (-1)^(p+1) det(other three vectors), D = sum
    const int rows[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
    int sp[4];
    Triple all[24];
    for (int p = 0; p < 4; p++) {
      const float* a = f[rows[p][0]];
      const float* b = f[rows[p][1]];
      const float* c = f[rows[p][2]];
      const double sg = (p & 1) ? 1.0 : -1.0;
      Triple tr[6] = {{a[0], b[1], c[2], sg}, {a[0], b[2], c[1], -sg}, {a[1], b[0], c[2], -sg},
                      {a[1], b[2], c[0], sg},  {a[2], b[0], c[1], sg},  {a[2], b[1], c[0], -sg}};
      sp[p] = sign_triples(tr, 6);
      for (int q = 0; q < 6; q++) all[6 * p + q] = tr[q];
    }
    bool mixed = false;
    for (int p = 0; p < 4; p++)
      for (int q = p + 1; q < 4; q++) mixed = mixed || (sp[p] * sp[q] < 0);
    if (mixed) continue;
    const int sd = sign_triples(all, 24);
    if (sd == 0) continue;
    bool in = true;
    for (int p = 0; p < 4; p++) in = in && (sp[p] * sd >= 0);
    if (!in) continue;
    const int X[4][3] = {{0, 0, 0}, {D[3], D[2], D[1]}, {D[6], D[5], D[4]}, {1, 1, 1}};
    double v[4][3];
    for (int p = 0; p < 4; p++)
      for (int q = 0; q < 3; q++) v[p][q] = f[p][q];
    out.push(6ll * (long long)base + D[0], type_3d(v, X));
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

}  // namespace

void critical_points_2d(Ctx& ctx, const float* dU, const float* dV, size_t r1, size_t r2,
                        int64_t* ids, int32_t* types, size_t cap, size_t* count) {
  CPSZ_REQUIRE(dU && dV && ids && types && count, kErrArg, "critical_points_2d: bad arguments");
  const size_t cells = (r1 > 3 && r2 > 3) ? (r1 - 3) * (r2 - 3) : 0;
  unsigned long long* dc;
  CpOut o = make_out(ctx, std::max<size_t>(std::min<size_t>(2 * cells, 1u << 24), 1), dc);
  if (cells) {
    k_cp2d<<<(int)div_up(cells, 256), 256, 0, ctx.stream()>>>(dU, dV, (int)r1, (int)r2, o);
    CPSZ_HIP(hipGetLastError());
  }
  collect(ctx, o, dc, ids, types, cap, count);
}

void critical_points_3d(Ctx& ctx, const float* dU, const float* dV, const float* dW, size_t r1,
                        size_t r2, size_t r3, int64_t* ids, int32_t* types, size_t cap,
                        size_t* count) {
  CPSZ_REQUIRE(dU && dV && dW && ids && types && count, kErrArg, "critical_points_3d: bad arguments");
  const size_t cells = (r1 > 1 && r2 > 1 && r3 > 1) ? (r1 - 1) * (r2 - 1) * (r3 - 1) : 0;
  unsigned long long* dc;
  CpOut o = make_out(ctx, std::max<size_t>(std::min<size_t>(6 * cells, 1u << 24), 1), dc);
  if (cells) {
    k_cp3d<<<(int)div_up(cells, 256), 256, 0, ctx.stream()>>>(dU, dV, dW, (int)r1, (int)r2, (int)r3, o);
    CPSZ_HIP(hipGetLastError());
  }
  collect(ctx, o, dc, ids, types, cap, count);
}

}  // namespace cpsz
