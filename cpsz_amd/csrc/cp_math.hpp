// Exact 2D and 3D critical-point predicates and types, shared by the GPU
// judge, the learned-bound labelers and their host emulations.
//
// A triangle of the reference's triangulation (test/cp_extraction_2d.cpp:62-147)
// holds a critical point when the zero vector lies in the closed triangle of
// its three vertex vectors: the barycentric numerators cross(b,c), cross(c,a),
// cross(a,b) all have the sign of their sum and the sum is non-zero. Products
// of two floats are exact in double; the sum of six of them is evaluated as an
// error-free expansion (TwoSum, Shewchuk 1997). Triangles with a zero vertex
// vector are skipped as in the reference. The type (codes of
// test/cp_extraction_2d.cpp:32-39) comes from the signs of the eigenvalues of
// the linear interpolant's Jacobian, read off its trace and determinant with
// +, -, *, / only, so host and device classify identically.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace cpsz {

struct Expansion {
  double t[56];
  int n = 0;
  __host__ __device__ void add(double x) {
    double q = x;
    int m = 0;
    for (int i = 0; i < n; i++) {
      const double s = q + t[i];
      const double bv = s - q;
      const double av = s - bv;
      const double e = (q - av) + (t[i] - bv);
      if (e != 0) t[m++] = e;
      q = s;
    }
    if (q != 0) t[m++] = q;
    n = m;
  }
  __host__ __device__ int sign() const {
    for (int i = n - 1; i >= 0; i--)
      if (t[i] != 0) return t[i] > 0 ? 1 : -1;
    return 0;
  }
};

// sign(a*b - c*d) for floats (both products exact in double)
__host__ __device__ inline int sign_cross(float a, float b, float c, float d) {
  const double x = (double)a * b, y = (double)c * d;
  return (x > y) - (x < y);
}

// Type of the critical point of a triangle with vertex vectors v[p] = (u, v);
// lower = 0: coordinates {(0,0),(0,1),(1,1)}, lower = 1: {(0,0),(1,0),(1,1)}.
__host__ __device__ inline int cp_type_2d(const double v[3][2], int lower) {
  const double X[2][3][2] = {{{0, 0}, {0, 1}, {1, 1}}, {{0, 0}, {1, 0}, {1, 1}}};
  const double(*x)[2] = X[lower];
  const double xa = x[0][0] - x[2][0], xb = x[1][0] - x[2][0];
  const double ya = x[0][1] - x[2][1], yb = x[1][1] - x[2][1];
  const double det = xa * yb - xb * ya;
  const double ix[2][2] = {{yb / det, -xb / det}, {-ya / det, xa / det}};
  const double F[2][2] = {{v[0][0] - v[2][0], v[1][0] - v[2][0]}, {v[0][1] - v[2][1], v[1][1] - v[2][1]}};
  double J[2][2];
  for (int r = 0; r < 2; r++)
    for (int c = 0; c < 2; c++) J[r][c] = F[r][0] * ix[0][c] + F[r][1] * ix[1][c];
  const double tr = J[0][0] + J[1][1];
  const double dt = J[0][0] * J[1][1] - J[1][0] * J[0][1];
  const double delta = tr * tr - 4 * dt;
  if (delta >= 0) {
    if (dt < 0) return 3;               // roots of opposite sign
    if (dt > 0) return tr < 0 ? 1 : 2;  // same sign as their sum
    return tr > 0 ? 2 : 0;              // a zero root: the other is tr
  }
  return tr < 0 ? 4 : (tr > 0 ? 5 : 6);
}

// ------------------------------------------------------------------- 3D --
// sum_i s_i * a_i * b_i * c_i (floats, s_i = +-1): its sign, filtered by a
// forward error bound, else exact (FMA-based TwoProduct into an expansion).
struct Triple {
  float a, b, c;
  double s;
};

__host__ __device__ inline void add_triple(Expansion& e, double s, float a, float b, float c) {
  const double ab = (double)a * b;  // exact
  const double p = ab * (double)c;
  const double q = fma(ab, (double)c, -p);
  e.add(s * p);
  if (q != 0) e.add(s * q);
}

__host__ __device__ inline int sign_triples(const Triple* t, int m) {
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

// Type code (test/cp_extraction_3d.cpp:30-38) of the critical point of a
// tetrahedron with vertex vectors v[p] at coordinates X[p].
__host__ __device__ inline int type_3d(const double v[4][3], const int X[4][3]) {
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

// The six tetrahedra of a cell in the judge's order
// (test/cp_extraction_3d.cpp:127-209): vertex 0 is the cell origin, vertex 3
// the opposite corner (1,1,1), vertices 1 and 2 the corners ca(t), cb(t),
// coded di << 2 | dj << 1 | dk; the simplex id is 6 * origin + sid(t).
__host__ __device__ inline int judge_tet_a(int t) { return (int)((0x121424u >> (4 * t)) & 7u); }
__host__ __device__ inline int judge_tet_b(int t) { return (int)((0x335566u >> (4 * t)) & 7u); }
__host__ __device__ inline int judge_tet_sid(int t) { return (int)((0x534120u >> (4 * t)) & 7u); }

// -1 when the tetrahedron (vertex vectors f[p] = (u, v, w), vertices 1 and 2
// at corners ca, cb) holds no critical point, else its type code.
__host__ __device__ inline int tet_cp(const float f[4][3], int ca, int cb) {
  for (int p = 0; p < 4; p++)
    if (f[p][0] == 0 && f[p][1] == 0 && f[p][2] == 0) return -1;
  // numerators: D_p = (-1)^(p+1) det(other three vectors), D = sum
  int sp[4];
  Triple all[24];
  for (int p = 0; p < 4; p++) {
    const float* a = f[p == 0 ? 1 : 0];
    const float* b = f[p <= 1 ? 2 : 1];
    const float* c = f[p <= 2 ? 3 : 2];
    const double sg = (p & 1) ? 1.0 : -1.0;
    Triple* tr = all + 6 * p;
    tr[0] = Triple{a[0], b[1], c[2], sg};
    tr[1] = Triple{a[0], b[2], c[1], -sg};
    tr[2] = Triple{a[1], b[0], c[2], -sg};
    tr[3] = Triple{a[1], b[2], c[0], sg};
    tr[4] = Triple{a[2], b[0], c[1], sg};
    tr[5] = Triple{a[2], b[1], c[0], -sg};
    sp[p] = sign_triples(tr, 6);
  }
  for (int p = 0; p < 4; p++)
    for (int q = p + 1; q < 4; q++)
      if (sp[p] * sp[q] < 0) return -1;
  const int sd = sign_triples(all, 24);
  if (sd == 0) return -1;
  for (int p = 0; p < 4; p++)
    if (sp[p] * sd < 0) return -1;
  // coordinates (x, y, z) = (dk, dj, di)
  const int X[4][3] = {{0, 0, 0}, {ca & 1, (ca >> 1) & 1, ca >> 2}, {cb & 1, (cb >> 1) & 1, cb >> 2}, {1, 1, 1}};
  double v[4][3];
  for (int p = 0; p < 4; p++)
    for (int q = 0; q < 3; q++) v[p][q] = f[p][q];
  return type_3d(v, X);
}

// -1 when the triangle (vertex vectors f[p] = (u, v)) holds no critical point,
// else its type code.
__host__ __device__ inline int tri_cp(const float f[3][2], int lower) {
  for (int p = 0; p < 3; p++)
    if (f[p][0] == 0 && f[p][1] == 0) return -1;
  const int s0 = sign_cross(f[1][0], f[2][1], f[2][0], f[1][1]);
  const int s1 = sign_cross(f[2][0], f[0][1], f[0][0], f[2][1]);
  const int s2 = sign_cross(f[0][0], f[1][1], f[1][0], f[0][1]);
  if (s0 * s1 < 0 || s0 * s2 < 0 || s1 * s2 < 0) return -1;  // mixed signs: zero outside
  Expansion e;
  e.add((double)f[1][0] * f[2][1]);
  e.add(-((double)f[2][0] * f[1][1]));
  e.add((double)f[2][0] * f[0][1]);
  e.add(-((double)f[0][0] * f[2][1]));
  e.add((double)f[0][0] * f[1][1]);
  e.add(-((double)f[1][0] * f[0][1]));
  const int s = e.sign();
  if (s == 0 || s0 * s < 0 || s1 * s < 0 || s2 * s < 0) return -1;
  double v[3][2];
  for (int p = 0; p < 3; p++) v[p][0] = f[p][0], v[p][1] = f[p][1];
  return cp_type_2d(v, lower);
}

// ------------------------------------------------ Simulation of Simplicity --
// (Edelsbrunner & Muecke, "Simulation of Simplicity", ACM TOG 9(1), 1990.)
// Every vertex vector is perturbed symbolically: component c of the vertex of
// global index g moves by eps^(2^(d*rank + c)), where rank orders the vertices
// of a determinant by global index. A determinant of perturbed rows is a
// polynomial in eps whose monomials all have distinct exponents, so its sign
// is the sign of the first nonzero coefficient in increasing exponent order:
// the plain determinant, then cofactors, entries and finally +-1, which makes
// the sign never 0. The zero vector then never lies on a simplex face, and a
// zero of the piecewise-linear field lies in exactly one simplex of the mesh,
// decided consistently by the simplices that share a face (they evaluate the
// same determinant). Zero vertex vectors are perturbed like any other.

// det [a; b] with a of lower global index than b: the sequence det, b_v,
// -b_u, -a_v, -1 (monomials eps_a0, eps_a1, eps_b0, eps_a1 * eps_b0).
__host__ __device__ inline int sos_det2(float au, float av, float bu, float bv) {
  const int s = sign_cross(au, bv, av, bu);
  if (s) return s;
  if (bv != 0) return bv > 0 ? 1 : -1;
  if (bu != 0) return bu > 0 ? -1 : 1;
  if (av != 0) return av > 0 ? -1 : 1;
  return -1;
}

// det [p; q] under the perturbation, for vertices of global indices ip != iq
__host__ __device__ inline int sos_cross(const float* p, const float* q, long long ip, long long iq) {
  return ip < iq ? sos_det2(p[0], p[1], q[0], q[1]) : -sos_det2(q[0], q[1], p[0], p[1]);
}

// det [r0; r1; r2] with rows in increasing global index: the coefficient
// sequence in increasing exponent order (exponent sum of 2^(3 rank + c) over
// the perturbed entries): kind 0 the determinant, 1 the cofactor of (r, c),
// 2 the remaining entry (r, c) times the permutation sign s, 3 the constant s.
struct SosTerm3 {
  signed char kind, r, c, s;
};
constexpr SosTerm3 kSos3[18] = {{0, 0, 0, 1},  {1, 0, 0, 1},  {1, 0, 1, -1}, {1, 0, 2, 1},  {1, 1, 0, -1},
                                {2, 2, 2, -1}, {2, 2, 1, 1},  {1, 1, 1, 1},  {2, 2, 2, 1},  {2, 2, 0, -1},
                                {1, 1, 2, -1}, {2, 2, 1, -1}, {2, 2, 0, 1},  {1, 2, 0, 1},  {2, 1, 2, 1},
                                {2, 1, 1, -1}, {2, 0, 2, -1}, {3, 0, 0, -1}};

__host__ __device__ inline int sos_det3_sorted(const float* m0, const float* m1, const float* m2) {
  const float* M[3] = {m0, m1, m2};
  for (int k = 0; k < 18; k++) {
    const SosTerm3 t = kSos3[k];
    int sg = 0;
    if (t.kind == 0) {
      const Triple tr[6] = {{m0[0], m1[1], m2[2], 1}, {m0[0], m1[2], m2[1], -1}, {m0[1], m1[0], m2[2], -1},
                            {m0[1], m1[2], m2[0], 1}, {m0[2], m1[0], m2[1], 1},  {m0[2], m1[1], m2[0], -1}};
      sg = sign_triples(tr, 6);
    } else if (t.kind == 1) {
      const int ra = t.r == 0 ? 1 : 0, rb = t.r == 2 ? 1 : 2;
      const int ca = t.c == 0 ? 1 : 0, cb = t.c == 2 ? 1 : 2;
      sg = t.s * sign_cross(M[ra][ca], M[rb][cb], M[ra][cb], M[rb][ca]);
    } else if (t.kind == 2) {
      const float x = M[t.r][t.c];
      sg = x > 0 ? t.s : (x < 0 ? -t.s : 0);
    } else {
      sg = t.s;
    }
    if (sg) return sg;
  }
  return -1;  // not reached: the last term is a constant
}

// det [a; b; c] under the perturbation, vertices of distinct global indices
__host__ __device__ inline int sos_det3(const float* a, const float* b, const float* c, long long ia, long long ib,
                                        long long ic) {
  // sort by index, tracking the permutation's parity
  const float* r[3] = {a, b, c};
  long long g[3] = {ia, ib, ic};
  int par = 1;
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2 - i; j++)
      if (g[j] > g[j + 1]) {
        const long long tg = g[j];
        g[j] = g[j + 1];
        g[j + 1] = tg;
        const float* tr = r[j];
        r[j] = r[j + 1];
        r[j + 1] = tr;
        par = -par;
      }
  return par * sos_det3_sorted(r[0], r[1], r[2]);
}

// tri_cp under Simulation of Simplicity: the zero of the perturbed field lies
// inside the triangle (vertex global indices idx) iff its three barycentric
// numerators have one sign; never on an edge or a vertex.
__host__ __device__ inline int tri_cp_sos(const float f[3][2], const long long idx[3], int lower) {
  const int s0 = sos_cross(f[1], f[2], idx[1], idx[2]);
  const int s1 = sos_cross(f[2], f[0], idx[2], idx[0]);
  const int s2 = sos_cross(f[0], f[1], idx[0], idx[1]);
  if (s0 != s1 || s0 != s2) return -1;
  double v[3][2];
  for (int p = 0; p < 3; p++) v[p][0] = f[p][0], v[p][1] = f[p][1];
  return cp_type_2d(v, lower);
}

// tet_cp under Simulation of Simplicity (vertex global indices idx).
__host__ __device__ inline int tet_cp_sos(const float f[4][3], const long long idx[4], int ca, int cb) {
  int sp[4];
  for (int p = 0; p < 4; p++) {
    const int a = p == 0 ? 1 : 0, b = p <= 1 ? 2 : 1, c = p <= 2 ? 3 : 2;
    const int sg = (p & 1) ? 1 : -1;  // D_p = (-1)^(p+1) det(other three)
    sp[p] = sg * sos_det3(f[a], f[b], f[c], idx[a], idx[b], idx[c]);
  }
  if (sp[0] != sp[1] || sp[0] != sp[2] || sp[0] != sp[3]) return -1;
  const int X[4][3] = {{0, 0, 0}, {ca & 1, (ca >> 1) & 1, ca >> 2}, {cb & 1, (cb >> 1) & 1, cb >> 2}, {1, 1, 1}};
  double v[4][3];
  for (int p = 0; p < 4; p++)
    for (int q = 0; q < 3; q++) v[p][q] = f[p][q];
  return type_3d(v, X);
}

}  // namespace cpsz
