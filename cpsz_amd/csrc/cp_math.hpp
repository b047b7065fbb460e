// Exact 2D critical-point predicate and type, shared by the GPU judge, the
// learned-bound labeler and their host emulations.
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

}  // namespace cpsz
