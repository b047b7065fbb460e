// Critical-point extraction — TEST INFRASTRUCTURE (see orc.hpp).
//
// Simplex enumeration, ids and type codes follow test/cp_extraction_2d.cpp:62-147
// (2D: interior cells i in [1, r1-3], j in [1, r2-3]; ids 2*(i*(r2-1)+j)(+1);
// types 0..6 per :32-39,83-100) and test/cp_extraction_3d.cpp:79-209 (3D: all
// cells; id 6*(vertex index)+{0,2,1,4,3,5}; types 0..8 per :30-38,79-117).
// The reference decides "zero inside the simplex" with FTK's floating-point
// inverse interpolation (absent here): this restatement decides it with exact
// signs of the barycentric numerators (error-free products and expansion
// sums), closed simplex, so it is parity unpinned against FTK by construction.
#include <algorithm>
#include <cmath>
#include <complex>
#include <limits>
#include <vector>

#include "orc.hpp"

namespace orc {

namespace {

// ---- exact expansion arithmetic (Shewchuk, "Adaptive precision floating-point
// arithmetic and fast robust geometric predicates", 1997) ----
inline void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  double bv = s - a;
  double av = s - bv;
  e = (a - av) + (b - bv);
}
inline void two_prod(double a, double b, double& p, double& e) {
  p = a * b;
  e = std::fma(a, b, -p);
}
struct Expansion {
  double t[64];
  int n = 0;
  void add(double x) {  // grow-expansion
    double q = x;
    int m = 0;
    for (int i = 0; i < n; i++) {
      double s, e;
      two_sum(q, t[i], s, e);
      if (e != 0) t[m++] = e;
      q = s;
    }
    if (q != 0) t[m++] = q;
    n = m;
  }
  int sign() const {  // the largest-magnitude component is last
    for (int i = n - 1; i >= 0; i--)
      if (t[i] != 0) return t[i] > 0 ? 1 : -1;
    return 0;
  }
};

// sign of a*b - c*d for floats (products exact in double)
inline int sign_cross(float a, float b, float c, float d) {
  double x = (double)a * b, y = (double)c * d;
  return (x > y) - (x < y);
}

// add s * (a*b*c) exactly
inline void add_triple(Expansion& e, double s, float a, float b, float c) {
  double ab = (double)a * b;  // exact
  double p, q;
  two_prod(ab, (double)c, p, q);
  e.add(s * p);
  if (q != 0) e.add(s * q);
}

// sign of det [a; b; c] (rows are 3-vectors of floats), exactly
int sign_det3(const float* a, const float* b, const float* c) {
  Expansion e;
  add_triple(e, 1, a[0], b[1], c[2]);
  add_triple(e, -1, a[0], b[2], c[1]);
  add_triple(e, -1, a[1], b[0], c[2]);
  add_triple(e, 1, a[1], b[2], c[0]);
  add_triple(e, 1, a[2], b[0], c[1]);
  add_triple(e, -1, a[2], b[1], c[0]);
  return e.sign();
}

// Type classification uses only +, -, * and / on doubles in a fixed order (no
// sqrt or transcendental functions), so the device extractor, written with the
// same operation sequence, agrees with it bit for bit; eigenvalue signs come
// from the characteristic polynomial's coefficients rather than from roots.

// 2D type: eigenvalues of the Jacobian of the linear interpolant
// (test/cp_extraction_2d.cpp:32-39 codes). Real pair (delta >= 0): saddle when
// det < 0, else the sign of the trace; complex pair: the sign of the trace.
int type_2d(const double v[3][2], const double X[3][2]) {
  // J (X_p - X_2) = v_p - v_2 for p = 0, 1
  const double xa = X[0][0] - X[2][0], xb = X[1][0] - X[2][0];
  const double ya = X[0][1] - X[2][1], yb = X[1][1] - X[2][1];
  const double det = xa * yb - xb * ya;
  const double ix[2][2] = {{yb / det, -xb / det}, {-ya / det, xa / det}};
  double F[2][2] = {{v[0][0] - v[2][0], v[1][0] - v[2][0]}, {v[0][1] - v[2][1], v[1][1] - v[2][1]}};
  double J[2][2];
  for (int r = 0; r < 2; r++)
    for (int c = 0; c < 2; c++) J[r][c] = F[r][0] * ix[0][c] + F[r][1] * ix[1][c];
  const double tr = J[0][0] + J[1][1];
  const double dt = J[0][0] * J[1][1] - J[1][0] * J[0][1];
  const double delta = tr * tr - 4 * dt;
  if (delta >= 0) {
    if (dt < 0) return 3;  // roots of opposite sign
    if (dt > 0) return tr < 0 ? 1 : 2;  // same sign as their sum (tr != 0 here)
    return tr > 0 ? 2 : 0;  // a zero root: the other is tr
  }
  if (tr < 0) return 4;
  if (tr > 0) return 5;
  return 6;
}

// 3D type from the monic characteristic cubic p(x) = x^3 + b x^2 + c x + d
// (test/cp_extraction_3d.cpp:30-38 codes: 2 * #negative real parts, +2 for a
// complex pair, +1 for three real roots). Discriminant of Cardano's reduced
// form decides the root structure. One real root r and a pair s +- iw: r < 0
// iff p(0) = d > 0, and b c - d = -2 s ((r + s)^2 + w^2), so s < 0 iff
// b c - d > 0. Three real roots: Descartes' rule of signs on p(-x), exact when
// every root is real.
int type_3d(const double v[4][3], const int X[4][3]) {
  double Xm[3][3], Um[3][3];
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) {
      Xm[r][c] = X[c][r] - X[3][r];
      Um[r][c] = v[c][r] - v[3][r];
    }
  // J = inv(Xm) * Um (similar to the true Jacobian Um*inv(Xm): same spectrum)
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
  // Cardano: q, r and the discriminant q^3 + r^2 (> 0: one real root and a
  // complex pair; < 0: three real roots).
  const double q = (3 * c - b * b) / 9;
  const double r = (-27 * dd + b * (9 * c - 2 * b * b)) / 54;
  const double disc = q * q * q + r * r;
  if (std::fabs(disc) < std::numeric_limits<double>::epsilon()) return 0;
  int neg;
  if (disc > 0) {
    neg = (dd > 0) + 2 * (b * c - dd > 0);
    return 2 * neg + 2;
  }
  // coefficients of p(-x) = -x^3 + b x^2 - c x + d; zeros are skipped
  const double co[4] = {-1.0, b, -c, dd};
  neg = 0;
  double last = -1.0;
  for (int k = 1; k < 4; k++) {
    if (co[k] == 0) continue;
    if ((co[k] > 0) != (last > 0)) neg++;
    last = co[k];
  }
  return 2 * neg + 1;
}

// zero inside the closed triangle of vectors a, b, c (exact)
bool zero_in_triangle(const float a[2], const float b[2], const float c[2]) {
  int s0 = sign_cross(b[0], c[1], c[0], b[1]);  // cross(b, c)
  int s1 = sign_cross(c[0], a[1], a[0], c[1]);  // cross(c, a)
  int s2 = sign_cross(a[0], b[1], b[0], a[1]);  // cross(a, b)
  Expansion e;
  auto addp = [&](double s, float x, float y) { e.add(s * ((double)x * y)); };
  addp(1, b[0], c[1]); addp(-1, c[0], b[1]);
  addp(1, c[0], a[1]); addp(-1, a[0], c[1]);
  addp(1, a[0], b[1]); addp(-1, b[0], a[1]);
  int s = e.sign();
  if (s == 0) return false;
  return s0 * s >= 0 && s1 * s >= 0 && s2 * s >= 0;
}

bool zero_in_tet(const float v[4][3]) {
  // barycentric numerators: D_p = det of the other three vectors with
  // alternating sign so that sum_p D_p v_p = 0 and sum_p D_p = D.
  int s[4];
  s[0] = -sign_det3(v[1], v[2], v[3]);
  s[1] = sign_det3(v[0], v[2], v[3]);
  s[2] = -sign_det3(v[0], v[1], v[3]);
  s[3] = sign_det3(v[0], v[1], v[2]);
  // D = sum of the four numerators = det of [v1-v0, v2-v0, v3-v0] up to sign;
  // its sign: evaluate exactly as an expansion of all four determinants.
  Expansion e;
  const int rows[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
  const double sg[4] = {-1, 1, -1, 1};
  for (int p = 0; p < 4; p++) {
    const float* a = v[rows[p][0]];
    const float* b = v[rows[p][1]];
    const float* c = v[rows[p][2]];
    add_triple(e, sg[p], a[0], b[1], c[2]);
    add_triple(e, -sg[p], a[0], b[2], c[1]);
    add_triple(e, -sg[p], a[1], b[0], c[2]);
    add_triple(e, sg[p], a[1], b[2], c[0]);
    add_triple(e, sg[p], a[2], b[0], c[1]);
    add_triple(e, -sg[p], a[2], b[1], c[0]);
  }
  const int sd = e.sign();
  if (sd == 0) return false;
  for (int p = 0; p < 4; p++)
    if (s[p] * sd < 0) return false;
  return true;
}

// ---- Simulation of Simplicity (Edelsbrunner & Muecke 1990): checker
// restatement, written generically. Entry (r, c) of a d x d determinant, r
// the row's rank by global vertex index, is perturbed by eps^(2^(d r + c));
// the perturbed determinant is expanded into monomials (partial matchings of
// rows to columns), ordered by exponent, and the sign is the first nonzero
// coefficient's. A coefficient is the signed sum over the permutations that
// extend the matching of the products of the unperturbed remaining entries,
// evaluated exactly.
int perm_parity(const int* p, int d) {
  int s = 1;
  for (int i = 0; i < d; i++)
    for (int j = i + 1; j < d; j++)
      if (p[i] > p[j]) s = -s;
  return s;
}

int sos_sorted(int d, const float* const* M) {
  struct Mono {
    uint32_t e;
    int col[3];  // column of each row, -1 = unperturbed
  };
  std::vector<Mono> ms;
  // every assignment row -> (unperturbed or a column), columns distinct
  const int choices = d + 1;
  int total = 1;
  for (int r = 0; r < d; r++) total *= choices;
  for (int code = 0; code < total; code++) {
    Mono m{0, {-1, -1, -1}};
    int x = code, used = 0;
    bool ok = true;
    for (int r = 0; r < d; r++) {
      const int c = x % choices - 1;
      x /= choices;
      m.col[r] = c;
      if (c >= 0) {
        if (used >> c & 1) ok = false;
        used |= 1 << c;
        m.e += 1u << (d * r + c);
      }
    }
    if (ok) ms.push_back(m);
  }
  std::sort(ms.begin(), ms.end(), [](const Mono& a, const Mono& b) { return a.e < b.e; });
  for (const Mono& m : ms) {
    Expansion ex;
    int perm[3] = {0, 1, 2};
    // every permutation of the columns extending the matching
    do {
      bool ext = true;
      for (int r = 0; r < d; r++)
        if (m.col[r] >= 0 && perm[r] != m.col[r]) ext = false;
      if (!ext) continue;
      const double sg = perm_parity(perm, d);
      float f[3];
      int nf = 0;
      for (int r = 0; r < d; r++)
        if (m.col[r] < 0) f[nf++] = M[r][perm[r]];
      if (nf == 0) ex.add(sg);
      else if (nf == 1) ex.add(sg * (double)f[0]);
      else if (nf == 2) ex.add(sg * ((double)f[0] * f[1]));  // exact
      else add_triple(ex, sg, f[0], f[1], f[2]);
    } while (std::next_permutation(perm, perm + d));
    const int s = ex.sign();
    if (s) return s;
  }
  return 0;  // not reached: the full matchings have coefficient +-1
}

}  // namespace

int sos_sign_det(int d, const float* rows, const int64_t* idx) {
  int ord[3] = {0, 1, 2};
  std::sort(ord, ord + d, [&](int a, int b) { return idx[a] < idx[b]; });
  const float* M[3];
  for (int r = 0; r < d; r++) M[r] = rows + d * ord[r];
  return perm_parity(ord, d) * sos_sorted(d, M);
}

int tri_cp_2d_sos(const float f[3][2], const int64_t idx[3], int lower) {
  const double X1[3][2] = {{0, 0}, {0, 1}, {1, 1}};
  const double X2[3][2] = {{0, 0}, {1, 0}, {1, 1}};
  // barycentric numerators det[f1; f2], det[f2; f0], det[f0; f1]
  int s[3];
  for (int p = 0; p < 3; p++) {
    const int a = (p + 1) % 3, b = (p + 2) % 3;
    const float rows[4] = {f[a][0], f[a][1], f[b][0], f[b][1]};
    const int64_t ix[2] = {idx[a], idx[b]};
    s[p] = sos_sign_det(2, rows, ix);
  }
  if (s[0] != s[1] || s[0] != s[2]) return -1;
  double v[3][2];
  for (int p = 0; p < 3; p++) v[p][0] = f[p][0], v[p][1] = f[p][1];
  return type_2d(v, lower ? X2 : X1);
}

int tet_cp_3d_sos(const float f[4][3], const int64_t idx[4], int t) {
  static const int A[6][3] = {{1, 0, 0}, {0, 1, 0}, {1, 0, 0}, {0, 0, 1}, {0, 1, 0}, {0, 0, 1}};
  static const int B[6][3] = {{1, 1, 0}, {1, 1, 0}, {1, 0, 1}, {1, 0, 1}, {0, 1, 1}, {0, 1, 1}};
  // D_p = (-1)^(p+1) det of the other three vectors in vertex order
  int s[4];
  for (int p = 0; p < 4; p++) {
    float rows[9];
    int64_t ix[3];
    int k = 0;
    for (int q = 0; q < 4; q++) {
      if (q == p) continue;
      for (int c = 0; c < 3; c++) rows[3 * k + c] = f[q][c];
      ix[k++] = idx[q];
    }
    s[p] = ((p & 1) ? 1 : -1) * sos_sign_det(3, rows, ix);
  }
  if (s[0] != s[1] || s[0] != s[2] || s[0] != s[3]) return -1;
  const int X[4][3] = {{0, 0, 0}, {A[t][2], A[t][1], A[t][0]}, {B[t][2], B[t][1], B[t][0]}, {1, 1, 1}};
  double v[4][3];
  for (int p = 0; p < 4; p++)
    for (int q = 0; q < 3; q++) v[p][q] = f[p][q];
  return type_3d(v, X);
}

std::vector<CP> cp_extract_2d_sos(const float* U, const float* V, size_t r1, size_t r2) {
  std::vector<CP> out;
  for (size_t i = 1; i + 2 < r1; i++)
    for (size_t j = 1; j + 2 < r2; j++) {
      const int64_t cell = 2 * (int64_t)(i * (r2 - 1) + j);
      const int64_t idx[2][3] = {{(int64_t)(i * r2 + j), (int64_t)((i + 1) * r2 + j), (int64_t)((i + 1) * r2 + j + 1)},
                                 {(int64_t)(i * r2 + j), (int64_t)(i * r2 + j + 1), (int64_t)((i + 1) * r2 + j + 1)}};
      for (int t = 0; t < 2; t++) {
        float f[3][2];
        for (int p = 0; p < 3; p++) f[p][0] = U[idx[t][p]], f[p][1] = V[idx[t][p]];
        const int ty = tri_cp_2d_sos(f, idx[t], t);
        if (ty >= 0) out.push_back(CP{cell + t, ty});
      }
    }
  return out;
}

std::vector<CP> cp_extract_3d_sos(const float* U, const float* V, const float* W, size_t r1, size_t r2,
                                  size_t r3) {
  std::vector<CP> out;
  const size_t d0 = r2 * r3, d1 = r3;
  static const int A[6][3] = {{1, 0, 0}, {0, 1, 0}, {1, 0, 0}, {0, 0, 1}, {0, 1, 0}, {0, 0, 1}};
  static const int B[6][3] = {{1, 1, 0}, {1, 1, 0}, {1, 0, 1}, {1, 0, 1}, {0, 1, 1}, {0, 1, 1}};
  static const int SID[6] = {0, 2, 1, 4, 3, 5};
  for (size_t i = 0; i + 1 < r1; i++)
    for (size_t j = 0; j + 1 < r2; j++)
      for (size_t k = 0; k + 1 < r3; k++) {
        const size_t base = i * d0 + j * d1 + k;
        for (int t = 0; t < 6; t++) {
          const int64_t vi[4] = {(int64_t)base, (int64_t)((i + A[t][0]) * d0 + (j + A[t][1]) * d1 + k + A[t][2]),
                                 (int64_t)((i + B[t][0]) * d0 + (j + B[t][1]) * d1 + k + B[t][2]),
                                 (int64_t)((i + 1) * d0 + (j + 1) * d1 + k + 1)};
          float f[4][3];
          for (int p = 0; p < 4; p++) f[p][0] = U[vi[p]], f[p][1] = V[vi[p]], f[p][2] = W[vi[p]];
          const int ty = tet_cp_3d_sos(f, vi, t);
          if (ty >= 0) out.push_back(CP{6 * (int64_t)base + SID[t], ty});
        }
      }
  return out;
}

namespace {
}  // namespace

int tri_cp_2d(const float f[3][2], int lower) {
  const double X1[3][2] = {{0, 0}, {0, 1}, {1, 1}};
  const double X2[3][2] = {{0, 0}, {1, 0}, {1, 1}};
  for (int p = 0; p < 3; p++)
    if (f[p][0] == 0 && f[p][1] == 0) return -1;
  if (!zero_in_triangle(f[0], f[1], f[2])) return -1;
  double v[3][2];
  for (int p = 0; p < 3; p++) v[p][0] = f[p][0], v[p][1] = f[p][1];
  return type_2d(v, lower ? X2 : X1);
}

int tet_cp_3d(const float f[4][3], int t) {
  // (vertex 1, vertex 2) of tetrahedron t as (di, dj, dk), the judge's order
  // (test/cp_extraction_3d.cpp:127-209)
  static const int A[6][3] = {{1, 0, 0}, {0, 1, 0}, {1, 0, 0}, {0, 0, 1}, {0, 1, 0}, {0, 0, 1}};
  static const int B[6][3] = {{1, 1, 0}, {1, 1, 0}, {1, 0, 1}, {1, 0, 1}, {0, 1, 1}, {0, 1, 1}};
  for (int p = 0; p < 4; p++)
    if (f[p][0] == 0 && f[p][1] == 0 && f[p][2] == 0) return -1;
  if (!zero_in_tet(f)) return -1;
  const int X[4][3] = {{0, 0, 0}, {A[t][2], A[t][1], A[t][0]}, {B[t][2], B[t][1], B[t][0]}, {1, 1, 1}};
  double v[4][3];
  for (int p = 0; p < 4; p++)
    for (int q = 0; q < 3; q++) v[p][q] = f[p][q];
  return type_3d(v, X);
}

std::vector<CP> cp_extract_2d(const float* U, const float* V, size_t r1, size_t r2) {
  std::vector<CP> out;
  const double X1[3][2] = {{0, 0}, {0, 1}, {1, 1}};
  const double X2[3][2] = {{0, 0}, {1, 0}, {1, 1}};
  for (size_t i = 1; i + 2 < r1; i++)
    for (size_t j = 1; j + 2 < r2; j++) {
      const int64_t cell = 2 * (int64_t)(i * (r2 - 1) + j);
      const size_t idx[2][3] = {{i * r2 + j, (i + 1) * r2 + j, (i + 1) * r2 + j + 1},
                                {i * r2 + j, i * r2 + j + 1, (i + 1) * r2 + j + 1}};
      for (int t = 0; t < 2; t++) {
        float f[3][2];
        bool zero = false;
        for (int p = 0; p < 3; p++) {
          f[p][0] = U[idx[t][p]];
          f[p][1] = V[idx[t][p]];
          if (f[p][0] == 0 && f[p][1] == 0) zero = true;
        }
        if (zero || !zero_in_triangle(f[0], f[1], f[2])) continue;
        double v[3][2];
        for (int p = 0; p < 3; p++) v[p][0] = f[p][0], v[p][1] = f[p][1];
        out.push_back(CP{cell + t, type_2d(v, t == 0 ? X1 : X2)});
      }
    }
  return out;
}

std::vector<CP> cp_extract_3d(const float* U, const float* V, const float* W, size_t r1, size_t r2,
                              size_t r3) {
  std::vector<CP> out;
  const size_t d0 = r2 * r3, d1 = r3;
  // (id offset, vertex 1 and 2 as (di, dj, dk), coordinates as (dk, dj, di))
  struct TetDef { int sid; int a[3]; int b[3]; };
  const TetDef defs[6] = {
      {0, {1, 0, 0}, {1, 1, 0}}, {2, {0, 1, 0}, {1, 1, 0}}, {1, {1, 0, 0}, {1, 0, 1}},
      {4, {0, 0, 1}, {1, 0, 1}}, {3, {0, 1, 0}, {0, 1, 1}}, {5, {0, 0, 1}, {0, 1, 1}},
  };
  for (size_t i = 0; i + 1 < r1; i++)
    for (size_t j = 0; j + 1 < r2; j++)
      for (size_t k = 0; k + 1 < r3; k++) {
        const size_t base = i * d0 + j * d1 + k;
        for (const TetDef& t : defs) {
          const size_t vi[4] = {base, (i + t.a[0]) * d0 + (j + t.a[1]) * d1 + k + t.a[2],
                                (i + t.b[0]) * d0 + (j + t.b[1]) * d1 + k + t.b[2],
                                (i + 1) * d0 + (j + 1) * d1 + k + 1};
          float f[4][3];
          bool zero = false;
          for (int p = 0; p < 4; p++) {
            f[p][0] = U[vi[p]], f[p][1] = V[vi[p]], f[p][2] = W[vi[p]];
            if (f[p][0] == 0 && f[p][1] == 0 && f[p][2] == 0) zero = true;
          }
          if (zero || !zero_in_tet(f)) continue;
          const int X[4][3] = {{0, 0, 0}, {t.a[2], t.a[1], t.a[0]}, {t.b[2], t.b[1], t.b[0]}, {1, 1, 1}};
          double v[4][3];
          for (int p = 0; p < 4; p++)
            for (int q = 0; q < 3; q++) v[p][q] = f[p][q];
          out.push_back(CP{6 * (int64_t)base + t.sid, type_3d(v, X)});
        }
      }
  return out;
}

}  // namespace orc
