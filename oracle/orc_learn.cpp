// Learned error-bound proposer checkers — TEST INFRASTRUCTURE (see orc.hpp).
//
// The proposer has no counterpart in the reference; these CPU restatements
// define what the device features and bisection labels must be:
//  * features: 5x5 patch of u then v (zero outside) divided by the patch's
//    largest magnitude, log2(analytic / max_pwr_eb) / 16 (-4 for a zero
//    bound), rounded to BF16 (nearest even), zero padded to 64;
//  * labels: the largest integer gain g in [0, gmax] (gmax: first g with
//    analytic * 2^g >= max_pwr_eb, at most 24) such that moving the vertex
//    alone to any corner of its box u +- |u| eb, v +- |v| eb (eb = min(max,
//    analytic * 2^g), corners rounded to float) keeps the critical point of
//    every incident triangle (existence and type, tri_cp_2d), found by
//    bisection as the device does.
//  * 3D features: 3x3x3 patch of u, then v, then w (i slowest, k fastest;
//    zero outside) divided by the patch's largest magnitude, then the same
//    log feature, rounded to BF16, zero padded to 96;
//  * 3D labels: as in 2D over the tetrahedra incident to the vertex (judge
//    order, tet_cp_3d), with the corners x * (1 + eb) and x / (1 + eb) per
//    component (rounded to float): the offline_log codec keeps each value's
//    sign and |log2|x'| - log2|x|| < log2(1 + eb).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "orc.hpp"

namespace orc {

namespace {
uint16_t f2bf(float x) {
  uint32_t b;
  std::memcpy(&b, &x, 4);
  b += 0x7FFFu + ((b >> 16) & 1u);
  return (uint16_t)(b >> 16);
}
// incident triangles: cell offset (di, dj), kind (0 upper, 1 lower), position
const int kInc[6][4] = {{-1, -1, 0, 2}, {-1, -1, 1, 2}, {-1, 0, 0, 1}, {0, -1, 1, 1}, {0, 0, 0, 0}, {0, 0, 1, 0}};
// vertex p of a triangle of kind t: upper (0,0),(1,0),(1,1); lower (0,0),(0,1),(1,1)
void tri_vertex(int t, int p, int& di, int& dj) {
  static const int up[3][2] = {{0, 0}, {1, 0}, {1, 1}}, lo[3][2] = {{0, 0}, {0, 1}, {1, 1}};
  di = t == 0 ? up[p][0] : lo[p][0];
  dj = t == 0 ? up[p][1] : lo[p][1];
}
}  // namespace

void learn_features_2d(const float* U, const float* V, const double* E, size_t r1, size_t r2,
                       double max_pwr_eb, const uint32_t* idx, size_t m, uint16_t* X) {
  for (size_t q = 0; q < m; q++) {
    const size_t v = idx ? idx[q] : q;
    const long i = (long)(v / r2), j = (long)(v % r2);
    float pu[25], pv[25], s = 0.f;
    for (int a = 0; a < 5; a++)
      for (int b = 0; b < 5; b++) {
        const long ii = i + a - 2, jj = j + b - 2;
        const bool in = ii >= 0 && ii < (long)r1 && jj >= 0 && jj < (long)r2;
        pu[a * 5 + b] = in ? U[ii * r2 + jj] : 0.f;
        pv[a * 5 + b] = in ? V[ii * r2 + jj] : 0.f;
        s = std::max(s, std::max(std::fabs(pu[a * 5 + b]), std::fabs(pv[a * 5 + b])));
      }
    if (!(s > 0.f)) s = 1.f;
    uint16_t* x = X + q * 64;
    for (int k = 0; k < 25; k++) {
      x[k] = f2bf(pu[k] / s);
      x[25 + k] = f2bf(pv[k] / s);
    }
    const double e = E[v];
    x[50] = f2bf(e > 0 ? std::max((float)std::log2(e / max_pwr_eb), -64.f) / 16.f : -4.f);
    for (int k = 51; k < 64; k++) x[k] = 0;
  }
}

void learn_labels_2d(const float* U, const float* V, const double* E, size_t r1, size_t r2,
                     double max_pwr_eb, const uint32_t* idx, size_t m, uint8_t* L) {
  for (size_t q = 0; q < m; q++) {
    const size_t v = idx ? idx[q] : q;
    const long i = (long)(v / r2), j = (long)(v % r2);
    const double e = E[v];
    if (!(e > 0)) {
      L[q] = 0;
      continue;
    }
    float f[6][3][2] = {};
    int base[6], kind[6], pos[6];
    bool valid[6];
    for (int k = 0; k < 6; k++) {
      const long ci = i + kInc[k][0], cj = j + kInc[k][1];
      kind[k] = kInc[k][2];
      pos[k] = kInc[k][3];
      valid[k] = ci >= 0 && ci < (long)r1 - 1 && cj >= 0 && cj < (long)r2 - 1;
      base[k] = -1;
      if (!valid[k]) continue;
      for (int p = 0; p < 3; p++) {
        int di, dj;
        tri_vertex(kind[k], p, di, dj);
        f[k][p][0] = U[(ci + di) * r2 + cj + dj];
        f[k][p][1] = V[(ci + di) * r2 + cj + dj];
      }
      base[k] = tri_cp_2d(f[k], kind[k]);
    }
    const float u = U[v], w = V[v];
    int gmax = 0;
    while (gmax < 24 && std::ldexp(e, gmax) < max_pwr_eb) gmax++;
    auto ok = [&](int g) {
      const double eb = std::min(max_pwr_eb, std::ldexp(e, g));
      const double du = std::fabs((double)u) * eb, dv = std::fabs((double)w) * eb;
      for (int c = 0; c < 4; c++) {
        const float cu = (float)((double)u + ((c & 1) ? du : -du));
        const float cv = (float)((double)w + ((c & 2) ? dv : -dv));
        for (int k = 0; k < 6; k++) {
          if (!valid[k]) continue;
          float g3[3][2];
          for (int p = 0; p < 3; p++) {
            g3[p][0] = p == pos[k] ? cu : f[k][p][0];
            g3[p][1] = p == pos[k] ? cv : f[k][p][1];
          }
          if (tri_cp_2d(g3, kind[k]) != base[k]) return false;
        }
      }
      return true;
    };
    int lo = 0, hi = gmax;
    while (lo < hi) {
      const int mid = (lo + hi + 1) / 2;
      if (ok(mid)) lo = mid;
      else hi = mid - 1;
    }
    L[q] = (uint8_t)lo;
  }
}

void learn_features_3d(const float* U, const float* V, const float* W, const double* E, size_t r1, size_t r2,
                       size_t r3, double max_pwr_eb, const uint32_t* idx, size_t m, uint16_t* X) {
  const float* src[3] = {U, V, W};
  for (size_t q = 0; q < m; q++) {
    const size_t v = idx ? idx[q] : q;
    const long i = (long)(v / (r2 * r3)), j = (long)(v / r3 % r2), k = (long)(v % r3);
    float pt[81], s = 0.f;
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++)
        for (int c = 0; c < 3; c++) {
          const long ii = i + a - 1, jj = j + b - 1, kk = k + c - 1;
          const bool in = ii >= 0 && ii < (long)r1 && jj >= 0 && jj < (long)r2 && kk >= 0 && kk < (long)r3;
          for (int p = 0; p < 3; p++) {
            const float x = in ? src[p][(ii * r2 + jj) * r3 + kk] : 0.f;
            pt[p * 27 + a * 9 + b * 3 + c] = x;
            s = std::max(s, std::fabs(x));
          }
        }
    if (!(s > 0.f)) s = 1.f;
    uint16_t* x = X + q * 96;
    for (int t = 0; t < 81; t++) x[t] = f2bf(pt[t] / s);
    const double e = E[v];
    x[81] = f2bf(e > 0 ? std::max((float)std::log2(e / max_pwr_eb), -64.f) / 16.f : -4.f);
    for (int t = 82; t < 96; t++) x[t] = 0;
  }
}

void learn_labels_3d(const float* U, const float* V, const float* W, const double* E, size_t r1, size_t r2,
                     size_t r3, double max_pwr_eb, const uint32_t* idx, size_t m, uint8_t* L) {
  // corners of tetrahedron t (judge order) as di << 2 | dj << 1 | dk
  static const int C[6][4] = {{0, 4, 6, 7}, {0, 2, 6, 7}, {0, 4, 5, 7}, {0, 1, 5, 7}, {0, 2, 3, 7}, {0, 1, 3, 7}};
  const float* src[3] = {U, V, W};
  const size_t d0 = r2 * r3, d1 = r3;
  struct Inc {
    size_t org;
    int t, pos, base;
  };
  for (size_t q = 0; q < m; q++) {
    const size_t v = idx ? idx[q] : q;
    const long i = (long)(v / d0), j = (long)(v / d1 % r2), k = (long)(v % r3);
    const double e = E[v];
    if (!(e > 0)) {
      L[q] = 0;
      continue;
    }
    auto tet = [&](const Inc& in, const float* xv) {
      float f[4][3];
      for (int p = 0; p < 4; p++) {
        const int c = C[in.t][p];
        const size_t o = in.org + (c >> 2) * d0 + ((c >> 1) & 1) * d1 + (c & 1);
        for (int r = 0; r < 3; r++) f[p][r] = (xv && p == in.pos) ? xv[r] : src[r][o];
      }
      return tet_cp_3d(f, in.t);
    };
    std::vector<Inc> inc;
    for (int c = 0; c < 8; c++) {
      const long ci = i - (c >> 2), cj = j - ((c >> 1) & 1), ck = k - (c & 1);
      if (ci < 0 || ci >= (long)r1 - 1 || cj < 0 || cj >= (long)r2 - 1 || ck < 0 || ck >= (long)r3 - 1) continue;
      for (int t = 0; t < 6; t++)
        for (int p = 0; p < 4; p++)
          if (C[t][p] == c) {
            Inc in{(size_t)ci * d0 + (size_t)cj * d1 + (size_t)ck, t, p, 0};
            in.base = tet(in, nullptr);
            inc.push_back(in);
          }
    }
    const float x0[3] = {U[v], V[v], W[v]};
    int gmax = 0;
    while (gmax < 24 && std::ldexp(e, gmax) < max_pwr_eb) gmax++;
    auto ok = [&](int g) {
      const double up = 1.0 + std::min(max_pwr_eb, std::ldexp(e, g));
      for (int corner = 0; corner < 8; corner++) {
        float xv[3];
        for (int r = 0; r < 3; r++)
          xv[r] = ((corner >> r) & 1) ? (float)((double)x0[r] * up) : (float)((double)x0[r] / up);
        for (const Inc& in : inc)
          if (tet(in, xv) != in.base) return false;
      }
      return true;
    };
    int lo = 0, hi = gmax;
    while (lo < hi) {
      const int mid = (lo + hi + 1) / 2;
      if (ok(mid)) lo = mid;
      else hi = mid - 1;
    }
    L[q] = (uint8_t)lo;
  }
}

}  // namespace orc
