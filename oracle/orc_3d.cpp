// 3D offline_log codec restatement — TEST INFRASTRUCTURE (see orc.hpp).
//
// Compressor: src/sz_compress_cp_preserve_3d.cpp:9-46 (first six tetrahedra of
// the per-cell table), :264-358 (per-tet bound), :360-581 (driver), with
// include/sz_cp_preserve_utils.hpp:14-22 (eb quantization), :41-57
// (same_direction), :167-191 (acbrt1 bound), :205-221 (log transform).
// Decompressor: src/sz_decompress_cp_preserve_3d.cpp:7-110.
//
// ORC_MUTANT (default 0) plants one known mistake per value, so that a CPU
// test can show each is caught by at least one parity pin:
//   1 MIN/MAX swapped in the per-tet combination (:339-354)
//   2 cofactor sums kept in double instead of float (:334-335)
//   3 '>' instead of '>=' in the unpredictable test (:507)
//   4 first tetrahedron's vertex order permuted (:9-16)
//   5 cbrtf in place of acbrt1 in the sign bound (sz_cp_preserve_utils.hpp:188-189)
//   6 first tetrahedron replaced by one outside the cell's Kuhn split (:9-16)
// Builds with ORC_MUTANT != 0 exist only inside that test.
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif
#include <cfloat>
#include <cmath>
#include <cstring>
#include <limits>

#include "orc.hpp"

namespace orc {

namespace {

template <typename A>
inline A minm(const A& a, const A& b) { return (a < b) ? a : b; }  // MIN, sz_def.hpp:148
template <typename A>
inline A maxm(const A& a, const A& b) { return (a > b) ? a : b; }  // MAX, sz_def.hpp:147

template <typename V>
inline void acc(V value, double& pos, double& neg) {
  if (value >= 0) pos += value;
  else neg += -value;
}

template <typename T>
inline int sgn(T v) { return (T(0) < v) - (v < T(0)); }

// same_direction (sz_cp_preserve_utils.hpp:52-57)
template <typename T>
inline bool same_direction(T a, T b, T c, T d) {
  int s = sgn(a);
  if (s == 0) return false;
  return s == sgn(b) && s == sgn(c) && s == sgn(d);
}

inline float cube_sign_bound(float positive, float negative) { return sign_bound_3d(positive, negative); }

#if ORC_MUTANT == 2
using CofSum = double;
#else
using CofSum = float;  // float p_m0 = positive_m0, ... (:334-335)
#endif

// eb = MAX(eb, MIN(eb1, eb2)) (:339-354)
inline float combine(float eb, float e1, float e2) {
#if ORC_MUTANT == 1
  return minm(eb, maxm(e1, e2));
#else
  return maxm(eb, minm(e1, e2));
#endif
}

template <typename T>
double tet_bound(T u0, T u1, T u2, T u3, T v0, T v1, T v2, T v3, T w0, T w1, T w2, T w3) {
  double u3v1w2 = u3 * v1 * w2, u3v2w1 = u3 * v2 * w1, u3v2w0 = u3 * v2 * w0,
         u3v0w2 = u3 * v0 * w2, u3v0w1 = u3 * v0 * w1, u3v1w0 = u3 * v1 * w0;
  double u1v3w2 = u1 * v3 * w2, u2v3w1 = u2 * v3 * w1, u2v3w0 = u2 * v3 * w0,
         u0v3w2 = u0 * v3 * w2, u0v3w1 = u0 * v3 * w1, u1v3w0 = u1 * v3 * w0;
  double u1v2w3 = u1 * v2 * w3, u2v1w3 = u2 * v1 * w3, u0v2w3 = u0 * v2 * w3,
         u2v0w3 = u2 * v0 * w3, u0v1w3 = u0 * v1 * w3, u1v0w3 = u1 * v0 * w3;
  double u0v1w2 = u0 * v1 * w2, u0v2w1 = u0 * v2 * w1, u1v2w0 = u1 * v2 * w0,
         u1v0w2 = u1 * v0 * w2, u2v0w1 = u2 * v0 * w1, u2v1w0 = u2 * v1 * w0;
  double u3_0 = -u3v1w2 + u3v2w1, u3_1 = -u3v2w0 + u3v0w2, u3_2 = -u3v0w1 + u3v1w0;
  double v3_0 = u1v3w2 - u2v3w1, v3_1 = u2v3w0 - u0v3w2, v3_2 = u0v3w1 - u1v3w0;
  double w3_0 = -u1v2w3 + u2v1w3, w3_1 = u0v2w3 - u2v0w3, w3_2 = -u0v1w3 + u1v0w3;
  double c_4 = u0v1w2 - u0v2w1 + u1v2w0 - u1v0w2 + u2v0w1 - u2v1w0;
  double M0 = u3_0 + v3_0 + w3_0;
  double M1 = u3_1 + v3_1 + w3_1;
  double M2 = u3_2 + v3_2 + w3_2;
  double M3 = c_4;
  double M = M0 + M1 + M2 + M3;
  if (M == 0) return 0;
  bool flag[4];
  flag[0] = (M0 == 0) || (M / M0 > 1);
  flag[1] = (M1 == 0) || (M / M1 > 1);
  flag[2] = (M2 == 0) || (M / M2 > 1);
  flag[3] = (M3 == 0) || (M / M3 > 1);
  if (flag[0] && flag[1] && flag[2] && flag[3]) {
    if (same_direction(u0, u1, u2, u3) || same_direction(v0, v1, v2, v3) ||
        same_direction(w0, w1, w2, w3))
      return 1;
    return 0;
  }
  float eb = 0;
  double p0 = 0, n0 = 0, p1 = 0, n1 = 0, p2 = 0, n2 = 0, p3 = 0, n3 = 0;
  acc(-u3v1w2, p0, n0); acc(u3v2w1, p0, n0); acc(u1v3w2, p0, n0);
  acc(-u2v3w1, p0, n0); acc(-u1v2w3, p0, n0); acc(u2v1w3, p0, n0);
  acc(-u3v2w0, p1, n1); acc(u3v0w2, p1, n1); acc(u2v3w0, p1, n1);
  acc(-u0v3w2, p1, n1); acc(u0v2w3, p1, n1); acc(-u2v0w3, p1, n1);
  acc(-u3v0w1, p2, n2); acc(u3v1w0, p2, n2); acc(u0v3w1, p2, n2);
  acc(-u1v3w0, p2, n2); acc(-u0v1w3, p2, n2); acc(u1v0w3, p2, n2);
  acc(u0v1w2, p3, n3); acc(-u0v2w1, p3, n3); acc(u1v2w0, p3, n3);
  acc(-u1v0w2, p3, n3); acc(u2v0w1, p3, n3); acc(-u2v1w0, p3, n3);
  CofSum pm0 = p0, pm1 = p1, pm2 = p2, pm3 = p3;
  CofSum nm0 = n0, nm1 = n1, nm2 = n2, nm3 = n3;
  if (!flag[0]) {
    float e1 = cube_sign_bound(pm0, nm0), e2 = cube_sign_bound(pm1 + pm2 + pm3, nm1 + nm2 + nm3);
    eb = combine(eb, e1, e2);
  }
  if (!flag[1]) {
    float e1 = cube_sign_bound(pm1, nm1), e2 = cube_sign_bound(pm0 + pm2 + pm3, nm0 + nm2 + nm3);
    eb = combine(eb, e1, e2);
  }
  if (!flag[2]) {
    float e1 = cube_sign_bound(pm2, nm2), e2 = cube_sign_bound(pm0 + pm1 + pm3, nm0 + nm1 + nm3);
    eb = combine(eb, e1, e2);
  }
  if (!flag[3]) {
    float e1 = cube_sign_bound(pm3, nm3), e2 = cube_sign_bound(pm0 + pm1 + pm2, nm0 + nm1 + nm2);
    eb = combine(eb, e1, e2);
  }
  return eb;
}

// Vertex offsets (k, j, i) of the six tetrahedra of a cell; the cell's origin
// vertex is the fourth vertex of each (src/sz_compress_cp_preserve_3d.cpp:9-46).
const int kTet[6][3][3] = {
#if ORC_MUTANT == 4
    {{0, 1, 1}, {0, 0, 1}, {1, 1, 1}}, {{0, 1, 0}, {0, 1, 1}, {1, 1, 1}},
#elif ORC_MUTANT == 6
    {{0, 0, 1}, {1, 1, 0}, {1, 1, 1}}, {{0, 1, 0}, {0, 1, 1}, {1, 1, 1}},
#else
    {{0, 0, 1}, {0, 1, 1}, {1, 1, 1}}, {{0, 1, 0}, {0, 1, 1}, {1, 1, 1}},
#endif
    {{0, 0, 1}, {1, 0, 1}, {1, 1, 1}}, {{1, 0, 0}, {1, 0, 1}, {1, 1, 1}},
    {{0, 1, 0}, {1, 1, 0}, {1, 1, 1}}, {{1, 0, 0}, {1, 1, 0}, {1, 1, 1}},
};

template <typename V>
void put(std::vector<uint8_t>& out, const V& v) {
  const size_t at = out.size();
  out.resize(at + sizeof(V));
  std::memcpy(out.data() + at, &v, sizeof(V));
}
template <typename V>
bool get(const uint8_t*& p, const uint8_t* end, V& v) {
  if ((size_t)(end - p) < sizeof(V)) return false;
  std::memcpy(&v, p, sizeof(V));
  p += sizeof(V);
  return true;
}

void put_bits(std::vector<uint8_t>& out, const std::vector<uint8_t>& flags) {
  // convertIntArray2ByteArray_fast_1b_to_result_sz (src/sz_compression_utils.cpp:31-52)
  const size_t n = flags.size();
  const size_t nb = (n - 1) / 8 + 1;
  const size_t at = out.size();
  out.resize(at + nb, 0);
  for (size_t i = 0; i < n; i++)
    if (flags[i] == 1) out[at + i / 8] |= (uint8_t)(1u << (7 - (i % 8)));
}

constexpr int kCapacity = 65536;
constexpr int kRadius = 32768;

}  // namespace

float acbrt1(float x0) {
  // include/sz_cp_preserve_utils.hpp:173-184: integer first guess, two Newton steps.
  int32_t ix;
  std::memcpy(&ix, &x0, 4);
  ix = ix / 4 + ix / 16;
  ix = ix + ix / 16;
  ix = ix + ix / 256;
  ix = 0x2a5137a0 + ix;
  float x;
  std::memcpy(&x, &ix, 4);
  x = 0.33333333f * (2.0f * x + x0 / (x * x));
  x = 0.33333333f * (2.0f * x + x0 / (x * x));
  return x;
}

// max_eb_to_keep_sign_3d_offline (sz_cp_preserve_utils.hpp:186-191)
float sign_bound_3d(float positive, float negative) {
  if (positive == 0 && negative == 0) return 0;
#if ORC_MUTANT == 5
  float P = cbrtf(positive);
  float N = cbrtf(negative);
#else
  float P = acbrt1(positive);
  float N = acbrt1(negative);
#endif
  return std::fabs(P - N) / (P + N);
}

bool same_direction_3d(float a, float b, float c, float d) { return same_direction(a, b, c, d); }

int eb_exp_quantize(double& eb, int base, double log_of_base, double threshold) {
  // eb_exponential_quantize (sz_cp_preserve_utils.hpp:14-22)
  if (eb <= threshold) {
    eb = 0;
    return 0;
  }
  int id = std::log2(eb / threshold) / log_of_base;
  eb = std::pow((double)base, (double)id) * threshold;
  return id;
}

template <typename T>
void log_transform_3d(const T* x, size_t n, T* logv, uint8_t* sign) {
  // log_transform (sz_cp_preserve_utils.hpp:205-221): log2f of |x| (a double
  // input is converted to float by the log2f call), -100 for zeros
  for (size_t i = 0; i < n; i++) {
    if (x[i] != 0) {
      sign[i] = (x[i] > 0);
      logv[i] = (x[i] > 0) ? log2f(x[i]) : log2f(-x[i]);
    } else {
      sign[i] = 0;
      logv[i] = -100;
    }
  }
}

double tet_bound_3d(const float u[4], const float v[4], const float w[4]) {
  return tet_bound<float>(u[0], u[1], u[2], u[3], v[0], v[1], v[2], v[3], w[0], w[1], w[2], w[3]);
}

template <typename T>
void eb_3d_offline(const T* U, const T* V, const T* W, size_t r1, size_t r2, size_t r3,
                   double max_pwr_eb, double* eb) {
  const size_t n = r1 * r2 * r3, d0 = r2 * r3, d1 = r3;
  for (size_t i = 0; i < n; i++) eb[i] = max_pwr_eb;
  ptrdiff_t off[6][3];
  for (int t = 0; t < 6; t++)
    for (int x = 0; x < 3; x++)
      off[t][x] = kTet[t][x][0] + kTet[t][x][1] * (ptrdiff_t)d1 + kTet[t][x][2] * (ptrdiff_t)d0;
  for (size_t i = 0; i + 1 < r1; i++)
    for (size_t j = 0; j + 1 < r2; j++)
      for (size_t k = 0; k + 1 < r3; k++) {
        const size_t c = i * d0 + j * d1 + k;
        for (int t = 0; t < 6; t++) {
          const size_t a = c + off[t][0], b = c + off[t][1], e = c + off[t][2];
          double cur = tet_bound<T>(U[a], U[b], U[e], U[c], V[a], V[b], V[e], V[c], W[a], W[b],
                                    W[e], W[c]);
          eb[c] = minm(eb[c], cur);
          eb[a] = minm(eb[a], cur);
          eb[b] = minm(eb[b], cur);
          eb[e] = minm(eb[e], cur);
        }
      }
}

int eb_id_3d(double required_eb) {
  if (!(required_eb > 0)) return 0;
  double abs_eb = std::log2(1 + required_eb);
  return eb_exp_quantize(abs_eb, 2, std::log2(2.0), std::numeric_limits<float>::epsilon());
}

template <typename T>
std::vector<uint8_t> compress_3d_offline_log(const T* U, const T* V, const T* W, size_t r1,
                                             size_t r2, size_t r3, double max_pwr_eb) {
  const size_t n = r1 * r2 * r3, d0 = r2 * r3, d1 = r3;
  std::vector<double> eb(n);
  eb_3d_offline(U, V, W, r1, r2, r3, max_pwr_eb, eb.data());
  const T* src[3] = {U, V, W};
  std::vector<T> logv[3];
  std::vector<uint8_t> out;
  const int base = 2;
  put(out, base);
  put(out, (int)kRadius);
  for (int p = 0; p < 3; p++) {
    std::vector<uint8_t> sign(n);
    logv[p].resize(n);
    log_transform_3d(src[p], n, logv[p].data(), sign.data());
    put_bits(out, sign);
  }
  const double threshold = std::numeric_limits<float>::epsilon();
  const double log_of_base = std::log2((double)base);
  const int id_max = (int)(std::log2(1.0 / threshold) / log_of_base) + 1;
  std::vector<int> ids(n);
  std::vector<int> codes(3 * n);
  size_t cpos = 0;
  std::vector<T> unpred;
  size_t idx = 0;
  for (size_t i = 0; i < r1; i++)
    for (size_t j = 0; j < r2; j++)
      for (size_t k = 0; k < r3; k++, idx++) {
        const double required = eb[idx];
        if (required > 0) {
          bool unp = false;
          T dec[3];
          double abs_eb = std::log2(1 + required);
          const int id = eb_exp_quantize(abs_eb, base, log_of_base, threshold);
          if (id > 0) {
            for (int p = 0; p < 3; p++) {
              const T* L = logv[p].data() + idx;
              const T cur = *L;
              T e0 = (i && j && k) ? L[-(ptrdiff_t)(d0 + d1 + 1)] : 0;
              T e1 = (i && j) ? L[-(ptrdiff_t)(d0 + d1)] : 0;
              T e2 = (i && k) ? L[-(ptrdiff_t)(d0 + 1)] : 0;
              T e3 = i ? L[-(ptrdiff_t)d0] : 0;
              T e4 = (j && k) ? L[-(ptrdiff_t)(d1 + 1)] : 0;
              T e5 = j ? L[-(ptrdiff_t)d1] : 0;
              T e6 = k ? L[-1] : 0;
              T pred = e0 + e3 + e5 + e6 - e1 - e2 - e4;
              double diff = cur - pred;
              double qd = std::fabs(diff) / abs_eb + 1;
              if (qd < kCapacity) {
                qd = (diff > 0) ? qd : -qd;
                int qi = (int)(qd / 2) + kRadius;
                codes[cpos + p] = qi;
                dec[p] = pred + 2 * (qi - kRadius) * abs_eb;
#if ORC_MUTANT == 3
                if (std::fabs(dec[p] - cur) > abs_eb) {
#else
                if (std::fabs(dec[p] - cur) >= abs_eb) {
#endif
                  unp = true;
                  break;
                }
              } else {
                unp = true;
                break;
              }
            }
          } else {
            unp = true;
          }
          if (unp) {
            ids[idx] = id_max;
            unpred.push_back(U[idx]);
            unpred.push_back(V[idx]);
            unpred.push_back(W[idx]);
          } else {
            ids[idx] = id;
            cpos += 3;
            for (int p = 0; p < 3; p++) logv[p][idx] = dec[p];
          }
        } else {
          ids[idx] = 0;
          unpred.push_back(U[idx]);
          unpred.push_back(V[idx]);
          unpred.push_back(W[idx]);
        }
      }
  put(out, (size_t)unpred.size());
  {
    const size_t at = out.size();
    out.resize(at + unpred.size() * sizeof(T));
    if (!unpred.empty()) std::memcpy(out.data() + at, unpred.data(), unpred.size() * sizeof(T));
  }
  put(out, (size_t)n);
  huff_encode_section(2 * 256, ids.data(), n, out);
  put(out, (size_t)cpos);
  huff_encode_section(2 * kCapacity, codes.data(), cpos, out);
  return out;
}

template <typename T>
bool decompress_3d_online_log(const uint8_t* s, size_t len, size_t r1, size_t r2, size_t r3,
                              std::vector<T>& U, std::vector<T>& V, std::vector<T>& W) {
  const uint8_t* p = s;
  const uint8_t* end = s + len;
  const size_t n = r1 * r2 * r3, d0 = r2 * r3, d1 = r3;
  int base, radius;
  if (!get(p, end, base) || !get(p, end, radius)) return false;
  const size_t smap = (n - 1) / 8 + 1;
  if ((size_t)(end - p) < 3 * smap) return false;
  const uint8_t* sign[3] = {p, p + smap, p + 2 * smap};
  p += 3 * smap;
  size_t unpred_count;
  if (!get(p, end, unpred_count)) return false;
  if ((size_t)(end - p) / sizeof(T) < unpred_count) return false;
  const uint8_t* unpred = p;
  p += unpred_count * sizeof(T);
  size_t eb_num, data_num;
  std::vector<int> ids, codes;
  if (!get(p, end, eb_num) || eb_num != n) return false;
  if (!huff_decode_section(p, end, eb_num, ids)) return false;
  if (!get(p, end, data_num)) return false;
  if (!huff_decode_section(p, end, data_num, codes)) return false;
  const double threshold = std::numeric_limits<float>::epsilon();
  const double log_of_base = std::log2((double)base);
  const int id_max = (int)(std::log2(1.0 / threshold) / log_of_base) + 1;
  std::vector<T>* outs[3] = {&U, &V, &W};
  for (int q = 0; q < 3; q++) outs[q]->assign(n, 0);
  std::vector<uint8_t> is_unpred(n, 0);
  size_t ui = 0, ci = 0, idx = 0;
  auto unpred_at = [&](size_t k) {
    T v;
    std::memcpy(&v, unpred + k * sizeof(T), sizeof(T));
    return v;
  };
  for (size_t i = 0; i < r1; i++)
    for (size_t j = 0; j < r2; j++)
      for (size_t k = 0; k < r3; k++, idx++) {
        const int id = ids[idx];
        if (id == 0 || id == id_max) {
          is_unpred[idx] = 1;
          for (int q = 0; q < 3; q++) {
            if (ui >= unpred_count) return false;
            const T cur = unpred_at(ui++);
            (*outs[q])[idx] = (cur == 0) ? -100 : log2f(std::fabs(cur));
          }
        } else {
          const double eb = std::pow((double)base, (double)id) * threshold;
          if (ci + 3 > data_num) return false;
          for (int q = 0; q < 3; q++) {
            const T* L = outs[q]->data() + idx;
            T e0 = (i && j && k) ? L[-(ptrdiff_t)(d0 + d1 + 1)] : 0;
            T e1 = (i && j) ? L[-(ptrdiff_t)(d0 + d1)] : 0;
            T e2 = (i && k) ? L[-(ptrdiff_t)(d0 + 1)] : 0;
            T e3 = i ? L[-(ptrdiff_t)d0] : 0;
            T e4 = (j && k) ? L[-(ptrdiff_t)(d1 + 1)] : 0;
            T e5 = j ? L[-(ptrdiff_t)d1] : 0;
            T e6 = k ? L[-1] : 0;
            T pred = e0 + e3 + e5 + e6 - e1 - e2 - e4;
            (*outs[q])[idx] = pred + 2 * (codes[ci + q] - radius) * eb;
          }
          ci += 3;
        }
      }
  ui = 0;
  for (size_t i = 0; i < n; i++) {
    if (is_unpred[i]) {
      for (int q = 0; q < 3; q++) (*outs[q])[i] = unpred_at(ui++);
    } else {
      for (int q = 0; q < 3; q++) {
        T& x = (*outs[q])[i];
        const bool pos = (sign[q][i / 8] >> (7 - (i % 8))) & 1;
        if (x < -99) x = 0;
        else x = pos ? std::exp2(x) : -std::exp2(x);
      }
    }
  }
  return true;
}

template void log_transform_3d<float>(const float*, size_t, float*, uint8_t*);
template void log_transform_3d<double>(const double*, size_t, double*, uint8_t*);
template void eb_3d_offline<float>(const float*, const float*, const float*, size_t, size_t, size_t,
                                   double, double*);
template std::vector<uint8_t> compress_3d_offline_log<float>(const float*, const float*,
                                                             const float*, size_t, size_t, size_t,
                                                             double);
template std::vector<uint8_t> compress_3d_offline_log<double>(const double*, const double*,
                                                              const double*, size_t, size_t,
                                                              size_t, double);
template bool decompress_3d_online_log<float>(const uint8_t*, size_t, size_t, size_t, size_t,
                                              std::vector<float>&, std::vector<float>&,
                                              std::vector<float>&);
template bool decompress_3d_online_log<double>(const uint8_t*, size_t, size_t, size_t, size_t,
                                               std::vector<double>&, std::vector<double>&,
                                               std::vector<double>&);

}  // namespace orc
