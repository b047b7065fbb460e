// 2D offline codec restatement — TEST INFRASTRUCTURE (see orc.hpp).
//
// Compressor: src/sz_compress_cp_preserve_2d.cpp:17-274 with the helpers in
// include/sz_cp_preserve_utils.hpp:14-22,59-90,193-197 and include/sz_def.hpp:147-152.
// Lorenzo codec: src/sz_compress_3d.cpp:126-236,
// include/sz_compress_block_processing.hpp:104-139, include/sz_prediction.hpp:10-14,
// include/sz_compression_utils.hpp:35-65, src/sz_compression_utils.cpp:31-52.
// Decompressor: src/sz_decompress_cp_preserve_2d.cpp:8-44,
// src/sz_decompress_3d.cpp:83-155, include/sz_decompress_block_processing.hpp:65-86,
// include/sz_decompression_utils.hpp:34-73, src/sz_decompression_utils.cpp:62-89.
#include <cfloat>
#include <cmath>
#include <cstring>
#include <limits>

#include "orc.hpp"

namespace orc {

namespace {

template <typename A>
inline A minf(const A& a, const A& b) { return (a < b) ? a : b; }  // MINF, sz_def.hpp:149-152
template <typename A>
inline A maxm(const A& a, const A& b) { return (a > b) ? a : b; }  // MAX, sz_def.hpp:147

// accumulate (sz_cp_preserve_utils.hpp:193-197): the value keeps its own type
// (float products stay float) and is added into double accumulators.
template <typename V>
inline void acc(V value, double& pos, double& neg) {
  if (value >= 0) pos += value;
  else neg += -value;
}

// max_eb_to_keep_sign (sz_cp_preserve_utils.hpp:59-90) for degrees 1 and 2.
inline double sign_bound(double positive, double negative, int degree) {
  if (negative == 0 || positive == 0) return 1;
  double P, N;
  if (degree == 1) {
    P = positive;
    N = negative;
  } else {
    P = std::sqrt(positive);
    N = std::sqrt(negative);
  }
  return std::fabs(P - N) / (P + N);
}

// max_eb_to_keep_sign_2d_offline_2 / _4 (:31-49): double inputs, degree 2.
inline double bound2(double a, double b) {
  double p = 0, n = 0;
  acc(a, p, n);
  acc(b, p, n);
  return sign_bound(p, n, 2);
}
inline double bound4(double a, double b, double c, double d) {
  double p = 0, n = 0;
  acc(a, p, n);
  acc(b, p, n);
  acc(c, p, n);
  acc(d, p, n);
  return sign_bound(p, n, 2);
}

// max_eb_to_keep_sign_det2x2 (:17-29): products in T.
template <typename T>
double det_bound(T u0, T u1, T u2, T v0, T v1, T v2) {
  double p = 0, n = 0;
  acc(u0 * v1, p, n);
  acc(-u0 * v2, p, n);
  acc(u1 * v2, p, n);
  acc(-u1 * v0, p, n);
  acc(u2 * v0, p, n);
  acc(-u2 * v1, p, n);
  return sign_bound(p, n, 2);
}

// max_eb_to_keep_sign_eigen_delta_2 (:75-123) with get_adjugate_matrix_for_position (:54-63).
template <typename T>
double eigen_bound(T u0, T u1, T u2, T v0, T v1, T v2, T x0, T x1, T x2, T y0, T y1, T y2) {
  double eb = 1;
  T c[4];
  {
    T determinant = (x0 - x2) * (y1 - y2) - (x1 - x2) * (y0 - y2);
    c[0] = (y1 - y2) / determinant;
    c[1] = -(y0 - y2) / determinant;
    c[2] = -(x1 - x2) / determinant;
    c[3] = (x0 - x2) / determinant;
    double p = 0, n = 0;
    acc(c[0] * u0, p, n);
    acc(c[1] * u1, p, n);
    acc(-(c[0] + c[1]) * u2, p, n);
    acc(c[2] * v0, p, n);
    acc(c[3] * v1, p, n);
    acc(-(c[2] + c[3]) * v2, p, n);
    eb = sign_bound(p, n, 1);
    eb = minf(eb, det_bound(u0, u1, u2, v0, v1, v2));
  }
  T m = c[1] * c[2] - c[0] * c[3];
  T C = (-m) * (u0 * v1 - u0 * v2 + u1 * v2 - u1 * v0 + u2 * v0 - u2 * v1);
  if (C == 0) return 0;
  if (C < 0) return eb;
  T coeff[6] = {c[0] * u0, c[1] * u1, -(c[1] + c[0]) * u2, c[2] * v0, c[3] * v1, -(c[3] + c[2]) * v2};
  double p = 0, n = 0;
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 6; j++) acc(coeff[i] * coeff[j], p, n);  // accumulate_in_square, :66-73
  acc(-4 * m * u1 * v0, p, n);
  acc(4 * m * u2 * v0, p, n);
  acc(4 * m * u0 * v1, p, n);
  acc(-4 * m * u2 * v1, p, n);
  acc(-4 * m * u0 * v2, p, n);
  acc(4 * m * u1 * v2, p, n);
  eb = minf(eb, sign_bound(p, n, 2));
  return eb;
}

// max_eb_to_keep_position_and_type (:125-172).
template <typename T>
double tri_bound(T u0, T u1, T u2, T v0, T v1, T v2, T x0, T x1, T x2, T y0, T y1, T y2) {
  double u0v1 = u0 * v1, u1v0 = u1 * v0, u0v2 = u0 * v2, u2v0 = u2 * v0, u1v2 = u1 * v2,
         u2v1 = u2 * v1;
  double det = u0v1 - u1v0 + u1v2 - u2v1 + u2v0 - u0v2;
  double eb = 0;
  if (det != 0) {
    bool f1 = (det / (u2v0 - u0v2) >= 1);
    bool f2 = (det / (u1v2 - u2v1) >= 1);
    bool f3 = (det / (u0v1 - u1v0) >= 1);
    double k1 = minf(bound2(u2v0, -u0v2), bound4(u0v1, -u1v0, u1v2, -u2v1));
    double k2 = minf(bound2(u1v2, -u2v1), bound4(u0v1, -u1v0, u2v0, -u0v2));
    double k3 = minf(bound2(u0v1, -u1v0), bound4(u1v2, -u2v1, u2v0, -u0v2));
    if (f1 && f2 && f3) {
      double e4 = minf(k3, eigen_bound(u0, u1, u2, v0, v1, v2, x0, x1, x2, y0, y1, y2));
      eb = minf(k1, k2);
      eb = minf(eb, e4);
    } else {
      eb = 0;
      if (!f1) eb = maxm(eb, k1);
      if (!f2) eb = maxm(eb, k2);
      if (!f3) eb = maxm(eb, k3);
    }
  }
  return eb;
}

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

constexpr int kBlock = 12;        // BSIZE default (include/sz_compress_3d.hpp:12)
constexpr int kCapacity = 65536;  // src/sz_compress_3d.cpp:202
constexpr int kRadius = 32768;

}  // namespace

double tri_bound_2d(float u0, float u1, float u2, float v0, float v1, float v2, int lower) {
  // Triangle coordinates as passed by the compressor ("reversed order", :187-206).
  const float Xu[3][2] = {{0, 0}, {1, 0}, {1, 1}};
  const float Xl[3][2] = {{0, 0}, {0, 1}, {1, 1}};
  const float(*X)[2] = lower ? Xl : Xu;
  return tri_bound<float>(u0, u1, u2, v0, v1, v2, X[0][1], X[1][1], X[2][1], X[0][0], X[1][0],
                          X[2][0]);
}

template <typename T>
void eb_2d_offline(const T* U, const T* V, size_t r1, size_t r2, double max_pwr_eb, double* eb) {
  const size_t n = r1 * r2;
  for (size_t i = 0; i < n; i++) eb[i] = max_pwr_eb;
  const T Xu[3][2] = {{0, 0}, {1, 0}, {1, 1}};
  const T Xl[3][2] = {{0, 0}, {0, 1}, {1, 1}};
  const size_t off_u[3] = {0, r2, r2 + 1};
  const size_t off_l[3] = {0, 1, r2 + 1};
  for (size_t i = 0; i + 1 < r1; i++) {
    for (size_t j = 0; j + 1 < r2; j++) {
      const size_t base = i * r2 + j;
      for (int k = 0; k < 2; k++) {
        const T(*X)[2] = k == 0 ? Xu : Xl;
        const size_t* o = k == 0 ? off_u : off_l;
        double cur = tri_bound<T>(U[base + o[0]], U[base + o[1]], U[base + o[2]], V[base + o[0]],
                                  V[base + o[1]], V[base + o[2]], X[0][1], X[1][1], X[2][1],
                                  X[0][0], X[1][0], X[2][0]);
        for (int m = 0; m < 3; m++) eb[base + o[m]] = minf(eb[base + o[m]], cur);
      }
    }
  }
}

int eb_quantize_2d(double& eb) {
  // eb_exponential_quantize(eb, 4, log2(4), DBL_EPSILON) (sz_cp_preserve_utils.hpp:14-22)
  const double threshold = std::numeric_limits<double>::epsilon();
  const int base = 4;
  const double log_of_base = std::log2((double)base);
  if (eb <= threshold) {
    eb = 0;
    return 0;
  }
  int id = std::log2(eb / threshold) / log_of_base;
  eb = std::pow((double)base, (double)id) * threshold;
  return id;
}

template <typename T>
std::vector<uint8_t> lorenzo_2d_with_eb(const T* data, const double* prec, size_t r1, size_t r2,
                                        std::vector<int>* codes_out, std::vector<T>* recon_out) {
  const size_t n = r1 * r2;
  const size_t nbx = (r1 - 1) / kBlock + 1, nby = (r2 - 1) / kBlock + 1;
  const size_t W = r2 + 1;
  // Reconstructed values with a zero row above and a zero column to the left;
  // equivalent to the reference's (BSIZE+1)x(r2+1) rolling buffer (:144-147,186-187).
  std::vector<T> R((r1 + 1) * W, (T)0);
  std::vector<int> codes;
  codes.reserve(n);
  std::vector<T> unpred;
  for (size_t bi = 0; bi < nbx; bi++) {
    const int sx = ((bi + 1) * kBlock < r1) ? kBlock : (int)(r1 - bi * kBlock);
    for (size_t bj = 0; bj < nby; bj++) {
      const int sy = ((bj + 1) * kBlock < r2) ? kBlock : (int)(r2 - bj * kBlock);
      for (int ii = 0; ii < sx; ii++) {
        for (int jj = 0; jj < sy; jj++) {
          const size_t i = bi * kBlock + ii, j = bj * kBlock + jj;
          const double precision = prec[i * r2 + j];
          const T cur = data[i * r2 + j];
          T* rp = &R[(i + 1) * W + (j + 1)];
          if (precision == 0) {
            codes.push_back(0);
            unpred.push_back(cur);
            *rp = cur;
            continue;
          }
          // lorenzo_predict_2d returns T; the caller stores it in a float.
          const float pred = rp[-1] + rp[-(ptrdiff_t)W] - rp[-(ptrdiff_t)W - 1];
          // quantize (include/sz_compression_utils.hpp:50-65)
          double diff = cur - pred;
          double quant_diff = std::fabs(diff) / precision + 1;
          if (quant_diff < kCapacity) {
            quant_diff = (diff > 0) ? quant_diff : -quant_diff;
            int qi = (int)(quant_diff / 2) + kRadius;
            T dec = pred + 2 * (qi - kRadius) * precision;
            *rp = dec;
            if (std::fabs(dec - cur) <= precision) {
              codes.push_back(qi);
              continue;
            }
          }
          *rp = cur;
          unpred.push_back(cur);
          codes.push_back(0);
        }
      }
    }
  }
  std::vector<uint8_t> out;
  put(out, (int)kBlock);
  put(out, (char)0);  // block_independant
  put(out, (int)kRadius);
  {
    // meanInfo<T>{use_mean=false, mean=0}; padding bytes written as zero
    // (uninitialised in the reference, src/sz_compress_3d.cpp:222).
    uint8_t mi[sizeof(T) * 2] = {0};
    const size_t at = out.size();
    out.resize(at + sizeof(T) * 2);
    std::memcpy(out.data() + at, mi, sizeof(mi));
  }
  put(out, (size_t)0);  // reg_count
  put(out, (size_t)unpred.size());
  {
    const size_t at = out.size();
    out.resize(at + unpred.size() * sizeof(T));
    if (!unpred.empty()) std::memcpy(out.data() + at, unpred.data(), unpred.size() * sizeof(T));
  }
  out.resize(out.size() + (nbx * nby - 1) / 8 + 1, 0);  // indicator bits, all zero
  huff_encode_section(2 * kCapacity, codes.data(), n, out);
  if (codes_out) *codes_out = codes;
  if (recon_out) {
    recon_out->resize(n);
    for (size_t i = 0; i < r1; i++)
      for (size_t j = 0; j < r2; j++) (*recon_out)[i * r2 + j] = R[(i + 1) * W + j + 1];
  }
  return out;
}

template <typename T>
std::vector<uint8_t> compress_2d_offline(const T* U, const T* V, size_t r1, size_t r2,
                                         double max_pwr_eb) {
  const size_t n = r1 * r2;
  std::vector<double> eb(n);
  eb_2d_offline(U, V, r1, r2, max_pwr_eb, eb.data());
  std::vector<double> eb_u(n), eb_v(n);
  std::vector<int> ids(2 * n);
  const double threshold = std::numeric_limits<double>::epsilon();
  for (size_t i = 0; i < n; i++) {
    eb_u[i] = std::fabs(U[i]) * eb[i];
    ids[i] = eb_quantize_2d(eb_u[i]);
    if (eb_u[i] < threshold) eb_u[i] = 0;
  }
  for (size_t i = 0; i < n; i++) {
    eb_v[i] = std::fabs(V[i]) * eb[i];
    ids[n + i] = eb_quantize_2d(eb_v[i]);
    if (eb_v[i] < threshold) eb_v[i] = 0;
  }
  std::vector<uint8_t> sec_eb;
  huff_encode_section(2 * 1024, ids.data(), 2 * n, sec_eb);
  std::vector<uint8_t> sec_u = lorenzo_2d_with_eb<T>(U, eb_u.data(), r1, r2, nullptr, nullptr);
  std::vector<uint8_t> sec_v = lorenzo_2d_with_eb<T>(V, eb_v.data(), r1, r2, nullptr, nullptr);
  std::vector<uint8_t> out;
  put(out, (int)4);
  put(out, threshold);
  put(out, (size_t)sec_eb.size());
  put(out, (size_t)sec_u.size());
  put(out, (size_t)sec_v.size());
  out.insert(out.end(), sec_eb.begin(), sec_eb.end());
  out.insert(out.end(), sec_u.begin(), sec_u.end());
  out.insert(out.end(), sec_v.begin(), sec_v.end());
  return out;
}

namespace {
template <typename T>
bool decode_component(const uint8_t* p, const uint8_t* end, const std::vector<double>& prec,
                      size_t r1, size_t r2, std::vector<T>& out) {
  int block_size;
  char bi;
  int radius;
  if (!get(p, end, block_size) || !get(p, end, bi) || !get(p, end, radius)) return false;
  if (block_size <= 0) return false;
  uint8_t mi[sizeof(T) * 2];
  if ((size_t)(end - p) < sizeof(mi)) return false;
  std::memcpy(mi, p, sizeof(mi));
  p += sizeof(mi);
  const bool use_mean = mi[0] != 0;
  T mean;
  std::memcpy(&mean, mi + sizeof(T), sizeof(T));
  size_t reg_count, unpred_count;
  if (!get(p, end, reg_count) || !get(p, end, unpred_count)) return false;
  if (reg_count != 0) return false;  // regression blocks are never written by the 2D compressor
  if ((size_t)(end - p) / sizeof(T) < unpred_count) return false;
  const uint8_t* unpred = p;
  p += unpred_count * sizeof(T);
  const size_t nbx = (r1 - 1) / block_size + 1, nby = (r2 - 1) / block_size + 1;
  const size_t ind_bytes = (nbx * nby - 1) / 8 + 1;
  if ((size_t)(end - p) < ind_bytes) return false;
  for (size_t b = 0; b < ind_bytes; b++)
    if (p[b]) return false;  // regression indicator set
  p += ind_bytes;
  std::vector<int> codes;
  if (!huff_decode_section(p, end, r1 * r2, codes)) return false;
  const size_t W = r2 + 1;
  std::vector<T> R((r1 + 1) * W, (T)0);
  size_t ci = 0, ui = 0;
  for (size_t bx = 0; bx < nbx; bx++) {
    const int sx = ((bx + 1) * block_size < r1) ? block_size : (int)(r1 - bx * block_size);
    for (size_t by = 0; by < nby; by++) {
      const int sy = ((by + 1) * block_size < r2) ? block_size : (int)(r2 - by * block_size);
      for (int ii = 0; ii < sx; ii++)
        for (int jj = 0; jj < sy; jj++) {
          const size_t i = bx * block_size + ii, j = by * block_size + jj;
          T* rp = &R[(i + 1) * W + j + 1];
          const float pred = rp[-1] + rp[-(ptrdiff_t)W] - rp[-(ptrdiff_t)W - 1];
          const int code = codes[ci++];
          T v;
          if (code == 0) {
            if (ui >= unpred_count) return false;
            std::memcpy(&v, unpred + sizeof(T) * ui++, sizeof(T));
          } else if (code == 1 && use_mean) {
            v = mean;
          } else {
            v = pred + 2 * (code - radius) * prec[i * r2 + j];
          }
          *rp = v;
        }
    }
  }
  out.resize(r1 * r2);
  for (size_t i = 0; i < r1; i++)
    for (size_t j = 0; j < r2; j++) out[i * r2 + j] = R[(i + 1) * W + j + 1];
  return true;
}
}  // namespace

template <typename T>
bool decompress_2d_offline(const uint8_t* s, size_t len, size_t r1, size_t r2, std::vector<T>& U,
                           std::vector<T>& V) {
  const uint8_t* p = s;
  const uint8_t* end = s + len;
  int base;
  double threshold;
  size_t eb_size, u_size, v_size;
  if (!get(p, end, base) || !get(p, end, threshold) || !get(p, end, eb_size) ||
      !get(p, end, u_size) || !get(p, end, v_size))
    return false;
  const size_t n = r1 * r2;
  std::vector<int> ids;
  const uint8_t* q = p;
  if (!huff_decode_section(q, end, 2 * n, ids)) return false;
  // The reader continues right after the id section (the stored eb_size is
  // not consulted, src/sz_decompress_cp_preserve_2d.cpp:27,35).
  std::vector<double> eb(n);
  for (size_t i = 0; i < n; i++) eb[i] = ids[i] == 0 ? 0 : std::pow((double)base, (double)ids[i]) * threshold;
  if (!decode_component<T>(q, end, eb, r1, r2, U)) return false;
  q += u_size;
  for (size_t i = 0; i < n; i++)
    eb[i] = ids[n + i] == 0 ? 0 : std::pow((double)base, (double)ids[n + i]) * threshold;
  if (!decode_component<T>(q, end, eb, r1, r2, V)) return false;
  return true;
}

// ---- offline_log: bounds in the log domain -------------------------------
// sz_compress_cp_preserve_2d_offline_log (src/sz_compress_cp_preserve_2d.cpp:285-374):
// the offline vertex bound eb, then eb := log2(1 + eb) quantized base 2 over
// DBL_EPSILON (one id per vertex, shared by u and v), sign maps of u and v,
// and the Lorenzo codec run on log2f|x| with that bound.
template <typename T>
std::vector<uint8_t> compress_2d_offline_log(const T* U, const T* V, size_t r1, size_t r2, double max_pwr_eb) {
  const size_t n = r1 * r2;
  std::vector<double> eb(n);
  eb_2d_offline(U, V, r1, r2, max_pwr_eb, eb.data());
  const size_t smap = (n - 1) / 8 + 1;
  std::vector<uint8_t> signs(2 * smap, 0);
  std::vector<T> logu(n), logv(n);
  {
    std::vector<uint8_t> sign(n);
    const T* src[2] = {U, V};
    T* dst[2] = {logu.data(), logv.data()};
    for (int c = 0; c < 2; c++) {
      log_transform_3d(src[c], n, dst[c], sign.data());  // log_transform (sz_cp_preserve_utils.hpp:205-221)
      // convertIntArray2ByteArray_fast_1b_to_result_sz (src/sz_compression_utils.cpp:31-52)
      for (size_t i = 0; i < n; i++)
        if (sign[i] == 1) signs[c * smap + i / 8] |= (uint8_t)(1u << (7 - (i % 8)));
    }
  }
  const double threshold = std::numeric_limits<double>::epsilon();
  std::vector<int> ids(n);
  for (size_t i = 0; i < n; i++) {
    eb[i] = std::log2(1 + eb[i]);
    ids[i] = eb_exp_quantize(eb[i], 2, std::log2(2.0), threshold);
  }
  std::vector<uint8_t> sec_eb;
  huff_encode_section(2 * 1024, ids.data(), n, sec_eb);
  std::vector<uint8_t> sec_u = lorenzo_2d_with_eb<T>(logu.data(), eb.data(), r1, r2, nullptr, nullptr);
  std::vector<uint8_t> sec_v = lorenzo_2d_with_eb<T>(logv.data(), eb.data(), r1, r2, nullptr, nullptr);
  std::vector<uint8_t> out;
  put(out, (int)2);
  put(out, threshold);
  put(out, (size_t)sec_eb.size());
  put(out, (size_t)sec_u.size());
  put(out, (size_t)sec_v.size());
  out.insert(out.end(), sec_eb.begin(), sec_eb.end());
  out.insert(out.end(), signs.begin(), signs.end());
  out.insert(out.end(), sec_u.begin(), sec_u.end());
  out.insert(out.end(), sec_v.begin(), sec_v.end());
  return out;
}

// sz_decompress_cp_preserve_2d_offline_log (src/sz_decompress_cp_preserve_2d.cpp:54-99):
// id 0 -> bound 0, else base^id * threshold; the log values decode with that
// bound; x = 0 below -99, else +-exp2 by the sign map.
template <typename T>
bool decompress_2d_offline_log(const uint8_t* s, size_t len, size_t r1, size_t r2, std::vector<T>& U,
                               std::vector<T>& V) {
  const uint8_t* p = s;
  const uint8_t* end = s + len;
  int base;
  double threshold;
  size_t eb_size, u_size, v_size;
  if (!get(p, end, base) || !get(p, end, threshold) || !get(p, end, eb_size) || !get(p, end, u_size) ||
      !get(p, end, v_size))
    return false;
  const size_t n = r1 * r2;
  std::vector<int> ids;
  if (!huff_decode_section(p, end, n, ids)) return false;
  std::vector<double> eb(n);
  for (size_t i = 0; i < n; i++) eb[i] = ids[i] == 0 ? 0 : std::pow((double)base, (double)ids[i]) * threshold;
  const size_t smap = (n - 1) / 8 + 1;
  if ((size_t)(end - p) < 2 * smap) return false;
  const uint8_t* signs = p;
  p += 2 * smap;
  if (!decode_component<T>(p, end, eb, r1, r2, U)) return false;
  p += u_size;
  if (!decode_component<T>(p, end, eb, r1, r2, V)) return false;
  std::vector<T>* out[2] = {&U, &V};
  for (int c = 0; c < 2; c++)
    for (size_t i = 0; i < n; i++) {
      T& x = (*out[c])[i];
      const bool pos = (signs[c * smap + i / 8] >> (7 - (i % 8))) & 1;
      if (x < -99) x = 0;
      else x = pos ? std::exp2(x) : -std::exp2(x);
    }
  return true;
}

template std::vector<uint8_t> compress_2d_offline_log<float>(const float*, const float*, size_t, size_t, double);
template std::vector<uint8_t> compress_2d_offline_log<double>(const double*, const double*, size_t, size_t,
                                                              double);
template bool decompress_2d_offline_log<float>(const uint8_t*, size_t, size_t, size_t, std::vector<float>&,
                                               std::vector<float>&);
template bool decompress_2d_offline_log<double>(const uint8_t*, size_t, size_t, size_t, std::vector<double>&,
                                                std::vector<double>&);
template void eb_2d_offline<float>(const float*, const float*, size_t, size_t, double, double*);
template void eb_2d_offline<double>(const double*, const double*, size_t, size_t, double, double*);
template std::vector<uint8_t> lorenzo_2d_with_eb<float>(const float*, const double*, size_t, size_t,
                                                        std::vector<int>*, std::vector<float>*);
template std::vector<uint8_t> lorenzo_2d_with_eb<double>(const double*, const double*, size_t,
                                                         size_t, std::vector<int>*,
                                                         std::vector<double>*);
template std::vector<uint8_t> compress_2d_offline<float>(const float*, const float*, size_t, size_t,
                                                         double);
template std::vector<uint8_t> compress_2d_offline<double>(const double*, const double*, size_t,
                                                          size_t, double);
template bool decompress_2d_offline<float>(const uint8_t*, size_t, size_t, size_t,
                                           std::vector<float>&, std::vector<float>&);
template bool decompress_2d_offline<double>(const uint8_t*, size_t, size_t, size_t,
                                            std::vector<double>&, std::vector<double>&);

}  // namespace orc
