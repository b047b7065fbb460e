// extern "C" entry points of the CPU oracle for ctypes — TEST INFRASTRUCTURE.
// Buffers returned through uint8_t** are malloc'd; release with orc_free.
#include <cstdlib>
#include <cstring>
#include <exception>

#include "orc.hpp"

using namespace orc;

namespace {
int emit(const std::vector<uint8_t>& v, uint8_t** out, size_t* size) {
  *size = v.size();
  *out = (uint8_t*)std::malloc(v.size() ? v.size() : 1);
  if (!*out) return -2;
  if (!v.empty()) std::memcpy(*out, v.data(), v.size());
  return 0;
}
}  // namespace

extern "C" {

void orc_free(void* p) { std::free(p); }

int orc_compress_2d_f32(const float* U, const float* V, size_t r1, size_t r2, double eb,
                        uint8_t** out, size_t* size) {
  try { return emit(compress_2d_offline<float>(U, V, r1, r2, eb), out, size); } catch (...) { return -1; }
}
int orc_compress_2d_f64(const double* U, const double* V, size_t r1, size_t r2, double eb,
                        uint8_t** out, size_t* size) {
  try { return emit(compress_2d_offline<double>(U, V, r1, r2, eb), out, size); } catch (...) { return -1; }
}
int orc_decompress_2d_f32(const uint8_t* s, size_t len, size_t r1, size_t r2, float* U, float* V) {
  std::vector<float> u, v;
  if (!decompress_2d_offline<float>(s, len, r1, r2, u, v)) return -1;
  std::memcpy(U, u.data(), u.size() * 4);
  std::memcpy(V, v.data(), v.size() * 4);
  return 0;
}
int orc_decompress_2d_f64(const uint8_t* s, size_t len, size_t r1, size_t r2, double* U, double* V) {
  std::vector<double> u, v;
  if (!decompress_2d_offline<double>(s, len, r1, r2, u, v)) return -1;
  std::memcpy(U, u.data(), u.size() * 8);
  std::memcpy(V, v.data(), v.size() * 8);
  return 0;
}
int orc_compress_2d_log_f32(const float* U, const float* V, size_t r1, size_t r2, double eb, uint8_t** out,
                            size_t* size) {
  try { return emit(compress_2d_offline_log<float>(U, V, r1, r2, eb), out, size); } catch (...) { return -1; }
}
int orc_decompress_2d_log_f32(const uint8_t* s, size_t len, size_t r1, size_t r2, float* U, float* V) {
  std::vector<float> u, v;
  if (!decompress_2d_offline_log<float>(s, len, r1, r2, u, v)) return -1;
  std::memcpy(U, u.data(), 4 * r1 * r2);
  std::memcpy(V, v.data(), 4 * r1 * r2);
  return 0;
}

int orc_eb_2d_f32(const float* U, const float* V, size_t r1, size_t r2, double eb, double* out) {
  eb_2d_offline<float>(U, V, r1, r2, eb, out);
  return 0;
}
// ids and Lorenzo precisions from per-vertex eb_u values (|U|*eb).
int orc_eb_quantize_2d(const double* eb_u, size_t n, int* ids, double* prec) {
  for (size_t i = 0; i < n; i++) {
    double e = eb_u[i];
    ids[i] = eb_quantize_2d(e);
    if (e < 2.220446049250313080847263336181640625e-16) e = 0;
    prec[i] = e;
  }
  return 0;
}
int orc_tri_bound_2d(const float* uv6, size_t n, int lower, double* out) {
  for (size_t i = 0; i < n; i++) {
    const float* t = uv6 + 6 * i;
    out[i] = tri_bound_2d(t[0], t[1], t[2], t[3], t[4], t[5], lower);
  }
  return 0;
}
int orc_lorenzo_2d_f32(const float* data, const double* prec, size_t r1, size_t r2, int* codes,
                       float* recon, uint8_t** out, size_t* size) {
  std::vector<int> c;
  std::vector<float> r;
  try {
    auto s = lorenzo_2d_with_eb<float>(data, prec, r1, r2, &c, &r);
    std::memcpy(codes, c.data(), c.size() * 4);
    std::memcpy(recon, r.data(), r.size() * 4);
    return emit(s, out, size);
  } catch (...) { return -1; }
}
int orc_huff_encode(uint32_t state_num, const int* syms, size_t n, uint8_t** out, size_t* size) {
  try {
    std::vector<uint8_t> v;
    huff_encode_section(state_num, syms, n, v);
    return emit(v, out, size);
  } catch (...) { return -1; }
}
int orc_huff_decode(const uint8_t* s, size_t len, size_t n, int* out, size_t* consumed) {
  std::vector<int> v;
  const uint8_t* p = s;
  if (!huff_decode_section(p, s + len, n, v)) return -1;
  std::memcpy(out, v.data(), n * 4);
  *consumed = (size_t)(p - s);
  return 0;
}
int orc_compress_3d_f32(const float* U, const float* V, const float* W, size_t r1, size_t r2,
                        size_t r3, double eb, uint8_t** out, size_t* size) {
  try { return emit(compress_3d_offline_log<float>(U, V, W, r1, r2, r3, eb), out, size); } catch (...) { return -1; }
}
int orc_compress_3d_f64(const double* U, const double* V, const double* W, size_t r1, size_t r2,
                        size_t r3, double eb, uint8_t** out, size_t* size) {
  try { return emit(compress_3d_offline_log<double>(U, V, W, r1, r2, r3, eb), out, size); } catch (...) { return -1; }
}
int orc_decompress_3d_f32(const uint8_t* s, size_t len, size_t r1, size_t r2, size_t r3, float* U,
                          float* V, float* W) {
  std::vector<float> u, v, w;
  if (!decompress_3d_online_log<float>(s, len, r1, r2, r3, u, v, w)) return -1;
  std::memcpy(U, u.data(), u.size() * 4);
  std::memcpy(V, v.data(), v.size() * 4);
  std::memcpy(W, w.data(), w.size() * 4);
  return 0;
}
int orc_decompress_3d_f64(const uint8_t* s, size_t len, size_t r1, size_t r2, size_t r3, double* U,
                          double* V, double* W) {
  std::vector<double> u, v, w;
  if (!decompress_3d_online_log<double>(s, len, r1, r2, r3, u, v, w)) return -1;
  std::memcpy(U, u.data(), u.size() * 8);
  std::memcpy(V, v.data(), v.size() * 8);
  std::memcpy(W, w.data(), w.size() * 8);
  return 0;
}
int orc_eb_3d_f32(const float* U, const float* V, const float* W, size_t r1, size_t r2, size_t r3,
                  double eb, double* out) {
  eb_3d_offline<float>(U, V, W, r1, r2, r3, eb, out);
  return 0;
}
int orc_eb_id_3d(const double* eb, size_t n, int* ids) {
  for (size_t i = 0; i < n; i++) ids[i] = eb_id_3d(eb[i]);
  return 0;
}
int orc_acbrt1(const float* x, size_t n, float* out) {
  for (size_t i = 0; i < n; i++) out[i] = acbrt1(x[i]);
  return 0;
}
float orc_acbrt1_one(float x) { return orc::acbrt1(x); }
int orc_sign_bound_3d(const float* pos, const float* neg, size_t n, float* out) {
  for (size_t i = 0; i < n; i++) out[i] = orc::sign_bound_3d(pos[i], neg[i]);
  return 0;
}
int orc_same_direction_3d(const float* q4, size_t n, int* out) {
  for (size_t i = 0; i < n; i++) out[i] = orc::same_direction_3d(q4[4 * i], q4[4 * i + 1], q4[4 * i + 2], q4[4 * i + 3]);
  return 0;
}
int orc_eb_exp_quantize(const double* eb, size_t n, int base, double log_of_base, double threshold, int* ids,
                        double* eb_out) {
  for (size_t i = 0; i < n; i++) {
    double e = eb[i];
    ids[i] = orc::eb_exp_quantize(e, base, log_of_base, threshold);
    eb_out[i] = e;
  }
  return 0;
}
int orc_log_transform_3d_f32(const float* x, size_t n, float* logv, uint8_t* sign) {
  orc::log_transform_3d(x, n, logv, sign);
  return 0;
}
int orc_tet_bound_3d(const float* uvw12, size_t n, double* out) {
  for (size_t i = 0; i < n; i++) {
    const float* t = uvw12 + 12 * i;
    out[i] = tet_bound_3d(t, t + 4, t + 8);
  }
  return 0;
}
// Critical points: writes up to cap (simplex, type) pairs, returns the count.
int orc_learn_features_2d(const float* U, const float* V, const double* E, size_t r1, size_t r2,
                          double max_pwr_eb, const uint32_t* idx, size_t m, uint16_t* X) {
  learn_features_2d(U, V, E, r1, r2, max_pwr_eb, idx, m, X);
  return 0;
}

int orc_learn_labels_2d(const float* U, const float* V, const double* E, size_t r1, size_t r2,
                        double max_pwr_eb, const uint32_t* idx, size_t m, uint8_t* L) {
  learn_labels_2d(U, V, E, r1, r2, max_pwr_eb, idx, m, L);
  return 0;
}

int orc_learn_features_3d(const float* U, const float* V, const float* W, const double* E, size_t r1, size_t r2,
                          size_t r3, double max_pwr_eb, const uint32_t* idx, size_t m, uint16_t* X) {
  learn_features_3d(U, V, W, E, r1, r2, r3, max_pwr_eb, idx, m, X);
  return 0;
}

int orc_learn_labels_3d(const float* U, const float* V, const float* W, const double* E, size_t r1, size_t r2,
                        size_t r3, double max_pwr_eb, const uint32_t* idx, size_t m, uint8_t* L) {
  learn_labels_3d(U, V, W, E, r1, r2, r3, max_pwr_eb, idx, m, L);
  return 0;
}

int orc_tet_cp_3d(const float* f12, int t) {
  float f[4][3];
  for (int p = 0; p < 4; p++)
    for (int q = 0; q < 3; q++) f[p][q] = f12[3 * p + q];
  return tet_cp_3d(f, t);
}

long orc_cp_2d(const float* U, const float* V, size_t r1, size_t r2, int64_t* simplex, int* type,
               size_t cap) {
  auto cps = cp_extract_2d(U, V, r1, r2);
  for (size_t i = 0; i < cps.size() && i < cap; i++) simplex[i] = cps[i].simplex, type[i] = cps[i].type;
  return (long)cps.size();
}
long orc_cp_3d(const float* U, const float* V, const float* W, size_t r1, size_t r2, size_t r3,
               int64_t* simplex, int* type, size_t cap) {
  auto cps = cp_extract_3d(U, V, W, r1, r2, r3);
  for (size_t i = 0; i < cps.size() && i < cap; i++) simplex[i] = cps[i].simplex, type[i] = cps[i].type;
  return (long)cps.size();
}

int orc_sos_det(int d, const float* rows, const int64_t* idx) { return sos_sign_det(d, rows, idx); }

long orc_cp_2d_sos(const float* U, const float* V, size_t r1, size_t r2, int64_t* simplex, int* type, size_t cap) {
  auto cps = cp_extract_2d_sos(U, V, r1, r2);
  for (size_t i = 0; i < cps.size() && i < cap; i++) simplex[i] = cps[i].simplex, type[i] = cps[i].type;
  return (long)cps.size();
}
long orc_cp_3d_sos(const float* U, const float* V, const float* W, size_t r1, size_t r2, size_t r3,
                   int64_t* simplex, int* type, size_t cap) {
  auto cps = cp_extract_3d_sos(U, V, W, r1, r2, r3);
  for (size_t i = 0; i < cps.size() && i < cap; i++) simplex[i] = cps[i].simplex, type[i] = cps[i].type;
  return (long)cps.size();
}

}  // extern "C"
