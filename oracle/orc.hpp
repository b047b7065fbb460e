// CPU oracle for the cpSZ critical-point-preserving codec.
//
// TEST INFRASTRUCTURE ONLY. This is a plain C++ restatement of the reference
// CPU library's algorithm (bzhan166/cpSZ, src/ and include/), used by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.
// The product (cpsz_amd/) never includes, links or calls anything here.
//
// Every function cites the reference file:line it restates. Build flags are
// fixed by oracle/Makefile: -O2 -ffp-contract=off, baseline x86-64, so that
// float/double operation order and rounding match the reference build.
//
// Parity status of each part (pinning evidence lives in tests/):
//   Huffman build/serialize/encode/decode ... pinned (golden streams from the
//                                             compiled reference, round trips)
//   2D offline eb derivation + eb ids ....... pinned (golden streams)
//   2D Lorenzo codec (compress/decompress) .. pinned (golden streams, decoded
//                                             fields from the compiled reference)
//   3D offline_log compressor ............... pinned without the reference
//                                             compressor (its file includes FTK):
//                                             acbrt1 (every float), the sign
//                                             bound, same_direction, eb
//                                             quantization and log_transform bit
//                                             for bit against the reference
//                                             header compiled into _ref; the
//                                             per-tet bound against exact
//                                             cofactors on integer tetrahedra;
//                                             vertex bounds against the Kuhn
//                                             split; the stream through the
//                                             compiled reference decoder.
//                                             Planted mistakes (ORC_MUTANT in
//                                             orc_3d.cpp) each fail a pin.
//   3D online_log decompressor .............. pinned (compiled reference decoder)
//   CP extraction (2D/3D) ................... parity unpinned against FTK
//                                             (absent); pinned by exact
//                                             predicates and closed forms.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace orc {

// ---------------------------------------------------------------- Huffman --
// Restates src/sz_huffman.cpp:60-355,357-499,570-649 and the section wrappers
// src/sz_compression_utils.cpp:61-82, src/sz_decompression_utils.cpp:92-108.
struct HuffNode {
  uint64_t freq;
  uint32_t sym;     // symbol for a leaf, 0 for an internal node
  uint8_t leaf;     // 1 leaf, 0 internal
  int32_t left;     // child node index or -1
  int32_t right;
};

struct HuffCodebook {
  std::vector<HuffNode> nodes;
  int root = -1;
  std::vector<uint64_t> code_hi;  // first 64 code bits, left aligned
  std::vector<uint64_t> code_lo;  // bits 64..127 (codes longer than 64 bits)
  std::vector<uint8_t> len;       // code length per symbol (0 = absent)
  std::vector<uint8_t> present;   // symbol occurs
};

// Builds the tree exactly as createHuffmanTree+init (sz_huffman.cpp:60,195).
// Returns false if a symbol is out of range [0, 2*state_num).
bool huff_build(const int* syms, size_t n, uint32_t state_num, HuffCodebook& cb);
// Serialized tree bytes (convert_HuffTree_to_bytes_anyStates, :417-499).
std::vector<uint8_t> huff_serialize_tree(const HuffCodebook& cb, size_t node_count);
// MSB-first bitstream (encode, :217-320).
std::vector<uint8_t> huff_encode_bits(const HuffCodebook& cb, const int* syms, size_t n);
// Full section: size_t node_count | u32 tree_size | tree | size_t nbytes | bits
void huff_encode_section(uint32_t state_num, const int* syms, size_t n, std::vector<uint8_t>& out);
// Parses a section at p (advancing p); returns false on malformed input.
bool huff_decode_section(const uint8_t*& p, const uint8_t* end, size_t n, std::vector<int>& out);

// -------------------------------------------------------------- 2D codec --
// sz_compress_cp_preserve_2d_offline (src/sz_compress_cp_preserve_2d.cpp:175-274)
template <typename T>
std::vector<uint8_t> compress_2d_offline(const T* U, const T* V, size_t r1, size_t r2,
                                         double max_pwr_eb);
// sz_decompress_cp_preserve_2d_offline (src/sz_decompress_cp_preserve_2d.cpp:8-44)
template <typename T>
bool decompress_2d_offline(const uint8_t* s, size_t len, size_t r1, size_t r2,
                           std::vector<T>& U, std::vector<T>& V);

// Stage functions exposed for tests.
// 2D offline_log pair (src/sz_compress_cp_preserve_2d.cpp:285-374,
// src/sz_decompress_cp_preserve_2d.cpp:54-99).
template <typename T>
std::vector<uint8_t> compress_2d_offline_log(const T* U, const T* V, size_t r1, size_t r2, double max_pwr_eb);
template <typename T>
bool decompress_2d_offline_log(const uint8_t* s, size_t len, size_t r1, size_t r2, std::vector<T>& U,
                               std::vector<T>& V);
template <typename T>
void eb_2d_offline(const T* U, const T* V, size_t r1, size_t r2, double max_pwr_eb,
                   double* eb);
// eb_exponential_quantize with base 4 over DBL_EPSILON (sz_cp_preserve_utils.hpp:14-22);
// returns id, writes the Lorenzo precision into *eb_io.
int eb_quantize_2d(double& eb_io);
// sz_compress_2d_with_eb (src/sz_compress_3d.cpp:196-236); also returns the
// block-major codes and the reconstructed field for tests.
template <typename T>
std::vector<uint8_t> lorenzo_2d_with_eb(const T* data, const double* prec, size_t r1,
                                        size_t r2, std::vector<int>* codes_out,
                                        std::vector<T>* recon_out);

// -------------------------------------------------------------- 3D codec --
// sz_compress_cp_preserve_3d_offline_log (src/sz_compress_cp_preserve_3d.cpp:360-581)
template <typename T>
std::vector<uint8_t> compress_3d_offline_log(const T* U, const T* V, const T* W, size_t r1,
                                             size_t r2, size_t r3, double max_pwr_eb);
// sz_decompress_cp_preserve_3d_online_log (src/sz_decompress_cp_preserve_3d.cpp:7-110)
template <typename T>
bool decompress_3d_online_log(const uint8_t* s, size_t len, size_t r1, size_t r2, size_t r3,
                              std::vector<T>& U, std::vector<T>& V, std::vector<T>& W);
template <typename T>
void eb_3d_offline(const T* U, const T* V, const T* W, size_t r1, size_t r2, size_t r3,
                   double max_pwr_eb, double* eb);
// eb id for a vertex bound as the 3D compressor derives it (:470-476):
// 0 when required_eb <= 0, else eb_exponential_quantize(log2(1+eb), 2, 1, FLT_EPSILON).
int eb_id_3d(double required_eb);
// eb_exponential_quantize (include/sz_cp_preserve_utils.hpp:14-22): id, and eb
// replaced by base^id * threshold (0 at or below the threshold).
int eb_exp_quantize(double& eb, int base, double log_of_base, double threshold);
// log_transform (include/sz_cp_preserve_utils.hpp:205-221): log2f|x| and the
// sign flag (1 = positive) per value, -100 and 0 for zeros.
template <typename T>
void log_transform_3d(const T* x, size_t n, T* logv, uint8_t* sign);
float acbrt1(float x);
// max_eb_to_keep_sign_3d_offline (include/sz_cp_preserve_utils.hpp:186-191)
float sign_bound_3d(float positive, float negative);
// same_direction (include/sz_cp_preserve_utils.hpp:52-57)
bool same_direction_3d(float a, float b, float c, float d);
// max_eb_to_keep_position_and_type_3d_offline (:264-358)
double tet_bound_3d(const float u[4], const float v[4], const float w[4]);
// max_eb_to_keep_position_and_type (src/sz_compress_cp_preserve_2d.cpp:125-172)
double tri_bound_2d(float u0, float u1, float u2, float v0, float v1, float v2, int lower);

// ---------------------------------------------------------- CP judge ----
// Critical points per simplex (test/cp_extraction_2d.cpp:62-147,
// test/cp_extraction_3d.cpp:79-209), existence decided by exact signs.
struct CP {
  int64_t simplex;
  int type;
};
std::vector<CP> cp_extract_2d(const float* U, const float* V, size_t r1, size_t r2);
// -1 if the triangle of vertex vectors f[p] = (u, v) holds no critical point,
// else its type; lower selects the coordinates of the lower triangle.
int tri_cp_2d(const float f[3][2], int lower);
// -1 or the type of tetrahedron t (0..5, judge order) with vertex vectors f.
int tet_cp_3d(const float f[4][3], int t);
// Simulation of Simplicity forms (no reference counterpart): rows[0..d*d) of
// vertices with distinct global indices idx; +-1, never 0.
int sos_sign_det(int d, const float* rows, const int64_t* idx);
int tri_cp_2d_sos(const float f[3][2], const int64_t idx[3], int lower);
int tet_cp_3d_sos(const float f[4][3], const int64_t idx[4], int t);
std::vector<CP> cp_extract_2d_sos(const float* U, const float* V, size_t r1, size_t r2);
std::vector<CP> cp_extract_3d_sos(const float* U, const float* V, const float* W, size_t r1, size_t r2,
                                  size_t r3);

// Learned-bound proposer checkers (no reference counterpart; see orc_learn.cpp).
void learn_features_2d(const float* U, const float* V, const double* E, size_t r1, size_t r2,
                       double max_pwr_eb, const uint32_t* idx, size_t m, uint16_t* X);
void learn_labels_2d(const float* U, const float* V, const double* E, size_t r1, size_t r2,
                     double max_pwr_eb, const uint32_t* idx, size_t m, uint8_t* L);
void learn_features_3d(const float* U, const float* V, const float* W, const double* E, size_t r1, size_t r2,
                       size_t r3, double max_pwr_eb, const uint32_t* idx, size_t m, uint16_t* X);
void learn_labels_3d(const float* U, const float* V, const float* W, const double* E, size_t r1, size_t r2,
                     size_t r3, double max_pwr_eb, const uint32_t* idx, size_t m, uint8_t* L);
std::vector<CP> cp_extract_3d(const float* U, const float* V, const float* W, size_t r1,
                              size_t r2, size_t r3);

}  // namespace orc
