// Host side of the Huffman stage: tree construction from a device-computed
// histogram, code assignment, tree serialization, and the parsed decode form.
//
// The stream format and the tree shape must match the reference coder
// (src/sz_huffman.cpp): the tree is built with the reference's priority-queue
// discipline (:119-144, :195-215), codes are assigned left=0/right=1
// (:152-187), and the tree is serialized in preorder with u8/u16/u32 child
// indices chosen by node count (:357-499).
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

namespace cpsz {

struct HuffEncodeTable {
  uint32_t state_num = 0;
  size_t node_count = 0;              // 2 * leaves - 1
  std::vector<uint8_t> tree;          // serialized tree bytes
  // present symbols with their codes (left-aligned in 64 bits) and lengths
  std::vector<uint32_t> syms;
  std::vector<uint64_t> codes;
  std::vector<uint8_t> lens;
  int max_len = 0;
  bool single = false;                // one-symbol alphabet: empty bitstream
};

// Build from frequencies over symbols [0, nbins). Symbols must be < 2*state_num.
void huff_build(const uint64_t* freq, size_t nbins, uint32_t state_num, HuffEncodeTable& t);
// The same from the present symbols only: syms[0..k) strictly increasing,
// freq[i] > 0 the count of syms[i] (the tree depends only on these pairs in
// symbol order, so the result is identical to the dense form's).
void huff_build_sparse(const uint32_t* syms, const uint32_t* freq, size_t k, uint32_t state_num,
                       HuffEncodeTable& t);

// Parsed tree for decoding: node arrays in serialized (preorder) numbering.
struct HuffDecodeTree {
  size_t node_count = 0;
  std::vector<uint32_t> left, right, sym;  // child index 0 = none
  std::vector<uint8_t> leaf;
  bool root_leaf = false;                  // one-symbol alphabet
  uint32_t root_sym = 0;
  int max_len = 0;
  // depth + 1 of each node the parse reached (0 = unreached), read and then
  // overwritten by huff_build_lut
  mutable std::vector<uint8_t> scratch;
};

// Parses `tree_size` bytes of a serialized tree for `node_count` nodes.
// Returns false on malformed input (index out of range, cycles, bad leaves).
bool huff_parse_tree(const uint8_t* tree, size_t tree_size, size_t node_count, HuffDecodeTree& d);

// Lookup table of 2^bits entries for the first `bits` stream bits:
// entry = sym << 8 | len for codes of length <= bits; for longer codes,
// entry = 0x80 | node (internal node reached after `bits` bits) << 8 with len 0.
// Two-level decode tables. Level 1 (2^bits entries) maps the next `bits`
// bits to sym << 8 | length for codes of at most `bits` bits; a prefix that
// ends inside the tree maps to off << 8 | 0x80 | K: level-2 table lut2[off ..
// off + 2^K) over the next K bits (K = min(height of the subtree, bits2)),
// whose entries are sym << 8 | remaining length, or node << 8 (length 0) for
// codes still longer, which continue bit by bit down the tree from `node`.
// Builds level 1 and, per continued prefix, (node, K, off) into `pfx`;
// returns the level-2 entry count (the device fills level 2 from `pfx`).
size_t huff_build_lut(const HuffDecodeTree& d, int bits, int bits2, std::vector<uint32_t>& lut,
                      std::vector<uint32_t>& pfx);

// Host decode of n symbols of an MSB-first bitstream with the same two-level
// tables the device decoder uses (huff_gpu.hip; for tests of the tables).
void decode_host(const uint8_t* tree, size_t tree_size, size_t node_count, const uint8_t* bits, size_t nbytes,
                 size_t n, uint32_t* out);

}  // namespace cpsz
