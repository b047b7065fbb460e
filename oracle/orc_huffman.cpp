// Huffman coder restatement — TEST INFRASTRUCTURE (see orc.hpp).
//
// Follows src/sz_huffman.cpp (tree build :195-215 with the priority queue
// :119-144, code assignment :152-187, tree serialization :357-499, bitstream
// :217-320, decode :322-355, tree parsing :501-649) and the section framing in
// src/sz_compression_utils.cpp:61-82 / src/sz_decompression_utils.cpp:92-108.
#include <cstring>
#include <stdexcept>

#include "orc.hpp"

namespace orc {

namespace {

// The reference queue is a 1-based binary heap whose remove() sifts the hole
// to a leaf and drops the last element there without sifting it back up
// (sz_huffman.cpp:130-144); the resulting tie-breaks define the tree shape.
struct RefQueue {
  const std::vector<HuffNode>& nodes;
  std::vector<int> q;  // q[1..end-1]
  int end = 1;
  explicit RefQueue(const std::vector<HuffNode>& n, size_t cap) : nodes(n), q(cap + 2, -1) {}
  void insert(int nd) {  // qinsert, :119-128
    int i = end++;
    int j;
    while ((j = i >> 1)) {
      if (nodes[q[j]].freq <= nodes[nd].freq) break;
      q[i] = q[j];
      i = j;
    }
    q[i] = nd;
  }
  int remove() {  // qremove, :130-144
    if (end < 2) return -1;
    int top = q[1];
    int i = 1, l;
    --end;
    while ((l = i << 1) < end) {
      if (l + 1 < end && nodes[q[l + 1]].freq < nodes[q[l]].freq) ++l;
      q[i] = q[l];
      i = l;
    }
    q[i] = q[end];
    return top;
  }
};

void assign_codes(HuffCodebook& cb, int nd, int len, uint64_t bits) {
  // build_code, :152-187 (codes longer than 64 bits are rejected earlier).
  const HuffNode& n = cb.nodes[nd];
  if (n.leaf) {
    if (len > 64) throw std::runtime_error("huffman code longer than 64 bits");
    // len == 0 (single-symbol alphabet): the reference shifts by 64, which the
    // x86 shifter masks to a shift by 0, leaving code 0 of length 0.
    cb.code_hi[n.sym] = (len == 0) ? 0 : (bits << (64 - len));
    cb.len[n.sym] = (uint8_t)len;
    return;
  }
  assign_codes(cb, n.left, len + 1, bits << 1);
  assign_codes(cb, n.right, len + 1, (bits << 1) | 1);
}

template <typename Idx>
void put_index(std::vector<uint8_t>& out, size_t off, Idx v) {
  std::memcpy(out.data() + off, &v, sizeof(Idx));
}

}  // namespace

bool huff_build(const int* syms, size_t n, uint32_t state_num, HuffCodebook& cb) {
  const size_t all = 2 * (size_t)state_num;  // allNodes, :65
  std::vector<uint64_t> freq(all, 0);
  for (size_t i = 0; i < n; i++) {
    if (syms[i] < 0 || (size_t)syms[i] >= all) return false;
    freq[syms[i]]++;
  }
  cb = HuffCodebook();
  cb.nodes.reserve(2 * all);
  RefQueue q(cb.nodes, 2 * all);
  for (size_t s = 0; s < all; s++) {
    if (!freq[s]) continue;
    cb.nodes.push_back(HuffNode{freq[s], (uint32_t)s, 1, -1, -1});
    q.insert((int)cb.nodes.size() - 1);
  }
  if (cb.nodes.empty()) return false;  // the reference dereferences NULL here
  while (q.end > 2) {
    // new_node(h, 0, 0, qremove(h), qremove(h)) (:211): GCC on x86-64
    // evaluates the call's arguments right to left, so the first removal
    // becomes the right child and the second the left child.
    int right = q.remove();
    int left = q.remove();
    cb.nodes.push_back(HuffNode{cb.nodes[left].freq + cb.nodes[right].freq, 0, 0, left, right});
    q.insert((int)cb.nodes.size() - 1);
  }
  cb.root = q.q[1];
  cb.code_hi.assign(all, 0);
  cb.code_lo.assign(all, 0);
  cb.len.assign(all, 0);
  cb.present.assign(all, 0);
  for (size_t s = 0; s < all; s++) cb.present[s] = freq[s] != 0;
  assign_codes(cb, cb.root, 0, 0);
  return true;
}

std::vector<uint8_t> huff_serialize_tree(const HuffCodebook& cb, size_t node_count) {
  // Preorder numbering: root 0, then left subtree, then right (pad_tree_*, :357-415).
  const size_t w = node_count <= 256 ? 1 : (node_count <= 65536 ? 2 : 4);
  const size_t total = 1 + 2 * node_count * w + 4 * node_count + node_count;
  std::vector<uint8_t> out(total, 0);
  out[0] = 0;  // sysEndianType, a static 0 in the reference (:13)
  const size_t offL = 1, offR = 1 + node_count * w, offC = 1 + 2 * node_count * w,
               offT = offC + 4 * node_count;
  uint32_t counter = 0;
  // explicit stack of (node, slot)
  std::vector<std::pair<int, uint32_t>> stack;
  // recursive lambda keeps the exact visit order: node, left subtree, right subtree
  struct Rec {
    const HuffCodebook& cb;
    std::vector<uint8_t>& out;
    size_t w, offL, offR, offC, offT;
    uint32_t& counter;
    void put(size_t base, uint32_t slot, uint32_t v) {
      if (w == 1) put_index<uint8_t>(out, base + slot, (uint8_t)v);
      else if (w == 2) put_index<uint16_t>(out, base + 2 * (size_t)slot, (uint16_t)v);
      else put_index<uint32_t>(out, base + 4 * (size_t)slot, v);
    }
    void go(int nd, uint32_t slot) {
      const HuffNode& n = cb.nodes[nd];
      uint32_t c = n.leaf ? n.sym : 0;
      std::memcpy(out.data() + offC + 4 * (size_t)slot, &c, 4);
      out[offT + slot] = n.leaf;
      if (n.left >= 0) {
        ++counter;
        put(offL, slot, counter);
        go(n.left, counter);
      }
      if (n.right >= 0) {
        ++counter;
        put(offR, slot, counter);
        go(n.right, counter);
      }
    }
  } rec{cb, out, w, offL, offR, offC, offT, counter};
  rec.go(cb.root, 0);
  return out;
}

std::vector<uint8_t> huff_encode_bits(const HuffCodebook& cb, const int* syms, size_t n) {
  size_t total_bits = 0;
  for (size_t i = 0; i < n; i++) total_bits += cb.len[syms[i]];
  std::vector<uint8_t> out((total_bits + 7) / 8, 0);
  size_t bit = 0;
  for (size_t i = 0; i < n; i++) {
    const int s = syms[i];
    const int L = cb.len[s];
    const uint64_t c = cb.code_hi[s];
    for (int b = 0; b < L; b++, bit++) {
      if ((c >> (63 - b)) & 1) out[bit >> 3] |= (uint8_t)(0x80u >> (bit & 7));
    }
  }
  return out;
}

void huff_encode_section(uint32_t state_num, const int* syms, size_t n, std::vector<uint8_t>& out) {
  HuffCodebook cb;
  if (!huff_build(syms, n, state_num, cb)) throw std::runtime_error("huffman: bad input");
  size_t leaves = 0;
  for (uint32_t s = 0; s < state_num; s++) leaves += cb.present[s];  // counts code[i] for i < stateNum
  const size_t node_count = leaves * 2 - 1;
  std::vector<uint8_t> tree = huff_serialize_tree(cb, node_count);
  std::vector<uint8_t> bits = huff_encode_bits(cb, syms, n);
  const uint32_t tree_size = (uint32_t)tree.size();
  const size_t nbytes = bits.size();
  const size_t at = out.size();
  out.resize(at + 8 + 4 + tree.size() + 8 + bits.size());
  uint8_t* p = out.data() + at;
  std::memcpy(p, &node_count, 8);
  std::memcpy(p + 8, &tree_size, 4);
  std::memcpy(p + 12, tree.data(), tree.size());
  std::memcpy(p + 12 + tree.size(), &nbytes, 8);
  if (!bits.empty()) std::memcpy(p + 20 + tree.size(), bits.data(), bits.size());
}

bool huff_decode_section(const uint8_t*& p, const uint8_t* end, size_t n, std::vector<int>& out) {
  if (end - p < 12) return false;
  size_t node_count;
  uint32_t tree_size;
  std::memcpy(&node_count, p, 8);
  std::memcpy(&tree_size, p + 8, 4);
  p += 12;
  if (node_count == 0 || node_count > (1u << 28)) return false;
  const size_t w = node_count <= 256 ? 1 : (node_count <= 65536 ? 2 : 4);
  const size_t need = 1 + 2 * node_count * w + 5 * node_count;
  if ((size_t)(end - p) < need || tree_size < need) return false;
  const uint8_t* tr = p;
  auto idx = [&](size_t base, size_t i) -> uint32_t {
    if (w == 1) return tr[base + i];
    if (w == 2) { uint16_t v; std::memcpy(&v, tr + base + 2 * i, 2); return v; }
    uint32_t v; std::memcpy(&v, tr + base + 4 * i, 4); return v;
  };
  const size_t offL = 1, offR = 1 + node_count * w, offC = 1 + 2 * node_count * w,
               offT = offC + 4 * node_count;
  auto C = [&](size_t i) { uint32_t v; std::memcpy(&v, tr + offC + 4 * i, 4); return v; };
  auto Tt = [&](size_t i) { return tr[offT + i]; };
  p += tree_size;
  if (end - p < 8) return false;
  size_t nbytes;
  std::memcpy(&nbytes, p, 8);
  p += 8;
  if ((size_t)(end - p) < nbytes) return false;
  const uint8_t* bits = p;
  p += nbytes;
  out.assign(n, 0);
  // Root: uchar trees take C[0], t[0]; wider trees force an internal root
  // (reconstruct_HuffTree_from_bytes_anyStates, :588 vs :615,641).
  const bool root_leaf = (w == 1) && Tt(0);
  if (root_leaf) {  // decode(), :328-333
    for (size_t i = 0; i < n; i++) out[i] = (int)C(0);
    return true;
  }
  size_t cur = 0, bit = 0;
  for (size_t k = 0; k < n;) {
    if ((bit >> 3) >= nbytes) return false;
    const int b = (bits[bit >> 3] >> (7 - (bit & 7))) & 1;
    bit++;
    const uint32_t nx = b ? idx(offR, cur) : idx(offL, cur);
    if (nx == 0 || nx >= node_count) return false;  // a NULL child in the reference
    if (Tt(nx)) {
      out[k++] = (int)C(nx);
      cur = 0;
    } else {
      cur = nx;
    }
  }
  return true;
}

}  // namespace orc
