#include "huff_host.hpp"

#include <algorithm>
#include <cstring>
#include <string>

#include "common.hpp"

namespace cpsz {

namespace {

// Priority queue with the reference's exact discipline: 1-based heap; insert
// stops at a parent with freq <= new; remove moves the smaller child up along
// one path (right child only if strictly smaller) and drops the last element
// into the vacated leaf slot without re-sifting it (src/sz_huffman.cpp:119-144).
// Entries pack (frequency << 24 | node id) so the sift loops touch one
// 8-byte word per level and a removal hands back its frequency with its id;
// comparisons use the frequency bits only, keeping the reference's tie
// behaviour.
class Queue {
 public:
  // `h` holds at least (entries + 2) words: the spare slot past the end lets
  // the child choice read h[c + 1] unconditionally
  explicit Queue(uint64_t* h) : h_(h) {}
  int size() const { return end_ - 1; }
  void push(uint64_t e) {
    const uint64_t freq = e >> 24;
    uint64_t* h = h_;
    int i = end_++;
    for (int p = i >> 1; p; p = i >> 1) {
      if ((h[p] >> 24) <= freq) break;
      h[i] = h[p];
      i = p;
    }
    h[i] = e;
  }
  uint64_t top() const { return h_[1]; }
  // Removes the two smallest entries as two consecutive qremove() calls would
  // (child choice branch-free). The second pop's descent trails the first's
  // by one level, so their two chains of dependent loads overlap. Every node
  // the second pop reads at a level was last written by the first pop one
  // level further down, earlier in program order; the second pop fills its
  // hole only after the first has filled its own (the last slots can be
  // shared). Each step prefetches the four grandchildren of the node it moves
  // to, the only entries the step after next can read.
  void pop2(uint64_t& a, uint64_t& b) {
    uint64_t* h = h_;
    a = h[1];
    const int e1 = --end_;
    int i1 = 1;
    bool done1 = false;
    auto step1 = [&] {
      const int c = 2 * i1;
      if (c < e1) {
        const int cc = c + (int)((c + 1 < e1) & ((h[c + 1] >> 24) < (h[c] >> 24)));
        __builtin_prefetch(h + 4 * (size_t)cc);
        h[i1] = h[cc];
        i1 = cc;
      } else {
        h[i1] = h[e1];
        done1 = true;
      }
    };
    step1();
    b = h[1];
    const int e2 = --end_;
    int i2 = 1;
    bool walk2 = true;
    while (!done1) {
      step1();
      if (walk2) {
        const int c = 2 * i2;
        if (c < e2) {
          const int cc = c + (int)((c + 1 < e2) & ((h[c + 1] >> 24) < (h[c] >> 24)));
          h[i2] = h[cc];
          i2 = cc;
        } else {
          walk2 = false;
        }
      }
    }
    for (int c = 2 * i2; walk2 && c < e2; c = 2 * i2) {
      const int cc = c + (int)((c + 1 < e2) & ((h[c + 1] >> 24) < (h[c] >> 24)));
      h[i2] = h[cc];
      i2 = cc;
    }
    h[i2] = h[e2];
  }

 private:
  uint64_t* h_;
  int end_ = 1;
};

// Working storage of a tree build, kept per thread: fresh vectors of a few
// hundred KiB would cost a page fault per 4 KiB on every call.
struct BuildWork {
  std::vector<uint64_t> heap;
  std::vector<uint32_t> sym;   // leaf id -> symbol (leaves take ids [0, leaves))
  std::vector<int32_t> kids;   // internal node leaves + k: kids[2k] left, kids[2k + 1] right
};

struct Pending {
  int32_t id;
  uint32_t parent_slot;
  int len;
  uint64_t bits;
};

// Preorder serialization: node, left subtree, right subtree (pad_tree_*,
// :357-415), each child taking the next slot number when it is reached;
// codes: left appends 0, right appends 1 (build_code, :152-187). Walked with
// an explicit stack of right children still to visit.
template <typename I>
void serialize(const BuildWork& wk, uint32_t leaves, int32_t root, size_t N, HuffEncodeTable& t) {
  uint8_t* L = t.tree.data() + 1;
  uint8_t* R = L + N * sizeof(I);
  uint8_t* C = R + N * sizeof(I);
  uint8_t* T = C + 4 * N;
  uint32_t* syms = t.syms.data();
  uint64_t* codes = t.codes.data();
  uint8_t* lens = t.lens.data();
  Pending stack[64];
  int sp = 0;
  uint32_t counter = 0, k = 0;
  int32_t id = root;
  uint32_t slot = 0;
  int len = 0, max_len = 0;
  uint64_t bits = 0;
  for (;;) {
    while (id >= (int32_t)leaves) {
      // children sit at depth len + 1: a code longer than 64 bits follows
      CPSZ_REQUIRE(len < 64, kErrUnsupported, "huffman code longer than 64 bits");
      const int32_t* kd = &wk.kids[2 * (size_t)(id - leaves)];
      const I ls = (I)++counter;
      std::memcpy(L + sizeof(I) * slot, &ls, sizeof(I));
      stack[sp++] = Pending{kd[1], slot, len + 1, (bits << 1) | 1};
      id = kd[0];
      slot = ls;
      len++;
      bits <<= 1;
    }
    const uint32_t s = wk.sym[id];
    std::memcpy(C + 4 * (size_t)slot, &s, 4);
    T[slot] = 1;
    syms[k] = s;
    codes[k] = len ? (bits << (64 - len)) : 0;
    lens[k] = (uint8_t)len;
    k++;
    if (len > max_len) max_len = len;
    if (!sp) break;
    const Pending p = stack[--sp];
    const I rs = (I)++counter;
    std::memcpy(R + sizeof(I) * p.parent_slot, &rs, sizeof(I));
    id = p.id;
    slot = rs;
    len = p.len;
    bits = p.bits;
  }
  CPSZ_REQUIRE(counter + 1 == N && k == leaves, kErrInternal, "huffman tree size mismatch");
  t.max_len = max_len;
}

}  // namespace

namespace {

// The tree from the present symbols in increasing order: sym(i), freq(i) for
// i in [0, k).
template <class SymAt, class FreqAt>
void build_from(size_t k, SymAt sym_at, FreqAt freq_at, uint32_t state_num, HuffEncodeTable& t) {
  // fields are overwritten, not reassigned: a reused table keeps its allocations
  t.state_num = state_num;
  t.max_len = 0;
  const size_t leaves = k;
  uint64_t total = 0;
  for (size_t i = 0; i < k; i++) {
    CPSZ_REQUIRE(sym_at(i) < 2 * (size_t)state_num, kErrInternal, "huffman symbol out of range");
    CPSZ_REQUIRE(i == 0 || sym_at(i) > sym_at(i - 1), kErrInternal, "huffman symbols out of order");
    total += freq_at(i);
  }
  CPSZ_REQUIRE(leaves > 0, kErrInternal, "huffman: empty input");
  CPSZ_REQUIRE(2 * leaves < (1u << 24), kErrUnsupported, "huffman alphabet too large");
  CPSZ_REQUIRE(total < (1ull << 40), kErrUnsupported, "huffman symbol count too large");
  thread_local BuildWork wk;
  wk.heap.resize(2 * leaves + 2);
  wk.sym.resize(leaves);
  wk.kids.resize(2 * leaves);
  Queue q(wk.heap.data());
  uint32_t next = 0;
  size_t counted = 0;  // node_count counts leaves below state_num (src/sz_compression_utils.cpp:66-68)
  for (size_t i = 0; i < k; i++) {
    const uint64_t f = freq_at(i);
    CPSZ_REQUIRE(f > 0, kErrInternal, "huffman: zero frequency listed");
    wk.sym[next] = (uint32_t)sym_at(i);
    counted += sym_at(i) < state_num;
    q.push(f << 24 | next);
    next++;
  }
  // Merge: the reference passes two qremove() calls as arguments to
  // new_node(.., left, right) (:211); its GCC build evaluates them right to
  // left, so the first removal becomes the right child.
  int32_t* kids = wk.kids.data();
  while (q.size() > 1) {
    uint64_t er, el;
    q.pop2(er, el);
    kids[2 * (size_t)(next - leaves)] = (int32_t)(el & 0xFFFFFF);
    kids[2 * (size_t)(next - leaves) + 1] = (int32_t)(er & 0xFFFFFF);
    q.push(((er >> 24) + (el >> 24)) << 24 | next);
    next++;
  }
  const int32_t root = (int32_t)(q.top() & 0xFFFFFF);
  CPSZ_REQUIRE(counted == leaves, kErrUnsupported, "symbol beyond the coder's state count");
  t.node_count = 2 * counted - 1;
  t.single = root < (int32_t)leaves;
  const size_t N = t.node_count;
  const size_t w = N <= 256 ? 1 : (N <= 65536 ? 2 : 4);
  t.tree.assign(1 + 2 * N * w + 5 * N, 0);
  t.syms.resize(leaves);
  t.codes.resize(leaves);
  t.lens.resize(leaves);
  if (w == 1) serialize<uint8_t>(wk, (uint32_t)leaves, root, N, t);
  else if (w == 2) serialize<uint16_t>(wk, (uint32_t)leaves, root, N, t);
  else serialize<uint32_t>(wk, (uint32_t)leaves, root, N, t);
}

}  // namespace

void huff_build(const uint64_t* freq, size_t nbins, uint32_t state_num, HuffEncodeTable& t) {
  thread_local std::vector<uint32_t> present;
  present.clear();
  for (size_t s = 0; s < nbins; s++)
    if (freq[s]) present.push_back((uint32_t)s);
  build_from(
      present.size(), [&](size_t i) { return (size_t)present[i]; },
      [&](size_t i) { return freq[present[i]]; }, state_num, t);
}

void huff_build_sparse(const uint32_t* syms, const uint32_t* freq, size_t k, uint32_t state_num,
                       HuffEncodeTable& t) {
  build_from(
      k, [&](size_t i) { return (size_t)syms[i]; }, [&](size_t i) { return (uint64_t)freq[i]; }, state_num, t);
}

namespace {
template <typename I>
void read_indices(const uint8_t* b, size_t N, uint32_t* out) {
  for (size_t i = 0; i < N; i++) {
    I v;
    std::memcpy(&v, b + sizeof(I) * i, sizeof(I));
    out[i] = v;
  }
}
}  // namespace

bool huff_parse_tree(const uint8_t* tree, size_t tree_size, size_t N, HuffDecodeTree& d) {
  // fields are overwritten, not reassigned: a reused object keeps its
  // allocations (fresh large vectors cost page faults on every call)
  d.node_count = 0;
  d.root_leaf = false;
  d.root_sym = 0;
  d.max_len = 0;
  if (N == 0 || N > (1u << 24)) return false;
  const size_t w = N <= 256 ? 1 : (N <= 65536 ? 2 : 4);
  if (tree_size < 1 + 2 * N * w + 5 * N) return false;
  const uint8_t* L = tree + 1;
  const uint8_t* R = L + N * w;
  const uint8_t* C = R + N * w;
  const uint8_t* T = C + 4 * N;
  d.node_count = N;
  d.left.resize(N);
  d.right.resize(N);
  d.sym.resize(N);
  d.leaf.resize(N);
  if (w == 1) {
    read_indices<uint8_t>(L, N, d.left.data());
    read_indices<uint8_t>(R, N, d.right.data());
  } else if (w == 2) {
    read_indices<uint16_t>(L, N, d.left.data());
    read_indices<uint16_t>(R, N, d.right.data());
  } else {
    std::memcpy(d.left.data(), L, 4 * N);
    std::memcpy(d.right.data(), R, 4 * N);
  }
  std::memcpy(d.sym.data(), C, 4 * N);
  for (size_t i = 0; i < N; i++) d.leaf[i] = T[i] != 0;
  // Root: a u8-indexed tree takes its root's leaf flag from the stream; wider
  // trees always start from an internal root
  // (reconstruct_HuffTree_from_bytes_anyStates, src/sz_huffman.cpp:588,615,641).
  d.root_leaf = (w == 1) && d.leaf[0];
  d.root_sym = d.sym[0];
  if (d.root_leaf) return true;
  d.leaf[0] = 0;
  // Validate the nodes reachable from the root in one forward pass: children
  // point forward (preorder numbering), so a node's depth is known before its
  // index is reached; internal nodes need two children, no node may have two
  // parents, and codes are at most 64 bits.
  // (Raw pointers in locals: the byte stores into the depths may alias
  // anything, which would reload the vectors' data pointers every node.)
  d.scratch.assign(N, 0);
  uint8_t* const depth = d.scratch.data();  // depth + 1 of reached nodes, 0 = unreached
  const uint32_t* const left = d.left.data();
  const uint32_t* const right = d.right.data();
  const uint8_t* const is_leaf = d.leaf.data();
  depth[0] = 1;
  int max_len = 0;
  for (size_t i = 0; i < N; i++) {
    const int di = depth[i];
    if (!di) continue;
    if (is_leaf[i]) {
      if (di - 1 > max_len) max_len = di - 1;
      continue;
    }
    const uint32_t l = left[i], r = right[i];
    if (l >= N || r >= N || l <= i || r <= i || l == r || di > 64) return false;
    if (depth[l] || depth[r]) return false;
    depth[l] = depth[r] = (uint8_t)(di + 1);
  }
  d.max_len = max_len;
  return true;
}

size_t huff_build_lut(const HuffDecodeTree& d, int bits, int bits2, std::vector<uint32_t>& lut,
                      std::vector<uint32_t>& pfx) {
  const size_t N = d.node_count;
  // Level-2 widths: K = the height of a prefix node's subtree, capped at
  // bits2 (any K in [1, bits2] decodes correctly; K only sizes the table). In
  // preorder a subtree is its root followed by the run of deeper nodes, so one
  // forward pass over the depths the parse recorded (depth + 1 of each
  // reachable node, 0 for the others) finds every prefix node's height; K is
  // written over the root's own entry, which the pass has already read.
  std::vector<uint8_t>& h = d.scratch;
  CPSZ_REQUIRE(h.size() == N, kErrInternal, "huffman decode tables need a parsed tree");
  {
    const int droot = bits + 1;
    size_t root = N;
    int maxd = 0;
    auto close = [&] {
      if (root < N) h[root] = (uint8_t)std::max(1, std::min(bits2, maxd - droot));
      root = N;
    };
    for (size_t i = 0; i < N; i++) {
      const int di = h[i];
      if (!di) continue;
      if (di <= droot) {
        close();
        if (di == droot && !d.leaf[i]) {
          root = i;
          maxd = di;
        }
      } else if (di > maxd) {
        maxd = di;
      }
    }
    close();
  }
  lut.assign((size_t)1 << bits, 0);
  pfx.clear();
  size_t total = 0;
  // Depth-first over the nodes of depth <= bits, left before right, so the
  // level-1 ranges (and the level-2 offsets) come out in increasing order: a
  // leaf at depth dl fills the 2^(bits - dl) entries starting with its code; an
  // internal node at depth `bits` starts a level-2 table.
  struct Item {
    uint32_t node, code;
    int depth;
  };
  std::vector<Item> st;
  st.push_back({0, 0, 0});
  while (!st.empty()) {
    const Item it = st.back();
    st.pop_back();
    if (it.depth > 0 && d.leaf[it.node]) {
      const uint32_t e = (d.sym[it.node] << 8) | (uint32_t)it.depth;
      const uint32_t lo = it.code << (bits - it.depth), n = 1u << (bits - it.depth);
      for (uint32_t k = 0; k < n; k++) lut[lo + k] = e;
      continue;
    }
    if (it.depth == bits) {
      const uint32_t K = (uint32_t)std::max(1, std::min<int>(h[it.node], bits2));
      CPSZ_REQUIRE(total < (1u << 24), kErrUnsupported, "huffman decode tables too large");
      pfx.push_back(it.node);
      pfx.push_back(K);
      pfx.push_back((uint32_t)total);
      lut[it.code] = ((uint32_t)total << 8) | 0x80u | K;
      total += (size_t)1 << K;
      continue;
    }
    st.push_back({d.right[it.node], (it.code << 1) | 1u, it.depth + 1});
    st.push_back({d.left[it.node], it.code << 1, it.depth + 1});
  }
  return total;
}

}  // namespace cpsz
