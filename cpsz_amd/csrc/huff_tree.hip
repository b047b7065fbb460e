// Huffman trees on the GPU with the reference coder's exact priority-queue
// discipline (src/sz_huffman.cpp:119-144 insert/remove, :195-215 build,
// :152-187 codes, :357-499 serialization), one wave per tree.
//
// The heap is the reference's 1-based binary heap ordered by frequency; an
// entry is frequency << 24 | node, where node is 0x800000 | symbol for a leaf
// and the merge number for an internal node (the low bits never take part in
// a comparison). Its two operations are walked by the whole wave instead of
// one lane:
//  * remove: the hole left at the root descends along the smaller child (the
//    right one only if strictly smaller) to the last level, then takes the
//    last element without re-sifting it. Each step of the wave covers the 7
//    levels below the hole at once: every lane takes two of the 127 nodes of
//    that window, loads their children and decides which child the hole would
//    take there; two ballots turn the decisions into bit masks that the
//    scalar unit follows down the 7 levels, and the lanes whose nodes lie on
//    the path move their chosen child up, all in one parallel store;
//  * insert: the new element's ancestors are loaded by one lane each; the
//    first ancestor (from below) whose frequency is <= the new one stops the
//    sift, found with one ballot; the ancestors below it move down one level
//    and the element takes the freed slot, all in one parallel store.
// A merge's two children are kept in the heap slot the second removal frees
// (slot k - merge number, which the shrinking heap never uses again), so the
// serialization walk reads them from LDS. The walk runs on one lane in
// preorder (node, left subtree, right subtree; a child gets the next slot
// number when it is reached; left appends 0, right 1), writing the tree bytes
// and the dense codebook the encoder reads.
//
// Within one wave, LDS accesses complete in issue order, so the heap steps
// need no barrier when the heap lives in LDS; trees too large for LDS keep it
// in global memory, with a barrier (and its wait for outstanding accesses)
// between dependent steps.
#include "huff_gpu.hpp"

namespace cpsz {

namespace {

constexpr int kTreeLds = 8190;        // heap slots 1..8190 in LDS (64 KiB): trees of up to 8190 leaves
constexpr uint32_t kLeaf = 0x800000;  // node field of a leaf entry: kLeaf | symbol
constexpr uint64_t kNode = 0xFFFFFF;

__device__ __forceinline__ uint64_t rl64(uint64_t v, int l) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
  return (uint64_t)hi << 32 | lo;
}

template <bool LDS>
__device__ __forceinline__ void step_sync() {
  if constexpr (LDS) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// qinsert (:119-128): slot `end` (1-based), sift up while the parent's
// frequency is larger.
template <bool LDS>
__device__ __forceinline__ void heap_push(uint64_t* H, int& end, uint64_t x) {
  const int lane = threadIdx.x;
  const int slot = end++;
  const uint64_t f = x >> 24;
  const int depth = 31 - __clz(slot);  // ancestors slot >> 1 .. slot >> depth (= 1)
  uint64_t a = 0;
  bool stop = false;
  if (lane < depth) {
    a = H[slot >> (lane + 1)];
    stop = (a >> 24) <= f;
  }
  const uint64_t m = __ballot(stop);
  const int k = m ? (int)__ffsll((long long)m) - 1 : depth;
  step_sync<LDS>();
  if (lane < k) H[slot >> lane] = a;
  if (lane == 0) H[slot >> k] = x;
  step_sync<LDS>();
}

// One window node of the removal's descent: node t (1-based within the
// window, level lt below the window's root i) at heap slot (i << lt) + off.
struct WinNode {
  int lt, off;
  bool valid;
};

template <bool LDS>
__device__ __forceinline__ void win_decide(const uint64_t* H, const WinNode& w, int i, int e, bool& ex,
                                           bool& right, uint64_t& chosen) {
  ex = right = false;
  chosen = 0;
  if (!w.valid) return;
  const int c = 2 * ((i << w.lt) + w.off);
  if (c >= e) return;
  ex = true;
  const uint64_t vc = H[c];
  if (c + 1 < e) {
    const uint64_t vd = H[c + 1];
    right = (vd >> 24) < (vc >> 24);
    chosen = right ? vd : vc;
  } else {
    chosen = vc;
  }
}

// qremove (:130-144): returns the root; the hole descends along the smaller
// child while a child exists, then takes the last element.
template <bool LDS>
__device__ __forceinline__ uint64_t heap_pop(uint64_t* H, int& end, const WinNode& w0, const WinNode& w1) {
  const int lane = threadIdx.x;
  const uint64_t top = H[1];
  const int e = --end;  // the last element's slot
  const uint64_t last = H[e];
  int i = 1;
  for (;;) {
    bool ex0, r0, ex1, r1;
    uint64_t ch0, ch1;
    win_decide<LDS>(H, w0, i, e, ex0, r0, ch0);
    win_decide<LDS>(H, w1, i, e, ex1, r1, ch1);
    const uint64_t E0 = __ballot(ex0), R0 = __ballot(r0), E1 = __ballot(ex1), R1 = __ballot(r1);
    // follow the decisions from the window's root (t = 1) down at most 7 levels
    int t = 1, lv = 0;
    for (; lv < 7; lv++) {
      const int b = t - 1;
      const uint64_t em = b < 64 ? E0 : E1, rm = b < 64 ? R0 : R1;
      if (!((em >> (b & 63)) & 1)) break;
      t = 2 * t + (int)((rm >> (b & 63)) & 1);
    }
    // the path's nodes above the hole take their chosen child's value
    if (w0.valid && w0.lt < lv && (t >> (lv - w0.lt)) == (1 << w0.lt) + w0.off)
      H[(i << w0.lt) + w0.off] = ch0;
    if (w1.valid && w1.lt < lv && (t >> (lv - w1.lt)) == (1 << w1.lt) + w1.off)
      H[(i << w1.lt) + w1.off] = ch1;
    i = (i << lv) + t - (1 << lv);
    if (lv < 7) break;
    step_sync<LDS>();
  }
  step_sync<LDS>();
  if (lane == 0) H[i] = last;
  step_sync<LDS>();
  return top;
}

__device__ __forceinline__ void put_le(uint8_t* p, uint32_t v, int w) {
  for (int b = 0; b < w; b++) p[b] = (uint8_t)(v >> (8 * b));
}

template <bool LDS>
__device__ void build_tree(const TreeJob& J, uint64_t* H, uint32_t k, const uint32_t* syms, const uint32_t* freqs,
                           int32_t* stk_id, uint32_t* stk_slot, uint32_t* stk_len, uint64_t* stk_bits) {
  const int lane = threadIdx.x;
  uint64_t* info = J.info;
  WinNode w0, w1;
  {
    const int t0 = lane + 1, t1 = lane + 65;
    w0.lt = 31 - __clz(t0);
    w0.off = t0 - (1 << w0.lt);
    w0.valid = true;
    w1.lt = 31 - __clz(t1);
    w1.off = t1 - (1 << w1.lt);
    w1.valid = t1 < 128;
  }
  // ---- heap of the leaves in symbol order (:205-207), 64 leaves per load
  int end = 1;
  for (uint32_t q0 = 0; q0 < k; q0 += 64) {
    uint64_t mine = 0;
    if (q0 + lane < k) mine = (uint64_t)freqs[q0 + lane] << 24 | kLeaf | syms[q0 + lane];
    const uint32_t cnt = k - q0 < 64 ? k - q0 : 64;
    for (uint32_t j = 0; j < cnt; j++) heap_push<LDS>(H, end, rl64(mine, (int)j));
  }
  // ---- the merges (:208-212); the first removal becomes the right child
  // (the reference's GCC build evaluates new_node's two qremove() arguments
  // right to left)
  for (uint32_t j = 0; end > 2; j++) {
    const uint64_t er = heap_pop<LDS>(H, end, w0, w1);
    const uint64_t el = heap_pop<LDS>(H, end, w0, w1);
    if (lane == 0) H[end + 1] = (el & kNode) | (er & kNode) << 32;  // slot k - j
    heap_push<LDS>(H, end, ((er >> 24) + (el >> 24)) << 24 | j);
  }
  const uint32_t root = (uint32_t)(H[1] & kNode);
  // ---- serialization: zeroed tree bytes, then the preorder walk
  const uint32_t N = 2 * k - 1;
  const int w = N <= 256 ? 1 : (N <= 65536 ? 2 : 4);
  const size_t tree_bytes = 1 + 2 * (size_t)N * w + 5 * (size_t)N;
  for (size_t b = lane; b < tree_bytes; b += 64) J.tree[b] = 0;
  __syncthreads();
  if (lane != 0) return;
  uint8_t* L = J.tree + 1;
  uint8_t* R = L + (size_t)N * w;
  uint8_t* C = R + (size_t)N * w;
  uint8_t* T = C + 4 * (size_t)N;
  int sp = 0;
  uint32_t counter = 0, slot = 0, len = 0, max_len = 0;
  uint32_t id = root;
  uint64_t bits = 0;
  bool bad = false;
  for (;;) {
    while (!(id & kLeaf)) {
      if (len >= 64) {  // children at depth len + 1: codes longer than 64 bits are refused
        bad = true;
        break;
      }
      const uint64_t kd = H[k - id];
      const uint32_t ls = ++counter;
      put_le(L + (size_t)w * slot, ls, w);
      stk_id[sp] = (int32_t)(kd >> 32);
      stk_slot[sp] = slot;
      stk_len[sp] = len + 1;
      stk_bits[sp] = bits << 1 | 1;
      sp++;
      id = (uint32_t)(kd & kNode);
      slot = ls;
      len++;
      bits <<= 1;
    }
    if (bad) break;
    const uint32_t s = id & ~kLeaf;
    put_le(C + 4 * (size_t)slot, s, 4);
    T[slot] = 1;
    J.code[s] = len ? bits << (64 - len) : 0;
    J.len[s] = (uint8_t)len;
    if (len > max_len) max_len = len;
    if (!sp) break;
    sp--;
    const uint32_t rs = ++counter;
    put_le(R + (size_t)w * stk_slot[sp], rs, w);
    id = (uint32_t)stk_id[sp];
    slot = rs;
    len = stk_len[sp];
    bits = stk_bits[sp];
  }
  info[0] = N;
  info[1] = tree_bytes;
  info[2] = max_len;
  info[3] = bad ? 4 : (counter + 1 == N ? 0 : 5);
}

}  // namespace

__global__ void __launch_bounds__(64) k_huff_tree(const TreeJob* __restrict__ jobs) {
  __shared__ uint64_t lheap[kTreeLds + 1];
  __shared__ int32_t stk_id[64];
  __shared__ uint32_t stk_slot[64], stk_len[64];
  __shared__ uint64_t stk_bits[64];
  const TreeJob& J = jobs[blockIdx.x];
  const int lane = threadIdx.x;
  const uint32_t k = *J.count;
  const uint32_t* syms = J.syms + *J.base;
  const uint32_t* freqs = J.freqs + *J.base;
  // node_count counts leaves below state_num (src/sz_compression_utils.cpp:66-68)
  uint32_t counted = 0;
  for (uint32_t q = lane; q < k; q += 64) counted += syms[q] < J.state_num;
  for (int o = 32; o > 0; o >>= 1) counted += __shfl_xor(counted, o);
  if (k == 0 || k > J.max_leaves || k >= kLeaf || J.state_num > kLeaf || counted != k) {
    if (lane == 0) J.info[3] = k == 0 ? 1 : (k > J.max_leaves || k >= kLeaf || J.state_num > kLeaf ? 2 : 3);
    return;
  }
  if (k <= (uint32_t)kTreeLds)
    build_tree<true>(J, lheap, k, syms, freqs, stk_id, stk_slot, stk_len, stk_bits);
  else
    build_tree<false>(J, J.heap, k, syms, freqs, stk_id, stk_slot, stk_len, stk_bits);
}

void build_trees_device(Ctx& ctx, const TreeJob* d_jobs, int ntrees) {
  if (ntrees <= 0) return;
  k_huff_tree<<<ntrees, 64, 0, ctx.stream()>>>(d_jobs);
  CPSZ_HIP(hipGetLastError());
}

}  // namespace cpsz
