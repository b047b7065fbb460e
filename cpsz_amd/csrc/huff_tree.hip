// Huffman trees on the GPU with the reference coder's exact priority-queue
// discipline (src/sz_huffman.cpp:119-144 insert/remove, :195-215 build,
// :152-187 codes, :357-499 serialization), one wave per tree.
//
// The heap is the reference's 1-based binary heap of (frequency, node) words
// (frequency << 24 | node id, as the host builder packs them). Its two
// operations are walked by the whole wave instead of one lane:
//  * remove: the hole left at the root descends along the smaller child
//    (the right one only if strictly smaller) to the last level, then takes
//    the last element without re-sifting it. The wave loads the 62 entries of
//    the next 5 levels below the hole at once and follows the path through
//    them with lane reads, so a descent costs one memory round trip per 5
//    levels instead of one per level;
//  * insert: the new element's ancestors are loaded by one lane each; the
//    first ancestor (from below) whose frequency is <= the new one stops the
//    sift, found with one ballot; the ancestors below it move down one level
//    and the element takes the freed slot, all in one parallel store.
// The serialization walks the finished tree in preorder on one lane (node,
// left subtree, right subtree; a child gets the next slot number when it is
// reached; left appends 0, right 1), writing the tree bytes and the dense
// codebook the encoder reads.
#include "huff_gpu.hpp"

namespace cpsz {

namespace {

constexpr int kTreeLds = 8190;  // heap entries kept in LDS (64 KiB); larger heaps use TreeJob::heap

__device__ __forceinline__ uint64_t rl64(uint64_t v, int l) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
  return (uint64_t)hi << 32 | lo;
}

// qinsert (:119-128): slot `end` (1-based), sift up while the parent's
// frequency is larger.
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
  __syncthreads();
  if (lane < k) H[slot >> lane] = a;
  if (lane == 0) H[slot >> k] = x;
  __syncthreads();
}

// qremove (:130-144): returns the root; the hole descends along the smaller
// child while a child exists, then takes the last element.
__device__ __forceinline__ uint64_t heap_pop(uint64_t* H, int& end) {
  const int lane = threadIdx.x;
  const uint64_t top = H[1];
  const int e = --end;  // the last element's slot
  int i = 1;
  for (bool done = false; !done;) {
    const int t = lane + 2;
    const int lvl = 31 - __clz(t);  // lanes 0..61: levels 1..5 below i
    const int node = (i << lvl) + (t - (1 << lvl));
    uint64_t v = ~0ull;
    if (lane < 62 && node < e) v = H[node];
    int cur = i;
    for (int l = 1; l <= 5; l++) {
      const int c = 2 * cur;
      if (c >= e) {
        done = true;
        break;
      }
      const int tc = c - (i << l) + (1 << l) - 2;  // the lane holding node c
      const uint64_t vc = rl64(v, tc), vd = rl64(v, tc + 1);
      const bool right = (c + 1 < e) && ((vd >> 24) < (vc >> 24));
      if (lane == 0) H[cur] = right ? vd : vc;
      cur = c + (right ? 1 : 0);
    }
    i = cur;
    __syncthreads();
  }
  if (lane == 0) H[i] = H[e];
  __syncthreads();
  return top;
}

__device__ __forceinline__ void put_le(uint8_t* p, uint32_t v, int w) {
  for (int b = 0; b < w; b++) p[b] = (uint8_t)(v >> (8 * b));
}

}  // namespace

__global__ void __launch_bounds__(64) k_huff_tree(const TreeJob* __restrict__ jobs) {
  __shared__ uint64_t lheap[kTreeLds + 2];
  __shared__ int32_t stk_id[64];
  __shared__ uint32_t stk_slot[64], stk_len[64];
  __shared__ uint64_t stk_bits[64];
  const TreeJob& J = jobs[blockIdx.x];
  const int lane = threadIdx.x;
  const uint32_t k = *J.count;
  const uint32_t* syms = J.syms + *J.base;
  const uint32_t* freqs = J.freqs + *J.base;
  uint64_t* info = J.info;
  // node_count counts leaves below state_num (src/sz_compression_utils.cpp:66-68)
  uint32_t counted = 0;
  for (uint32_t q = lane; q < k; q += 64) counted += syms[q] < J.state_num;
  for (int o = 32; o > 0; o >>= 1) counted += __shfl_xor(counted, o);
  if (k == 0 || k > J.max_leaves || counted != k) {
    if (lane == 0) info[3] = k == 0 ? 1 : (k > J.max_leaves ? 2 : 3);
    return;
  }
  uint64_t* H = k + 2 <= (uint32_t)(kTreeLds + 2) ? lheap : J.heap;
  // ---- heap of the leaves in symbol order, then the merges (:205-212); the
  // first removal becomes the right child (the reference's GCC build
  // evaluates new_node's two qremove() arguments right to left)
  int end = 1;
  for (uint32_t q = 0; q < k; q++) heap_push(H, end, (uint64_t)freqs[q] << 24 | q);
  uint32_t next = k;
  while (end > 2) {
    const uint64_t er = heap_pop(H, end);
    const uint64_t el = heap_pop(H, end);
    if (lane == 0) {
      J.kids[2 * (size_t)(next - k)] = (int32_t)(el & 0xFFFFFF);
      J.kids[2 * (size_t)(next - k) + 1] = (int32_t)(er & 0xFFFFFF);
    }
    heap_push(H, end, ((er >> 24) + (el >> 24)) << 24 | next);
    next++;
  }
  const int32_t root = (int32_t)(H[1] & 0xFFFFFF);
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
  int32_t id = root;
  uint64_t bits = 0;
  bool bad = false;
  for (;;) {
    while (id >= (int32_t)k) {
      if (len >= 64) {  // children at depth len + 1: codes longer than 64 bits are refused
        bad = true;
        break;
      }
      const int32_t* kd = J.kids + 2 * (size_t)(id - (int32_t)k);
      const uint32_t ls = ++counter;
      put_le(L + (size_t)w * slot, ls, w);
      stk_id[sp] = kd[1];
      stk_slot[sp] = slot;
      stk_len[sp] = len + 1;
      stk_bits[sp] = bits << 1 | 1;
      sp++;
      id = kd[0];
      slot = ls;
      len++;
      bits <<= 1;
    }
    if (bad) break;
    const uint32_t s = syms[id];
    put_le(C + 4 * (size_t)slot, s, 4);
    T[slot] = 1;
    J.code[s] = len ? bits << (64 - len) : 0;
    J.len[s] = (uint8_t)len;
    if (len > max_len) max_len = len;
    if (!sp) break;
    sp--;
    const uint32_t rs = ++counter;
    put_le(R + (size_t)w * stk_slot[sp], rs, w);
    id = stk_id[sp];
    slot = rs;
    len = stk_len[sp];
    bits = stk_bits[sp];
  }
  info[0] = N;
  info[1] = tree_bytes;
  info[2] = max_len;
  info[3] = bad ? 4 : (counter + 1 == N ? 0 : 5);
}

void build_trees_device(Ctx& ctx, const TreeJob* d_jobs, int ntrees) {
  if (ntrees <= 0) return;
  k_huff_tree<<<ntrees, 64, 0, ctx.stream()>>>(d_jobs);
  CPSZ_HIP(hipGetLastError());
}

}  // namespace cpsz
