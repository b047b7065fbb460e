#include "huff_gpu.hpp"
#include "pool.hpp"

#include <cstring>
#include <functional>

namespace cpsz {

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_scatter_codebook(const uint32_t* syms, const uint64_t* codes, const uint8_t* lens,
                                   int k, uint64_t* code, uint8_t* len) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k) return;
  code[syms[i]] = codes[i];
  len[syms[i]] = lens[i];
}

__device__ __forceinline__ void dec_resync(const DecTree& t, const uint32_t* slut, const uint16_t* ms,
                                           const uint32_t* __restrict__ w, const uint32_t* sw, uint64_t lbase,
                                           uint32_t lcount, uint64_t end, uint64_t s0, uint64_t os, uint64_t oe,
                                           uint32_t oc, uint64_t& ex, uint32_t& c, int* __restrict__ err);

// Block `blk` of one stream's speculative / fix-up pass (k_dec_spec runs one
// stream's blocks, k_dec_spec_multi every stream's blocks in one grid).
__device__ __forceinline__ void dec_spec_block(const DecTree& t, const uint32_t* __restrict__ w, uint64_t nbits,
                                               uint32_t S, uint32_t nseg, const uint64_t* __restrict__ prev_exit,
                                               const uint64_t* __restrict__ prev_start,
                                               uint64_t* __restrict__ start, uint64_t* __restrict__ exitp,
                                               uint32_t* __restrict__ cnt, int* __restrict__ err, int first,
                                               uint32_t blk, uint32_t* slut, uint32_t* sw, uint16_t* ms) {
  const uint32_t i = blk * blockDim.x + threadIdx.x;
  bool work = i < nseg;
  uint64_t s0 = 0;
  if (work) {
    if (first) {
      s0 = (uint64_t)i * S;
    } else {
      s0 = i == 0 ? 0 : prev_exit[i - 1];
      if (s0 == prev_start[i]) {  // unchanged start: keep the previous result (and count)
        start[i] = s0;
        exitp[i] = prev_exit[i];
        work = false;
      }
    }
  }
  if (!__syncthreads_or(work)) return;
  for (int k = threadIdx.x; k < (1 << kLutBits); k += blockDim.x) slut[k] = t.lut[k];
  uint64_t lbase;
  const uint32_t lcount = stage_words(sw, w, nbits, S, blk * blockDim.x, lbase);
  __syncthreads();
  build_mstep(slut, ms);
  __syncthreads();
  if (!work) return;
  uint64_t end = (uint64_t)(i + 1) * S;
  if (end > nbits) end = nbits;
  BitReader br;
  uint32_t c = 0;
  if (first) {
    // Warm-up: start kDecWarm bits before the segment and take as its start
    // the first code boundary at or after i * S. A prefix-code path usually
    // resyncs within a few code words, so this path is often already the true
    // one there: its start then equals the previous segment's exit, and the
    // fix-up passes skip the segment.
    br.init(w, s0 - (s0 < kDecWarm ? s0 : kDecWarm), sw, lbase, lcount);
    uint32_t skipped = 0;
    if (!dec_until(t, slut, ms, br, s0, skipped)) *err = 1;
    s0 = br.pos;
    if (!dec_until(t, slut, ms, br, end, c)) *err = 1;
    start[i] = s0;
    exitp[i] = br.pos;
    cnt[i] = c;
    return;
  }
  // New start: decode the new path and the previous pass's path of this
  // segment in lockstep (always advancing the one behind) until they reach
  // the same code boundary; from there both paths coincide, so the previous
  // pass's exit and the rest of its count carry over. Prefix codes resync
  // within a few code words, so a fix-up costs a few symbols, not a segment.
  uint64_t ex = 0;
  dec_resync(t, slut, ms, w, sw, lbase, lcount, end, s0, prev_start[i], prev_exit[i], cnt[i], ex, c, err);
  start[i] = s0;
  exitp[i] = ex;
  cnt[i] = c;
}

__global__ void __launch_bounds__(256) k_dec_spec(DecTree t, const uint32_t* __restrict__ w, uint64_t nbits,
                                                  uint32_t S, uint32_t nseg,
                                                  const uint64_t* __restrict__ prev_exit,
                                                  const uint64_t* __restrict__ prev_start,
                                                  uint64_t* __restrict__ start, uint64_t* __restrict__ exitp,
                                                  uint32_t* __restrict__ cnt, int* __restrict__ changed,
                                                  int* __restrict__ err, int first) {
  __shared__ uint32_t slut[1 << kLutBits];
  __shared__ uint32_t sw[kStageWords];
  __shared__ uint16_t ms[1 << kLutBits];
  (void)changed;
  dec_spec_block(t, w, nbits, S, nseg, prev_exit, prev_start, start, exitp, cnt, err, first, blockIdx.x, slut, sw,
                 ms);
}

// ---- block-settled passes (the multi-stream decoder). A pass leaves every
// block internally consistent: each segment but the block's first starts at
// its predecessor's exit. So only the links between blocks need host rounds.

// Decodes from s0 along the new path and along the old path of the segment
// (start os, exit oe, count oc) in lockstep, always advancing the one behind,
// until the two meet at a code boundary (from there they coincide, so the old
// exit and the rest of its count carry over) or the new path reaches `end`.
__device__ __forceinline__ void dec_resync(const DecTree& t, const uint32_t* slut, const uint16_t* ms,
                                           const uint32_t* __restrict__ w, const uint32_t* sw, uint64_t lbase,
                                           uint32_t lcount, uint64_t end, uint64_t s0, uint64_t os, uint64_t oe,
                                           uint32_t oc, uint64_t& ex, uint32_t& c, int* __restrict__ err) {
  BitReader br, old;
  br.init(w, s0, sw, lbase, lcount);
  old.init(w, os, sw, lbase, lcount);
  uint32_t cn = 0, co = 0;
  // A bounded lockstep: paths that have not met after kLock codes rarely meet
  // soon (the codes resynchronize slowly there), so the new path then runs
  // alone to `end` with multi-code steps, and its own count is the count.
  constexpr int kLock = 48;
  int steps = 0;
  for (;;) {
    if (br.pos >= end) {
      ex = br.pos;
      break;
    }
    if (br.pos == old.pos) {
      ex = oe;
      cn += oc - co;
      break;
    }
    if (++steps > kLock) {
      if (!dec_until(t, slut, ms, br, end, cn)) *err = 1;
      ex = br.pos;
      break;
    }
    const bool adv_new = br.pos < old.pos;
    if (dec_one(t, slut, adv_new ? br : old) == 0xFFFFFFFFu) {
      *err = 1;
      ex = br.pos;
      break;
    }
    cn += adv_new;
    co += !adv_new;
  }
  c = cn;
}

// Settles a block: while some segment does not start at its predecessor's
// exit, every such segment re-decodes from that exit (in parallel: a serial
// fix of one segment at a time would leave all but one lane of the wave idle).
// After iteration k the first k + 1 segments are final (the first one is
// given), so blockDim iterations suffice; a segment that adopts a still-wrong
// predecessor exit is corrected in a later iteration.
// lex: blockDim LDS words.
__device__ __forceinline__ void dec_settle(const DecTree& t, const uint32_t* slut, const uint16_t* ms,
                                           const uint32_t* __restrict__ w, const uint32_t* sw, uint64_t lbase,
                                           uint32_t lcount, uint64_t end, bool valid, uint64_t& st, uint64_t& ex,
                                           uint32_t& c, uint64_t* lex, int* __restrict__ err) {
  for (uint32_t it = 0; it < blockDim.x; it++) {
    __syncthreads();
    lex[threadIdx.x] = ex;
    __syncthreads();
    bool redo = false;
    uint64_t want = 0;
    if (valid && threadIdx.x > 0) {
      want = lex[threadIdx.x - 1];
      redo = want != st;
    }
    if (!__syncthreads_or(redo)) break;
    if (redo) {
      uint64_t e2;
      uint32_t c2;
      dec_resync(t, slut, ms, w, sw, lbase, lcount, end, want, st, ex, c, e2, c2, err);
      st = want;
      ex = e2;
      c = c2;
    }
  }
}

__device__ __forceinline__ void dec_pass_block(const DecTree& t, const uint32_t* __restrict__ w, uint64_t nbits,
                                               uint32_t S, uint32_t nseg, const uint64_t* __restrict__ prev_exit,
                                               const uint64_t* __restrict__ prev_start,
                                               uint64_t* __restrict__ start, uint64_t* __restrict__ exitp,
                                               uint32_t* __restrict__ cnt, int* __restrict__ err, int first,
                                               uint32_t blk, uint32_t* slut, uint32_t* sw, uint64_t* lex,
                                               uint16_t* ms, uint64_t warm) {
  const uint32_t i = blk * blockDim.x + threadIdx.x;
  const bool valid = i < nseg;
  uint64_t st = 0, ex = 0, s0 = 0;
  uint32_t c = 0;
  bool work;
  if (first) {
    work = valid;
  } else {
    // only a block whose first segment gets a new start changes
    if (valid) {
      st = prev_start[i];
      ex = prev_exit[i];
      c = cnt[i];
    }
    if (threadIdx.x == 0 && valid) s0 = i == 0 ? 0 : prev_exit[i - 1];
    work = threadIdx.x == 0 && valid && s0 != st;
    if (!__syncthreads_or(work)) {
      if (valid) {
        start[i] = st;
        exitp[i] = ex;
      }
      return;
    }
  }
  for (int k = threadIdx.x; k < (1 << kLutBits); k += blockDim.x) slut[k] = t.lut[k];
  uint64_t lbase;
  const uint32_t lcount = stage_words(sw, w, nbits, S, blk * blockDim.x, lbase);
  __syncthreads();
  build_mstep(slut, ms);
  __syncthreads();
  uint64_t end = (uint64_t)(i + 1) * S;
  if (end > nbits) end = nbits;
  if (first && valid) {
    // Warm-up: start `warm` bits before the segment and take as its start
    // the first code boundary at or after i * S. A prefix-code path usually
    // resyncs within a few code words, so this path is often already the true
    // one there.
    const uint64_t a = (uint64_t)i * S;
    BitReader br;
    br.init(w, a - (a < warm ? a : warm), sw, lbase, lcount);
    uint32_t skipped = 0;
    if (!dec_until(t, slut, ms, br, a, skipped)) *err = 1;
    st = br.pos;
    if (!dec_until(t, slut, ms, br, end, c)) *err = 1;
    ex = br.pos;
  } else if (work) {
    uint64_t e2;
    uint32_t c2;
    dec_resync(t, slut, ms, w, sw, lbase, lcount, end, s0, st, ex, c, e2, c2, err);
    st = s0;
    ex = e2;
    c = c2;
  }
  dec_settle(t, slut, ms, w, sw, lbase, lcount, end, valid, st, ex, c, lex, err);
  if (valid) {
    start[i] = st;
    exitp[i] = ex;
    cnt[i] = c;
  }
}

// ---- multi-stream passes: block b belongs to the job j with
// blk0[j] <= b < blk0[j + 1]; `cur` selects the ping-pong buffers.
__device__ __forceinline__ int mdec_job(const uint32_t* __restrict__ blk0, int nj, uint32_t b) {
  int lo = 0, hi = nj - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (blk0[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ void __launch_bounds__(256) k_dec_spec_multi(const MDecJob* __restrict__ jobs,
                                                        const uint32_t* __restrict__ blk0, int nj, int cur,
                                                        int first, int* __restrict__ err) {
  __shared__ uint32_t slut[1 << kLutBits];
  __shared__ uint32_t sw[kStageWords];
  __shared__ uint64_t lex[256];
  __shared__ uint16_t ms[1 << kLutBits];
  const int j = mdec_job(blk0, nj, blockIdx.x);
  const MDecJob& J = jobs[j];
  if (first)
    dec_pass_block(J.t, J.w, J.nbits, J.S, J.nseg, nullptr, nullptr, J.start[0], J.ex[0], J.cnt, err + j, 1,
                   blockIdx.x - blk0[j], slut, sw, lex, ms, J.warm);
  else
    dec_pass_block(J.t, J.w, J.nbits, J.S, J.nseg, J.ex[cur], J.start[cur], J.start[cur ^ 1], J.ex[cur ^ 1],
                   J.cnt, err + j, 0, blockIdx.x - blk0[j], slut, sw, lex, ms, J.warm);
}

// changed[j] = 1 when a segment of job j does not start at its predecessor's exit
__global__ void __launch_bounds__(256) k_dec_check_multi(const MDecJob* __restrict__ jobs,
                                                         const uint32_t* __restrict__ blk0, int nj, int cur,
                                                         int* __restrict__ changed) {
  const int j = mdec_job(blk0, nj, blockIdx.x);
  const MDecJob& J = jobs[j];
  const uint32_t i = (blockIdx.x - blk0[j]) * blockDim.x + threadIdx.x;
  if (i >= J.nseg) return;
  const uint64_t want = i == 0 ? 0 : J.ex[cur][i - 1];
  if (J.start[cur][i] != want) changed[j] = 1;
}

// Exclusive scan of the per-segment counts of every job, in three passes:
// block sums, a scan of each job's block sums (one workgroup per job, which
// also writes the job's total), then the in-block scans.
__global__ void __launch_bounds__(256) k_dec_bsum_multi(const MDecJob* __restrict__ jobs,
                                                        const uint32_t* __restrict__ blk0, int nj) {
  __shared__ uint64_t red[256];
  const int j = mdec_job(blk0, nj, blockIdx.x);
  const MDecJob& J = jobs[j];
  const uint32_t b = blockIdx.x - blk0[j];
  const uint32_t i = b * blockDim.x + threadIdx.x;
  red[threadIdx.x] = i < J.nseg ? J.cnt[i] : 0;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) J.bsum[b] = red[0];
}

__global__ void __launch_bounds__(256) k_dec_bscan_multi(const MDecJob* __restrict__ jobs, uint64_t* __restrict__ totals) {
  __shared__ uint64_t sc[256];
  const MDecJob& J = jobs[blockIdx.x];
  uint64_t carry = 0;
  for (uint32_t base = 0; base < J.nblk; base += 256) {
    const uint32_t k = base + threadIdx.x;
    const uint64_t v = k < J.nblk ? J.bsum[k] : 0;
    sc[threadIdx.x] = v;
    __syncthreads();
    for (int s = 1; s < 256; s <<= 1) {  // inclusive Hillis-Steele scan
      const uint64_t a = (int)threadIdx.x >= s ? sc[threadIdx.x - s] : 0;
      __syncthreads();
      sc[threadIdx.x] += a;
      __syncthreads();
    }
    if (k < J.nblk) J.bsum[k] = carry + sc[threadIdx.x] - v;
    carry += sc[255];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    J.off[J.nseg] = carry;
    totals[blockIdx.x] = carry;
  }
}

__global__ void __launch_bounds__(256) k_dec_off_multi(const MDecJob* __restrict__ jobs,
                                                       const uint32_t* __restrict__ blk0, int nj) {
  __shared__ uint64_t sc[256];
  const int j = mdec_job(blk0, nj, blockIdx.x);
  const MDecJob& J = jobs[j];
  const uint32_t b = blockIdx.x - blk0[j];
  const uint32_t i = b * blockDim.x + threadIdx.x;
  const uint64_t v = i < J.nseg ? J.cnt[i] : 0;
  sc[threadIdx.x] = v;
  __syncthreads();
  for (int s = 1; s < 256; s <<= 1) {
    const uint64_t a = (int)threadIdx.x >= s ? sc[threadIdx.x - s] : 0;
    __syncthreads();
    sc[threadIdx.x] += a;
    __syncthreads();
  }
  if (i < J.nseg) J.off[i] = J.bsum[b] + sc[threadIdx.x] - v;
}

__global__ void k_dec_check(uint32_t nseg, const uint64_t* exitp, const uint64_t* start,
                            int* changed) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nseg) return;
  const uint64_t want = i == 0 ? 0 : exitp[i - 1];
  if (start[i] != want) *changed = 1;
}

// Level-2 entry for the K bits u below `node`: sym << 8 | length of the leaf
// reached, or the node reached after K bits, node << 8.
__host__ __device__ inline uint32_t lut2_entry(const uint32_t* left, const uint32_t* right, const uint32_t* sym,
                                               const uint8_t* leaf, uint32_t node, int K, uint32_t u) {
  for (int len = 0; len < K; len++) {
    node = ((u >> (K - 1 - len)) & 1) ? right[node] : left[node];
    if (leaf[node]) return (sym[node] << 8) | (uint32_t)(len + 1);
  }
  return node << 8;
}

// Level-2 decode tables: entry e of prefix p (pfx triples (node, K, off),
// offsets ascending) walks the K bits of e - off down from node.
__global__ void k_fill_lut2(const uint32_t* __restrict__ left, const uint32_t* __restrict__ right,
                            const uint32_t* __restrict__ sym, const uint8_t* __restrict__ leaf,
                            const uint32_t* __restrict__ pfx, uint32_t npfx, uint32_t n2,
                            uint32_t* __restrict__ lut2) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n2) return;
  uint32_t lo = 0, hi = npfx - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (pfx[3 * mid + 2] <= e) lo = mid;
    else hi = mid - 1;
  }
  lut2[e] = lut2_entry(left, right, sym, leaf, pfx[3 * lo], (int)pfx[3 * lo + 1], e - pfx[3 * lo + 2]);
}

void decode_host(const uint8_t* tree, size_t tree_size, size_t node_count, const uint8_t* bits, size_t nbytes,
                 size_t n, uint32_t* out) {
  HuffDecodeTree d;
  CPSZ_REQUIRE(huff_parse_tree(tree, tree_size, node_count, d), kErrStream, "malformed huffman tree");
  if (d.root_leaf) {
    for (size_t i = 0; i < n; i++) out[i] = d.root_sym;
    return;
  }
  std::vector<uint32_t> lut, pfx;
  const size_t n2 = huff_build_lut(d, kLutBits, kLut2Bits, lut, pfx);
  std::vector<uint32_t> lut2(n2);
  for (size_t p = 0; p < pfx.size() / 3; p++)
    for (uint32_t u = 0; u < (1u << pfx[3 * p + 1]); u++)
      lut2[pfx[3 * p + 2] + u] = lut2_entry(d.left.data(), d.right.data(), d.sym.data(), d.leaf.data(),
                                            pfx[3 * p], (int)pfx[3 * p + 1], u);
  const uint64_t nbits = 8 * (uint64_t)nbytes;
  uint64_t pos = 0;
  auto peek = [&](int k) -> uint32_t {  // next k <= 32 bits, zeros past the end
    uint32_t v = 0;
    for (int b = 0; b < k; b++) {
      const uint64_t q = pos + b;
      v = (v << 1) | (q < nbits ? (bits[q >> 3] >> (7 - (q & 7))) & 1u : 0u);
    }
    return v;
  };
  for (size_t i = 0; i < n; i++) {
    const uint32_t e = lut[peek(kLutBits)];
    const uint32_t L = e & 0xFF;
    if (L && !(L & 0x80u)) {
      pos += L;
      out[i] = e >> 8;
      continue;
    }
    uint32_t node = e >> 8;
    pos += kLutBits;
    if (L & 0x80u) {
      const int K = (int)(L & 0x7F);
      const uint32_t e2 = lut2[node + peek(K)];
      if (e2 & 0xFF) {
        pos += e2 & 0xFF;
        out[i] = e2 >> 8;
        continue;
      }
      node = e2 >> 8;
      pos += K;
    }
    for (;;) {
      CPSZ_REQUIRE(pos < nbits, kErrStream, "huffman bitstream exhausted");
      node = peek(1) ? d.right[node] : d.left[node];
      pos++;
      if (d.leaf[node]) break;
    }
    out[i] = d.sym[node];
  }
}

namespace {
__global__ void k_total(const uint64_t* out, const uint64_t* in, size_t n, uint64_t* total) {
  *total = n ? out[n - 1] + in[n - 1] : 0;
}
__global__ void k_total32(const uint64_t* out, const uint32_t* in, size_t n, uint64_t* total) {
  *total = n ? out[n - 1] + in[n - 1] : 0;
}
struct U32ToU64 {
  __device__ uint64_t operator()(uint32_t x) const { return x; }
};
}  // namespace

void scan_u64(Ctx& ctx, const uint64_t* in, uint64_t* out, size_t n, uint64_t* d_total) {
  if (n == 0) {
    CPSZ_HIP(hipMemsetAsync(d_total, 0, 8, ctx.stream()));
    return;
  }
  size_t tmp = 0;
  CPSZ_HIP(rocprim::exclusive_scan(nullptr, tmp, in, out, (uint64_t)0, n, rocprim::plus<uint64_t>(),
                                   ctx.stream()));
  void* d_tmp = ctx.raw_buf("scan.tmp", tmp);
  CPSZ_HIP(rocprim::exclusive_scan(d_tmp, tmp, in, out, (uint64_t)0, n, rocprim::plus<uint64_t>(),
                                   ctx.stream()));
  k_total<<<1, 1, 0, ctx.stream()>>>(out, in, n, d_total);
  CPSZ_HIP(hipGetLastError());
}

void scan_u32_to_u64(Ctx& ctx, const uint32_t* in, uint64_t* out, size_t n, uint64_t* d_total) {
  if (n == 0) {
    CPSZ_HIP(hipMemsetAsync(d_total, 0, 8, ctx.stream()));
    return;
  }
  auto it = rocprim::make_transform_iterator(in, U32ToU64());
  size_t tmp = 0;
  CPSZ_HIP(rocprim::exclusive_scan(nullptr, tmp, it, out, (uint64_t)0, n, rocprim::plus<uint64_t>(),
                                   ctx.stream()));
  void* d_tmp = ctx.raw_buf("scan.tmp", tmp);
  CPSZ_HIP(rocprim::exclusive_scan(d_tmp, tmp, it, out, (uint64_t)0, n, rocprim::plus<uint64_t>(),
                                   ctx.stream()));
  k_total32<<<1, 1, 0, ctx.stream()>>>(out, in, n, d_total);
  CPSZ_HIP(hipGetLastError());
}

void upload_codebook(Ctx& ctx, const HuffEncodeTable& t, uint64_t* d_code, uint8_t* d_len,
                     const char* tag) {
  const int k = (int)t.syms.size();
  std::string s(tag);
  const size_t bytes = k * (4 + 8 + 1);
  uint8_t* h = (uint8_t*)ctx.host_buf((s + ".cb").c_str(), bytes);
  std::memcpy(h, t.syms.data(), 4 * k);
  std::memcpy(h + 4 * k, t.codes.data(), 8 * k);
  std::memcpy(h + 12 * k, t.lens.data(), k);
  uint8_t* d = ctx.buf<uint8_t>((s + ".cb").c_str(), bytes + 16);
  // keep the u64 array 8-byte aligned on the device: place codes first
  uint64_t* dc = (uint64_t*)d;
  uint32_t* ds = (uint32_t*)(d + 8 * (size_t)k);
  uint8_t* dl = d + 12 * (size_t)k;
  CPSZ_HIP(hipMemcpyAsync(dc, h + 4 * k, 8 * (size_t)k, hipMemcpyHostToDevice, ctx.stream()));
  CPSZ_HIP(hipMemcpyAsync(ds, h, 4 * (size_t)k, hipMemcpyHostToDevice, ctx.stream()));
  CPSZ_HIP(hipMemcpyAsync(dl, h + 12 * k, (size_t)k, hipMemcpyHostToDevice, ctx.stream()));
  k_scatter_codebook<<<(int)div_up(k, 256), 256, 0, ctx.stream()>>>(ds, dc, dl, k, d_code, d_len);
  CPSZ_HIP(hipGetLastError());
}

void prepare_decoder(const uint8_t* tree, size_t tree_size, size_t node_count, HostDecoder& h) {
  CPSZ_REQUIRE(huff_parse_tree(tree, tree_size, node_count, h.d), kErrStream,
               "malformed huffman tree in stream");
  h.n2 = h.d.root_leaf ? 0 : huff_build_lut(h.d, kLutBits, kLut2Bits, h.lut, h.pfx);
}

DeviceDecoder upload_decoder(Ctx& ctx, const HostDecoder& hd, const char* tag) {
  const HuffDecodeTree& d = hd.d;
  DeviceDecoder out;
  if (d.root_leaf) {
    out.single = true;
    out.single_sym = d.root_sym;
    return out;
  }
  const std::vector<uint32_t>& lut = hd.lut;
  const std::vector<uint32_t>& pfx = hd.pfx;
  const size_t n2 = hd.n2;
  const size_t N = d.node_count;
  const uint32_t npfx = (uint32_t)(pfx.size() / 3);
  std::string s(tag);
  const size_t lut_b = lut.size() * 4, pfx_b = pfx.size() * 4, arr_b = N * 4;
  const size_t bytes = lut_b + pfx_b + 3 * arr_b + N;
  uint8_t* h = (uint8_t*)ctx.host_buf((s + ".dec").c_str(), bytes);
  std::memcpy(h, lut.data(), lut_b);
  if (pfx_b) std::memcpy(h + lut_b, pfx.data(), pfx_b);
  uint8_t* q = h + lut_b + pfx_b;
  std::memcpy(q, d.left.data(), arr_b);
  std::memcpy(q + arr_b, d.right.data(), arr_b);
  std::memcpy(q + 2 * arr_b, d.sym.data(), arr_b);
  std::memcpy(q + 3 * arr_b, d.leaf.data(), N);
  uint8_t* dv = ctx.buf<uint8_t>((s + ".dec").c_str(), bytes);
  CPSZ_HIP(hipMemcpyAsync(dv, h, bytes, hipMemcpyHostToDevice, ctx.stream()));
  out.tree.lut = (const uint32_t*)dv;
  const uint32_t* dpfx = (const uint32_t*)(dv + lut_b);
  const uint8_t* dq = dv + lut_b + pfx_b;
  out.tree.left = (const uint32_t*)dq;
  out.tree.right = (const uint32_t*)(dq + arr_b);
  out.tree.sym = (const uint32_t*)(dq + 2 * arr_b);
  out.tree.leaf = dq + 3 * arr_b;
  uint32_t* lut2 = ctx.buf<uint32_t>((s + ".lut2").c_str(), n2 + 1);
  out.tree.lut2 = lut2;
  if (n2) {
    k_fill_lut2<<<(int)div_up(n2, 256), 256, 0, ctx.stream()>>>(out.tree.left, out.tree.right, out.tree.sym,
                                                                 out.tree.leaf, dpfx, npfx, (uint32_t)n2, lut2);
    CPSZ_HIP(hipGetLastError());
  }
  out.tree.nodes = (uint32_t)N;
  // The pinned staging buffer is reused by the next decoder upload of the same
  // tag only after the stream has consumed it (callers synchronize per call).
  return out;
}

// ---- many decoders at once: one packed host-to-device copy and one level-2
// fill launch for all of them
namespace {
struct Lut2Fill {
  const uint32_t *left, *right, *sym, *pfx;
  const uint8_t* leaf;
  uint32_t* lut2;
  uint32_t npfx, n2;
  uint64_t e0;  // first entry of this decoder in the launch
};

__global__ void k_fill_lut2_multi(const Lut2Fill* __restrict__ f, int nf, uint64_t total) {
  const uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (g >= total) return;
  int lo = 0, hi = nf - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (f[mid].e0 <= g) lo = mid;
    else hi = mid - 1;
  }
  const Lut2Fill& d = f[lo];
  const uint32_t e = (uint32_t)(g - d.e0);
  uint32_t a = 0, b = d.npfx - 1;
  while (a < b) {
    const uint32_t mid = (a + b + 1) >> 1;
    if (d.pfx[3 * mid + 2] <= e) a = mid;
    else b = mid - 1;
  }
  d.lut2[e] = lut2_entry(d.left, d.right, d.sym, d.leaf, d.pfx[3 * a], (int)d.pfx[3 * a + 1], e - d.pfx[3 * a + 2]);
}

__global__ void k_stage_bits_multi(const BitsCopy* __restrict__ c, const uint32_t* __restrict__ blk0, int nc) {
  int lo = 0, hi = nc - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (blk0[mid] <= blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  const BitsCopy& j = c[lo];
  const uint64_t k0 = (uint64_t)(blockIdx.x - blk0[lo]) * kStageBlockWords;
  for (uint32_t t = threadIdx.x; t < kStageBlockWords; t += blockDim.x) {
    const uint64_t k = k0 + t;
    if (k >= j.nwords) break;
    const uint64_t b = 4 * k;
    uint32_t v = 0;
    if (b + 4 <= j.nbytes) {
      v = (uint32_t)j.src[b] | (uint32_t)j.src[b + 1] << 8 | (uint32_t)j.src[b + 2] << 16 |
          (uint32_t)j.src[b + 3] << 24;
    } else {
      for (int q = 0; q < 4; q++)
        if (b + q < j.nbytes) v |= (uint32_t)j.src[b + q] << (8 * q);
    }
    j.dst[k] = v;
  }
}
}  // namespace

void upload_decoders(Ctx& ctx, const HostDecoder* const* hd, int n, DeviceDecoder* out, const char* tag,
                     WorkerPool* pool, int workers) {
  hipStream_t st = ctx.stream();
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  std::vector<size_t> base(n + 1, 0), l2base(n + 1, 0);
  for (int k = 0; k < n; k++) {
    const HostDecoder& h = *hd[k];
    const size_t N = h.d.root_leaf ? 0 : h.d.node_count;
    base[k + 1] = base[k] + (h.d.root_leaf ? 0 : al(h.lut.size() * 4 + h.pfx.size() * 4 + 13 * N));
    l2base[k + 1] = l2base[k] + (h.d.root_leaf ? 0 : al(4 * (h.n2 + 1)));
  }
  const std::string t(tag);
  const size_t nfill = sizeof(Lut2Fill) * (size_t)(n ? n : 1);
  uint8_t* hbuf = (uint8_t*)ctx.host_buf((t + ".h").c_str(), base[n] + nfill);
  uint8_t* dbuf = ctx.buf<uint8_t>((t + ".d").c_str(), base[n] + nfill);
  uint32_t* dl2 = ctx.buf<uint32_t>((t + ".lut2").c_str(), l2base[n] / 4 + 1);
  Lut2Fill* hf = (Lut2Fill*)(hbuf + base[n]);
  // a decoder's tables into its slot of the pinned staging buffer
  auto copy_one = [&](int k) {
    const HostDecoder& h = *hd[k];
    if (h.d.root_leaf) return;
    const size_t N = h.d.node_count, lut_b = h.lut.size() * 4, pfx_b = h.pfx.size() * 4, arr_b = N * 4;
    uint8_t* q = hbuf + base[k];
    std::memcpy(q, h.lut.data(), lut_b);
    if (pfx_b) std::memcpy(q + lut_b, h.pfx.data(), pfx_b);
    uint8_t* a = q + lut_b + pfx_b;
    std::memcpy(a, h.d.left.data(), arr_b);
    std::memcpy(a + arr_b, h.d.right.data(), arr_b);
    std::memcpy(a + 2 * arr_b, h.d.sym.data(), arr_b);
    std::memcpy(a + 3 * arr_b, h.d.leaf.data(), N);
  };
  if (pool) {
    // the copies of many decoders (megabytes) split over the host workers
    const int parts = std::max(1, std::min(workers, n));
    std::vector<std::function<void(Ctx&)>> jobs(parts);
    for (int j = 0; j < parts; j++)
      jobs[j] = [&, j](Ctx&) {
        for (int k = j; k < n; k += parts) copy_one(k);
      };
    pool_run(*pool, jobs, parts);
  }
  int nf = 0;
  uint64_t total2 = 0;
  for (int k = 0; k < n; k++) {
    const HostDecoder& h = *hd[k];
    DeviceDecoder& o = out[k];
    o = DeviceDecoder();
    if (h.d.root_leaf) {
      o.single = true;
      o.single_sym = h.d.root_sym;
      continue;
    }
    const size_t N = h.d.node_count, lut_b = h.lut.size() * 4, pfx_b = h.pfx.size() * 4, arr_b = N * 4;
    if (!pool) copy_one(k);
    const uint8_t* dq = dbuf + base[k];
    const uint8_t* da = dq + lut_b + pfx_b;
    o.tree.lut = (const uint32_t*)dq;
    o.tree.left = (const uint32_t*)da;
    o.tree.right = (const uint32_t*)(da + arr_b);
    o.tree.sym = (const uint32_t*)(da + 2 * arr_b);
    o.tree.leaf = da + 3 * arr_b;
    o.tree.lut2 = dl2 + l2base[k] / 4;
    o.tree.nodes = (uint32_t)N;
    if (h.n2) {
      Lut2Fill& f = hf[nf++];
      f.left = o.tree.left;
      f.right = o.tree.right;
      f.sym = o.tree.sym;
      f.leaf = o.tree.leaf;
      f.pfx = (const uint32_t*)(dq + lut_b);
      f.lut2 = dl2 + l2base[k] / 4;
      f.npfx = (uint32_t)(h.pfx.size() / 3);
      f.n2 = (uint32_t)h.n2;
      f.e0 = total2;
      total2 += h.n2;
    }
  }
  CPSZ_HIP(hipMemcpyAsync(dbuf, hbuf, base[n] + sizeof(Lut2Fill) * nf, hipMemcpyHostToDevice, st));
  if (total2) {
    k_fill_lut2_multi<<<(unsigned)div_up(total2, (uint64_t)256), 256, 0, st>>>((const Lut2Fill*)(dbuf + base[n]),
                                                                                nf, total2);
    CPSZ_HIP(hipGetLastError());
  }
}

void stage_bits_multi(Ctx& ctx, BitsCopy* c, int n, const char* tag) {
  hipStream_t st = ctx.stream();
  std::vector<uint32_t> blk0(n + 1, 0);
  size_t words = 0;
  for (int k = 0; k < n; k++) {
    c[k].nwords = div_up(c[k].nbytes, (uint64_t)4) + 8;  // zero slack past the last bit
    words += (c[k].nwords + 63) & ~(uint64_t)63;
  }
  const std::string t(tag);
  uint32_t* dw = ctx.buf<uint32_t>((t + ".words").c_str(), words + 1);
  size_t w0 = 0;
  for (int k = 0; k < n; k++) {
    c[k].dst = dw + w0;
    w0 += (c[k].nwords + 63) & ~(uint64_t)63;
    blk0[k + 1] = blk0[k] + (uint32_t)div_up(c[k].nwords, (uint64_t)kStageBlockWords);
  }
  const size_t tb = sizeof(BitsCopy) * n, bb = 4 * (size_t)(n + 1);
  uint8_t* h = (uint8_t*)ctx.host_buf((t + ".tab").c_str(), tb + bb);
  std::memcpy(h, c, tb);
  std::memcpy(h + tb, blk0.data(), bb);
  uint8_t* d = ctx.buf<uint8_t>((t + ".tab").c_str(), tb + bb);
  CPSZ_HIP(hipMemcpyAsync(d, h, tb + bb, hipMemcpyHostToDevice, st));
  if (blk0[n]) {
    k_stage_bits_multi<<<blk0[n], 256, 0, st>>>((const BitsCopy*)d, (const uint32_t*)(d + tb), n);
    CPSZ_HIP(hipGetLastError());
  }
}

DeviceDecoder upload_decoder(Ctx& ctx, const uint8_t* tree, size_t tree_size, size_t node_count,
                             const char* tag) {
  thread_local HostDecoder hd;  // keeps its allocations from call to call
  prepare_decoder(tree, tree_size, node_count, hd);
  return upload_decoder(ctx, hd, tag);
}

}  // namespace cpsz
