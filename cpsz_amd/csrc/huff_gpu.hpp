// GPU Huffman stage shared by the 2D and 3D codecs.
//
// Symbols are produced by a "source" that exposes one chunk as items in
// stream order (`get(chunk, item, sym[])`), so the block-major
// (2D) or predictable-vertex (3D) orders of the reference stream are walked in
// place without materializing a permuted copy. One wave takes one chunk, its
// lanes consecutive items.
//
//   histogram : LDS-privatized window + global atomics outside the window
//   encode    : per-chunk bit counts -> exclusive scan -> each wave packs its
//               chunk's codes MSB-first in LDS (lane offsets from a wave
//               prefix sum) and stores 32-bit big-endian words coalesced; only
//               the first and last (shared) word of a chunk use atomicOr
//   decode    : self-synchronizing parallel decode of the single unchunked
//               bitstream (speculative segment decode, fix-up rounds until every
//               segment starts on its predecessor's exit, scan, final decode)
#pragma once
#include <hip/hip_runtime.h>

#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/transform_iterator.hpp>
#include <cstring>
#include <functional>
#include <string>
#include <type_traits>
#include <vector>

#include "common.hpp"
#include "ctx.hpp"
#include "huff_host.hpp"

namespace cpsz {

constexpr int kLutBits = 11;
constexpr int kLut2Bits = 12;  // width cap of the level-2 decode tables
constexpr int kFixPerCheck = 2;   // fix-up passes queued per host convergence check
constexpr uint64_t kDecWarm = 384;  // bits decoded before a segment in the speculative pass

__device__ __forceinline__ uint32_t bswap32(uint32_t x) { return __builtin_bswap32(x); }

// --------------------------------------------------------------- histogram --
// One wave per chunk (sources as for the encode kernels below), grid-stride
// over chunks; symbols in [win_lo, win_lo + win_n) count in LDS, the rest in
// global memory directly.
template <class Src>
__global__ void __launch_bounds__(256) k_hist(Src src, size_t nchunks, uint32_t* __restrict__ hist,
                                              uint32_t win_lo, uint32_t win_n);

// ------------------------------------------------------------------ encode --
// The encode-side kernels run one wave per chunk (4 chunks per 256-thread
// block). A source exposes its chunk as items taken by lanes in stream order:
//   Src::kMaxItems            items per chunk at most
//   Src::kPer                 symbols per item at most
//   int get(c, idx, sym[])    the symbols of item idx, in stream order, and
//                             their count; idx in [0, kMaxItems), 0 symbols
//                             past the chunk's end (branch-free, so that a
//                             wave issues every round's loads together)
constexpr int kEncWaves = 4;
constexpr int kEncThreads = kEncWaves * 64;  // block size of the chunk-bits and encode kernels

inline size_t encode_blocks(size_t nchunks) { return (nchunks + kEncWaves - 1) / kEncWaves; }

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  return x;
}

// A source may also carry the stream's unpredictable values (kUnpred): the
// values at its symbol-0 items, gathered in stream order while encoding.
template <class S, class = void>
struct HasUnpred : std::false_type {};
template <class S>
struct HasUnpred<S, std::void_t<decltype(S::kUnpred)>> : std::bool_constant<S::kUnpred> {};

// zc (optional): the chunk's count of symbol 0 (the 2D codes' unpredictable
// marker) for the caller's offsets of the gathered values
template <class Src>
__device__ __forceinline__ void chunk_bits_wave(const Src& src, size_t c, const uint8_t* __restrict__ len,
                                                uint64_t* __restrict__ bits, uint32_t* __restrict__ zc = nullptr) {
  constexpr int kRounds = (Src::kMaxItems + 63) / 64;
  const int lane = threadIdx.x & 63;
  // every round's symbols are fetched before any is used (overlapped latency)
  uint32_t sym[kRounds][Src::kPer];
  int ns[kRounds];
#pragma unroll
  for (int r = 0; r < kRounds; r++) ns[r] = src.get(c, r * 64 + lane, sym[r]);
  uint32_t b = 0;
#pragma unroll
  for (int r = 0; r < kRounds; r++)
#pragma unroll
    for (int q = 0; q < Src::kPer; q++)
      if (q < ns[r]) b += len[sym[r][q]];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) b += __shfl_xor(b, o);
  if (lane == 0) bits[c] = b;
  if (zc) {
    uint32_t z = 0;
#pragma unroll
    for (int r = 0; r < kRounds; r++)
#pragma unroll
      for (int q = 0; q < Src::kPer; q++) z += (uint32_t)__popcll(__ballot(q < ns[r] && sym[r][q] == 0));
    if (lane == 0) zc[c] = z;
  }
}

template <class Src>
__global__ void __launch_bounds__(kEncThreads) k_chunk_bits(Src src, size_t nchunks,
                                                    const uint8_t* __restrict__ len,
                                                    uint64_t* __restrict__ bits) {
  const size_t c = (size_t)xcd_block(blockIdx.x, gridDim.x) * kEncWaves + (threadIdx.x >> 6);
  if (c >= nchunks) return;  // whole wave
  chunk_bits_wave(src, c, len, bits);
}

// ---- many streams per launch: a job table sorted by first block; block b
// belongs to the job with the largest blk0 <= b.
template <class Job>
__device__ __forceinline__ int job_of(const Job* __restrict__ jobs, int nj, uint32_t b) {
  int lo = 0, hi = nj - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].blk0 <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <class Src>
struct EncJob {
  Src src;
  size_t nchunks;
  const uint64_t* code;  // dense codebook of this stream's tree
  const uint8_t* len;
  const uint64_t* off;   // global exclusive scan of chunk bits, at this stream's first chunk
  uint64_t* bits;        // chunk bits, at this stream's first chunk
  uint32_t* words;       // this stream's zero-filled word array
  uint32_t blk0;         // first block (kEncWaves chunks per block)
  uint32_t* zc;          // optional: per-chunk symbol-0 counts (with the bits)
};

template <class Src>
__global__ void __launch_bounds__(kEncThreads) k_chunk_bits_multi(const EncJob<Src>* __restrict__ jobs, int nj) {
  const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
  const EncJob<Src>& J = jobs[job_of(jobs, nj, b)];
  const size_t c = (size_t)(b - J.blk0) * kEncWaves + (threadIdx.x >> 6);
  if (c >= J.nchunks) return;
  chunk_bits_wave(J.src, c, J.len, J.bits, J.zc);
}

template <class Src>
__device__ __forceinline__ void hist_block(const Src& src, size_t nchunks, uint32_t* __restrict__ hist,
                                           uint32_t win_lo, uint32_t win_n, uint32_t blk, uint32_t nblk,
                                           uint32_t* sh) {
  constexpr int kRounds = (Src::kMaxItems + 63) / 64;
  for (uint32_t i = threadIdx.x; i < win_n; i += blockDim.x) sh[i] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  // symbol 0 (the unpredictable marker of the 2D code streams, often frequent)
  // counts in a register when it lies outside the window: one global atomic
  // per wave instead of one per occurrence on a single address
  const bool zero_out = win_lo > 0;
  uint32_t zeros = 0;
  const uint32_t waves = blockDim.x >> 6;
  for (size_t c = (size_t)blk * waves + (threadIdx.x >> 6); c < nchunks; c += (size_t)nblk * waves) {
    uint32_t sym[kRounds][Src::kPer];
    int ns[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; r++) ns[r] = src.get(c, r * 64 + lane, sym[r]);
#pragma unroll
    for (int r = 0; r < kRounds; r++)
#pragma unroll
      for (int q = 0; q < Src::kPer; q++) {
        if (q < ns[r]) {
          const uint32_t o = sym[r][q] - win_lo;
          if (o < win_n) atomicAdd(&sh[o], 1u);
          else if (zero_out && sym[r][q] == 0) zeros++;
          else atomicAdd(&hist[sym[r][q]], 1u);
        }
      }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) zeros += __shfl_xor(zeros, o);
  if (lane == 0 && zeros) atomicAdd(&hist[0], zeros);
  __syncthreads();
  // blocks start their flush at different bins so that their global atomics
  // on the popular bins do not all arrive together
  const uint32_t rot = (uint32_t)(((uint64_t)blk * 2654435761u) % win_n);
  for (uint32_t i = threadIdx.x; i < win_n; i += blockDim.x) {
    const uint32_t b = i + rot < win_n ? i + rot : i + rot - win_n;
    if (sh[b]) atomicAdd(&hist[win_lo + b], sh[b]);
  }
}

template <class Src>
__global__ void __launch_bounds__(256) k_hist(Src src, size_t nchunks, uint32_t* __restrict__ hist,
                                              uint32_t win_lo, uint32_t win_n) {
  extern __shared__ uint32_t sh[];
  hist_block(src, nchunks, hist, win_lo, win_n, xcd_block(blockIdx.x, gridDim.x), gridDim.x, sh);
}

template <class Src>
struct HistJob {
  Src src;
  size_t nchunks;
  uint32_t* hist;
  uint32_t blk0, nblk;
};

template <class Src>
__global__ void __launch_bounds__(256) k_hist_multi(const HistJob<Src>* __restrict__ jobs, int nj, uint32_t win_lo,
                                                    uint32_t win_n) {
  extern __shared__ uint32_t sh[];
  const HistJob<Src>& J = jobs[job_of(jobs, nj, blockIdx.x)];
  hist_block(J.src, J.nchunks, J.hist, win_lo, win_n, blockIdx.x - J.blk0, J.nblk, sh);
}

// Writes the MSB-aligned code cw (L bits) at bit p of an MSB-first word array.
__device__ __forceinline__ void put_bits(uint32_t* st, uint32_t p, uint64_t cw, uint32_t L) {
  const uint32_t wi = p >> 5, o = p & 31;
  atomicOr(&st[wi], (uint32_t)(cw >> (32 + o)));
  if (L > 32 - o) atomicOr(&st[wi + 1], (uint32_t)((cw << (32 - o)) >> 32));
  if (L > 64 - o) atomicOr(&st[wi + 2], (uint32_t)((cw << (64 - o)) >> 32));  // o > 0 here
}

// Each wave assembles its chunk's bits in LDS (lanes place consecutive
// symbols at offsets from a wave prefix sum) and then stores the chunk's
// words with consecutive lanes on consecutive words. The words a chunk
// shares with its neighbours (a partial first or last word) are merged with
// atomicOr into the zero-filled output; the others are plain stores. b0 is
// the chunk's first bit within its stream. Called by whole blocks: the
// block's waves (adjacent chunks, whose rows share cache lines) stay in step
// at its barriers, which measured faster than letting them drift apart.
template <class Src>
__device__ __forceinline__ void encode_wave(const Src& src, size_t c, bool live, uint64_t b0, uint64_t nbits,
                                            const uint64_t* __restrict__ code, const uint8_t* __restrict__ len,
                                            uint32_t* __restrict__ words, uint32_t* st) {
  const int lane = threadIdx.x & 63;
  const uint32_t sh = (uint32_t)(b0 & 31), nw = (uint32_t)((sh + nbits + 31) >> 5);
  for (uint32_t k = lane; k < nw; k += 64) st[k] = 0;
  __syncthreads();
  if (live) {
    constexpr int kRounds = (Src::kMaxItems + 63) / 64;
    // every round's symbols, lengths and codes are fetched before any is used
    uint32_t sym[kRounds][Src::kPer], L[kRounds][Src::kPer];
    uint64_t cw[kRounds][Src::kPer];
    int ns[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; r++) ns[r] = src.get(c, r * 64 + lane, sym[r]);
    if constexpr (HasUnpred<Src>::value) {
      if (src.unp) {
        // the chunk's unpredictable values, at its offset among the stream's
        uint64_t o = src.uoff[c] - src.uoff[0];
        const uint64_t below = (1ull << lane) - 1;
#pragma unroll
        for (int r = 0; r < kRounds; r++) {
          const bool z = ns[r] && sym[r][0] == 0;
          const uint64_t m = __ballot(z);
          if (z) src.unp[o + __popcll(m & below)] = src.x[src.elem(c, r * 64 + lane)];
          o += __popcll(m);
        }
      }
    }
#pragma unroll
    for (int r = 0; r < kRounds; r++)
#pragma unroll
      for (int q = 0; q < Src::kPer; q++) {
        const uint32_t l = len[sym[r][q]];  // symbols are always valid table indices
        L[r][q] = q < ns[r] ? l : 0u;
        cw[r][q] = code[sym[r][q]];
      }
    uint32_t pos = sh;
#pragma unroll
    for (int r = 0; r < kRounds; r++) {
      uint32_t b = 0;
#pragma unroll
      for (int q = 0; q < Src::kPer; q++) b += L[r][q];
      const uint32_t incl = wave_incl_scan(b);
      uint32_t p = pos + incl - b;
#pragma unroll
      for (int q = 0; q < Src::kPer; q++) {
        if (L[r][q]) put_bits(st, p, cw[r][q], L[r][q]);
        p += L[r][q];
      }
      pos += __shfl(incl, 63);
    }
  }
  __syncthreads();
  if (live) {
    const size_t w0 = b0 >> 5;
    const bool tail_shared = ((b0 + nbits) & 31) != 0;
    for (uint32_t k = lane; k < nw; k += 64) {
      const uint32_t word = bswap32(st[k]);
      if ((k == 0 && sh != 0) || (k == nw - 1 && tail_shared)) atomicOr(&words[w0 + k], word);
      else words[w0 + k] = word;
    }
  }
}

template <class Src>
__global__ void __launch_bounds__(kEncThreads) k_encode(Src src, size_t nchunks,
                                                const uint64_t* __restrict__ code,
                                                const uint8_t* __restrict__ len,
                                                const uint64_t* __restrict__ off,
                                                const uint64_t* __restrict__ bits,
                                                uint32_t* __restrict__ words) {
  constexpr int kStage = Src::kMaxItems * Src::kPer * 2 + 2;  // 64-bit codes at most
  __shared__ uint32_t stage[kEncWaves][kStage];
  const int wv = threadIdx.x >> 6;
  const size_t c = (size_t)xcd_block(blockIdx.x, gridDim.x) * kEncWaves + wv;
  const bool live = c < nchunks;  // uniform per wave
  encode_wave(src, c, live, live ? off[c] : 0, live ? bits[c] : 0, code, len, words, stage[wv]);
}

template <class Src>
__global__ void __launch_bounds__(kEncThreads) k_encode_multi(const EncJob<Src>* __restrict__ jobs, int nj) {
  constexpr int kStage = Src::kMaxItems * Src::kPer * 2 + 2;
  __shared__ uint32_t stage[kEncWaves][kStage];
  const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
  const EncJob<Src>& J = jobs[job_of(jobs, nj, b)];
  if (!J.words) return;  // a stream without an output array is skipped (whole block)
  const int wv = threadIdx.x >> 6;
  const size_t c = (size_t)(b - J.blk0) * kEncWaves + wv;
  const bool live = c < J.nchunks;
  const uint64_t base = J.off[0];
  encode_wave(J.src, c, live, live ? J.off[c] - base : 0, live ? J.bits[c] : 0, J.code, J.len, J.words,
              stage[wv]);
}

// One Huffman tree built on the device (huff_tree.hip) from a histogram's
// present symbols: syms/freqs + *base hold *count (symbol, count) pairs in
// increasing symbol order. Outputs: the serialized tree (13 * (2k - 1) + 1
// bytes at most), the dense codebook code[sym] / len[sym], and info = {node
// count, tree bytes, longest code, error (0 = none)}. heap (max_leaves + 1
// words) is scratch for trees too large for LDS (more than 8190 leaves).
// Errors: 1 empty, 2 more leaves than max_leaves (or beyond 2^23), 3 a symbol
// at or beyond state_num, 4 a code longer than 64 bits, 5 internal.
struct TreeJob {
  const uint32_t* syms;
  const uint32_t* freqs;
  const uint32_t* count;
  const uint64_t* base;
  uint32_t state_num, max_leaves;
  uint64_t* heap;
  uint8_t* tree;
  uint64_t* code;
  uint8_t* len;
  uint64_t* info;
};
__global__ void k_huff_tree(const TreeJob* jobs);
void build_trees_device(Ctx& ctx, const TreeJob* d_jobs, int ntrees);

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v);
__global__ void k_scatter_codebook(const uint32_t* syms, const uint64_t* codes, const uint8_t* lens,
                                   int k, uint64_t* code, uint8_t* len);

// ------------------------------------------------------------------ decode --
struct DecTree {
  const uint32_t* lut;    // 2^kLutBits entries
  const uint32_t* lut2;   // level-2 tables (huff_build_lut)
  const uint32_t* left;   // node arrays (preorder numbering)
  const uint32_t* right;
  const uint32_t* sym;
  const uint8_t* leaf;
  uint32_t nodes;
};

// MSB-first reader over big-endian 32-bit words; the word array needs at
// least two zero words past the last bit. Words [lbase, lbase + lcount) are
// also staged in LDS (lw) and read from there: a refill is then an LDS read,
// not a global-memory round trip on the decoder's serial dependency chain.
typedef __attribute__((address_space(3))) const uint32_t lds_u32;  // an LDS word

struct BitReader {
  const uint32_t* w;
  lds_u32* lw;
  uint64_t lbase;
  uint32_t lcount;
  uint64_t buf;   // upcoming bits, MSB-aligned
  int nb;         // valid bits in buf (> 32 after every refill)
  uint64_t widx;  // next word to load
  uint64_t pos;   // bit position of buf's MSB
  __device__ __forceinline__ uint32_t word(uint64_t k) const {
    const uint64_t d = k - lbase;
    uint32_t x;
    if (d < lcount) x = lw[d];
    else x = w[k];
    return bswap32(x);
  }
  __device__ __forceinline__ void init(const uint32_t* __restrict__ words, uint64_t p,
                                       const uint32_t* staged = nullptr, uint64_t sbase = 0, uint32_t scount = 0) {
    w = words;
    lw = (lds_u32*)staged;
    lbase = sbase;
    lcount = scount;
    pos = p;
    widx = p >> 5;
    const int s = (int)(p & 31);
    buf = (((uint64_t)word(widx) << 32) | word(widx + 1)) << s;
    nb = 64 - s;
    widx += 2;
    refill();
  }
  __device__ __forceinline__ void refill() {
    if (nb <= 32) {
      buf |= (uint64_t)word(widx++) << (32 - nb);
      nb += 32;
    }
  }
  __device__ __forceinline__ uint32_t peek() const { return (uint32_t)(buf >> 32); }
  __device__ __forceinline__ void skip(int L) {
    buf <<= L;
    nb -= L;
    pos += (uint64_t)L;
    refill();
  }
};

// A block's 256 segments of 256 bits span 2048 words; the few words past them
// cover the last segment's overrun to its first boundary past the end.
constexpr uint32_t kStageWords = 2048 + 8;

// Stages the words of block segments [b0, b0 + 256) (segment size S bits) into
// sw; returns how many were staged (0 unless S == 256). Callers synchronize.
__device__ __forceinline__ uint32_t stage_words(uint32_t* sw, const uint32_t* __restrict__ w, uint64_t nbits,
                                                uint32_t S, uint32_t b0, uint64_t& lbase) {
  lbase = (uint64_t)b0 * (S / 32);
  if (S != 256) return 0;
  const uint64_t avail = (nbits + 31) / 32 + 8;  // words allocated, zero padded
  uint32_t cnt = kStageWords;
  if (lbase + cnt > avail) cnt = avail > lbase ? (uint32_t)(avail - lbase) : 0u;
  for (uint32_t k = threadIdx.x; k < cnt; k += blockDim.x) sw[k] = w[lbase + k];
  return cnt;
}

// Decodes one symbol; returns it or 0xFFFFFFFF on a bad code. `lut` is the
// level-1 table (in LDS): sym << 8 | length for short codes, off << 8 | 0x80 |
// K for prefixes continued by a level-2 table (t.lut2, global memory) over the
// next K bits; codes longer still continue bit by bit down the tree.
__device__ __forceinline__ uint32_t dec_one(const DecTree& t, const uint32_t* lut, BitReader& br) {
  const uint32_t e = lut[br.peek() >> (32 - kLutBits)];
  const uint32_t L = e & 0xFF;
  if (L && !(L & 0x80u)) {
    br.skip((int)L);
    return e >> 8;
  }
  uint32_t node = e >> 8;
  br.skip(kLutBits);
  if (L & 0x80u) {
    const int K = (int)(L & 0x7F);
    const uint32_t e2 = t.lut2[node + (br.peek() >> (32 - K))];
    const uint32_t L2 = e2 & 0xFF;
    if (L2) {
      br.skip((int)L2);
      return e2 >> 8;
    }
    node = e2 >> 8;
    br.skip(K);
  }
  for (int d = 0; d < 64; d++) {
    const uint32_t b = br.peek() >> 31;
    br.skip(1);
    node = b ? t.right[node] : t.left[node];
    if (node == 0 || node >= t.nodes) return 0xFFFFFFFFu;
    if (t.leaf[node]) return t.sym[node];
  }
  return 0xFFFFFFFFu;
}

// Multi-code steps for the passes that only locate code boundaries: for each
// kLutBits-bit window, the number n of whole codes (each at most kLutBits
// bits, decoded greedily from the window's start) and their total length,
// stored as n << 8 | length (0 when the first code is longer). Built from the
// level-1 table in LDS by the whole block; callers synchronize before use.
__device__ __forceinline__ void build_mstep(const uint32_t* slut, uint16_t* ms) {
  constexpr uint32_t kMask = (1u << kLutBits) - 1;
  for (uint32_t w = threadIdx.x; w <= kMask; w += blockDim.x) {
    uint32_t pos = 0, n = 0;
    for (;;) {
      const uint32_t e = slut[(w << pos) & kMask];
      const uint32_t L = e & 0xFF;
      if (!L || (L & 0x80u) || pos + L > (uint32_t)kLutBits) break;
      pos += L;
      n++;
    }
    ms[w] = (uint16_t)(n << 8 | pos);
  }
}

// Decodes from br's position up to the first code boundary at or after
// `stop`, adding the number of codes to c: a window's run of short codes at a
// time while the run ends at or before `stop` (so no boundary at or after
// `stop` is stepped over), one code at a time otherwise. False on a bad code.
__device__ __forceinline__ bool dec_until(const DecTree& t, const uint32_t* slut, const uint16_t* ms, BitReader& br,
                                          uint64_t stop, uint32_t& c) {
  while (br.pos < stop) {
    const uint32_t m = ms[br.peek() >> (32 - kLutBits)];
    const uint32_t len = m & 0xFF;
    if (len && br.pos + len <= stop) {
      br.skip((int)len);
      c += m >> 8;
      continue;
    }
    if (dec_one(t, slut, br) == 0xFFFFFFFFu) return false;
    c++;
  }
  return true;
}

// Multi-code decode for streams whose symbols fit a byte: for each window,
// up to 3 whole short codes: length | n << 4 | sym0 << 8 | sym1 << 16 | sym2 << 24
// (n = 0 when the first code is longer than the window).
__device__ __forceinline__ void build_mdec8(const uint32_t* slut, uint32_t* md) {
  constexpr uint32_t kMask = (1u << kLutBits) - 1;
  for (uint32_t w = threadIdx.x; w <= kMask; w += blockDim.x) {
    uint32_t pos = 0, n = 0, syms = 0;
    while (n < 3) {
      const uint32_t e = slut[(w << pos) & kMask];
      const uint32_t L = e & 0xFF;
      if (!L || (L & 0x80u) || pos + L > (uint32_t)kLutBits || (e >> 8) > 0xFFu) break;
      syms |= (e >> 8) << (8 * n);
      pos += L;
      n++;
    }
    md[w] = pos | n << 4 | syms << 8;
  }
}

// Phase 1 / fix-up: decode from start to the first boundary >= seg_end.
__global__ void __launch_bounds__(256) k_dec_spec(DecTree t, const uint32_t* __restrict__ w, uint64_t nbits,
                                                  uint32_t S, uint32_t nseg, const uint64_t* __restrict__ prev_exit,
                           const uint64_t* __restrict__ prev_start, uint64_t* __restrict__ start,
                           uint64_t* __restrict__ exitp, uint32_t* __restrict__ cnt,
                           int* __restrict__ changed, int* __restrict__ err, int first);

__global__ void k_dec_check(uint32_t nseg, const uint64_t* exitp, const uint64_t* start,
                            int* changed);

// One bitstream of a multi-stream decode (every stream's segments in one grid):
// its grid blocks are [blk0, blk0 + nblk); start/ex are the ping-pong buffers of
// the passes, cnt the per-segment counts, off their exclusive scan (nseg + 1
// entries) and bsum the per-block sums, then block offsets.
struct MDecJob {
  DecTree t;
  const uint32_t* w;
  uint64_t nbits, n;
  uint32_t S, nseg, blk0, nblk;
  uint32_t warm;  // bits decoded before a segment in the speculative pass
  uint64_t* start[2];
  uint64_t* ex[2];
  uint32_t* cnt;
  uint64_t* off;
  uint64_t* bsum;
};
template <class Sink>
struct MDecJobS {
  MDecJob j;
  Sink sink;
};
__global__ void k_dec_spec_multi(const MDecJob* jobs, const uint32_t* blk0, int nj, int cur, int first, int* err);
__global__ void k_dec_check_multi(const MDecJob* jobs, const uint32_t* blk0, int nj, int cur, int* changed);
__global__ void k_dec_bsum_multi(const MDecJob* jobs, const uint32_t* blk0, int nj);
__global__ void k_dec_bscan_multi(const MDecJob* jobs, uint64_t* totals);
__global__ void k_dec_off_multi(const MDecJob* jobs, const uint32_t* blk0, int nj);

// Phase 3: decode every segment into its final position. The symbols of a
// block's 256 consecutive segments are contiguous in the output: they are
// gathered in LDS and written out by consecutive threads (coalesced), instead
// of each thread storing its own run one symbol at a time (64 scattered
// 2-byte stores per instruction, which reach HBM as ~15x the bytes). A block
// whose run exceeds the staging buffer writes directly. Sink::kBytes (1 or 2)
// is the width of a staged symbol.
template <int kBytes>
struct DecStage {
  using T = typename std::conditional<kBytes == 1, uint8_t, uint16_t>::type;
  static constexpr uint32_t kCap = kBytes == 1 ? 32768 : 12288;  // symbols per block
};

template <class Sink>
__device__ __forceinline__ void dec_write_block(const DecTree& t, const uint32_t* __restrict__ w, uint64_t nbits,
                                                uint32_t S, uint32_t nseg, const uint64_t* __restrict__ start,
                                                const uint64_t* __restrict__ exitp,
                                                const uint64_t* __restrict__ off, uint64_t n, const Sink& sink,
                                                uint32_t blk, uint32_t* slut, uint32_t* sw,
                                                typename DecStage<Sink::kBytes>::T* stage, uint32_t* md) {
  using Stage = DecStage<Sink::kBytes>;
  for (int k = threadIdx.x; k < (1 << kLutBits); k += blockDim.x) slut[k] = t.lut[k];
  uint64_t lbase;
  const uint32_t lcount = stage_words(sw, w, nbits, S, blk * blockDim.x, lbase);
  __syncthreads();
  if (Sink::kBytes == 1) {
    build_mdec8(slut, md);
    __syncthreads();
  }
  const uint32_t b0 = blk * blockDim.x;
  const uint32_t b1 = min(b0 + blockDim.x, nseg);  // off[] holds nseg + 1 entries
  const uint64_t o_begin = min(off[b0], n), o_end = min(off[b1], n);
  const bool staged = o_end - o_begin <= (uint64_t)Stage::kCap;
  const uint32_t i = b0 + threadIdx.x;
  if (i < nseg) {
    BitReader br;
    br.init(w, start[i], sw, lbase, lcount);
    const uint64_t end = exitp[i];
    uint64_t o = off[i];
    if (staged) {
      if (Sink::kBytes == 1) {
        // up to 3 codes per step while they all start before `end` and fit
        while (br.pos < end && o < o_end) {
          const uint32_t m = md[br.peek() >> (32 - kLutBits)];
          const uint32_t len = m & 0xF, k = (m >> 4) & 3;
          if (k && br.pos + len <= end && o + k <= o_end) {
            typename Stage::T* d = stage + (o - o_begin);
            d[0] = (typename Stage::T)(m >> 8);
            if (k > 1) d[1] = (typename Stage::T)(m >> 16);
            if (k > 2) d[2] = (typename Stage::T)(m >> 24);
            o += k;
            br.skip((int)len);
          } else {
            stage[o++ - o_begin] = (typename Stage::T)dec_one(t, slut, br);
          }
        }
      } else {
        while (br.pos < end && o < o_end) stage[o++ - o_begin] = (typename Stage::T)dec_one(t, slut, br);
      }
    } else {
      while (br.pos < end && o < n) sink.put(o++, dec_one(t, slut, br));
    }
  }
  if (!staged) return;  // uniform across the block
  __syncthreads();
  const uint32_t cnt = (uint32_t)(o_end - o_begin);
  for (uint32_t x = threadIdx.x; x < cnt; x += blockDim.x) sink.put(o_begin + x, stage[x]);
}

template <class Sink>
__global__ void __launch_bounds__(256) k_dec_write(DecTree t, const uint32_t* __restrict__ w, uint64_t nbits,
                                                   uint32_t S, uint32_t nseg, const uint64_t* __restrict__ start,
                                                   const uint64_t* __restrict__ exitp,
                                                   const uint64_t* __restrict__ off, uint64_t n,
                                                   Sink sink) {
  __shared__ uint32_t slut[1 << kLutBits];
  __shared__ uint32_t sw[kStageWords];
  __shared__ typename DecStage<Sink::kBytes>::T stage[DecStage<Sink::kBytes>::kCap];
  __shared__ uint32_t md[Sink::kBytes == 1 ? (1 << kLutBits) : 1];
  dec_write_block(t, w, nbits, S, nseg, start, exitp, off, n, sink, blockIdx.x, slut, sw, stage, md);
}

template <class Sink>
__global__ void __launch_bounds__(256) k_dec_write_multi(const MDecJobS<Sink>* __restrict__ jobs,
                                                         const uint32_t* __restrict__ blk0, int nj, int cur) {
  __shared__ uint32_t slut[1 << kLutBits];
  __shared__ uint32_t sw[kStageWords];
  __shared__ typename DecStage<Sink::kBytes>::T stage[DecStage<Sink::kBytes>::kCap];
  __shared__ uint32_t md[Sink::kBytes == 1 ? (1 << kLutBits) : 1];
  int lo = 0, hi = nj - 1;  // the job owning this block
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (blk0[mid] <= blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  const MDecJob& J = jobs[lo].j;
  dec_write_block(J.t, J.w, J.nbits, J.S, J.nseg, J.start[cur], J.ex[cur], J.off, J.n, jobs[lo].sink,
                  blockIdx.x - J.blk0, slut, sw, stage, md);
}

template <class Sink>
__global__ void k_dec_const(uint64_t n, uint32_t s, Sink sink) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    sink.put(i, s);
}

// --------------------------------------------------------- host helpers ----
// Exclusive scan of u64 counts (rocPRIM), writing the total to *d_total.
void scan_u64(Ctx& ctx, const uint64_t* in, uint64_t* out, size_t n, uint64_t* d_total);
void scan_u32_to_u64(Ctx& ctx, const uint32_t* in, uint64_t* out, size_t n, uint64_t* d_total);

// Uploads an encode table into dense device tables (code, len) over nbins.
void upload_codebook(Ctx& ctx, const HuffEncodeTable& t, uint64_t* d_code, uint8_t* d_len,
                     const char* tag);

struct DeviceDecoder {
  DecTree tree{};
  bool single = false;
  uint32_t single_sym = 0;
};
// Parses a serialized tree and uploads the decode tables (tag names the buffers).
// The host half of a decoder: the parsed tree and the level-1 / level-2
// tables. prepare_decoder makes no HIP call, so it can run on another thread
// while the device works.
struct HostDecoder {
  HuffDecodeTree d;
  std::vector<uint32_t> lut, pfx;
  size_t n2 = 0;  // level-2 entries
};
void prepare_decoder(const uint8_t* tree, size_t tree_size, size_t node_count, HostDecoder& h);
DeviceDecoder upload_decoder(Ctx& ctx, const HostDecoder& h, const char* tag);
DeviceDecoder upload_decoder(Ctx& ctx, const uint8_t* tree, size_t tree_size, size_t node_count,
                             const char* tag);

// Many decoders in one packed copy and one level-2 fill launch (out[k] for
// hd[k]); `tag` names the buffers.
// pool (optional): host workers that share the table copies (the caller must
// not itself be running on that pool)
class WorkerPool;
void upload_decoders(Ctx& ctx, const HostDecoder* const* hd, int n, DeviceDecoder* out, const char* tag,
                     WorkerPool* pool = nullptr, int workers = 0);

// Bitstreams at arbitrary byte offsets copied into aligned word arrays with
// zero slack (the decoder's input form), all in one launch. Fills dst/nwords.
struct BitsCopy {
  const uint8_t* src;
  uint64_t nbytes;
  uint32_t* dst;
  uint64_t nwords;
};
constexpr uint32_t kStageBlockWords = 2048;
void stage_bits_multi(Ctx& ctx, BitsCopy* c, int n, const char* tag);

// One bitstream to decode: n symbols of words already copied to aligned
// device words (padded with at least 8 zero bytes) into `sink`; `tag` names
// its buffers.
template <class Sink>
struct DecJob {
  const DeviceDecoder* dd;
  const uint32_t* words;
  uint64_t nbytes, n;
  Sink sink;
  const char* tag;
  // speculative warm-up in bits (0: kDecWarm); codes of a few bits that
  // resynchronize slowly want a longer one (multi-code steps keep it cheap)
  uint32_t warm = 0;
};

// Decodes several independent bitstreams in lockstep on the context's stream:
// their passes are queued together, so one host check serves every stream of
// a round, and one read of the totals serves all. `overlap` (may be empty) is
// host work run once the first round is queued, before the first wait.
template <class Sink>
void decode_streams(Ctx& ctx, DecJob<Sink>* jobs, int nj, const std::function<void()>& overlap = {}) {
  hipStream_t st = ctx.stream();
  struct State {
    uint64_t nbits = 0;
    uint32_t S = 256, nseg = 0;
    int blocks = 0, cur = 0;
    bool active = false, done = false;
    uint64_t *start[2] = {nullptr, nullptr}, *ex[2] = {nullptr, nullptr}, *off = nullptr;
    uint32_t* cnt = nullptr;
    int* flags = nullptr;
  };
  std::vector<State> sv(nj);
  bool overlapped = false;
  auto run_overlap = [&] {
    if (!overlapped && overlap) {
      overlapped = true;
      overlap();
    }
  };
  for (int k = 0; k < nj; k++) {
    const DecJob<Sink>& j = jobs[k];
    State& s = sv[k];
    if (j.n == 0) continue;
    if (j.dd->single) {
      k_dec_const<<<1024, 256, 0, st>>>(j.n, j.dd->single_sym, j.sink);
      CPSZ_HIP(hipGetLastError());
      continue;
    }
    s.nbits = j.nbytes * 8;
    CPSZ_REQUIRE(s.nbits > 0, kErrStream, "empty huffman bitstream for a multi-symbol tree");
    while (div_up(s.nbits, s.S) > (1u << 21)) s.S *= 2;
    s.nseg = (uint32_t)div_up(s.nbits, s.S);
    const std::string t(j.tag);
    s.start[0] = ctx.buf<uint64_t>((t + ".st0").c_str(), s.nseg);
    s.start[1] = ctx.buf<uint64_t>((t + ".st1").c_str(), s.nseg);
    s.ex[0] = ctx.buf<uint64_t>((t + ".ex0").c_str(), s.nseg);
    s.ex[1] = ctx.buf<uint64_t>((t + ".ex1").c_str(), s.nseg);
    s.cnt = ctx.buf<uint32_t>((t + ".cnt").c_str(), s.nseg);
    s.off = ctx.buf<uint64_t>((t + ".off").c_str(), s.nseg + 1);
    s.flags = ctx.buf<int>((t + ".flags").c_str(), 2);
    s.blocks = (int)div_up(s.nseg, 256);
    s.active = true;
    CPSZ_HIP(hipMemsetAsync(s.flags, 0, 2 * sizeof(int), st));
    k_dec_spec<<<s.blocks, 256, 0, st>>>(j.dd->tree, j.words, s.nbits, s.S, s.nseg, nullptr, nullptr, s.start[0],
                                         s.ex[0], s.cnt, s.flags, s.flags + 1, 1);
    CPSZ_HIP(hipGetLastError());
  }
  int* hflags = (int*)ctx.host_buf("dec.hflags", 2 * sizeof(int) * (size_t)(nj > 0 ? nj : 1));
  for (int round = 0; round < 4096; round++) {
    bool any = false;
    for (int k = 0; k < nj; k++) {
      State& s = sv[k];
      if (!s.active || s.done) continue;
      const DecJob<Sink>& j = jobs[k];
      // kFixPerCheck fix-up rounds per host check
      for (int r = 0; r < kFixPerCheck; r++) {
        k_dec_spec<<<s.blocks, 256, 0, st>>>(j.dd->tree, j.words, s.nbits, s.S, s.nseg, s.ex[s.cur],
                                             s.start[s.cur], s.start[s.cur ^ 1], s.ex[s.cur ^ 1], s.cnt, s.flags,
                                             s.flags + 1, 0);
        s.cur ^= 1;
      }
      CPSZ_HIP(hipMemsetAsync(s.flags, 0, sizeof(int), st));
      k_dec_check<<<s.blocks, 256, 0, st>>>(s.nseg, s.ex[s.cur], s.start[s.cur], s.flags);
      CPSZ_HIP(hipMemcpyAsync(hflags + 2 * k, s.flags, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
      any = true;
    }
    run_overlap();
    if (!any) break;
    CPSZ_HIP(hipStreamSynchronize(st));
    for (int k = 0; k < nj; k++) {
      State& s = sv[k];
      if (!s.active || s.done) continue;
      CPSZ_REQUIRE(hflags[2 * k + 1] == 0, kErrStream, "invalid huffman code in stream");
      if (!hflags[2 * k]) s.done = true;
    }
  }
  run_overlap();
  // a stream whose segment starts have not settled within the round limit
  // would decode from inconsistent starts: refuse it
  for (int k = 0; k < nj; k++)
    CPSZ_REQUIRE(!sv[k].active || sv[k].done, kErrStream,
                 "huffman decode did not converge (bitstream does not resynchronize)");
  uint64_t* htot = (uint64_t*)ctx.host_buf("dec.htotals", sizeof(uint64_t) * (size_t)(nj > 0 ? nj : 1));
  bool any = false;
  for (int k = 0; k < nj; k++) {
    State& s = sv[k];
    if (!s.active) continue;
    scan_u32_to_u64(ctx, s.cnt, s.off, s.nseg, s.off + s.nseg);
    CPSZ_HIP(hipMemcpyAsync(htot + k, s.off + s.nseg, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    any = true;
  }
  if (!any) return;
  CPSZ_HIP(hipStreamSynchronize(st));
  for (int k = 0; k < nj; k++)
    if (sv[k].active)
      CPSZ_REQUIRE(htot[k] >= jobs[k].n, kErrStream, "huffman bitstream holds fewer symbols than expected");
  for (int k = 0; k < nj; k++) {
    const State& s = sv[k];
    if (!s.active) continue;
    const DecJob<Sink>& j = jobs[k];
    k_dec_write<<<s.blocks, 256, 0, st>>>(j.dd->tree, j.words, s.nbits, s.S, s.nseg, s.start[s.cur], s.ex[s.cur],
                                          s.off, j.n, j.sink);
    CPSZ_HIP(hipGetLastError());
  }
}

// Decodes many independent bitstreams with every stream's segments in one grid
// per pass: one launch per pass for all streams, one host check per round
// (the streams whose segments have settled pass through unchanged), a
// three-launch scan of all streams' counts and one write launch. `tag` names
// the buffers (one arena for all streams).
template <class Sink>
void decode_streams_multi(Ctx& ctx, const DecJob<Sink>* jobs, int nj, const char* tag) {
  hipStream_t st = ctx.stream();
  std::vector<int> act;
  for (int k = 0; k < nj; k++) {
    const DecJob<Sink>& j = jobs[k];
    if (j.n == 0) continue;
    if (j.dd->single) {
      k_dec_const<<<1024, 256, 0, st>>>(j.n, j.dd->single_sym, j.sink);
      CPSZ_HIP(hipGetLastError());
      continue;
    }
    CPSZ_REQUIRE(j.nbytes > 0, kErrStream, "empty huffman bitstream for a multi-symbol tree");
    act.push_back(k);
  }
  const int na = (int)act.size();
  if (!na) return;
  std::vector<MDecJobS<Sink>> h(na);
  std::vector<uint32_t> blk0(na + 1, 0);
  size_t arena = 0;
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  for (int a = 0; a < na; a++) {
    const DecJob<Sink>& j = jobs[act[a]];
    MDecJob& m = h[a].j;
    m = MDecJob{};
    m.t = j.dd->tree;
    m.w = j.words;
    m.nbits = j.nbytes * 8;
    m.n = j.n;
    m.S = 256;
    m.warm = j.warm ? j.warm : (uint32_t)kDecWarm;
    while (div_up(m.nbits, (uint64_t)m.S) > (1u << 21)) m.S *= 2;
    m.nseg = (uint32_t)div_up(m.nbits, (uint64_t)m.S);
    m.nblk = (uint32_t)div_up(m.nseg, 256u);
    m.blk0 = blk0[a];
    CPSZ_REQUIRE((uint64_t)blk0[a] + m.nblk < (1ull << 31), kErrUnsupported, "huffman decode: too many segments");
    blk0[a + 1] = blk0[a] + m.nblk;
    h[a].sink = j.sink;
    arena += 4 * al(8 * (size_t)m.nseg) + al(4 * (size_t)m.nseg) + al(8 * ((size_t)m.nseg + 1)) +
             al(8 * (size_t)m.nblk);
  }
  const uint32_t nblocks = blk0[na];
  uint8_t* d = ctx.buf<uint8_t>((std::string(tag) + ".arena").c_str(), arena);
  for (int a = 0; a < na; a++) {
    MDecJob& m = h[a].j;
    auto take = [&](size_t bytes) {
      uint8_t* q = d;
      d += al(bytes);
      return q;
    };
    m.start[0] = (uint64_t*)take(8 * (size_t)m.nseg);
    m.start[1] = (uint64_t*)take(8 * (size_t)m.nseg);
    m.ex[0] = (uint64_t*)take(8 * (size_t)m.nseg);
    m.ex[1] = (uint64_t*)take(8 * (size_t)m.nseg);
    m.cnt = (uint32_t*)take(4 * (size_t)m.nseg);
    m.off = (uint64_t*)take(8 * ((size_t)m.nseg + 1));
    m.bsum = (uint64_t*)take(8 * (size_t)m.nblk);
  }
  // job table, block starts and flags travel in one copy each way
  const size_t tb = al(na * sizeof(MDecJobS<Sink>)), bb = al((na + 1) * 4), fb = al(2 * na * 4), ob = na * 8;
  const std::string t(tag);
  uint8_t* hh = (uint8_t*)ctx.host_buf((t + ".tab").c_str(), tb + bb + fb + ob);
  std::memcpy(hh, h.data(), na * sizeof(MDecJobS<Sink>));
  std::memcpy(hh + tb, blk0.data(), (na + 1) * 4);
  uint8_t* dt = ctx.buf<uint8_t>((t + ".tab").c_str(), tb + bb + fb + ob);
  CPSZ_HIP(hipMemcpyAsync(dt, hh, tb + bb, hipMemcpyHostToDevice, st));
  const MDecJobS<Sink>* djs = (const MDecJobS<Sink>*)dt;
  // MDecJobS<Sink> starts with its MDecJob, so a stride-aware view is needed
  // for the Sink-agnostic kernels: they take an MDecJob table of their own
  std::vector<MDecJob> hj(na);
  for (int a = 0; a < na; a++) hj[a] = h[a].j;
  MDecJob* dj = ctx.buf<MDecJob>((t + ".jobs").c_str(), na);
  uint8_t* hjb = (uint8_t*)ctx.host_buf((t + ".jobs").c_str(), na * sizeof(MDecJob));
  std::memcpy(hjb, hj.data(), na * sizeof(MDecJob));
  CPSZ_HIP(hipMemcpyAsync(dj, hjb, na * sizeof(MDecJob), hipMemcpyHostToDevice, st));
  const uint32_t* dblk0 = (const uint32_t*)(dt + tb);
  int* changed = (int*)(dt + tb + bb);
  int* err = changed + na;
  uint64_t* totals = (uint64_t*)(dt + tb + bb + fb);
  int* hflags = (int*)(hh + tb + bb);
  uint64_t* htot = (uint64_t*)(hh + tb + bb + fb);
  CPSZ_HIP(hipMemsetAsync(err, 0, na * 4, st));
  k_dec_spec_multi<<<nblocks, 256, 0, st>>>(dj, dblk0, na, 0, 1, err);
  CPSZ_HIP(hipGetLastError());
  ctx.mark((t + ".spec").c_str());
  int cur = 0;
  bool settled = false;
  int rounds = 0;
  for (int round = 0; round < 4096 && !settled; round++, rounds++) {
    for (int r = 0; r < kFixPerCheck; r++) {
      k_dec_spec_multi<<<nblocks, 256, 0, st>>>(dj, dblk0, na, cur, 0, err);
      cur ^= 1;
    }
    CPSZ_HIP(hipMemsetAsync(changed, 0, na * 4, st));
    k_dec_check_multi<<<nblocks, 256, 0, st>>>(dj, dblk0, na, cur, changed);
    CPSZ_HIP(hipGetLastError());
    CPSZ_HIP(hipMemcpyAsync(hflags, changed, 2 * na * 4, hipMemcpyDeviceToHost, st));
    CPSZ_HIP(hipStreamSynchronize(st));
    settled = true;
    for (int a = 0; a < na; a++) {
      CPSZ_REQUIRE(hflags[na + a] == 0, kErrStream, "invalid huffman code in stream");
      if (hflags[a]) settled = false;
    }
  }
  // a stream whose segment starts have not settled within the round limit
  // would decode from inconsistent starts: refuse it
  CPSZ_REQUIRE(settled, kErrStream, "huffman decode did not converge (bitstream does not resynchronize)");
  ctx.mark((t + ".fix" + std::to_string(rounds)).c_str());
  k_dec_bsum_multi<<<nblocks, 256, 0, st>>>(dj, dblk0, na);
  k_dec_bscan_multi<<<na, 256, 0, st>>>(dj, totals);
  k_dec_off_multi<<<nblocks, 256, 0, st>>>(dj, dblk0, na);
  CPSZ_HIP(hipGetLastError());
  CPSZ_HIP(hipMemcpyAsync(htot, totals, na * 8, hipMemcpyDeviceToHost, st));
  CPSZ_HIP(hipStreamSynchronize(st));
  for (int a = 0; a < na; a++)
    CPSZ_REQUIRE(htot[a] >= h[a].j.n, kErrStream, "huffman bitstream holds fewer symbols than expected");
  ctx.mark((t + ".scan").c_str());
  k_dec_write_multi<Sink><<<nblocks, 256, 0, st>>>(djs, dblk0, na, cur);
  CPSZ_HIP(hipGetLastError());
  ctx.mark((t + ".write").c_str());
}

// Decodes n symbols of one bitstream (see decode_streams).
template <class Sink>
void decode_stream(Ctx& ctx, const DeviceDecoder& dd, const uint32_t* d_words, uint64_t nbytes, uint64_t n,
                   Sink sink, const char* tag, const std::function<void()>& overlap = {}) {
  DecJob<Sink> j{&dd, d_words, nbytes, n, sink, tag};
  decode_streams(ctx, &j, 1, overlap);
}

}  // namespace cpsz
