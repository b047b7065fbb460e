// GPU Huffman stage shared by the 2D and 3D codecs.
//
// Symbols are produced by a "source" functor that visits the symbols of one
// chunk in stream order (`each(chunk, f)`), so the block-major (2D) or
// predictable-vertex (3D) orders of the reference stream are walked in place
// without materializing a permuted copy.
//
//   histogram : LDS-privatized window + global atomics outside the window
//   encode    : per-chunk bit counts -> exclusive scan -> each chunk packs its
//               codes MSB-first into 32-bit big-endian words; only the first
//               and last (shared) word of a chunk use atomicOr
//   decode    : self-synchronizing parallel decode of the single unchunked
//               bitstream (speculative segment decode, fix-up rounds until every
//               segment starts on its predecessor's exit, scan, final decode)
#pragma once
#include <hip/hip_runtime.h>

#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/transform_iterator.hpp>
#include <vector>

#include "common.hpp"
#include "ctx.hpp"
#include "huff_host.hpp"

namespace cpsz {

constexpr int kLutBits = 12;
constexpr int kLut2Bits = 12;  // width cap of the level-2 decode tables

__device__ __forceinline__ uint32_t bswap32(uint32_t x) { return __builtin_bswap32(x); }

// --------------------------------------------------------------- histogram --
template <class Src>
__global__ void __launch_bounds__(256) k_hist(Src src, size_t nchunks, uint32_t* __restrict__ hist,
                                              uint32_t win_lo, uint32_t win_n) {
  extern __shared__ uint32_t sh[];
  for (uint32_t i = threadIdx.x; i < win_n; i += blockDim.x) sh[i] = 0;
  __syncthreads();
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < nchunks;
       c += (size_t)gridDim.x * blockDim.x) {
    src.each(c, [&](uint32_t s) {
      const uint32_t o = s - win_lo;
      if (o < win_n) atomicAdd(&sh[o], 1u);
      else atomicAdd(&hist[s], 1u);
    });
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < win_n; i += blockDim.x)
    if (sh[i]) atomicAdd(&hist[win_lo + i], sh[i]);
}

// ------------------------------------------------------------------ encode --
template <class Src>
__global__ void __launch_bounds__(256) k_chunk_bits(Src src, size_t nchunks,
                                                    const uint8_t* __restrict__ len,
                                                    uint64_t* __restrict__ bits) {
  const size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (c >= nchunks) return;
  uint64_t b = 0;
  src.each(c, [&](uint32_t s) { b += len[s]; });
  bits[c] = b;
}

template <class Src>
__global__ void __launch_bounds__(256) k_encode(Src src, size_t nchunks,
                                                const uint64_t* __restrict__ code,
                                                const uint8_t* __restrict__ len,
                                                const uint64_t* __restrict__ off,
                                                uint32_t* __restrict__ words) {
  const size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (c >= nchunks) return;
  const uint64_t pos0 = off[c];
  size_t w = pos0 >> 5;
  int nb = (int)(pos0 & 31);  // bits held in acc (leading ones belong to the previous chunk)
  uint64_t acc = 0;           // MSB-aligned pending bits
  bool shared = nb != 0;      // the current word is shared with the previous chunk
  src.each(c, [&](uint32_t s) {
    const int L = len[s];
    uint64_t cw = code[s];
    int rem = L;
    while (rem > 0) {
      const int take = rem < (64 - nb) ? rem : (64 - nb);
      acc |= cw >> nb;  // bits of cw past `take` are zero or shifted out
      cw <<= take;
      nb += take;
      rem -= take;
      while (nb >= 32) {
        const uint32_t word = bswap32((uint32_t)(acc >> 32));
        if (shared) atomicOr(&words[w], word);
        else words[w] = word;
        shared = false;
        acc <<= 32;
        nb -= 32;
        w++;
      }
    }
  });
  if (nb > 0) atomicOr(&words[w], bswap32((uint32_t)(acc >> 32)));
}

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

// Phase 1 / fix-up: decode from start to the first boundary >= seg_end.
__global__ void __launch_bounds__(256) k_dec_spec(DecTree t, const uint32_t* __restrict__ w, uint64_t nbits,
                                                  uint32_t S, uint32_t nseg, const uint64_t* __restrict__ prev_exit,
                           const uint64_t* __restrict__ prev_start, uint64_t* __restrict__ start,
                           uint64_t* __restrict__ exitp, uint32_t* __restrict__ cnt,
                           int* __restrict__ changed, int* __restrict__ err, int first);

__global__ void k_dec_check(uint32_t nseg, const uint64_t* exitp, const uint64_t* start,
                            int* changed);

// Phase 3: decode every segment into its final position. The symbols of a
// block's 256 consecutive segments are contiguous in the output: they are
// gathered in LDS and written out by consecutive threads (coalesced), instead
// of each thread storing its own run one symbol at a time (64 scattered
// 2-byte stores per instruction, which reach HBM as ~15x the bytes). A block
// whose run exceeds the staging buffer writes directly.
constexpr uint32_t kDecStage = 12288;  // u16 symbols staged per block

template <class Sink>
__global__ void __launch_bounds__(256) k_dec_write(DecTree t, const uint32_t* __restrict__ w, uint64_t nbits,
                                                   uint32_t S, uint32_t nseg, const uint64_t* __restrict__ start,
                                                   const uint64_t* __restrict__ exitp,
                                                   const uint64_t* __restrict__ off, uint64_t n,
                                                   Sink sink) {
  __shared__ uint32_t slut[1 << kLutBits];
  __shared__ uint32_t sw[kStageWords];
  __shared__ uint16_t stage[kDecStage];
  for (int k = threadIdx.x; k < (1 << kLutBits); k += blockDim.x) slut[k] = t.lut[k];
  uint64_t lbase;
  const uint32_t lcount = stage_words(sw, w, nbits, S, blockIdx.x * blockDim.x, lbase);
  __syncthreads();
  const uint32_t b0 = blockIdx.x * blockDim.x;
  const uint32_t b1 = min(b0 + blockDim.x, nseg);  // off[] holds nseg + 1 entries
  const uint64_t o_begin = min(off[b0], n), o_end = min(off[b1], n);
  const bool staged = o_end - o_begin <= (uint64_t)kDecStage;
  const uint32_t i = b0 + threadIdx.x;
  if (i < nseg) {
    BitReader br;
    br.init(w, start[i], sw, lbase, lcount);
    const uint64_t end = exitp[i];
    uint64_t o = off[i];
    if (staged) {
      while (br.pos < end && o < o_end) stage[o++ - o_begin] = (uint16_t)dec_one(t, slut, br);
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
DeviceDecoder upload_decoder(Ctx& ctx, const uint8_t* tree, size_t tree_size, size_t node_count,
                             const char* tag);

// Decodes n symbols of a bitstream already copied to aligned device words
// (padded with at least 8 zero bytes) into `sink`.
template <class Sink>
void decode_stream(Ctx& ctx, const DeviceDecoder& dd, const uint32_t* d_words, uint64_t nbytes,
                   uint64_t n, Sink sink, const char* tag);

template <class Sink>
void decode_stream(Ctx& ctx, const DeviceDecoder& dd, const uint32_t* d_words, uint64_t nbytes,
                   uint64_t n, Sink sink, const char* tag) {
  hipStream_t st = ctx.stream();
  if (n == 0) return;
  if (dd.single) {
    k_dec_const<<<1024, 256, 0, st>>>(n, dd.single_sym, sink);
    CPSZ_HIP(hipGetLastError());
    return;
  }
  const uint64_t nbits = nbytes * 8;
  CPSZ_REQUIRE(nbits > 0, kErrStream, "empty huffman bitstream for a multi-symbol tree");
  uint32_t S = 256;
  while (div_up(nbits, S) > (1u << 21)) S *= 2;
  const uint32_t nseg = (uint32_t)div_up(nbits, S);
  std::string t(tag);
  uint64_t* start[2] = {ctx.buf<uint64_t>((t + ".st0").c_str(), nseg), ctx.buf<uint64_t>((t + ".st1").c_str(), nseg)};
  uint64_t* ex[2] = {ctx.buf<uint64_t>((t + ".ex0").c_str(), nseg), ctx.buf<uint64_t>((t + ".ex1").c_str(), nseg)};
  uint32_t* cnt = ctx.buf<uint32_t>((t + ".cnt").c_str(), nseg);
  uint64_t* off = ctx.buf<uint64_t>((t + ".off").c_str(), nseg + 1);
  int* flags = ctx.buf<int>("dec.flags", 2);
  int* hflags = (int*)ctx.host_buf("dec.flags", 2 * sizeof(int));
  const int blocks = (int)div_up(nseg, 256);
  CPSZ_HIP(hipMemsetAsync(flags, 0, 2 * sizeof(int), st));
  k_dec_spec<<<blocks, 256, 0, st>>>(dd.tree, d_words, nbits, S, nseg, nullptr, nullptr, start[0],
                                     ex[0], cnt, flags, flags + 1, 1);
  CPSZ_HIP(hipGetLastError());
  int cur = 0;
  for (int round = 0; round < 4096; round++) {
    // two fix-up rounds per host check
    for (int k = 0; k < 2; k++) {
      k_dec_spec<<<blocks, 256, 0, st>>>(dd.tree, d_words, nbits, S, nseg, ex[cur], start[cur],
                                         start[cur ^ 1], ex[cur ^ 1], cnt, flags, flags + 1, 0);
      cur ^= 1;
    }
    CPSZ_HIP(hipMemsetAsync(flags, 0, sizeof(int), st));
    k_dec_check<<<blocks, 256, 0, st>>>(nseg, ex[cur], start[cur], flags);
    CPSZ_HIP(hipMemcpyAsync(hflags, flags, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
    CPSZ_HIP(hipStreamSynchronize(st));
    CPSZ_REQUIRE(hflags[1] == 0, kErrStream, "invalid huffman code in stream");
    if (!hflags[0]) break;
  }
  scan_u32_to_u64(ctx, cnt, off, nseg, off + nseg);
  uint64_t total = 0;
  CPSZ_HIP(hipMemcpyAsync(&total, off + nseg, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
  CPSZ_HIP(hipStreamSynchronize(st));
  CPSZ_REQUIRE(total >= n, kErrStream, "huffman bitstream holds fewer symbols than expected");
  k_dec_write<<<blocks, 256, 0, st>>>(dd.tree, d_words, nbits, S, nseg, start[cur], ex[cur], off, n, sink);
  CPSZ_HIP(hipGetLastError());
}

}  // namespace cpsz
