#include "huff_gpu.hpp"

#include <cstring>

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

__global__ void __launch_bounds__(256) k_dec_spec(DecTree t, const uint32_t* __restrict__ w, uint64_t nbits,
                                                  uint32_t S, uint32_t nseg,
                                                  const uint64_t* __restrict__ prev_exit,
                                                  const uint64_t* __restrict__ prev_start,
                                                  uint64_t* __restrict__ start, uint64_t* __restrict__ exitp,
                                                  uint32_t* __restrict__ cnt, int* __restrict__ changed,
                                                  int* __restrict__ err, int first) {
  __shared__ uint32_t slut[1 << kLutBits];
  __shared__ uint32_t sw[kStageWords];
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
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
  const uint32_t lcount = stage_words(sw, w, nbits, S, blockIdx.x * blockDim.x, lbase);
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
    while (br.pos < s0) {
      if (dec_one(t, slut, br) == 0xFFFFFFFFu) {
        *err = 1;
        break;
      }
    }
    s0 = br.pos;
    while (br.pos < end) {
      if (dec_one(t, slut, br) == 0xFFFFFFFFu) {
        *err = 1;
        break;
      }
      c++;
    }
    start[i] = s0;
    exitp[i] = br.pos;
    cnt[i] = c;
    return;
  }
  br.init(w, s0, sw, lbase, lcount);
  // New start: decode the new path and the previous pass's path of this
  // segment in lockstep (always advancing the one behind) until they reach
  // the same code boundary; from there both paths coincide, so the previous
  // pass's exit and the rest of its count carry over. Prefix codes resync
  // within a few code words, so a fix-up costs a few symbols, not a segment.
  BitReader old;
  old.init(w, prev_start[i], sw, lbase, lcount);
  const uint64_t old_exit = prev_exit[i];
  const uint32_t old_cnt = cnt[i];
  uint32_t c_old = 0;
  uint64_t ex = 0;
  for (;;) {
    if (br.pos >= end) {
      ex = br.pos;
      break;
    }
    if (br.pos == old.pos) {
      ex = old_exit;
      c += old_cnt - c_old;
      break;
    }
    const bool adv_new = br.pos < old.pos;
    if (dec_one(t, slut, adv_new ? br : old) == 0xFFFFFFFFu) {
      *err = 1;
      ex = br.pos;
      break;
    }
    c += adv_new;
    c_old += !adv_new;
  }
  start[i] = s0;
  exitp[i] = ex;
  cnt[i] = c;
  (void)changed;
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

DeviceDecoder upload_decoder(Ctx& ctx, const uint8_t* tree, size_t tree_size, size_t node_count,
                             const char* tag) {
  thread_local HostDecoder hd;  // keeps its allocations from call to call
  prepare_decoder(tree, tree_size, node_count, hd);
  return upload_decoder(ctx, hd, tag);
}

}  // namespace cpsz
