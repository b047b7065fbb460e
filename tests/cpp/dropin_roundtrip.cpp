// Caller written against the reference's C++ API (the flow of
// test/sz_compress_cp_preserve_2d_test.cpp:20-49 and
// test/sz_compress_cp_preserve_3d_test.cpp:26-42), compiled against the
// drop-in header and linked with libcpsz_hip.so.
//
// usage: dropin_roundtrip <2d|2dlog|3d> <in.bin> <out_prefix> r1 r2 [r3] eb
//   (2dlog: the 2D offline_log pair)
//   in.bin holds the components back to back (float32, row-major); writes
//   <out_prefix>.stream (pre-ZSTD stream), <out_prefix>.zst and
//   <out_prefix>.out (decompressed components back to back). Prints
//   "ok <stream bytes> <zstd bytes>" or "fail <message>".
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "sz_cp_preserve_dropin.hpp"

static bool write_file(const std::string& p, const void* d, size_t n) {
  FILE* f = std::fopen(p.c_str(), "wb");
  if (!f) return false;
  const bool ok = std::fwrite(d, 1, n, f) == n;
  std::fclose(f);
  return ok;
}

int main(int argc, char** argv) {
  if (argc < 7) {
    std::printf("fail usage\n");
    return 2;
  }
  const bool is3d = std::strcmp(argv[1], "3d") == 0;
  const bool islog = std::strcmp(argv[1], "2dlog") == 0;
  // int dims and atof, as the reference drivers parse them
  // (test/sz_compress_cp_preserve_2d_test.cpp:14-16)
  const int r1 = std::atoi(argv[4]), r2 = std::atoi(argv[5]);
  const int r3 = is3d ? std::atoi(argv[6]) : 1;
  const double max_eb = std::atof(argv[is3d ? 7 : 6]);
  const size_t n = (size_t)r1 * r2 * r3, ncomp = is3d ? 3 : 2;
  std::vector<float> in(ncomp * n);
  FILE* f = std::fopen(argv[2], "rb");
  if (!f || std::fread(in.data(), 4, in.size(), f) != in.size()) {
    std::printf("fail read\n");
    return 2;
  }
  std::fclose(f);
  const std::string out = argv[3];
  size_t size = 0;
  unsigned char* stream =
      is3d ? sz_compress_cp_preserve_3d_offline_log(in.data(), in.data() + n, in.data() + 2 * n, r1, r2, r3,
                                                    size, false, max_eb)
           : islog ? sz_compress_cp_preserve_2d_offline_log(in.data(), in.data() + n, r1, r2, size, false, max_eb)
                   : sz_compress_cp_preserve_2d_offline(in.data(), in.data() + n, r1, r2, size, false, max_eb);
  if (!stream) {
    std::printf("fail compress: %s\n", cpsz_last_error());
    return 1;
  }
  // test/sz_compress_cp_preserve_2d_test.cpp:29,36
  unsigned char* z = nullptr;
  const size_t zs = sz_lossless_compress(ZSTD_COMPRESSOR, 3, stream, size, &z);
  unsigned char* back = nullptr;
  const size_t bs = sz_lossless_decompress(ZSTD_COMPRESSOR, z, zs, &back, size);
  if (zs == 0 || bs != size || std::memcmp(back, stream, size) != 0) {
    std::printf("fail zstd round trip\n");
    return 1;
  }
  // the rest of include/sz_lossless.hpp: frame check and 64 KiB header read
  unsigned char* head = nullptr;
  if (is_lossless_compressed_data(z, zs) != ZSTD_COMPRESSOR || is_lossless_compressed_data(stream, 4) != -1 ||
      sz_lossless_decompress65536bytes(ZSTD_COMPRESSOR, z, zs, &head) != 65536 ||
      std::memcmp(head, stream, size < 65536 ? size : 65536) != 0) {
    std::printf("fail lossless header helpers\n");
    return 1;
  }
  std::free(head);
  float *U = nullptr, *V = nullptr, *W = nullptr;
  if (is3d) sz_decompress_cp_preserve_3d_online_log<float>(back, r1, r2, r3, U, V, W);
  else if (islog) sz_decompress_cp_preserve_2d_offline_log<float>(back, r1, r2, U, V);
  else sz_decompress_cp_preserve_2d_offline<float>(back, r1, r2, U, V);
  if (!U || !V || (is3d && !W)) {
    std::printf("fail decompress: %s\n", cpsz_last_error());
    return 1;
  }
  std::vector<float> dec(ncomp * n);
  std::memcpy(dec.data(), U, 4 * n);
  std::memcpy(dec.data() + n, V, 4 * n);
  if (is3d) std::memcpy(dec.data() + 2 * n, W, 4 * n);
  const bool wrote = write_file(out + ".stream", stream, size) && write_file(out + ".zst", z, zs) &&
                     write_file(out + ".out", dec.data(), 4 * dec.size());
  std::free(stream);
  std::free(z);
  std::free(back);
  std::free(U);
  std::free(V);
  std::free(W);
  if (!wrote) {
    std::printf("fail write\n");
    return 1;
  }
  std::printf("ok %zu %zu\n", size, zs);
  return 0;
}
