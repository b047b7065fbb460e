// Exception tables that make log2_dd / exp2_dd (libm_math.hpp) reproduce the
// host C library's log2f / exp2f bit for bit (see libm_math.hpp).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "libm_math.hpp"

namespace cpsz {

struct LibmTable {
  std::vector<float> band;
  std::vector<uint8_t> kind;
  std::vector<uint32_t> off, cnt;
  std::vector<uint32_t> list_x, list_y;
  std::vector<uint32_t> bitmap;
  LibmView view() const {
    return LibmView{band.data(), kind.data(), off.data(), cnt.data(),
                    list_x.data(), list_y.data(), bitmap.data()};
  }
};

struct LibmTables {
  LibmTable log2f, exp2f;
  std::string libc_version;
};

// Process-wide tables: loaded from the cache file next to the library when it
// matches this host's C library, otherwise rebuilt (multi-threaded sweep of
// every float in the domains) and written back to the cache when possible.
const LibmTables& libm_tables();
// Rebuilds the tables and writes them to `path` (used by the build step).
void libm_prepare(const std::string& path);
// Default cache path (next to libcpsz_hip.so).
std::string libm_cache_path();

}  // namespace cpsz
