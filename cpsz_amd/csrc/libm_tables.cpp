#include "libm_tables.hpp"

#include <dlfcn.h>
#include <gnu/libc-version.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <random>
#include <thread>

#include "common.hpp"

namespace cpsz {

namespace {

// Domains of the two functions as the codec uses them.
//   log2f: every finite x > 0 (slots = exponent field 0..254)
//   exp2f: y in [-99, 128): positive bit patterns [0, 0x43000000), negative
//          patterns [0x80000000, 0xC2C60000] (-99.0f is 0xC2C60000)
struct Domain {
  int nslots;
  bool is_log;
  // mantissa range [lo, hi) of a slot, empty if outside the domain
  void range(int slot, uint32_t& lo, uint32_t& hi) const {
    lo = 0;
    hi = 1u << 23;
    if (is_log) {
      if (slot >= 255) hi = 0;
      if (slot == 0) lo = 1;  // skip +0
      return;
    }
    const uint32_t base = (uint32_t)slot << 23;
    if (slot < 256) {
      if (base >= 0x43000000u) hi = 0;
    } else {
      if (base > 0xC2C60000u) hi = 0;
      else if (base == (0xC2C60000u & 0xFF800000u)) hi = (0xC2C60000u & 0x7FFFFFu) + 1;
    }
  }
};

struct Exc {
  uint32_t x, y;
  float dist;
  bool rule;
};

void sweep_slot(const Domain& d, int slot, std::vector<Exc>& out) {
  uint32_t lo, hi;
  d.range(slot, lo, hi);
  out.clear();
  for (uint32_t m = lo; m < hi; m++) {
    const uint32_t xb = ((uint32_t)slot << 23) | m;
    const float x = ffrom(xb);
    double y;
    float g;
    if (d.is_log) {
      y = log2_dd(x);
      g = log2f(x);
    } else {
      y = exp2_dd(x);
      g = exp2f(x);
    }
    const float f = (float)y;
    if (fbits(f) == fbits(g)) continue;
    out.push_back(Exc{xb, fbits(g), midpoint_distance(y, f), fbits(side_neighbour(y, f)) == fbits(g)});
  }
}

LibmTable build_table(const Domain& d) {
  std::vector<std::vector<Exc>> per(d.nslots);
  std::atomic<int> next{0};
  unsigned nt = std::thread::hardware_concurrency();
  if (nt == 0) nt = 4;
  if (nt > 16) nt = 16;
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; t++)
    th.emplace_back([&] {
      std::vector<Exc> tmp;
      for (int s; (s = next.fetch_add(1)) < d.nslots;) {
        sweep_slot(d, s, tmp);
        per[s] = tmp;
      }
    });
  for (auto& t : th) t.join();
  LibmTable t;
  t.band.assign(d.nslots, -1.f);
  t.kind.assign(d.nslots, 0);
  t.off.assign(d.nslots, 0);
  t.cnt.assign(d.nslots, 0);
  for (int s = 0; s < d.nslots; s++) {
    const auto& e = per[s];
    if (e.empty()) continue;
    float band = 0;
    bool rule = true;
    for (const Exc& x : e) {
      band = std::max(band, x.dist);
      rule = rule && x.rule;
    }
    t.band[s] = band;
    // dense binades (only log2f near 1 has tens of thousands) get a 1 MB
    // bitmap; the rest a sorted list searched only inside the band
    if (rule && e.size() > 20000) {
      t.kind[s] = 2;
      t.off[s] = (uint32_t)t.bitmap.size();
      t.bitmap.resize(t.bitmap.size() + (1u << 18), 0);
      for (const Exc& x : e) {
        const uint32_t m = x.x & 0x7FFFFFu;
        t.bitmap[t.off[s] + (m >> 5)] |= 1u << (m & 31);
      }
    } else {
      t.kind[s] = 1;
      t.off[s] = (uint32_t)t.list_x.size();
      t.cnt[s] = (uint32_t)e.size();
      for (const Exc& x : e) {  // sweep order is ascending in x
        t.list_x.push_back(x.x);
        t.list_y.push_back(x.y);
      }
    }
  }
  return t;
}

const Domain kLog{256, true};
const Domain kExp{512, false};

template <typename V>
void wvec(FILE* f, const std::vector<V>& v) {
  uint64_t n = v.size();
  fwrite(&n, 8, 1, f);
  if (n) fwrite(v.data(), sizeof(V), n, f);
}
template <typename V>
bool rvec(FILE* f, std::vector<V>& v) {
  uint64_t n;
  if (fread(&n, 8, 1, f) != 1 || n > (1ull << 30)) return false;
  v.resize(n);
  return !n || fread(v.data(), sizeof(V), n, f) == n;
}

constexpr char kMagic[8] = {'C', 'P', 'S', 'Z', 'L', 'B', 'M', '1'};

bool save(const std::string& path, const LibmTables& t) {
  const std::string tmp = path + ".tmp";
  FILE* f = fopen(tmp.c_str(), "wb");
  if (!f) return false;
  fwrite(kMagic, 1, 8, f);
  char ver[64] = {0};
  std::strncpy(ver, t.libc_version.c_str(), 63);
  fwrite(ver, 1, 64, f);
  for (const LibmTable* x : {&t.log2f, &t.exp2f}) {
    wvec(f, x->band);
    wvec(f, x->kind);
    wvec(f, x->off);
    wvec(f, x->cnt);
    wvec(f, x->list_x);
    wvec(f, x->list_y);
    wvec(f, x->bitmap);
  }
  const bool ok = fclose(f) == 0;
  return ok && std::rename(tmp.c_str(), path.c_str()) == 0;
}

bool load(const std::string& path, LibmTables& t) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) return false;
  char mg[8], ver[64];
  bool ok = fread(mg, 1, 8, f) == 8 && std::memcmp(mg, kMagic, 8) == 0 && fread(ver, 1, 64, f) == 64;
  if (ok) {
    ver[63] = 0;
    t.libc_version = ver;
    for (LibmTable* x : {&t.log2f, &t.exp2f})
      ok = ok && rvec(f, x->band) && rvec(f, x->kind) && rvec(f, x->off) && rvec(f, x->cnt) &&
           rvec(f, x->list_x) && rvec(f, x->list_y) && rvec(f, x->bitmap);
  }
  fclose(f);
  return ok && t.log2f.band.size() == 256 && t.exp2f.band.size() == 512;
}

// Spot check against the live host libm: every listed exception of a few
// slots plus random inputs over both domains.
bool validate(const LibmTables& t) {
  if (t.libc_version != gnu_get_libc_version()) return false;
  const LibmView lv = t.log2f.view(), ev = t.exp2f.view();
  std::mt19937 rng(12345);
  for (int i = 0; i < 200000; i++) {
    const uint32_t xb = 1 + (rng() % 0x7F7FFFFFu);
    const float x = ffrom(xb);
    if (fbits(host_log2f(lv, x)) != fbits(log2f(x))) return false;
    const float y = -99.0f + (float)(rng() % 2270000) * 1e-4f;
    if (fbits(host_exp2f(ev, y)) != fbits(exp2f(y))) return false;
  }
  for (size_t i = 0; i < t.log2f.list_x.size(); i += 7) {
    const float x = ffrom(t.log2f.list_x[i]);
    if (fbits(host_log2f(lv, x)) != fbits(log2f(x))) return false;
  }
  for (size_t i = 0; i < t.exp2f.list_x.size(); i += 7) {
    const float y = ffrom(t.exp2f.list_x[i]);
    if (fbits(host_exp2f(ev, y)) != fbits(exp2f(y))) return false;
  }
  return true;
}

LibmTables build_all() {
  LibmTables t;
  t.libc_version = gnu_get_libc_version();
  t.log2f = build_table(kLog);
  t.exp2f = build_table(kExp);
  return t;
}

}  // namespace

std::string libm_cache_path() {
  if (const char* e = std::getenv("CPSZ_LIBM_CACHE")) return e;
  Dl_info info;
  if (dladdr((void*)&libm_cache_path, &info) && info.dli_fname) {
    std::string p(info.dli_fname);
    const size_t k = p.rfind('/');
    return (k == std::string::npos ? std::string(".") : p.substr(0, k)) + "/libm_tables.bin";
  }
  return "libm_tables.bin";
}

void libm_prepare(const std::string& path) {
  LibmTables t = build_all();
  CPSZ_REQUIRE(validate(t), kErrInternal, "libm table self-check failed");
  CPSZ_REQUIRE(save(path, t), kErrInternal, "cannot write " + path);
}

const LibmTables& libm_tables() {
  static std::once_flag once;
  static LibmTables* t = nullptr;
  std::call_once(once, [] {
    auto* x = new LibmTables();
    const std::string path = libm_cache_path();
    if (!(load(path, *x) && validate(*x))) {
      *x = build_all();
      CPSZ_REQUIRE(validate(*x), kErrInternal, "libm table self-check failed");
      save(path, *x);  // best effort
    }
    t = x;
  });
  return *t;
}

}  // namespace cpsz
