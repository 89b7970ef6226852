// serial.hip -- binary caches of parsed inputs (host only, no device work).
//
// The reference serializes a loaded fragment once under --serialization_prefix and
// reads that file on later runs instead of parsing the text again (loader.h:30-61,
// SetSerialize; the file "<prefix>/<digest>/frag_<fid>.s" with digest = efile +
// vfile + worker count, '/' replaced by '_', config.h:49).  This library caches
// what it parses -- the edge list of an edge file and the records of an update
// file -- under the same path rule, in a format of its own (little-endian):
//
//   edges    "SUMINCE1", u32 weight_kind, u32 n, u64 m, u64 skipped lines,
//            u64 offsets[n + 1], u32 cols[m], u32 weight bits[m] (weighted only),
//            u64 FNV-1a 64 of every byte before it
//   updates  "SUMINCU1", u64 count, count 24-byte suminc_update records, u64 FNV-1a
//
// A load checks the magic, every size against the file size, the checksum, and for
// an edge list that the offsets start at 0, never decrease, end at m, and that
// every column is below n.
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ctx.hpp"
#include "io.hpp"

namespace suminc {
namespace {

constexpr char kEdgeMagic[8] = {'S', 'U', 'M', 'I', 'N', 'C', 'E', '1'};
constexpr char kUpdateMagic[8] = {'S', 'U', 'M', 'I', 'N', 'C', 'U', '1'};

struct Fnv {
  uint64_t h = 0xcbf29ce484222325ull;
  void add(const void* p, size_t bytes) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < bytes; ++i) {
      h ^= b[i];
      h *= 0x100000001b3ull;
    }
  }
};

struct File {
  FILE* f = nullptr;
  std::string path;
  File(const char* p, const char* mode) {
    SUMINC_REQUIRE(p, SUMINC_E_INVALID, "null path");
    path = p;
    f = std::fopen(p, mode);
    SUMINC_REQUIRE(f, SUMINC_E_INVALID, "cannot open " + path + ": " + std::strerror(errno));
  }
  ~File() {
    if (f) std::fclose(f);
  }
  void put(const void* p, size_t bytes, Fnv& h) {
    if (!bytes) return;
    SUMINC_REQUIRE(std::fwrite(p, 1, bytes, f) == bytes, SUMINC_E_INVALID, "write failed: " + path);
    h.add(p, bytes);
  }
  void get(void* p, size_t bytes, Fnv* h) {
    if (!bytes) return;
    SUMINC_REQUIRE(std::fread(p, 1, bytes, f) == bytes, SUMINC_E_INVALID, "truncated file: " + path);
    if (h) h->add(p, bytes);
  }
  uint64_t size() {
    const long at = std::ftell(f);
    std::fseek(f, 0, SEEK_END);
    const long end = std::ftell(f);
    std::fseek(f, at, SEEK_SET);
    return end < 0 ? 0 : (uint64_t)end;
  }
  void close_checked() {
    const int rc = std::fclose(f);
    f = nullptr;
    SUMINC_REQUIRE(rc == 0, SUMINC_E_INVALID, "close failed: " + path);
  }
};

void save_edges(const suminc_edges* e, const char* path) {
  SUMINC_REQUIRE(e, SUMINC_E_INVALID, "null edge list");
  File f(path, "wb");
  Fnv h;
  const uint32_t wk = (uint32_t)e->wkind, n = e->n;
  const uint64_t m = e->col.size(), skipped = e->skipped;
  f.put(kEdgeMagic, 8, h);
  f.put(&wk, 4, h);
  f.put(&n, 4, h);
  f.put(&m, 8, h);
  f.put(&skipped, 8, h);
  f.put(e->off.data(), e->off.size() * 8, h);
  f.put(e->col.data(), m * 4, h);
  if (wk != SUMINC_W_NONE) f.put(e->w.data(), m * 4, h);
  Fnv unused;
  f.put(&h.h, 8, unused);
  f.close_checked();
}

suminc_edges* load_edges(const char* path) {
  File f(path, "rb");
  const uint64_t fsize = f.size();
  Fnv h;
  char magic[8];
  uint32_t wk = 0, n = 0;
  uint64_t m = 0, skipped = 0;
  SUMINC_REQUIRE(fsize >= 40, SUMINC_E_INVALID, "not an edge cache (too short): " + f.path);
  f.get(magic, 8, &h);
  SUMINC_REQUIRE(std::memcmp(magic, kEdgeMagic, 8) == 0, SUMINC_E_INVALID, "not an edge cache: " + f.path);
  f.get(&wk, 4, &h);
  f.get(&n, 4, &h);
  f.get(&m, 8, &h);
  f.get(&skipped, 8, &h);
  SUMINC_REQUIRE(wk == SUMINC_W_NONE || wk == SUMINC_W_U32 || wk == SUMINC_W_F32, SUMINC_E_INVALID,
                 "bad weight kind in " + f.path);
  SUMINC_REQUIRE(m < 0xFFFFFFFFull, SUMINC_E_TOO_LARGE, "more than 2^32 - 1 edges in " + f.path);
  const uint64_t want = 32 + ((uint64_t)n + 1) * 8 + m * 4 + (wk ? m * 4 : 0) + 8;
  SUMINC_REQUIRE(fsize == want, SUMINC_E_INVALID, "edge cache size does not match its header: " + f.path);
  suminc_edges* e = new suminc_edges();
  try {
    e->n = n;
    e->wkind = (int)wk;
    e->skipped = skipped;
    e->off.resize((size_t)n + 1);
    e->col.resize(m);
    f.get(e->off.data(), e->off.size() * 8, &h);
    f.get(e->col.data(), m * 4, &h);
    if (wk) {
      e->w.resize(m);
      f.get(e->w.data(), m * 4, &h);
    }
    uint64_t sum = 0;
    f.get(&sum, 8, nullptr);
    SUMINC_REQUIRE(sum == h.h, SUMINC_E_INVALID, "edge cache checksum mismatch: " + f.path);
    SUMINC_REQUIRE(e->off[0] == 0 && e->off[n] == m, SUMINC_E_INVALID, "edge cache offsets out of range: " + f.path);
    for (uint32_t v = 0; v < n; ++v)
      SUMINC_REQUIRE(e->off[v + 1] >= e->off[v], SUMINC_E_INVALID, "edge cache offsets decrease: " + f.path);
    for (uint64_t k = 0; k < m; ++k)
      SUMINC_REQUIRE(e->col[k] < n, SUMINC_E_INVALID, "edge cache column out of range: " + f.path);
  } catch (...) {
    delete e;
    throw;
  }
  return e;
}

void save_updates(const char* path, const suminc_update* ops, uint64_t count) {
  SUMINC_REQUIRE(ops || count == 0, SUMINC_E_INVALID, "null updates");
  File f(path, "wb");
  Fnv h;
  f.put(kUpdateMagic, 8, h);
  f.put(&count, 8, h);
  f.put(ops, count * sizeof(suminc_update), h);
  Fnv unused;
  f.put(&h.h, 8, unused);
  f.close_checked();
}

void load_updates(const char* path, suminc_update* ops, uint64_t cap, uint64_t* count) {
  SUMINC_REQUIRE(count, SUMINC_E_INVALID, "null count");
  File f(path, "rb");
  const uint64_t fsize = f.size();
  Fnv h;
  char magic[8];
  uint64_t cnt = 0;
  SUMINC_REQUIRE(fsize >= 24, SUMINC_E_INVALID, "not an update cache (too short): " + f.path);
  f.get(magic, 8, &h);
  SUMINC_REQUIRE(std::memcmp(magic, kUpdateMagic, 8) == 0, SUMINC_E_INVALID, "not an update cache: " + f.path);
  f.get(&cnt, 8, &h);
  SUMINC_REQUIRE(fsize == 24 + cnt * sizeof(suminc_update), SUMINC_E_INVALID,
                 "update cache size does not match its header: " + f.path);
  *count = cnt;
  if (!ops) return;
  SUMINC_REQUIRE(cap >= cnt, SUMINC_E_INVALID, "update buffer too small");
  f.get(ops, cnt * sizeof(suminc_update), &h);
  uint64_t sum = 0;
  f.get(&sum, 8, nullptr);
  SUMINC_REQUIRE(sum == h.h, SUMINC_E_INVALID, "update cache checksum mismatch: " + f.path);
}

// "<prefix>/<efile + vfile + workers, '/' -> '_'>/frag_<fid>.s" (loader.h:30-45,
// config.h:49)
std::string serialization_path(const char* prefix, const char* efile, const char* vfile, uint32_t workers,
                               uint32_t fid) {
  SUMINC_REQUIRE(prefix && efile, SUMINC_E_INVALID, "null prefix or edge file name");
  std::string digest = std::string(efile) + (vfile ? vfile : "") + std::to_string(workers);
  for (char& ch : digest)
    if (ch == '/') ch = '_';
  return std::string(prefix) + "/" + digest + "/frag_" + std::to_string(fid) + ".s";
}

}  // namespace
}  // namespace suminc

extern "C" {

int suminc_serialization_path(const char* prefix, const char* efile, const char* vfile, uint32_t workers,
                              uint32_t fid, char* out, uint64_t cap, uint64_t* length) {
  return suminc::host_call([&] {
    const std::string p = suminc::serialization_path(prefix, efile, vfile, workers, fid);
    if (length) *length = p.size();
    if (out) {
      SUMINC_REQUIRE(cap > p.size(), SUMINC_E_INVALID, "path buffer too small");
      std::memcpy(out, p.c_str(), p.size() + 1);
    }
  });
}

int suminc_edges_save(const suminc_edges* edges, const char* path) {
  return suminc::host_call([&] { suminc::save_edges(edges, path); });
}

int suminc_edges_load(const char* path, suminc_edges** out) {
  return suminc::host_call([&] {
    SUMINC_REQUIRE(out, SUMINC_E_INVALID, "null output");
    *out = suminc::load_edges(path);
  });
}

int suminc_updates_save(const char* path, const suminc_update* ops, uint64_t count) {
  return suminc::host_call([&] { suminc::save_updates(path, ops, count); });
}

int suminc_updates_load(const char* path, suminc_update* ops, uint64_t cap, uint64_t* count) {
  return suminc::host_call([&] { suminc::load_updates(path, ops, cap, count); });
}

}  // extern "C"
