// io.hip -- the reference's text formats, parsed on the host (no device work):
//
//   edge file     "u v [w]" per line; empty lines, lines starting with '#' and lines
//                 that do not parse are skipped
//                 (grape/fragment/ev_fragment_loader.h:112, e_fragment_loader.h:113;
//                 tokens as grape/io/tsv_line_parser.h reads them); undirected files
//                 give both directions (the fragment's directed flag).
//   update file   "a u v [w]" / "d u v" per line (inc_fragment_builder.h:38-72).
//   cluster file  "<size> <id> ... <id>" per cluster, whitespace-separated
//                 (compressor_base.h:3153-3250, the "<efile>.c_<max_node_num>" file).
//
// Large files are read with one pass per thread over newline-aligned chunks of a
// memory map; the edge list becomes a CSR with sorted rows by a counting scatter
// and a per-row sort on (column, line order), keeping the first line of every
// (u, v) pair.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "ctx.hpp"
#include "io.hpp"

namespace suminc {

void upload_csr(suminc_ctx* c, uint32_t n, uint64_t m, const uint64_t* off, const uint32_t* col, int wkind,
                const void* w);

namespace {

thread_local std::string t_err;  // message of the last failed call without a context

// Read-only memory map of a whole file.
struct MappedFile {
  const char* p = nullptr;
  size_t size = 0;
  int fd = -1;
  explicit MappedFile(const char* path) {
    SUMINC_REQUIRE(path, SUMINC_E_INVALID, "null path");
    fd = ::open(path, O_RDONLY);
    SUMINC_REQUIRE(fd >= 0, SUMINC_E_INVALID, std::string("cannot open ") + path + ": " + std::strerror(errno));
    struct stat st;
    if (::fstat(fd, &st) != 0) {
      ::close(fd);
      throw Error(SUMINC_E_INVALID, std::string("cannot stat ") + path);
    }
    size = (size_t)st.st_size;
    if (size) {
      void* m = ::mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
      if (m == MAP_FAILED) {
        ::close(fd);
        throw Error(SUMINC_E_INVALID, std::string("cannot map ") + path);
      }
      p = static_cast<const char*>(m);
    }
  }
  ~MappedFile() {
    if (p) ::munmap(const_cast<char*>(p), size);
    if (fd >= 0) ::close(fd);
  }
  MappedFile(const MappedFile&) = delete;
  MappedFile& operator=(const MappedFile&) = delete;
};

unsigned io_threads() {
  const unsigned hw = std::thread::hardware_concurrency();
  return std::max(1u, std::min(16u, hw ? hw : 1u));
}

// [begin, end) byte ranges of about equal size, each starting at a line start.
std::vector<std::pair<size_t, size_t>> line_chunks(const MappedFile& f, unsigned parts) {
  std::vector<std::pair<size_t, size_t>> out;
  size_t b = 0;
  for (unsigned i = 0; i < parts && b < f.size; ++i) {
    size_t e = i + 1 == parts ? f.size : std::max(b, f.size / parts * (i + 1));
    while (e < f.size && f.p[e - 1] != '\n') ++e;
    out.emplace_back(b, e);
    b = e;
  }
  return out;
}

template <class F>
void run_parallel(size_t parts, F&& f) {
  std::vector<std::thread> th;
  std::vector<std::string> errs(parts);
  std::vector<int> codes(parts, SUMINC_OK);
  for (size_t i = 0; i < parts; ++i)
    th.emplace_back([&, i] {
      try {
        f(i);
      } catch (const Error& e) {
        codes[i] = e.code;
        errs[i] = e.what();
      } catch (const std::bad_alloc&) {
        codes[i] = SUMINC_E_NOMEM;
        errs[i] = "host allocation failed";
      }
    });
  for (auto& t : th) t.join();
  for (size_t i = 0; i < parts; ++i)
    if (codes[i] != SUMINC_OK) throw Error(codes[i], errs[i]);
}

inline bool is_space(char ch) { return ch == ' ' || ch == '\t' || ch == '\n' || ch == '\r' || ch == '\v' || ch == '\f'; }

// Tokenizer over one line [p, e).
struct Line {
  const char* p;
  const char* e;
  void skip() {
    while (p < e && is_space(*p)) ++p;
  }
  bool more() {
    skip();
    return p < e;
  }
  bool token(const char*& b, const char*& t) {
    skip();
    if (p >= e) return false;
    b = p;
    while (p < e && !is_space(*p)) ++p;
    t = p;
    return true;
  }
};

std::string where(size_t line_no) { return " (line " + std::to_string(line_no) + ")"; }

uint64_t parse_u64(const char* b, const char* t, size_t line_no, const char* what) {
  SUMINC_REQUIRE(b < t, SUMINC_E_INVALID, std::string("missing ") + what + where(line_no));
  uint64_t v = 0;
  for (const char* q = b; q < t; ++q) {
    SUMINC_REQUIRE(*q >= '0' && *q <= '9', SUMINC_E_INVALID,
                   std::string("bad ") + what + " '" + std::string(b, t) + "'" + where(line_no));
    v = v * 10 + (uint64_t)(*q - '0');
    SUMINC_REQUIRE(v <= 0xFFFFFFFFull, SUMINC_E_INVALID, std::string(what) + " too large" + where(line_no));
  }
  return v;
}

uint32_t parse_weight(const char* b, const char* t, int wkind, size_t line_no) {
  if (wkind == SUMINC_W_U32) {
    const uint64_t v = parse_u64(b, t, line_no, "weight");
    SUMINC_REQUIRE(v >= 1, SUMINC_E_INVALID, "integer weights must be at least 1" + where(line_no));
    return (uint32_t)v;
  }
  std::string s(b, t);
  char* endp = nullptr;
  const float f = std::strtof(s.c_str(), &endp);
  SUMINC_REQUIRE(endp && *endp == '\0' && !s.empty(), SUMINC_E_INVALID, "bad weight '" + s + "'" + where(line_no));
  uint32_t bits;
  std::memcpy(&bits, &f, 4);
  return bits;
}

// Per-chunk parse result of an edge or update file.
struct Parsed {
  std::vector<uint32_t> u, v, w;
  std::vector<uint8_t> type;
  uint64_t skipped = 0;
};

// Calls fn(line_begin, line_end, line_number_within_chunk) for every line of a chunk.
template <class F>
void for_lines(const MappedFile& f, size_t b, size_t e, F&& fn) {
  size_t i = b, no = 0;
  while (i < e) {
    size_t j = i;
    while (j < e && f.p[j] != '\n') ++j;
    fn(f.p + i, f.p + j, no++);
    i = j + 1;
  }
}

// Line counts before every chunk, for 1-based line numbers in messages.
std::vector<size_t> chunk_line_bases(const MappedFile& f, const std::vector<std::pair<size_t, size_t>>& ch) {
  std::vector<size_t> nl(ch.size(), 0);
  run_parallel(ch.size(), [&](size_t i) {
    size_t c = 0;
    for (size_t k = ch[i].first; k < ch[i].second; ++k) c += f.p[k] == '\n';
    nl[i] = c;
  });
  std::vector<size_t> base(ch.size(), 1);
  for (size_t i = 1; i < ch.size(); ++i) base[i] = base[i - 1] + nl[i - 1];
  return base;
}

suminc_edges* read_edges(const char* path, int wkind, int directed) {
  SUMINC_REQUIRE(wkind >= SUMINC_W_NONE && wkind <= SUMINC_W_F32, SUMINC_E_INVALID, "bad weight kind");
  MappedFile f(path);
  const auto ch = line_chunks(f, io_threads());
  const auto base = chunk_line_bases(f, ch);
  std::vector<Parsed> parts(ch.size());
  run_parallel(ch.size(), [&](size_t i) {
    Parsed& P = parts[i];
    for_lines(f, ch[i].first, ch[i].second, [&](const char* b, const char* e, size_t no) {
      if (b == e || *b == '#') return;  // empty and comment lines
      Line L{b, e};
      if (!L.more()) return;
      const size_t line_no = base[i] + no;
      const char *tb, *te;
      uint64_t u = 0, v = 0;
      uint32_t w = 0;
      try {  // a line that does not parse is skipped, as the loaders skip it
        L.token(tb, te);
        u = parse_u64(tb, te, line_no, "source id");
        SUMINC_REQUIRE(L.token(tb, te), SUMINC_E_INVALID, "missing destination id" + where(line_no));
        v = parse_u64(tb, te, line_no, "destination id");
        SUMINC_REQUIRE(u < 0xFFFFFFFFull && v < 0xFFFFFFFFull, SUMINC_E_INVALID, "vertex id too large" + where(line_no));
        if (wkind != SUMINC_W_NONE) {
          SUMINC_REQUIRE(L.token(tb, te), SUMINC_E_INVALID, "missing weight" + where(line_no));
          w = parse_weight(tb, te, wkind, line_no);
        }
      } catch (const Error&) {
        ++P.skipped;
        return;
      }
      P.u.push_back((uint32_t)u);
      P.v.push_back((uint32_t)v);
      if (wkind != SUMINC_W_NONE) P.w.push_back(w);
      if (!directed && u != v) {
        P.u.push_back((uint32_t)v);
        P.v.push_back((uint32_t)u);
        if (wkind != SUMINC_W_NONE) P.w.push_back(w);
      }
    });
  });
  // global edge order = file order (chunk by chunk)
  std::vector<uint64_t> first(parts.size() + 1, 0);
  for (size_t i = 0; i < parts.size(); ++i) first[i + 1] = first[i] + parts[i].u.size();
  const uint64_t raw = first.back();
  uint64_t skipped = 0;
  for (auto& P : parts) skipped += P.skipped;
  uint32_t nmax = 0;
  for (auto& P : parts)
    for (size_t k = 0; k < P.u.size(); ++k) nmax = std::max(nmax, std::max(P.u[k], P.v[k]));
  const uint32_t n = raw ? nmax + 1 : 0;
  SUMINC_REQUIRE(raw < 0xFFFFFFFFull, SUMINC_E_TOO_LARGE, "edge count must stay below 2^32");
  // counting scatter by source, then each row sorted by (column, file order)
  std::vector<std::atomic<uint64_t>> cnt(n + 1);
  for (auto& x : cnt) x.store(0, std::memory_order_relaxed);
  run_parallel(parts.size(), [&](size_t i) {
    for (uint32_t u : parts[i].u) cnt[u + 1].fetch_add(1, std::memory_order_relaxed);
  });
  std::vector<uint64_t> rb(n + 1, 0);
  for (uint32_t x = 0; x < n; ++x) rb[x + 1] = rb[x] + cnt[x + 1].load(std::memory_order_relaxed);
  for (uint32_t x = 0; x <= n; ++x) cnt[x].store(rb[x], std::memory_order_relaxed);
  std::vector<uint64_t> slot_key(raw);  // (column << 32) | position within the chunk order, per row
  std::vector<uint32_t> slot_w(wkind ? raw : 0);
  std::vector<uint32_t> slot_idx(raw);
  run_parallel(parts.size(), [&](size_t i) {
    const Parsed& P = parts[i];
    for (size_t k = 0; k < P.u.size(); ++k) {
      const uint64_t s = cnt[P.u[k]].fetch_add(1, std::memory_order_relaxed);
      slot_key[s] = ((uint64_t)P.v[k] << 32) | (uint32_t)(first[i] + k);
      if (wkind) slot_w[s] = P.w[k];
    }
  });
  parts.clear();
  parts.shrink_to_fit();
  std::vector<uint64_t> keep_cnt(n + 1, 0);
  const unsigned T = io_threads();
  run_parallel(T, [&](size_t t) {
    for (uint64_t x = t; x < n; x += T) {
      const uint64_t b = rb[x], e = rb[x + 1];
      for (uint64_t s = b; s < e; ++s) slot_idx[s] = (uint32_t)(s - b);
      std::sort(slot_idx.begin() + b, slot_idx.begin() + e,
                [&](uint32_t a, uint32_t c) { return slot_key[b + a] < slot_key[b + c]; });
      uint64_t kept = 0;
      uint64_t prev = ~0ull;
      for (uint64_t s = b; s < e; ++s) {
        const uint64_t col = slot_key[b + slot_idx[s]] >> 32;
        if (col != prev) ++kept;
        prev = col;
      }
      keep_cnt[x + 1] = kept;
    }
  });
  suminc_edges* out = new suminc_edges();
  out->n = n;
  out->wkind = wkind;
  out->skipped = skipped;
  out->off.assign((size_t)n + 1, 0);
  for (uint32_t x = 0; x < n; ++x) out->off[x + 1] = out->off[x] + keep_cnt[x + 1];
  const uint64_t m = n ? out->off[n] : 0;
  out->col.resize(m);
  if (wkind) out->w.resize(m);
  run_parallel(T, [&](size_t t) {
    for (uint64_t x = t; x < n; x += T) {
      const uint64_t b = rb[x], e = rb[x + 1];
      uint64_t d = out->off[x];
      uint64_t prev = ~0ull;
      for (uint64_t s = b; s < e; ++s) {
        const uint64_t k = b + slot_idx[s];
        const uint64_t col = slot_key[k] >> 32;
        if (col == prev) continue;  // a later line of the same pair
        prev = col;
        out->col[d] = (uint32_t)col;
        if (wkind) out->w[d] = slot_w[k];
        ++d;
      }
    }
  });
  return out;
}

void read_updates(const char* path, int wkind, int directed, suminc_update* ops, uint64_t cap, uint64_t* count) {
  SUMINC_REQUIRE(count, SUMINC_E_INVALID, "null count");
  SUMINC_REQUIRE(wkind >= SUMINC_W_NONE && wkind <= SUMINC_W_F32, SUMINC_E_INVALID, "bad weight kind");
  MappedFile f(path);
  const auto ch = line_chunks(f, io_threads());
  const auto base = chunk_line_bases(f, ch);
  std::vector<Parsed> parts(ch.size());
  run_parallel(ch.size(), [&](size_t i) {
    Parsed& P = parts[i];
    for_lines(f, ch[i].first, ch[i].second, [&](const char* b, const char* e, size_t no) {
      if (b == e || *b == '#') return;
      Line L{b, e};
      if (!L.more()) return;
      const size_t line_no = base[i] + no;
      const char *tb, *te;
      L.token(tb, te);
      const std::string type(tb, te);
      SUMINC_REQUIRE(type == "a" || type == "d", SUMINC_E_INVALID, "invalid op '" + type + "'" + where(line_no));
      SUMINC_REQUIRE(L.token(tb, te), SUMINC_E_INVALID, "missing source id" + where(line_no));
      const uint64_t u = parse_u64(tb, te, line_no, "source id");
      SUMINC_REQUIRE(L.token(tb, te), SUMINC_E_INVALID, "missing destination id" + where(line_no));
      const uint64_t v = parse_u64(tb, te, line_no, "destination id");
      uint32_t w = type == "a" && wkind == SUMINC_W_U32 ? 1u : 0u;
      if (type == "a" && wkind != SUMINC_W_NONE) {
        SUMINC_REQUIRE(L.token(tb, te), SUMINC_E_INVALID, "missing weight" + where(line_no));
        w = parse_weight(tb, te, wkind, line_no);
      }
      const uint8_t t = type == "a" ? SUMINC_OP_ADD : SUMINC_OP_DEL;
      P.type.push_back(t);
      P.u.push_back((uint32_t)u);
      P.v.push_back((uint32_t)v);
      P.w.push_back(w);
      if (!directed && u != v) {
        P.type.push_back(t);
        P.u.push_back((uint32_t)v);
        P.v.push_back((uint32_t)u);
        P.w.push_back(w);
      }
    });
  });
  uint64_t total = 0;
  for (auto& P : parts) total += P.u.size();
  *count = total;
  if (!ops) return;
  SUMINC_REQUIRE(cap >= total, SUMINC_E_INVALID, "op capacity too small: " + std::to_string(total) + " ops");
  uint64_t k = 0;
  for (auto& P : parts)
    for (size_t i = 0; i < P.u.size(); ++i, ++k) {
      suminc_update& op = ops[k];
      std::memset(&op, 0, sizeof(op));
      op.type = P.type[i];
      op.u = P.u[i];
      op.v = P.v[i];
      if (wkind == SUMINC_W_F32) std::memcpy(&op.w_f32, &P.w[i], 4);
      else op.w_u32 = P.w[i];
    }
}

void read_clusters(const char* path, uint32_t n, uint32_t min_node, uint32_t* cluster_of, uint32_t* n_clusters) {
  SUMINC_REQUIRE(cluster_of, SUMINC_E_INVALID, "null output");
  MappedFile f(path);
  std::fill(cluster_of, cluster_of + n, SUMINC_INF_U32);
  Line L{f.p, f.p + f.size};
  uint32_t kept = 0;
  uint64_t record = 0;
  std::vector<uint32_t> mem;
  const char *tb, *te;
  while (L.token(tb, te)) {
    ++record;
    const uint64_t size = parse_u64(tb, te, 0, "cluster size");
    mem.clear();
    for (uint64_t i = 0; i < size; ++i) {
      SUMINC_REQUIRE(L.token(tb, te), SUMINC_E_INVALID,
                     "cluster " + std::to_string(record) + " ends before its " + std::to_string(size) + " ids");
      const uint64_t v = parse_u64(tb, te, 0, "vertex id");
      SUMINC_REQUIRE(v < n, SUMINC_E_INVALID,
                     "vertex id " + std::to_string(v) + " >= n in cluster " + std::to_string(record));
      mem.push_back((uint32_t)v);
    }
    std::sort(mem.begin(), mem.end());
    mem.erase(std::unique(mem.begin(), mem.end()), mem.end());  // the reference collects a std::set
    if (mem.size() < min_node) continue;
    for (uint32_t v : mem) cluster_of[v] = kept;
    ++kept;
  }
  if (n_clusters) *n_clusters = kept;
}

void write_clusters(const char* path, uint32_t n, const uint32_t* cluster_of) {
  SUMINC_REQUIRE(path && cluster_of, SUMINC_E_INVALID, "null argument");
  uint32_t K = 0;
  for (uint32_t v = 0; v < n; ++v)
    if (cluster_of[v] != SUMINC_INF_U32) {
      SUMINC_REQUIRE(cluster_of[v] < n, SUMINC_E_INVALID, "cluster id must be < n or SUMINC_INF_U32");
      K = std::max(K, cluster_of[v] + 1);
    }
  std::vector<uint64_t> off((size_t)K + 1, 0);
  for (uint32_t v = 0; v < n; ++v)
    if (cluster_of[v] != SUMINC_INF_U32) ++off[cluster_of[v] + 1];
  for (uint32_t k = 0; k < K; ++k) off[k + 1] += off[k];
  std::vector<uint32_t> mem(off[K]);
  std::vector<uint64_t> cur(off.begin(), off.end() - 1);
  for (uint32_t v = 0; v < n; ++v)
    if (cluster_of[v] != SUMINC_INF_U32) mem[cur[cluster_of[v]]++] = v;  // ascending ids
  FILE* fp = std::fopen(path, "w");
  SUMINC_REQUIRE(fp, SUMINC_E_INVALID, std::string("cannot write ") + path + ": " + std::strerror(errno));
  std::string buf;
  bool ok = true;
  for (uint32_t k = 0; k < K && ok; ++k) {
    if (off[k + 1] == off[k]) continue;
    buf = std::to_string(off[k + 1] - off[k]);
    for (uint64_t i = off[k]; i < off[k + 1]; ++i) {
      buf += ' ';
      buf += std::to_string(mem[i]);
    }
    buf += '\n';
    ok = std::fwrite(buf.data(), 1, buf.size(), fp) == buf.size();
  }
  ok = (std::fclose(fp) == 0) && ok;
  SUMINC_REQUIRE(ok, SUMINC_E_INVALID, std::string("write failed: ") + path);
}

template <class F>
int host_guarded(F&& f) {
  try {
    f();
    t_err.clear();
    return SUMINC_OK;
  } catch (const Error& e) {
    t_err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    t_err = "host allocation failed";
    return SUMINC_E_NOMEM;
  } catch (const std::exception& e) {
    t_err = e.what();
    return SUMINC_E_INVALID;
  }
}

}  // namespace

const char* io_last_error() { return t_err.c_str(); }

int host_call(const std::function<void()>& f) { return host_guarded(f); }

void upload_edges(suminc_ctx* c, const suminc_edges* e) {
  SUMINC_REQUIRE(e, SUMINC_E_INVALID, "null edge list");
  SUMINC_REQUIRE(e->n >= 1, SUMINC_E_INVALID, "empty edge list");
  upload_csr(c, e->n, e->col.size(), e->off.data(), e->col.data(), e->wkind, e->wkind ? e->w.data() : nullptr);
}

}  // namespace suminc

extern "C" {

int suminc_edges_read_tsv(const char* path, int weight_kind, int directed, suminc_edges** out) {
  return suminc::host_guarded([&] {
    SUMINC_REQUIRE(out, SUMINC_E_INVALID, "null output");
    *out = suminc::read_edges(path, weight_kind, directed);
  });
}

int suminc_edges_info(const suminc_edges* e, uint32_t* n, uint64_t* m, int* weight_kind, uint64_t* skipped_lines) {
  return suminc::host_guarded([&] {
    SUMINC_REQUIRE(e, SUMINC_E_INVALID, "null edge list");
    if (n) *n = e->n;
    if (m) *m = e->col.size();
    if (weight_kind) *weight_kind = e->wkind;
    if (skipped_lines) *skipped_lines = e->skipped;
  });
}

int suminc_edges_copy(const suminc_edges* e, uint64_t* offsets, uint32_t* cols, void* weights) {
  return suminc::host_guarded([&] {
    SUMINC_REQUIRE(e, SUMINC_E_INVALID, "null edge list");
    if (offsets) std::memcpy(offsets, e->off.data(), e->off.size() * 8);
    if (cols && !e->col.empty()) std::memcpy(cols, e->col.data(), e->col.size() * 4);
    if (weights && e->wkind && !e->w.empty()) std::memcpy(weights, e->w.data(), e->w.size() * 4);
  });
}

void suminc_edges_free(suminc_edges* e) { delete e; }

int suminc_updates_read_tsv(const char* path, int weight_kind, int directed, suminc_update* ops, uint64_t cap,
                            uint64_t* count) {
  return suminc::host_guarded([&] { suminc::read_updates(path, weight_kind, directed, ops, cap, count); });
}

int suminc_clusters_read(const char* path, uint32_t n, uint32_t min_node_num, uint32_t* cluster_of,
                         uint32_t* n_clusters) {
  return suminc::host_guarded([&] { suminc::read_clusters(path, n, min_node_num, cluster_of, n_clusters); });
}

int suminc_clusters_write(const char* path, uint32_t n, const uint32_t* cluster_of) {
  return suminc::host_guarded([&] { suminc::write_clusters(path, n, cluster_of); });
}

}  // extern "C"
