// oracle.cpp -- CPU restatement of the reference's hot path.
//
// THIS IS TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only tests/, the smoke check
// in __graft_entry__.py and the cpu_baseline leg of bench.py load this library,
// and only as the checker / the timed CPU baseline.  The product library
// (suminc_amd/csrc) never links, loads or calls anything in this directory.
//
// Every function restates an algorithm of tang0113/SumInc (paths relative to the
// reference root) in plain C++17 + OpenMP with a C ABI for ctypes.  Arithmetic is
// FP64 unless a function says otherwise.  Pinning (what each function is checked
// against, beyond "the GPU agrees with it") is listed in the header comment of
// each function and in tests/test_oracle.py:
//   * orc_philox4x32_10  -- Random123 known-answer vectors.
//   * orc_rmat_*         -- brute-force recount of a tiny graph; dedupe/self-loop
//                           invariants; quadrant frequencies vs a,b,c,d.
//   * orc_apply_batch    -- brute-force set semantics on tiny inputs.
//   * orc_pr_ingress     -- scipy sparse direct solve of (I - d P^T) r = (1-d)/N
//                           (closed form), mass invariant, error bound.
//   * orc_pr_amend       -- mass invariant, reseeded fixed point == scratch solve.
//   * orc_sssp_dijkstra  -- scipy.sparse.csgraph.dijkstra, edge invariants.
//   * orc_sum_pr         -- equals the closed form (same fixed point), shortcut
//                           rows equal a dense per-cluster solve.
//   * orc_sum_sssp*      -- equals orc_sssp_dijkstra / scipy Dijkstra (distances
//                           and smallest-id parents); shortcut rows equal scipy
//                           Dijkstra on the cluster's induced subgraph.
//   * orc_wpr_*          -- dense numpy solve + central finite differences.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <queue>
#include <unordered_map>
#include <unordered_set>
#include <vector>
#include <omp.h>

namespace {

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11).  Multipliers and Weyl key increments from that paper.
// ---------------------------------------------------------------------------
inline void philox(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Stream ids of the generator (the product's device generator uses the same
// numbers; they are part of the synthetic-input definition, not of SumInc).
constexpr uint32_t STREAM_RMAT = 0x524D4154u;   // "RMAT"
constexpr uint32_t STREAM_WEIGHT = 0x57474854u; // "WGHT"
constexpr uint32_t STREAM_DEL = 0x44454C45u;    // "DELE"
constexpr uint32_t STREAM_INS = 0x494E5345u;    // "INSE"

inline void seed_key(uint64_t seed, uint32_t key[2]) {
  key[0] = (uint32_t)seed;
  key[1] = (uint32_t)(seed >> 32);
}

inline uint32_t mulhi32(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) >> 32); }

struct Graph {
  uint32_t n = 0;
  uint64_t m = 0;
  std::vector<uint64_t> off;   // n+1
  std::vector<uint32_t> col;   // m
  std::vector<uint32_t> wu;    // m or empty
  std::vector<float> wf;       // m or empty
  uint32_t deg(uint32_t u) const { return (uint32_t)(off[u + 1] - off[u]); }
};

// Weight of edge (u,v): a pure function of (seed, u, v) so that it does not
// depend on duplicate order.  kind 1: u32 uniform in [wmin,wmax]; kind 2: f32 in
// (0,1] with 24 random bits.
inline void edge_weight(uint64_t seed, uint32_t u, uint32_t v, int kind, uint32_t wmin,
                        uint32_t wmax, uint32_t* wu, float* wf) {
  uint32_t key[2], ctr[4] = {u, v, 0u, STREAM_WEIGHT}, o[4];
  seed_key(seed, key);
  philox(ctr, key, o);
  if (kind == 1) *wu = wmin + mulhi32(o[0], wmax - wmin + 1u);
  if (kind == 2) *wf = (float)((o[0] >> 8) + 1u) * (1.0f / 16777216.0f);
}

void build_csr_from_sorted_keys(Graph& g, uint32_t n, const std::vector<uint64_t>& keys) {
  g.n = n;
  g.m = keys.size();
  g.off.assign((size_t)n + 1, 0);
  g.col.resize(g.m);
  for (uint64_t i = 0; i < g.m; ++i) {
    g.off[(keys[i] >> 32) + 1]++;
    g.col[i] = (uint32_t)keys[i];
  }
  for (uint32_t i = 0; i < n; ++i) g.off[i + 1] += g.off[i];
}

}  // namespace

extern "C" {

struct orc_update {  // one line of an update file: "a u v w" (type 0) or "d u v" (type 1)
  uint32_t type, u, v, w_u32;
  float w_f32;
  uint32_t pad;
};

void orc_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  philox(ctr, key, out);
}

int orc_num_threads(void) { return omp_get_max_threads(); }
void orc_set_threads(int t) { if (t > 0) omp_set_num_threads(t); }

// ---------------------------------------------------------------------------
// Graph handle
// ---------------------------------------------------------------------------
void* orc_graph_from_csr(uint32_t n, uint64_t m, const uint64_t* off, const uint32_t* col,
                         const uint32_t* wu, const float* wf) {
  Graph* g = new Graph();
  g->n = n;
  g->m = m;
  g->off.assign(off, off + n + 1);
  g->col.assign(col, col + m);
  if (wu) g->wu.assign(wu, wu + m);
  if (wf) g->wf.assign(wf, wf + m);
  return g;
}

void orc_graph_info(void* h, uint32_t* n, uint64_t* m, int* has_wu, int* has_wf) {
  Graph* g = (Graph*)h;
  *n = g->n;
  *m = g->m;
  *has_wu = !g->wu.empty();
  *has_wf = !g->wf.empty();
}

void orc_graph_copy(void* h, uint64_t* off, uint32_t* col, uint32_t* wu, float* wf) {
  Graph* g = (Graph*)h;
  if (off) std::copy(g->off.begin(), g->off.end(), off);
  if (col) std::copy(g->col.begin(), g->col.end(), col);
  if (wu && !g->wu.empty()) std::copy(g->wu.begin(), g->wu.end(), wu);
  if (wf && !g->wf.empty()) std::copy(g->wf.begin(), g->wf.end(), wf);
}

void orc_graph_free(void* h) { delete (Graph*)h; }

// ---------------------------------------------------------------------------
// R-MAT (Chakrabarti, Zhan, Faloutsos, SDM'04) without noise or permutation:
// edge i descends `scale` levels; at level l one Philox word (ctr = {i_lo, i_hi,
// l/4, STREAM_RMAT}, word l%4) picks quadrant a|b|c|d by comparing against
// floor(a*2^32), floor((a+b)*2^32), floor((a+b+c)*2^32).  Source bit first is the
// MSB.  Self-loops dropped, duplicates merged, CSR sorted by (u, v).
// ---------------------------------------------------------------------------
void* orc_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
               int weight_kind, uint32_t wmin, uint32_t wmax) {
  const uint32_t n = 1u << scale;
  const uint64_t raw = (uint64_t)edge_factor << scale;
  const uint64_t t1 = (uint64_t)(a * 4294967296.0);
  const uint64_t t2 = (uint64_t)((a + b) * 4294967296.0);
  const uint64_t t3 = (uint64_t)((a + b + c) * 4294967296.0);
  uint32_t key[2];
  seed_key(seed, key);
  std::vector<uint64_t> keys(raw);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)raw; ++i) {
    uint32_t u = 0, v = 0, o[4] = {0, 0, 0, 0};
    for (int l = 0; l < scale; ++l) {
      if ((l & 3) == 0) {
        uint32_t ctr[4] = {(uint32_t)i, (uint32_t)((uint64_t)i >> 32), (uint32_t)(l >> 2), STREAM_RMAT};
        philox(ctr, key, o);
      }
      uint64_t r = o[l & 3];
      uint32_t bu = 0, bv = 0;
      if (r < t1) { bu = 0; bv = 0; }
      else if (r < t2) { bu = 0; bv = 1; }
      else if (r < t3) { bu = 1; bv = 0; }
      else { bu = 1; bv = 1; }
      u |= bu << (scale - 1 - l);
      v |= bv << (scale - 1 - l);
    }
    keys[i] = (u == v) ? ~0ull : (((uint64_t)u << 32) | v);
  }
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  if (!keys.empty() && keys.back() == ~0ull) keys.pop_back();
  Graph* g = new Graph();
  build_csr_from_sorted_keys(*g, n, keys);
  if (weight_kind == 1) g->wu.resize(g->m);
  if (weight_kind == 2) g->wf.resize(g->m);
  if (weight_kind) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)g->m; ++i) {
      uint32_t u = (uint32_t)(keys[i] >> 32), v = (uint32_t)keys[i];
      edge_weight(seed, u, v, weight_kind, wmin, wmax, weight_kind == 1 ? &g->wu[i] : nullptr,
                  weight_kind == 2 ? &g->wf[i] : nullptr);
    }
  }
  return g;
}

// ---------------------------------------------------------------------------
// Update batch generator (synthetic input definition, mirrors the device one):
// n_del deletes of existing edges picked by edge index (64-bit Philox draw scaled
// to m), then n_ins inserts with uniform endpoints (v bumped by one when it equals
// u) and a fresh weight.  Ops are emitted deletes first, in draw order.
// ---------------------------------------------------------------------------
void orc_gen_batch(void* h, uint64_t seed, uint32_t batch_id, uint64_t n_del, uint64_t n_ins,
                   int weight_kind, uint32_t wmin, uint32_t wmax, orc_update* ops) {
  Graph* g = (Graph*)h;
  uint32_t key[2];
  seed_key(seed, key);
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < (int64_t)n_del; ++j) {
    uint32_t ctr[4] = {(uint32_t)j, (uint32_t)((uint64_t)j >> 32), batch_id, STREAM_DEL}, o[4];
    philox(ctr, key, o);
    uint64_t r = ((uint64_t)o[0] << 32) | o[1];
    uint64_t idx = (uint64_t)(((unsigned __int128)r * g->m) >> 64);
    uint32_t u = (uint32_t)(std::upper_bound(g->off.begin(), g->off.end(), idx) - g->off.begin() - 1);
    orc_update& op = ops[j];
    op.type = 1; op.u = u; op.v = g->col[idx]; op.w_u32 = 0; op.w_f32 = 0.f; op.pad = 0;
  }
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < (int64_t)n_ins; ++j) {
    uint32_t ctr[4] = {(uint32_t)j, (uint32_t)((uint64_t)j >> 32), batch_id, STREAM_INS}, o[4];
    philox(ctr, key, o);
    uint32_t u = mulhi32(o[0], g->n), v = mulhi32(o[1], g->n);
    if (v == u) v = (v + 1u == g->n) ? 0u : v + 1u;
    orc_update& op = ops[n_del + j];
    op.type = 0; op.u = u; op.v = v; op.pad = 0;
    op.w_u32 = (weight_kind == 1) ? wmin + mulhi32(o[2], wmax - wmin + 1u) : 0u;
    op.w_f32 = (weight_kind == 2) ? (float)((o[2] >> 8) + 1u) * (1.0f / 16777216.0f) : 0.f;
  }
}

// ---------------------------------------------------------------------------
// IncFragmentBuilder::Init/Build (grape/fragment/inc_fragment_builder.h:29-133):
// the first "a u v w" of a pair wins (unordered_map::emplace, :57); every added
// pair is emitted (:83-93); an old edge survives iff it is neither deleted nor
// re-added (:105-116).  So an add beats a delete of the same pair and an add of
// an existing pair replaces its weight.  touched[u] = 1 for every u named as a
// source by any op (sum_sync_iter_worker.h:247-283 marks is_update the same way).
// ---------------------------------------------------------------------------
void* orc_apply_batch(void* h, const orc_update* ops, uint64_t n_ops, uint8_t* touched) {
  Graph* g = (Graph*)h;
  std::unordered_map<uint64_t, std::pair<uint32_t, float>> added;
  std::unordered_set<uint64_t> deleted;
  added.reserve(n_ops * 2);
  deleted.reserve(n_ops * 2);
  for (uint64_t i = 0; i < n_ops; ++i) {
    uint64_t k = ((uint64_t)ops[i].u << 32) | ops[i].v;
    if (ops[i].type == 0) added.emplace(k, std::make_pair(ops[i].w_u32, ops[i].w_f32));
    else deleted.insert(k);
    if (touched) touched[ops[i].u] = 1;
  }
  const bool hu = !g->wu.empty(), hf = !g->wf.empty();
  std::vector<std::pair<uint64_t, std::pair<uint32_t, float>>> edges;
  edges.reserve(g->m + added.size());
  for (auto& kv : added) edges.push_back(kv);
  for (uint32_t u = 0; u < g->n; ++u) {
    for (uint64_t e = g->off[u]; e < g->off[u + 1]; ++e) {
      uint64_t k = ((uint64_t)u << 32) | g->col[e];
      if (deleted.count(k) || added.count(k)) continue;
      edges.push_back({k, {hu ? g->wu[e] : 0u, hf ? g->wf[e] : 0.f}});
    }
  }
  std::sort(edges.begin(), edges.end(),
            [](const auto& x, const auto& y) { return x.first < y.first; });
  std::vector<uint64_t> keys(edges.size());
  for (size_t i = 0; i < edges.size(); ++i) keys[i] = edges[i].first;
  Graph* ng = new Graph();
  build_csr_from_sorted_keys(*ng, g->n, keys);
  if (hu) { ng->wu.resize(ng->m); for (size_t i = 0; i < edges.size(); ++i) ng->wu[i] = edges[i].second.first; }
  if (hf) { ng->wf.resize(ng->m); for (size_t i = 0; i < edges.size(); ++i) ng->wf[i] = edges[i].second.second; }
  return ng;
}

// ---------------------------------------------------------------------------
// Plain delta-accumulative PageRank, the reference's CPU Ingress branch
// (grape/worker/sum_sync_iter_worker.h:857-882) with
//   init_c: delta = (1-d)/N, init_v: r = 0   (pagerank/pagerank_ingress.h:27-32)
//   g_function: if outdeg>0, delta_w += delta*d/outdeg   (pagerank_ingress.h:50-72)
//   isChange: |delta| > T/N                  (sum_sync_iter_worker.h:2914-2925)
//   termCheck: sum_v |delta_v| < T (FP64)    (sum_sync_iter_worker.h:2873-2912)
//   stop when termCheck holds or step > max_rounds (:1372, pr_mr default 100).
// The vertex loop is an OpenMP parallel for with CAS-loop atomic adds like the
// reference's cilk_for + atomic_add (grape/utils/atomic_ops.h:171-187), so the
// rounding order is run-dependent, as in the reference.  If init != 0 the
// caller's r/delta are used as the starting state (incremental use).
// stats[0]=rounds, [1]=edges traversed (f_send_delta_num), [2]=final sum|delta|.
// ---------------------------------------------------------------------------
static inline void atomic_add_f64(double* p, double v) {
  std::atomic_ref<double> a(*p);
  double cur = a.load(std::memory_order_relaxed);
  while (!a.compare_exchange_weak(cur, cur + v, std::memory_order_relaxed)) {}
}

int orc_pr_ingress(void* h, double d, double T, int max_rounds, int init, double* r,
                   double* delta, double* stats) {
  Graph* g = (Graph*)h;
  const uint32_t n = g->n;
  if (init) {
    for (uint32_t v = 0; v < n; ++v) { r[v] = 0.0; delta[v] = (1.0 - d) / n; }
  }
  const double eps = T / n;
  int step = 0;
  double edges = 0, sum = 0;
  while (true) {
    ++step;
    double e_round = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : e_round)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
      uint32_t u = (uint32_t)i;
      std::atomic_ref<double> du(delta[u]);
      if (std::fabs(du.load(std::memory_order_relaxed)) > eps) {
        double dv = du.exchange(0.0, std::memory_order_relaxed);
        uint32_t deg = g->deg(u);
        if (deg > 0) {
          double outv = dv * d / deg;
          for (uint64_t e = g->off[u]; e < g->off[u + 1]; ++e) atomic_add_f64(&delta[g->col[e]], outv);
          e_round += deg;
        }
        r[u] += dv;  // only this iteration touches r[u]
      }
    }
    edges += e_round;
    sum = 0;
#pragma omp parallel for reduction(+ : sum)
    for (int64_t i = 0; i < (int64_t)n; ++i) sum += std::fabs(delta[i]);
    if (sum < T || step > max_rounds) break;
  }
  if (stats) { stats[0] = step; stats[1] = edges; stats[2] = sum; }
  return step;
}

// Jacobi form of the same iteration with no per-vertex threshold: used to reach
// a tight FP64 fixed point (the "FP64 run") quickly.  Stops when sum|delta| < tol.
int orc_pr_solve(void* h, double d, double tol, int max_rounds, double* r, double* delta_io,
                 int init) {
  Graph* g = (Graph*)h;
  const uint32_t n = g->n;
  std::vector<double> cur(n), nxt(n);
  if (init) { for (uint32_t v = 0; v < n; ++v) { r[v] = 0; cur[v] = (1.0 - d) / n; } }
  else { for (uint32_t v = 0; v < n; ++v) cur[v] = delta_io[v]; }
  int step = 0;
  for (; step < max_rounds; ++step) {
    std::fill(nxt.begin(), nxt.end(), 0.0);
    double sum = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : sum)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
      uint32_t u = (uint32_t)i;
      double dv = cur[u];
      if (dv == 0.0) continue;
      sum += std::fabs(dv);
      r[u] += dv;
      uint32_t deg = g->deg(u);
      if (!deg) continue;
      double outv = dv * d / deg;
      for (uint64_t e = g->off[u]; e < g->off[u + 1]; ++e) atomic_add_f64(&nxt[g->col[e]], outv);
    }
    std::swap(cur, nxt);
    if (sum < tol) { ++step; break; }
  }
  if (delta_io) for (uint32_t v = 0; v < n; ++v) delta_io[v] = cur[v];
  return step;
}

// AmendValue(type) (sum_sync_iter_worker.h:139-180): for every source u (or only
// those with touched[u] != 0 when touched is given -- the exact-arithmetic
// equivalent the reference left commented out at :187-208), push
// g_function(delta = type * r_u) along u's out-edges of graph h.
void orc_pr_amend(void* h, double d, int type, const double* r, double* delta,
                  const uint8_t* touched) {
  Graph* g = (Graph*)h;
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < (int64_t)g->n; ++i) {
    uint32_t u = (uint32_t)i;
    if (touched && !touched[u]) continue;
    uint32_t deg = g->deg(u);
    if (!deg || r[u] == 0.0) continue;
    double outv = type * r[u] * d / deg;
    for (uint64_t e = g->off[u]; e < g->off[u + 1]; ++e) atomic_add_f64(&delta[g->col[e]], outv);
  }
}

// ---------------------------------------------------------------------------
// SSSP.  Dijkstra with a binary heap: dist[v] (0xFFFFFFFF = unreachable, the
// reference's INT32_MAX sentinel mapped to u32) and parent[v] = the smallest id u
// with dist[u] + w(u,v) == dist[v] (source: itself; unreachable: itself, like
// GetInitDelta's GenDelta(v, inf) in sssp/sssp_ingress.h:30-41).
// ---------------------------------------------------------------------------
void orc_sssp_dijkstra(void* h, uint32_t src, uint32_t* dist, uint32_t* parent) {
  Graph* g = (Graph*)h;
  const uint32_t INF = 0xFFFFFFFFu;
  std::vector<uint64_t> best(g->n, ~0ull);
  for (uint32_t v = 0; v < g->n; ++v) { dist[v] = INF; parent[v] = v; }
  typedef std::pair<uint64_t, uint32_t> item;
  std::priority_queue<item, std::vector<item>, std::greater<item>> pq;
  best[src] = 0;
  pq.push({0, src});
  while (!pq.empty()) {
    auto [du, u] = pq.top();
    pq.pop();
    if (du != best[u] || dist[u] != INF) continue;
    dist[u] = (uint32_t)du;
    for (uint64_t e = g->off[u]; e < g->off[u + 1]; ++e) {
      uint32_t v = g->col[e];
      uint64_t nd = du + (g->wu.empty() ? 1u : g->wu[e]);
      if (nd < best[v]) { best[v] = nd; pq.push({nd, v}); }
    }
  }
  // parents: min-id predecessor on a shortest path
  for (uint32_t v = 0; v < g->n; ++v) if (v != src && dist[v] != INF) parent[v] = INF;
  for (uint32_t u = 0; u < g->n; ++u) {
    if (dist[u] == INF) continue;
    for (uint64_t e = g->off[u]; e < g->off[u + 1]; ++e) {
      uint32_t v = g->col[e];
      if (v == src) continue;
      uint64_t nd = (uint64_t)dist[u] + (g->wu.empty() ? 1u : g->wu[e]);
      if (nd == dist[v] && u < parent[v]) parent[v] = u;
    }
  }
}

// The reference's weighted SSSP Ingress loop (the -cilk=false branch of
// sssp/sssp_ingress.h:131-175 driven by sum_sync_traversal_worker.h:1146-1159):
// value/delta pairs, CombineValueDelta(value, delta) then push delta.value + w
// with SetIfLessThanAtomic (grape/utils/dependency_data.h:21-34, a CAS on the
// {parent, value} pair) to every out-neighbour, activating it on success.
// Frontier-synchronous like the reference's curr_modified_/next_modified_.
// stats[0] = rounds, stats[1] = edges traversed.
int orc_sssp_ingress(void* h, uint32_t src, uint32_t* dist, uint32_t* parent, double* stats) {
  Graph* g = (Graph*)h;
  const uint32_t n = g->n, INF = 0xFFFFFFFFu;
  // packed {value:hi32, parent:lo32}; atomic min == SetIfLessThanAtomic with the
  // smaller parent id winning ties (the reference keeps whichever CAS lands first).
  std::vector<uint64_t> delta(n);
  std::vector<uint32_t> value(n, INF);
  for (uint32_t v = 0; v < n; ++v) delta[v] = ((uint64_t)INF << 32) | v;
  delta[src] = ((uint64_t)0 << 32) | src;
  std::vector<uint8_t> cur(n, 0), nxt(n, 0);
  cur[src] = 1;
  int rounds = 0;
  double edges = 0;
  bool any = true;
  while (any) {
    ++rounds;
    double e_round = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : e_round)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
      uint32_t u = (uint32_t)i;
      if (!cur[u]) continue;
      cur[u] = 0;
      uint64_t du = std::atomic_ref<uint64_t>(delta[u]).load(std::memory_order_relaxed);
      uint32_t dv = (uint32_t)(du >> 32);
      if (dv >= value[u]) continue;  // CombineValueDelta false
      value[u] = dv;
      for (uint64_t e = g->off[u]; e < g->off[u + 1]; ++e) {
        uint32_t v = g->col[e];
        uint64_t cand = ((uint64_t)(dv + (g->wu.empty() ? 1u : g->wu[e])) << 32) | u;
        std::atomic_ref<uint64_t> a(delta[v]);
        uint64_t old = a.load(std::memory_order_relaxed);
        while (cand < old && !a.compare_exchange_weak(old, cand, std::memory_order_relaxed)) {}
        if (cand < old) nxt[v] = 1;
      }
      e_round += (double)g->deg(u);
    }
    edges += e_round;
    any = false;
    for (uint32_t v = 0; v < n; ++v) if (nxt[v]) { any = true; break; }
    std::swap(cur, nxt);
  }
  for (uint32_t v = 0; v < n; ++v) {
    dist[v] = value[v];
    parent[v] = (value[v] == INF) ? v : (uint32_t)delta[v];
  }
  if (stats) { stats[0] = rounds; stats[1] = edges; }
  return rounds;
}

// The reference's incremental traversal step restated (deltaCompute,
// sum_sync_traversal_worker.h:113-352, then the Ingress rounds of Query):
//  1. roots: every v whose parent edge (parent(v), v) is gone from the new graph
//     (:145-173 walks the deleted pairs; a pair whose weight changed is a root too
//     here, since keeping its old distance would be stale);
//  2. reset: frontier walk over the OLD out-edges, children x with parent(x) == u
//     join; reset vertices get value/delta = identity (:180-260);
//  3. resend: every vertex with a finite value pushes value + w along its NEW
//     out-edges once, unconditionally (:316-329);
//  4. Ingress rounds from the vertices that improved, as orc_sssp_ingress.
// dist/parent: in = state on the old graph, out = state on the new graph.
// stats: [0] rounds, [1] edges traversed (resend + rounds), [2] reset vertices.
void orc_sssp_inc(void* h_old, void* h_new, uint32_t src, uint32_t* dist, uint32_t* parent, double* stats) {
  Graph* go = (Graph*)h_old;
  Graph* g = (Graph*)h_new;
  const uint32_t n = g->n, INF = 0xFFFFFFFFu;
  auto find = [](const Graph* gr, uint32_t u, uint32_t v) -> int64_t {
    const uint32_t* b = gr->col.data() + gr->off[u];
    const uint32_t* e = gr->col.data() + gr->off[u + 1];
    const uint32_t* it = std::lower_bound(b, e, v);
    return (it != e && *it == v) ? (int64_t)(gr->off[u] + (it - b)) : -1;
  };
  std::vector<uint8_t> cur(n, 0), nxt(n, 0), reset(n, 0);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const uint32_t v = (uint32_t)i, p = parent[v];
    if (v == src || dist[v] == INF || p == v) continue;
    const int64_t eo = find(go, p, v), en = find(g, p, v);
    const uint32_t wo = eo < 0 ? 1u : (go->wu.empty() ? 1u : go->wu[eo]);
    const uint32_t wn = en < 0 ? 0u : (g->wu.empty() ? 1u : g->wu[en]);
    if (en < 0 || wn != wo) cur[v] = 1;
  }
  double n_reset = 0;
  bool any = true;
  while (any) {
    any = false;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : n_reset)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
      const uint32_t u = (uint32_t)i;
      if (!cur[u]) continue;
      for (uint64_t e = go->off[u]; e < go->off[u + 1]; ++e) {
        const uint32_t x = go->col[e];
        if (x != u && !reset[x] && !cur[x] && parent[x] == u && dist[x] != INF) nxt[x] = 1;
      }
      n_reset += 1;
    }
    for (uint32_t u = 0; u < n; ++u)
      if (cur[u]) { reset[u] = 1; cur[u] = 0; }
    for (uint32_t u = 0; u < n; ++u)
      if (nxt[u] && !reset[u]) { cur[u] = 1; any = true; }
    std::fill(nxt.begin(), nxt.end(), 0);
  }
  std::vector<uint64_t> delta(n);
  std::vector<uint32_t> value(n);
  for (uint32_t v = 0; v < n; ++v) {
    if (reset[v]) { value[v] = INF; delta[v] = ((uint64_t)INF << 32) | v; }
    else { value[v] = dist[v]; delta[v] = ((uint64_t)dist[v] << 32) | parent[v]; }
  }
  auto wt = [&](uint64_t e) -> uint64_t { return g->wu.empty() ? 1u : g->wu[e]; };
  double edges = 0;
  // resend
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : edges)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const uint32_t u = (uint32_t)i;
    if (value[u] == INF) continue;
    for (uint64_t e = g->off[u]; e < g->off[u + 1]; ++e) {
      const uint32_t v = g->col[e];
      const uint64_t nd = (uint64_t)value[u] + wt(e);
      if (nd >= INF) continue;
      const uint64_t cand = (nd << 32) | u;
      std::atomic_ref<uint64_t> a(delta[v]);
      uint64_t old = a.load(std::memory_order_relaxed);
      while (cand < old && !a.compare_exchange_weak(old, cand, std::memory_order_relaxed)) {}
      if (cand < old && (old >> 32) > nd) cur[v] = 1;
    }
    edges += (double)g->deg(u);
  }
  // Ingress rounds
  int rounds = 0;
  any = true;
  while (any) {
    any = false;
    for (uint32_t v = 0; v < n; ++v) if (cur[v]) { any = true; break; }
    if (!any) break;
    ++rounds;
    double e_round = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : e_round)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
      const uint32_t u = (uint32_t)i;
      if (!cur[u]) continue;
      cur[u] = 0;
      const uint64_t du = std::atomic_ref<uint64_t>(delta[u]).load(std::memory_order_relaxed);
      const uint32_t dv = (uint32_t)(du >> 32);
      if (dv >= value[u]) continue;
      value[u] = dv;
      for (uint64_t e = g->off[u]; e < g->off[u + 1]; ++e) {
        const uint32_t v = g->col[e];
        const uint64_t nd = (uint64_t)dv + wt(e);
        if (nd >= INF) continue;
        const uint64_t cand = (nd << 32) | u;
        std::atomic_ref<uint64_t> a(delta[v]);
        uint64_t old = a.load(std::memory_order_relaxed);
        while (cand < old && !a.compare_exchange_weak(old, cand, std::memory_order_relaxed)) {}
        if (cand < old && (old >> 32) > nd) nxt[v] = 1;
      }
      e_round += (double)g->deg(u);
    }
    edges += e_round;
    std::swap(cur, nxt);
  }
  for (uint32_t v = 0; v < n; ++v) {
    const uint32_t dv = (uint32_t)(delta[v] >> 32);
    dist[v] = std::min(value[v], dv);
    parent[v] = dist[v] == INF ? v : (uint32_t)delta[v];
  }
  if (stats) { stats[0] = rounds; stats[1] = edges; stats[2] = n_reset; }
}

// ---------------------------------------------------------------------------
// Supernode compression (mirror-free form, mirror_k = 1e8)
// ---------------------------------------------------------------------------

// Leaf clustering (the build's stand-in for the reference's external Louvain
// file): v joins h when h is v's only neighbour in or out; a leaf pair anchors on
// its smaller id; clusters below min_node are dropped, larger ones keep the first
// max_node - 1 leaves by id.  Cluster ids are dense in anchor order.
void orc_leaf_clusters(void* h, uint32_t max_node, uint32_t min_node, uint32_t* raw) {
  Graph* g = (Graph*)h;
  const uint32_t n = g->n, INF = 0xFFFFFFFFu;
  std::vector<std::vector<uint32_t>> in(n);
  for (uint32_t u = 0; u < n; ++u)
    for (uint64_t e = g->off[u]; e < g->off[u + 1]; ++e) in[g->col[e]].push_back(u);
  std::vector<uint32_t> anchor(n, INF);
  for (uint32_t v = 0; v < n; ++v) {
    std::vector<uint32_t> nb;
    for (uint64_t e = g->off[v]; e < g->off[v + 1]; ++e) nb.push_back(g->col[e]);
    for (uint32_t u : in[v]) nb.push_back(u);
    std::sort(nb.begin(), nb.end());
    nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
    if (nb.size() == 1) anchor[v] = nb[0];
  }
  std::vector<std::vector<uint32_t>> leaves(n);
  for (uint32_t v = 0; v < n; ++v) {
    uint32_t a = anchor[v];
    if (a == INF) continue;
    if (anchor[a] == v && v < a) continue;  // v anchors the pair
    leaves[a].push_back(v);                 // v ascending
  }
  for (uint32_t v = 0; v < n; ++v) raw[v] = INF;
  uint32_t id = 0;
  for (uint32_t a = 0; a < n; ++a) {
    if (leaves[a].empty()) continue;
    size_t take = std::min<size_t>(leaves[a].size(), max_node - 1);
    if (take + 1 < min_node) continue;
    raw[a] = id;
    for (size_t i = 0; i < take; ++i) raw[leaves[a][i]] = id;
    ++id;
  }
}

// compress_by_cluster + get_init_supernode_by_clusters (mirror-free branch) +
// judge_out_bound_node + get_nodetype_mirror (compressor_base.h:3160-3568,
// 3924-3956, 4144-4182).  Clusters are read in id order; a cluster smaller than
// min_node is skipped; with is_use_mirror == false all four benefits equal
// benefit[0] = inner_edges - |old_entry| * |old_exit| and the cluster is kept iff
// (float)benefit > compress_threshold.  out: accepted cluster id (dense, in raw
// order) per vertex, node type per vertex.
void orc_sum_structure(void* h, const uint32_t* raw, uint32_t min_node, uint32_t max_node, double threshold,
                       uint32_t* cid, uint8_t* type) {
  Graph* g = (Graph*)h;
  const uint32_t n = g->n, INF = 0xFFFFFFFFu;
  uint32_t K = 0;
  for (uint32_t v = 0; v < n; ++v) if (raw[v] != INF) K = std::max(K, raw[v] + 1);
  std::vector<std::vector<uint32_t>> cl(K);
  for (uint32_t v = 0; v < n; ++v) if (raw[v] != INF) cl[raw[v]].push_back(v);
  std::vector<std::vector<uint32_t>> in(n);
  for (uint32_t u = 0; u < n; ++u)
    for (uint64_t e = g->off[u]; e < g->off[u + 1]; ++e) in[g->col[e]].push_back(u);
  std::vector<uint8_t> entry(n, 0), exit_(n, 0);
  std::vector<uint32_t> newid(K, INF);
  uint32_t acc = 0;
  for (uint32_t j = 0; j < K; ++j) {
    const auto& P = cl[j];
    if (P.size() < min_node || P.size() > max_node) continue;
    long long inner = 0;
    std::vector<uint32_t> S, B;
    for (uint32_t u : P) {
      bool is_in = false, is_out = false;
      for (uint32_t w : in[u]) { if (raw[w] != j) is_in = true; else ++inner; }
      for (uint64_t e = g->off[u]; e < g->off[u + 1]; ++e) if (raw[g->col[e]] != j) is_out = true;
      if (is_in) S.push_back(u);
      if (is_out) B.push_back(u);
    }
    float benefit = (float)((double)inner - (double)S.size() * (double)B.size());
    if (benefit > (float)threshold) {
      newid[j] = acc++;
      for (uint32_t u : S) entry[u] = 1;
      for (uint32_t u : B) exit_[u] = 1;
    }
  }
  for (uint32_t v = 0; v < n; ++v) {
    uint32_t c = raw[v] == INF ? INF : newid[raw[v]];
    cid[v] = c;
    if (c == INF) type[v] = 0;                              // SingleNode
    else if (entry[v] && exit_[v]) type[v] = 3;             // BothOutInNode
    else if (entry[v]) type[v] = 2;                         // OnlyInNode
    else if (exit_[v]) type[v] = 1;                         // OnlyOutNode
    else type[v] = 4;                                       // InnerNode
  }
}

// build_iter_index_mirror / inc_run_to_convergence_mirror (iter_compressor.h:
// 318-431): delta = 1 at the entry, r = 0; sweeps over the cluster's members in
// id order (in place, Gauss-Seidel like the reference's serial loop): to_send =
// delta_v, delta_v = 0, push to_send*d/outdeg(v) along v's edges inside the
// cluster, r_v += to_send; stop when the sweep's sum |dr| <= thr or after maxit
// sweeps.  Outputs r (reach: bound_delta for exits, inner_value otherwise) and the
// leftover delta (inner_delta) for every vertex (0 outside the cluster).
int orc_sum_shortcut(void* h, const uint32_t* cid, uint32_t entry, double d, double thr, int maxit, double* r,
                     double* delta) {
  Graph* g = (Graph*)h;
  const uint32_t n = g->n;
  const uint32_t k = cid[entry];
  std::vector<uint32_t> mem;
  for (uint32_t v = 0; v < n; ++v) { r[v] = 0; delta[v] = 0; if (cid[v] == k) mem.push_back(v); }
  delta[entry] = 1.0;
  int step = 0;
  while (true) {
    ++step;
    double diff = 0;
    for (uint32_t v : mem) {
      double to_send = delta[v];
      delta[v] = 0;
      if (to_send != 0) {
        double outv = to_send * d / (double)g->deg(v);
        for (uint64_t e = g->off[v]; e < g->off[v + 1]; ++e)
          if (cid[g->col[e]] == k) delta[g->col[e]] += outv;
        r[v] += to_send;
      }
      diff += std::fabs(to_send);
    }
    if (diff <= thr || step > maxit) break;
  }
  return step;
}

// The SumInc PageRank query restated (sum_sync_iter_worker.h:348-438 first_step,
// 949-1127 the compressed rounds by node type, 1385-1491 correction, then the
// plain Ingress rounds), FP64, with the reference's serial-in-place order inside
// a round.  Shortcuts use orc_sum_shortcut with thr = T/N*|C|, 100 sweeps.
// Deviations shared with the device path: the sketch phase stops on |delta| of
// Single/entry vertices + |bound| of exits (Inner residuals are left to the plain
// phase), pending exit inflow is flushed before the correction, and the
// correction spreads every non-zero entry accumulator.
// r/delta: in = state (init != 0: start from delta = (1-d)/N, r = 0), out = result.
// stats: [0] sketch rounds, [1] plain rounds, [2] edges, [3] final residual.
void orc_sum_pr(void* h, const uint32_t* cid, const uint8_t* type, double d, double T, int max_rounds, int init,
                const uint8_t* dirty_vertex, double* r, double* delta, double* stats) {
  Graph* g = (Graph*)h;
  const uint32_t n = g->n, INF = 0xFFFFFFFFu;
  if (init) for (uint32_t v = 0; v < n; ++v) { r[v] = 0; delta[v] = (1.0 - d) / n; }
  uint32_t K = 0;
  for (uint32_t v = 0; v < n; ++v) if (cid[v] != INF) K = std::max(K, cid[v] + 1);
  std::vector<std::vector<uint32_t>> mem(K);
  for (uint32_t v = 0; v < n; ++v) if (cid[v] != INF) mem[cid[v]].push_back(v);
  const double eps = T / n;
  double edges = 0;
  // shortcuts of every entry
  std::vector<std::vector<std::pair<uint32_t, double>>> reach(n), resid(n);
  {
    std::vector<double> rr(n), dd(n);
    for (uint32_t v = 0; v < n; ++v) {
      if (type[v] != 2 && type[v] != 3) continue;
      orc_sum_shortcut(h, cid, v, d, eps * mem[cid[v]].size(), 100, rr.data(), dd.data());
      for (uint32_t x : mem[cid[v]]) {
        if (rr[x] != 0) reach[v].push_back({x, rr[x]});
        if (dd[x] != 0) resid[v].push_back({x, dd[x]});
      }
    }
  }
  auto is_exit = [&](uint32_t v) { return type[v] == 1 || type[v] == 3; };
  auto push_external = [&](uint32_t v, double amount) {  // d*amount/outdeg over edges leaving v's cluster
    uint32_t deg = g->deg(v);
    if (!deg) return;
    double outv = amount * d / deg;
    for (uint64_t e = g->off[v]; e < g->off[v + 1]; ++e)
      if (cid[g->col[e]] != cid[v]) { delta[g->col[e]] += outv; edges += 1; }
  };
  // precompute (precompute_spnode_one + first_step exit sends + precompute_spnode_two)
  std::vector<double> rloc(n, 0.0);
  for (uint32_t k = 0; k < K; ++k) {
    bool dirty = !dirty_vertex;
    if (dirty_vertex) for (uint32_t v : mem[k]) if (dirty_vertex[v]) dirty = true;
    if (!dirty) continue;
    double thr = eps * mem[k].size();
    int step = 0;
    while (true) {
      ++step;
      double diff = 0;
      for (uint32_t v : mem[k]) {
        double to_send = delta[v];
        if (to_send == 0) continue;
        delta[v] = 0;
        double outv = to_send * d / (double)g->deg(v);
        for (uint64_t e = g->off[v]; e < g->off[v + 1]; ++e)
          if (cid[g->col[e]] == k) { delta[g->col[e]] += outv; edges += 1; }
        rloc[v] += to_send;
        diff += std::fabs(to_send);
      }
      if (diff <= thr || step > 100) break;
    }
    for (uint32_t v : mem[k]) {
      if (is_exit(v) && rloc[v] != 0) push_external(v, rloc[v]);
      r[v] += rloc[v];
      rloc[v] = 0;
    }
  }
  // sketch rounds
  std::vector<double> bound(n, 0.0), spnode(n, 0.0);
  int sk_rounds = 0;
  while (true) {
    double res = 0;
    for (uint32_t v = 0; v < n; ++v) {
      if (type[v] != 4 && type[v] != 1) res += std::fabs(delta[v]);
      if (is_exit(v)) res += std::fabs(bound[v]);
    }
    if (res < T || sk_rounds >= max_rounds) break;
    ++sk_rounds;
    for (uint32_t v = 0; v < n; ++v) {
      uint8_t t = type[v];
      if (t == 0 && std::fabs(delta[v]) > eps) {
        double dv = delta[v];
        delta[v] = 0;
        uint32_t deg = g->deg(v);
        if (deg) for (uint64_t e = g->off[v]; e < g->off[v + 1]; ++e) delta[g->col[e]] += dv * d / deg;
        edges += deg;
        r[v] += dv;
      }
      if ((t == 2 || t == 3) && std::fabs(delta[v]) > eps) {
        double dv = delta[v];
        delta[v] = 0;
        spnode[v] += dv;
        for (auto& xw : reach[v]) if (is_exit(xw.first)) { bound[xw.first] += xw.second * dv; edges += 1; }
      }
      if (is_exit(v) && std::fabs(bound[v]) > eps) {
        double b = bound[v];
        bound[v] = 0;
        r[v] += b;
        push_external(v, b);
      }
    }
  }
  // flush + correction
  for (uint32_t v = 0; v < n; ++v)
    if (is_exit(v) && bound[v] != 0) { double b = bound[v]; bound[v] = 0; r[v] += b; push_external(v, b); }
  for (uint32_t v = 0; v < n; ++v) {
    if (spnode[v] == 0) continue;
    double s = spnode[v];
    for (auto& xw : reach[v]) if (!is_exit(xw.first)) r[xw.first] += s * xw.second;
    for (auto& xw : resid[v]) delta[xw.first] += s * xw.second;
    spnode[v] = 0;
  }
  // plain residual phase
  double st[3];
  orc_pr_ingress(h, d, T, max_rounds, 0, r, delta, st);
  if (stats) { stats[0] = sk_rounds; stats[1] = st[0]; stats[2] = edges + st[1]; stats[3] = st[2]; }
}

// In-cluster shortest distances from `s` (the traversal index of
// TravCompressor::build_trav_index_mirror + run_to_convergence_mirror,
// trav_compressor.h:244-300: value/delta worklist over the edges that stay inside
// the cluster).  out: dist[x] for every vertex (INF outside the cluster or
// unreachable), dist[s] = 0.
void orc_sum_sssp_shortcut(void* h, const uint32_t* cid, uint32_t s, uint32_t* dist) {
  Graph* g = (Graph*)h;
  const uint32_t n = g->n, INF = 0xFFFFFFFFu;
  const uint32_t k = cid[s];
  for (uint32_t v = 0; v < n; ++v) dist[v] = INF;
  dist[s] = 0;
  std::vector<uint32_t> work{s};
  std::vector<uint8_t> in_work(n, 0);
  in_work[s] = 1;
  while (!work.empty()) {
    std::vector<uint32_t> next;
    for (uint32_t u : work) {
      in_work[u] = 0;
      for (uint64_t e = g->off[u]; e < g->off[u + 1]; ++e) {
        const uint32_t v = g->col[e];
        if (cid[v] != k) continue;
        const uint64_t nd = (uint64_t)dist[u] + (g->wu.empty() ? 1u : g->wu[e]);
        if (nd < dist[v]) {
          dist[v] = (uint32_t)nd;
          if (!in_work[v]) { in_work[v] = 1; next.push_back(v); }
        }
      }
    }
    work.swap(next);
  }
}

// The SumInc SSSP query restated from scratch, mirror-free
// (sum_sync_traversal_worker.h:1160-1233 compressed rounds by node type,
// 1322-1406 correction).  Shortcut rows from orc_sum_sssp_shortcut.  If the
// source is clustered its cluster is first relaxed in-cluster from it (the
// reference seeds the source's supernode through GetInitDelta, trav_compressor.h:
// 244-262).  Rounds are frontier-synchronous: a Single relaxes all out-edges, an
// exit its edges that leave the cluster, an entry pushes dist + shortcut to every
// exit of its cluster.  Correction: every member takes min over entries e of
// dist(e) + shortcut(e, x).  Parents: smallest-id predecessor on a shortest path
// (as orc_sssp_dijkstra).  stats: [0] rounds, [1] sketch edges relaxed.
void orc_sum_sssp(void* h, const uint32_t* cid, const uint8_t* type, uint32_t src, uint32_t* dist,
                  uint32_t* parent, double* stats) {
  Graph* g = (Graph*)h;
  const uint32_t n = g->n, INF = 0xFFFFFFFFu;
  auto w = [&](uint64_t e) -> uint64_t { return g->wu.empty() ? 1u : g->wu[e]; };
  uint32_t K = 0;
  for (uint32_t v = 0; v < n; ++v) if (cid[v] != INF) K = std::max(K, cid[v] + 1);
  std::vector<std::vector<uint32_t>> mem(K);
  for (uint32_t v = 0; v < n; ++v) if (cid[v] != INF) mem[cid[v]].push_back(v);
  auto is_exit = [&](uint32_t v) { return type[v] == 1 || type[v] == 3; };
  auto is_entry = [&](uint32_t v) { return type[v] == 2 || type[v] == 3; };
  // shortcut rows of every entry (distance to every member of its cluster)
  std::vector<std::vector<uint32_t>> sc(n);
  std::vector<uint32_t> tmp(n);
  for (uint32_t v = 0; v < n; ++v) {
    if (!is_entry(v)) continue;
    orc_sum_sssp_shortcut(h, cid, v, tmp.data());
    for (uint32_t x : mem[cid[v]]) sc[v].push_back(tmp[x]);
  }
  std::vector<uint64_t> D(n, INF);
  D[src] = 0;
  std::vector<uint8_t> in_next(n, 0);
  std::vector<uint32_t> front{src};
  if (cid[src] != INF) {
    orc_sum_sssp_shortcut(h, cid, src, tmp.data());
    for (uint32_t x : mem[cid[src]])
      if (tmp[x] != INF && x != src) { D[x] = tmp[x]; if (is_exit(x)) front.push_back(x); }
  }
  double rounds = 0, edges = 0;
  std::vector<uint32_t> next;
  auto relax = [&](uint32_t v, uint64_t nd) {
    if (nd < D[v]) {
      D[v] = nd;
      if (!in_next[v]) { in_next[v] = 1; next.push_back(v); }
    }
  };
  while (!front.empty()) {
    ++rounds;
    next.clear();
    for (uint32_t u : front) {
      const uint64_t du = D[u];
      const uint8_t t = type[u];
      if (t == 0 || is_exit(u))
        for (uint64_t e = g->off[u]; e < g->off[u + 1]; ++e) {
          const uint32_t v = g->col[e];
          if (t != 0 && cid[v] == cid[u]) continue;  // exits: external edges only
          relax(v, du + w(e));
          edges += 1;
        }
      if (is_entry(u)) {
        const auto& m = mem[cid[u]];
        for (size_t j = 0; j < m.size(); ++j)
          if (is_exit(m[j]) && sc[u][j] != INF) { relax(m[j], du + sc[u][j]); edges += 1; }
      }
    }
    for (uint32_t v : next) in_next[v] = 0;
    front = next;
  }
  for (uint32_t k = 0; k < K; ++k)
    for (uint32_t e : mem[k]) {
      if (!is_entry(e) || D[e] == INF) continue;
      for (size_t j = 0; j < mem[k].size(); ++j)
        if (sc[e][j] != INF) D[mem[k][j]] = std::min(D[mem[k][j]], D[e] + sc[e][j]);
    }
  for (uint32_t v = 0; v < n; ++v) {
    dist[v] = D[v] >= INF ? INF : (uint32_t)D[v];
    parent[v] = v;
  }
  for (uint32_t v = 0; v < n; ++v) if (v != src && dist[v] != INF) parent[v] = INF;
  for (uint32_t u = 0; u < n; ++u) {
    if (dist[u] == INF) continue;
    for (uint64_t e = g->off[u]; e < g->off[u + 1]; ++e) {
      const uint32_t v = g->col[e];
      if (v != src && (uint64_t)dist[u] + w(e) == dist[v] && u < parent[v]) parent[v] = u;
    }
  }
  if (stats) { stats[0] = rounds; stats[1] = edges; }
}

// ---------------------------------------------------------------------------
// Weighted personalized PageRank and its adjoint (no reference counterpart; the
// nearest app is PHPIngress, php/php_ingress.h:15-120).  FP64 Jacobi series over
// the out-CSR: r = (1-d) p + d W^T r with W[u,v] = w_uv / s_u (dangling rows
// zero); lambda = c + d W lambda; gradients of L = c.r:
//   dL/dp = (1-d) lambda,  dL/dw_uv = (d r_u / s_u)(lambda_v - sum_k W[u,k] lambda_k).
// Stops when sum|delta| < tol (forward) / sum|mu| < tol * sum|c| (adjoint).
// Returns the number of rounds.
// ---------------------------------------------------------------------------
int orc_wpr_forward(void* h, const float* w, const double* p, double d, double tol, int max_rounds, double* r) {
  Graph* g = (Graph*)h;
  const uint32_t n = g->n;
  std::vector<double> s(n, 0.0), x(n), y(n);
  for (uint32_t u = 0; u < n; ++u)
    for (uint64_t e = g->off[u]; e < g->off[u + 1]; ++e) s[u] += (double)w[e];
  for (uint32_t v = 0; v < n; ++v) { x[v] = (1.0 - d) * p[v]; r[v] = 0.0; }
  int rounds = 0;
  while (true) {
    std::fill(y.begin(), y.end(), 0.0);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
      const uint32_t u = (uint32_t)i;
      if (x[u] == 0.0 || !(s[u] > 0)) continue;
      const double f = d * x[u] / s[u];
      for (uint64_t e = g->off[u]; e < g->off[u + 1]; ++e) {
        std::atomic_ref<double> a(y[g->col[e]]);
        double old = a.load(std::memory_order_relaxed);
        while (!a.compare_exchange_weak(old, old + f * (double)w[e], std::memory_order_relaxed)) {}
      }
    }
    double res = 0;
    for (uint32_t v = 0; v < n; ++v) { r[v] += x[v]; x[v] = y[v]; res += std::fabs(y[v]); }
    ++rounds;
    if (res < tol || rounds >= max_rounds) break;
  }
  for (uint32_t v = 0; v < n; ++v) r[v] += x[v];
  return rounds;
}

int orc_wpr_adjoint(void* h, const float* w, const double* r, const double* c, double d, double tol, int max_rounds,
                    double* lam, double* grad_p, double* grad_w) {
  Graph* g = (Graph*)h;
  const uint32_t n = g->n;
  std::vector<double> s(n, 0.0), x(n), y(n);
  for (uint32_t u = 0; u < n; ++u)
    for (uint64_t e = g->off[u]; e < g->off[u + 1]; ++e) s[u] += (double)w[e];
  double l1 = 0;
  for (uint32_t v = 0; v < n; ++v) { x[v] = c[v]; lam[v] = 0.0; l1 += std::fabs(c[v]); }
  int rounds = 0;
  while (true) {
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
      const uint32_t u = (uint32_t)i;
      double acc = 0;
      if (s[u] > 0)
        for (uint64_t e = g->off[u]; e < g->off[u + 1]; ++e) acc += (double)w[e] * x[g->col[e]];
      y[u] = s[u] > 0 ? d * acc / s[u] : 0.0;
    }
    double res = 0;
    for (uint32_t v = 0; v < n; ++v) { lam[v] += x[v]; x[v] = y[v]; res += std::fabs(y[v]); }
    ++rounds;
    if (res < tol * l1 || rounds >= max_rounds) break;
  }
  for (uint32_t v = 0; v < n; ++v) lam[v] += x[v];
  for (uint32_t v = 0; v < n; ++v) grad_p[v] = (1.0 - d) * lam[v];
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const uint32_t u = (uint32_t)i;
    double a = 0;
    if (s[u] > 0)
      for (uint64_t e = g->off[u]; e < g->off[u + 1]; ++e) a += (double)w[e] / s[u] * lam[g->col[e]];
    for (uint64_t e = g->off[u]; e < g->off[u + 1]; ++e)
      grad_w[e] = s[u] > 0 ? d * r[u] / s[u] * (lam[g->col[e]] - a) : 0.0;
  }
  return rounds;
}

}  // extern "C"
