// sssp.hip -- weighted min-plus SSSP on the device (the reference's traversal
// Ingress loop, sum_sync_traversal_worker.h:1146-1159 with the weighted
// SSSPIngress::Compute of sssp/sssp_ingress.h:131-175; replaces the unit-weight
// tjnsssp kernels, cudaCode/my_ssspworker.cu:374-495).
//
// State: dp[v] = dist << 32 | parent (u64).  One 64-bit atomicMin both lowers
// the distance and, on equal distance, picks the smaller parent id, so the fixed
// point is unique: dist = shortest distance, parent = smallest-id predecessor on
// a shortest path (the reference's DependencyData CAS keeps whichever proposal
// lands first, dependency_data.h:21-34; distances are identical, parents are a
// deterministic choice among the reference's valid ones).
//
// Frontier: double-buffered vertex lists + bitmaps.  A vertex whose distance
// drops is appended to the next list once (atomicOr on its bitmap word, then a
// wave64-aggregated append).  Per round: k_sssp_push (light tiles, heavy
// vertices deferred to 2048-edge chunks), k_sssp_heavy, k_sssp_end.  The lists,
// bitmaps and the count of pending seeds belong to the SSSP state: batches append
// seeds, and each query takes them at its start (k_sssp_begin).
#include <algorithm>

#include "ctx.hpp"
#include "device.hpp"
#include "pr.hpp"

namespace suminc {

struct SsspArgs {
  const uint32_t* off;
  const uint32_t* col;
  const uint32_t* w;  // null: unit weights
  unsigned long long* dp;
  uint32_t* wl[2];
  uint32_t* bits[2];
  uint2* heavy;
  Counters* cnt;
  int app;  // SUMINC_APP_*
};

__device__ __forceinline__ void sssp_relax(const SsspArgs& a, int nxt, uint32_t u, uint32_t du, uint32_t e) {
  uint32_t v = a.col[e];
  uint32_t w = a.w ? a.w[e] : 1u;
  unsigned long long nd = trav_extend(a.app, du, w);  // < 2^32 - 1 unless the path saturates
  bool add = false;
  if (nd < INF32) {
    unsigned long long cand = (nd << 32) | u;
    // words only decrease while the rounds run: a candidate no smaller than a
    // plain (possibly stale) read cannot win, so it skips the atomic.  SSWP takes
    // strictly better words only (no parent change on a tie: no parent cycles)
    if (a.app == SUMINC_APP_SSWP ? nd < (a.dp[v] >> 32) : cand < a.dp[v]) {
      unsigned long long old = atomicMin(&a.dp[v], cand);
      add = (old >> 32) > nd;
    }
  }
  if (add) {
    uint32_t bit = 1u << (v & 31);
    uint32_t prev = atomicOr(&a.bits[nxt][v >> 5], bit);
    add = !(prev & bit);
  }
  // every lane reaches the ballot, active or not
  unsigned long long slot = wave_append(add, &a.cnt->wl_count[nxt]);
  if (add) a.wl[nxt][slot] = v;
}

__global__ void __launch_bounds__(kBlock) k_sssp_push(SsspArgs a) {
  __shared__ uint32_t s_scan[kBlock + 1];
  __shared__ uint32_t s_start[kBlock];
  __shared__ uint32_t s_warp[kBlock / 64 + 1];
  __shared__ uint32_t s_u[kBlock];
  __shared__ uint32_t s_du[kBlock];
  __shared__ unsigned long long s_edges;
  if (a.cnt->done) return;
  const int cur = (int)(a.cnt->aux[1] & 1), nxt = cur ^ 1;
  const unsigned long long nl = a.cnt->wl_count[cur];
  const uint32_t vt = frontier_tile<kBlock>(nl);
  const unsigned long long tiles = (nl + vt - 1) / vt;
  if (threadIdx.x == 0) s_edges = 0;
  __syncthreads();
  for (unsigned long long t = blockIdx.x; t < tiles; t += gridDim.x) {
    unsigned long long i = t * vt + threadIdx.x;
    uint32_t start = 0, len = 0;
    if (threadIdx.x < vt && i < nl) {
      uint32_t u = a.wl[cur][i];
      atomicAnd(&a.bits[cur][u >> 5], ~(1u << (u & 31)));
      uint32_t du = (uint32_t)(a.dp[u] >> 32);
      uint32_t b = a.off[u], e = a.off[u + 1];
      if (du != INF32 && e > b) {
        if (e - b <= kHeavyDeg) {
          start = b;
          len = e - b;
        } else {
          uint32_t nch = (e - b + kChunk - 1) / kChunk;
          unsigned long long slot = atomicAdd(&a.cnt->heavy_count, (unsigned long long)nch);
          for (uint32_t k = 0; k < nch; ++k) a.heavy[slot + k] = make_uint2(u, k);
        }
        atomicAdd(&s_edges, (unsigned long long)(e - b));
      }
      s_u[threadIdx.x] = u;
      s_du[threadIdx.x] = du;
    }
    block_adjacency<kBlock>(start, len, s_scan, s_start, s_warp,
                            [&](int slot, uint32_t, uint32_t e) { sssp_relax(a, nxt, s_u[slot], s_du[slot], e); });
  }
  __syncthreads();
  if (threadIdx.x == 0 && s_edges) atomicAdd(&a.cnt->edges, s_edges);
}

__global__ void __launch_bounds__(kBlock) k_sssp_heavy(SsspArgs a) {
  if (a.cnt->done) return;
  const int cur = (int)(a.cnt->aux[1] & 1), nxt = cur ^ 1;
  const unsigned long long nh = a.cnt->heavy_count;
  for (unsigned long long h = blockIdx.x; h < nh; h += gridDim.x) {
    uint2 hv = a.heavy[h];
    uint32_t u = hv.x;
    uint32_t du = (uint32_t)(a.dp[u] >> 32);
    uint32_t b = a.off[u] + hv.y * kChunk;
    uint32_t e = min(a.off[u + 1], b + kChunk);
    // all lanes iterate the same trip count so wave_append sees full waves
    for (uint32_t k0 = b; k0 < e; k0 += kBlock) {
      uint32_t k = k0 + threadIdx.x;
      if (k < e) sssp_relax(a, nxt, u, du, k);
    }
  }
}

__global__ void k_sssp_end(Counters* cnt, int max_rounds) {
  if (threadIdx.x != 0 || cnt->done) return;
  const int cur = (int)(cnt->aux[1] & 1), nxt = cur ^ 1;
  cnt->aux[6] += cnt->wl_count[cur];  // frontier vertices processed
  cnt->wl_count[cur] = 0;
  cnt->heavy_count = 0;
  cnt->rounds += 1;
  cnt->aux[1] ^= 1ull;
  if (cnt->wl_count[nxt] == 0 || (long long)cnt->rounds >= (long long)max_rounds) cnt->done = 1;
}

__global__ void k_sssp_init(uint32_t n, unsigned long long* dp, uint32_t src) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  dp[v] = (v == src) ? (unsigned long long)src : ((unsigned long long)INF32 << 32) | v;
}

__global__ void k_sssp_seed(unsigned long long* seeds, uint32_t* wl0, uint32_t* bits0, uint32_t src) {
  if (threadIdx.x) return;
  *seeds = 1;
  wl0[0] = src;
  bits0[src >> 5] = 1u << (src & 31);
}

// Start of a query: list 0 holds the pending seeds; the round state starts clear.
__global__ void k_sssp_begin(Counters* cnt, unsigned long long* seeds) {
  if (threadIdx.x) return;
  cnt->wl_count[0] = *seeds;
  *seeds = 0;
  cnt->wl_count[1] = 0;
  cnt->heavy_count = 0;
  cnt->aux[1] = 0;
  cnt->aux[6] = 0;
  cnt->done = 0;
  cnt->rounds = 0;
  cnt->edges = 0;
}

__global__ void k_sssp_unpack(uint32_t n, const unsigned long long* dp, int app, uint32_t* dist, uint32_t* parent) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  unsigned long long x = dp[v];
  dist[v] = app == SUMINC_APP_SSWP ? ~(uint32_t)(x >> 32) : (uint32_t)(x >> 32);
  parent[v] = (uint32_t)x;
}

SsspArgs sssp_args(suminc_ctx* c) {
  SsspArgs a;
  a.off = c->out_off.as<uint32_t>();
  a.col = c->out_col.as<uint32_t>();
  a.w = c->wkind == SUMINC_W_U32 && c->sssp_app != SUMINC_APP_BFS ? c->out_w.as<uint32_t>() : nullptr;
  a.dp = c->sssp_dp.as<unsigned long long>();
  a.app = c->sssp_app;
  a.wl[0] = c->sssp_wl0.as<uint32_t>();
  a.wl[1] = c->sssp_wl1.as<uint32_t>();
  a.bits[0] = c->sssp_flag0.as<uint32_t>();
  a.bits[1] = c->sssp_flag1.as<uint32_t>();
  a.heavy = c->heavy.as<uint2>();
  a.cnt = dev_counters(c);
  return a;
}

void sssp_alloc(suminc_ctx* c) {
  const size_t n = c->n;
  c->sssp_dp.ensure(n * 8);
  const size_t words = (n + 31) / 32 + 1;
  c->sssp_flag0.ensure(words * 4);
  c->sssp_flag1.ensure(words * 4);
  c->sssp_wl0.ensure(n * 4 + 4);
  c->sssp_wl1.ensure(n * 4 + 4);
  c->sssp_seeds.ensure(8);
  c->heavy.ensure_slack((c->m / kChunk + c->m / kHeavyDeg + 2) * sizeof(uint4));
}

void sssp_clear_flags(suminc_ctx* c) {
  const size_t words = ((size_t)c->n + 31) / 32 + 1;
  SUMINC_HIP(hipMemsetAsync(c->sssp_flag0.p, 0, words * 4, c->stream));
  SUMINC_HIP(hipMemsetAsync(c->sssp_flag1.p, 0, words * 4, c->stream));
}

void traversal_init(suminc_ctx* c, int app, uint32_t source) {
  SUMINC_REQUIRE(app == SUMINC_APP_SSSP || app == SUMINC_APP_BFS || app == SUMINC_APP_SSWP, SUMINC_E_INVALID,
                 "unknown traversal app");
  c->sssp_app = app;
  sssp_init(c, source);
}

void sssp_init(suminc_ctx* c, uint32_t source) {
  SUMINC_REQUIRE(c->has_graph, SUMINC_E_STATE, "no graph");
  SUMINC_REQUIRE(source < c->n, SUMINC_E_INVALID, "source out of range");
  sssp_alloc(c);
  c->sssp_src = source;
  k_sssp_init<<<grid_for(c->n), kBlock, 0, c->stream>>>(c->n, c->sssp_dp.as<unsigned long long>(), source);
  sssp_clear_flags(c);
  k_sssp_seed<<<1, 64, 0, c->stream>>>(c->sssp_seeds.as<unsigned long long>(), c->sssp_wl0.as<uint32_t>(),
                                       c->sssp_flag0.as<uint32_t>(), source);
  SUMINC_HIP(hipGetLastError());
  c->sssp_ready = true;
  c->sssp_pending = true;
}

// Runs rounds from the frontier currently held in list 0 (set by sssp_init or
// by the batch repair) until it empties.
void sssp_run(suminc_ctx* c, suminc_stats* st) {
  SUMINC_REQUIRE(c->sssp_ready, SUMINC_E_STATE, "SSSP state not initialised");
  const int max_rounds = c->prm.max_rounds > 0 ? c->prm.max_rounds : 1000000;
  SsspArgs a = sssp_args(c);
  c->sssp_pending = false;
  SUMINC_HIP(hipEventRecord(c->ev_start, c->stream));
  k_sssp_begin<<<1, 64, 0, c->stream>>>(a.cnt, c->sssp_seeds.as<unsigned long long>());
  int group = c->hint_sssp;  // rounds of the previous query + 1 first
  size_t slot = 0;
  while (true) {
    for (int i = 0; i < group; ++i, ++slot) {
      if (c->prof_on) SUMINC_HIP(hipEventRecord(prof_event(c, 2 * slot), c->stream));
      k_sssp_push<<<push_grid(), kBlock, 0, c->stream>>>(a);
      k_sssp_heavy<<<push_grid(), kBlock, 0, c->stream>>>(a);
      if (c->prof_on) SUMINC_HIP(hipEventRecord(prof_event(c, 2 * slot + 1), c->stream));
      k_sssp_end<<<1, 64, 0, c->stream>>>(a.cnt, max_rounds);
    }
    SUMINC_HIP(hipGetLastError());
    read_counters(c);
    if (c->host_counters->done) break;
    if (group < 64) group *= 2;
  }
  SUMINC_HIP(hipEventRecord(c->ev_stop, c->stream));
  SUMINC_HIP(hipEventSynchronize(c->ev_stop));
  float ms = 0;
  SUMINC_HIP(hipEventElapsedTime(&ms, c->ev_start, c->ev_stop));
  Counters* h = c->host_counters;
  c->hint_sssp = std::max(2, std::min(64, (int)h->rounds + 1));
  if (c->prof_on) {
    prof_collect(c, SUMINC_KERNEL_SSSP_PUSH, (size_t)h->rounds, 0, 2);
    suminc_kernel_profile& p = c->prof[SUMINC_KERNEL_SSSP_PUSH];
    // algorithmic bytes of the min-plus push: column + weight (8 B) and a 32-bit
    // distance read-modify-write (8 B) per edge, 16 B per frontier vertex (list
    // entry, row pointers, distance); the packed 64-bit {dist, parent} word moves
    // 8 B more per edge than this count
    p.edges += (double)h->edges;
    p.vertices += (double)h->aux[6];
    p.alg_bytes += 16.0 * (double)h->edges + 16.0 * (double)h->aux[6];
  }
  c->stats = suminc_stats{};
  c->stats.rounds = (double)h->rounds;
  c->stats.edges_traversed = (double)h->edges;
  c->stats.kernel_ms = ms;
  c->stats.push_rounds = (double)h->rounds;
  if (st) *st = c->stats;
  // leave list 0 as the (empty) current list for the next repair
  const int cur = (int)(h->aux[1] & 1);
  SUMINC_REQUIRE(h->wl_count[cur] == 0 || h->rounds < (unsigned long long)max_rounds, SUMINC_E_NOT_CONVERGED,
                 "round cap reached with a non-empty frontier");
  if (cur != 0) {
    // swap buffers so that list/bitmap 0 is current
    c->sssp_wl0.swap(c->sssp_wl1);
    c->sssp_flag0.swap(c->sssp_flag1);
  }
}

// the next frontier (list 1 after a round) becomes the pending seeds of list 0
__global__ void k_sssp_carry(const Counters* cnt, unsigned long long* seeds) {
  if (threadIdx.x == 0) *seeds = cnt->wl_count[1];
}

// One min-plus round from the pending frontier (the seeds of init or a batch, or
// the previous step's output): tjnsssp::g_function + setCurSize/getAllSize
// (sum_sync_traversal_worker.h:1090, 1136-1142).  The frontier the round produced
// becomes the pending one; *frontier_out gets its size (0: fixed point).
void sssp_step(suminc_ctx* c, uint32_t* frontier_out) {
  SUMINC_REQUIRE(c->sssp_ready, SUMINC_E_STATE, "SSSP state not initialised");
  SsspArgs a = sssp_args(c);
  c->sssp_pending = false;
  k_sssp_begin<<<1, 64, 0, c->stream>>>(a.cnt, c->sssp_seeds.as<unsigned long long>());
  k_sssp_push<<<push_grid(), kBlock, 0, c->stream>>>(a);
  k_sssp_heavy<<<push_grid(), kBlock, 0, c->stream>>>(a);
  k_sssp_end<<<1, 64, 0, c->stream>>>(a.cnt, 1 << 30);
  k_sssp_carry<<<1, 64, 0, c->stream>>>(a.cnt, c->sssp_seeds.as<unsigned long long>());
  SUMINC_HIP(hipGetLastError());
  read_counters(c);
  const Counters* h = c->host_counters;
  c->stats = suminc_stats{};
  c->stats.rounds = 1;
  c->stats.edges_traversed = (double)h->edges;
  c->stats.push_rounds = 1;
  // list 1 holds the next frontier with its bitmap: make it list 0
  c->sssp_wl0.swap(c->sssp_wl1);
  c->sssp_flag0.swap(c->sssp_flag1);
  if (frontier_out) *frontier_out = (uint32_t)h->wl_count[1];
}

void sssp_download(suminc_ctx* c, uint32_t* dist, uint32_t* parent) {
  SUMINC_REQUIRE(c->sssp_ready, SUMINC_E_STATE, "SSSP state not initialised");
  DevBuf dd, pp;
  dd.ensure((size_t)c->n * 4);
  pp.ensure((size_t)c->n * 4);
  k_sssp_unpack<<<grid_for(c->n), kBlock, 0, c->stream>>>(c->n, c->sssp_dp.as<unsigned long long>(), c->sssp_app,
                                                        dd.as<uint32_t>(), pp.as<uint32_t>());
  SUMINC_HIP(hipGetLastError());
  if (dist) SUMINC_HIP(hipMemcpyAsync(dist, dd.p, (size_t)c->n * 4, hipMemcpyDeviceToHost, c->stream));
  if (parent) SUMINC_HIP(hipMemcpyAsync(parent, pp.p, (size_t)c->n * 4, hipMemcpyDeviceToHost, c->stream));
  SUMINC_HIP(hipStreamSynchronize(c->stream));
}

}  // namespace suminc
