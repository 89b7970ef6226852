// sketch_sssp.hip -- SSSP on the supernode sketch (the reference's
// SumSyncTraversalWorker compressed phase, mirror-free): shortcuts are in-cluster
// shortest paths from every entry (TravCompressor::build_trav_index_mirror +
// run_to_convergence_mirror + fianl_build_trav_index_mirror,
// trav_compressor.h:244-431); sketch rounds relax a Single's out-edges, an
// exit's external edges and an entry's shortcut row to its exits
// (sum_sync_traversal_worker.h:1160-1233, SSSPIngress::Compute /
// ComputeByIndexDelta, sssp/sssp_ingress.h:131-261); the correction gives every
// member min over entries of dist(entry) + in-cluster distance
// (sum_sync_traversal_worker.h:1322-1403).
//
// State is the plain path's packed {dist:hi32, parent:lo32} array, so parents are
// real graph edges and the batch repair (subtree reset + re-pull) applies
// unchanged.  Exactness of parents: after the distances are final, one pass over
// every member's internal in-edges min-reduces (dist + w, src) into the member,
// so every vertex ends with the smallest-id predecessor on a shortest path.
#include <algorithm>
#include <cstring>

#include "ctx.hpp"
#include "device.hpp"
#include "pr.hpp"
#include "sketch.hpp"

namespace suminc {

constexpr int kMaxClusterS = 2048;
constexpr uint64_t kInf64 = ~0ull;

// Relaxes the internal edge i -> j (local indices) of weight w on packed
// {dist, parent} words in LDS; true when it lowered j.
__device__ __forceinline__ bool relax_lds(uint32_t i, uint32_t j, uint32_t w, const uint32_t* s_id,
                                          unsigned long long* s_dp) {
  const unsigned long long di = s_dp[i] >> 32;
  if (di == 0xFFFFFFFFull) return false;
  const unsigned long long nd = di + w;
  if (nd >= 0xFFFFFFFFull) return false;
  const unsigned long long cand = (nd << 32) | s_id[i];
  if (cand < s_dp[j]) {
    atomicMin(&s_dp[j], cand);
    return true;
  }
  return false;
}

// The first kEdgeCache * kBlock internal edges of a cluster, kept in registers
// across the Bellman-Ford rounds of a workgroup (from -> to: the relaxation
// direction; the reversed edges give distances towards a member).
constexpr int kEdgeCache = 8;
struct EdgeCache {
  uint32_t a[kEdgeCache], b[kEdgeCache], w[kEdgeCache];
  __device__ __forceinline__ void load(uint32_t pb, uint32_t pe, const uint32_t* __restrict__ from,
                                       const uint32_t* __restrict__ to, const uint32_t* __restrict__ wt) {
#pragma unroll
    for (int r = 0; r < kEdgeCache; ++r) {
      const uint32_t p = pb + r * kBlock + threadIdx.x;
      a[r] = p < pe ? from[p] : INF32;
      b[r] = p < pe ? to[p] : 0u;
      w[r] = p < pe ? wt[p] : 0u;
    }
  }
};

// In-cluster Bellman-Ford rounds on packed {dist, parent} words in LDS,
// edge-parallel over the cluster's internal edges with LDS atomic min, until a
// round changes nothing.  s_id holds the members' vertex ids; edges past the
// cache are read from memory.  One barrier per round: the round's change flag
// rotates over three LDS words (the word a round sets was last read two rounds
// earlier).  Returns rounds run.
__device__ uint32_t cluster_bf(uint32_t size, uint32_t pb, uint32_t pe, const EdgeCache& ec,
                               const uint32_t* __restrict__ from, const uint32_t* __restrict__ to,
                               const uint32_t* __restrict__ wt, const uint32_t* s_id, unsigned long long* s_dp,
                               int* s_flag) {
  if (threadIdx.x == 0) s_flag[0] = 0;
  __syncthreads();
  const uint32_t pc = pb + kEdgeCache * kBlock;  // first edge past the cache
  uint32_t rounds = 0;
  while (true) {
    const int cur = (int)(rounds % 3);
    if (threadIdx.x == 0) s_flag[(rounds + 1) % 3] = 0;
    bool changed = false;
#pragma unroll
    for (int r = 0; r < kEdgeCache; ++r)
      if (ec.a[r] != INF32) changed |= relax_lds(ec.a[r], ec.b[r], ec.w[r], s_id, s_dp);
    for (uint32_t p = pc + threadIdx.x; p < pe; p += kBlock) changed |= relax_lds(from[p], to[p], wt[p], s_id, s_dp);
    if (changed) s_flag[cur] = 1;
    __syncthreads();
    ++rounds;
    if (!s_flag[cur] || rounds > size + 1) break;
  }
  return rounds;
}

constexpr unsigned kClusterGrid = 2048;
constexpr uint32_t kSmallCluster = 64;  // clusters up to this size run one per wave

// Workers (workgroups or waves) per cluster when W workers share nk clusters:
// one when the clusters fill the workers, else W / nk, each taking every P-th
// source (entry or exit) of its cluster.  A cluster's Bellman-Ford runs are
// independent, so its sources spread over P workers instead of queueing on one.
__device__ __forceinline__ uint32_t workers_per_cluster(uint32_t nk, uint32_t W) {
  return nk == 0 || nk >= W ? 1u : W / nk;
}

// Clusters whose flag equals `want`, split by size: clusters of at most
// kSmallCluster members at list[0, counts[0]), larger ones at
// list[K - counts[1], K) (filled from the back).  One counter atomic per
// workgroup and list (every cluster is visited, so per-wave atomics on the two
// counters would queue behind each other).
__global__ void __launch_bounds__(kBlock) k_cluster_lists(uint32_t K, const uint8_t* __restrict__ flag, uint8_t want,
                                                          const uint32_t* __restrict__ cl_off, uint32_t* list,
                                                          unsigned long long* counts) {
  __shared__ uint32_t s_cnt[2][kBlock / 64];
  __shared__ unsigned long long s_base[2];
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  const int wv = threadIdx.x >> 6, lane = lane_id();
  const bool take = k < K && flag[k] == want;
  const bool small = take && cl_off[k + 1] - cl_off[k] <= kSmallCluster;
  const bool large = take && !small;
  const unsigned long long ms = __ballot(small), ml = __ballot(large);
  if (lane == 0) {
    s_cnt[0][wv] = (uint32_t)__popcll(ms);
    s_cnt[1][wv] = (uint32_t)__popcll(ml);
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    uint32_t tot = 0;
    for (int w = 0; w < kBlock / 64; ++w) tot += s_cnt[threadIdx.x][w];
    s_base[threadIdx.x] = tot ? atomicAdd(&counts[threadIdx.x], (unsigned long long)tot) : 0ull;
  }
  __syncthreads();
  unsigned long long bs = s_base[0], bl = s_base[1];
  for (int w = 0; w < wv; ++w) {
    bs += s_cnt[0][w];
    bl += s_cnt[1][w];
  }
  const unsigned long long below = (1ull << lane) - 1ull;
  if (small) list[bs + __popcll(ms & below)] = k;
  else if (large) list[K - 1 - (bl + __popcll(ml & below))] = k;
}

// Orders this wave's LDS traffic before what follows (wave-local exchange).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// cluster_bf for a cluster of at most 64 members run by one wave (s_id, s_dp are
// this wave's LDS slots): up to 128 internal edges stay in registers across the
// rounds; a round ends with a wave vote instead of workgroup barriers.
__device__ uint32_t cluster_bf_wave(uint32_t size, uint32_t pb, uint32_t pe, const uint32_t* __restrict__ src,
                                    const uint32_t* __restrict__ dst, const uint32_t* __restrict__ wt,
                                    const uint32_t* s_id, unsigned long long* s_dp) {
  const int lane = lane_id();
  auto relax = [&](uint32_t i, uint32_t j, uint32_t w) -> bool { return relax_lds(i, j, w, s_id, s_dp); };
  uint32_t rounds = 0;
  if (pe - pb <= 128) {
    uint32_t ei[2] = {0, 0}, ej[2] = {0, 0}, ew[2] = {0, 0};
    bool ev[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const uint32_t p = pb + lane + 64 * r;
      ev[r] = p < pe;
      if (ev[r]) {
        ei[r] = src[p];
        ej[r] = dst[p];
        ew[r] = wt[p];
      }
    }
    while (true) {
      bool changed = false;
#pragma unroll
      for (int r = 0; r < 2; ++r)
        if (ev[r]) changed |= relax(ei[r], ej[r], ew[r]);
      wave_lds_sync();
      ++rounds;
      if (!__any(changed) || rounds > size + 1) break;
    }
  } else {
    while (true) {
      bool changed = false;
      for (uint32_t p = pb + lane; p < pe; p += 64) changed |= relax(src[p], dst[p], wt[p]);
      wave_lds_sync();
      ++rounds;
      if (!__any(changed) || rounds > size + 1) break;
    }
  }
  return rounds;
}

// Shortcut rows: for every entry e of cluster k, the in-cluster distance to every
// exit of k (the first nexit members; 0 for the entry itself if it is an exit).
// The rows hold distances only: an exit reached through a row takes parent INF32
// for the moment, and the correction pass, which re-relaxes the cluster's
// internal edges once the distances are final, gives it its smallest-id
// in-cluster predecessor.  A cluster runs one Bellman-Ford per entry, or one per
// exit over the reversed internal edges when it has fewer exits than entries.
__global__ void __launch_bounds__(kBlock) k_ss_shortcuts(const uint32_t* __restrict__ cl_off,
                                                         const uint32_t* __restrict__ cl_nexit,
                                                         const uint32_t* __restrict__ cl_mem,
                                                         const uint32_t* __restrict__ cl_eoff,
                                                         const uint32_t* __restrict__ cl_ent,
                                                         const uint32_t* __restrict__ lidx,
                                                         const unsigned long long* __restrict__ ss_base,
                                                         Internal ie, uint32_t* __restrict__ sd,
                                                         const uint32_t* __restrict__ list, uint32_t K,
                                                         const unsigned long long* __restrict__ counts,
                                                         uint8_t* valid, unsigned long long* work) {
  __shared__ unsigned long long s_dp[kMaxClusterS];
  __shared__ uint32_t s_id[kMaxClusterS];
  __shared__ int s_flag[3];
  const uint32_t nk = (uint32_t)counts[1];  // the large clusters, at the back of the list
  const uint32_t P = workers_per_cluster(nk, gridDim.x);
  unsigned long long edges = 0;
  for (uint32_t wi = blockIdx.x; wi < nk * P; wi += gridDim.x) {  // nk * P <= max(nk, grid)
    const uint32_t li = wi % nk, part = wi / nk;
    const uint32_t k = list[K - 1 - li];
    const uint32_t slot0 = cl_off[k], size = cl_off[k + 1] - slot0, nexit = cl_nexit[k];
    const uint32_t pb = ie.off[k], pe = ie.off[k + 1];
    const uint32_t eb = cl_eoff[k], ee = cl_eoff[k + 1], nent = ee - eb;
    if (threadIdx.x == 0 && part == 0) valid[k] = 1;
    if (nexit == 0 || nent == 0 || part >= (nent <= nexit ? nent : nexit)) continue;
    for (uint32_t j = threadIdx.x; j < size; j += kBlock) s_id[j] = cl_mem[slot0 + j];
    const bool fwd = nent <= nexit;
    EdgeCache ec;
    ec.load(pb, pe, fwd ? ie.src : ie.dst, fwd ? ie.dst : ie.src, ie.wt);
    if (fwd) {
      for (uint32_t e = eb + part; e < ee; e += P) {
        const uint32_t src = lidx[cl_ent[e]];
        const unsigned long long base = ss_base[e];
        if (nexit == 1 && src == 0) {  // the only exit is the entry itself
          if (threadIdx.x == 0) sd[base] = 0;
          continue;
        }
        for (uint32_t j = threadIdx.x; j < size; j += kBlock) s_dp[j] = j == src ? 0xFFFFFFFFull : kInf64;
        const uint32_t rounds = cluster_bf(size, pb, pe, ec, ie.src, ie.dst, ie.wt, s_id, s_dp, s_flag);
        edges += (unsigned long long)rounds * (pe - pb);
        for (uint32_t j = threadIdx.x; j < nexit; j += kBlock) sd[base + j] = (uint32_t)(s_dp[j] >> 32);
        __syncthreads();
      }
    } else {
      for (uint32_t x = part; x < nexit; x += P) {
        for (uint32_t j = threadIdx.x; j < size; j += kBlock) s_dp[j] = j == x ? 0xFFFFFFFFull : kInf64;
        const uint32_t rounds = cluster_bf(size, pb, pe, ec, ie.dst, ie.src, ie.wt, s_id, s_dp, s_flag);
        edges += (unsigned long long)rounds * (pe - pb);
        for (uint32_t e = eb + threadIdx.x; e < ee; e += kBlock)
          sd[ss_base[e] + x] = (uint32_t)(s_dp[lidx[cl_ent[e]]] >> 32);
        __syncthreads();
      }
    }
    __syncthreads();  // s_id of the next cluster
  }
  if (threadIdx.x == 0 && edges) atomicAdd(work, edges);
}

// k_ss_shortcuts for the small clusters (list[0, counts[0])), one wave each: the
// entries come from the slot flags, the row of the q-th entry starts at the
// cluster's row base + q * nexit.
__global__ void __launch_bounds__(kBlock) k_ss_shortcuts_small(const uint32_t* __restrict__ cl_off,
                                                               const uint32_t* __restrict__ cl_nexit,
                                                               const uint32_t* __restrict__ cl_mem,
                                                               const uint8_t* __restrict__ slot_ent,
                                                               const ClBase* __restrict__ clb, Internal ie,
                                                               uint32_t* __restrict__ sd,
                                                               const uint32_t* __restrict__ list,
                                                               const unsigned long long* __restrict__ counts,
                                                               uint8_t* valid, unsigned long long* work) {
  __shared__ unsigned long long s_dp_w[kBlock / 64][kSmallCluster];
  __shared__ uint32_t s_id_w[kBlock / 64][kSmallCluster];
  __shared__ unsigned long long s_edges;
  const int wv = threadIdx.x >> 6, lane = lane_id();
  const unsigned long long below = (1ull << lane) - 1ull;
  unsigned long long* s_dp = s_dp_w[wv];
  uint32_t* s_id = s_id_w[wv];
  if (threadIdx.x == 0) s_edges = 0;
  __syncthreads();
  const uint32_t nk = (uint32_t)counts[0];
  const uint32_t nwaves = gridDim.x * (kBlock / 64);
  const uint32_t P = workers_per_cluster(nk, nwaves);
  unsigned long long edges = 0;
  for (uint32_t wi = blockIdx.x * (kBlock / 64) + wv; wi < nk * P; wi += nwaves) {  // nk * P <= max(nk, nwaves)
    const uint32_t li = wi % nk, part = wi / nk;
    const uint32_t k = list[li];
    const uint32_t slot0 = cl_off[k], size = cl_off[k + 1] - slot0, nexit = cl_nexit[k];
    const uint32_t pb = ie.off[k], pe = ie.off[k + 1];
    const unsigned long long rbase = clb[k].ss;
    if (lane == 0 && part == 0) valid[k] = 1;
    bool en = false;
    if (lane < (int)size) {
      s_id[lane] = cl_mem[slot0 + lane];
      en = slot_ent[slot0 + lane] != 0;
    }
    const unsigned long long ents = __ballot(en);
    const uint32_t nent = (uint32_t)__popcll(ents);
    if (nexit == 0 || nent == 0 || part >= (nent <= nexit ? nent : nexit)) continue;
    if (nent <= nexit) {
      unsigned long long left = ents;
      for (uint32_t q = 0; left; ++q) {
        const uint32_t src = (uint32_t)__ffsll((long long)left) - 1;
        left &= left - 1;
        if (q % P != part) continue;
        const unsigned long long r0 = rbase + (unsigned long long)q * nexit;
        if (nexit == 1 && src == 0) {  // the only exit is the entry itself
          if (lane == 0) sd[r0] = 0;
          continue;
        }
        if (lane < (int)size) s_dp[lane] = (uint32_t)lane == src ? 0xFFFFFFFFull : kInf64;
        wave_lds_sync();
        const uint32_t rounds = cluster_bf_wave(size, pb, pe, ie.src, ie.dst, ie.wt, s_id, s_dp);
        edges += (unsigned long long)rounds * (pe - pb);
        if (lane < (int)nexit) sd[r0 + lane] = (uint32_t)(s_dp[lane] >> 32);
        wave_lds_sync();
      }
    } else {
      const unsigned long long q = __popcll(ents & below);  // this lane's entry rank
      for (uint32_t x = part; x < nexit; x += P) {
        if (lane < (int)size) s_dp[lane] = (uint32_t)lane == x ? 0xFFFFFFFFull : kInf64;
        wave_lds_sync();
        const uint32_t rounds = cluster_bf_wave(size, pb, pe, ie.dst, ie.src, ie.wt, s_id, s_dp);
        edges += (unsigned long long)rounds * (pe - pb);
        if (en) sd[rbase + q * nexit + x] = (uint32_t)(s_dp[lane] >> 32);
        wave_lds_sync();
      }
    }
  }
  if (lane == 0 && edges) atomicAdd(&s_edges, edges);
  __syncthreads();
  if (threadIdx.x == 0 && s_edges) atomicAdd(work, s_edges);
}

// One shortcut row with parents, for inspection: forward Bellman-Ford from the
// entry in slot src of cluster k; distance and parent of every exit.
__global__ void __launch_bounds__(kBlock) k_ss_row_one(const uint32_t* __restrict__ cl_off,
                                                       const uint32_t* __restrict__ cl_nexit,
                                                       const uint32_t* __restrict__ cl_mem, Internal ie, uint32_t k,
                                                       uint32_t src, uint32_t* out_d, uint32_t* out_p) {
  __shared__ unsigned long long s_dp[kMaxClusterS];
  __shared__ uint32_t s_id[kMaxClusterS];
  __shared__ int s_flag[3];
  const uint32_t slot0 = cl_off[k], size = cl_off[k + 1] - slot0, nexit = cl_nexit[k];
  for (uint32_t j = threadIdx.x; j < size; j += kBlock) {
    s_id[j] = cl_mem[slot0 + j];
    s_dp[j] = j == src ? 0xFFFFFFFFull : kInf64;
  }
  EdgeCache ec;
  ec.load(ie.off[k], ie.off[k + 1], ie.src, ie.dst, ie.wt);
  cluster_bf(size, ie.off[k], ie.off[k + 1], ec, ie.src, ie.dst, ie.wt, s_id, s_dp, s_flag);
  for (uint32_t j = threadIdx.x; j < nexit; j += kBlock) {
    out_d[j] = (uint32_t)(s_dp[j] >> 32);
    out_p[j] = (uint32_t)s_dp[j];
  }
}

struct SsArgs {
  const uint32_t* off;   // CSR (an exit's sketch row skips its edges inside the cluster)
  const uint32_t* col;
  const uint32_t* w;     // null: unit weights
  const uint8_t* ntype;
  const uint32_t* cid;
  const uint32_t* ent_of;
  const uint32_t* cl_off;
  const uint32_t* cl_mem;
  const uint32_t* cl_nexit;
  const unsigned long long* ss_base;
  const uint32_t* sd;
  unsigned long long* dp;
  uint32_t* wl[2];
  uint32_t* bits[2];
  uint2* heavy;          // two halves of heavy_half entries: chunks made in even / odd rounds
  uint64_t heavy_half;
  Counters* cnt;
  uint8_t* dirty;
};

// Appends v to frontier list `list` (bitmap dedup); ctr counts the list.
__device__ __forceinline__ void ss_append(const SsArgs& a, int list, unsigned long long* ctr, uint32_t v,
                                          bool add_in) {
  bool add = false;
  if (add_in) {
    uint32_t bit = 1u << (v & 31);
    add = !(atomicOr(&a.bits[list][v >> 5], bit) & bit);
  }
  unsigned long long slot = wave_append(add, ctr);
  if (add) a.wl[list][slot] = v;
}

// Marks the clusters of frontier members (exit or inner) and of the source dirty.
__global__ void k_ss_mark(SsArgs a, uint32_t src, int mark_src) {
  const unsigned long long nl = a.cnt->wl_count[0];
  for (unsigned long long i = (unsigned long long)blockIdx.x * kBlock + threadIdx.x; i < nl;
       i += (unsigned long long)gridDim.x * kBlock) {
    const uint32_t v = a.wl[0][i];
    const uint8_t t = a.ntype[v];
    if (t == 1 || t == 3 || t == 4) a.dirty[a.cid[v]] = 1;
  }
  if (mark_src && blockIdx.x == 0 && threadIdx.x == 0 && a.cid[src] != INF32) a.dirty[a.cid[src]] = 1;
}

// In-cluster relaxation of the listed clusters from the members' current state.
// Precompute (final = 0): improved exits join the frontier (list 0).
// Correction (final = 1, after the sketch rounds): the entries and exits hold their
// final distances, so the relaxation gives every member its distance, and the
// packed {dist, parent} minimum over the internal in-edges (together with what the
// external relaxations left in the member's word) its smallest-id parent.
__global__ void __launch_bounds__(kBlock) k_ss_relax_clusters(SsArgs a, Internal ie,
                                                              const uint32_t* __restrict__ list, uint32_t K,
                                                              const unsigned long long* __restrict__ counts,
                                                              unsigned long long* work, int final_pass) {
  __shared__ unsigned long long s_dp[kMaxClusterS];
  __shared__ uint32_t s_id[kMaxClusterS];
  __shared__ int s_flag[3];
  const uint32_t nk = (uint32_t)counts[1];  // the large clusters, at the back of the list
  for (uint32_t li = blockIdx.x; li < nk; li += gridDim.x) {
  const uint32_t k = list[K - 1 - li];
  const uint32_t slot0 = a.cl_off[k], size = a.cl_off[k + 1] - slot0, nexit = a.cl_nexit[k];
  const uint32_t pb = ie.off[k], pe = ie.off[k + 1];
  for (uint32_t j = threadIdx.x; j < size; j += kBlock) {
    const uint32_t v = a.cl_mem[slot0 + j];
    s_id[j] = v;
    s_dp[j] = a.dp[v];
  }
  __syncthreads();
  EdgeCache ec;
  ec.load(pb, pe, ie.src, ie.dst, ie.wt);
  uint32_t rounds = cluster_bf(size, pb, pe, ec, ie.src, ie.dst, ie.wt, s_id, s_dp, s_flag);
  for (uint32_t j0 = 0; j0 < size; j0 += kBlock) {  // uniform trip count for the wave ballots
    const uint32_t j = j0 + threadIdx.x;
    bool improved = false;
    uint32_t v = 0;
    if (j < size) {
      v = a.cl_mem[slot0 + j];
      const unsigned long long x = s_dp[j];
      if (x < a.dp[v]) {
        a.dp[v] = x;
        improved = !final_pass && j < nexit && (x >> 32) != 0xFFFFFFFFull;
      }
    }
    ss_append(a, 0, &a.cnt->wl_count[0], v, improved);
  }
  if (threadIdx.x == 0) {
    atomicAdd(work, (unsigned long long)rounds * (pe - pb));
    if (final_pass) a.dirty[k] = 0;
  }
  __syncthreads();  // LDS of the next cluster
  }
}

// k_ss_relax_clusters for the small clusters (list[0, counts[0])), one wave each.
__global__ void __launch_bounds__(kBlock) k_ss_relax_small(SsArgs a, Internal ie, const uint32_t* __restrict__ list,
                                                           const unsigned long long* __restrict__ counts,
                                                           unsigned long long* work, int final_pass) {
  __shared__ unsigned long long s_dp_w[kBlock / 64][kSmallCluster];
  __shared__ uint32_t s_id_w[kBlock / 64][kSmallCluster];
  __shared__ unsigned long long s_edges;
  const int wv = threadIdx.x >> 6, lane = lane_id();
  unsigned long long* s_dp = s_dp_w[wv];
  uint32_t* s_id = s_id_w[wv];
  if (threadIdx.x == 0) s_edges = 0;
  __syncthreads();
  const uint32_t nk = (uint32_t)counts[0];
  unsigned long long edges = 0;
  for (uint32_t li = blockIdx.x * (kBlock / 64) + wv; li < nk; li += gridDim.x * (kBlock / 64)) {
    const uint32_t k = list[li];
    const uint32_t slot0 = a.cl_off[k], size = a.cl_off[k + 1] - slot0, nexit = a.cl_nexit[k];
    const uint32_t pb = ie.off[k], pe = ie.off[k + 1];
    uint32_t v = 0;
    unsigned long long cur = kInf64;
    if (lane < (int)size) {
      v = a.cl_mem[slot0 + lane];
      cur = a.dp[v];
      s_id[lane] = v;
      s_dp[lane] = cur;
    }
    wave_lds_sync();
    const uint32_t rounds = cluster_bf_wave(size, pb, pe, ie.src, ie.dst, ie.wt, s_id, s_dp);
    edges += (unsigned long long)rounds * (pe - pb);
    bool improved = false;
    if (lane < (int)size) {
      const unsigned long long x = s_dp[lane];
      if (x < cur) {
        a.dp[v] = x;
        improved = !final_pass && (uint32_t)lane < nexit && (x >> 32) != 0xFFFFFFFFull;
      }
    }
    ss_append(a, 0, &a.cnt->wl_count[0], v, improved);
    if (lane == 0 && final_pass) a.dirty[k] = 0;
    wave_lds_sync();
  }
  if (lane == 0 && edges) atomicAdd(&s_edges, edges);
  __syncthreads();
  if (threadIdx.x == 0 && s_edges) atomicAdd(work, s_edges);
}

__device__ __forceinline__ void ss_relax(const SsArgs& a, int nxt, unsigned long long* ctr, uint32_t v,
                                         unsigned long long nd, uint32_t parent) {
  bool add = false;
  // words only decrease during a query, so a candidate no smaller than a plain
  // read of the word (possibly stale, never below the current value) cannot win:
  // only the candidates that may improve pay for the atomic
  if (nd < 0xFFFFFFFFull) {
    const unsigned long long cand = (nd << 32) | parent;
    if (cand < a.dp[v]) {
      const unsigned long long old = atomicMin(&a.dp[v], cand);
      add = (old >> 32) > nd;
    }
  }
  ss_append(a, nxt, ctr, v, add);
}

// Rounds are numbered on the host and passed to the kernels, so no kernel has to
// close a round: round r reads list r % 2 counted by rc[r % 3] (round 0: the seeds
// in list 0, counted by wl_count[0]), appends to list (r + 1) % 2 counted by
// rc[(r + 1) % 3] and clears rc[(r + 2) % 3] for round r + 1 (last read by round
// r - 1).  Rows longer than kHeavyDeg become kChunk-edge chunks that the next
// round relaxes (label-correcting rounds converge in any order; a chunk reads its
// row's distance when it runs): round r writes chunks to heavy half r % 2 counted
// by hc[r % 3], relaxes the previous round's from the other half (hc[(r - 1) % 3])
// and clears hc[(r + 1) % 3].  An empty frontier with no pending chunks, or the
// round cap, sets done in round r's first workgroup; every later launch returns.
__device__ __forceinline__ unsigned long long* round_in(Counters* c, uint32_t r) {
  return r == 0 ? &c->wl_count[0] : &c->rc[r % 3];
}

// One sketch round over the current frontier: Singles relax their CSR row,
// exits their external edges, entries their shortcut row to the exits.
__global__ void __launch_bounds__(kBlock) k_ss_push(SsArgs a, uint32_t r, uint32_t max_rounds) {
  __shared__ uint32_t s_scan[kBlock + 1];
  __shared__ uint32_t s_start[kBlock];
  __shared__ uint32_t s_warp[kBlock / 64 + 1];
  __shared__ uint32_t s_u[kBlock], s_du[kBlock], s_iblen[kBlock], s_mbase[kBlock], s_rb[kBlock], s_k[kBlock];
  __shared__ unsigned long long s_rbase[kBlock];
  __shared__ unsigned long long s_edges, s_sc;
  Counters* cnt = a.cnt;
  if (cnt->done) return;
  const int cur = (int)(r & 1), nxt = cur ^ 1;
  const unsigned long long nl = *round_in(cnt, r);
  const unsigned long long nh = r ? cnt->hc[(r - 1) % 3] : 0ull;  // chunks the previous round left
  unsigned long long* out = &cnt->rc[(r + 1) % 3];
  unsigned long long* hcount = &cnt->hc[r % 3];
  uint2* hnew = a.heavy + (r & 1) * a.heavy_half;
  const uint2* hold = a.heavy + ((r + 1) & 1) * a.heavy_half;
  const bool stop = (nl == 0 && nh == 0) || r >= max_rounds;  // counts of earlier launches: all agree
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    cnt->rc[(r + 2) % 3] = 0;
    cnt->hc[(r + 1) % 3] = 0;
    if (r == 1) cnt->wl_count[0] = 0;  // the seeds were round 0's input
    if (stop) {
      cnt->done = 1;
      if (nl || nh) cnt->aux[1] = 1;  // stopped by the round cap
    } else {
      cnt->aux[6] += nl;
      cnt->rounds = r + 1;
    }
  }
  if (stop) return;
  const uint32_t vt = frontier_tile<kBlock>(nl);
  const unsigned long long tiles = (nl + vt - 1) / vt;
  if (threadIdx.x == 0) s_edges = s_sc = 0;
  __syncthreads();
  for (unsigned long long t = blockIdx.x; t < tiles; t += gridDim.x) {
    const unsigned long long i = t * vt + threadIdx.x;
    uint32_t len = 0;
    if (threadIdx.x < vt && i < nl) {
      const uint32_t u = a.wl[cur][i];
      atomicAnd(&a.bits[cur][u >> 5], ~(1u << (u & 31)));
      const uint32_t du = (uint32_t)(a.dp[u] >> 32);
      const uint8_t ty = a.ntype[u];
      uint32_t iblen = 0, sclen = 0;
      s_k[threadIdx.x] = ty == 0 ? INF32 : a.cid[u];
      if (du != INF32) {
        if (ty == 0 || ty == 1 || ty == 3) {
          s_rb[threadIdx.x] = a.off[u];
          iblen = a.off[u + 1] - a.off[u];
        }
        if (ty == 2 || ty == 3) {
          const uint32_t k = a.cid[u];
          a.dirty[k] = 1;  // the correction revisits this cluster
          sclen = a.cl_nexit[k];
          s_mbase[threadIdx.x] = a.cl_off[k];
          s_rbase[threadIdx.x] = a.ss_base[a.ent_of[u]];
        }
        if (iblen > kHeavyDeg) {  // long rows go to 2048-edge chunks
          const uint32_t nch = (iblen + kChunk - 1) / kChunk;
          const unsigned long long slot = atomicAdd(hcount, (unsigned long long)nch);
          for (uint32_t c = 0; c < nch; ++c) hnew[slot + c] = make_uint2(u, c);
          atomicAdd(&s_edges, (unsigned long long)iblen);
          iblen = 0;
        }
      }
      s_u[threadIdx.x] = u;
      s_du[threadIdx.x] = du;
      s_iblen[threadIdx.x] = iblen;
      len = iblen + sclen;
      if (len) atomicAdd(&s_edges, (unsigned long long)len);
      if (sclen) atomicAdd(&s_sc, (unsigned long long)sclen);
    }
    block_adjacency<kBlock>(0, len, s_scan, s_start, s_warp, [&](int slot, uint32_t k, uint32_t) {
      const uint32_t u = s_u[slot];
      const unsigned long long du = s_du[slot];
      if (k < s_iblen[slot]) {
        const uint32_t e = s_rb[slot] + k;
        const uint32_t x = a.col[e];
        const bool inside = s_k[slot] != INF32 && a.cid[x] == s_k[slot];  // an exit's edge inside its cluster
        ss_relax(a, nxt, out, x, inside ? 0xFFFFFFFFull : du + (a.w ? a.w[e] : 1u), u);
      } else {
        const uint32_t j = k - s_iblen[slot];
        const unsigned long long r = s_rbase[slot] + j;
        const uint32_t dj = a.sd[r];
        ss_relax(a, nxt, out, a.cl_mem[s_mbase[slot] + j], dj == INF32 ? 0xFFFFFFFFull : du + dj, INF32);
      }
    });
  }
  // the previous round's heavy chunks
  for (unsigned long long h = blockIdx.x; h < nh; h += gridDim.x) {
    const uint2 hv = hold[h];
    const uint32_t u = hv.x;
    const unsigned long long du = a.dp[u] >> 32;
    const uint32_t kc = a.ntype[u] == 0 ? INF32 : a.cid[u];
    const uint32_t b = a.off[u] + hv.y * kChunk;
    const uint32_t e = min(a.off[u + 1], b + kChunk);
    for (uint32_t k0 = b; k0 < e; k0 += kBlock) {  // same trip count for every lane (wave ballots)
      const uint32_t k = k0 + threadIdx.x;
      if (k < e) {
        const uint32_t x = a.col[k];
        const bool inside = kc != INF32 && a.cid[x] == kc;
        ss_relax(a, nxt, out, x, inside ? 0xFFFFFFFFull : du + (a.w ? a.w[k] : 1u), u);
      } else {
        ss_append(a, nxt, out, 0, false);
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0 && s_edges) atomicAdd(&a.cnt->edges, s_edges);
  if (threadIdx.x == 0 && s_sc) atomicAdd(&a.cnt->aux[7], s_sc);
}

// Correction: (a) every member takes min over entries of dist(entry) + in-cluster
// distance (parent = in-cluster predecessor); (b) with final distances, every
// member min-reduces (dist(src) + w, src) over its internal in-edges, which fixes
// equal-distance parents to the smallest id.
static SsArgs ss_args(suminc_ctx* c) {
  Sketch& s = c->sk;
  SsArgs a;
  a.off = c->out_off.as<uint32_t>();
  a.col = c->out_col.as<uint32_t>();
  a.w = c->wkind == SUMINC_W_U32 ? c->out_w.as<uint32_t>() : nullptr;
  a.ntype = s.ntype.as<uint8_t>();
  a.cid = s.cid.as<uint32_t>();
  a.ent_of = s.ent_of.as<uint32_t>();
  a.cl_off = s.cl_off.as<uint32_t>();
  a.cl_mem = s.cl_mem.as<uint32_t>();
  a.cl_nexit = s.cl_nexit.as<uint32_t>();
  a.ss_base = s.ss_base.as<unsigned long long>();
  a.sd = s.ss_dist.as<uint32_t>();
  a.dp = c->sssp_dp.as<unsigned long long>();
  a.wl[0] = c->sssp_wl0.as<uint32_t>();
  a.wl[1] = c->sssp_wl1.as<uint32_t>();
  a.bits[0] = c->sssp_flag0.as<uint32_t>();
  a.bits[1] = c->sssp_flag1.as<uint32_t>();
  a.heavy = c->heavy.as<uint2>();
  a.heavy_half = c->heavy.cap / sizeof(uint2) / 2;
  a.cnt = dev_counters(c);
  a.dirty = s.ss_dirty.as<uint8_t>();
  return a;
}

static unsigned small_grid(uint32_t K) {
  return std::max(1u, std::min(kClusterGrid, (K + kBlock / 64 - 1) / (kBlock / 64)));
}

// In-cluster relaxation of the listed clusters: small ones one per wave, large
// ones one per workgroup.
static void relax_clusters(suminc_ctx* c, const SsArgs& a, const unsigned long long* lcounts, int final_pass) {
  Sketch& s = c->sk;
  k_ss_relax_small<<<small_grid(s.K), kBlock, 0, c->stream>>>(a, internal_of(s), s.ss_list.as<uint32_t>(), lcounts,
                                                             &a.cnt->aux[8], final_pass);
  k_ss_relax_clusters<<<std::min(s.K, kClusterGrid), kBlock, 0, c->stream>>>(
      a, internal_of(s), s.ss_list.as<uint32_t>(), s.K, lcounts, &a.cnt->aux[8], final_pass);
}

// Recomputes the SSSP rows of the clusters whose rows are not current; the row
// kernels mark the clusters they filled valid.  `zeroed`: the list counts in
// ss_work[1..2] are already zero.
static void compute_sssp_rows(suminc_ctx* c, bool zeroed) {
  Sketch& s = c->sk;
  if (!s.ss_need && s.ss_rows) return;
  if (!s.ss_rows) SUMINC_HIP(hipMemsetAsync(s.ss_valid.p, 0, (size_t)s.K + 1, c->stream));
  if (!s.ss_rows || s.ss_dist.cap < s.ss_len * 4 + 4) {
    s.ss_dist.ensure_slack(s.ss_len * 4 + 4);
  }
  s.ss_work.ensure(64);
  s.ss_list.ensure((size_t)s.K * 4 + 4);
  if (!zeroed) SUMINC_HIP(hipMemsetAsync(s.ss_work.p, 0, 64, c->stream));
  unsigned long long* lcounts = s.ss_work.as<unsigned long long>() + 1;
  // internal edges read by the row kernels: kept apart from the query's traversed edges
  unsigned long long* work = s.ss_work.as<unsigned long long>();
  if (s.K) {
    k_cluster_lists<<<grid_for(s.K), kBlock, 0, c->stream>>>(s.K, s.ss_valid.as<uint8_t>(), 0,
                                                             s.cl_off.as<uint32_t>(), s.ss_list.as<uint32_t>(),
                                                             lcounts);
    k_ss_shortcuts_small<<<small_grid(s.K), kBlock, 0, c->stream>>>(
        s.cl_off.as<uint32_t>(), s.cl_nexit.as<uint32_t>(), s.cl_mem.as<uint32_t>(), s.slot_ent.as<uint8_t>(),
        s.lay_base.as<ClBase>(), internal_of(s), s.ss_dist.as<uint32_t>(), s.ss_list.as<uint32_t>(), lcounts,
        s.ss_valid.as<uint8_t>(), work);
    k_ss_shortcuts<<<std::min(s.K, kClusterGrid), kBlock, 0, c->stream>>>(
        s.cl_off.as<uint32_t>(), s.cl_nexit.as<uint32_t>(), s.cl_mem.as<uint32_t>(), s.cl_eoff.as<uint32_t>(),
        s.cl_ent.as<uint32_t>(), s.lidx.as<uint32_t>(), s.ss_base.as<unsigned long long>(), internal_of(s),
        s.ss_dist.as<uint32_t>(), s.ss_list.as<uint32_t>(), s.K, lcounts, s.ss_valid.as<uint8_t>(), work);
  }
  SUMINC_HIP(hipGetLastError());
  s.ss_rows = true;
  s.ss_need = false;
}

__global__ void k_ss_begin(Counters* cnt, unsigned long long* lists, unsigned long long* seeds) {
  if (threadIdx.x < 8) lists[threadIdx.x] = 0;  // work slot and the three list-count pairs
  if (threadIdx.x) return;
  cnt->wl_count[0] = *seeds;  // list 0 holds the pending seeds
  *seeds = 0;
  cnt->aux[8] = 0;  // internal edges read by the in-cluster kernels
  cnt->wl_count[1] = 0;
  cnt->heavy_count = 0;
  cnt->rc[0] = cnt->rc[1] = cnt->rc[2] = 0;
  cnt->hc[0] = cnt->hc[1] = cnt->hc[2] = 0;
  cnt->aux[1] = 0;
  cnt->aux[6] = 0;
  cnt->aux[7] = 0;
  cnt->done = 0;
  cnt->rounds = 0;
  cnt->edges = 0;
}

void sum_sssp_run(suminc_ctx* c, suminc_stats* st) {
  SUMINC_REQUIRE(c->sssp_ready, SUMINC_E_STATE, "SSSP state not initialised");
  SUMINC_REQUIRE(c->sssp_app == SUMINC_APP_SSSP, SUMINC_E_STATE,
                 "the supernode query runs SSSP only (the state was set up for BFS or SSWP)");
  sketch_refresh(c);
  Sketch& s = c->sk;
  s.ss_work.ensure(64);
  k_ss_begin<<<1, 64, 0, c->stream>>>(dev_counters(c), s.ss_work.as<unsigned long long>(),
                                      c->sssp_seeds.as<unsigned long long>());
  compute_sssp_rows(c, true);
  if (!s.ss_dirty.p || s.ss_dirty.cap < (size_t)s.K + 1 || !s.ss_dirty_clear) {
    s.ss_dirty.ensure((size_t)s.K + 1);
    SUMINC_HIP(hipMemsetAsync(s.ss_dirty.p, 0, (size_t)s.K + 1, c->stream));
  }
  s.ss_dirty_clear = false;
  const int max_rounds = c->prm.max_rounds > 0 ? c->prm.max_rounds : 1000000;
  const bool first = c->sssp_pending;
  c->sssp_pending = false;
  SsArgs a = ss_args(c);
  Counters* cnt = a.cnt;
  s.ss_list.ensure((size_t)s.K * 4 + 4);
  unsigned long long* lcounts = s.ss_work.as<unsigned long long>() + 3;
  SUMINC_HIP(hipEventRecord(c->ev_start, c->stream));
  // frontier list 0 holds the seeds (source after init, repair seeds after a batch)
  k_ss_mark<<<256, kBlock, 0, c->stream>>>(a, c->sssp_src, first ? 1 : 0);
  if (s.K) {
    k_cluster_lists<<<grid_for(s.K), kBlock, 0, c->stream>>>(s.K, s.ss_dirty.as<uint8_t>(), 1,
                                                             s.cl_off.as<uint32_t>(), s.ss_list.as<uint32_t>(),
                                                             lcounts);
    relax_clusters(c, a, lcounts, 0);
  }
  SUMINC_HIP(hipGetLastError());
  int group = c->hint_sssp_sketch;  // rounds of the previous query + 1 first
  int more = 4;                     // then a few more rounds at a time, widening
  size_t slot = 0;
  while (true) {
    for (int i = 0; i < group; ++i, ++slot) {
      if (c->prof_on) SUMINC_HIP(hipEventRecord(prof_event(c, 2 * slot), c->stream));
      k_ss_push<<<push_grid(), kBlock, 0, c->stream>>>(a, (uint32_t)slot, (uint32_t)max_rounds);
      if (c->prof_on) SUMINC_HIP(hipEventRecord(prof_event(c, 2 * slot + 1), c->stream));
    }
    SUMINC_HIP(hipGetLastError());
    read_counters(c);
    if (c->host_counters->done) break;
    group = more;
    more = std::min(more * 2, 64);
  }
  const Counters* h = c->host_counters;
  const int cur = (int)(h->rounds & 1);  // the list the last (empty) round read
  SUMINC_REQUIRE(h->aux[1] == 0, SUMINC_E_NOT_CONVERGED, "round cap reached with a non-empty frontier");
  const double rounds = (double)h->rounds, edges = (double)h->edges;
  c->hint_sssp_sketch = std::max(2, std::min(64, (int)h->rounds + 1));
  if (c->prof_on) {
    prof_collect(c, SUMINC_KERNEL_SSSP_SKETCH, (size_t)h->rounds, 0, 2);
    suminc_kernel_profile& p = c->prof[SUMINC_KERNEL_SSSP_SKETCH];
    // algorithmic bytes: a sketch edge moves column + weight (8 B) and a 32-bit
    // distance read-modify-write (8 B); a shortcut slot moves distance + parent +
    // member id (12 B) and the read-modify-write (8 B); a frontier vertex 16 B
    const double sc = (double)h->aux[7];
    p.edges += edges;
    p.vertices += (double)h->aux[6];
    p.alg_bytes += 16.0 * (edges - sc) + 20.0 * sc + 16.0 * (double)h->aux[6];
  }
  if (s.K) {
    unsigned long long* fcounts = s.ss_work.as<unsigned long long>() + 5;
    k_cluster_lists<<<grid_for(s.K), kBlock, 0, c->stream>>>(s.K, s.ss_dirty.as<uint8_t>(), 1,
                                                             s.cl_off.as<uint32_t>(), s.ss_list.as<uint32_t>(),
                                                             fcounts);
    relax_clusters(c, a, fcounts, 1);  // clears the dirty flags it visits
  }
  SUMINC_HIP(hipGetLastError());
  SUMINC_HIP(hipEventRecord(c->ev_stop, c->stream));
  read_counters(c);  // synchronises the stream
  s.ss_dirty_clear = true;
  float ms = 0;
  SUMINC_HIP(hipEventElapsedTime(&ms, c->ev_start, c->ev_stop));
  const unsigned long long wk = c->host_counters->aux[8];
  if (cur != 0) {  // keep list/bitmap 0 as the current (empty) frontier
    c->sssp_wl0.swap(c->sssp_wl1);
    c->sssp_flag0.swap(c->sssp_flag1);
  }
  suminc_stats out{};
  out.rounds = rounds;
  out.edges_traversed = edges + (double)wk;
  out.kernel_ms = ms;
  out.push_rounds = rounds;
  out.precompute_edges = (double)wk;  // internal edges read by the precompute
  c->stats = out;
  if (st) *st = out;
}

}  // namespace suminc

namespace suminc {

void sketch_sssp_shortcut(suminc_ctx* c, uint32_t v, uint32_t cap, uint32_t* members, uint32_t* dist,
                          uint32_t* parent, uint32_t* count, uint32_t* nexit) {
  SUMINC_REQUIRE(c->has_graph, SUMINC_E_STATE, "no graph");
  sketch_refresh(c);
  Sketch& s = c->sk;
  SUMINC_REQUIRE(v < c->n, SUMINC_E_INVALID, "vertex out of range");
  compute_sssp_rows(c, false);
  uint32_t e = INF32, k = INF32;
  SUMINC_HIP(hipMemcpy(&e, s.ent_of.as<uint32_t>() + v, 4, hipMemcpyDeviceToHost));
  SUMINC_REQUIRE(e != INF32, SUMINC_E_INVALID, "vertex is not an entry of a supernode");
  SUMINC_HIP(hipMemcpy(&k, s.cid.as<uint32_t>() + v, 4, hipMemcpyDeviceToHost));
  uint32_t co[2], ne = 0;
  unsigned long long base = 0;
  SUMINC_HIP(hipMemcpy(co, s.cl_off.as<uint32_t>() + k, 8, hipMemcpyDeviceToHost));
  SUMINC_HIP(hipMemcpy(&ne, s.cl_nexit.as<uint32_t>() + k, 4, hipMemcpyDeviceToHost));
  SUMINC_HIP(hipMemcpy(&base, s.ss_base.as<unsigned long long>() + e, 8, hipMemcpyDeviceToHost));
  const uint32_t size = co[1] - co[0];
  if (count) *count = size;
  if (nexit) *nexit = ne;
  SUMINC_REQUIRE(cap >= size || !(members || dist || parent), SUMINC_E_INVALID, "output capacity too small");
  if (members) SUMINC_HIP(hipMemcpy(members, s.cl_mem.as<uint32_t>() + co[0], (size_t)size * 4, hipMemcpyDeviceToHost));
  // rows hold the exits (the first nexit members) only
  if (dist) {
    for (uint32_t j = ne; j < size; ++j) dist[j] = INF32;
    if (ne) SUMINC_HIP(hipMemcpy(dist, s.ss_dist.as<uint32_t>() + base, (size_t)ne * 4, hipMemcpyDeviceToHost));
  }
  if (parent) {  // the rows keep distances only: the parents come from one forward run
    for (uint32_t j = ne; j < size; ++j) parent[j] = INF32;
    if (ne) {
      uint32_t src = 0;
      SUMINC_HIP(hipMemcpy(&src, s.lidx.as<uint32_t>() + v, 4, hipMemcpyDeviceToHost));
      DevBuf out;
      out.ensure((size_t)ne * 8);
      k_ss_row_one<<<1, kBlock, 0, c->stream>>>(s.cl_off.as<uint32_t>(), s.cl_nexit.as<uint32_t>(),
                                                 s.cl_mem.as<uint32_t>(), internal_of(s), k, src,
                                                 out.as<uint32_t>(), out.as<uint32_t>() + ne);
      SUMINC_HIP(hipGetLastError());
      SUMINC_HIP(hipStreamSynchronize(c->stream));
      SUMINC_HIP(hipMemcpy(parent, out.as<uint32_t>() + ne, (size_t)ne * 4, hipMemcpyDeviceToHost));
    }
  }
}

}  // namespace suminc
