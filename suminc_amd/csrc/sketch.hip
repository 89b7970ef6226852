// sketch.hip -- supernode compression on the device (the mirror-free form of
// the reference's compressor and SumSyncIterWorker phases):
//
//   clustering   built-in leaf clustering (each hub with its private
//                neighbours), or a caller-supplied assignment like the
//                reference's "<efile>.c_<max_node_num>" cluster file
//                (compressor_base.h:3160-3250).
//   acceptance   a cluster becomes a supernode iff its members are at least
//                min_node_num and inner_edges - |entries|*|exits| > compress_threshold
//                (get_init_supernode_by_clusters, compressor_base.h:3253-3568, the
//                benefit[0] branch that mirror_k = 1e8 leaves; the other three
//                benefits collapse onto it without mirrors).
//   node types   Single / OnlyOut (exit) / OnlyIn (entry) / BothOutIn / Inner
//                (get_nodetype_mirror, compressor_base.h:4144-4182, types.h:70-78).
//   layout       members of cluster k at cl_mem[cl_off[k]..], exits first; one
//                dense shortcut row per entry with one slot per member (reach mass
//                and residual for PageRank; distance and parent for SSSP); the
//                sketch out-edges ib (a Single's whole out-list, an exit's edges
//                that leave its cluster; sketch2csr_mirror, compressor_base.h:
//                4636-4899); the internal in-edges of every member.
//   shortcuts    one workgroup per cluster runs the unit injection of every
//                entry inside the cluster in LDS (build_iter_index_mirror +
//                inc_run_to_convergence_mirror, iter_compressor.h:318-431): Jacobi
//                rounds r += delta, delta <- internal pull with weight d/outdeg,
//                until the mass processed in a round is <= T/N*|C| or 100 rounds.
//   PR query     precompute of dirty clusters (precompute_spnode_one/two +
//                first_step exit sends, iter_compressor.h:534-785,
//                sum_sync_iter_worker.h:348-438), sketch rounds by node type
//                (sum_sync_iter_worker.h:949-1111), flush of the exits' pending
//                inflow, correction through inner_value/inner_delta
//                (:1414-1469), then the plain residual phase.
#include <cstring>
#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <vector>

#include "ctx.hpp"
#include "device.hpp"
#include "pr.hpp"
#include "push.hpp"
#include "sketch.hpp"

namespace suminc {

constexpr int kMaxCluster = 2048;  // LDS-resident cluster bound of the shortcut kernels

__device__ __forceinline__ uint32_t row_of(const uint32_t* off, uint32_t n, uint32_t e) {
  uint32_t lo = 0, hi = n - 1;
  while (lo < hi) {
    uint32_t mid = lo + ((hi - lo + 1) >> 1);
    if (off[mid] <= e) lo = mid; else hi = mid - 1;
  }
  return lo;
}

template <class F>
static void cub(suminc_ctx* c, F&& f) {
  size_t bytes = 0;
  SUMINC_HIP(f((void*)nullptr, bytes));
  c->tmp.ensure(bytes);
  SUMINC_HIP(f(c->tmp.p, bytes));
}

__global__ void k_fill32(uint32_t* p, uint64_t n, uint32_t v) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// leaf clustering
// ---------------------------------------------------------------------------
// anchor[v] = the single distinct neighbour of v (in or out), else INF
__global__ void k_leaf_anchor(uint32_t n, const uint32_t* ooff, const uint32_t* ocol, const uint32_t* ioff,
                              const uint32_t* icol, uint32_t* anchor) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  uint32_t dout = ooff[v + 1] - ooff[v], din = ioff[v + 1] - ioff[v];
  uint32_t h = INF32;
  bool leaf = dout + din > 0 && dout <= 1 && din <= 1;
  if (leaf && dout == 1) h = ocol[ooff[v]];
  if (leaf && din == 1) {
    uint32_t x = icol[ioff[v]];
    if (h != INF32 && x != h) leaf = false;
    h = x;
  }
  anchor[v] = leaf ? h : INF32;
}

// Leaves keyed by (anchor, leaf); a leaf pair anchors on its smaller id.
__global__ void k_leaf_keys(uint32_t n, const uint32_t* anchor, uint64_t* keys) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  uint32_t h = anchor[v];
  bool member = h != INF32 && !(anchor[h] == v && v < h);
  keys[v] = member ? (((uint64_t)h << 32) | v) : ~0ull;
}

__global__ void k_seg_heads(const uint64_t* keys, uint32_t L, uint8_t* head) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L) return;
  head[i] = (i == 0) || (keys[i] >> 32) != (keys[i - 1] >> 32);
}

__global__ void k_seg_ok(const uint32_t* seg_start, const unsigned* nseg_p, uint32_t L, uint32_t max_node,
                         uint32_t min_node, uint32_t* ok) {
  uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t S = *nseg_p;
  if (s > S) return;
  if (s == S) { ok[S] = 0; return; }
  uint32_t len = (s + 1 < S ? seg_start[s + 1] : L) - seg_start[s];
  uint32_t take = min(len, max_node - 1);
  ok[s] = (take + 1 >= min_node) ? 1u : 0u;
}

__global__ void k_leaf_assign(const uint64_t* keys, uint32_t L, const uint32_t* seg_start, const unsigned* nseg_p,
                              const uint32_t* ok, const uint32_t* okpre, uint32_t max_node, uint32_t* raw) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L) return;
  const uint32_t S = *nseg_p;
  uint32_t lo = 0, hi = S - 1;  // segment of position i
  while (lo < hi) {
    uint32_t mid = lo + ((hi - lo + 1) >> 1);
    if (seg_start[mid] <= i) lo = mid; else hi = mid - 1;
  }
  if (!ok[lo]) return;
  uint32_t rank = i - seg_start[lo];
  if (rank >= max_node - 1) return;
  uint32_t id = okpre[lo];
  raw[(uint32_t)keys[i]] = id;
  if (rank == 0) raw[(uint32_t)(keys[i] >> 32)] = id;
}

void cluster_leaves(suminc_ctx* c, uint32_t max_node, uint32_t min_node, uint32_t* out) {
  SUMINC_REQUIRE(c->has_graph, SUMINC_E_STATE, "no graph");
  SUMINC_REQUIRE(max_node >= 2 && max_node <= (uint32_t)kMaxCluster, SUMINC_E_INVALID,
                 "max_node_num must be in [2, 2048]");
  const uint32_t n = c->n;
  Sketch& s = c->sk;
  DevBuf anchor, ka, kb, head, seg, nseg, ok, okpre;
  anchor.ensure((size_t)n * 4); ka.ensure((size_t)n * 8); kb.ensure((size_t)n * 8); head.ensure(n + 1);
  seg.ensure((size_t)n * 4 + 4); nseg.ensure(8); ok.ensure(((size_t)n + 1) * 4); okpre.ensure(((size_t)n + 1) * 4);
  k_leaf_anchor<<<grid_for(n), kBlock, 0, c->stream>>>(n, c->out_off.as<uint32_t>(), c->out_col.as<uint32_t>(),
                                                       c->in_off.as<uint32_t>(), c->in_col.as<uint32_t>(),
                                                       anchor.as<uint32_t>());
  k_leaf_keys<<<grid_for(n), kBlock, 0, c->stream>>>(n, anchor.as<uint32_t>(), ka.as<uint64_t>());
  cub(c, [&](void* p, size_t& b) {
    return rocprim::radix_sort_keys(p, b, ka.as<uint64_t>(), kb.as<uint64_t>(), (size_t)n, 0, 64, c->stream);
  });
  // leaves are the keys before the first sentinel
  cub(c, [&](void* p, size_t& b) {
    return rocprim::lower_bound(p, b, kb.as<uint64_t>(), rocprim::constant_iterator<uint64_t>(~0ull),
                                okpre.as<uint32_t>(), (size_t)n, (size_t)1, rocprim::less<uint64_t>(), c->stream);
  });
  uint32_t L = 0;
  SUMINC_HIP(hipMemcpyAsync(&L, okpre.p, 4, hipMemcpyDeviceToHost, c->stream));
  SUMINC_HIP(hipStreamSynchronize(c->stream));
  s.raw_cid.ensure((size_t)n * 4);
  k_fill32<<<grid_for(n), kBlock, 0, c->stream>>>(s.raw_cid.as<uint32_t>(), n, INF32);
  uint32_t K = 0;
  if (L) {
    k_seg_heads<<<grid_for(L), kBlock, 0, c->stream>>>(kb.as<uint64_t>(), L, head.as<uint8_t>());
    cub(c, [&](void* p, size_t& b) {
      return rocprim::select(p, b, rocprim::counting_iterator<uint32_t>(0), head.as<uint8_t>(), seg.as<uint32_t>(),
                             nseg.as<unsigned>(), (size_t)L, c->stream);
    });
    k_seg_ok<<<grid_for((uint64_t)L + 1), kBlock, 0, c->stream>>>(seg.as<uint32_t>(), nseg.as<unsigned>(), L, max_node,
                                                                  min_node, ok.as<uint32_t>());
    cub(c, [&](void* p, size_t& b) {
      return rocprim::exclusive_scan(p, b, ok.as<uint32_t>(), okpre.as<uint32_t>(), 0u, (size_t)L + 1,
                                     rocprim::plus<uint32_t>(), c->stream);
    });
    k_leaf_assign<<<grid_for(L), kBlock, 0, c->stream>>>(kb.as<uint64_t>(), L, seg.as<uint32_t>(), nseg.as<unsigned>(),
                                                         ok.as<uint32_t>(), okpre.as<uint32_t>(), max_node,
                                                         s.raw_cid.as<uint32_t>());
    unsigned S = 0;
    SUMINC_HIP(hipMemcpyAsync(&S, nseg.p, 4, hipMemcpyDeviceToHost, c->stream));
    SUMINC_HIP(hipStreamSynchronize(c->stream));
    SUMINC_HIP(hipMemcpyAsync(&K, okpre.as<uint32_t>() + S, 4, hipMemcpyDeviceToHost, c->stream));
  }
  SUMINC_HIP(hipGetLastError());
  SUMINC_HIP(hipStreamSynchronize(c->stream));
  s.Kraw = K;
  s.built = false;
  s.raw_acc.release();
  s.rlist.release();
  if (out) SUMINC_HIP(hipMemcpy(out, s.raw_cid.p, (size_t)n * 4, hipMemcpyDeviceToHost));
}

void set_clusters(suminc_ctx* c, const uint32_t* cluster_of) {
  SUMINC_REQUIRE(c->has_graph, SUMINC_E_STATE, "no graph");
  SUMINC_REQUIRE(cluster_of, SUMINC_E_INVALID, "null cluster assignment");
  const uint32_t n = c->n;
  uint32_t K = 0;
  for (uint32_t v = 0; v < n; ++v)
    if (cluster_of[v] != INF32) {
      SUMINC_REQUIRE(cluster_of[v] < n, SUMINC_E_INVALID, "cluster id must be < n or SUMINC_INF_U32");
      K = std::max(K, cluster_of[v] + 1);
    }
  Sketch& s = c->sk;
  s.raw_cid.ensure((size_t)n * 4);
  SUMINC_HIP(hipMemcpy(s.raw_cid.p, cluster_of, (size_t)n * 4, hipMemcpyHostToDevice));
  s.Kraw = K;
  s.built = false;
  s.raw_acc.release();
  s.rlist.release();
}

// ---------------------------------------------------------------------------
// structure.  The first build scans the rows of the clustered vertices once (one
// wave per vertex); afterwards every batch updates the per-member edge counts and
// the internal edge list from its groups, so a rebuild costs O(M + internal
// edges) and never reads a row.
// ---------------------------------------------------------------------------
constexpr int kWaves = kBlock / 64;

__global__ void k_raw_flag(uint32_t n, const uint32_t* raw, uint8_t* flag) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) flag[v] = raw[v] != INF32;
}

// Edges leaving / entering the raw cluster of each listed vertex and its internal
// in-edges (judge_out_bound_node / get_nodetype_mirror, compressor_base.h:
// 3924-3956, 4144-4182); `inner` gets the internal edge count of every raw cluster.
__global__ void __launch_bounds__(kBlock) k_member_scan(const uint32_t* __restrict__ list, uint32_t L,
                                                        const uint32_t* __restrict__ raw,
                                                        const uint32_t* __restrict__ ooff,
                                                        const uint32_t* __restrict__ ocol,
                                                        const uint32_t* __restrict__ ioff,
                                                        const uint32_t* __restrict__ icol, uint32_t* ext_out,
                                                        uint32_t* ext_in, uint32_t* inner) {
  const uint32_t w = blockIdx.x * kWaves + (threadIdx.x >> 6);
  if (w >= L) return;  // whole wave
  const int lane = lane_id();
  const uint32_t v = list[w], k = raw[v];
  uint32_t ext = 0, inn = 0, exin = 0;
  for (uint32_t e = ooff[v] + lane; e < ooff[v + 1]; e += 64) ext += raw[ocol[e]] != k;
  for (uint32_t e = ioff[v] + lane; e < ioff[v + 1]; e += 64) {
    const bool same = raw[icol[e]] == k;
    inn += same;
    exin += !same;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    ext += __shfl_xor(ext, o);
    inn += __shfl_xor(inn, o);
    exin += __shfl_xor(exin, o);
  }
  if (lane == 0) {
    ext_out[v] = ext;
    ext_in[v] = exin;
    if (inn) atomicAdd(&inner[k], inn);
  }
}

__global__ void k_raw_counts(const uint32_t* list, uint32_t L, const uint32_t* raw, const uint32_t* ext_out,
                             const uint32_t* ext_in, uint32_t* size, uint32_t* ns, uint32_t* nb) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L) return;
  uint32_t v = list[i], k = raw[v];
  atomicAdd(&size[k], 1u);
  if (ext_in[v]) atomicAdd(&ns[k], 1u);
  if (ext_out[v]) atomicAdd(&nb[k], 1u);
}

// benefit[0] of get_init_supernode_by_clusters, computed in double and compared
// as float like the reference's std::vector<float> benefit (compressor_base.h:3463-3481)
__global__ void k_accept(uint32_t K, const uint32_t* size, const uint32_t* inner, const uint32_t* ns,
                         const uint32_t* nb, uint32_t min_node, uint32_t max_node, float threshold, uint8_t* acc) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  float benefit = (float)((double)inner[k] - (double)ns[k] * (double)nb[k]);
  acc[k] = size[k] >= min_node && size[k] <= max_node && benefit > threshold;
}

__global__ void k_acc_u32(uint32_t K, const uint8_t* acc, uint32_t* out) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k <= K) out[k] = k < K ? acc[k] : 0u;
}

// accepted cluster id of every vertex (fixed after the first build)
__global__ void k_assign_cid(uint32_t n, const uint32_t* raw, const uint8_t* acc, const uint32_t* newid,
                             uint32_t* cid, uint8_t* ntype) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  uint32_t k = raw[v];
  uint32_t id = (k != INF32 && acc[k]) ? newid[k] : INF32;
  cid[v] = id;
  ntype[v] = 0;  // SingleNode until typed below
}

// ---------------------------------------------------------------------------
// Member order.  The first build sorts the accepted members by (cluster, id)
// once; that base order, the cluster offsets cl_off and every member's base slot
// stay fixed, because membership never changes after the first build.  Every
// layout then partitions each cluster stably into its exits followed by the
// other members (the shortcut rows of SSSP hold the exit slots only), and lays
// out the entries and the shortcut-row bases from one scan over the clusters.
// ---------------------------------------------------------------------------
struct ClBasePlus {
  __host__ __device__ ClBase operator()(const ClBase& a, const ClBase& b) const {
    return ClBase{a.sc + b.sc, a.ss + b.ss, a.ent + b.ent, 0u};
  }
};

__global__ void k_base_keys(const uint32_t* list, uint32_t L, const uint32_t* cid, unsigned long long* keys) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L) return;
  const uint32_t v = list[i], k = cid[v];
  keys[i] = k == INF32 ? ~0ull : ((unsigned long long)k << 32) | v;
}

__global__ void k_base_order(const unsigned long long* keys, uint32_t M, uint32_t* base_mem, uint32_t* bslot) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  const uint32_t v = (uint32_t)keys[i];
  base_mem[i] = v;
  bslot[v] = i;
}

template <class K>
__global__ void k_row_ptr32(const K* rowkey, uint32_t M, uint32_t K_, uint32_t* off, int shift) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > M) return;
  int64_t row = (i == M) ? (int64_t)K_ : (int64_t)(rowkey[i] >> shift);
  int64_t prev = (i == 0) ? -1 : (int64_t)(rowkey[i - 1] >> shift);
  for (int64_t r = prev + 1; r <= row; ++r) off[r] = i;
}

// Clusters of at most kWaveCluster members are laid out one wave each, larger ones
// one workgroup each (from the list of large clusters).
constexpr uint32_t kWaveCluster = 64;
constexpr int kLargeRegs = kMaxCluster / kBlock;  // member slots one thread of a large cluster holds

// One wave per cluster of at most kWaveCluster members touched since the last
// layout (untouched clusters keep their types, order and counts): node types from
// the edge counts (BothOutIn / OnlyIn / OnlyOut / Inner), the stable exits-first
// member order with an entry flag per slot, local indices and the per-cluster
// counts for the scan.  The members stay in registers between the two passes.
// A touched cluster whose benefit on the current graph, the acceptance rule of the
// first build (internal edges - entries * exits > compress_threshold, compared as
// float), no longer holds is dissolved: from this layout on its members are
// Singles (no entries, no exits, no shortcut rows), and it stays dissolved.  The
// internal edges of cluster k are the keys [cl_off[k] << 32, cl_off[k + 1] << 32)
// of the sorted internal edge list.
struct Dissolve {
  const unsigned long long* key;  // internal edge list, sorted
  uint64_t n;
  float threshold;
  uint8_t* dead;                  // per cluster: dissolved
};

__device__ __forceinline__ uint64_t key_lower_bound(const unsigned long long* __restrict__ a, uint64_t n,
                                                    unsigned long long x) {
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (a[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ bool dissolves(const Dissolve& dz, uint32_t k, uint32_t b, uint32_t size, uint32_t nen,
                                          uint32_t nex) {
  if (dz.dead[k]) return true;
  const uint64_t inner = key_lower_bound(dz.key, dz.n, (unsigned long long)(b + size) << 32) -
                         key_lower_bound(dz.key, dz.n, (unsigned long long)b << 32);
  const float benefit = (float)((double)inner - (double)nen * (double)nex);
  return !(benefit > dz.threshold);
}

__global__ void __launch_bounds__(kBlock) k_member_partition(uint32_t K, const uint32_t* __restrict__ cl_off,
                                                             const uint32_t* __restrict__ base_mem,
                                                             const uint32_t* __restrict__ ext_out,
                                                             const uint32_t* __restrict__ ext_in,
                                                             const uint8_t* __restrict__ touched, uint8_t* ntype,
                                                             uint32_t* cl_mem, uint8_t* slot_ent, uint32_t* lidx,
                                                             uint32_t* cl_nexit, ClBase* cl_cnt, Dissolve dz) {
  const int lane = lane_id();
  const unsigned long long below = (1ull << lane) - 1ull;
  const uint32_t nw = gridDim.x * kWaves;
  for (uint32_t k = blockIdx.x * kWaves + (threadIdx.x >> 6); k < K; k += nw) {
    if (!touched[k]) continue;  // whole wave
    const uint32_t b = cl_off[k], size = cl_off[k + 1] - b;
    if (size > kWaveCluster) continue;  // k_member_partition_large
    if (dz.dead[k]) {  // whole wave: members stay Singles in base order
      for (uint32_t j = lane; j < size; j += 64) {
        const uint32_t v = base_mem[b + j];
        ntype[v] = 0;
        cl_mem[b + j] = v;
        slot_ent[b + j] = 0;
        lidx[v] = j;
      }
      if (lane == 0) {
        cl_nexit[k] = 0;
        cl_cnt[k] = ClBase{0ull, 0ull, 0u, 0u};
      }
      continue;
    }
    uint32_t nex = 0, nen = 0;
    uint32_t v0 = 0;
    bool en0 = false, ex0 = false;
    for (uint32_t j0 = 0; j0 < size; j0 += 64) {
      const uint32_t j = j0 + lane;
      bool en = false, ex = false;
      uint32_t v = 0;
      if (j < size) {
        v = base_mem[b + j];
        en = ext_in[v] != 0;
        ex = ext_out[v] != 0;
        const uint8_t t = en && ex ? 3 : en ? 2 : ex ? 1 : 4;
        if (ntype[v] != t) ntype[v] = t;
      }
      if (j0 == 0) {
        v0 = v;
        en0 = en;
        ex0 = ex;
      }
      const unsigned long long mex = __ballot(ex), men = __ballot(en);
      nex += __popcll(mex);
      nen += __popcll(men);
    }
    if (dissolves(dz, k, b, size, nen, nex)) {  // wave-uniform: every lane has the same counts
      for (uint32_t j = lane; j < size; j += 64) {
        const uint32_t v = base_mem[b + j];
        ntype[v] = 0;
        cl_mem[b + j] = v;
        slot_ent[b + j] = 0;
        lidx[v] = j;
      }
      if (lane == 0) {
        dz.dead[k] = 1;
        cl_nexit[k] = 0;
        cl_cnt[k] = ClBase{0ull, 0ull, 0u, 0u};
      }
      continue;
    }
    uint32_t pe = 0, pi = nex;
    for (uint32_t j0 = 0; j0 < size; j0 += 64) {
      const uint32_t j = j0 + lane;
      const bool in = j < size;
      uint32_t v = v0;
      bool ex = ex0, en = en0;
      if (in && j0 != 0) {
        v = base_mem[b + j];
        ex = ext_out[v] != 0;
        en = ext_in[v] != 0;
      }
      const unsigned long long mex = __ballot(in && ex), mrest = __ballot(in && !ex);
      if (in) {
        const uint32_t pos = ex ? pe + __popcll(mex & below) : pi + __popcll(mrest & below);
        cl_mem[b + pos] = v;
        slot_ent[b + pos] = en;
        lidx[v] = pos;
      }
      pe += __popcll(mex);
      pi += __popcll(mrest);
    }
    if (lane == 0) {
      cl_nexit[k] = nex;
      cl_cnt[k] = ClBase{(unsigned long long)nen * size, (unsigned long long)nen * nex, nen, 0u};
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) cl_cnt[K] = ClBase{0ull, 0ull, 0u, 0u};
}

// k_member_partition for the clusters of more than kWaveCluster members, one
// workgroup each: slot r * kBlock + t sits in register r of thread t, every load
// of a pass is issued before any of its results is used (a 1024-member cluster
// costs a few memory round trips, not a few per 64 members), and the stable
// exits-first order comes from workgroup scans.
__global__ void __launch_bounds__(kBlock) k_member_partition_large(const uint32_t* __restrict__ large, uint32_t L,
                                                                   const uint32_t* __restrict__ cl_off,
                                                                   const uint32_t* __restrict__ base_mem,
                                                                   const uint32_t* __restrict__ ext_out,
                                                                   const uint32_t* __restrict__ ext_in,
                                                                   const uint8_t* __restrict__ touched, uint8_t* ntype,
                                                                   uint32_t* cl_mem, uint8_t* slot_ent, uint32_t* lidx,
                                                                   uint32_t* cl_nexit, ClBase* cl_cnt, Dissolve dz) {
  __shared__ uint32_t s_w0[kBlock / 64 + 1], s_w1[kBlock / 64 + 1];
  const uint32_t tid = threadIdx.x;
  for (uint32_t li = blockIdx.x; li < L; li += gridDim.x) {
    const uint32_t k = large[li];
    if (!touched[k]) continue;  // whole workgroup
    const uint32_t b = cl_off[k], size = cl_off[k + 1] - b;
    const uint32_t nr = (size + kBlock - 1) / kBlock;  // <= kLargeRegs: at most kMaxCluster members
    // threads past the end re-read the last member and are masked out below
    uint32_t v[kLargeRegs], ein[kLargeRegs], eout[kLargeRegs], old[kLargeRegs];
#pragma unroll
    for (int r = 0; r < kLargeRegs; ++r)
      if ((uint32_t)r < nr) v[r] = base_mem[b + min(r * kBlock + tid, size - 1)];
#pragma unroll
    for (int r = 0; r < kLargeRegs; ++r)
      if ((uint32_t)r < nr) {
        ein[r] = ext_in[v[r]];
        eout[r] = ext_out[v[r]];
        old[r] = ntype[v[r]];
      }
    uint32_t enb = 0, exb = 0, myex = 0, myen = 0;  // enb, exb: bit r = slot r * kBlock + tid
#pragma unroll
    for (int r = 0; r < kLargeRegs; ++r)
      if ((uint32_t)r < nr) {
        const bool in = r * kBlock + tid < size;
        const bool en = in && ein[r] != 0, ex = in && eout[r] != 0;
        if (in) {
          const uint32_t t = en && ex ? 3 : en ? 2 : ex ? 1 : 4;
          if (old[r] != t) ntype[v[r]] = (uint8_t)t;
        }
        enb |= (en ? 1u : 0u) << r;
        exb |= (ex ? 1u : 0u) << r;
        myex += ex ? 1u : 0u;
        myen += en ? 1u : 0u;
      }
    uint32_t nex, nen;
    block_exclusive_scan<kBlock>(myex, s_w0, &nex);
    block_exclusive_scan<kBlock>(myen, s_w1, &nen);
    __syncthreads();  // the scans below reuse s_w0 and s_w1
    if (dissolves(dz, k, b, size, nen, nex)) {  // workgroup-uniform: every thread has the same counts
      for (uint32_t j = tid; j < size; j += kBlock) {
        const uint32_t u = base_mem[b + j];
        ntype[u] = 0;
        cl_mem[b + j] = u;
        slot_ent[b + j] = 0;
        lidx[u] = j;
      }
      __syncthreads();  // every thread has read dead[k] before it is set
      if (tid == 0) {
        dz.dead[k] = 1;
        cl_nexit[k] = 0;
        cl_cnt[k] = ClBase{0ull, 0ull, 0u, 0u};
      }
      continue;
    }
    uint32_t pe = 0, pi = nex;
#pragma unroll
    for (int r = 0; r < kLargeRegs; ++r)
      if ((uint32_t)r < nr) {
        const bool in = r * kBlock + tid < size;
        const bool ex = (exb >> r) & 1u, en = (enb >> r) & 1u;
        uint32_t te, ti;
        const uint32_t oe = block_exclusive_scan<kBlock>(ex ? 1u : 0u, s_w0, &te);
        const uint32_t oi = block_exclusive_scan<kBlock>(in && !ex ? 1u : 0u, s_w1, &ti);
        if (in) {
          const uint32_t pos = ex ? pe + oe : pi + oi;
          cl_mem[b + pos] = v[r];
          slot_ent[b + pos] = en;
          lidx[v[r]] = pos;
        }
        pe += te;
        pi += ti;
        __syncthreads();
      }
    if (tid == 0) {
      cl_nexit[k] = nex;
      cl_cnt[k] = ClBase{(unsigned long long)nen * size, (unsigned long long)nen * nex, nen, 0u};
    }
  }
}

// Node-type histogram of the members (for the structure report only).
__global__ void k_type_hist(uint32_t M, const uint32_t* __restrict__ base_mem, const uint8_t* __restrict__ ntype,
                            uint32_t* hist) {
  __shared__ uint32_t s_h[8];
  if (threadIdx.x < 8) s_h[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < M) atomicAdd(&s_h[ntype[base_mem[i]] & 7], 1u);
  __syncthreads();
  if (threadIdx.x < 8 && s_h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], s_h[threadIdx.x]);
}

// One wave per cluster of at most kWaveCluster members: entries in member order,
// their shortcut-row bases and ent_of (INF32 for the other members of touched
// clusters; untouched clusters have the same non-entries as before).  Clears the
// touched flags.
__global__ void __launch_bounds__(kBlock) k_entry_layout(uint32_t K, const uint32_t* __restrict__ cl_off,
                                                         const uint32_t* __restrict__ cl_mem,
                                                         const uint32_t* __restrict__ cl_nexit,
                                                         const uint8_t* __restrict__ slot_ent, uint8_t* touched,
                                                         const ClBase* __restrict__ base, uint32_t* cl_eoff,
                                                         uint32_t* cl_ent, uint32_t* ent_of,
                                                         unsigned long long* sc_base, unsigned long long* ss_base) {
  const int lane = lane_id();
  const unsigned long long below = (1ull << lane) - 1ull;
  const uint32_t nw = gridDim.x * kWaves;
  for (uint32_t k = blockIdx.x * kWaves + (threadIdx.x >> 6); k < K; k += nw) {
    const uint32_t b = cl_off[k], size = cl_off[k + 1] - b, nex = cl_nexit[k];
    if (size > kWaveCluster) continue;  // k_entry_layout_large
    const bool tk = touched[k] != 0;
    const ClBase cb = base[k];
    uint32_t r = 0;
    for (uint32_t j0 = 0; j0 < size; j0 += 64) {
      const uint32_t j = j0 + lane;
      const bool en = j < size && slot_ent[b + j] != 0;
      const unsigned long long men = __ballot(en);
      if (en) {
        const uint32_t v = cl_mem[b + j];
        const uint32_t q = r + __popcll(men & below), e = cb.ent + q;
        cl_ent[e] = v;
        ent_of[v] = e;
        sc_base[e] = cb.sc + (unsigned long long)q * size;
        ss_base[e] = cb.ss + (unsigned long long)q * nex;
      } else if (tk && j < size) {
        ent_of[cl_mem[b + j]] = INF32;
      }
      r += __popcll(men);
    }
    if (lane == 0) {
      cl_eoff[k] = cb.ent;
      if (tk) touched[k] = 0;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const ClBase t = base[K];
    cl_eoff[K] = t.ent;
    sc_base[t.ent] = t.sc;
    ss_base[t.ent] = t.ss;
  }
}

// k_entry_layout for the clusters of more than kWaveCluster members, one
// workgroup each, the slots in registers as in k_member_partition_large.
__global__ void __launch_bounds__(kBlock) k_entry_layout_large(const uint32_t* __restrict__ large, uint32_t L,
                                                               const uint32_t* __restrict__ cl_off,
                                                               const uint32_t* __restrict__ cl_mem,
                                                               const uint32_t* __restrict__ cl_nexit,
                                                               const uint8_t* __restrict__ slot_ent, uint8_t* touched,
                                                               const ClBase* __restrict__ base, uint32_t* cl_eoff,
                                                               uint32_t* cl_ent, uint32_t* ent_of,
                                                               unsigned long long* sc_base, unsigned long long* ss_base) {
  __shared__ uint32_t s_w[kBlock / 64 + 1];
  const uint32_t tid = threadIdx.x;
  for (uint32_t li = blockIdx.x; li < L; li += gridDim.x) {
    const uint32_t k = large[li];
    const uint32_t b = cl_off[k], size = cl_off[k + 1] - b, nex = cl_nexit[k];
    const bool tk = touched[k] != 0;
    const ClBase cb = base[k];
    const uint32_t nr = (size + kBlock - 1) / kBlock;  // <= kLargeRegs
    uint32_t se[kLargeRegs], mv[kLargeRegs];
#pragma unroll
    for (int r = 0; r < kLargeRegs; ++r)
      if ((uint32_t)r < nr) {
        const uint32_t j = min(r * kBlock + tid, size - 1);
        se[r] = slot_ent[b + j];
        mv[r] = cl_mem[b + j];
      }
    uint32_t q0 = 0;
#pragma unroll
    for (int r = 0; r < kLargeRegs; ++r)
      if ((uint32_t)r < nr) {
        const bool in = r * kBlock + tid < size;
        const bool en = in && se[r] != 0;
        uint32_t tot;
        const uint32_t o = block_exclusive_scan<kBlock>(en ? 1u : 0u, s_w, &tot);
        if (en) {
          const uint32_t v = mv[r], q = q0 + o, e = cb.ent + q;
          cl_ent[e] = v;
          ent_of[v] = e;
          sc_base[e] = cb.sc + (unsigned long long)q * size;
          ss_base[e] = cb.ss + (unsigned long long)q * nex;
        } else if (tk && in) {
          ent_of[mv[r]] = INF32;
        }
        q0 += tot;
        __syncthreads();  // the next scan reuses s_w
      }
    if (tid == 0) {
      cl_eoff[k] = cb.ent;
      if (tk) touched[k] = 0;
    }
  }
}

__global__ void k_large_list(uint32_t K, const uint32_t* __restrict__ cl_off, uint32_t* list,
                             unsigned long long* count) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  const bool large = k < K && cl_off[k + 1] - cl_off[k] > kWaveCluster;
  const unsigned long long slot = wave_append(large, count);
  if (large) list[slot] = k;
}

// First build: the internal in-edges of every accepted member, one wave per
// member, appended as keys (base slot of dst << 32 | src) with their weights.
__global__ void __launch_bounds__(kBlock) k_internal_collect(const uint32_t* __restrict__ list, uint32_t L,
                                                             const uint32_t* __restrict__ cid,
                                                             const uint32_t* __restrict__ bslot,
                                                             const uint32_t* __restrict__ ioff,
                                                             const uint32_t* __restrict__ icol,
                                                             const uint32_t* __restrict__ iw,
                                                             unsigned long long* key, uint32_t* wout,
                                                             unsigned long long* count) {
  const uint32_t i = blockIdx.x * kWaves + (threadIdx.x >> 6);
  if (i >= L) return;  // whole wave
  const int lane = lane_id();
  const uint32_t v = list[i], k = cid[v];
  if (k == INF32) return;  // whole wave
  const unsigned long long hi = (unsigned long long)bslot[v] << 32;
  const uint32_t b = ioff[v], e = ioff[v + 1];
  for (uint32_t e0 = b; e0 < e; e0 += 64) {
    const uint32_t x = e0 + lane;
    uint32_t src = 0;
    bool in = false;
    if (x < e) {
      src = icol[x];
      in = cid[src] == k;
    }
    const unsigned long long slot = wave_append(in, count);
    if (in) {
      key[slot] = hi | src;
      wout[slot] = iw ? iw[x] : 1u;
    }
  }
}

// The internal edge list is sorted by (base slot of dst, src), hence grouped by
// cluster: the in-cluster arrays are a direct map of it, ie_off per cluster.
__global__ void k_internal_fill(uint64_t N, uint32_t K, const unsigned long long* __restrict__ key,
                                const uint32_t* __restrict__ w, const uint32_t* __restrict__ base_mem,
                                const uint32_t* __restrict__ cid, const uint32_t* __restrict__ lidx,
                                const uint32_t* __restrict__ ooff, uint32_t* ie_off, uint32_t* ie_src,
                                uint32_t* ie_dst, uint32_t* ie_deg, uint32_t* ie_wt) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p > N) return;
  const int64_t row = p == N ? (int64_t)K : (int64_t)cid[base_mem[key[p] >> 32]];
  const int64_t prev = p == 0 ? -1 : (int64_t)cid[base_mem[key[p - 1] >> 32]];
  for (int64_t r = prev + 1; r <= row; ++r) ie_off[r] = (uint32_t)p;
  if (p == N) return;
  const uint32_t src = (uint32_t)key[p], dst = base_mem[key[p] >> 32];
  ie_src[p] = lidx[src];
  ie_dst[p] = lidx[dst];
  ie_deg[p] = ooff[src + 1] - ooff[src];
  ie_wt[p] = w[p];
}

// Moves the shortcut rows of the clusters the batches did not touch from the
// previous layout.  A cluster's rows are contiguous in both layouts and keep
// their length (its members' entry/exit status can only change through an edge
// with an endpoint in it).
template <class T, bool kSsspRows>
__global__ void k_rows_move(uint64_t len, uint32_t K, const ClBase* __restrict__ nb, const ClBase* __restrict__ ob,
                            const uint8_t* __restrict__ valid, const T* __restrict__ oldA, T* __restrict__ newA,
                            const T* __restrict__ oldB, T* __restrict__ newB) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < len; p += stride) {
    uint32_t lo = 0, hi = K;  // last cluster whose base is <= p (nb[K] = len > p)
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) >> 1;
      const unsigned long long bm = kSsspRows ? nb[mid].ss : nb[mid].sc;
      if (bm <= p) lo = mid; else hi = mid - 1;
    }
    if (!valid[lo]) continue;
    const unsigned long long nbase = kSsspRows ? nb[lo].ss : nb[lo].sc;
    const unsigned long long obase = kSsspRows ? ob[lo].ss : ob[lo].sc;
    const unsigned long long q = obase + (p - nbase);
    newA[p] = oldA[q];
    if (newB) newB[p] = oldB[q];
  }
}

// Grows a buffer keeping its first `keep` bytes.
static void grow_keep(suminc_ctx* c, DevBuf& b, size_t bytes, size_t keep) {
  if (bytes <= b.cap && b.p) return;
  DevBuf t;
  t.ensure(bytes + bytes / 16);
  if (keep && b.p) SUMINC_HIP(hipMemcpyAsync(t.p, b.p, keep, hipMemcpyDeviceToDevice, c->stream));
  SUMINC_HIP(hipStreamSynchronize(c->stream));
  b.swap(t);
}

static uint32_t layout_grid(uint32_t K) { return std::max<uint32_t>(1u, (K + kWaves - 1) / kWaves); }

static uint32_t move_grid(uint64_t len) {
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(4096, (len + kBlock - 1) / kBlock));
}

// Layout of the current structure from the per-member edge counts and the
// internal edge list: node types, member order, entries and their shortcut-row
// bases, the in-cluster edge arrays, and the rows of untouched clusters moved
// over from the previous layout.
static void sketch_layout(suminc_ctx* c, bool rebuild) {
  Sketch& s = c->sk;
  const uint32_t n = c->n, K = s.K, M = s.M;
  if (rebuild) s.lay_base.swap(s.lay_base_old);  // the previous bases locate the rows to move
  s.lay_hist.ensure(32);
  s.lay_cnt.ensure(((size_t)K + 1) * sizeof(ClBase));
  s.lay_base.ensure(((size_t)K + 1) * sizeof(ClBase));
  s.cl_mem.ensure((size_t)M * 4 + 4); s.slot_ent.ensure((size_t)M + 1); s.cl_nexit.ensure((size_t)K * 4 + 4);
  s.cl_eoff.ensure(((size_t)K + 1) * 4);
  s.cl_ent.ensure((size_t)M * 4 + 4);
  s.sc_base.ensure(((size_t)M + 1) * 8); s.ss_base.ensure(((size_t)M + 1) * 8);
  if (!rebuild || !s.ent_of.p) {
    s.ent_of.ensure((size_t)n * 4);
    k_fill32<<<grid_for(n), kBlock, 0, c->stream>>>(s.ent_of.as<uint32_t>(), n, INF32);
  }
  ClBase* cnt = s.lay_cnt.as<ClBase>();
  ClBase* base = s.lay_base.as<ClBase>();
  if (!rebuild) {  // every cluster is laid out by the first layout
    s.touched.ensure((size_t)K + 1);
    SUMINC_HIP(hipMemsetAsync(s.touched.p, 1, (size_t)K + 1, c->stream));
    // the clusters laid out one workgroup each (membership is fixed from here on)
    s.large.ensure((size_t)K * 4 + 4);
    DevBuf lc;
    lc.ensure(8);
    SUMINC_HIP(hipMemsetAsync(lc.p, 0, 8, c->stream));
    if (K)
      k_large_list<<<grid_for(K), kBlock, 0, c->stream>>>(K, s.cl_off.as<uint32_t>(), s.large.as<uint32_t>(),
                                                          lc.as<unsigned long long>());
    unsigned long long nl = 0;
    SUMINC_HIP(hipMemcpyAsync(&nl, lc.p, 8, hipMemcpyDeviceToHost, c->stream));
    SUMINC_HIP(hipStreamSynchronize(c->stream));
    s.n_large = (uint32_t)nl;
  }
  const unsigned large_grid = std::max<uint32_t>(1u, std::min<uint32_t>(s.n_large, 4096u));
  // the sorted internal edge list (merged after every batch) gives each cluster's
  // internal edge count for the dissolution test
  s.dead.ensure((size_t)K + 1);
  if (!rebuild) SUMINC_HIP(hipMemsetAsync(s.dead.p, 0, (size_t)K + 1, c->stream));
  const Dissolve dz{s.iem_key.as<unsigned long long>(), s.iem_n, (float)s.threshold, s.dead.as<uint8_t>()};
  k_member_partition<<<layout_grid(K), kBlock, 0, c->stream>>>(
      K, s.cl_off.as<uint32_t>(), s.base_mem.as<uint32_t>(), s.ext_out.as<uint32_t>(), s.ext_in.as<uint32_t>(),
      s.touched.as<uint8_t>(), s.ntype.as<uint8_t>(), s.cl_mem.as<uint32_t>(), s.slot_ent.as<uint8_t>(),
      s.lidx.as<uint32_t>(), s.cl_nexit.as<uint32_t>(), cnt, dz);
  if (s.n_large)
    k_member_partition_large<<<large_grid, kBlock, 0, c->stream>>>(
        s.large.as<uint32_t>(), s.n_large, s.cl_off.as<uint32_t>(), s.base_mem.as<uint32_t>(),
        s.ext_out.as<uint32_t>(), s.ext_in.as<uint32_t>(), s.touched.as<uint8_t>(), s.ntype.as<uint8_t>(),
        s.cl_mem.as<uint32_t>(), s.slot_ent.as<uint8_t>(), s.lidx.as<uint32_t>(), s.cl_nexit.as<uint32_t>(), cnt, dz);
  cub(c, [&](void* p, size_t& b) {
    return rocprim::exclusive_scan(p, b, cnt, base, ClBase{0ull, 0ull, 0u, 0u}, (size_t)K + 1, ClBasePlus(),
                                   c->stream);
  });
  // the totals (entries, row lengths) size the row buffers below; the host reads
  // them while the entry layout and the in-cluster arrays are built
  check_begin(c, base + K, sizeof(ClBase));
  k_entry_layout<<<layout_grid(K), kBlock, 0, c->stream>>>(
      K, s.cl_off.as<uint32_t>(), s.cl_mem.as<uint32_t>(), s.cl_nexit.as<uint32_t>(), s.slot_ent.as<uint8_t>(),
      s.touched.as<uint8_t>(), base, s.cl_eoff.as<uint32_t>(), s.cl_ent.as<uint32_t>(), s.ent_of.as<uint32_t>(),
      s.sc_base.as<unsigned long long>(), s.ss_base.as<unsigned long long>());
  if (s.n_large)
    k_entry_layout_large<<<large_grid, kBlock, 0, c->stream>>>(
        s.large.as<uint32_t>(), s.n_large, s.cl_off.as<uint32_t>(), s.cl_mem.as<uint32_t>(),
        s.cl_nexit.as<uint32_t>(), s.slot_ent.as<uint8_t>(), s.touched.as<uint8_t>(), base, s.cl_eoff.as<uint32_t>(),
        s.cl_ent.as<uint32_t>(), s.ent_of.as<uint32_t>(), s.sc_base.as<unsigned long long>(),
        s.ss_base.as<unsigned long long>());
  // in-cluster edge arrays
  const uint64_t ie_m = s.iem_n;
  s.ie_m = ie_m;
  s.ie_off.ensure(((size_t)K + 1) * 4);
  for (DevBuf* b : {&s.ie_src, &s.ie_dst, &s.ie_deg, &s.ie_wt}) b->ensure_slack(ie_m * 4 + 4);
  k_internal_fill<<<grid_for(ie_m + 1), kBlock, 0, c->stream>>>(
      ie_m, K, s.iem_key.as<unsigned long long>(), s.iem_w.as<uint32_t>(), s.base_mem.as<uint32_t>(),
      s.cid.as<uint32_t>(), s.lidx.as<uint32_t>(), c->out_off.as<uint32_t>(), s.ie_off.as<uint32_t>(),
      s.ie_src.as<uint32_t>(), s.ie_dst.as<uint32_t>(), s.ie_deg.as<uint32_t>(), s.ie_wt.as<uint32_t>());
  SUMINC_HIP(hipGetLastError());
  ClBase tot;
  std::memcpy(&tot, check_wait(c), sizeof(ClBase));
  s.E = tot.ent;
  s.sc_len = tot.sc;
  s.ss_len = tot.ss;
  // shortcut rows: validity per cluster (PageRank, SSSP)
  s.pr_valid.ensure((size_t)K + 1); s.ss_valid.ensure((size_t)K + 1);
  if (!rebuild) {
    SUMINC_HIP(hipMemsetAsync(s.pr_valid.p, 0, (size_t)K + 1, c->stream));
    SUMINC_HIP(hipMemsetAsync(s.ss_valid.p, 0, (size_t)K + 1, c->stream));
    s.pr_rows_for = -1;
    s.ss_rows = false;
  } else {
    const ClBase* old = s.lay_base_old.as<ClBase>();
    if (s.pr_rows_for >= 0) {
      s.sc_R.swap(s.sc_R_old); s.sc_D.swap(s.sc_D_old);
      const size_t es = s.pr_rows_for == SUMINC_PR_FP64 ? 8 : 4;
      s.sc_R.ensure_slack(s.sc_len * es + es); s.sc_D.ensure_slack(s.sc_len * es + es);
      if (s.sc_len && es == 8)
        k_rows_move<double, false><<<move_grid(s.sc_len), kBlock, 0, c->stream>>>(
            s.sc_len, K, base, old, s.pr_valid.as<uint8_t>(), s.sc_R_old.as<double>(), s.sc_R.as<double>(),
            s.sc_D_old.as<double>(), s.sc_D.as<double>());
      else if (s.sc_len)
        k_rows_move<float, false><<<move_grid(s.sc_len), kBlock, 0, c->stream>>>(
            s.sc_len, K, base, old, s.pr_valid.as<uint8_t>(), s.sc_R_old.as<float>(), s.sc_R.as<float>(),
            s.sc_D_old.as<float>(), s.sc_D.as<float>());
    }
    if (s.ss_rows) {
      s.ss_dist.swap(s.ss_dist_old);
      s.ss_dist.ensure_slack(s.ss_len * 4 + 4);
      if (s.ss_len)
        k_rows_move<uint32_t, true><<<move_grid(s.ss_len), kBlock, 0, c->stream>>>(
            s.ss_len, K, base, old, s.ss_valid.as<uint8_t>(), s.ss_dist_old.as<uint32_t>(),
            s.ss_dist.as<uint32_t>(), nullptr, nullptr);
    }
    SUMINC_HIP(hipGetLastError());
  }
  s.pr_need = s.ss_need = true;
  s.version = c->graph_version;
}

// First build from the raw cluster assignment: scans the rows of the raw members,
// decides acceptance (kept across batches, as the reference keeps its supernodes)
// and collects the internal edge list; then lays the structure out.
void sketch_build(suminc_ctx* c, uint32_t min_node, uint32_t max_node, double threshold) {
  SUMINC_REQUIRE(c->has_graph, SUMINC_E_STATE, "no graph");
  Sketch& s = c->sk;
  SUMINC_REQUIRE(s.raw_cid.p, SUMINC_E_STATE, "no cluster assignment (call suminc_cluster_* first)");
  const uint32_t n = c->n;
  const uint32_t Kr = s.Kraw;
  DevBuf inner, size, ns, nb, acc32, newid, rflag, rcount, icount;
  // raw member list
  rflag.ensure(n + 1); rcount.ensure(8);
  s.rlist.ensure((size_t)n * 4 + 4);
  k_raw_flag<<<grid_for(n), kBlock, 0, c->stream>>>(n, s.raw_cid.as<uint32_t>(), rflag.as<uint8_t>());
  cub(c, [&](void* p, size_t& b) {
    return rocprim::select(p, b, rocprim::counting_iterator<uint32_t>(0), rflag.as<uint8_t>(), s.rlist.as<uint32_t>(),
                           rcount.as<unsigned>(), (size_t)n, c->stream);
  });
  unsigned rm = 0;
  SUMINC_HIP(hipMemcpyAsync(&rm, rcount.p, 4, hipMemcpyDeviceToHost, c->stream));
  SUMINC_HIP(hipStreamSynchronize(c->stream));
  s.RM = rm;
  const uint32_t RM = rm;
  s.ext_out.ensure((size_t)n * 4); s.ext_in.ensure((size_t)n * 4);
  inner.ensure((size_t)Kr * 4 + 4);
  SUMINC_HIP(hipMemsetAsync(inner.p, 0, (size_t)Kr * 4 + 4, c->stream));
  if (RM)
    k_member_scan<<<(RM + kWaves - 1) / kWaves, kBlock, 0, c->stream>>>(
        s.rlist.as<uint32_t>(), RM, s.raw_cid.as<uint32_t>(), c->out_off.as<uint32_t>(), c->out_col.as<uint32_t>(),
        c->in_off.as<uint32_t>(), c->in_col.as<uint32_t>(), s.ext_out.as<uint32_t>(), s.ext_in.as<uint32_t>(),
        inner.as<uint32_t>());
  size.ensure((size_t)Kr * 4 + 4); ns.ensure((size_t)Kr * 4 + 4); nb.ensure((size_t)Kr * 4 + 4);
  for (DevBuf* b : {&size, &ns, &nb}) SUMINC_HIP(hipMemsetAsync(b->p, 0, (size_t)Kr * 4 + 4, c->stream));
  if (RM)
    k_raw_counts<<<grid_for(RM), kBlock, 0, c->stream>>>(s.rlist.as<uint32_t>(), RM, s.raw_cid.as<uint32_t>(),
                                                         s.ext_out.as<uint32_t>(), s.ext_in.as<uint32_t>(),
                                                         size.as<uint32_t>(), ns.as<uint32_t>(), nb.as<uint32_t>());
  s.raw_acc.ensure(Kr + 1);
  if (Kr)
    k_accept<<<grid_for(Kr), kBlock, 0, c->stream>>>(Kr, size.as<uint32_t>(), inner.as<uint32_t>(), ns.as<uint32_t>(),
                                                     nb.as<uint32_t>(), min_node, max_node, (float)threshold,
                                                     s.raw_acc.as<uint8_t>());
  acc32.ensure(((size_t)Kr + 1) * 4); newid.ensure(((size_t)Kr + 1) * 4);
  k_acc_u32<<<grid_for((uint64_t)Kr + 1), kBlock, 0, c->stream>>>(Kr, s.raw_acc.as<uint8_t>(), acc32.as<uint32_t>());
  cub(c, [&](void* p, size_t& b) {
    return rocprim::exclusive_scan(p, b, acc32.as<uint32_t>(), newid.as<uint32_t>(), 0u, (size_t)Kr + 1,
                                   rocprim::plus<uint32_t>(), c->stream);
  });
  uint32_t K = 0;
  SUMINC_HIP(hipMemcpyAsync(&K, newid.as<uint32_t>() + Kr, 4, hipMemcpyDeviceToHost, c->stream));
  s.cid.ensure((size_t)n * 4); s.ntype.ensure(n); s.lidx.ensure((size_t)n * 4);
  k_assign_cid<<<grid_for(n), kBlock, 0, c->stream>>>(n, s.raw_cid.as<uint32_t>(), s.raw_acc.as<uint8_t>(),
                                                      newid.as<uint32_t>(), s.cid.as<uint32_t>(), s.ntype.as<uint8_t>());
  // internal edges of the accepted clusters, sorted by (src, dst)
  icount.ensure(8);
  SUMINC_HIP(hipMemsetAsync(icount.p, 0, 8, c->stream));
  SUMINC_HIP(hipStreamSynchronize(c->stream));
  s.K = K;
  s.threshold = threshold;
  {
    // the internal edges of the accepted clusters size the list
    std::vector<uint32_t> hin(Kr + 1), hsize(Kr + 1);
    std::vector<uint8_t> hacc(Kr + 1);
    SUMINC_HIP(hipMemcpy(hin.data(), inner.p, (size_t)Kr * 4, hipMemcpyDeviceToHost));
    SUMINC_HIP(hipMemcpy(hsize.data(), size.p, (size_t)Kr * 4, hipMemcpyDeviceToHost));
    SUMINC_HIP(hipMemcpy(hacc.data(), s.raw_acc.p, Kr, hipMemcpyDeviceToHost));
    uint64_t tot = 0, M = 0;
    uint32_t maxc = 0;
    for (uint32_t k = 0; k < Kr; ++k)
      if (hacc[k]) {
        tot += hin[k];
        M += hsize[k];
        maxc = std::max(maxc, hsize[k]);
      }
    SUMINC_REQUIRE(maxc <= (uint32_t)kMaxCluster, SUMINC_E_TOO_LARGE, "an accepted cluster has more than 2048 members");
    s.M = (uint32_t)M;
    s.max_cluster = maxc;
    // base member order (cluster, id), cluster offsets, base slots
    DevBuf keys_a, keys_b;
    keys_a.ensure((size_t)RM * 8 + 8); keys_b.ensure((size_t)RM * 8 + 8);
    s.base_mem.ensure(M * 4 + 4); s.bslot.ensure((size_t)n * 4); s.cl_off.ensure(((size_t)K + 1) * 4);
    if (RM) {
      k_base_keys<<<grid_for(RM), kBlock, 0, c->stream>>>(s.rlist.as<uint32_t>(), RM, s.cid.as<uint32_t>(),
                                                          keys_a.as<unsigned long long>());
      cub(c, [&](void* p, size_t& b) {
        return rocprim::radix_sort_keys(p, b, keys_a.as<uint64_t>(), keys_b.as<uint64_t>(), (size_t)RM, 0, 64,
                                        c->stream);
      });
    }
    if (M)
      k_base_order<<<grid_for(M), kBlock, 0, c->stream>>>(keys_b.as<unsigned long long>(), (uint32_t)M,
                                                          s.base_mem.as<uint32_t>(), s.bslot.as<uint32_t>());
    k_row_ptr32<unsigned long long><<<grid_for(M + 1), kBlock, 0, c->stream>>>(
        keys_b.as<unsigned long long>(), (uint32_t)M, K, s.cl_off.as<uint32_t>(), 32);
    // internal edges, sorted by (base slot of dst, src)
    const size_t cap = tot + 4096;
    s.iem_key.ensure(cap * 8); s.iem_w.ensure(cap * 4); s.iem_key2.ensure(cap * 8); s.iem_w2.ensure(cap * 4);
    if (RM)
      k_internal_collect<<<(RM + kWaves - 1) / kWaves, kBlock, 0, c->stream>>>(
          s.rlist.as<uint32_t>(), RM, s.cid.as<uint32_t>(), s.bslot.as<uint32_t>(), c->in_off.as<uint32_t>(),
          c->in_col.as<uint32_t>(), c->wkind == SUMINC_W_U32 ? c->in_w.as<uint32_t>() : nullptr,
          s.iem_key2.as<unsigned long long>(), s.iem_w2.as<uint32_t>(), icount.as<unsigned long long>());
    SUMINC_HIP(hipGetLastError());
    unsigned long long got = 0;
    SUMINC_HIP(hipMemcpyAsync(&got, icount.p, 8, hipMemcpyDeviceToHost, c->stream));
    SUMINC_HIP(hipStreamSynchronize(c->stream));
    SUMINC_REQUIRE(got == tot, SUMINC_E_STATE, "internal edge count mismatch");
    if (got)
      cub(c, [&](void* p, size_t& b) {
        return rocprim::radix_sort_pairs(p, b, s.iem_key2.as<uint64_t>(), s.iem_key.as<uint64_t>(),
                                         s.iem_w2.as<uint32_t>(), s.iem_w.as<uint32_t>(), (size_t)got, 0, 64,
                                         c->stream);
      });
    s.iem_n = got;
    s.iem_dead.ensure(cap);
    SUMINC_HIP(hipMemsetAsync(s.iem_dead.p, 0, cap, c->stream));
  }
  sketch_layout(c, false);
  // every cluster needs a precompute after the first build of the PageRank side
  s.dirty.ensure((size_t)K + 1);
  SUMINC_HIP(hipMemsetAsync(s.dirty.p, 1, (size_t)K + 1, c->stream));
  s.built = true;
  SUMINC_HIP(hipStreamSynchronize(c->stream));
}

// ---------------------------------------------------------------------------
// batches (the reference's inc_compress_mirror bookkeeping, compressor_base.h:
// 5634-5852, and inc_judge_out_bound_node, :3958-3990): every group of a batch
// marks the clusters of its endpoints dirty (precompute) and their shortcut rows
// stale, updates the edge counts that decide entry/exit status, and updates the
// internal edge list.  The layout is redone on the next query.
// ---------------------------------------------------------------------------
__global__ void k_sk_batch(const uint32_t* __restrict__ row, const uint32_t* __restrict__ col,
                           const uint8_t* __restrict__ act, const uint8_t* __restrict__ ex,
                           const uint32_t* __restrict__ gw, const unsigned* countp, uint32_t N, int wu32,
                           const uint32_t* __restrict__ cid, uint8_t* dirty, uint8_t* pr_valid, uint8_t* ss_valid,
                           uint8_t* touched,
                           uint32_t* ext_out, uint32_t* ext_in, const uint32_t* __restrict__ bslot,
                           const unsigned long long* __restrict__ iem_key,
                           uint64_t iem_n, uint8_t* iem_dead, unsigned long long* add_key, uint32_t* add_w,
                           unsigned long long* add_cnt, unsigned long long* delta) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  bool app = false;
  unsigned long long key = 0;
  uint32_t w = 1;
  if (g < N && g < *countp) {
    const uint32_t u = row[g], v = col[g];
    const uint32_t cu = cid[u], cv = cid[v];
    if (cu != INF32) { dirty[cu] = 1; pr_valid[cu] = 0; ss_valid[cu] = 0; touched[cu] = 1; }
    if (cv != INF32) { dirty[cv] = 1; pr_valid[cv] = 0; ss_valid[cv] = 0; touched[cv] = 1; }
    const bool existed = ex[g], add = act[g] == 0;
    const bool removed = existed && !add, added = !existed && add;
    if ((removed || added) && cu != cv) {
      if (cu != INF32) atomicAdd(&ext_out[u], added ? 1u : 0xFFFFFFFFu);
      if (cv != INF32) atomicAdd(&ext_in[v], added ? 1u : 0xFFFFFFFFu);
    }
    if (cu != INF32 && cu == cv) {
      key = ((unsigned long long)bslot[v] << 32) | u;
      if (existed) {  // deleted or re-weighted: the old entry goes
        uint64_t lo = 0, hi = iem_n;
        while (lo < hi) {
          const uint64_t mid = (lo + hi) >> 1;
          if (iem_key[mid] < key) lo = mid + 1; else hi = mid;
        }
        if (lo < iem_n && iem_key[lo] == key) {
          iem_dead[lo] = 1;
          atomicAdd(delta, 0xFFFFFFFFFFFFFFFFull);
        }
      }
      if (add) {
        app = true;
        w = wu32 ? gw[g] : 1u;
      }
    }
  }
  const unsigned long long slot = wave_append(app, add_cnt);
  if (app) {
    add_key[slot] = key;
    add_w[slot] = w;
    atomicAdd(delta, 1ull);
  }
}

__global__ void k_sk_batch_prep(unsigned long long* add_cnt, unsigned long long* delta) {
  if (threadIdx.x == 0) {
    *add_cnt = 0;
    *delta = 0;
  }
}

void sketch_batch(suminc_ctx* c, const uint32_t* rows, const uint32_t* cols, const uint8_t* act, const uint8_t* ex,
                  const uint32_t* w, const unsigned* countp, uint32_t N) {
  Sketch& s = c->sk;
  s.batch_pending = false;
  if (!s.built || !N) return;
  const uint64_t n0 = s.iem_n, tot = n0 + N;
  grow_keep(c, s.iem_key, tot * 8 + 8, n0 * 8);
  grow_keep(c, s.iem_w, tot * 4 + 4, n0 * 4);
  grow_keep(c, s.iem_dead, tot + 1, n0);
  s.iem_key2.ensure_slack(tot * 8 + 8);
  s.iem_w2.ensure_slack(tot * 4 + 4);
  Counters* cnt = dev_counters(c);
  unsigned long long* add_cnt = &cnt->aux[7];
  unsigned long long* delta = &cnt->aux[8];
  k_sk_batch_prep<<<1, 64, 0, c->stream>>>(add_cnt, delta);
  k_sk_batch<<<grid_for(N), kBlock, 0, c->stream>>>(
      rows, cols, act, ex, w, countp, N, c->wkind == SUMINC_W_U32, s.cid.as<uint32_t>(), s.dirty.as<uint8_t>(),
      s.pr_valid.as<uint8_t>(), s.ss_valid.as<uint8_t>(), s.touched.as<uint8_t>(), s.ext_out.as<uint32_t>(),
      s.ext_in.as<uint32_t>(), s.bslot.as<uint32_t>(), s.iem_key.as<unsigned long long>(), n0, s.iem_dead.as<uint8_t>(),
      s.iem_key.as<unsigned long long>() + n0, s.iem_w.as<uint32_t>() + n0, add_cnt, delta);
  SUMINC_HIP(hipGetLastError());
  s.batch_pending = true;
  s.pr_need = s.ss_need = true;
}

// Sorts up to kSmallSort (key, weight) pairs in one workgroup (bitonic network in LDS).
constexpr uint32_t kSmallSort = 2048;
__global__ void __launch_bounds__(1024) k_sort_small(const unsigned long long* __restrict__ key,
                                                     const uint32_t* __restrict__ w, uint32_t n,
                                                     unsigned long long* okey, uint32_t* ow) {
  __shared__ unsigned long long s_k[kSmallSort];
  __shared__ uint32_t s_w[kSmallSort];
  uint32_t P = 64;
  while (P < n) P <<= 1;
  for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) {
    s_k[i] = i < n ? key[i] : ~0ull;
    s_w[i] = i < n ? w[i] : 0u;
  }
  __syncthreads();
  for (uint32_t k = 2; k <= P; k <<= 1)
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) {
        const uint32_t l = i ^ j;
        if (l > i) {
          const unsigned long long x = s_k[i], y = s_k[l];
          if ((i & k) == 0 ? x > y : x < y) {
            s_k[i] = y;
            s_k[l] = x;
            const uint32_t t = s_w[i];
            s_w[i] = s_w[l];
            s_w[l] = t;
          }
        }
      }
      __syncthreads();
    }
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    okey[i] = s_k[i];
    ow[i] = s_w[i];
  }
}

struct LiveOne {
  __host__ __device__ uint32_t operator()(uint8_t dead) const { return dead ? 0u : 1u; }
};

// Merges the sorted list without its dead entries with the sorted adds: an old
// entry lands after the live entries and the adds below it, an add after the
// adds and the live entries below it (keys are unique among live entries and adds).
__global__ void k_iem_merge(uint64_t n0, uint64_t A, uint64_t live, const unsigned long long* __restrict__ key,
                            const uint32_t* __restrict__ w, const uint8_t* __restrict__ dead,
                            const uint32_t* __restrict__ pre, const unsigned long long* __restrict__ akey,
                            const uint32_t* __restrict__ aw, unsigned long long* okey, uint32_t* ow) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n0) {
    if (dead[i]) return;
    const unsigned long long x = key[i];
    uint64_t lo = 0, hi = A;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (akey[mid] < x) lo = mid + 1; else hi = mid;
    }
    const uint64_t q = pre[i] + lo;
    okey[q] = x;
    ow[q] = w[i];
  } else if (i < n0 + A) {
    const uint64_t j = i - n0;
    const unsigned long long x = akey[j];
    uint64_t lo = 0, hi = n0;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (key[mid] < x) lo = mid + 1; else hi = mid;
    }
    const uint64_t q = j + (lo < n0 ? pre[lo] : live);
    okey[q] = x;
    ow[q] = aw[j];
  }
}

// After the batch's final synchronisation (the add and dead counts are on the
// host): merges the batch's internal adds into the sorted internal edge list and
// drops its dead entries.
void sketch_batch_done(suminc_ctx* c) {
  Sketch& s = c->sk;
  if (!s.built || !s.batch_pending) return;
  s.batch_pending = false;
  const uint64_t n0 = s.iem_n;
  const uint64_t A = c->host_counters->aux[7];
  const long long delta = (long long)c->host_counters->aux[8];
  const uint64_t n1 = (uint64_t)((long long)n0 + delta);
  const uint64_t D = (uint64_t)((long long)A - delta);
  if (!A && !D) return;
  s.add_key.ensure_slack(A * 8 + 8); s.add_w.ensure_slack(A * 4 + 4);
  s.iem_pre.ensure_slack(n0 * 4 + 4);
  const unsigned long long* tail_k = s.iem_key.as<unsigned long long>() + n0;
  const uint32_t* tail_w = s.iem_w.as<uint32_t>() + n0;
  if (A && A <= kSmallSort)
    k_sort_small<<<1, 1024, 0, c->stream>>>(tail_k, tail_w, (uint32_t)A, s.add_key.as<unsigned long long>(),
                                            s.add_w.as<uint32_t>());
  else if (A)
    cub(c, [&](void* p, size_t& b) {
      return rocprim::radix_sort_pairs(p, b, tail_k, s.add_key.as<unsigned long long>(), tail_w,
                                       s.add_w.as<uint32_t>(), (size_t)A, 0, 64, c->stream);
    });
  if (n0)
    cub(c, [&](void* p, size_t& b) {
      return rocprim::exclusive_scan(p, b, rocprim::make_transform_iterator(s.iem_dead.as<uint8_t>(), LiveOne()),
                                     s.iem_pre.as<uint32_t>(), 0u, (size_t)n0, rocprim::plus<uint32_t>(), c->stream);
    });
  k_iem_merge<<<grid_for(n0 + A), kBlock, 0, c->stream>>>(
      n0, A, n0 - D, s.iem_key.as<unsigned long long>(), s.iem_w.as<uint32_t>(), s.iem_dead.as<uint8_t>(),
      s.iem_pre.as<uint32_t>(), s.add_key.as<unsigned long long>(), s.add_w.as<uint32_t>(),
      s.iem_key2.as<unsigned long long>(), s.iem_w2.as<uint32_t>());
  SUMINC_HIP(hipGetLastError());
  s.iem_key.swap(s.iem_key2);
  s.iem_w.swap(s.iem_w2);
  SUMINC_HIP(hipMemsetAsync(s.iem_dead.p, 0, std::max(n0, n1) + 1, c->stream));  // grown tails were never cleared
  s.iem_n = n1;
}

void sketch_refresh(suminc_ctx* c) {
  Sketch& s = c->sk;
  SUMINC_REQUIRE(s.built, SUMINC_E_STATE, "no sketch (call suminc_sum_build first)");
  if (s.version != c->graph_version) sketch_layout(c, true);  // acceptance and membership kept
}

// ---------------------------------------------------------------------------
// PageRank shortcuts: unit injection of every entry inside its cluster (LDS)
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ T block_sum(T v, T* s_red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  __syncthreads();
  if (lane_id() == 0) s_red[threadIdx.x >> 6] = v;
  __syncthreads();
  T t = T(0);
  for (int w = 0; w < kBlock / 64; ++w) t += s_red[w];
  return t;
}

// Jacobi rounds inside one cluster held in LDS: r += delta; delta <- sum over
// internal edges of delta[src] * d / outdeg(src), edge-parallel with LDS atomics
// (a hub's in-edges spread over the workgroup); stops when the mass processed in
// a round is <= thr or after maxit rounds.  Returns the rounds run.
template <class T>
__device__ uint32_t cluster_iterate(uint32_t size, uint32_t pb, uint32_t pe, const Internal& ie, T d, double thr,
                                    int maxit, T* s_r, T* s_d, T* s_n, T* s_red) {
  uint32_t rounds = 0;
  for (int it = 0; it < maxit; ++it) {
    T processed = T(0);
    for (uint32_t j = threadIdx.x; j < size; j += kBlock) {
      s_n[j] = T(0);
      processed += fabs_t(s_d[j]);
    }
    __syncthreads();
    for (uint32_t p = pb + threadIdx.x; p < pe; p += kBlock) {
      const T x = s_d[ie.src[p]];
      if (x != T(0)) atomicAdd(&s_n[ie.dst[p]], x * d / (T)ie.deg[p]);
    }
    T tot = block_sum(processed, s_red);  // synchronises the workgroup
    for (uint32_t j = threadIdx.x; j < size; j += kBlock) {
      s_r[j] += s_d[j];
      s_d[j] = s_n[j];
    }
    __syncthreads();
    ++rounds;
    if ((double)tot <= thr) break;
  }
  return rounds;
}

// One workgroup per entry (grid-stride over the E entries): a cluster's entries
// are independent injections, so a cluster with many entries spreads over many
// workgroups.  Entries of clusters whose rows are valid are skipped.
template <class T>
__global__ void __launch_bounds__(kBlock) k_pr_shortcuts(uint32_t E, const uint32_t* __restrict__ cid,
                                                         const uint32_t* __restrict__ cl_off,
                                                         const uint32_t* __restrict__ cl_ent,
                                                         const uint32_t* __restrict__ lidx,
                                                         const unsigned long long* __restrict__ sc_base,
                                                         Internal ie, T d, double thr_per_vertex,
                                                         int maxit, T* __restrict__ R, T* __restrict__ D,
                                                         const uint8_t* __restrict__ valid, unsigned long long* work) {
  __shared__ T s_r[kMaxCluster], s_d[kMaxCluster], s_n[kMaxCluster];
  __shared__ T s_red[kBlock / 64];
  unsigned long long edges = 0;
  for (uint32_t e = blockIdx.x; e < E; e += gridDim.x) {
    const uint32_t v = cl_ent[e];
    const uint32_t k = cid[v];
    if (valid[k]) continue;
    const uint32_t slot0 = cl_off[k], size = cl_off[k + 1] - slot0;
    const double thr = thr_per_vertex * size;
    const uint32_t src = lidx[v];
    for (uint32_t j = threadIdx.x; j < size; j += kBlock) {
      s_r[j] = T(0);
      s_d[j] = j == src ? T(1) : T(0);
    }
    __syncthreads();
    const uint32_t pb = ie.off[k], pe = ie.off[k + 1];
    uint32_t rounds = cluster_iterate<T>(size, pb, pe, ie, d, thr, maxit, s_r, s_d, s_n, s_red);
    edges += (unsigned long long)rounds * (pe - pb);
    const unsigned long long base = sc_base[e];
    for (uint32_t j = threadIdx.x; j < size; j += kBlock) {
      R[base + j] = s_r[j];
      D[base + j] = s_d[j];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && edges) atomicAdd(work, edges);
}

// Recomputes the PageRank rows of the clusters whose rows are not current (all of
// them after the first build or a precision change, the touched ones after a batch).
template <class T>
static void compute_pr_rows(suminc_ctx* c) {
  Sketch& s = c->sk;
  if (!s.pr_need && s.pr_rows_for == c->pr_prec) return;
  const size_t es = sizeof(T);
  if (s.pr_rows_for != c->pr_prec) SUMINC_HIP(hipMemsetAsync(s.pr_valid.p, 0, (size_t)s.K + 1, c->stream));
  if (s.pr_rows_for != c->pr_prec || s.sc_R.cap < s.sc_len * es + es) {
    s.sc_R.ensure_slack(s.sc_len * es + es);
    s.sc_D.ensure_slack(s.sc_len * es + es);
  }
  DevBuf work;
  work.ensure(8);
  SUMINC_HIP(hipMemsetAsync(work.p, 0, 8, c->stream));
  const double thr_pv = c->prm.termcheck_threshold / (double)c->n;  // T / N_old * |C| per cluster
  const int maxit = c->prm.shortcut_max_rounds > 0 ? c->prm.shortcut_max_rounds : 100;
  if (s.E)
    k_pr_shortcuts<T><<<std::min<uint32_t>(s.E, 16384), kBlock, 0, c->stream>>>(
        s.E, s.cid.as<uint32_t>(), s.cl_off.as<uint32_t>(), s.cl_ent.as<uint32_t>(), s.lidx.as<uint32_t>(),
        s.sc_base.as<unsigned long long>(), internal_of(s), (T)c->prm.damping, thr_pv, maxit, s.sc_R.as<T>(), s.sc_D.as<T>(), s.pr_valid.as<uint8_t>(),
        work.as<unsigned long long>());
  SUMINC_HIP(hipGetLastError());
  SUMINC_HIP(hipMemsetAsync(s.pr_valid.p, 1, (size_t)s.K + 1, c->stream));
  unsigned long long w = 0;
  SUMINC_HIP(hipMemcpyAsync(&w, work.p, 8, hipMemcpyDeviceToHost, c->stream));
  SUMINC_HIP(hipStreamSynchronize(c->stream));
  c->stats.shortcut_edges = (double)w;  // internal edges read by the shortcut build
  s.pr_rows_for = c->pr_prec;
  s.pr_need = false;
}

// ---------------------------------------------------------------------------
// PageRank query on the sketch
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ void kahan_add(T* rank, float* comp, uint32_t v, T x) {
  if constexpr (sizeof(T) == 4) {
    float r = rank[v], cc = comp[v];
    float y = (float)x - cc;
    float t = r + y;
    comp[v] = (t - r) - y;
    rank[v] = t;
  } else {
    rank[v] += x;
  }
}

// Appends a push item (row v of the sketch out-edges, contribution cval).
template <class T>
__device__ __forceinline__ void push_item(bool want, uint32_t v, T cval, uint32_t rowlen, uint32_t* wl_v, T* wl_c,
                                          uint4* heavy, Counters* cnt) {
  bool light = want && rowlen > 0 && rowlen <= kHeavyDeg;
  if (want && rowlen > kHeavyDeg) {
    uint32_t nch = (rowlen + kChunk - 1) / kChunk;
    unsigned long long slot = atomicAdd(&cnt->heavy_count, (unsigned long long)nch);
    for (uint32_t k = 0; k < nch; ++k) {
      uint4 h;
      h.x = v;
      h.y = k;
      if constexpr (sizeof(T) == 4) {
        h.z = __float_as_uint((float)cval);
        h.w = 0;
      } else {
        unsigned long long bits = __double_as_longlong((double)cval);
        h.z = (uint32_t)bits;
        h.w = (uint32_t)(bits >> 32);
      }
      heavy[slot + k] = h;
    }
  }
  unsigned long long slot = wave_append(light, &cnt->wl_count[0]);
  if (light) {
    wl_v[slot] = v;
    wl_c[slot] = cval;
  }
}

struct SkArgs {
  uint32_t n;
  const uint32_t* pos;     // compact source position of each vertex for the pull's contributions (null: ids)
  const uint32_t* off;     // full out-CSR row pointers (degrees)
  const uint8_t* ntype;
  const uint32_t* ent_of;
  uint32_t* wl_v;
  uint4* heavy;
  Counters* cnt;
  double* partial;
  const uint32_t* cid;
  const uint32_t* cl_nexit;
};

// residual of the sketch phase: |delta| of Single/entry vertices + |bound| of exits,
// and the out-edges the round's senders will scatter (the push / pull choice).
// kSkPer vertices a thread (vertex base + j * kBlock + t, coalesced).
constexpr int kSkPer = 4;
template <class T>
__global__ void __launch_bounds__(kBlock) k_sk_count(SkArgs a, const T* __restrict__ delta,
                                                     const T* __restrict__ bound, double eps) {
  __shared__ double s_sum[kBlock / 64];
  __shared__ unsigned long long s_edges[kBlock / 64];
  if (a.cnt->done) return;
  double x = 0;
  unsigned long long ed = 0;
  const uint32_t base = blockIdx.x * (kBlock * kSkPer);
  // the kSkPer vertices' type, delta and row bounds loaded first (delta of a type
  // that holds none is read and ignored), then the sums
  uint8_t ty[kSkPer];
  T dl[kSkPer];
  uint32_t o0[kSkPer], o1[kSkPer];
#pragma unroll
  for (int j = 0; j < kSkPer; ++j) {
    const uint32_t v = base + j * kBlock + threadIdx.x;
    const bool in = v < a.n;
    ty[j] = in ? a.ntype[v] : (uint8_t)4;
    dl[j] = in ? delta[v] : T(0);
    o0[j] = in ? a.off[v] : 0u;
    o1[j] = in ? a.off[v + 1] : 0u;
  }
#pragma unroll
  for (int j = 0; j < kSkPer; ++j) {
    const uint32_t v = base + j * kBlock + threadIdx.x;
    const uint8_t t = ty[j];
    if (t == 4) continue;
    const uint32_t deg = o1[j] - o0[j];
    double dv = 0;
    if (t != 1) {
      dv = (double)fabs_t(dl[j]);
      x += dv;
      if (t == 0 && dv > eps) ed += deg;
    }
    if (t == 1 || t == 3) {
      const double b = (double)fabs_t(bound[v]);
      x += b;
      // an exit sends in the round its cluster's entries inject: its own entry's
      // mass (a BothOutIn hub's) leaves it this round
      if (b > eps || (t == 3 && dv > eps)) ed += deg;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    x += __shfl_xor(x, o);
    ed += __shfl_xor(ed, o);
  }
  if (lane_id() == 0) {
    s_sum[threadIdx.x >> 6] = x;
    s_edges[threadIdx.x >> 6] = ed;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    unsigned long long te = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
      t += s_sum[w];
      te += s_edges[w];
    }
    a.partial[blockIdx.x] = t;
    reinterpret_cast<unsigned long long*>(a.partial)[2 * gridDim.x + blockIdx.x] = te;
  }
}

// Stop rule of the sketch phase and the round's scatter: pull over the CSC when
// the senders' out-edges exceed m / div (as the plain rounds decide), else push.
__global__ void __launch_bounds__(1024) k_sk_decide(const double* __restrict__ partial, uint32_t nparts, Counters* cnt,
                                                    double T, int max_rounds, uint64_t m, int div,
                                                    uint8_t* __restrict__ mode_hist) {
  __shared__ double s[1024];
  __shared__ unsigned long long se[1024];
  if (cnt->done) return;
  const unsigned long long* pe = reinterpret_cast<const unsigned long long*>(partial) + 2 * nparts;
  double acc = 0;
  unsigned long long acc_e = 0;
  for (uint32_t i = threadIdx.x; i < nparts; i += 1024) {
    acc += partial[i];
    acc_e += pe[i];
  }
  s[threadIdx.x] = acc;
  se[threadIdx.x] = acc_e;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      s[threadIdx.x] += s[threadIdx.x + o];
      se[threadIdx.x] += se[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double r = s[0];
    cnt->residual = r;
    if (r != r || r > 1e300) cnt->aux[0] = 1;
    if (!(r >= T) || (long long)cnt->rounds >= (long long)max_rounds || cnt->aux[0]) {
      cnt->done = 1;
      return;
    }
    cnt->mode = div > 0 && se[0] * (unsigned long long)div > m ? 1 : 0;
    if (cnt->rounds < 65536) mode_hist[cnt->rounds] = (uint8_t)cnt->mode;
    cnt->rounds += 1;
  }
}

// Node-type dispatch of one sketch round (sum_sync_iter_worker.h:949-1111), in
// two passes around the shortcut launch: k_sk_send lets Singles push and entries
// inject into their shortcut rows; k_sk_exits, after the rows moved that mass to
// the exits, lets every exit turn its inflow into rank and send d/outdeg of it out
// of the cluster, so mass crosses a cluster within the round it entered (an order
// the reference's parallel loop may take as well).  The flush at the end of the
// sketch phase is the exit pass over every non-zero inflow.
// Block sums of the round's edges: k_sk_send in partial[g, 2g) (u64 view, g its
// grid), k_sk_exits in partial[2g, 2g + its grid) (the count's slots, read by then).
template <class T>
__device__ __forceinline__ void sk_contrib(const SkArgs& a, T* contrib, uint32_t v, T x) {
  const uint32_t slot = a.pos ? a.pos[v] : v;  // vertices without out-edges have no position and never send
  if (slot != INF32) contrib[slot] = x;
}

template <class T>
__device__ __forceinline__ void sk_block_edges(unsigned long long ed, unsigned long long* out) {
  __shared__ unsigned long long s_e[kBlock / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) ed += __shfl_xor(ed, o);
  if (lane_id() == 0) s_e[threadIdx.x >> 6] = ed;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < kBlock / 64; ++w) t += s_e[w];
    *out = t;
  }
}

template <class T>
__global__ void __launch_bounds__(kBlock) k_sk_send(SkArgs a, T* __restrict__ delta, T* __restrict__ spnode,
                                                    T* __restrict__ rank, float* __restrict__ comp,
                                                    T* __restrict__ wl_c, T* __restrict__ sc_x,
                                                    T* __restrict__ contrib, double eps, T d) {
  if (a.cnt->done) return;
  const bool pull = a.cnt->mode != 0;  // contributions for the CSC pull instead of push items
  const uint32_t base = blockIdx.x * (kBlock * kSkPer);
  unsigned long long ed = 0;
  // the kSkPer vertices' type, delta and row bounds loaded first
  uint8_t ty[kSkPer];
  T dl[kSkPer];
  uint32_t o0[kSkPer], o1[kSkPer], ps[kSkPer];
#pragma unroll
  for (int j = 0; j < kSkPer; ++j) {
    const uint32_t v = base + j * kBlock + threadIdx.x;
    const bool in = v < a.n;
    ty[j] = in ? a.ntype[v] : (uint8_t)4;
    dl[j] = in ? delta[v] : T(0);
    o0[j] = in ? a.off[v] : 0u;
    o1[j] = in ? a.off[v + 1] : 0u;
    // the contribution's slot (vertices without out-edges have no position and never send)
    ps[j] = in && pull && a.pos ? a.pos[v] : v;
  }
#pragma unroll
  for (int j = 0; j < kSkPer; ++j) {  // uniform trip count: push_item appends per wave
    const uint32_t v = base + j * kBlock + threadIdx.x;
    bool want_push = false;
    T cval = T(0);
    uint32_t deg = 0;
    if (v < a.n) {
      const uint8_t t = ty[j];
      deg = o1[j] - o0[j];
      if (t == 0) {  // SingleNode: plain push over its out-list
        const T dv = dl[j];
        if ((double)fabs_t(dv) > eps) {
          delta[v] = T(0);
          kahan_add(rank, comp, v, dv);
          if (deg) { want_push = true; cval = d * dv / (T)deg; }
        }
      } else if (t == 2 || t == 3) {  // entry: inject into the shortcut
        const T dv = dl[j];
        T sval = T(0);
        if ((double)fabs_t(dv) > eps) {
          delta[v] = T(0);
          spnode[v] += dv;
          sval = dv;
          ed += a.cl_nexit[a.cid[v]];
        }
        sc_x[a.ent_of[v]] = sval;  // every entry's slot is written each round (0: idle)
      }
      if (pull && ps[j] != INF32) contrib[ps[j]] = want_push ? cval : T(0);  // every slot: the exits' come after
    }
    push_item<T>(want_push && !pull, v, cval, deg, a.wl_v, wl_c, a.heavy, a.cnt);
    if (want_push) ed += deg;
  }
  sk_block_edges<T>(ed, reinterpret_cast<unsigned long long*>(a.partial) + gridDim.x + blockIdx.x);
}

// The exits of the current layout, one per thread (flush: every non-zero inflow).
template <class T>
__global__ void __launch_bounds__(kBlock) k_sk_exits(SkArgs a, const uint32_t* __restrict__ exits,
                                                     const unsigned long long* __restrict__ n_exits,
                                                     T* __restrict__ bound, T* __restrict__ rank,
                                                     float* __restrict__ comp, T* __restrict__ wl_c,
                                                     T* __restrict__ contrib, double eps, T d, int flush,
                                                     uint32_t part_base) {
  if (a.cnt->done) return;
  const bool pull = a.cnt->mode != 0;
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  bool want_push = false;
  T cval = T(0);
  uint32_t v = 0, deg = 0;
  if (i < *n_exits) {
    v = exits[i];
    deg = a.off[v + 1] - a.off[v];
    const T b = bound[v];
    if (flush ? b != T(0) : (double)fabs_t(b) > eps) {
      bound[v] = T(0);
      kahan_add(rank, comp, v, b);
      // the push skips an exit's edges inside its cluster; the pull takes them back
      if (deg) { want_push = true; cval = d * b / (T)deg; }
    }
    if (pull && want_push) sk_contrib(a, contrib, v, cval);
  }
  push_item<T>(want_push && !pull, v, cval, deg, a.wl_v, wl_c, a.heavy, a.cnt);
  sk_block_edges<T>(want_push ? deg : 0u, reinterpret_cast<unsigned long long*>(a.partial) + part_base + blockIdx.x);
}

// The exits of the current layout: the first cl_nexit[k] member slots of every
// cluster k (a thread per cluster: most clusters have one exit or a few).
__global__ void __launch_bounds__(kBlock) k_exit_list(uint32_t K, const uint32_t* __restrict__ cl_off,
                                                      const uint32_t* __restrict__ cl_mem,
                                                      const uint32_t* __restrict__ cl_nexit, uint32_t* __restrict__ list,
                                                      unsigned long long* __restrict__ count) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t nx = k < K ? cl_nexit[k] : 0u;
  const unsigned long long at = nx ? atomicAdd(count, (unsigned long long)nx) : 0ull;
  const uint32_t base = nx ? cl_off[k] : 0u;
  for (uint32_t j = 0; j < nx; ++j) list[at + j] = cl_mem[base + j];
}

// Shortcut items: bound[x] += R[e][x] * delta for the exits x of the entry's
// cluster (g_index_function, pagerank_ingress.h:102-125).  sc_x holds the delta
// each entry injected this round (0 when idle), indexed by entry: a fixed list
// instead of an appended one, so the gather needs no same-address atomics.
template <class T>
__global__ void __launch_bounds__(kBlock) k_sk_shortcut(uint32_t E, const T* __restrict__ sc_x,
                                                        const Counters* cnt, const uint32_t* __restrict__ cl_ent,
                                                        const uint32_t* __restrict__ cid,
                                                        const uint32_t* __restrict__ cl_off,
                                                        const uint32_t* __restrict__ cl_mem,
                                                        const uint32_t* __restrict__ cl_nexit,
                                                        const unsigned long long* __restrict__ sc_base,
                                                        const T* __restrict__ R, T* __restrict__ bound) {
  __shared__ uint32_t s_scan[kBlock + 1];
  __shared__ uint32_t s_start[kBlock];
  __shared__ uint32_t s_warp[kBlock / 64 + 1];
  __shared__ uint32_t s_mbase[kBlock];
  __shared__ unsigned long long s_rbase[kBlock];
  __shared__ T s_x[kBlock];
  if (cnt->done) return;
  const uint32_t tiles = (E + kBlock - 1) / kBlock;
  for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {
    const uint32_t e = t * kBlock + threadIdx.x;
    uint32_t len = 0;
    const T x = e < E ? sc_x[e] : T(0);
    if (x != T(0)) {
      uint32_t k = cid[cl_ent[e]];
      len = cl_nexit[k];
      s_mbase[threadIdx.x] = cl_off[k];
      s_rbase[threadIdx.x] = sc_base[e];
      s_x[threadIdx.x] = x;
    }
    block_adjacency<kBlock>(0, len, s_scan, s_start, s_warp, [&](int slot, uint32_t j, uint32_t) {
      T w = R[s_rbase[slot] + j];
      if (w != T(0)) unsafeAtomicAdd(&bound[cl_mem[s_mbase[slot] + j]], w * s_x[slot]);
    });
  }
}

__global__ void __launch_bounds__(1024) k_sk_end(Counters* cnt, const unsigned long long* __restrict__ pe,
                                                 uint32_t nparts) {
  __shared__ unsigned long long s[1024];
  if (cnt->done) return;
  unsigned long long acc = 0;
  for (uint32_t i = threadIdx.x; i < nparts; i += 1024) acc += pe[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x) return;
  cnt->edges += s[0];
  cnt->round_edges = 0;
  cnt->wl_count[0] = 0;
  cnt->wl_count[1] = 0;
  cnt->heavy_count = 0;
}

// push list of a flush/precompute pass: run once, then clear
__global__ void k_lists_clear(Counters* cnt) {
  if (threadIdx.x) return;
  cnt->wl_count[0] = 0;
  cnt->wl_count[1] = 0;
  cnt->heavy_count = 0;
  cnt->done = 0;
  cnt->mode = 0;
}

// Precompute of one dirty cluster per workgroup: the members' pending delta is
// iterated inside the cluster (run_to_convergence_for_precpt), the mass that
// arrived becomes rank, the residual stays in delta, and every exit sends d/outdeg
// of its arrived mass over its external edges (first_step, :389-434): the exit's
// arrived mass is left in bound, so the first sketch round's exit pass turns it
// into rank and sends it with that round's scatter.  Dissolved clusters are
// Singles and are not iterated.
template <class T>
__global__ void __launch_bounds__(kBlock) k_sk_precompute(SkArgs a, const uint32_t* __restrict__ cl_off,
                                                          const uint32_t* __restrict__ cl_mem,
                                                          const uint32_t* __restrict__ cl_nexit, Internal ie,
                                                          uint8_t* dirty, const uint8_t* __restrict__ dead,
                                                          T* __restrict__ delta, T* __restrict__ rank,
                                                          float* __restrict__ comp, T* __restrict__ bound, T d,
                                                          double thr_per_vertex, int maxit,
                                                          unsigned long long* work) {
  __shared__ T s_r[kMaxCluster], s_d[kMaxCluster], s_n[kMaxCluster];
  __shared__ T s_red[kBlock / 64];
  const uint32_t k = blockIdx.x;
  if (!dirty[k]) return;
  if (dead[k]) {
    if (threadIdx.x == 0) dirty[k] = 0;
    return;
  }
  const uint32_t slot0 = cl_off[k], size = cl_off[k + 1] - slot0, nexit = cl_nexit[k];
  for (uint32_t j = threadIdx.x; j < size; j += kBlock) {
    uint32_t v = cl_mem[slot0 + j];
    s_r[j] = T(0);
    s_d[j] = delta[v];
  }
  __syncthreads();
  const uint32_t pb = ie.off[k], pe = ie.off[k + 1];
  uint32_t rounds = cluster_iterate<T>(size, pb, pe, ie, d, thr_per_vertex * size, maxit, s_r, s_d, s_n, s_red);
  for (uint32_t j = threadIdx.x; j < size; j += kBlock) {
    const uint32_t v = cl_mem[slot0 + j];
    delta[v] = s_d[j];
    if (s_r[j] != T(0)) {
      if (j < nexit) bound[v] += s_r[j];  // sent by the next exit pass
      else kahan_add(rank, comp, v, s_r[j]);
    }
  }
  if (threadIdx.x == 0) {
    dirty[k] = 0;
    atomicAdd(work, (unsigned long long)rounds * (pe - pb));
  }
}

// Correction (sum_sync_iter_worker.h:1414-1469): the delta accumulated at every
// entry is spread with its shortcut row: rank of non-exits += s * R, delta of all
// members += s * D (inner_value / inner_delta).  One workgroup per cluster sums
// its entries in LDS so every member is updated once.
template <class T>
__global__ void __launch_bounds__(kBlock) k_sk_correct(const uint32_t* __restrict__ cl_off,
                                                       const uint32_t* __restrict__ cl_mem,
                                                       const uint32_t* __restrict__ cl_nexit,
                                                       const uint32_t* __restrict__ cl_eoff,
                                                       const uint32_t* __restrict__ cl_ent,
                                                       const unsigned long long* __restrict__ sc_base,
                                                       const T* __restrict__ R, const T* __restrict__ D,
                                                       T* __restrict__ spnode, T* __restrict__ delta,
                                                       T* __restrict__ rank, float* __restrict__ comp) {
  __shared__ T s_ar[kMaxCluster], s_ad[kMaxCluster];
  const uint32_t k = blockIdx.x;
  const uint32_t slot0 = cl_off[k], size = cl_off[k + 1] - slot0, nexit = cl_nexit[k];
  for (uint32_t j = threadIdx.x; j < size; j += kBlock) {
    s_ar[j] = T(0);
    s_ad[j] = T(0);
  }
  __syncthreads();
  for (uint32_t e = cl_eoff[k]; e < cl_eoff[k + 1]; ++e) {
    const uint32_t s = cl_ent[e];
    const T sp = spnode[s];
    if (sp == T(0)) continue;  // uniform across the block
    const unsigned long long base = sc_base[e];
    for (uint32_t j = threadIdx.x; j < size; j += kBlock) {
      if (j >= nexit) s_ar[j] += sp * R[base + j];
      s_ad[j] += sp * D[base + j];
    }
    __syncthreads();
    if (threadIdx.x == 0) spnode[s] = T(0);
  }
  __syncthreads();
  for (uint32_t j = threadIdx.x; j < size; j += kBlock) {
    const uint32_t v = cl_mem[slot0 + j];
    if (s_ar[j] != T(0)) kahan_add(rank, comp, v, s_ar[j]);
    if (s_ad[j] != T(0)) delta[v] += s_ad[j];
  }
}

// chk[v] = 1 iff an internal edge of v's cluster runs from one of its exits
// (local index < nexit) to v.  Only such in-edges carry a contribution the
// sketch pull must skip; every other vertex sums all its in-edges unchecked.
// chk is all zero on entry (the caller clears it).
// The internal edges leaving an exit of the current layout, as (contribution slot
// of the exit, destination): a thread per cluster.
__global__ void __launch_bounds__(kBlock) k_pull_internal(uint32_t K, const uint32_t* __restrict__ ie_off,
                                                          const uint32_t* __restrict__ ie_src,
                                                          const uint32_t* __restrict__ ie_dst,
                                                          const uint32_t* __restrict__ cl_off,
                                                          const uint32_t* __restrict__ cl_mem,
                                                          const uint32_t* __restrict__ cl_nexit,
                                                          const uint32_t* __restrict__ pos, uint2* __restrict__ list,
                                                          unsigned long long* __restrict__ count) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= K) return;
  const uint32_t nx = cl_nexit[k];
  if (!nx) return;
  const uint32_t base = cl_off[k], b = ie_off[k], e = ie_off[k + 1];
  uint32_t take = 0;
  for (uint32_t i = b; i < e; ++i) take += ie_src[i] < nx;
  if (!take) return;
  unsigned long long at = atomicAdd(count, (unsigned long long)take);
  for (uint32_t i = b; i < e; ++i) {
    if (ie_src[i] >= nx) continue;
    uint32_t su = cl_mem[base + ie_src[i]];
    if (pos) su = pos[su];
    list[at++] = make_uint2(su, cl_mem[base + ie_dst[i]]);
  }
}

// After a sketch round's pull: the pull gathered every in-edge, so the mass an
// exit sent along its internal edges (carried by the shortcut rows instead) is
// taken back.  list holds (contribution slot of the exit, destination) per
// internal edge leaving an exit.
template <class T>
__global__ void __launch_bounds__(kBlock) k_pull_unskip(const uint2* __restrict__ list,
                                                        const unsigned long long* __restrict__ count,
                                                        const T* __restrict__ contrib, T* __restrict__ delta,
                                                        const Counters* cnt) {
  if (cnt->done || cnt->mode == 0) return;
  const unsigned long long L = *count;
  for (unsigned long long i = (unsigned long long)blockIdx.x * kBlock + threadIdx.x; i < L;
       i += (unsigned long long)gridDim.x * kBlock) {
    const uint2 ed = list[i];
    const T x = contrib[ed.x];
    if (x != T(0)) unsafeAtomicAdd(&delta[ed.y], -x);
  }
}

template <class T>
static void sk_push_pass(suminc_ctx* c) {
  Sketch& s = c->sk;
  k_pr_push<T, true><<<push_grid(), kBlock, 0, c->stream>>>(
      c->out_off.as<uint32_t>(), c->out_col.as<uint32_t>(), c->pr_delta.as<T>(), c->wl0.as<uint32_t>(),
      c->wl1.as<T>(), c->heavy.as<uint4>(), dev_counters(c), s.ntype.as<uint8_t>(), s.cid.as<uint32_t>());
}

template <class T>
static void sum_pr_run_t(suminc_ctx* c, suminc_stats* st) {
  Sketch& s = c->sk;
  Counters* cnt = dev_counters(c);
  const uint32_t n = c->n;
  const double T_stop = c->prm.termcheck_threshold;
  const double eps = T_stop / (double)n;
  const int max_rounds = c->prm.max_rounds > 0 ? c->prm.max_rounds : 1000000;
  const int maxit = c->prm.shortcut_max_rounds > 0 ? c->prm.shortcut_max_rounds : 100;
  const T d = (T)c->prm.damping;
  const size_t es = sizeof(T);
  const int div = c->prm.pull_threshold_div;
  if (div > 0) ensure_pull_heavy(c);
  compute_pr_rows<T>(c);
  s.bound.ensure((size_t)n * es);
  s.spnode.ensure((size_t)n * es);
  SUMINC_HIP(hipMemsetAsync(s.bound.p, 0, (size_t)n * es, c->stream));
  SUMINC_HIP(hipMemsetAsync(s.spnode.p, 0, (size_t)n * es, c->stream));
  c->tmp3.ensure((size_t)n * es + 8);
  // the sketch rounds pull like the plain rounds (contributions at the plain pull's
  // compact positions when it uses them), then take back what exits sent along
  // their internal edges: the list of those edges for this layout
  SkArgs a;
  a.n = n;
  a.pos = div > 0 && c->pull_compact ? c->src_pos.as<uint32_t>() : nullptr;
  if (div > 0) {
    s.xi_list.ensure((s.ie_m + 1) * sizeof(uint2));
    s.xi_count.ensure(8);
    SUMINC_HIP(hipMemsetAsync(s.xi_count.p, 0, 8, c->stream));
    if (s.K)
      k_pull_internal<<<grid_for(s.K), kBlock, 0, c->stream>>>(
          s.K, s.ie_off.as<uint32_t>(), s.ie_src.as<uint32_t>(), s.ie_dst.as<uint32_t>(), s.cl_off.as<uint32_t>(),
          s.cl_mem.as<uint32_t>(), s.cl_nexit.as<uint32_t>(), a.pos, s.xi_list.as<uint2>(),
          s.xi_count.as<unsigned long long>());
  }
  a.off = c->out_off.as<uint32_t>();
  a.ntype = s.ntype.as<uint8_t>();
  a.ent_of = s.ent_of.as<uint32_t>();
  a.wl_v = c->wl0.as<uint32_t>();
  a.heavy = c->heavy.as<uint4>();
  a.cnt = cnt;
  a.partial = c->partial.as<double>();
  a.cid = s.cid.as<uint32_t>();
  a.cl_nexit = s.cl_nexit.as<uint32_t>();
  DevBuf work;
  work.ensure(8);
  SUMINC_HIP(hipMemsetAsync(work.p, 0, 8, c->stream));
  suminc_stats out{};
  SUMINC_HIP(hipEventRecord(c->ev_start, c->stream));
  // 1. precompute of dirty clusters + the exits' sends
  k_lists_clear<<<1, 64, 0, c->stream>>>(cnt);
  if (s.K)
    k_sk_precompute<T><<<s.K, kBlock, 0, c->stream>>>(a, s.cl_off.as<uint32_t>(), s.cl_mem.as<uint32_t>(),
                                                      s.cl_nexit.as<uint32_t>(), internal_of(s),
                                                      s.dirty.as<uint8_t>(), s.dead.as<uint8_t>(),
                                                      c->pr_delta.as<T>(), c->pr_rank.as<T>(),
                                                      c->pr_comp.as<float>(), s.bound.as<T>(), d, eps, maxit,
                                                      work.as<unsigned long long>());
  k_lists_clear<<<1, 64, 0, c->stream>>>(cnt);
  SUMINC_HIP(hipGetLastError());
  // the exits of this layout, for the rounds' exit pass
  s.exits.ensure((size_t)s.M * 4 + 4);
  s.exit_count.ensure(8);
  SUMINC_HIP(hipMemsetAsync(s.exit_count.p, 0, 8, c->stream));
  if (s.K)
    k_exit_list<<<grid_for(s.K), kBlock, 0, c->stream>>>(
        s.K, s.cl_off.as<uint32_t>(), s.cl_mem.as<uint32_t>(), s.cl_nexit.as<uint32_t>(), s.exits.as<uint32_t>(),
        s.exit_count.as<unsigned long long>());
  check_begin(c, s.exit_count.p, 8);
  const uint64_t n_exits = *reinterpret_cast<const unsigned long long*>(check_wait(c));
  // 2. sketch rounds
  const unsigned g4 = grid_for(n, kBlock * kSkPer);  // the vertex passes: kSkPer vertices a thread
  const unsigned ge = grid_for(n_exits);             // the exit pass
  c->tmp4.ensure(65536);  // per-round scatter mode (profile attribution)
  int group = 4;
  size_t slot = 0;
  while (true) {
    for (int i = 0; i < group; ++i, ++slot) {
      k_sk_count<T><<<g4, kBlock, 0, c->stream>>>(a, c->pr_delta.as<T>(), s.bound.as<T>(), eps);
      k_sk_decide<<<1, 1024, 0, c->stream>>>(a.partial, g4, cnt, T_stop, max_rounds, c->m, div,
                                             c->tmp4.as<uint8_t>());
      k_sk_send<T><<<g4, kBlock, 0, c->stream>>>(a, c->pr_delta.as<T>(), s.spnode.as<T>(), c->pr_rank.as<T>(),
                                                 c->pr_comp.as<float>(), c->wl1.as<T>(), c->tmp3.as<T>(),
                                                 c->pr_contrib.as<T>(), eps, d);
      k_sk_shortcut<T><<<grid_for(s.E, kBlock, push_grid()), kBlock, 0, c->stream>>>(
          s.E, c->tmp3.as<T>(), cnt, s.cl_ent.as<uint32_t>(), s.cid.as<uint32_t>(),
          s.cl_off.as<uint32_t>(), s.cl_mem.as<uint32_t>(), s.cl_nexit.as<uint32_t>(),
          s.sc_base.as<unsigned long long>(), s.sc_R.as<T>(), s.bound.as<T>());
      k_sk_exits<T><<<ge, kBlock, 0, c->stream>>>(a, s.exits.as<uint32_t>(), s.exit_count.as<unsigned long long>(),
                                                  s.bound.as<T>(), c->pr_rank.as<T>(), c->pr_comp.as<float>(),
                                                  c->wl1.as<T>(), c->pr_contrib.as<T>(), eps, d, 0, 2 * g4);
      sk_push_pass<T>(c);
      if (c->prof_on) SUMINC_HIP(hipEventRecord(prof_event(c, 2 * slot), c->stream));
      if (div > 0) pull_scatter<T>(c, 0);
      if (c->prof_on) SUMINC_HIP(hipEventRecord(prof_event(c, 2 * slot + 1), c->stream));
      if (div > 0)
        k_pull_unskip<T><<<grid_for(s.ie_m + 1, kBlock, 4096), kBlock, 0, c->stream>>>(
            s.xi_list.as<uint2>(), s.xi_count.as<unsigned long long>(), c->pr_contrib.as<T>(), c->pr_delta.as<T>(),
            cnt);
      k_sk_end<<<1, 1024, 0, c->stream>>>(cnt, reinterpret_cast<unsigned long long*>(a.partial) + g4, g4 + ge);
    }
    SUMINC_HIP(hipGetLastError());
    read_counters(c);
    if (c->host_counters->done) break;
    if (group < 32) group *= 2;
  }
  const Counters* h = c->host_counters;
  out.sketch_rounds = (double)h->rounds;  // sketch rounds
  if (c->prof_on && h->rounds) {
    // the sketch rounds' pull launches (the plain rounds' kernel): device time of
    // the rounds that pulled, and their algorithmic bytes (row index + contribution
    // per in-edge; row pointers and the delta read-modify-write per vertex)
    std::vector<uint8_t> mode(std::min<size_t>((size_t)h->rounds, 65536));
    SUMINC_HIP(hipMemcpy(mode.data(), c->tmp4.p, mode.size(), hipMemcpyDeviceToHost));
    suminc_kernel_profile& q = c->prof[SUMINC_KERNEL_PR_PULL];
    for (size_t i = 0; i < mode.size(); ++i) {
      if (!mode[i]) continue;
      float t = 0;
      SUMINC_HIP(hipEventElapsedTime(&t, c->prof_events[2 * i], c->prof_events[2 * i + 1]));
      q.launches += 1;
      q.ms += t;
      q.edges += (double)c->m;
      q.vertices += (double)n;
      q.alg_bytes += (double)c->m * (4.0 + es) + (double)n * (8.0 + 2.0 * es);
    }
  }
  SUMINC_REQUIRE(!h->aux[0], SUMINC_E_NUMERIC, "NaN/Inf in the sketch residual");
  // 3. flush the exits' pending inflow, 4. correction
  k_lists_clear<<<1, 64, 0, c->stream>>>(cnt);
  k_sk_exits<T><<<ge, kBlock, 0, c->stream>>>(a, s.exits.as<uint32_t>(), s.exit_count.as<unsigned long long>(),
                                              s.bound.as<T>(), c->pr_rank.as<T>(), c->pr_comp.as<float>(),
                                              c->wl1.as<T>(), c->pr_contrib.as<T>(), eps, d, 1, 2 * g4);
  sk_push_pass<T>(c);
  k_lists_clear<<<1, 64, 0, c->stream>>>(cnt);
  if (s.K)
    k_sk_correct<T><<<s.K, kBlock, 0, c->stream>>>(s.cl_off.as<uint32_t>(), s.cl_mem.as<uint32_t>(),
                                                   s.cl_nexit.as<uint32_t>(), s.cl_eoff.as<uint32_t>(),
                                                   s.cl_ent.as<uint32_t>(), s.sc_base.as<unsigned long long>(),
                                                   s.sc_R.as<T>(), s.sc_D.as<T>(), s.spnode.as<T>(),
                                                   c->pr_delta.as<T>(), c->pr_rank.as<T>(), c->pr_comp.as<float>());
  SUMINC_HIP(hipGetLastError());
  SUMINC_HIP(hipEventRecord(c->ev_stop, c->stream));
  SUMINC_HIP(hipEventSynchronize(c->ev_stop));
  float ms_sketch = 0;
  SUMINC_HIP(hipEventElapsedTime(&ms_sketch, c->ev_start, c->ev_stop));
  unsigned long long wk = 0;
  SUMINC_HIP(hipMemcpy(&wk, work.p, 8, hipMemcpyDeviceToHost));
  const double sketch_edges = (double)h->edges;
  // 5. residual phase: plain Ingress rounds on the full graph
  suminc_stats rs{};
  pr_run(c, &rs);
  out.rounds = rs.rounds + out.sketch_rounds;
  out.edges_traversed = rs.edges_traversed + sketch_edges + (double)wk;
  out.residual = rs.residual;
  out.kernel_ms = rs.kernel_ms + ms_sketch;
  out.push_rounds = rs.push_rounds;
  out.pull_rounds = rs.pull_rounds;
  out.mass_defect = rs.mass_defect;
  out.mass_total = rs.mass_total;
  out.sketch_ms = ms_sketch;
  out.residual_rounds = rs.rounds;  // residual-phase rounds
  out.precompute_edges = (double)wk;  // internal edges read by precompute
  c->stats = out;
  if (st) *st = out;
}

void sum_pr_run(suminc_ctx* c, suminc_stats* st) {
  SUMINC_REQUIRE(c->pr_ready, SUMINC_E_STATE, "PageRank state not initialised");
  sketch_refresh(c);
  if (c->pr_prec == SUMINC_PR_FP64) sum_pr_run_t<double>(c, st);
  else sum_pr_run_t<float>(c, st);
}

void sketch_info(suminc_ctx* c, suminc_sketch_info* out) {
  SUMINC_REQUIRE(out, SUMINC_E_INVALID, "null output");
  sketch_refresh(c);
  Sketch& s = c->sk;
  std::memset(out, 0, sizeof(*out));
  uint32_t hist[8] = {0};
  s.lay_hist.ensure(32);
  SUMINC_HIP(hipMemsetAsync(s.lay_hist.p, 0, 32, c->stream));
  if (s.M)
    k_type_hist<<<grid_for(s.M), kBlock, 0, c->stream>>>(s.M, s.base_mem.as<uint32_t>(), s.ntype.as<uint8_t>(),
                                                         s.lay_hist.as<uint32_t>());
  SUMINC_HIP(hipMemcpyAsync(hist, s.lay_hist.p, 32, hipMemcpyDeviceToHost, c->stream));
  SUMINC_HIP(hipStreamSynchronize(c->stream));
  uint32_t typed = 0;
  for (int t = 1; t < 5; ++t) {
    s.counts[t] = hist[t];
    typed += hist[t];
  }
  // members of dissolved clusters are Singles (hist[0] over the members)
  SUMINC_REQUIRE(typed + hist[0] == s.M, SUMINC_E_STATE, "supernode member count changed");
  s.counts[0] = c->n - typed;
  // dissolved clusters and the internal edges of the live ones
  std::vector<uint8_t> dead(s.K);
  std::vector<uint32_t> ieo((size_t)s.K + 1);
  if (s.K) {
    SUMINC_HIP(hipMemcpy(dead.data(), s.dead.p, s.K, hipMemcpyDeviceToHost));
    SUMINC_HIP(hipMemcpy(ieo.data(), s.ie_off.p, ((size_t)s.K + 1) * 4, hipMemcpyDeviceToHost));
  }
  uint64_t n_dead = 0, live_ie = 0;
  for (uint32_t k = 0; k < s.K; ++k) {
    if (dead[k]) ++n_dead;
    else live_ie += ieo[k + 1] - ieo[k];
  }
  out->clusters = s.K - n_dead;
  out->dissolved_clusters = n_dead;
  out->raw_clusters = s.Kraw;
  out->entries = s.E;
  out->clustered_vertices = typed;
  for (int t = 0; t < 5; ++t) out->node_type_count[t] = s.counts[t];
  out->sketch_edges = c->m - live_ie;  // Singles' CSR rows + the exits' external edges
  out->internal_edges = live_ie;
  out->shortcut_slots = s.sc_len;
  out->max_cluster = s.max_cluster;
}

// cluster of every vertex, SUMINC_INF_U32 for Singles and members of dissolved clusters
__global__ void k_live_cid(uint32_t n, const uint32_t* __restrict__ cid, const uint8_t* __restrict__ dead,
                           uint32_t* __restrict__ out) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const uint32_t k = cid[v];
  out[v] = k != INF32 && !dead[k] ? k : INF32;
}

void sketch_download(suminc_ctx* c, uint32_t* cluster_of, uint8_t* node_type) {
  sketch_refresh(c);
  Sketch& s = c->sk;
  if (cluster_of) {
    DevBuf live;
    live.ensure((size_t)c->n * 4 + 4);
    k_live_cid<<<grid_for(c->n), kBlock, 0, c->stream>>>(c->n, s.cid.as<uint32_t>(), s.dead.as<uint8_t>(),
                                                        live.as<uint32_t>());
    SUMINC_HIP(hipGetLastError());
    SUMINC_HIP(hipMemcpyAsync(cluster_of, live.p, (size_t)c->n * 4, hipMemcpyDeviceToHost, c->stream));
    SUMINC_HIP(hipStreamSynchronize(c->stream));
  }
  if (node_type) SUMINC_HIP(hipMemcpy(node_type, s.ntype.p, c->n, hipMemcpyDeviceToHost));
}

// Shortcut row of entry vertex v: members of its cluster (exits first) with the
// reach mass and residual (PageRank, FP64 copies).
void sketch_pr_shortcut(suminc_ctx* c, uint32_t v, uint32_t cap, uint32_t* members, double* reach, double* residual,
                        uint32_t* count, uint32_t* nexit) {
  SUMINC_REQUIRE(c->pr_ready, SUMINC_E_STATE, "PageRank state not initialised");
  sketch_refresh(c);
  Sketch& s = c->sk;
  SUMINC_REQUIRE(v < c->n, SUMINC_E_INVALID, "vertex out of range");
  if (c->pr_prec == SUMINC_PR_FP64) compute_pr_rows<double>(c);
  else compute_pr_rows<float>(c);
  uint32_t e = INF32, k = INF32;
  SUMINC_HIP(hipMemcpy(&e, s.ent_of.as<uint32_t>() + v, 4, hipMemcpyDeviceToHost));
  SUMINC_REQUIRE(e != INF32, SUMINC_E_INVALID, "vertex is not an entry of a supernode");
  SUMINC_HIP(hipMemcpy(&k, s.cid.as<uint32_t>() + v, 4, hipMemcpyDeviceToHost));
  uint32_t co[2], ne = 0;
  unsigned long long base = 0;
  SUMINC_HIP(hipMemcpy(co, s.cl_off.as<uint32_t>() + k, 8, hipMemcpyDeviceToHost));
  SUMINC_HIP(hipMemcpy(&ne, s.cl_nexit.as<uint32_t>() + k, 4, hipMemcpyDeviceToHost));
  SUMINC_HIP(hipMemcpy(&base, s.sc_base.as<unsigned long long>() + e, 8, hipMemcpyDeviceToHost));
  const uint32_t size = co[1] - co[0];
  if (count) *count = size;
  if (nexit) *nexit = ne;
  SUMINC_REQUIRE(cap >= size || !(members || reach || residual), SUMINC_E_INVALID, "output capacity too small");
  if (members) SUMINC_HIP(hipMemcpy(members, s.cl_mem.as<uint32_t>() + co[0], (size_t)size * 4, hipMemcpyDeviceToHost));
  if (reach || residual) {
    if (c->pr_prec == SUMINC_PR_FP64) {
      if (reach) SUMINC_HIP(hipMemcpy(reach, s.sc_R.as<double>() + base, (size_t)size * 8, hipMemcpyDeviceToHost));
      if (residual) SUMINC_HIP(hipMemcpy(residual, s.sc_D.as<double>() + base, (size_t)size * 8, hipMemcpyDeviceToHost));
    } else {
      std::vector<float> tmp(size);
      if (reach) {
        SUMINC_HIP(hipMemcpy(tmp.data(), s.sc_R.as<float>() + base, (size_t)size * 4, hipMemcpyDeviceToHost));
        for (uint32_t j = 0; j < size; ++j) reach[j] = tmp[j];
      }
      if (residual) {
        SUMINC_HIP(hipMemcpy(tmp.data(), s.sc_D.as<float>() + base, (size_t)size * 4, hipMemcpyDeviceToHost));
        for (uint32_t j = 0; j < size; ++j) residual[j] = tmp[j];
      }
    }
  }
}

}  // namespace suminc
