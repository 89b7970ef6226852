// compress.hip -- the supernode compressor with mirror vertices, and the device
// shortcut build over its clusters.
//
// The compressor stays on the host, as in the reference (the device path takes its
// products through the sketch upload); it restates, from a cluster assignment:
//   acceptance with mirror selection   compressor_base.h:3253-3611 (get_init_supernode_by_clusters)
//   mirrors, absorption, supernodes    :3616-3783 (final_build_supernode)
//   cluster subgraphs with mirrors     :3786-3856 (build_subgraph_mirror)
//   exits (supernode_out_bound)        :3924-4007 (judge_out_bound_node)
//   node types 0..6                    :4144-4182 (get_nodetype_mirror)
//   is_e_ / ib_e_ / sync_e_            :4636-4899 (sketch2csr_mirror)
// The shortcut rows (iter_compressor.h:318-491, build_iter_index_mirror) are
// computed on the device, one workgroup per supernode source, in global memory
// (no cluster-size cap).  The products are uploaded as the context's sketch in the
// reference's layout (rsketch.hip).
//
// Deliberate choices where the reference leaves an order open: clusters are taken
// in cluster-id order, a cluster's members and mirror candidates in id order
// (the reference iterates std::set / std::unordered_set), and mirror ids are
// numbered cluster by cluster, in-mirrors before out-mirrors, by master id.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "ctx.hpp"
#include "device.hpp"
#include "pr.hpp"

namespace suminc {

namespace {

struct HostGraph {
  uint32_t n = 0;
  std::vector<uint32_t> ooff, ocol, ioff, icol;
  uint32_t outdeg(uint32_t v) const { return ooff[v + 1] - ooff[v]; }
};

inline bool has(const std::vector<uint32_t>& sorted, uint32_t x) {
  return std::binary_search(sorted.begin(), sorted.end(), x);
}

inline void erase_sorted(std::vector<uint32_t>& sorted, uint32_t x) {
  auto it = std::lower_bound(sorted.begin(), sorted.end(), x);
  if (it != sorted.end() && *it == x) sorted.erase(it);
}

constexpr uint32_t kNone = 0xFFFFFFFFu;

}  // namespace

// Host products of one compression (also what suminc_sketch_download returns).
struct MirrorSketch {
  uint32_t n = 0, all = 0, K = 0;
  std::vector<uint8_t> type;                    // n
  std::vector<uint32_t> master;                 // all - n
  std::vector<uint32_t> cluster_of;             // all: accepted cluster per id (kNone: Single)
  std::vector<uint8_t> out_bound;               // n: supernode_out_bound
  std::vector<std::vector<uint32_t>> members;   // K: supernode_ids (sorted, then absorbed)
  std::vector<std::vector<uint32_t>> ids;       // K: cluster_ids = members + in-mirror ids + out-mirror ids
  std::vector<std::vector<uint32_t>> in_m, out_m;  // K: masters of in- / out-mirrors (sorted)
  std::vector<uint32_t> sources;                // supernode sources, numbered as the reference does
  std::vector<uint32_t> source_cluster;
  std::vector<std::vector<uint32_t>> sub;       // all: subgraph rows
  uint64_t in_mirrors = 0, out_mirrors = 0, absorbed = 0;
};

// The accepted clusters before the final build: what the reference writes to its
// cluster cache (serialize_cluster, compressor_base.h:3081-3150).
struct Acceptance {
  std::vector<std::vector<uint32_t>> members;       // supernode_ids (= cluster_ids here), ascending
  std::vector<std::vector<uint32_t>> sources;       // supernode_source
  std::vector<std::vector<uint32_t>> in_m, out_m;   // supernode_in_mirror / _out_mirror (masters, ascending)
};

static void accept_clusters(const HostGraph& g, const uint32_t* cluster_of, uint32_t min_node, uint32_t max_node,
                            double threshold, double mirror_k, Acceptance& A) {
  const uint32_t n = g.n;
  const bool use_mirror = mirror_k != 1e8;  // the reference's switch (compressor_base.h:3302)
  const long long k = (long long)mirror_k;
  const float obj = (float)threshold;
  // kept clusters in id order (the cluster file's records, compress_by_cluster)
  uint32_t Kr = 0;
  for (uint32_t v = 0; v < n; ++v)
    if (cluster_of[v] != kNone) Kr = std::max(Kr, cluster_of[v] + 1);
  std::vector<std::vector<uint32_t>> raw(Kr);
  for (uint32_t v = 0; v < n; ++v)
    if (cluster_of[v] != kNone) raw[cluster_of[v]].push_back(v);
  std::vector<std::vector<uint32_t>> clusters;
  std::vector<uint32_t> id2cl(n, kNone);  // id2clusterid over the kept clusters
  for (auto& c : raw)
    if (c.size() >= min_node && c.size() <= max_node) {
      for (uint32_t v : c) id2cl[v] = (uint32_t)clusters.size();
      clusters.push_back(std::move(c));
    }
  // acceptance with mirrors, cluster by cluster (compressor_base.h:3316-3600)
  std::vector<uint32_t> id2sp(n, kNone);        // accepted cluster so far
  for (uint32_t j = 0; j < clusters.size(); ++j) {
    const std::vector<uint32_t>& P = clusters[j];
    std::unordered_map<uint32_t, long long> in_freq, out_freq;
    std::vector<uint32_t> old_entry, old_exit;
    long long inner = 0;
    for (uint32_t u : P) {
      bool ext_in = false, ext_out = false;
      for (uint32_t e = g.ioff[u]; e < g.ioff[u + 1]; ++e) {
        const uint32_t w = g.icol[e];
        if (id2cl[w] != j) {
          if (use_mirror) {
            const uint32_t sp = id2sp[w];
            if (sp == kNone || !has(A.out_m[sp], u)) in_freq[w] += 1;
          }
          ext_in = true;
        } else {
          ++inner;
        }
      }
      for (uint32_t e = g.ooff[u]; e < g.ooff[u + 1]; ++e) {
        const uint32_t w = g.ocol[e];
        if (id2cl[w] != j) {
          if (use_mirror) {
            const uint32_t sp = id2sp[w];
            if (sp == kNone || !has(A.in_m[sp], u)) out_freq[w] += 1;
          }
          ext_out = true;
        }
      }
      if (ext_in) old_entry.push_back(u);
      if (ext_out) old_exit.push_back(u);
    }
    std::vector<uint32_t> in_mirror, out_mirror;
    long long in_edges = 0, out_edges = 0;
    for (const auto& f : in_freq)
      if (f.second > k) { in_mirror.push_back(f.first); in_edges += f.second; }
    for (const auto& f : out_freq)
      if (f.second > k) { out_mirror.push_back(f.first); out_edges += f.second; }
    std::sort(in_mirror.begin(), in_mirror.end());
    std::sort(out_mirror.begin(), out_mirror.end());
    // new exits B (an out-edge leaving P u out-mirrors) and entries S (an in-edge
    // from outside P u in-mirrors)
    std::vector<uint32_t> B, Sn;
    for (uint32_t u : P) {
      for (uint32_t e = g.ooff[u]; e < g.ooff[u + 1]; ++e) {
        const uint32_t w = g.ocol[e];
        if (id2cl[w] != j && !has(out_mirror, w)) { B.push_back(u); break; }
      }
      for (uint32_t e = g.ioff[u]; e < g.ioff[u + 1]; ++e) {
        const uint32_t w = g.icol[e];
        if (id2cl[w] != j && !has(in_mirror, w)) { Sn.push_back(u); break; }
      }
    }
    const long long oX = (long long)old_exit.size(), oE = (long long)old_entry.size();
    const long long nX = (long long)B.size() + (long long)out_mirror.size();
    const long long nE = (long long)Sn.size() + (long long)in_mirror.size();
    float benefit[4];
    benefit[0] = (float)((double)inner - (double)(oX * oE));
    benefit[1] = (float)((double)(inner + in_edges) - (double)(oX * nE + (long long)in_mirror.size()));
    benefit[2] = (float)((double)(inner + out_edges) - (double)(nX * oE + (long long)out_mirror.size()));
    benefit[3] = (float)((double)(inner + in_edges + out_edges) -
                         (double)(nX * nE + (long long)in_mirror.size() + (long long)out_mirror.size()));
    int mi = 0;
    for (int i = 1; i < 4; ++i)
      if (benefit[mi] < benefit[i]) mi = i;
    if (!(benefit[mi] > obj)) continue;
    if (mi == 0) {
      Sn = old_entry;
      in_mirror.clear();
      out_mirror.clear();
    } else if (mi == 1) {
      out_mirror.clear();
    } else if (mi == 2) {
      Sn = old_entry;
      in_mirror.clear();
    }
    const uint32_t id = (uint32_t)A.members.size();
    A.members.push_back(P);
    A.sources.push_back(Sn);
    A.in_m.push_back(in_mirror);
    A.out_m.push_back(out_mirror);
    for (uint32_t u : P) id2sp[u] = id;
  }
}

// The final build from the accepted clusters (final_build_supernode and what
// follows it, compressor_base.h:3616-4182).
static void final_build(const HostGraph& g, const Acceptance& A, MirrorSketch& S) {
  const uint32_t n = g.n;
  S = MirrorSketch();
  S.members = A.members;
  S.in_m = A.in_m;
  S.out_m = A.out_m;
  std::vector<std::vector<uint32_t>> src_sets = A.sources;
  std::vector<uint32_t> id2sp(n, kNone);
  std::vector<int8_t> fc(n, 0);  // 0: FC_default, -1: member, 1: supernode source (Fc >= 0)
  for (uint32_t i = 0; i < A.members.size(); ++i)
    for (uint32_t u : A.members[i]) {
      fc[u] = -1;
      id2sp[u] = i;
    }
  const uint32_t K = (uint32_t)S.members.size();
  S.K = K;
  S.n = n;
  // mirrors, absorption of unclustered candidates, supernodes (:3616-3717)
  S.ids.assign(K, {});
  std::vector<std::vector<uint32_t>> in_ids(K), out_ids(K);
  uint32_t all = n;
  S.master.clear();
  for (uint32_t i = 0; i < K; ++i) {
    std::vector<uint32_t>& P = S.members[i];
    std::vector<uint32_t>& Sn = src_sets[i];
    std::vector<uint32_t>& im = S.in_m[i];
    std::vector<uint32_t>& om = S.out_m[i];
    S.ids[i] = P;
    for (int pass = 0; pass < 2; ++pass) {
      std::vector<uint32_t> cand = pass == 0 ? im : om, del;
      for (uint32_t v : cand) {
        if (fc[v] != 0) continue;  // clustered: stays a mirror
        P.push_back(v);
        S.ids[i].push_back(v);
        for (uint32_t e = g.ioff[v]; e < g.ioff[v + 1]; ++e) {
          const uint32_t w = g.icol[e];
          if (id2sp[w] != i && !has(im, w)) { Sn.push_back(v); break; }
        }
        fc[v] = -1;
        id2sp[v] = i;
        del.push_back(v);
        ++S.absorbed;
      }
      for (uint32_t v : del) {
        erase_sorted(im, v);
        erase_sorted(om, v);
      }
    }
    for (uint32_t v : im) {
      S.master.push_back(v);
      S.ids[i].push_back(all);
      in_ids[i].push_back(all++);
    }
    for (uint32_t v : om) {
      S.master.push_back(v);
      S.ids[i].push_back(all);
      out_ids[i].push_back(all++);
    }
    S.in_mirrors += im.size();
    S.out_mirrors += om.size();
    for (uint32_t s : Sn) {
      fc[s] = 1;
      S.sources.push_back(s);
      S.source_cluster.push_back(i);
    }
    for (uint32_t m : in_ids[i]) {
      S.sources.push_back(m);
      S.source_cluster.push_back(i);
    }
  }
  S.all = all;
  S.cluster_of.assign(all, kNone);
  for (uint32_t v = 0; v < n; ++v) S.cluster_of[v] = id2sp[v];
  for (uint32_t i = 0; i < K; ++i) {
    for (uint32_t m : in_ids[i]) S.cluster_of[m] = i;
    for (uint32_t m : out_ids[i]) S.cluster_of[m] = i;
  }
  // cluster subgraphs with mirrors (:3786-3856)
  S.sub.assign(all, {});
  for (uint32_t i = 0; i < K; ++i) {
    for (uint32_t v : S.members[i])
      for (uint32_t e = g.ooff[v]; e < g.ooff[v + 1]; ++e)
        if (S.cluster_of[g.ocol[e]] == i) S.sub[v].push_back(g.ocol[e]);
    for (uint32_t m : in_ids[i]) {
      const uint32_t v = S.master[m - n];
      for (uint32_t e = g.ooff[v]; e < g.ooff[v + 1]; ++e)
        if (S.cluster_of[g.ocol[e]] == i) S.sub[m].push_back(g.ocol[e]);
    }
    for (uint32_t m : out_ids[i]) {
      const uint32_t v = S.master[m - n];
      const std::vector<uint32_t>& v_in = S.in_m[S.cluster_of[v]];  // entry priority
      for (uint32_t e = g.ioff[v]; e < g.ioff[v + 1]; ++e) {
        const uint32_t w = g.icol[e];
        if (S.cluster_of[w] == i && !has(v_in, w)) S.sub[w].push_back(m);
      }
    }
  }
  // exits (:3924-4007): an out-edge leaving the cluster that no out-mirror takes;
  // every in-mirror's master is an exit
  S.out_bound.assign(n, 0);
  for (uint32_t i = 0; i < K; ++i)
    for (uint32_t v : S.members[i])
      for (uint32_t e = g.ooff[v]; e < g.ooff[v + 1]; ++e) {
        const uint32_t w = g.ocol[e];
        if (S.cluster_of[w] != i && !has(S.out_m[i], w)) {
          S.out_bound[v] = 1;
          break;
        }
      }
  for (uint32_t i = 0; i < K; ++i)
    for (uint32_t v : S.in_m[i]) S.out_bound[v] = 1;
  // node types (:4144-4182)
  S.type.assign(n, 0);
  for (uint32_t v = 0; v < n; ++v) {
    uint8_t t = 0;
    if (fc[v] == 0) t = 0;
    else if (fc[v] > 0) t = 2;
    else if (!S.out_bound[v]) t = 4;
    if (S.out_bound[v]) t = fc[v] > 0 ? 3 : 1;
    S.type[v] = t;
  }
  for (uint32_t i = 0; i < K; ++i)
    for (uint32_t v : S.in_m[i]) {
      if (S.type[v] == 1) S.type[v] = 5;
      else if (S.type[v] == 3) S.type[v] = 6;
    }
}

// ---------------------------------------------------------------------------------
// Device shortcut build: one workgroup per supernode source, the cluster's round
// state in global memory (three values per id of the cluster list).  A unit of mass
// starts at the source; an in-mirror source first pushes it along its master's
// in-cluster edges (inc_run_to_convergence_mirror, iter_compressor.h:348-387);
// then every member with mass keeps it (reach) and sends d / outdeg of it along its
// internal edges (Jacobi rounds), until the mass the members moved in a round is at
// most thr * members or max_rounds.  Out-mirrors only collect.  Out: per slot of
// the cluster list, the reach (members: rank; out-mirrors: collected mass) and the
// residual (members' pending delta).
// ---------------------------------------------------------------------------------
struct ScArgs {
  const uint32_t* src;       // sources
  const uint32_t* src_cl;    // cluster of each source
  const uint64_t* src_base;  // first output / state slot of each source
  const uint32_t* cl_off;    // cluster list offsets (K + 1)
  const uint32_t* cl_ids;    // cluster lists: members, then in-mirrors, then out-mirrors
  const uint32_t* cl_inner;  // members per cluster
  const uint32_t* lslot;     // all: slot of an id within its cluster list
  const uint32_t* sub_off;   // all + 1
  const uint32_t* sub_col;
  const uint32_t* outdeg;    // all: full out-degree (a mirror: its master's)
  double d, thr;
  int max_rounds;
};

__global__ void __launch_bounds__(kBlock) k_mirror_shortcuts(ScArgs a, uint32_t n, double* __restrict__ r,
                                                             double* __restrict__ dl, double* __restrict__ dn,
                                                             unsigned long long* work) {
  __shared__ double s_red[kBlock / 64];
  __shared__ double s_tot;
  const uint32_t s = a.src[blockIdx.x], k = a.src_cl[blockIdx.x];
  const uint64_t base = a.src_base[blockIdx.x];
  const uint32_t c0 = a.cl_off[k], size = a.cl_off[k + 1] - c0, inner = a.cl_inner[k];
  double* R = r + base;
  double* D = dl + base;
  double* N = dn + base;
  for (uint32_t j = threadIdx.x; j < size; j += kBlock) {
    R[j] = 0.0;
    D[j] = 0.0;
    N[j] = 0.0;
  }
  __syncthreads();
  const uint32_t ss = a.lslot[s];
  if (threadIdx.x == 0) {
    if (s < n) {
      D[ss] = 1.0;
    } else {  // in-mirror source: its master's in-cluster edges first
      const double out = a.d / (double)a.outdeg[s];
      for (uint32_t e = a.sub_off[s]; e < a.sub_off[s + 1]; ++e) D[a.lslot[a.sub_col[e]]] += out;
    }
  }
  __syncthreads();
  unsigned long long edges = 0;
  int round = 0;
  while (true) {
    ++round;
    // members send their pending mass; N collects this round's inflow
    double moved = 0;
    for (uint32_t j = threadIdx.x; j < inner; j += kBlock) {
      const double x = D[j];
      if (x == 0.0) continue;
      D[j] = 0.0;
      R[j] += x;
      moved += fabs(x);
      const uint32_t v = a.cl_ids[c0 + j];
      const double out = a.d * x / (double)a.outdeg[v];
      for (uint32_t e = a.sub_off[v]; e < a.sub_off[v + 1]; ++e) {
        atomicAdd(&N[a.lslot[a.sub_col[e]]], out);
        ++edges;
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) moved += __shfl_xor(moved, o);
    if (lane_id() == 0) s_red[threadIdx.x >> 6] = moved;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0;
      for (int w = 0; w < kBlock / 64; ++w) t += s_red[w];
      s_tot = t;
    }
    __syncthreads();
    // members take the inflow as their next pending mass; out-mirrors keep collecting
    for (uint32_t j = threadIdx.x; j < size; j += kBlock) {
      D[j] += N[j];
      N[j] = 0.0;
    }
    __syncthreads();
    if (s_tot <= a.thr * (double)inner || round >= a.max_rounds) break;
  }
  if (edges) atomicAdd(work, edges);
}

// ---------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------
static void download_graph(suminc_ctx* c, HostGraph& g) {
  g.n = c->n;
  g.ooff.resize(c->n + 1);
  g.ioff.resize(c->n + 1);
  g.ocol.resize(c->m);
  g.icol.resize(c->m);
  SUMINC_HIP(hipMemcpy(g.ooff.data(), c->out_off.p, (size_t)(c->n + 1) * 4, hipMemcpyDeviceToHost));
  SUMINC_HIP(hipMemcpy(g.ioff.data(), c->in_off.p, (size_t)(c->n + 1) * 4, hipMemcpyDeviceToHost));
  if (c->m) {
    SUMINC_HIP(hipMemcpy(g.ocol.data(), c->out_col.p, c->m * 4, hipMemcpyDeviceToHost));
    SUMINC_HIP(hipMemcpy(g.icol.data(), c->in_col.p, c->m * 4, hipMemcpyDeviceToHost));
  }
}

template <class V>
static void up(DevBuf& b, const std::vector<V>& h) {
  b.ensure(h.size() * sizeof(V) + sizeof(V));
  if (!h.empty()) SUMINC_HIP(hipMemcpy(b.p, h.data(), h.size() * sizeof(V), hipMemcpyHostToDevice));
}

// Runs the device shortcut build over the cluster lists of S; reach / resid per
// source, |cluster list| values each, at src_base.
static void mirror_rows(suminc_ctx* c, const MirrorSketch& S, const std::vector<uint32_t>& outdeg, double thr,
                        int max_rounds, std::vector<uint64_t>& src_base, std::vector<double>& reach,
                        std::vector<double>& resid, uint64_t* work_out) {
  const uint32_t K = S.K, all = S.all;
  std::vector<uint32_t> cl_off(K + 1, 0), cl_ids, cl_inner(K), lslot(all, 0);
  for (uint32_t i = 0; i < K; ++i) {
    cl_off[i + 1] = cl_off[i] + (uint32_t)S.ids[i].size();
    cl_inner[i] = (uint32_t)S.members[i].size();
    for (uint32_t j = 0; j < S.ids[i].size(); ++j) {
      cl_ids.push_back(S.ids[i][j]);
      lslot[S.ids[i][j]] = j;
    }
  }
  std::vector<uint32_t> sub_off(all + 1, 0), sub_col;
  for (uint32_t u = 0; u < all; ++u) {
    sub_off[u + 1] = sub_off[u] + (uint32_t)S.sub[u].size();
    sub_col.insert(sub_col.end(), S.sub[u].begin(), S.sub[u].end());
  }
  const uint32_t ns = (uint32_t)S.sources.size();
  src_base.assign(ns + 1, 0);
  for (uint32_t i = 0; i < ns; ++i) src_base[i + 1] = src_base[i] + S.ids[S.source_cluster[i]].size();
  const uint64_t slots = src_base[ns];
  reach.assign(slots, 0.0);
  resid.assign(slots, 0.0);
  if (!ns) return;
  DevBuf d_src, d_scl, d_base, d_off, d_ids, d_inner, d_lslot, d_soff, d_scol, d_deg, d_r, d_d, d_n, d_work;
  up(d_src, S.sources);
  up(d_scl, S.source_cluster);
  up(d_base, src_base);
  up(d_off, cl_off);
  up(d_ids, cl_ids);
  up(d_inner, cl_inner);
  up(d_lslot, lslot);
  up(d_soff, sub_off);
  up(d_scol, sub_col);
  up(d_deg, outdeg);
  d_r.ensure(slots * 8 + 8);
  d_d.ensure(slots * 8 + 8);
  d_n.ensure(slots * 8 + 8);
  d_work.ensure(8);
  SUMINC_HIP(hipMemset(d_work.p, 0, 8));
  ScArgs a;
  a.src = d_src.as<uint32_t>();
  a.src_cl = d_scl.as<uint32_t>();
  a.src_base = d_base.as<uint64_t>();
  a.cl_off = d_off.as<uint32_t>();
  a.cl_ids = d_ids.as<uint32_t>();
  a.cl_inner = d_inner.as<uint32_t>();
  a.lslot = d_lslot.as<uint32_t>();
  a.sub_off = d_soff.as<uint32_t>();
  a.sub_col = d_scol.as<uint32_t>();
  a.outdeg = d_deg.as<uint32_t>();
  a.d = c->prm.damping;
  a.thr = thr;
  a.max_rounds = max_rounds;
  k_mirror_shortcuts<<<ns, kBlock, 0, c->stream>>>(a, S.n, d_r.as<double>(), d_d.as<double>(), d_n.as<double>(),
                                                   d_work.as<unsigned long long>());
  SUMINC_HIP(hipGetLastError());
  SUMINC_HIP(hipStreamSynchronize(c->stream));
  SUMINC_HIP(hipMemcpy(reach.data(), d_r.p, slots * 8, hipMemcpyDeviceToHost));
  SUMINC_HIP(hipMemcpy(resid.data(), d_d.p, slots * 8, hipMemcpyDeviceToHost));
  unsigned long long w = 0;
  SUMINC_HIP(hipMemcpy(&w, d_work.p, 8, hipMemcpyDeviceToHost));
  if (work_out) *work_out = w;
}

// The products in the reference's CSR layout (sketch2csr_mirror, :4636-4899; the
// rows from fianl_build_iter_index_for_mirror, iter_compressor.h:433-491).
struct SketchRows {
  std::vector<uint64_t> is_off, ib_off, sync_off, iv_off, id_off, sub_off;
  std::vector<uint32_t> is_col, ib_col, sync_col, iv_col, id_col, sub_col;
  std::vector<double> is_w, iv_w, id_w;  // FP64 (the download converts to the reference's float)
};

static void sketch_rows(const HostGraph& g, const MirrorSketch& S, const std::vector<uint64_t>& src_base,
                        const std::vector<double>& reach, const std::vector<double>& resid, SketchRows& out) {
  const uint32_t n = S.n, all = S.all;
  // per source: bound_delta (exits + out-mirrors), inner_value, inner_delta
  std::vector<std::vector<std::pair<uint32_t, double>>> bd(all), iv(all), id(all);
  for (uint32_t i = 0; i < S.sources.size(); ++i) {
    const uint32_t s = S.sources[i], k = S.source_cluster[i];
    const std::vector<uint32_t>& L = S.ids[k];
    const uint32_t inner = (uint32_t)S.members[k].size();
    const uint32_t n_in = (uint32_t)S.in_m[k].size();
    for (uint32_t j = 0; j < L.size(); ++j) {
      const double rv = reach[src_base[i] + j], dv = resid[src_base[i] + j];
      const uint32_t x = L[j];
      if (j < inner) {
        if (rv != 0.0) (S.out_bound[x] ? bd[s] : iv[s]).emplace_back(x, rv);
        if (dv != 0.0) id[s].emplace_back(x, dv);
      } else if (j >= inner + n_in) {  // out-mirror: the mass it collected
        if (rv + dv != 0.0) bd[s].emplace_back(x, rv + dv);
      }
    }
    std::sort(bd[s].begin(), bd[s].end(),
              [](const std::pair<uint32_t, double>& a, const std::pair<uint32_t, double>& b) { return a.first < b.first; });
  }
  auto build = [&](std::vector<std::vector<std::pair<uint32_t, double>>>& rows, std::vector<uint64_t>& off,
                   std::vector<uint32_t>& col, std::vector<double>& w) {
    off.assign(all + 1, 0);
    for (uint32_t u = 0; u < all; ++u) {
      off[u + 1] = off[u] + rows[u].size();
      for (const auto& p : rows[u]) {
        col.push_back(p.first);
        w.push_back(p.second);
      }
    }
  };
  // is_e_: the bound_delta of in-mirror sources and of entry vertices
  std::vector<std::vector<std::pair<uint32_t, double>>> isr(all);
  for (uint32_t m = n; m < all; ++m) isr[m] = bd[m];
  for (uint32_t v = 0; v < n; ++v) {
    const uint8_t t = S.type[v];
    if (t == 2 || t == 3 || t == 6) isr[v] = bd[v];
  }
  build(isr, out.is_off, out.is_col, out.is_w);
  build(iv, out.iv_off, out.iv_col, out.iv_w);
  build(id, out.id_off, out.id_col, out.id_w);
  // ib_e_: Singles' out-lists; exits' edges that leave the cluster and that no
  // mirror carries
  out.ib_off.assign(n + 1, 0);
  for (uint32_t u = 0; u < n; ++u) {
    const uint8_t t = S.type[u];
    if (t == 0) {
      for (uint32_t e = g.ooff[u]; e < g.ooff[u + 1]; ++e) out.ib_col.push_back(g.ocol[e]);
    } else if (t == 1 || t == 3 || t == 5 || t == 6) {
      const uint32_t ids = S.cluster_of[u];
      for (uint32_t e = g.ooff[u]; e < g.ooff[u + 1]; ++e) {
        const uint32_t w = g.ocol[e], ow = S.cluster_of[w];
        if (ow == ids || has(S.out_m[ids], w)) continue;
        if (ow != kNone && has(S.in_m[ow], u)) continue;
        out.ib_col.push_back(w);
      }
    }
    out.ib_off[u + 1] = out.ib_col.size();
  }
  // sync_e_: master -> its in-mirrors, out-mirror -> its master
  std::vector<std::vector<uint32_t>> sync(all);
  for (uint32_t i = 0; i < S.K; ++i) {
    const uint32_t inner = (uint32_t)S.members[i].size(), n_in = (uint32_t)S.in_m[i].size();
    for (uint32_t j = inner; j < S.ids[i].size(); ++j) {
      const uint32_t m = S.ids[i][j], v = S.master[m - n];
      if (j < inner + n_in) sync[v].push_back(m);
      else sync[m].push_back(v);
    }
  }
  out.sync_off.assign(all + 1, 0);
  for (uint32_t u = 0; u < all; ++u) {
    out.sync_off[u + 1] = out.sync_off[u] + sync[u].size();
    out.sync_col.insert(out.sync_col.end(), sync[u].begin(), sync[u].end());
  }
  // the subgraph rows, for the precompute
  out.sub_off.assign(all + 1, 0);
  for (uint32_t u = 0; u < all; ++u) {
    out.sub_off[u + 1] = out.sub_off[u] + S.sub[u].size();
    out.sub_col.insert(out.sub_col.end(), S.sub[u].begin(), S.sub[u].end());
  }
}

// The compressor's state kept on the context for suminc_sketch_download.
struct MirrorState {
  Acceptance acc;
  MirrorSketch S;
  SketchRows rows;
};

void mirror_release(suminc_ctx* c) {
  delete reinterpret_cast<MirrorState*>(c->msk);
  c->msk = nullptr;
}

static void compress_finish(suminc_ctx* c, const HostGraph& g, MirrorState* M, suminc_compress_info* info);

void compress_mirror(suminc_ctx* c, const uint32_t* cluster_of, uint32_t min_node, uint32_t max_node,
                     double threshold, double mirror_k, suminc_compress_info* info) {
  SUMINC_REQUIRE(c->has_graph, SUMINC_E_STATE, "no graph");
  SUMINC_REQUIRE(cluster_of, SUMINC_E_INVALID, "null cluster assignment");
  SUMINC_REQUIRE(max_node >= 1 && min_node <= max_node, SUMINC_E_INVALID, "bad cluster size bounds");
  SUMINC_REQUIRE(mirror_k >= 0.0, SUMINC_E_INVALID, "mirror_k must be >= 0");
  for (uint32_t v = 0; v < c->n; ++v)
    SUMINC_REQUIRE(cluster_of[v] == kNone || cluster_of[v] < c->n, SUMINC_E_INVALID, "cluster id out of range");
  HostGraph g;
  download_graph(c, g);
  MirrorState* M = new MirrorState();
  mirror_release(c);
  c->msk = M;
  accept_clusters(g, cluster_of, min_node, max_node, threshold, mirror_k, M->acc);
  final_build(g, M->acc, M->S);
  compress_finish(c, g, M, info);
}

// ---- the cluster cache in the reference's layout ---------------------------------
// serialize_cluster / de_serialize_cluster (compressor_base.h:2987-3150), all
// little-endian: u32 init_mirror_num (in- plus out-mirrors), u32 cluster count; per
// cluster five lists, each u32 count + u32 ids: supernode_ids, cluster_ids (equal at
// this point), supernode_source, supernode_in_mirror, supernode_out_mirror (the
// masters); then i32 Fc[n] (-(cluster + 1) for members, INT32_MAX otherwise), u32
// id2spids[n] (cluster, or 0xFFFFFFFF), i32 -1.  The reference writes the mirror sets
// in hash order; this writer writes them ascending.
void compress_mirror_save(suminc_ctx* c, const char* path) {
  SUMINC_REQUIRE(c->msk, SUMINC_E_STATE, "no compression on this context (suminc_compress_mirror)");
  SUMINC_REQUIRE(path, SUMINC_E_INVALID, "null path");
  const Acceptance& A = reinterpret_cast<MirrorState*>(c->msk)->acc;
  const uint32_t n = c->n, K = (uint32_t)A.members.size();
  std::vector<uint32_t> out;
  uint64_t mirrors = 0;
  for (uint32_t i = 0; i < K; ++i) mirrors += A.in_m[i].size() + A.out_m[i].size();
  out.push_back((uint32_t)mirrors);
  out.push_back(K);
  auto list = [&](const std::vector<uint32_t>& l) {
    out.push_back((uint32_t)l.size());
    out.insert(out.end(), l.begin(), l.end());
  };
  std::vector<int32_t> fc(n, INT32_MAX);
  std::vector<uint32_t> sp(n, kNone);
  for (uint32_t i = 0; i < K; ++i) {
    list(A.members[i]);
    list(A.members[i]);
    list(A.sources[i]);
    list(A.in_m[i]);
    list(A.out_m[i]);
    for (uint32_t u : A.members[i]) {
      fc[u] = -(int32_t)(i + 1);
      sp[u] = i;
    }
  }
  FILE* f = std::fopen(path, "wb");
  SUMINC_REQUIRE(f, SUMINC_E_INVALID, std::string("cannot open ") + path);
  const int32_t eof = -1;
  bool ok = std::fwrite(out.data(), 4, out.size(), f) == out.size();
  ok = ok && std::fwrite(fc.data(), 4, n, f) == n;
  ok = ok && std::fwrite(sp.data(), 4, n, f) == n;
  ok = ok && std::fwrite(&eof, 4, 1, f) == 1;
  ok = (std::fclose(f) == 0) && ok;
  SUMINC_REQUIRE(ok, SUMINC_E_INVALID, std::string("write failed: ") + path);
}

void compress_mirror_load(suminc_ctx* c, const char* path, suminc_compress_info* info) {
  SUMINC_REQUIRE(c->has_graph, SUMINC_E_STATE, "no graph");
  SUMINC_REQUIRE(path, SUMINC_E_INVALID, "null path");
  const uint32_t n = c->n;
  FILE* f = std::fopen(path, "rb");
  SUMINC_REQUIRE(f, SUMINC_E_INVALID, std::string("cannot open ") + path);
  std::vector<uint32_t> words;
  {
    uint32_t buf[4096];
    size_t got;
    while ((got = std::fread(buf, 4, 4096, f)) > 0) words.insert(words.end(), buf, buf + got);
    std::fclose(f);
  }
  size_t at = 0;
  auto next = [&]() -> uint32_t {
    SUMINC_REQUIRE(at < words.size(), SUMINC_E_INVALID, std::string("truncated cluster cache: ") + path);
    return words[at++];
  };
  Acceptance A;
  const uint32_t mirrors = next(), K = next();
  SUMINC_REQUIRE(K <= n, SUMINC_E_INVALID, "cluster cache: more clusters than vertices");
  std::vector<uint32_t> owner(n, kNone);
  uint64_t mirror_count = 0;
  auto list = [&](std::vector<uint32_t>& l) {
    const uint32_t cnt = next();
    SUMINC_REQUIRE(cnt <= n, SUMINC_E_INVALID, "cluster cache: list longer than the vertex count");
    l.resize(cnt);
    for (uint32_t k = 0; k < cnt; ++k) {
      l[k] = next();
      SUMINC_REQUIRE(l[k] < n, SUMINC_E_INVALID, "cluster cache: vertex id out of range");
    }
    std::sort(l.begin(), l.end());
    SUMINC_REQUIRE(std::adjacent_find(l.begin(), l.end()) == l.end(), SUMINC_E_INVALID,
                   "cluster cache: repeated id in a list");
  };
  A.members.resize(K);
  A.sources.resize(K);
  A.in_m.resize(K);
  A.out_m.resize(K);
  for (uint32_t i = 0; i < K; ++i) {
    std::vector<uint32_t> cl;
    list(A.members[i]);
    list(cl);
    list(A.sources[i]);
    list(A.in_m[i]);
    list(A.out_m[i]);
    SUMINC_REQUIRE(cl == A.members[i], SUMINC_E_INVALID, "cluster cache: cluster_ids differ from supernode_ids");
    for (uint32_t u : A.members[i]) {
      SUMINC_REQUIRE(owner[u] == kNone, SUMINC_E_INVALID, "cluster cache: a vertex in two clusters");
      owner[u] = i;
    }
    mirror_count += A.in_m[i].size() + A.out_m[i].size();
  }
  for (uint32_t i = 0; i < K; ++i) {
    for (uint32_t s : A.sources[i])
      SUMINC_REQUIRE(owner[s] == i, SUMINC_E_INVALID, "cluster cache: a source outside its cluster");
    for (uint32_t v : A.in_m[i])
      SUMINC_REQUIRE(owner[v] != i, SUMINC_E_INVALID, "cluster cache: a mirror of its own cluster");
    for (uint32_t v : A.out_m[i])
      SUMINC_REQUIRE(owner[v] != i, SUMINC_E_INVALID, "cluster cache: a mirror of its own cluster");
  }
  SUMINC_REQUIRE(mirror_count == mirrors, SUMINC_E_INVALID, "cluster cache: init_mirror_num differs from the lists");
  for (uint32_t v = 0; v < n; ++v) {
    const int32_t fc = (int32_t)next();
    SUMINC_REQUIRE(owner[v] == kNone ? fc == INT32_MAX : fc == -(int32_t)(owner[v] + 1), SUMINC_E_INVALID,
                   "cluster cache: Fc differs from the lists");
  }
  for (uint32_t v = 0; v < n; ++v)
    SUMINC_REQUIRE(next() == owner[v], SUMINC_E_INVALID, "cluster cache: id2spids differs from the lists");
  SUMINC_REQUIRE(next() == 0xFFFFFFFFu && at == words.size(), SUMINC_E_INVALID,
                 "cluster cache: missing end marker or trailing data");
  HostGraph g;
  download_graph(c, g);
  MirrorState* M = new MirrorState();
  mirror_release(c);
  c->msk = M;
  M->acc = std::move(A);
  final_build(g, M->acc, M->S);
  compress_finish(c, g, M, info);
}

// the device shortcut rows, the CSR products and their upload; the info
static void compress_finish(suminc_ctx* c, const HostGraph& g, MirrorState* M, suminc_compress_info* info) {
  MirrorSketch& S = M->S;
  // full out-degrees (a mirror: its master's) and the device rows
  std::vector<uint32_t> outdeg(S.all);
  for (uint32_t v = 0; v < S.n; ++v) outdeg[v] = g.outdeg(v);
  for (uint32_t m = S.n; m < S.all; ++m) outdeg[m] = g.outdeg(S.master[m - S.n]);
  std::vector<uint64_t> base;
  std::vector<double> reach, resid;
  const int maxit = c->prm.shortcut_max_rounds > 0 ? c->prm.shortcut_max_rounds : 100;
  uint64_t work = 0;
  // the reference's per-shortcut stop rule: moved mass <= T / N * |members|
  mirror_rows(c, S, outdeg, c->prm.termcheck_threshold / (double)S.n, maxit, base, reach, resid, &work);
  sketch_rows(g, S, base, reach, resid, M->rows);
  SketchRows& R = M->rows;
  suminc_sketch_csr sk{};
  sk.n = S.n;
  sk.all_node_num = S.all;
  sk.node_type = S.type.data();
  sk.mirror_master = S.master.data();
  sk.is_off = R.is_off.data(); sk.is_col = R.is_col.data();
  sk.ib_off = R.ib_off.data(); sk.ib_col = R.ib_col.data();
  sk.sync_off = R.sync_off.data(); sk.sync_col = R.sync_col.data();
  sk.iv_off = R.iv_off.data(); sk.iv_col = R.iv_col.data();
  sk.id_off = R.id_off.data(); sk.id_col = R.id_col.data();
  sk.sub_off = R.sub_off.data(); sk.sub_col = R.sub_col.data();
  rsk_upload(c, &sk, R.is_w.data(), R.iv_w.data(), R.id_w.data());  // FP64 rows
  if (info) {
    std::memset(info, 0, sizeof(*info));
    info->clusters = S.K;
    info->all_node_num = S.all;
    info->in_mirrors = S.in_mirrors;
    info->out_mirrors = S.out_mirrors;
    info->absorbed = S.absorbed;
    info->supernodes = S.sources.size();
    for (uint32_t v = 0; v < S.n; ++v) info->node_type_count[S.type[v]] += 1;
    info->is_edges = R.is_col.size();
    info->ib_edges = R.ib_col.size();
    info->sync_edges = R.sync_col.size();
    info->inner_value_entries = R.iv_col.size();
    info->inner_delta_entries = R.id_col.size();
    info->subgraph_edges = R.sub_col.size();
    info->shortcut_work = work;
  }
}

void sketch_csr_download(suminc_ctx* c, suminc_sketch_buffers* b) {
  SUMINC_REQUIRE(c->msk, SUMINC_E_STATE, "no compression on this context (suminc_compress_mirror)");
  SUMINC_REQUIRE(b, SUMINC_E_INVALID, "null buffers");
  const MirrorState& M = *reinterpret_cast<MirrorState*>(c->msk);
  const MirrorSketch& S = M.S;
  const SketchRows& R = M.rows;
  auto cp = [](void* dst, const void* src, size_t bytes) {
    if (dst && bytes) std::memcpy(dst, src, bytes);
  };
  cp(b->node_type, S.type.data(), S.type.size());
  cp(b->mirror_master, S.master.data(), S.master.size() * 4);
  cp(b->cluster_of, S.cluster_of.data(), S.cluster_of.size() * 4);
  cp(b->is_off, R.is_off.data(), R.is_off.size() * 8);
  cp(b->is_col, R.is_col.data(), R.is_col.size() * 4);
  auto cpw = [](float* dst, const std::vector<double>& src) {
    if (dst)
      for (size_t i = 0; i < src.size(); ++i) dst[i] = (float)src[i];
  };
  cpw(b->is_w, R.is_w);
  cp(b->ib_off, R.ib_off.data(), R.ib_off.size() * 8);
  cp(b->ib_col, R.ib_col.data(), R.ib_col.size() * 4);
  cp(b->sync_off, R.sync_off.data(), R.sync_off.size() * 8);
  cp(b->sync_col, R.sync_col.data(), R.sync_col.size() * 4);
  cp(b->iv_off, R.iv_off.data(), R.iv_off.size() * 8);
  cp(b->iv_col, R.iv_col.data(), R.iv_col.size() * 4);
  cpw(b->iv_w, R.iv_w);
  cp(b->id_off, R.id_off.data(), R.id_off.size() * 8);
  cp(b->id_col, R.id_col.data(), R.id_col.size() * 4);
  cpw(b->id_w, R.id_w);
  cp(b->sub_off, R.sub_off.data(), R.sub_off.size() * 8);
  cp(b->sub_col, R.sub_col.data(), R.sub_col.size() * 4);
}

}  // namespace suminc
