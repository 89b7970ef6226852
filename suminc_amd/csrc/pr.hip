// pr.hip -- plain delta-accumulative PageRank on the device (the reference's
// Ingress phase: sum_sync_iter_worker.h:857-882 on CPU; tjn::pr_Ingress +
// tjn::deltaSum on GPU, cudaCode/my_worker.cu:124-142, 304-328).
//
// One round, all launches on the context stream, no host synchronisation:
//   k_pr_count   vertex pass over delta only: FP64 sum|delta| (deterministic
//                per-block partials) and the out-degree sum of active vertices
//                (|delta| > T/N, the reference's isChange).
//   k_pr_decide  one block: residual = fixed-order sum of the partials; stop if
//                residual < T (termCheck) or the round cap is reached; else pick
//                the scatter for this round: pull when the active edges exceed
//                m / pull_threshold_div, push otherwise.
//   k_pr_gather  exchange delta -> 0 for active vertices, rank += delta (Kahan
//                compensated in FP32 mode); push: append (v, d*delta/deg) to the
//                light list (wave64 ballot, one atomic per wave) or 2048-edge
//                chunks of high-degree vertices to the heavy list; pull: write the
//                contribution of every vertex (0 when inactive).
//   k_pr_push    persistent blocks over light tiles (block-cooperative adjacency
//                walk, coalesced column reads) then heavy chunks; scatter with the
//                hardware float/double atomic add.
//   k_pr_pull_*  CSC gather by in-degree class (lists built once per graph
//                version): a thread per row up to 32 in-edges, a wave per row up
//                to 256, above that 2048-edge chunks summed per workgroup and
//                added with one atomic.  Above 32 MiB of contributions the gather
//                reads a compact array of the vertices with out-edges, ordered by
//                decreasing out-degree (CSC columns renamed once per version);
//                an optional banded variant pulls one source band at a time.
//   k_pr_end     list counters reset, statistics.
// Kernels not selected for the round return at entry.
#include <cstring>
#include <rocprim/rocprim.hpp>

#include <algorithm>

#include "ctx.hpp"
#include "device.hpp"
#include "pr.hpp"
#include "push.hpp"

namespace suminc {

constexpr int kCountPer = 4;  // vertices per thread of k_pr_count and k_pr_gather

template <class T>
__global__ void __launch_bounds__(kBlock) k_pr_count(uint32_t n, const uint32_t* __restrict__ off,
                                                     const T* __restrict__ delta, double eps, Counters* cnt,
                                                     double* __restrict__ partial) {
  __shared__ double s_sum[kBlock / 64];
  __shared__ unsigned long long s_edges[kBlock / 64];
  if (cnt->done) return;
  // kCountPer vertices a thread (coalesced: vertex base + j * kBlock + t), their
  // residual loads issued before any is used
  const uint32_t base = blockIdx.x * (kBlock * kCountPer);
  T dl[kCountPer];
#pragma unroll
  for (int j = 0; j < kCountPer; ++j) {
    const uint32_t v = base + j * kBlock + threadIdx.x;
    dl[j] = v < n ? delta[v] : T(0);
  }
  double absd = 0.0;
  unsigned long long deg = 0;
#pragma unroll
  for (int j = 0; j < kCountPer; ++j) {
    const uint32_t v = base + j * kBlock + threadIdx.x;
    const double a = (double)fabs_t(dl[j]);
    absd += a;
    if (v < n && a > eps) deg += off[v + 1] - off[v];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    absd += __shfl_xor(absd, o);
    deg += __shfl_xor(deg, o);
  }
  const int wid = threadIdx.x >> 6;
  if (lane_id() == 0) {
    s_sum[wid] = absd;
    s_edges[wid] = deg;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    unsigned long long te = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
      t += s_sum[w];
      te += s_edges[w];
    }
    partial[blockIdx.x] = t;
    reinterpret_cast<unsigned long long*>(partial)[gridDim.x + blockIdx.x] = te;
  }
}

__global__ void __launch_bounds__(1024) k_pr_decide(const double* __restrict__ partial, uint32_t nparts,
                                                    Counters* cnt, double T, int max_rounds, uint64_t m, int div,
                                                    uint8_t* mode_hist) {
  __shared__ double s[1024];
  __shared__ unsigned long long se[1024];
  if (cnt->done) return;
  const unsigned long long* pe = reinterpret_cast<const unsigned long long*>(partial) + nparts;
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
    cnt->round_edges = se[0];
    cnt->residual = r;
    if (r != r || r > 1e300) cnt->aux[0] = 1;  // numeric failure flag
    if (!(r >= T) || (long long)cnt->rounds >= (long long)max_rounds || cnt->aux[0]) {
      cnt->done = 1;
      cnt->round_edges = 0;
      return;
    }
    const bool pull = div > 0 && cnt->round_edges * (unsigned long long)div > m;
    cnt->mode = pull ? 1 : 0;
    if (cnt->rounds < 65536) mode_hist[cnt->rounds] = pull ? 1 : 0;
    cnt->rounds += 1;
  }
}

template <class T>
__global__ void __launch_bounds__(kBlock) k_pr_gather(uint32_t n, const uint32_t* __restrict__ off, T* __restrict__ delta,
                                                      T* __restrict__ rank, float* __restrict__ comp, double eps, T d,
                                                      uint32_t* __restrict__ wl_v, T* __restrict__ wl_c,
                                                      uint4* __restrict__ heavy, T* __restrict__ contrib,
                                                      const uint32_t* __restrict__ pos, Counters* cnt,
                                                      unsigned long long* __restrict__ act_partial) {
  __shared__ unsigned long long s_act[kBlock / 64];
  if (cnt->done) return;
  const bool pull = cnt->mode != 0;
  // kCountPer vertices a thread (vertex base + j * kBlock + t, coalesced), the
  // residual loads issued first; every lane runs every j (the wave appends)
  const uint32_t base = blockIdx.x * (kBlock * kCountPer);
  T dl[kCountPer];
#pragma unroll
  for (int j = 0; j < kCountPer; ++j) {
    const uint32_t v = base + j * kBlock + threadIdx.x;
    dl[j] = v < n ? delta[v] : T(0);
  }
  unsigned long long nact = 0;
#pragma unroll
  for (int j = 0; j < kCountPer; ++j) {
    const uint32_t v = base + j * kBlock + threadIdx.x;
    bool light = false;
    T c = T(0);
    if (v < n) {
      const T dv = dl[j];
      if ((double)fabs_t(dv) > eps) {
        ++nact;
        delta[v] = T(0);
        if constexpr (sizeof(T) == 4) {
          float r = rank[v], cc = comp[v];
          float y = (float)dv - cc;
          float t = r + y;
          comp[v] = (t - r) - y;
          rank[v] = t;
        } else {
          rank[v] += dv;
        }
        uint32_t b = off[v], e = off[v + 1];
        uint32_t deg = e - b;
        if (deg > 0) {
          c = d * dv / (T)deg;
          if (!pull) {
            if (deg <= kHeavyDeg) {
              light = true;
            } else {
              uint32_t nch = (deg + kChunk - 1) / kChunk;
              unsigned long long slot = atomicAdd(&cnt->heavy_count, (unsigned long long)nch);
              for (uint32_t k = 0; k < nch; ++k) {
                uint4 h;
                h.x = v;
                h.y = k;
                if constexpr (sizeof(T) == 4) {
                  h.z = __float_as_uint((float)c);
                  h.w = 0;
                } else {
                  unsigned long long bits = __double_as_longlong((double)c);
                  h.z = (uint32_t)bits;
                  h.w = (uint32_t)(bits >> 32);
                }
                heavy[slot + k] = h;
              }
            }
          }
        }
      }
      if (pull) {  // the contribution of every vertex (pos: of every vertex with out-edges, at its compact position)
        const uint32_t p = pos ? pos[v] : v;
        if (p != INF32) contrib[p] = c;
      }
    }
    if (!pull) {
      unsigned long long slot = wave_append(light, &cnt->wl_count[0]);
      if (light) {
        wl_v[slot] = v;
        wl_c[slot] = c;
      }
    }
  }
  unsigned long long a = nact;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o);
  if (lane_id() == 0) s_act[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < kBlock / 64; ++w) t += s_act[w];
    act_partial[blockIdx.x] = t;
  }
}

// Whether the pull launches of this point in the stream run: gate 0, a round
// that chose pull (not stopped); gate 1, a reseed that chose pull.
__device__ __forceinline__ bool pull_runs(const Counters* cnt, int gate) {
  return gate ? cnt->reseed_mode == 1 : (!cnt->done && cnt->mode == 1);
}

// Sum of contrib over the in-edges k = k0 + r * stride (r < U, k < e).  All U
// column loads are issued before the U gathers, so a thread keeps U independent
// gathers in flight instead of one dependent (column, gather) pair per edge: the
// gathers are cache-miss latency bound, not bandwidth bound.
template <class T, int U>
__device__ __forceinline__ T gather_sum(const uint32_t* __restrict__ icol, const T* __restrict__ contrib, uint32_t k0,
                                        uint32_t stride, uint32_t e) {
  uint32_t u[U];
#pragma unroll
  for (int r = 0; r < U; ++r) {
    const uint32_t k = k0 + r * stride;
    u[r] = k < e ? icol[k] : INF32;
  }
  T x[U];
#pragma unroll
  for (int r = 0; r < U; ++r) x[r] = u[r] != INF32 ? contrib[u[r]] : T(0);
  T acc = T(0);
#pragma unroll
  for (int r = 0; r < U; ++r) acc += x[r];
  return acc;
}

// Pull, short rows (in-degree 1..kPullThread): one thread per vertex sums its
// in-edges in order; no atomics (each vertex has one owner).  bid / nb: this
// workgroup's index among the nb workgroups of the part.
template <class T>
__device__ __forceinline__ void pull_short_part(uint32_t bid, uint32_t nb, const uint32_t* __restrict__ list,
                                                uint32_t count, const uint32_t* __restrict__ ioff,
                                                const uint32_t* __restrict__ icol, const T* __restrict__ contrib,
                                                T* __restrict__ delta) {
  for (uint32_t i = bid * kBlock + threadIdx.x; i < count; i += nb * kBlock) {
    const uint32_t v = list[i];
    const uint32_t b = ioff[v], e = ioff[v + 1];
    T acc = T(0);
    for (uint32_t k = b; k < e; k += 4) acc += gather_sum<T, 4>(icol, contrib, k, 1, e);
    if (acc != T(0)) delta[v] += acc;
  }
}

// Pull, medium rows (kPullThread < in-degree <= kHeavyDeg): one wave per vertex,
// lanes stride the row (coalesced column reads), shuffle reduction.
template <class T>
__device__ __forceinline__ void pull_wave_part(uint32_t bid, uint32_t nb, const uint32_t* __restrict__ list,
                                               uint32_t count, const uint32_t* __restrict__ ioff,
                                               const uint32_t* __restrict__ icol, const T* __restrict__ contrib,
                                               T* __restrict__ delta) {
  const uint32_t waves = nb * (kBlock / 64);
  for (uint32_t i = bid * (kBlock / 64) + (threadIdx.x >> 6); i < count; i += waves) {
    const uint32_t v = list[i];
    const uint32_t b = ioff[v], e = ioff[v + 1];
    T acc = T(0);
    for (uint32_t k = b + lane_id(); k < e; k += 4 * 64)  // one pass for rows up to 256
      acc += gather_sum<T, 4>(icol, contrib, k, 64, e);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if (lane_id() == 0 && acc != T(0)) delta[v] += acc;
  }
}

// Pull, heavy part: 2048-edge chunks of vertices with in-degree > 256, one
// workgroup per chunk, one atomic per chunk.
template <class T>
__device__ __forceinline__ void pull_heavy_part(uint32_t bid, uint32_t nb, T* s_red, const uint2* __restrict__ chunks,
                                                uint64_t nchunks, const uint32_t* __restrict__ ioff,
                                                const uint32_t* __restrict__ icol, const T* __restrict__ contrib,
                                                T* __restrict__ delta) {
  for (uint64_t h = bid; h < nchunks; h += nb) {
    const uint2 ch = chunks[h];
    const uint32_t b = ioff[ch.x] + ch.y * kChunk;
    const uint32_t e = min(ioff[ch.x + 1], b + kChunk);
    T acc = gather_sum<T, kChunk / kBlock>(icol, contrib, b + threadIdx.x, kBlock, e);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if (lane_id() == 0) s_red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      T s = T(0);
      for (int w = 0; w < kBlock / 64; ++w) s += s_red[w];
      if (s != T(0)) unsafeAtomicAdd(&delta[ch.x], s);
    }
    __syncthreads();
  }
}

// One launch for the three row classes: workgroups [0, gh) take the heavy chunks
// (the longest work first), [gh, gh + gw) the wave rows, the rest the short rows,
// so the classes' tails overlap instead of following each other.
template <class T>
__global__ void __launch_bounds__(kBlock) k_pr_pull(const uint32_t* __restrict__ short_list,
                                                    const uint32_t* __restrict__ wave_list,
                                                    const uint2* __restrict__ chunks,
                                                    const unsigned long long* __restrict__ counts,
                                                    const uint32_t* __restrict__ ioff,
                                                    const uint32_t* __restrict__ icol, const T* __restrict__ contrib,
                                                    T* __restrict__ delta, const Counters* cnt, int gate, uint32_t gh,
                                                    uint32_t gw) {
  __shared__ T s_red[kBlock / 64];
  if (!pull_runs(cnt, gate)) return;
  const uint32_t bid = blockIdx.x;
  if (bid < gh)
    pull_heavy_part<T>(bid, gh, s_red, chunks, counts[2], ioff, icol, contrib, delta);
  else if (bid < gh + gw)
    pull_wave_part<T>(bid - gh, gw, wave_list, (uint32_t)counts[1], ioff, icol, contrib, delta);
  else
    pull_short_part<T>(bid - gh - gw, gridDim.x - gh - gw, short_list, (uint32_t)counts[0], ioff, icol, contrib,
                       delta);
}

__global__ void __launch_bounds__(1024) k_pr_end(Counters* cnt, const unsigned long long* __restrict__ act_partial,
                                                 uint32_t nparts) {
  __shared__ unsigned long long s[1024];
  if (cnt->done) return;
  unsigned long long a = 0;
  for (uint32_t i = threadIdx.x; i < nparts; i += 1024) a += act_partial[i];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x) return;
  cnt->active_vertices += s[0];
  cnt->edges += cnt->round_edges;
  if (cnt->mode) {
    cnt->pull_rounds += 1;
  } else {
    cnt->push_edges += cnt->round_edges;
    cnt->push_entries += cnt->wl_count[0] + cnt->heavy_count;
  }
  cnt->round_edges = 0;
  cnt->wl_count[0] = 0;
  cnt->heavy_count = 0;
}

// Static pull work lists, rebuilt per graph version, in vertex order and without
// atomics: vertices by in-degree class (short: 1..kPullThread, wave: ..kHeavyDeg)
// and the 2048-edge chunks of the longer rows.  Three passes: class counts per
// block, a one-block exclusive scan of the counts (totals to counts[0..2]), and
// the ordered fill.
__device__ __forceinline__ void pull_classes(uint32_t deg, uint32_t& s, uint32_t& w, uint32_t& h) {
  s = deg >= 1 && deg <= kPullThread;
  w = deg > kPullThread && deg <= kHeavyDeg;
  h = deg > kHeavyDeg ? (deg + kChunk - 1) / kChunk : 0u;
}

__global__ void __launch_bounds__(kBlock) k_pull_count(uint32_t n, const uint32_t* __restrict__ ioff,
                                                       uint32_t* __restrict__ parts) {
  __shared__ uint32_t s_w[3][kBlock / 64];
  const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
  uint32_t cs = 0, cw = 0, ch = 0;
  if (v < n) pull_classes(ioff[v + 1] - ioff[v], cs, cw, ch);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    cs += __shfl_xor(cs, o);
    cw += __shfl_xor(cw, o);
    ch += __shfl_xor(ch, o);
  }
  const int wid = threadIdx.x >> 6;
  if (lane_id() == 0) {
    s_w[0][wid] = cs;
    s_w[1][wid] = cw;
    s_w[2][wid] = ch;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    uint32_t t = 0;
    for (int w = 0; w < kBlock / 64; ++w) t += s_w[threadIdx.x][w];
    parts[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = t;
  }
}

// One block: each thread scans a contiguous run of the g block counts of a class.
__global__ void __launch_bounds__(1024) k_pull_scan(uint32_t g, uint32_t* __restrict__ parts,
                                                    unsigned long long* __restrict__ counts) {
  __shared__ uint32_t s_warp[1024 / 64 + 1];
  const uint32_t per = (g + 1023) / 1024;
  const uint32_t b = min(g, threadIdx.x * per), e = min(g, b + per);
  for (int k = 0; k < 3; ++k) {
    uint32_t* p = parts + (size_t)k * g;
    uint32_t sum = 0;
    for (uint32_t i = b; i < e; ++i) sum += p[i];
    uint32_t total;
    uint32_t run = block_exclusive_scan<1024>(sum, s_warp, &total);
    for (uint32_t i = b; i < e; ++i) {
      const uint32_t x = p[i];
      p[i] = run;
      run += x;
    }
    if (threadIdx.x == 0) counts[k] = total;
    __syncthreads();  // s_warp is reused by the next class
  }
}

__global__ void __launch_bounds__(kBlock) k_pull_fill(uint32_t n, const uint32_t* __restrict__ ioff,
                                                      const uint32_t* __restrict__ parts,
                                                      uint32_t* __restrict__ short_list,
                                                      uint32_t* __restrict__ wave_list, uint2* __restrict__ heavy) {
  __shared__ uint32_t s_warp[kBlock / 64 + 1];
  const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
  uint32_t cs = 0, cw = 0, ch = 0, tot;
  if (v < n) pull_classes(ioff[v + 1] - ioff[v], cs, cw, ch);
  const uint32_t ps = block_exclusive_scan<kBlock>(cs, s_warp, &tot);
  __syncthreads();
  const uint32_t pw = block_exclusive_scan<kBlock>(cw, s_warp, &tot);
  __syncthreads();
  const uint32_t ph = block_exclusive_scan<kBlock>(ch, s_warp, &tot);
  const uint32_t g = gridDim.x;
  if (cs) short_list[parts[blockIdx.x] + ps] = v;
  if (cw) wave_list[parts[g + blockIdx.x] + pw] = v;
  for (uint32_t k = 0; k < ch; ++k) heavy[parts[2 * g + blockIdx.x] + ph + k] = make_uint2(v, k);
}

template <class T>
__global__ void k_pr_init(uint32_t n, T* delta, T* rank, float* comp, T init_delta) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  delta[v] = init_delta;
  rank[v] = T(0);
  if (comp) comp[v] = 0.f;
}

template <class T>
__global__ void k_pr_to_f64(uint32_t n, const T* delta, const T* rank, const float* comp, double* d_out,
                            double* r_out) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  d_out[v] = (double)delta[v];
  double r = (double)rank[v];
  if (comp) r -= (double)comp[v];
  r_out[v] = r;
}

template <class T>
__global__ void k_pr_from_f64(uint32_t n, T* delta, T* rank, float* comp, const double* d_in, const double* r_in) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  delta[v] = (T)d_in[v];
  if (comp) {
    float hi = (float)r_in[v];
    rank[v] = (T)hi;
    comp[v] = (float)((double)hi - r_in[v]);  // rank - comp == r_in to ~2^-48 relative
  } else {
    rank[v] = (T)r_in[v];
  }
}

// FP64 mass balance of the state: per block of a fixed grid, the sums of
// r + delta and of d * r over the vertices with out-edges (r = rank - comp in
// FP32/Kahan mode); k_pr_mass_end adds the block sums in a fixed order.
constexpr unsigned kMassGrid = 1024;
template <class T>
__global__ void __launch_bounds__(kBlock) k_pr_mass(uint32_t n, const uint32_t* __restrict__ off,
                                                    const T* __restrict__ delta, const T* __restrict__ rank,
                                                    const float* __restrict__ comp, double d,
                                                    double* __restrict__ partial) {
  __shared__ double s_tot[kBlock / 64], s_out[kBlock / 64];
  double tot = 0.0, out = 0.0;
  for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < n; v += kMassGrid * kBlock) {
    double r = (double)rank[v];
    if (comp) r -= (double)comp[v];
    tot += r + (double)delta[v];
    if (off[v + 1] > off[v]) out += d * r;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    tot += __shfl_xor(tot, o);
    out += __shfl_xor(out, o);
  }
  if (lane_id() == 0) {
    s_tot[threadIdx.x >> 6] = tot;
    s_out[threadIdx.x >> 6] = out;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0, b = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
      a += s_tot[w];
      b += s_out[w];
    }
    partial[blockIdx.x] = a;
    partial[kMassGrid + blockIdx.x] = b;
  }
}

__global__ void __launch_bounds__(1024) k_pr_mass_end(const double* __restrict__ partial, double* __restrict__ out) {
  __shared__ double s[1024], so[1024];
  s[threadIdx.x] = partial[threadIdx.x];
  so[threadIdx.x] = partial[kMassGrid + threadIdx.x];
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      s[threadIdx.x] += s[threadIdx.x + o];
      so[threadIdx.x] += so[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = s[0];
    out[1] = so[0];
  }
}

__global__ void k_zero_counters(Counters* cnt) {
  if (threadIdx.x == 0) {
    Counters z{};
    *cnt = z;
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
unsigned push_grid() { return 2048; }

// Contribution index of the plain pull: compact source positions when in use
// (nullptr: vertex ids), and the CSC columns the plain pull reads.
const uint32_t* pull_pos(suminc_ctx* c) { return c->pull_compact ? c->src_pos.as<uint32_t>() : nullptr; }
const uint32_t* pull_cols(suminc_ctx* c) {
  return c->pull_compact ? c->in_col_c.as<uint32_t>() : c->in_col.as<uint32_t>();
}

void pr_alloc(suminc_ctx* c) {
  const size_t n = c->n;
  const size_t es = c->pr_prec == SUMINC_PR_FP64 ? 8 : 4;
  c->pr_delta.ensure(n * es);
  c->pr_rank.ensure(n * es);
  if (c->pr_prec == SUMINC_PR_FP32_KAHAN) c->pr_comp.ensure(n * 4);
  c->pr_contrib.ensure(n * es + 8);
  c->wl0.ensure(n * 4 + 4);
  c->wl1.ensure(n * es + 8);
  // heavy chunks: at most m / kChunk + (number of heavy vertices) entries
  c->heavy.ensure_slack((c->m / kChunk + c->m / kHeavyDeg + 2) * sizeof(uint4));
  c->partial.ensure(((size_t)grid_for(n) + 1) * 3 * 8);  // residual | edges | active vertices
  c->tmp4.ensure(65536);  // per-round scatter mode history
}

void pr_init(suminc_ctx* c) {
  SUMINC_REQUIRE(c->has_graph, SUMINC_E_STATE, "no graph");
  c->pr_prec = c->prm.pr_precision;
  pr_alloc(c);
  const double d = c->prm.damping;
  const double init = (1.0 - d) / (double)c->n;
  if (c->pr_prec == SUMINC_PR_FP64)
    k_pr_init<double><<<grid_for(c->n), kBlock, 0, c->stream>>>(c->n, c->pr_delta.as<double>(),
                                                               c->pr_rank.as<double>(), nullptr, init);
  else
    k_pr_init<float><<<grid_for(c->n), kBlock, 0, c->stream>>>(c->n, c->pr_delta.as<float>(), c->pr_rank.as<float>(),
                                                              c->pr_comp.as<float>(), (float)init);
  SUMINC_HIP(hipGetLastError());
  c->pr_ready = true;
  ++c->pr_state_id;
}

// ---------------------------------------------------------------------------
// Banded pull.  A dense round gathers contrib[u] for every in-edge; once the
// contributions outgrow the last-level cache each gather costs a DRAM sector.
// The CSC rows are sorted by source, so a source band [p*B, (p+1)*B) is one
// sub-range of every row: the wave and heavy rows are pulled band by band, each
// band's slice of contrib staying cache resident while all workgroups read it.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t row_lower_bound(const uint32_t* __restrict__ icol, uint32_t b, uint32_t e,
                                                    uint64_t x) {
  while (b < e) {
    const uint32_t mid = b + ((e - b) >> 1);
    if ((uint64_t)icol[mid] < x) b = mid + 1; else e = mid;
  }
  return b;
}

// split[p * W + i]: first in-edge of listed row i with source >= p * B (p = 0..P)
__global__ void k_band_split(const uint32_t* __restrict__ wave_list, uint32_t W, const uint32_t* __restrict__ ioff,
                             const uint32_t* __restrict__ icol, uint32_t P, uint32_t B, uint32_t* __restrict__ split) {
  const uint64_t total = (uint64_t)W * (P + 1);
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t p = (uint32_t)(t / W), i = (uint32_t)(t % W);
    const uint32_t v = wave_list[i];
    const uint32_t b = ioff[v], e = ioff[v + 1];
    split[t] = p == 0 ? b : (p == P ? e : row_lower_bound(icol, b, e, (uint64_t)p * B));
  }
}

// Band p of heavy row (chunk entry h with k == 0): its in-edge range and chunks.
__device__ __forceinline__ uint32_t band_range(const uint2* chunks, uint32_t h, uint32_t p, uint32_t P, uint32_t B,
                                               const uint32_t* ioff, const uint32_t* icol, uint32_t& v,
                                               uint32_t& lo, uint32_t& hi) {
  const uint2 ch = chunks[h];
  if (ch.y != 0) return 0;
  v = ch.x;
  const uint32_t b = ioff[v], e = ioff[v + 1];
  lo = p == 0 ? b : row_lower_bound(icol, b, e, (uint64_t)p * B);
  hi = p + 1 == P ? e : row_lower_bound(icol, b, e, (uint64_t)(p + 1) * B);
  return (hi - lo + kChunk - 1) / kChunk;
}

__global__ void k_band_heavy_count(const uint2* __restrict__ chunks, uint32_t Hc, const uint32_t* __restrict__ ioff,
                                   const uint32_t* __restrict__ icol, uint32_t P, uint32_t B,
                                   uint32_t* __restrict__ cnt) {
  const uint64_t total = (uint64_t)Hc * P;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t v, lo, hi;
    cnt[t] = band_range(chunks, (uint32_t)(t % Hc), (uint32_t)(t / Hc), P, B, ioff, icol, v, lo, hi);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) cnt[total] = 0;
}

// Band-major chunk list {row, first edge, end edge}; off[p] = first chunk of band p.
__global__ void k_band_heavy_fill(const uint2* __restrict__ chunks, uint32_t Hc, const uint32_t* __restrict__ ioff,
                                  const uint32_t* __restrict__ icol, uint32_t P, uint32_t B,
                                  const uint32_t* __restrict__ pre, uint4* __restrict__ out, uint32_t* __restrict__ off) {
  const uint64_t total = (uint64_t)Hc * P;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t h = (uint32_t)(t % Hc), p = (uint32_t)(t / Hc);
    if (h == 0) off[p] = pre[t];
    uint32_t v = 0, lo = 0, hi = 0;
    const uint32_t nch = band_range(chunks, h, p, P, B, ioff, icol, v, lo, hi);
    for (uint32_t j = 0; j < nch; ++j)
      out[pre[t] + j] = make_uint4(v, lo + j * kChunk, min(hi, lo + (j + 1) * kChunk), 0u);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) off[P] = pre[total];
}

// Band p of the wave rows: tiles of 256 rows, the block walks their band slices
// with coalesced column reads and sums per row in LDS, one delta update per row.
template <class T>
__global__ void __launch_bounds__(kBlock) k_pr_pull_band_rows(const uint32_t* __restrict__ list, uint32_t W,
                                                              const uint32_t* __restrict__ split, uint32_t p,
                                                              const uint32_t* __restrict__ icol,
                                                              const T* __restrict__ contrib, T* __restrict__ delta,
                                                              const Counters* cnt, int gate) {
  __shared__ uint32_t s_scan[kBlock + 1];
  __shared__ uint32_t s_start[kBlock];
  __shared__ uint32_t s_warp[kBlock / 64 + 1];
  __shared__ T s_acc[kBlock];
  if (!pull_runs(cnt, gate)) return;
  const uint32_t tiles = (W + kBlock - 1) / kBlock;
  for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {
    const uint32_t i = t * kBlock + threadIdx.x;
    uint32_t lo = 0, len = 0, v = 0;
    if (i < W) {
      v = list[i];
      lo = split[(size_t)p * W + i];
      len = split[(size_t)(p + 1) * W + i] - lo;
    }
    s_acc[threadIdx.x] = T(0);
    uint32_t total;
    s_scan[threadIdx.x] = block_exclusive_scan<kBlock>(len, s_warp, &total);  // syncs: s_acc is clear
    s_start[threadIdx.x] = lo;
    if (threadIdx.x == 0) s_scan[kBlock] = total;
    __syncthreads();
    // four edges per thread at a time, loads first (as gather_sum)
    for (uint32_t j0 = threadIdx.x; j0 < total; j0 += 4 * kBlock) {
      int slot[4];
      uint32_t u[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const uint32_t j = j0 + r * kBlock;
        slot[r] = -1;
        u[r] = INF32;
        if (j < total) {
          int a = 0, z = kBlock - 1;  // owner: last row whose first edge is <= j
          while (a < z) {
            const int mid = (a + z + 1) >> 1;
            if (s_scan[mid] <= j) a = mid; else z = mid - 1;
          }
          slot[r] = a;
          u[r] = icol[s_start[a] + (j - s_scan[a])];
        }
      }
      T x[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) x[r] = slot[r] >= 0 ? contrib[u[r]] : T(0);
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (slot[r] >= 0) atomicAdd(&s_acc[slot[r]], x[r]);
    }
    __syncthreads();
    if (i < W && s_acc[threadIdx.x] != T(0)) delta[v] += s_acc[threadIdx.x];
    __syncthreads();  // s_scan / s_start / s_acc are rewritten by the next tile
  }
}

// Band p of the heavy rows: one chunk (at most 2048 in-edges) per block at a time.
template <class T>
__global__ void __launch_bounds__(kBlock) k_pr_pull_band_heavy(const uint4* __restrict__ chunks,
                                                               const uint32_t* __restrict__ off, uint32_t p,
                                                               const uint32_t* __restrict__ icol,
                                                               const T* __restrict__ contrib, T* __restrict__ delta,
                                                               const Counters* cnt, int gate) {
  __shared__ T s_red[kBlock / 64];
  if (!pull_runs(cnt, gate)) return;
  const uint32_t h1 = off[p + 1];
  for (uint32_t h = off[p] + blockIdx.x; h < h1; h += gridDim.x) {
    const uint4 ch = chunks[h];
    T acc = gather_sum<T, kChunk / kBlock>(icol, contrib, ch.y + threadIdx.x, kBlock, ch.z);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if (lane_id() == 0) s_red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      T s = T(0);
      for (int w = 0; w < kBlock / 64; ++w) s += s_red[w];
      if (s != T(0)) unsafeAtomicAdd(&delta[ch.x], s);
    }
    __syncthreads();
  }
}

// Compact source positions of the current graph and the renamed CSC columns
// (every CSC column is a source with an out-edge, so its position is valid).
__global__ void k_src_flags(uint32_t n, const uint32_t* __restrict__ off, uint32_t* __restrict__ flag) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) flag[v] = off[v + 1] > off[v] ? 1u : 0u;
  else if (v == n) flag[n] = 0u;
}

__global__ void k_src_pos(uint32_t n, const uint32_t* __restrict__ off, const uint32_t* __restrict__ pre,
                          uint32_t* __restrict__ pos) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) pos[v] = off[v + 1] > off[v] ? pre[v] : INF32;
}

__global__ void k_col_pos(uint32_t n, const uint32_t* __restrict__ ioff, const uint32_t* __restrict__ icol,
                          const uint32_t* __restrict__ pos, uint32_t* __restrict__ icol_c) {
  const uint64_t m = ioff[n];
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (uint64_t)gridDim.x * blockDim.x)
    icol_c[e] = pos[icol[e]];
}

// Degree order: keys ~outdeg (a stable sort puts larger out-degrees first, ties
// by id, dangling vertices last), then pos[id] = rank.
__global__ void k_deg_keys(uint32_t n, const uint32_t* __restrict__ off, uint32_t* __restrict__ key,
                           uint32_t* __restrict__ id) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) {
    key[v] = 0xFFFFFFFFu - (off[v + 1] - off[v]);
    id[v] = v;
  }
}

__global__ void k_deg_pos(uint32_t n, const uint32_t* __restrict__ key, const uint32_t* __restrict__ id,
                          uint32_t* __restrict__ pos) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) pos[id[i]] = key[i] != 0xFFFFFFFFu ? i : INF32;
}

void build_positions_by_degree(suminc_ctx* c, const uint32_t* deg_off, const uint32_t* rows_off,
                               const uint32_t* cols, uint64_t m_cap, DevBuf& pos, DevBuf& scratch, DevBuf& cols_c) {
  const uint32_t n = c->n;
  pos.ensure((size_t)n * 4 + 4);
  scratch.ensure((size_t)n * 16 + 16);
  cols_c.ensure_slack(m_cap * 4 + 4);
  uint32_t* k1 = scratch.as<uint32_t>();
  uint32_t* i1 = k1 + n;
  uint32_t* k2 = i1 + n;
  uint32_t* i2 = k2 + n;
  size_t bytes = 0;
  k_deg_keys<<<grid_for(n), kBlock, 0, c->stream>>>(n, deg_off, k1, i1);
  SUMINC_HIP(rocprim::radix_sort_pairs(nullptr, bytes, k1, k2, i1, i2, (size_t)n, 0, 32, c->stream));
  c->tmp.ensure(bytes);
  SUMINC_HIP(rocprim::radix_sort_pairs(c->tmp.p, bytes, k1, k2, i1, i2, (size_t)n, 0, 32, c->stream));
  k_deg_pos<<<grid_for(n), kBlock, 0, c->stream>>>(n, k2, i2, pos.as<uint32_t>());
  k_col_pos<<<grid_for(m_cap, kBlock, 65536), kBlock, 0, c->stream>>>(n, rows_off, cols, pos.as<uint32_t>(),
                                                                      cols_c.as<uint32_t>());
  SUMINC_HIP(hipGetLastError());
}

bool compact_pull(const suminc_ctx* c, uint64_t es) {
  const int kib = c->prm.pull_compact_kib;
  return kib >= 0 && (uint64_t)c->n * es > (uint64_t)(kib ? kib : 32768) * 1024;
}

// by_degree: positions by decreasing out-degree, so the most gathered
// contributions share cache lines (measured on MI355X at 2^26 vertices: 9.0 ms per
// round, against 9.8 ms for id order with source bands and 12.2 ms without);
// otherwise id order, which keeps CSC rows sorted by position (the banded pull
// needs that).
static void build_src_positions(suminc_ctx* c, uint64_t m_cap, bool by_degree) {
  const uint32_t n = c->n;
  // measured on MI355X: at 2^22 vertices (16 MiB of FP32 contributions) renaming
  // costs more per batch than it saves; from 2^24 on it pays
  c->pull_compact = compact_pull(c, c->pr_prec == SUMINC_PR_FP64 ? 8 : 4);
  if (!c->pull_compact) return;
  if (by_degree) {
    build_positions_by_degree(c, c->out_off.as<uint32_t>(), c->in_off.as<uint32_t>(), c->in_col.as<uint32_t>(),
                              m_cap, c->src_pos, c->src_pre, c->in_col_c);
    return;
  }
  c->src_pos.ensure((size_t)n * 4 + 4);
  c->src_pre.ensure((size_t)n * 16 + 16);  // flags, then their exclusive prefix (pre[n] = n_src)
  c->in_col_c.ensure_slack(m_cap * 4 + 4);
  uint32_t* flag = c->src_pre.as<uint32_t>();
  uint32_t* pre = flag + n + 1;
  size_t bytes = 0;
  k_src_flags<<<grid_for((uint64_t)n + 1), kBlock, 0, c->stream>>>(n, c->out_off.as<uint32_t>(), flag);
  SUMINC_HIP(rocprim::exclusive_scan(nullptr, bytes, flag, pre, 0u, (size_t)n + 1, rocprim::plus<uint32_t>(),
                                     c->stream));
  c->tmp.ensure(bytes);
  SUMINC_HIP(rocprim::exclusive_scan(c->tmp.p, bytes, flag, pre, 0u, (size_t)n + 1, rocprim::plus<uint32_t>(),
                                     c->stream));
  k_src_pos<<<grid_for(n), kBlock, 0, c->stream>>>(n, c->out_off.as<uint32_t>(), pre, c->src_pos.as<uint32_t>());
  k_col_pos<<<grid_for(m_cap, kBlock, 65536), kBlock, 0, c->stream>>>(n, c->in_off.as<uint32_t>(),
                                                                      c->in_col.as<uint32_t>(),
                                                                      c->src_pos.as<uint32_t>(),
                                                                      c->in_col_c.as<uint32_t>());
  SUMINC_HIP(hipGetLastError());
}

// Band tables for the lists just built (sizes read back from the device).  Bands
// cut the compact source positions in id order (CSC rows are sorted by them too).
static bool banded_pull(const suminc_ctx* c) { return c->prm.pull_band_kib > 0; }

static void build_band_tables(suminc_ctx* c, uint64_t m_cap) {
  const int kib = c->prm.pull_band_kib;
  const uint64_t es = c->pr_prec == SUMINC_PR_FP64 ? 8 : 4;
  c->band_P = 1;
  if (!banded_pull(c)) return;
  // bands of at most kib KiB of contributions, equal in width (measured on MI355X
  // with id-order positions: 2 bands of 32 MiB at 2^24 vertices and 3-4 of 34-60
  // MiB at 2^26 are fastest; 4 MiB bands lose to the unbanded pull at 2^22
  // vertices, 128 MiB ones at 2^26)
  const uint64_t band_bytes = std::max<uint64_t>(4096, (uint64_t)kib * 1024);
  if ((uint64_t)c->n * es <= band_bytes) return;  // n_src <= n
  unsigned long long h[3];
  uint32_t n_src = c->n;
  SUMINC_HIP(hipMemcpyAsync(h, c->pull_cnt.p, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  if (c->pull_compact)
    SUMINC_HIP(hipMemcpyAsync(&n_src, c->src_pre.as<uint32_t>() + 2 * (size_t)c->n + 1, 4, hipMemcpyDeviceToHost,
                              c->stream));
  SUMINC_HIP(hipStreamSynchronize(c->stream));
  const uint64_t P = std::min<uint64_t>(64, ((uint64_t)n_src * es + band_bytes - 1) / band_bytes);
  if (P <= 1) return;
  const uint64_t B = ((uint64_t)n_src + P - 1) / P;
  const uint32_t S = (uint32_t)h[0], W = (uint32_t)h[1], Hc = (uint32_t)h[2];
  c->band_split.ensure_slack(((size_t)W * (P + 1) + 1) * 4);
  c->band_split_s.ensure_slack(((size_t)S * (P + 1) + 1) * 4);
  if (W)
    k_band_split<<<grid_for((uint64_t)W * (P + 1), kBlock, 65536), kBlock, 0, c->stream>>>(
        c->pull_wave.as<uint32_t>(), W, c->in_off.as<uint32_t>(), pull_cols(c), (uint32_t)P,
        (uint32_t)B, c->band_split.as<uint32_t>());
  if (S)
    k_band_split<<<grid_for((uint64_t)S * (P + 1), kBlock, 65536), kBlock, 0, c->stream>>>(
        c->pull_short.as<uint32_t>(), S, c->in_off.as<uint32_t>(), pull_cols(c), (uint32_t)P,
        (uint32_t)B, c->band_split_s.as<uint32_t>());
  const uint64_t nc = (uint64_t)Hc * P;
  c->band_cnt.ensure_slack((nc + 1) * 4);
  c->band_pre.ensure_slack((nc + 1) * 4);
  c->band_off.ensure((P + 1) * 4);
  // a row of len edges has at most len / kChunk + 1 chunks in each band
  c->band_chunks.ensure_slack(((size_t)m_cap / kChunk + (size_t)P * Hc + 1) * sizeof(uint4));
  SUMINC_HIP(hipMemsetAsync(c->band_off.p, 0, (P + 1) * 4, c->stream));
  const unsigned gc = grid_for(nc + 1, kBlock, 65536);
  k_band_heavy_count<<<gc, kBlock, 0, c->stream>>>(c->pull_heavy.as<uint2>(), Hc, c->in_off.as<uint32_t>(),
                                                   pull_cols(c), (uint32_t)P, (uint32_t)B,
                                                   c->band_cnt.as<uint32_t>());
  size_t bytes = 0;
  SUMINC_HIP(rocprim::exclusive_scan(nullptr, bytes, c->band_cnt.as<uint32_t>(), c->band_pre.as<uint32_t>(), 0u,
                                     (size_t)nc + 1, rocprim::plus<uint32_t>(), c->stream));
  c->tmp.ensure(bytes);
  SUMINC_HIP(rocprim::exclusive_scan(c->tmp.p, bytes, c->band_cnt.as<uint32_t>(), c->band_pre.as<uint32_t>(), 0u,
                                     (size_t)nc + 1, rocprim::plus<uint32_t>(), c->stream));
  k_band_heavy_fill<<<gc, kBlock, 0, c->stream>>>(c->pull_heavy.as<uint2>(), Hc, c->in_off.as<uint32_t>(),
                                                  pull_cols(c), (uint32_t)P, (uint32_t)B,
                                                  c->band_pre.as<uint32_t>(), c->band_chunks.as<uint4>(),
                                                  c->band_off.as<uint32_t>());
  SUMINC_HIP(hipGetLastError());
  c->band_P = (uint32_t)P;
  c->band_B = (uint32_t)B;
  c->band_W = W;
  c->band_S = S;
}

// ---------------------------------------------------------------------------------
// Tiled pull.  A dense round reads contrib at random once per in-edge; once the
// contributions outgrow L2 every such read fills a whole line from the Infinity
// Cache or HBM (the row-class pull above moves about 7x its algorithmic bytes at
// 2^26 vertices).  The tiled layout reorders the CSC once per graph version:
// destinations in blocks of kTileDst, each block's in-edges (cut into tiles of at
// most kTileEdges when longer) grouped by source window (2^win_log2 contribution
// slots, sized to a fraction of L2), each edge packed into 32 bits as (destination
// - block start) << kTileSrcBits | slot within the window.  A round launches one
// workgroup per tile: it sums its destinations in LDS while it walks the windows
// in order (every workgroup walks them in the same order, so a window's lines are
// shared in L2 by the workgroups on an XCD), and adds each destination's sum to
// delta once at the end (atomically for the tiles of a block cut in several).
// The edge stream is 4 B per edge instead of the CSC's column and the gather.
// ---------------------------------------------------------------------------------
constexpr uint32_t kTileDstBits = 13;
constexpr uint32_t kTileSrcBits = 32 - kTileDstBits;
constexpr uint32_t kTileEdges = 16u << kTileDstBits;  // in-edges per tile
constexpr uint32_t kTileMaxWindows = 1024;
// destinations per block (LDS sums): 8192 in FP32, 4096 in FP64 (32 KiB either way)
template <class T>
constexpr uint32_t kTileAcc = (32u << 10) / sizeof(T);
inline uint32_t tile_dst(int prec) { return prec == SUMINC_PR_FP64 ? kTileAcc<double> : kTileAcc<float>; }

// in_off at every block start (blocks + 1 values)
__global__ void k_tile_bounds(uint32_t n, uint32_t blocks, uint32_t bsize, const uint32_t* __restrict__ ioff,
                              uint32_t* __restrict__ out) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k <= blocks) out[k] = ioff[min((uint64_t)k * bsize, (uint64_t)n)];
}

// One workgroup per tile: window histogram of the tile's edges, their offsets, then
// every edge placed in its window's segment (order within a segment: arrival).
__global__ void __launch_bounds__(kBlock) k_tile_build(const uint4* __restrict__ tiles,
                                                       const uint32_t* __restrict__ ioff,
                                                       const uint32_t* __restrict__ icol, uint32_t win_log2,
                                                       uint32_t P, uint32_t* __restrict__ packed,
                                                       uint32_t* __restrict__ seg) {
  __shared__ uint32_t s_pre[kTileMaxWindows + 1];
  __shared__ uint32_t s_fill[kTileMaxWindows];
  const uint4 t = tiles[blockIdx.x];
  const uint32_t d0 = t.x, nd = t.y & 0x7FFFFFFFu, e0 = t.z, e1 = t.w;
  for (uint32_t i = threadIdx.x; i <= P; i += kBlock) {
    s_pre[i] = 0;
    if (i < P) s_fill[i] = 0;
  }
  __syncthreads();
  for (uint32_t e = e0 + threadIdx.x; e < e1; e += kBlock) atomicAdd(&s_pre[icol[e] >> win_log2], 1u);
  __syncthreads();
  if (threadIdx.x == 0) {  // exclusive scan (P <= kTileMaxWindows)
    uint32_t run = 0;
    for (uint32_t b = 0; b < P; ++b) {
      const uint32_t x = s_pre[b];
      s_pre[b] = run;
      run += x;
    }
    s_pre[P] = run;
  }
  __syncthreads();
  uint32_t* sg = seg + (size_t)blockIdx.x * (P + 1);
  for (uint32_t b = threadIdx.x; b <= P; b += kBlock) sg[b] = e0 + s_pre[b];
  const uint32_t mask = (1u << win_log2) - 1u;
  // the rows of the block in chunks of kBlock, each row clipped to the tile's edges,
  // their edges walked by the whole workgroup (the row of an edge from the LDS scan)
  __shared__ uint32_t s_scan[kBlock + 1];
  __shared__ uint32_t s_start[kBlock];
  __shared__ uint32_t s_warp[kBlock / 64 + 1];
  for (uint32_t r0 = d0; r0 < d0 + nd; r0 += kBlock) {
    const uint32_t v = r0 + threadIdx.x;
    uint32_t start = 0, len = 0;
    if (v < d0 + nd) {
      const uint32_t b = max(ioff[v], e0), e = min(ioff[v + 1], e1);
      if (e > b) {
        start = b;
        len = e - b;
      }
    }
    block_adjacency<kBlock>(start, len, s_scan, s_start, s_warp, [&](int slot, uint32_t, uint32_t e) {
      const uint32_t p = icol[e];
      const uint32_t b = p >> win_log2;
      const uint32_t at = atomicAdd(&s_fill[b], 1u);
      packed[e0 + s_pre[b] + at] = ((r0 + (uint32_t)slot - d0) << kTileSrcBits) | (p & mask);
    });
  }
}

// segment offsets of every (tile, window) in one array: flat[t * P + b] = the
// segment start, flat[tiles * P] = the last tile's end (tiles are consecutive)
__global__ void k_tile_flat(const uint32_t* __restrict__ seg, uint32_t tiles, uint32_t P, uint32_t* __restrict__ flat) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t total = (uint64_t)tiles * P;
  if (i < total) flat[i] = seg[(i / P) * (P + 1) + i % P];
  else if (i == total) flat[i] = seg[(uint64_t)(tiles - 1) * (P + 1) + P];
}

// One workgroup per tile: it walks the windows in order (their offsets held in
// LDS), four edges a thread at a time: the packed loads, the gathers from the
// window, the LDS sums.  Each window's run is sorted by slot, so neighbouring lanes
// read neighbouring contributions.
template <class T>
__global__ void __launch_bounds__(kBlock) k_pr_pull_tiled(const uint4* __restrict__ tiles,
                                                          const uint32_t* __restrict__ packed,
                                                          const uint32_t* __restrict__ seg, uint32_t P,
                                                          uint32_t win_log2, const T* __restrict__ contrib,
                                                          T* __restrict__ delta, const Counters* cnt, int gate) {
  __shared__ T s_acc[kTileAcc<T>];
  __shared__ uint32_t s_seg[kTileMaxWindows + 1];
  if (!pull_runs(cnt, gate)) return;
  const uint4 t = tiles[blockIdx.x];
  const uint32_t d0 = t.x, nd = t.y & 0x7FFFFFFFu;
  const bool split = (t.y >> 31) != 0;
  const uint32_t* sg = seg + (size_t)blockIdx.x * (P + 1);
  for (uint32_t i = threadIdx.x; i <= P; i += kBlock) s_seg[i] = sg[i];
  for (uint32_t i = threadIdx.x; i < nd; i += kBlock) s_acc[i] = T(0);
  __syncthreads();
  const uint32_t mask = (1u << win_log2) - 1u;
  for (uint32_t b = 0; b < P; ++b) {
    const uint32_t s = s_seg[b], e = s_seg[b + 1];
    const T* w = contrib + ((size_t)b << win_log2);
    for (uint32_t k = s + threadIdx.x; k < e; k += 4 * kBlock) {
      uint32_t q[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const uint32_t kk = k + r * kBlock;
        q[r] = kk < e ? __builtin_nontemporal_load(packed + kk) : 0u;  // streamed: keep the windows in L2
      }
      T x[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) x[r] = k + r * kBlock < e ? w[q[r] & mask] : T(0);
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (x[r] != T(0)) atomicAdd(&s_acc[q[r] >> kTileSrcBits], x[r]);
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < nd; i += kBlock) {
    const T x = s_acc[i];
    if (x != T(0)) {
      if (split) unsafeAtomicAdd(&delta[d0 + i], x);
      else delta[d0 + i] += x;
    }
  }
}

// Vertices with out-edges (the compact positions' count), added into *out (a
// grid of at most 1024 workgroups striding over the vertices: one atomic each).
__global__ void k_count_sources(uint32_t n, const uint32_t* __restrict__ off, unsigned long long* __restrict__ out) {
  __shared__ unsigned long long s[kBlock / 64];
  unsigned long long c = 0;
  for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < n; v += gridDim.x * kBlock)
    c += off[v + 1] > off[v] ? 1ull : 0ull;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if (lane_id() == 0) s[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long a = 0;
    for (int w = 0; w < kBlock / 64; ++w) a += s[w];
    if (a) atomicAdd(out, a);
  }
}

// The tiles of the current graph (after the positions and the band tables).
static void build_tiles(suminc_ctx* c, uint64_t m_cap) {
  c->tile_n = 0;
  const int kib = c->prm.pull_tile_kib;
  const uint64_t es = c->pr_prec == SUMINC_PR_FP64 ? 8 : 4;
  const uint32_t n = c->n;
  if (kib < 0 || c->band_P > 1 || n == 0) return;
  // measured on MI355X (FP32, 1 MiB windows): 7.1-7.3 against 8.6 ms per dense
  // round at 2^26 vertices, equal at 2^24, 0.67 against 0.38 ms at 2^22
  if (kib == 0 && (uint64_t)n * es <= (64ull << 20)) return;
  const uint64_t win_bytes = (uint64_t)(kib ? kib : 1024) * 1024;
  const uint32_t bsize = tile_dst(c->pr_prec);  // the kernel's LDS sums hold as many destinations
  const uint32_t blocks = (n + bsize - 1) / bsize;
  // the sources a column can name: the compact positions' count, or every vertex
  const size_t src_at = (((size_t)blocks + 1) * 4 + 7) / 8 * 8;  // the count after the bounds, 8-aligned
  c->tile_bounds.ensure(src_at + 8);
  unsigned long long* d_src = reinterpret_cast<unsigned long long*>(c->tile_bounds.as<uint8_t>() + src_at);
  SUMINC_HIP(hipMemsetAsync(d_src, 0, 8, c->stream));
  if (c->pull_compact)
    k_count_sources<<<grid_for(n, kBlock, 1024), kBlock, 0, c->stream>>>(n, c->out_off.as<uint32_t>(), d_src);
  k_tile_bounds<<<grid_for((uint64_t)blocks + 1), kBlock, 0, c->stream>>>(n, blocks, bsize, c->in_off.as<uint32_t>(),
                                                                         c->tile_bounds.as<uint32_t>());
  std::vector<uint32_t> hb(blocks + 1);
  unsigned long long n_src = n;
  SUMINC_HIP(hipMemcpyAsync(hb.data(), c->tile_bounds.p, hb.size() * 4, hipMemcpyDeviceToHost, c->stream));
  if (c->pull_compact) SUMINC_HIP(hipMemcpyAsync(&n_src, d_src, 8, hipMemcpyDeviceToHost, c->stream));
  SUMINC_HIP(hipStreamSynchronize(c->stream));
  n_src = std::max<unsigned long long>(n_src, 1);
  // the CSC's edge count from the device (during a batch the context's m is the
  // old graph's until the batch ends)
  const uint32_t m_now = hb[blocks];
  if (m_now == 0) return;
  uint32_t win_log2 = 8;
  while (win_log2 < kTileSrcBits && (2ull << win_log2) * es <= win_bytes) ++win_log2;
  while (win_log2 < kTileSrcBits && ((n_src + (1ull << win_log2) - 1) >> win_log2) > kTileMaxWindows) ++win_log2;
  const uint32_t P = (uint32_t)((n_src + (1ull << win_log2) - 1) >> win_log2);
  if (P > kTileMaxWindows) return;  // beyond 2^30 sources: the row-class pull
  std::vector<uint4> tiles;
  for (uint32_t k = 0; k < blocks; ++k) {
    const uint32_t e0 = hb[k], e1 = hb[k + 1];
    if (e1 == e0) continue;
    const uint32_t d0 = k * bsize, nd = std::min<uint32_t>(bsize, n - d0);
    const uint32_t parts = (e1 - e0 + kTileEdges - 1) / kTileEdges;
    for (uint32_t j = 0; j < parts; ++j) {
      uint4 t;
      t.x = d0;
      t.y = nd | (parts > 1 ? 0x80000000u : 0u);
      t.z = e0 + j * kTileEdges;
      t.w = std::min<uint32_t>(e1, t.z + kTileEdges);
      tiles.push_back(t);
    }
  }
  if (tiles.empty()) return;
  c->tile_tab.ensure(tiles.size() * sizeof(uint4));
  SUMINC_HIP(hipMemcpyAsync(c->tile_tab.p, tiles.data(), tiles.size() * sizeof(uint4), hipMemcpyHostToDevice,
                            c->stream));
  c->tile_edges.ensure_slack(m_cap * 4 + 4);
  c->tile_seg.ensure_slack(tiles.size() * ((size_t)P + 1) * 4);
  k_tile_build<<<(unsigned)tiles.size(), kBlock, 0, c->stream>>>(c->tile_tab.as<uint4>(), c->in_off.as<uint32_t>(),
                                                                 pull_cols(c), win_log2, P,
                                                                 c->tile_edges.as<uint32_t>(),
                                                                 c->tile_seg.as<uint32_t>());
  SUMINC_HIP(hipGetLastError());
  // each (tile, window) segment sorted by slot, so neighbouring lanes of a round
  // read neighbouring contributions (one line serves the edges of the hot sources)
  {
    const uint64_t T = tiles.size(), nseg = T * P;
    c->tile_flat.ensure((nseg + 1) * 4);
    c->tile_edges2.ensure_slack(m_cap * 4 + 4);
    k_tile_flat<<<grid_for(nseg + 1), kBlock, 0, c->stream>>>(c->tile_seg.as<uint32_t>(), (uint32_t)T, P,
                                                             c->tile_flat.as<uint32_t>());
    const uint32_t* flat = c->tile_flat.as<uint32_t>();
    size_t bytes = 0;
    SUMINC_HIP(rocprim::segmented_radix_sort_keys(nullptr, bytes, c->tile_edges.as<uint32_t>(),
                                                  c->tile_edges2.as<uint32_t>(), m_now, (unsigned)nseg,
                                                  flat, flat + 1, 0u, win_log2, c->stream));
    c->tmp.ensure(bytes);
    SUMINC_HIP(rocprim::segmented_radix_sort_keys(c->tmp.p, bytes, c->tile_edges.as<uint32_t>(),
                                                  c->tile_edges2.as<uint32_t>(), m_now, (unsigned)nseg,
                                                  flat, flat + 1, 0u, win_log2, c->stream));
    c->tile_edges.swap(c->tile_edges2);
  }
  SUMINC_HIP(hipStreamSynchronize(c->stream));  // `tiles` goes out of scope
  c->tile_n = (uint32_t)tiles.size();
  c->tile_P = P;
  c->tile_win_log2 = win_log2;
  c->tile_bsize = bsize;
}

void build_pull_lists(suminc_ctx* c, uint64_t m_cap) {
  const unsigned g = grid_for(c->n);
  c->pull_heavy.ensure_slack((m_cap / kChunk + m_cap / kHeavyDeg + 2) * sizeof(uint2));
  c->pull_short.ensure((size_t)c->n * 4 + 4);
  c->pull_wave.ensure((size_t)c->n * 4 + 4);
  c->pull_parts.ensure((size_t)g * 3 * 4);
  c->pull_cnt.ensure(4 * 8);
  k_pull_count<<<g, kBlock, 0, c->stream>>>(c->n, c->in_off.as<uint32_t>(), c->pull_parts.as<uint32_t>());
  k_pull_scan<<<1, 1024, 0, c->stream>>>(g, c->pull_parts.as<uint32_t>(), c->pull_cnt.as<unsigned long long>());
  k_pull_fill<<<g, kBlock, 0, c->stream>>>(c->n, c->in_off.as<uint32_t>(), c->pull_parts.as<uint32_t>(),
                                           c->pull_short.as<uint32_t>(), c->pull_wave.as<uint32_t>(),
                                           c->pull_heavy.as<uint2>());
  SUMINC_HIP(hipGetLastError());
  build_src_positions(c, m_cap, !banded_pull(c));
  build_band_tables(c, m_cap);
  build_tiles(c, m_cap);
  c->pull_prec = c->pr_prec;  // positions, bands and tiles depend on the element size
}

void ensure_pull_heavy(suminc_ctx* c) {
  if (c->pull_heavy_version == c->graph_version && c->pull_prec == c->pr_prec) return;
  build_pull_lists(c, c->m);
  c->pull_heavy_version = c->graph_version;
}

// The pull launches of a round (kernels return unless the gate opens, see
// pull_runs).  The list lengths stay on the device.  The contributions are at the
// compact positions when those are in use (the columns renamed to them).
template <class T>
void pull_scatter(suminc_ctx* c, int gate) {
  Counters* cnt = dev_counters(c);
  const unsigned long long* counts = c->pull_cnt.as<unsigned long long>();
  const uint32_t* icol = pull_cols(c);
  if (c->tile_n) {
    SUMINC_REQUIRE(c->tile_bsize <= kTileAcc<T>, SUMINC_E_STATE, "tiled pull built for another precision");
    k_pr_pull_tiled<T><<<c->tile_n, kBlock, 0, c->stream>>>(c->tile_tab.as<uint4>(), c->tile_edges.as<uint32_t>(),
                                                           c->tile_seg.as<uint32_t>(), c->tile_P, c->tile_win_log2,
                                                           c->pr_contrib.as<T>(), c->pr_delta.as<T>(), cnt, gate);
    return;
  }
  if (c->band_P > 1) {  // every row band by band
    const uint32_t W = (uint32_t)c->band_W, S = (uint32_t)c->band_S;
    for (uint32_t p = 0; p < c->band_P; ++p) {
      if (S)
        k_pr_pull_band_rows<T><<<grid_for(S, kBlock, 8192), kBlock, 0, c->stream>>>(
            c->pull_short.as<uint32_t>(), S, c->band_split_s.as<uint32_t>(), p, icol,
            c->pr_contrib.as<T>(), c->pr_delta.as<T>(), cnt, gate);
      if (W)
        k_pr_pull_band_rows<T><<<grid_for(W, kBlock, 8192), kBlock, 0, c->stream>>>(
            c->pull_wave.as<uint32_t>(), W, c->band_split.as<uint32_t>(), p, icol,
            c->pr_contrib.as<T>(), c->pr_delta.as<T>(), cnt, gate);
      k_pr_pull_band_heavy<T><<<push_grid(), kBlock, 0, c->stream>>>(
          c->band_chunks.as<uint4>(), c->band_off.as<uint32_t>(), p, icol,
          c->pr_contrib.as<T>(), c->pr_delta.as<T>(), cnt, gate);
    }
    return;
  }
  // grid: the heavy part as before (push_grid), the wave and short parts sized to
  // the vertex count (at most 8192 workgroups each)
  const uint32_t gh = push_grid(), gw = grid_for(c->n, kBlock / 64, 8192), gs = grid_for(c->n, kBlock, 8192);
  k_pr_pull<T><<<gh + gw + gs, kBlock, 0, c->stream>>>(c->pull_short.as<uint32_t>(), c->pull_wave.as<uint32_t>(),
                                                       c->pull_heavy.as<uint2>(), counts, c->in_off.as<uint32_t>(),
                                                       icol, c->pr_contrib.as<T>(), c->pr_delta.as<T>(), cnt, gate, gh,
                                                       gw);
}
template void pull_scatter<float>(suminc_ctx* c, int gate);
template void pull_scatter<double>(suminc_ctx* c, int gate);

// Enqueues one round; with profiling on, events 4*slot .. 4*slot+3 bracket the
// count+decide, gather and scatter (push or pull) launches.
template <class T>
static void enqueue_round(suminc_ctx* c, double eps, double T_stop, int max_rounds, int div, size_t slot) {
  Counters* cnt = dev_counters(c);
  const unsigned g = grid_for(c->n);
  // count and gather take kCountPer vertices a thread: gc (<= g) workgroups and partials
  const unsigned gc = grid_for(c->n, kBlock * kCountPer);
  if (c->prof_on) SUMINC_HIP(hipEventRecord(prof_event(c, 4 * slot), c->stream));
  k_pr_count<T><<<gc, kBlock, 0, c->stream>>>(c->n, c->out_off.as<uint32_t>(), c->pr_delta.as<T>(), eps, cnt,
                                              c->partial.as<double>());
  k_pr_decide<<<1, 1024, 0, c->stream>>>(c->partial.as<double>(), gc, cnt, T_stop, max_rounds, c->m, div,
                                         c->tmp4.as<uint8_t>());
  if (c->prof_on) SUMINC_HIP(hipEventRecord(prof_event(c, 4 * slot + 1), c->stream));
  k_pr_gather<T><<<gc, kBlock, 0, c->stream>>>(c->n, c->out_off.as<uint32_t>(), c->pr_delta.as<T>(),
                                              c->pr_rank.as<T>(), c->pr_comp.as<float>(), eps, (T)c->prm.damping,
                                              c->wl0.as<uint32_t>(), c->wl1.as<T>(), c->heavy.as<uint4>(),
                                              c->pr_contrib.as<T>(), pull_pos(c), cnt,
                                              reinterpret_cast<unsigned long long*>(c->partial.as<double>()) + 2 * g);
  if (c->prof_on) SUMINC_HIP(hipEventRecord(prof_event(c, 4 * slot + 2), c->stream));
  k_pr_push<T><<<push_grid(), kBlock, 0, c->stream>>>(c->out_off.as<uint32_t>(), c->out_col.as<uint32_t>(),
                                                      c->pr_delta.as<T>(), c->wl0.as<uint32_t>(), c->wl1.as<T>(),
                                                      c->heavy.as<uint4>(), cnt);
  if (div > 0) pull_scatter<T>(c, 0);
  if (c->prof_on) SUMINC_HIP(hipEventRecord(prof_event(c, 4 * slot + 3), c->stream));
  k_pr_end<<<1, 1024, 0, c->stream>>>(cnt, reinterpret_cast<unsigned long long*>(c->partial.as<double>()) + 2 * g,
                                      gc);
}

// Adds the profile of the first `rounds` enqueued rounds (the ones that did work).
static void pr_profile_rounds(suminc_ctx* c, size_t rounds) {
  if (!c->prof_on || rounds == 0) return;
  SUMINC_HIP(hipStreamSynchronize(c->stream));
  std::vector<uint8_t> mode(std::min<size_t>(rounds, 65536));
  SUMINC_HIP(hipMemcpy(mode.data(), c->tmp4.p, mode.size(), hipMemcpyDeviceToHost));
  double ms_push = 0, ms_pull = 0, ms_gather = 0;
  size_t n_push = 0, n_pull = 0;
  for (size_t i = 0; i < mode.size(); ++i) {
    float t_g = 0, t_s = 0;
    SUMINC_HIP(hipEventElapsedTime(&t_g, c->prof_events[4 * i + 1], c->prof_events[4 * i + 2]));
    SUMINC_HIP(hipEventElapsedTime(&t_s, c->prof_events[4 * i + 2], c->prof_events[4 * i + 3]));
    ms_gather += t_g;
    if (mode[i]) { ms_pull += t_s; ++n_pull; } else { ms_push += t_s; ++n_push; }
  }
  const Counters* h = c->host_counters;
  const double es = c->pr_prec == SUMINC_PR_FP64 ? 8.0 : 4.0;
  suminc_kernel_profile& p = c->prof[SUMINC_KERNEL_PR_PUSH];
  p.launches += (double)n_push;
  p.ms += ms_push;
  // push: column (4 B) + delta atomic read-modify-write (2 es) per edge; worklist
  // entry (4 + es) and two row pointers (8 B) per entry
  p.edges += (double)h->push_edges;
  p.vertices += (double)h->push_entries;
  p.alg_bytes += (double)h->push_edges * (4.0 + 2.0 * es) + (double)h->push_entries * (12.0 + es);
  suminc_kernel_profile& q = c->prof[SUMINC_KERNEL_PR_PULL];
  q.launches += (double)n_pull;
  q.ms += ms_pull;
  // pull: row index (4 B) + contribution gather (es) per in-edge of the graph;
  // two row pointers (8 B) + delta read-modify-write (2 es) per vertex
  q.edges += (double)n_pull * (double)c->m;
  q.vertices += (double)n_pull * (double)c->n;
  q.alg_bytes += (double)n_pull * ((double)c->m * (4.0 + es) + (double)c->n * (8.0 + 2.0 * es));
  suminc_kernel_profile& g = c->prof[SUMINC_KERNEL_PR_GATHER];
  g.launches += (double)mode.size();
  g.ms += ms_gather;
  // gather: delta read per vertex; per active vertex delta write, rank and Kahan
  // term read-modify-write, row pointers; pull rounds write every contribution
  const double active = (double)h->active_vertices;
  const double kahan = c->pr_prec == SUMINC_PR_FP64 ? 0.0 : 8.0;
  g.vertices += (double)mode.size() * c->n;
  g.alg_bytes += (double)mode.size() * c->n * es + active * (es + 2 * es + kahan + 8.0) +
                 (double)n_pull * c->n * es + (double)h->push_entries * (4.0 + es);
}

// FP64 mass balance of the current state into st->mass_total / mass_defect.
void pr_mass(suminc_ctx* c, suminc_stats* st) {
  if (!c->pr_ready || !st) return;
  c->mass_part.ensure((2 * kMassGrid + 2) * sizeof(double));
  double* part = c->mass_part.as<double>();
  const double d = c->prm.damping;
  if (c->pr_prec == SUMINC_PR_FP64)
    k_pr_mass<double><<<kMassGrid, kBlock, 0, c->stream>>>(c->n, c->out_off.as<uint32_t>(), c->pr_delta.as<double>(),
                                                          c->pr_rank.as<double>(), nullptr, d, part);
  else
    k_pr_mass<float><<<kMassGrid, kBlock, 0, c->stream>>>(c->n, c->out_off.as<uint32_t>(), c->pr_delta.as<float>(),
                                                         c->pr_rank.as<float>(), c->pr_comp.as<float>(), d, part);
  k_pr_mass_end<<<1, 1024, 0, c->stream>>>(part, part + 2 * kMassGrid);
  SUMINC_HIP(hipGetLastError());
  check_begin(c, part + 2 * kMassGrid, 2 * sizeof(double));
  const double* h = reinterpret_cast<const double*>(check_wait(c));
  st->mass_total = h[0];
  st->mass_defect = h[0] - h[1] - (1.0 - d);
}

// Runs rounds until the device sets the stop flag; host polls every `group` rounds.
void pr_run(suminc_ctx* c, suminc_stats* st) {
  SUMINC_REQUIRE(c->pr_ready, SUMINC_E_STATE, "PageRank state not initialised");
  const double T = c->prm.termcheck_threshold;
  const double eps = T / (double)c->n;
  const int max_rounds = c->prm.max_rounds > 0 ? c->prm.max_rounds : 1000000;
  const int div = c->prm.pull_threshold_div;
  if (div > 0) ensure_pull_heavy(c);
  k_zero_counters<<<1, 64, 0, c->stream>>>(dev_counters(c));
  SUMINC_HIP(hipEventRecord(c->ev_start, c->stream));
  int group = 4;
  size_t slot = 0;
  while (true) {
    for (int i = 0; i < group; ++i, ++slot) {
      if (c->pr_prec == SUMINC_PR_FP64) enqueue_round<double>(c, eps, T, max_rounds, div, slot);
      else enqueue_round<float>(c, eps, T, max_rounds, div, slot);
    }
    SUMINC_HIP(hipGetLastError());
    read_counters(c);
    if (c->host_counters->done) break;
    if (group < 32) group *= 2;
  }
  SUMINC_HIP(hipEventRecord(c->ev_stop, c->stream));
  SUMINC_HIP(hipEventSynchronize(c->ev_stop));
  float ms = 0;
  SUMINC_HIP(hipEventElapsedTime(&ms, c->ev_start, c->ev_stop));
  Counters* h = c->host_counters;
  pr_profile_rounds(c, (size_t)h->rounds);
  c->stats = suminc_stats{};
  c->stats.rounds = (double)h->rounds;
  c->stats.edges_traversed = (double)h->edges;
  c->stats.residual = h->residual;
  c->stats.kernel_ms = ms;
  c->stats.pull_rounds = (double)h->pull_rounds;
  c->stats.push_rounds = (double)h->rounds - (double)h->pull_rounds;
  const bool numeric_fail = h->aux[0] != 0;
  const double residual = h->residual;
  pr_mass(c, &c->stats);
  if (st) *st = c->stats;
  SUMINC_REQUIRE(!numeric_fail, SUMINC_E_NUMERIC, "NaN/Inf in PageRank residual");
  SUMINC_REQUIRE(!(residual >= T), SUMINC_E_NOT_CONVERGED, "round cap reached before sum|delta| < T");
}

void pr_step(suminc_ctx* c, double* residual) {
  SUMINC_REQUIRE(c->pr_ready, SUMINC_E_STATE, "PageRank state not initialised");
  const double T = c->prm.termcheck_threshold;
  const int div = c->prm.pull_threshold_div;
  if (div > 0) ensure_pull_heavy(c);
  k_zero_counters<<<1, 64, 0, c->stream>>>(dev_counters(c));
  // stop rule disabled for a single explicit step
  if (c->pr_prec == SUMINC_PR_FP64) enqueue_round<double>(c, T / c->n, -1.0, 1 << 30, div, 0);
  else enqueue_round<float>(c, T / c->n, -1.0, 1 << 30, div, 0);
  SUMINC_HIP(hipGetLastError());
  read_counters(c);
  pr_profile_rounds(c, 1);
  if (residual) *residual = c->host_counters->residual;
}

void pr_download(suminc_ctx* c, double* rank, double* delta) {
  SUMINC_REQUIRE(c->pr_ready, SUMINC_E_STATE, "PageRank state not initialised");
  DevBuf dd, dr;
  dd.ensure((size_t)c->n * 8);
  dr.ensure((size_t)c->n * 8);
  if (c->pr_prec == SUMINC_PR_FP64)
    k_pr_to_f64<double><<<grid_for(c->n), kBlock, 0, c->stream>>>(c->n, c->pr_delta.as<double>(),
                                                                 c->pr_rank.as<double>(), nullptr,
                                                                 dd.as<double>(), dr.as<double>());
  else
    k_pr_to_f64<float><<<grid_for(c->n), kBlock, 0, c->stream>>>(c->n, c->pr_delta.as<float>(),
                                                                c->pr_rank.as<float>(), c->pr_comp.as<float>(),
                                                                dd.as<double>(), dr.as<double>());
  SUMINC_HIP(hipGetLastError());
  if (rank) SUMINC_HIP(hipMemcpyAsync(rank, dr.p, (size_t)c->n * 8, hipMemcpyDeviceToHost, c->stream));
  if (delta) SUMINC_HIP(hipMemcpyAsync(delta, dd.p, (size_t)c->n * 8, hipMemcpyDeviceToHost, c->stream));
  SUMINC_HIP(hipStreamSynchronize(c->stream));
}

void pr_upload(suminc_ctx* c, const double* rank, const double* delta) {
  SUMINC_REQUIRE(c->has_graph, SUMINC_E_STATE, "no graph");
  SUMINC_REQUIRE(rank && delta, SUMINC_E_INVALID, "null state array");
  c->pr_prec = c->prm.pr_precision;
  pr_alloc(c);
  DevBuf dd, dr;
  dd.ensure((size_t)c->n * 8);
  dr.ensure((size_t)c->n * 8);
  SUMINC_HIP(hipMemcpyAsync(dr.p, rank, (size_t)c->n * 8, hipMemcpyHostToDevice, c->stream));
  SUMINC_HIP(hipMemcpyAsync(dd.p, delta, (size_t)c->n * 8, hipMemcpyHostToDevice, c->stream));
  if (c->pr_prec == SUMINC_PR_FP64)
    k_pr_from_f64<double><<<grid_for(c->n), kBlock, 0, c->stream>>>(c->n, c->pr_delta.as<double>(),
                                                                   c->pr_rank.as<double>(), nullptr,
                                                                   dd.as<double>(), dr.as<double>());
  else
    k_pr_from_f64<float><<<grid_for(c->n), kBlock, 0, c->stream>>>(c->n, c->pr_delta.as<float>(),
                                                                  c->pr_rank.as<float>(), c->pr_comp.as<float>(),
                                                                  dd.as<double>(), dr.as<double>());
  SUMINC_HIP(hipGetLastError());
  SUMINC_HIP(hipStreamSynchronize(c->stream));
  c->pr_ready = true;
  ++c->pr_state_id;
}

}  // namespace suminc
