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
// that chose pull (not stopped); gate 1, a reseed that chose pull; gate 2, a
// weighted series round (not stopped).
__device__ __forceinline__ bool pull_runs(const Counters* cnt, int gate) {
  if (gate == 2) return cnt == nullptr || !cnt->done;  // a series round (weighted PageRank)
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
  __shared__ double s_acc[kBlock];  // FP64: see kTileDst
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
    s_acc[threadIdx.x] = 0.0;
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
        if (slot[r] >= 0) atomicAdd(&s_acc[slot[r]], (double)x[r]);
    }
    __syncthreads();
    if (i < W && s_acc[threadIdx.x] != 0.0) delta[v] += (T)s_acc[threadIdx.x];
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

// Degree order by classes: four classes per octave of the degree (the leading bit
// and the two below it), 8-bit keys, larger degrees first, dangling vertices
// (degree 0) last.  The stable sort keeps id order inside a class, so a vertex
// pass writing pos[v] for consecutive v writes each class's positions in order
// (a few sequential streams instead of scattered 4-byte writes), while the hot
// sources still fill the first positions.  Then pos[id] = rank.
constexpr uint32_t kDegKeyDangling = 255;
__device__ __forceinline__ uint32_t degree_class_key(uint32_t deg) {
  if (deg == 0) return kDegKeyDangling;
  const uint32_t lead = 31u - (uint32_t)__clz(deg);  // floor(log2 deg)
  const uint32_t sub = lead >= 2 ? (deg >> (lead - 2)) & 3u : (deg << (2 - lead)) & 3u;
  return kDegKeyDangling - 1u - (lead * 4u + sub);  // in [126, 254]
}

__global__ void k_deg_keys(uint32_t n, const uint32_t* __restrict__ off, uint32_t* __restrict__ key,
                           uint32_t* __restrict__ id) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) {
    key[v] = degree_class_key(off[v + 1] - off[v]);
    id[v] = v;
  }
}

__global__ void k_deg_pos(uint32_t n, const uint32_t* __restrict__ key, const uint32_t* __restrict__ id,
                          uint32_t* __restrict__ pos) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) pos[id[i]] = key[i] != kDegKeyDangling ? i : INF32;
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
  SUMINC_HIP(rocprim::radix_sort_pairs(nullptr, bytes, k1, k2, i1, i2, (size_t)n, 0, 8, c->stream));
  c->tmp.ensure(bytes);
  SUMINC_HIP(rocprim::radix_sort_pairs(c->tmp.p, bytes, k1, k2, i1, i2, (size_t)n, 0, 8, c->stream));
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
  const bool was = c->pull_compact;
  c->pull_compact = compact_pull(c, c->pr_prec == SUMINC_PR_FP64 ? 8 : 4);
  if (!c->pull_compact) {
    c->pos_live = c->pos_carried = false;
    return;
  }
  if (by_degree) {
    if (c->pos_carried && was && c->pos_live) {
      // the batch renamed the new CSC's columns during its merge and appended the
      // positions of the new sources: only the count comes back
      unsigned long long cnt = 0;
      SUMINC_HIP(hipMemcpyAsync(&cnt, c->pos_ctr.p, 8, hipMemcpyDeviceToHost, c->stream));
      SUMINC_HIP(hipStreamSynchronize(c->stream));
      c->pos_appended += (uint32_t)(cnt - c->pos_count);
      c->pos_count = (uint32_t)cnt;
      c->pos_carried = false;
      return;
    }
    build_positions_by_degree(c, c->out_off.as<uint32_t>(), c->in_off.as<uint32_t>(), c->in_col.as<uint32_t>(),
                              m_cap, c->src_pos, c->src_pre, c->in_col_c);
    c->pos_count = (uint32_t)nonempty_rows(c, c->out_off.as<uint32_t>(), c->pos_ctr);  // positions [0, count)
    c->pos_appended = 0;
    c->pos_live = true;
    c->pos_carried = false;
    return;
  }
  c->pos_live = c->pos_carried = false;  // id order (the banded pull) is renumbered every time
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
  c->pos_count = (uint32_t)nonempty_rows(c, c->out_off.as<uint32_t>(), c->pos_ctr);
}

// New positions for the vertices that gained out-edges: flag them, an exclusive
// scan of the flags gives each its rank among them, and the count of positions in
// use (ctr) grows by their number.
__global__ void k_pos_new_flags(uint32_t n, const uint32_t* __restrict__ off, const uint32_t* __restrict__ pos,
                                uint32_t* __restrict__ flag) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) flag[v] = off[v + 1] > off[v] && pos[v] == INF32 ? 1u : 0u;
  else if (v == n) flag[n] = 0u;
}

__global__ void k_pos_new_assign(uint32_t n, const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pre,
                                 uint32_t base, uint32_t* __restrict__ pos, unsigned long long* __restrict__ ctr) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n && flag[v]) pos[v] = base + pre[v];
  else if (v == n) *ctr = (unsigned long long)base + pre[n];
}

bool pr_positions_carry(const suminc_ctx* c) {
  // the appended vertices are the coldest sources (they had no out-edge before the
  // batch), so appending keeps the degree-class order's hot windows; renumber once
  // the appended positions reach half of those renumbered (a 1% batch of uniform
  // inserts at 2^26 vertices appends about 2.4M to 37M)
  return c->pull_compact && c->pos_live && c->pull_prec == c->pr_prec && c->prm.pull_band_kib <= 0 &&
         (uint64_t)c->pos_appended * 2 < (uint64_t)c->pos_count - c->pos_appended;
}

void pr_positions_append(suminc_ctx* c) {
  const uint32_t n = c->n;
  c->pos_ctr.ensure(8);
  c->src_pre.ensure((size_t)n * 16 + 16);  // flags (n + 1), then their exclusive prefix
  uint32_t* flag = c->src_pre.as<uint32_t>();
  uint32_t* pre = flag + n + 1;
  k_pos_new_flags<<<grid_for((uint64_t)n + 1), kBlock, 0, c->stream>>>(n, c->out_off.as<uint32_t>(),
                                                                      c->src_pos.as<uint32_t>(), flag);
  size_t bytes = 0;
  SUMINC_HIP(rocprim::exclusive_scan(nullptr, bytes, flag, pre, 0u, (size_t)n + 1, rocprim::plus<uint32_t>(),
                                     c->stream));
  c->tmp.ensure(bytes);
  SUMINC_HIP(rocprim::exclusive_scan(c->tmp.p, bytes, flag, pre, 0u, (size_t)n + 1, rocprim::plus<uint32_t>(),
                                     c->stream));
  k_pos_new_assign<<<grid_for((uint64_t)n + 1), kBlock, 0, c->stream>>>(n, flag, pre, c->pos_count,
                                                                       c->src_pos.as<uint32_t>(),
                                                                       c->pos_ctr.as<unsigned long long>());
  SUMINC_HIP(hipGetLastError());
  c->pos_carried = true;
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
// destinations per block: 4096, their sums in LDS in FP64 (32 KiB) in both
// precisions.  The FP64 LDS add is the faster one on MI355X (a probe of random
// adds into 32 KiB: 1.5e12 FP64 against 2.0e11 FP32 lane additions a second, the
// FP32 add the bottleneck of a dense round), and it rounds each sum once.
constexpr uint32_t kTileDst = (32u << 10) / sizeof(double);
inline uint32_t tile_dst(int) { return kTileDst; }

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
  __shared__ double s_acc[kTileDst];
  __shared__ uint32_t s_seg[kTileMaxWindows + 1];
  if (!pull_runs(cnt, gate)) return;
  const uint4 t = tiles[blockIdx.x];
  const uint32_t d0 = t.x, nd = t.y & 0x7FFFFFFFu;
  const bool split = (t.y >> 31) != 0;
  const uint32_t* sg = seg + (size_t)blockIdx.x * (P + 1);
  for (uint32_t i = threadIdx.x; i <= P; i += kBlock) s_seg[i] = sg[i];
  for (uint32_t i = threadIdx.x; i < nd; i += kBlock) s_acc[i] = 0.0;
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
        if (x[r] != T(0)) atomicAdd(&s_acc[q[r] >> kTileSrcBits], (double)x[r]);
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < nd; i += kBlock) {
    const double x = s_acc[i];
    if (x != 0.0) {
      if (split) unsafeAtomicAdd(&delta[d0 + i], (T)x);
      else delta[d0 + i] += (T)x;
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
  // the sources a column can name: the compact positions in use, or every vertex
  c->tile_bounds.ensure(((size_t)blocks + 1) * 4);
  k_tile_bounds<<<grid_for((uint64_t)blocks + 1), kBlock, 0, c->stream>>>(n, blocks, bsize, c->in_off.as<uint32_t>(),
                                                                         c->tile_bounds.as<uint32_t>());
  std::vector<uint32_t> hb(blocks + 1);
  SUMINC_HIP(hipMemcpyAsync(hb.data(), c->tile_bounds.p, hb.size() * 4, hipMemcpyDeviceToHost, c->stream));
  SUMINC_HIP(hipStreamSynchronize(c->stream));
  const unsigned long long n_src = std::max<unsigned long long>(c->pull_compact ? c->pos_count : n, 1);
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

// ---------------------------------------------------------------------------------
// Binned pull (propagation blocking: Beamer, Asanovic and Patterson, "Reducing
// PageRank communication via propagation blocking", IPDPS 2017).  The tiled pull
// still reads each contribution where an in-edge names it, and at 2^26 vertices
// most of those reads miss L2: the tiles walk the windows at their own pace, so an
// XCD's L2 holds few of the windows its workgroups read.  Here a dense round makes
// two passes over the in-edges, held once per graph version in (source window,
// destination block) order, each packed into 32 bits as the tiles' are
// ((destination - block start) << kTileSrcBits | slot within the window):
//   k_pr_bin_gather  the value stream vals[i] = contrib[window(i) start + slot(i)]:
//                    workgroups over consecutive edges, so the workgroups in flight
//                    read one or two windows (2 MiB each), which stay in every
//                    XCD's L2; the packed reads and the value writes stream.
//   k_pr_bin_sum     one workgroup per part of a destination block: the block's
//                    segments of every window, taken as one run (their bounds in
//                    LDS), summed into LDS, then delta += the sums (atomically for
//                    a block cut into several parts).  The first `direct` windows
//                    (the hottest sources) are gathered here from contrib directly
//                    and left out of the value stream.
// Per streamed in-edge 16 bytes move (the packed edge twice, the value written and
// read back) and no line is filled for a single 4-byte gather.  Within the segment
// of one window and one build piece (at most kTileEdges in-edges of a block) the
// edges are sorted by slot.
// ---------------------------------------------------------------------------------
constexpr uint32_t kBinGatherEdges = 8192;  // in-edges per workgroup of the gather pass

// Rows with at least one entry in off (n + 1 row pointers), counted on the
// device into scratch (8 bytes).  Waits for the stream.
unsigned long long nonempty_rows(suminc_ctx* c, const uint32_t* off, DevBuf& scratch) {
  scratch.ensure(8);
  unsigned long long* d = scratch.as<unsigned long long>();
  SUMINC_HIP(hipMemsetAsync(d, 0, 8, c->stream));
  k_count_sources<<<grid_for(c->n, kBlock, 1024), kBlock, 0, c->stream>>>(c->n, off, d);
  unsigned long long h = 0;
  SUMINC_HIP(hipMemcpyAsync(&h, d, 8, hipMemcpyDeviceToHost, c->stream));
  SUMINC_HIP(hipStreamSynchronize(c->stream));
  return h;
}

// The build works on pieces of at most kTileEdges in-edges of one block (the
// tiles' table: first destination, destination count, CSC edge range), one
// workgroup each.  Counting: the piece's edges by source window into
// hist[w * J + j] (J pieces).  Window 0 (the hottest sources, most of the edges)
// is counted with one LDS atomic a wave; one counter taking most lanes' atomics
// would serialise them.
__global__ void __launch_bounds__(kBlock) k_bin_hist(const uint4* __restrict__ pieces, uint32_t J,
                                                     const uint32_t* __restrict__ icol, uint32_t wl, uint32_t P,
                                                     uint32_t* __restrict__ hist) {
  __shared__ uint32_t s_h[kTileMaxWindows];
  const uint4 t = pieces[blockIdx.x];
  for (uint32_t w = threadIdx.x; w < P; w += kBlock) s_h[w] = 0;
  __syncthreads();
  for (uint32_t e0 = t.z; e0 < t.w; e0 += kBlock) {  // the same trip count for every lane (ballots)
    const uint32_t e = e0 + threadIdx.x;
    const uint32_t w = e < t.w ? icol[e] >> wl : INF32;
    const unsigned long long hot = __ballot(w == 0);
    if (lane_id() == 0 && hot) atomicAdd(&s_h[0], (uint32_t)__popcll(hot));
    if (w != 0 && w != INF32) atomicAdd(&s_h[w], 1u);
  }
  __syncthreads();
  for (uint32_t w = threadIdx.x; w < P; w += kBlock) hist[(size_t)w * J + blockIdx.x] = s_h[w];
}

// Placing: the piece's edges packed into their (window, piece) segments, whose
// offsets F2[w * J + j] are the exclusive scan of hist (so the pieces of a block
// make up its (window, block) segments).
__global__ void __launch_bounds__(kBlock) k_bin_scatter(const uint4* __restrict__ pieces, uint32_t J,
                                                        const uint32_t* __restrict__ ioff,
                                                        const uint32_t* __restrict__ icol, uint32_t wl, uint32_t P,
                                                        const uint32_t* __restrict__ F2,
                                                        uint32_t* __restrict__ packed, uint32_t* __restrict__ eid) {
  __shared__ uint32_t s_at[kTileMaxWindows];
  __shared__ uint32_t s_scan[kBlock + 1];
  __shared__ uint32_t s_start[kBlock];
  __shared__ uint32_t s_warp[kBlock / 64 + 1];
  const uint4 t = pieces[blockIdx.x];
  const uint32_t d0 = t.x, nd = t.y & 0x7FFFFFFFu, e0 = t.z, e1 = t.w;
  for (uint32_t w = threadIdx.x; w < P; w += kBlock) s_at[w] = F2[(size_t)w * J + blockIdx.x];
  __syncthreads();
  const uint32_t mask = (1u << wl) - 1u;
  // the block's rows in chunks of kBlock, each clipped to the piece's edges
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
      const uint32_t w = p >> wl;
      // window 0's edges of the wave's active lanes take one LDS atomic together
      const unsigned long long hot = __ballot(w == 0);
      uint32_t at;
      if (w == 0) {
        const uint32_t below = (uint32_t)__popcll(hot & ((1ull << lane_id()) - 1ull));
        uint32_t base = 0;
        if (below == 0) base = atomicAdd(&s_at[0], (uint32_t)__popcll(hot));
        at = __builtin_amdgcn_readfirstlane(base) + below;  // the first active lane here took the atomic
      } else {
        at = atomicAdd(&s_at[w], 1u);
      }
      packed[at] = ((r0 + (uint32_t)slot - d0) << kTileSrcBits) | (p & mask);
      if (eid) eid[at] = e;  // the edge's index in row order, for per-edge coefficients
    });
  }
}

// The (window, block) segment offsets from the pieces': F[w * B + b] =
// F2[w * J + first[b]] (first[b]: the first piece of block b or after it), and
// F[P * B] = F2[P * J] (the edge count).
__global__ void k_bin_fold(const uint32_t* __restrict__ F2, uint32_t J, const uint32_t* __restrict__ first,
                           uint32_t B, uint32_t P, uint32_t* __restrict__ F) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t total = (uint64_t)P * B;
  if (i < total) F[i] = F2[(i / B) * J + first[i % B]];
  else if (i == total) F[i] = F2[(uint64_t)P * J];
}

// Four consecutive elements, loaded or stored as one access (the edge arrays and
// the value stream are read in 4-aligned groups).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
template <class T>
using vec4_t = T __attribute__((ext_vector_type(4)));

// The gather pass over the streamed edges [first, end) (first a multiple of 4; the
// edges before the stream's first that it also covers belong to direct windows,
// whose values nobody reads): kBinGatherEdges per workgroup, eight 4-edge groups a
// thread, all loads issued before the gathers.  A full workgroup inside one window
// (the window of its first edge by binary search over the window starts F[w * B])
// takes the vector path; the others (a window start or the end inside) go edge by
// edge, each thread's window only moving forward.  COEF: each value is multiplied
// by its edge's coefficient (coef, layout order).
template <class T, bool COEF>
__global__ void __launch_bounds__(kBlock) k_pr_bin_gather(const uint32_t* __restrict__ packed,
                                                          const uint32_t* __restrict__ F, uint32_t B, uint32_t P,
                                                          uint32_t first, uint32_t end, uint32_t wl,
                                                          const T* __restrict__ contrib, const T* __restrict__ coef,
                                                          T* __restrict__ vals, const Counters* cnt, int gate) {
  constexpr int G = kBinGatherEdges / (4 * kBlock);
  __shared__ uint32_t s_w;
  if (!pull_runs(cnt, gate)) return;
  const uint32_t i0 = (uint32_t)((uint64_t)first + (uint64_t)blockIdx.x * kBinGatherEdges);  // < end
  const uint32_t i1 = (uint32_t)min((uint64_t)i0 + kBinGatherEdges, (uint64_t)end);
  if (threadIdx.x == 0) {
    uint32_t lo = 0, hi = P - 1;
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) >> 1;
      if (F[(size_t)mid * B] <= i0) lo = mid;
      else hi = mid - 1;
    }
    s_w = lo;
  }
  __syncthreads();
  const uint32_t w0 = s_w;
  const uint32_t wend = w0 + 1 < P ? F[(size_t)(w0 + 1) * B] : end;
  const uint32_t mask = (1u << wl) - 1u;
  if (i1 - i0 == kBinGatherEdges && i1 <= wend) {
    const T* win = contrib + ((size_t)w0 << wl);
    const u32x4* p4 = reinterpret_cast<const u32x4*>(packed + i0);
    vec4_t<T>* v4 = reinterpret_cast<vec4_t<T>*>(vals + i0);
    u32x4 q[G];
#pragma unroll
    for (int r = 0; r < G; ++r) q[r] = __builtin_nontemporal_load(p4 + threadIdx.x + r * kBlock);
    vec4_t<T> a[G];
    if constexpr (COEF) {
      const vec4_t<T>* c4 = reinterpret_cast<const vec4_t<T>*>(coef + i0);
#pragma unroll
      for (int r = 0; r < G; ++r) a[r] = __builtin_nontemporal_load(c4 + threadIdx.x + r * kBlock);
    }
    vec4_t<T> x[G];
#pragma unroll
    for (int r = 0; r < G; ++r) {
      x[r].x = win[q[r].x & mask];
      x[r].y = win[q[r].y & mask];
      x[r].z = win[q[r].z & mask];
      x[r].w = win[q[r].w & mask];
      if constexpr (COEF) x[r] *= a[r];
    }
#pragma unroll
    for (int r = 0; r < G; ++r) __builtin_nontemporal_store(x[r], v4 + threadIdx.x + r * kBlock);
  } else {
    uint32_t w = w0;  // a thread's edges ascend, so its window only moves forward
    for (uint32_t k = i0 + threadIdx.x; k < i1; k += kBlock) {
      while (w + 1 < P && F[(size_t)(w + 1) * B] <= k) ++w;
      const T v = contrib[((size_t)w << wl) + (packed[k] & mask)];
      vals[k] = COEF ? v * coef[k] : v;
    }
  }
}

// The summing pass: one workgroup per part of a destination block.  The block's
// segment of window w is [s_w, e_w); it is read in 4-aligned groups of edges,
// those from s_w & ~3 to past e_w - 1, the edges of a group outside [s_w, e_w)
// masked off.  The groups of all windows in order form the block's run, a part is
// the run's range [z, zend); the segment bounds and the run position of every
// window go to LDS first (a workgroup scan), then every thread walks its groups in
// ascending order, its window pointer only moving forward, four groups (sixteen
// edges) an iteration with all their loads issued first.  A wave's 64 consecutive
// groups that lie in one window segment are taken across the wave instead: lane l
// takes the span's edges 64 c + l (c = 0..3), so that a gather instruction reads 64
// consecutive edges of the slot-sorted segment, and the vector L1 looks up fewer
// distinct lines an instruction (a lane's own group would spread an instruction
// over 256 edges).
constexpr uint32_t kBinBlock = 512;  // threads of a summing workgroup (four fit a CU's LDS)
constexpr uint32_t kBinMaxWindows = kBinBlock;
template <class T, bool COEF>
__global__ void __launch_bounds__(kBinBlock) k_pr_bin_sum(const uint4* __restrict__ parts,
                                                       const uint32_t* __restrict__ F, uint32_t B, uint32_t P,
                                                       uint32_t bsize, uint32_t n, const uint32_t* __restrict__ packed,
                                                       const T* __restrict__ vals, const T* __restrict__ contrib,
                                                       const T* __restrict__ coef, uint32_t wl, uint32_t direct,
                                                       T* __restrict__ delta, const Counters* cnt, int gate) {
  constexpr int U = 4;
  __shared__ double s_acc[kTileDst];
  __shared__ uint32_t s_beg[kBinMaxWindows];
  __shared__ uint32_t s_end[kBinMaxWindows];
  __shared__ uint32_t s_pre[kBinMaxWindows + 1];
  __shared__ uint32_t s_warp[kBinBlock / 64 + 1];
  if (!pull_runs(cnt, gate)) return;
  const uint4 t = parts[blockIdx.x];
  const uint32_t b = t.x & 0x7FFFFFFFu, z = t.z, zend = t.w;
  const bool split = (t.x >> 31) != 0;
  const uint32_t d0 = b * bsize, nd = min(bsize, n - d0);
  for (uint32_t i = threadIdx.x; i < nd; i += kBinBlock) s_acc[i] = 0.0;
  // every window's segment and group count, a window a thread (P <= kBinBlock)
  uint32_t ng = 0;
  const uint32_t wt = threadIdx.x;
  if (wt < P) {
    const uint32_t s = F[(size_t)wt * B + b], e = F[(size_t)wt * B + b + 1];
    s_beg[wt] = s;
    s_end[wt] = e;
    ng = e > s ? ((e + 3) >> 2) - (s >> 2) : 0u;
  }
  uint32_t total;
  const uint32_t run = block_exclusive_scan<kBinBlock>(ng, s_warp, &total);
  if (wt < P) s_pre[wt] = run;
  if (threadIdx.x == 0) s_pre[P] = total;
  __syncthreads();
  const uint32_t mask = (1u << wl) - 1u;
  uint32_t j[U];
#pragma unroll
  for (int r = 0; r < U; ++r) {
    const uint32_t g = z + threadIdx.x + r * kBinBlock;
    uint32_t lo = 0, hi = P - 1;
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) >> 1;
      if (s_pre[mid] <= g) lo = mid;
      else hi = mid - 1;
    }
    j[r] = lo;
  }
  for (uint32_t g = z + threadIdx.x; g < zend; g += U * kBinBlock) {
    // the groups' windows and positions first (LDS only), then every load in
    // straight-line code (an invalid group reads the arrays' first group, a direct
    // one the value stream's first group; both masked off below), so all of them
    // are in flight together
    uint32_t k0[U], lo[U], hi[U];
    bool dir[U];
#pragma unroll
    for (int r = 0; r < U; ++r) {
      const uint32_t gr = g + r * kBinBlock;
      const bool valid = gr < zend;
      if (valid)
        while (s_pre[j[r] + 1] <= gr) ++j[r];
      const uint32_t jr = j[r];
      k0[r] = valid ? (s_beg[jr] & ~3u) + 4 * (gr - s_pre[jr]) : 0u;
      lo[r] = valid ? s_beg[jr] : 1u;
      hi[r] = valid ? s_end[jr] : 0u;
      dir[r] = jr < direct;
    }
    // a span across the wave: all 64 lanes active with valid groups of one window
    // (then k0 = lane 0's k0 + 4 lane, and lo, hi and dir agree over the wave), for
    // all four groups of the iteration
    bool span = true;
#pragma unroll
    for (int r = 0; r < U; ++r) {
      const uint32_t j0 = __builtin_amdgcn_readfirstlane(j[r]);
      span = span && __ballot(g + r * kBinBlock < zend && j[r] == j0) == ~0ull;
    }
    u32x4 q[U];
    vec4_t<T> x[U];
    bool in[U][4];
    if (span) {
      const uint32_t l3 = 3u * (threadIdx.x & 63u);
#pragma unroll
      for (int r = 0; r < U; ++r) {
        const uint32_t* p = packed + (k0[r] - l3);  // lane 0's k0 + lane
        q[r].x = __builtin_nontemporal_load(p);
        q[r].y = __builtin_nontemporal_load(p + 64);
        q[r].z = __builtin_nontemporal_load(p + 128);
        q[r].w = __builtin_nontemporal_load(p + 192);
      }
#pragma unroll
      for (int r = 0; r < U; ++r) {
        if (!dir[r] || COEF) {
          const T* p = (dir[r] ? coef : vals) + (k0[r] - l3);
          x[r].x = __builtin_nontemporal_load(p);
          x[r].y = __builtin_nontemporal_load(p + 64);
          x[r].z = __builtin_nontemporal_load(p + 128);
          x[r].w = __builtin_nontemporal_load(p + 192);
        } else {
          x[r] = vec4_t<T>(T(0));
        }
      }
#pragma unroll
      for (int r = 0; r < U; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const uint32_t k = k0[r] - l3 + 64u * c;
          in[r][c] = k >= lo[r] && k < hi[r];
        }
    } else {
#pragma unroll
      for (int r = 0; r < U; ++r) q[r] = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(packed + k0[r]));
      // the streamed groups' values; a direct group reads its coefficients (COEF) or
      // the stream's first group (masked off below)
#pragma unroll
      for (int r = 0; r < U; ++r)
        x[r] = __builtin_nontemporal_load(
            reinterpret_cast<const vec4_t<T>*>(dir[r] ? (COEF ? coef + k0[r] : vals) : vals + k0[r]));
#pragma unroll
      for (int r = 0; r < U; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) in[r][c] = k0[r] + c >= lo[r] && k0[r] + c < hi[r];
    }
    if (dir[0] && dir[1] && dir[2] && dir[3]) {
      // all four groups in direct windows (the common case inside the hottest
      // window's long segments): the sixteen gathers issued together, unconditional
      // (an edge masked off reads its window's first slot)
      T a[U][4];
#pragma unroll
      for (int r = 0; r < U; ++r) {
        const T* win = contrib + ((size_t)j[r] << wl);
        a[r][0] = win[in[r][0] ? q[r].x & mask : 0u];
        a[r][1] = win[in[r][1] ? q[r].y & mask : 0u];
        a[r][2] = win[in[r][2] ? q[r].z & mask : 0u];
        a[r][3] = win[in[r][3] ? q[r].w & mask : 0u];
      }
#pragma unroll
      for (int r = 0; r < U; ++r) {
        x[r].x = in[r][0] ? (COEF ? a[r][0] * x[r].x : a[r][0]) : T(0);
        x[r].y = in[r][1] ? (COEF ? a[r][1] * x[r].y : a[r][1]) : T(0);
        x[r].z = in[r][2] ? (COEF ? a[r][2] * x[r].z : a[r][2]) : T(0);
        x[r].w = in[r][3] ? (COEF ? a[r][3] * x[r].w : a[r][3]) : T(0);
      }
    } else {
#pragma unroll
      for (int r = 0; r < U; ++r) {
        if (dir[r]) {
          const T* win = contrib + ((size_t)j[r] << wl);
          const T a0 = win[in[r][0] ? q[r].x & mask : 0u], a1 = win[in[r][1] ? q[r].y & mask : 0u];
          const T a2 = win[in[r][2] ? q[r].z & mask : 0u], a3 = win[in[r][3] ? q[r].w & mask : 0u];
          x[r].x = COEF ? a0 * x[r].x : a0;
          x[r].y = COEF ? a1 * x[r].y : a1;
          x[r].z = COEF ? a2 * x[r].z : a2;
          x[r].w = COEF ? a3 * x[r].w : a3;
        }
        if (!in[r][0]) x[r].x = T(0);
        if (!in[r][1]) x[r].y = T(0);
        if (!in[r][2]) x[r].z = T(0);
        if (!in[r][3]) x[r].w = T(0);
      }
    }
#pragma unroll
    for (int r = 0; r < U; ++r) {
      if (x[r].x != T(0)) atomicAdd(&s_acc[q[r].x >> kTileSrcBits], (double)x[r].x);
      if (x[r].y != T(0)) atomicAdd(&s_acc[q[r].y >> kTileSrcBits], (double)x[r].y);
      if (x[r].z != T(0)) atomicAdd(&s_acc[q[r].z >> kTileSrcBits], (double)x[r].z);
      if (x[r].w != T(0)) atomicAdd(&s_acc[q[r].w >> kTileSrcBits], (double)x[r].w);
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < nd; i += kBinBlock) {
    const double x = s_acc[i];
    if (x != 0.0) {
      if (split) unsafeAtomicAdd(&delta[d0 + i], (T)x);
      else delta[d0 + i] += (T)x;
    }
  }
}

// Builds L for the adjacency (rows_off, cols: n rows, columns < n_src) with source
// windows of win_bytes / es slots; leaves L.n_parts = 0 when the layout does not
// apply (no edges, more than kBinMaxWindows windows).  with_eid: L.eid holds the row
// order index of every edge, and L.coef has room for a coefficient per edge.
// m_cap: the edge capacity the buffers are sized for.  Waits for the stream.
void bin_layout_build(suminc_ctx* c, BinLayout& L, uint32_t n, const uint32_t* rows_off, const uint32_t* cols,
                      unsigned long long n_src, uint64_t es, uint64_t win_bytes, int direct_param, uint64_t m_cap,
                      bool with_eid) {
  L.n_parts = 0;
  if (n == 0) return;
  n_src = std::max<unsigned long long>(n_src, 1);
  const uint32_t bsize = kTileDst;  // the summing pass's LDS holds as many destinations
  const uint32_t B = (n + bsize - 1) / bsize;
  uint32_t wl = 8;
  while (wl < kTileSrcBits && (2ull << wl) * es <= win_bytes) ++wl;
  while (wl < kTileSrcBits && ((n_src + (1ull << wl) - 1) >> wl) > kBinMaxWindows) ++wl;
  const uint32_t P = (uint32_t)((n_src + (1ull << wl) - 1) >> wl);
  if (P > kBinMaxWindows) return;  // beyond 2^28 sources: the tiles or the row classes
  // the blocks' row bounds, cut into pieces of at most kTileEdges edges
  L.bounds.ensure(((size_t)B + 1) * 4);
  k_tile_bounds<<<grid_for((uint64_t)B + 1), kBlock, 0, c->stream>>>(n, B, bsize, rows_off, L.bounds.as<uint32_t>());
  std::vector<uint32_t> hb(B + 1);
  SUMINC_HIP(hipMemcpyAsync(hb.data(), L.bounds.p, hb.size() * 4, hipMemcpyDeviceToHost, c->stream));
  SUMINC_HIP(hipStreamSynchronize(c->stream));
  // the edge count from the device (during a batch the context's m is the old graph's)
  const uint32_t m_now = hb[B];
  if (m_now == 0) return;
  std::vector<uint4> pieces;
  std::vector<uint32_t> first(B);
  for (uint32_t k = 0; k < B; ++k) {
    first[k] = (uint32_t)pieces.size();
    const uint32_t e0 = hb[k], e1 = hb[k + 1];
    const uint32_t d0 = k * bsize, nd = std::min<uint32_t>(bsize, n - d0);
    for (uint64_t e = e0; e < e1; e += kTileEdges)
      pieces.push_back(make_uint4(d0, nd, (uint32_t)e, (uint32_t)std::min<uint64_t>(e1, e + kTileEdges)));
  }
  const uint32_t J = (uint32_t)pieces.size();
  const uint64_t nseg2 = (uint64_t)P * J, nseg = (uint64_t)P * B;
  L.pieces.ensure(((size_t)J + B) * sizeof(uint4));
  uint4* d_pieces = L.pieces.as<uint4>();
  uint32_t* d_first = reinterpret_cast<uint32_t*>(d_pieces + J);
  SUMINC_HIP(hipMemcpyAsync(d_pieces, pieces.data(), (size_t)J * sizeof(uint4), hipMemcpyHostToDevice, c->stream));
  SUMINC_HIP(hipMemcpyAsync(d_first, first.data(), (size_t)B * 4, hipMemcpyHostToDevice, c->stream));
  L.hist.ensure(nseg2 * 4);
  L.F2.ensure_slack((nseg2 + 1) * 4);  // the pieces' segment offsets
  L.F.ensure((nseg + 1) * 4);
  uint32_t* hist = L.hist.as<uint32_t>();
  uint32_t* F2 = L.F2.as<uint32_t>();
  uint32_t* F = L.F.as<uint32_t>();
  k_bin_hist<<<J, kBlock, 0, c->stream>>>(d_pieces, J, cols, wl, P, hist);
  SUMINC_HIP(hipGetLastError());
  SUMINC_HIP(hipMemsetAsync(F2, 0, 4, c->stream));
  size_t bytes = 0;
  SUMINC_HIP(rocprim::inclusive_scan(nullptr, bytes, hist, F2 + 1, (size_t)nseg2, rocprim::plus<uint32_t>(),
                                     c->stream));
  c->tmp.ensure(bytes);
  SUMINC_HIP(rocprim::inclusive_scan(c->tmp.p, bytes, hist, F2 + 1, (size_t)nseg2, rocprim::plus<uint32_t>(),
                                     c->stream));
  L.packed.ensure_slack(m_cap * 4 + 16);  // the summing pass reads 4-aligned groups
  if (with_eid) L.eid.ensure_slack(m_cap * 4 + 16);
  k_bin_scatter<<<J, kBlock, 0, c->stream>>>(d_pieces, J, rows_off, cols, wl, P, F2, L.packed.as<uint32_t>(),
                                             with_eid ? L.eid.as<uint32_t>() : nullptr);
  SUMINC_HIP(hipGetLastError());
  // each (window, piece) segment sorted by slot: neighbouring lanes read
  // neighbouring contributions (one line serves the edges of the hot sources), and
  // the edges of one destination (adjacent in arrival order) are spread apart, so
  // a wave's LDS additions seldom meet on one address
  L.packed2.ensure_slack(m_cap * 4 + 16);
  bytes = 0;
  if (with_eid) {
    L.eid2.ensure_slack(m_cap * 4 + 16);
    SUMINC_HIP(rocprim::segmented_radix_sort_pairs(nullptr, bytes, L.packed.as<uint32_t>(), L.packed2.as<uint32_t>(),
                                                   L.eid.as<uint32_t>(), L.eid2.as<uint32_t>(), m_now,
                                                   (unsigned)nseg2, F2, F2 + 1, 0u, wl, c->stream));
    c->tmp.ensure(bytes);
    SUMINC_HIP(rocprim::segmented_radix_sort_pairs(c->tmp.p, bytes, L.packed.as<uint32_t>(), L.packed2.as<uint32_t>(),
                                                   L.eid.as<uint32_t>(), L.eid2.as<uint32_t>(), m_now,
                                                   (unsigned)nseg2, F2, F2 + 1, 0u, wl, c->stream));
    L.eid.swap(L.eid2);
    L.coef.ensure_slack((m_cap + 4) * 8);
  } else {
    SUMINC_HIP(rocprim::segmented_radix_sort_keys(nullptr, bytes, L.packed.as<uint32_t>(), L.packed2.as<uint32_t>(),
                                                  m_now, (unsigned)nseg2, F2, F2 + 1, 0u, wl, c->stream));
    c->tmp.ensure(bytes);
    SUMINC_HIP(rocprim::segmented_radix_sort_keys(c->tmp.p, bytes, L.packed.as<uint32_t>(), L.packed2.as<uint32_t>(),
                                                  m_now, (unsigned)nseg2, F2, F2 + 1, 0u, wl, c->stream));
  }
  L.packed.swap(L.packed2);
  k_bin_fold<<<grid_for(nseg + 1), kBlock, 0, c->stream>>>(F2, J, d_first, B, P, F);
  SUMINC_HIP(hipGetLastError());
  std::vector<uint32_t> hF(nseg + 1);
  SUMINC_HIP(hipMemcpyAsync(hF.data(), F, hF.size() * 4, hipMemcpyDeviceToHost, c->stream));
  SUMINC_HIP(hipStreamSynchronize(c->stream));
  SUMINC_REQUIRE(hF[nseg] == m_now, SUMINC_E_STATE, "binned pull layout does not cover the adjacency");
  // the summing pass's parts: each block's run of 4-edge groups cut into parts of
  // at most `ep` edges, enough parts to fill the chip
  const uint64_t ep = std::min<uint64_t>(1u << 16, std::max<uint64_t>(1u << 12, m_now / 8192));
  std::vector<uint64_t> totals(B, 0);  // groups per block, window by window (sequential in hF)
  for (uint32_t w = 0; w < P; ++w) {
    const uint32_t* fw = hF.data() + (size_t)w * B;
    for (uint32_t b = 0; b < B; ++b) {
      const uint32_t s = fw[b], e = fw[b + 1];
      totals[b] += e > s ? ((e + 3) >> 2) - (s >> 2) : 0;
    }
  }
  std::vector<uint4> parts;
  for (uint32_t b = 0; b < B; ++b) {
    const uint64_t total = totals[b];
    if (!total) continue;
    const uint64_t np = (total + ep - 1) / ep, per = (total + np - 1) / np;
    const uint32_t flag = np > 1 ? 0x80000000u : 0u;
    for (uint64_t z = 0; z < total; z += per)
      parts.push_back(make_uint4(b | flag, 0u, (uint32_t)z, (uint32_t)std::min(total, z + per)));
  }
  L.parts.ensure(parts.size() * sizeof(uint4));
  SUMINC_HIP(hipMemcpyAsync(L.parts.p, parts.data(), parts.size() * sizeof(uint4), hipMemcpyHostToDevice,
                            c->stream));
  const uint32_t direct = direct_param < 0 ? 1u : std::min<uint32_t>(P, (uint32_t)direct_param);
  L.vals.ensure_slack((m_cap + 4) * es);
  SUMINC_HIP(hipStreamSynchronize(c->stream));  // `parts`, `pieces` and `first` go out of scope
  L.n_parts = (uint32_t)parts.size();
  L.P = P;
  L.B = B;
  L.wl = wl;
  L.bsize = bsize;
  L.direct = direct;
  L.first = hF[(size_t)direct * B];
  L.m = m_now;
}

// y[v] += sum over the edges (u, v) of L of x[u] (times the edge's coefficient,
// L.coef, when COEF): the gather pass, then the summing pass.  gate as for the
// pull launches (2: unless cnt is done).
template <class T, bool COEF>
void bin_layout_spmv(suminc_ctx* c, const BinLayout& L, const T* x, T* y, const Counters* cnt, int gate) {
  SUMINC_REQUIRE(L.bsize <= kTileDst, SUMINC_E_STATE, "binned pull blocks exceed the LDS sums");
  SUMINC_REQUIRE(L.vals.cap >= (L.m + 4) * sizeof(T) && (!COEF || L.coef.cap >= (L.m + 4) * sizeof(T)),
                 SUMINC_E_STATE, "binned pull layout built for a smaller element");
  const uint32_t* F = L.F.as<uint32_t>();
  const T* coef = COEF ? L.coef.as<T>() : nullptr;
  const uint64_t first = L.first & ~3ull;  // 4-aligned, for the vector path
  if (L.m > L.first)
    k_pr_bin_gather<T, COEF><<<(unsigned)((L.m - first + kBinGatherEdges - 1) / kBinGatherEdges), kBlock, 0,
                               c->stream>>>(L.packed.as<uint32_t>(), F, L.B, L.P, (uint32_t)first, (uint32_t)L.m,
                                            L.wl, x, coef, L.vals.as<T>(), cnt, gate);
  k_pr_bin_sum<T, COEF><<<L.n_parts, kBinBlock, 0, c->stream>>>(L.parts.as<uint4>(), F, L.B, L.P, L.bsize, c->n,
                                                               L.packed.as<uint32_t>(), L.vals.as<T>(), x, coef,
                                                               L.wl, L.direct, y, cnt, gate);
}
template void bin_layout_spmv<float, true>(suminc_ctx*, const BinLayout&, const float*, float*, const Counters*, int);
template void bin_layout_spmv<double, true>(suminc_ctx*, const BinLayout&, const double*, double*, const Counters*,
                                            int);

// The binned layout of the current graph (after the positions).
static void build_bins(suminc_ctx* c, uint64_t m_cap) {
  c->bins.n_parts = 0;
  const int kib = c->prm.pull_bin_kib;
  const uint64_t es = c->pr_prec == SUMINC_PR_FP64 ? 8 : 4;
  const uint32_t n = c->n;
  if (kib < 0 || c->band_P > 1 || n == 0) return;
  // measured on MI355X (FP32, 2 MiB windows, the row classes beside it): 0.25
  // against 0.38 ms per dense round at 2^22 vertices, 0.86 against 1.70 at 2^24,
  // 4.65 against 8.6 at 2^26
  if (kib == 0 && (uint64_t)n * es <= (8ull << 20)) return;
  const unsigned long long n_src = c->pull_compact ? c->pos_count : n;  // positions in use
  bin_layout_build(c, c->bins, n, c->in_off.as<uint32_t>(), pull_cols(c), n_src, es,
                   (uint64_t)(kib ? kib : 2048) * 1024, c->prm.pull_bin_direct, m_cap, false);
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
  build_bins(c, m_cap);
  c->tile_n = 0;
  if (!c->bins.n_parts) build_tiles(c, m_cap);
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
  if (c->bins.n_parts) {
    bin_layout_spmv<T, false>(c, c->bins, c->pr_contrib.as<T>(), c->pr_delta.as<T>(), cnt, gate);
    return;
  }
  if (c->tile_n) {
    SUMINC_REQUIRE(c->tile_bsize <= kTileDst, SUMINC_E_STATE, "tiled pull blocks exceed the LDS sums");
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
