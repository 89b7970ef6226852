// wpr.hip -- weighted personalized PageRank and its adjoint on the device.
//
// No reference counterpart: the reference's nearest app is the weighted
// personalized PHP Ingress (php/php_ingress.h:15-120), and nothing in it has
// gradients.  The forward keeps the delta-accumulative form of the plain path:
//   r = (1-d) p + d W^T r,  W[u,v] = w_uv / s_u,  s_u = sum_k w_uk (rows of
//   dangling vertices are zero, their mass leaks like the plain path's),
// computed as r = sum_t delta_t with delta_0 = (1-d) p, delta_{t+1} = d W^T delta_t:
// a pull over the in-CSR (CSC) with per-edge coefficient a_e = d w_e / s_src.
// The adjoint of L = sum_i c_i r_i is the same delta iteration on the transpose
// graph (the out-CSR): lambda = sum_t mu_t, mu_0 = c, mu_{t+1} = d W mu_t, and
//   dL/dp_v    = (1-d) lambda_v
//   dL/dw_uv   = (d r_u / s_u) (lambda_v - sum_k W[u,k] lambda_k).
//
// Rows are classified once per graph version and direction (short rows: one
// thread; medium: one wave; long: 2048-edge chunks with one atomic each), so a
// round has no per-edge atomics.  State: FP32 deltas with Kahan-compensated FP32
// sums, or FP64; the stop rule and every residual are FP64 block partials.
#include <cstring>
#include <mutex>
#include <rocprim/rocprim.hpp>

#include "ctx.hpp"
#include "device.hpp"
#include "pr.hpp"

namespace suminc {

// ---- row classes of one direction -----------------------------------------------
__global__ void k_wpr_lists(uint32_t n, const uint32_t* off, uint32_t* short_list, uint32_t* wave_list, uint2* heavy,
                            unsigned long long* counts) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t deg = v < n ? off[v + 1] - off[v] : 0;
  const bool is_short = deg >= 1 && deg <= kPullThread;
  const bool is_wave = deg > kPullThread && deg <= kHeavyDeg;
  unsigned long long s1 = wave_append(is_short, &counts[0]);
  if (is_short) short_list[s1] = v;
  unsigned long long s2 = wave_append(is_wave, &counts[1]);
  if (is_wave) wave_list[s2] = v;
  if (deg > kHeavyDeg) {
    uint32_t nch = (deg + kChunk - 1) / kChunk;
    unsigned long long slot = atomicAdd(&counts[2], (unsigned long long)nch);
    for (uint32_t k = 0; k < nch; ++k) heavy[slot + k] = make_uint2(v, k);
  }
}

struct RowLists {
  const uint32_t* off;
  const uint32_t* col;
  const uint32_t* shrt;
  const uint32_t* wave;
  const uint2* heavy;
  uint32_t n_short, n_wave;
  uint64_t n_heavy;
};

// Sum of a[k] * x[col[k]] over k = k0 + r * stride (r < U, k < e).  The U column
// and coefficient loads are issued before the U gathers, so a thread keeps U
// independent gathers in flight.
template <class T, int U>
__device__ __forceinline__ T spmv_sum(const uint32_t* __restrict__ col, const T* __restrict__ a,
                                      const T* __restrict__ x, uint32_t k0, uint32_t stride, uint32_t e) {
  uint32_t u[U];
  T w[U];
#pragma unroll
  for (int r = 0; r < U; ++r) {
    const uint32_t k = k0 + r * stride;
    const bool in = k < e;
    u[r] = in ? col[k] : 0u;
    w[r] = in ? a[k] : T(0);
  }
  T acc = T(0);
#pragma unroll
  for (int r = 0; r < U; ++r)
    if (w[r] != T(0)) acc += w[r] * x[u[r]];
  return acc;
}

// y[v] = sum_{e in row v} a[e] x[col[e]]; y must be zero before (heavy rows add).
// One launch for the three row classes: workgroups [0, gh) take the heavy
// chunks (one workgroup per chunk, one atomic), [gh, gh + gw) the wave rows (one
// wave per row), the rest the short rows (one thread per row).
template <class T>
__global__ void __launch_bounds__(kBlock) k_spmv(RowLists L, const T* __restrict__ a, const T* __restrict__ x,
                                                 T* __restrict__ y, const Counters* cnt, uint32_t gh, uint32_t gw) {
  __shared__ T s_red[kBlock / 64];
  if (cnt && cnt->done) return;
  const uint32_t bid = blockIdx.x;
  if (bid < gh) {
    for (uint64_t h = bid; h < L.n_heavy; h += gh) {
      const uint2 ch = L.heavy[h];
      const uint32_t b = L.off[ch.x] + ch.y * kChunk;
      const uint32_t e = min(L.off[ch.x + 1], b + kChunk);
      T acc = spmv_sum<T, kChunk / kBlock>(L.col, a, x, b + threadIdx.x, kBlock, e);
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
      if (lane_id() == 0) s_red[threadIdx.x >> 6] = acc;
      __syncthreads();
      if (threadIdx.x == 0) {
        T s = T(0);
        for (int w = 0; w < kBlock / 64; ++w) s += s_red[w];
        if (s != T(0)) atomicAdd(&y[ch.x], s);
      }
      __syncthreads();
    }
  } else if (bid < gh + gw) {
    const uint32_t waves = gw * (kBlock / 64);
    for (uint32_t i = (bid - gh) * (kBlock / 64) + (threadIdx.x >> 6); i < L.n_wave; i += waves) {
      const uint32_t v = L.wave[i];
      const uint32_t b = L.off[v], e = L.off[v + 1];
      T acc = T(0);
      for (uint32_t k = b + lane_id(); k < e; k += 4 * 64) acc += spmv_sum<T, 4>(L.col, a, x, k, 64, e);
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
      if (lane_id() == 0) y[v] = acc;
    }
  } else {
    const uint32_t nb = gridDim.x - gh - gw;
    for (uint32_t i = (bid - gh - gw) * kBlock + threadIdx.x; i < L.n_short; i += nb * kBlock) {
      const uint32_t v = L.shrt[i];
      const uint32_t b = L.off[v], e = L.off[v + 1];
      T acc = T(0);
      for (uint32_t k = b; k < e; k += 4) acc += spmv_sum<T, 4>(L.col, a, x, k, 1, e);
      y[v] = acc;
    }
  }
}

template <class T>
__device__ __forceinline__ void kahan(T* sum, float* comp, uint32_t v, T x) {
  if constexpr (sizeof(T) == 4) {
    float r = sum[v], cc = comp[v];
    float yv = (float)x - cc;
    float t = r + yv;
    comp[v] = (t - r) - yv;
    sum[v] = t;
  } else {
    sum[v] += x;
  }
}

// One vertex pass of a round: sum += x (compensated), x <- y, y <- 0, and the
// FP64 block partial of sum |x_new|.  kStepPer vertices a thread (vertex base +
// j * kBlock + t, coalesced), their x and y loads issued first.
constexpr int kStepPer = 4;

template <class T>
__global__ void __launch_bounds__(kBlock) k_wpr_step(uint32_t n, T* __restrict__ x, T* __restrict__ y,
                                                     T* __restrict__ sum, float* __restrict__ comp,
                                                     double* __restrict__ partial, const Counters* cnt,
                                                     const uint32_t* __restrict__ pos, T* __restrict__ xc) {
  __shared__ double s_p[kBlock / 64];
  if (cnt->done) return;
  const uint32_t base = blockIdx.x * (kBlock * kStepPer);
  T xo[kStepPer], xn[kStepPer];
#pragma unroll
  for (int j = 0; j < kStepPer; ++j) {
    const uint32_t v = base + j * kBlock + threadIdx.x;
    xo[j] = v < n ? x[v] : T(0);
    xn[j] = v < n ? y[v] : T(0);
  }
  double a = 0;
#pragma unroll
  for (int j = 0; j < kStepPer; ++j) {
    const uint32_t v = base + j * kBlock + threadIdx.x;
    if (v < n) {
      if (xo[j] != T(0)) kahan(sum, comp, v, xo[j]);
      x[v] = xn[j];
      y[v] = T(0);
      if (pos && pos[v] != INF32) xc[pos[v]] = xn[j];  // the next round gathers the compact copy
      a += fabs((double)xn[j]);
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o);
  if (lane_id() == 0) s_p[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int w = 0; w < kBlock / 64; ++w) t += s_p[w];
    partial[blockIdx.x] = t;
  }
}

__global__ void __launch_bounds__(1024) k_wpr_decide(const double* __restrict__ partial, uint32_t nparts,
                                                     Counters* cnt, double tol, int max_rounds,
                                                     unsigned long long edges_per_round) {
  __shared__ double s[1024];
  if (cnt->done) return;
  double acc = 0;
  for (uint32_t i = threadIdx.x; i < nparts; i += 1024) acc += partial[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x) return;
  const double r = s[0];
  cnt->residual = r;
  cnt->rounds += 1;
  cnt->edges += edges_per_round;
  if (r != r || r > 1e300) cnt->aux[0] = 1;
  if (!(r >= tol) || (long long)cnt->rounds >= (long long)max_rounds || cnt->aux[0]) cnt->done = 1;
}

// x = scale * seed, y = sum = 0; block partials of sum |x|
template <class T>
__global__ void __launch_bounds__(kBlock) k_wpr_seed(uint32_t n, const float* __restrict__ src, T scale,
                                                     T* __restrict__ x, T* __restrict__ y, T* __restrict__ sum,
                                                     float* __restrict__ comp, double* __restrict__ partial,
                                                     const uint32_t* __restrict__ pos, T* __restrict__ xc) {
  __shared__ double s_p[kBlock / 64];
  const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
  double a = 0;
  if (v < n) {
    const T s = scale * (T)src[v];
    x[v] = s;
    if (pos && pos[v] != INF32) xc[pos[v]] = s;
    y[v] = T(0);
    sum[v] = T(0);
    comp[v] = 0.f;
    a = fabs((double)s);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o);
  if (lane_id() == 0) s_p[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int w = 0; w < kBlock / 64; ++w) t += s_p[w];
    partial[blockIdx.x] = t;
  }
}

// zeroes the counters block (a kernel, not a memset: callers may capture the series)
__global__ void k_wpr_clear(Counters* cnt) {
  if (threadIdx.x == 0) {
    Counters z{};
    *cnt = z;
  }
}

__global__ void __launch_bounds__(1024) k_sum_partials(const double* __restrict__ partial, uint32_t nparts,
                                                       double* out) {
  __shared__ double s[1024];
  double acc = 0;
  for (uint32_t i = threadIdx.x; i < nparts; i += 1024) acc += partial[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

// sum += x at the stop (the last delta belongs to the series)
template <class T>
__global__ void k_wpr_finish(uint32_t n, const T* __restrict__ x, T* __restrict__ sum, float* __restrict__ comp,
                             float* __restrict__ out32, double* __restrict__ out64) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const T xo = x[v];
  if (xo != T(0)) kahan(sum, comp, v, xo);
  double val = (double)sum[v];
  if constexpr (sizeof(T) == 4) val -= (double)comp[v];
  if (out32) out32[v] = (float)val;
  if (out64) out64[v] = val;
}

// Wave-per-row kernels stride over the rows with at most this many workgroups (a
// grid of one wave per row would pass 2^32 threads at 2^26 rows).
constexpr unsigned kWaveGrid = 65536;

// ---- weights: row sums and the per-edge coefficients of both directions ---------
__global__ void __launch_bounds__(kBlock) k_wpr_rowsum(uint32_t n, const uint32_t* __restrict__ off,
                                                       const float* __restrict__ w, double* __restrict__ s) {
  const uint32_t waves = gridDim.x * (kBlock / 64);
  for (uint32_t v = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); v < n; v += waves) {  // whole waves
    double acc = 0;
    for (uint32_t e = off[v] + lane_id(); e < off[v + 1]; e += 64) acc += (double)w[e];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if (lane_id() == 0) s[v] = acc;
  }
}

// a_e = d w_e / s_src for the out-CSR (edge-parallel over rows via the tile rows)
template <class T>
__global__ void __launch_bounds__(kBlock) k_wpr_coef_csr(uint32_t n, const uint32_t* __restrict__ off,
                                                         const float* __restrict__ w, const double* __restrict__ s,
                                                         double d, T* __restrict__ a) {
  const uint32_t waves = gridDim.x * (kBlock / 64);
  for (uint32_t v = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); v < n; v += waves) {
    const double inv = s[v] > 0 ? d / s[v] : 0.0;
    for (uint32_t e = off[v] + lane_id(); e < off[v + 1]; e += 64) a[e] = (T)((double)w[e] * inv);
  }
}

// the in-CSR copy: a_csc[i] = a_csr[eid[i]]
template <class T>
__global__ void k_wpr_gather(uint64_t m, const uint32_t* __restrict__ eid, const T* __restrict__ src,
                             T* __restrict__ dst) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) dst[i] = src[eid[i]];
}

__global__ void k_wpr_gather_f(uint64_t m, const uint32_t* __restrict__ eid, const float* __restrict__ src,
                               uint32_t* __restrict__ dst) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) dst[i] = __float_as_uint(src[eid[i]]);
}

// CSR edge id of every CSC position: sort (dst, src) keys with the edge index
__global__ void k_wpr_edge_keys(const uint32_t* __restrict__ off, uint32_t n, uint64_t m,
                                const uint32_t* __restrict__ col, unsigned long long* keys, uint32_t* idx) {
  const uint32_t waves = gridDim.x * (kBlock / 64);
  for (uint32_t v = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); v < n; v += waves) {
    for (uint32_t e = off[v] + lane_id(); e < off[v + 1]; e += 64) {
      keys[e] = ((unsigned long long)col[e] << 32) | v;
      idx[e] = e;
    }
  }
  (void)m;
}

// ---- gradients -------------------------------------------------------------------
// grad_w[e] for CSR edge e = (u -> v): (d r_u / s_u) lambda_v - (r_u / s_u) y_u with
// y_u = sum_k a_uk lambda_k = d sum_k W[u,k] lambda_k.  One wave per row.
__global__ void __launch_bounds__(kBlock) k_wpr_grad_w(uint32_t n, const uint32_t* __restrict__ off,
                                                       const uint32_t* __restrict__ col, const double* __restrict__ s,
                                                       const double* __restrict__ r, const double* __restrict__ lam,
                                                       const double* __restrict__ yl, double d,
                                                       float* __restrict__ g32, double* __restrict__ g64) {
  const uint32_t waves = gridDim.x * (kBlock / 64);
  for (uint32_t u = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); u < n; u += waves) {
    const double ru = s[u] > 0 ? r[u] / s[u] : 0.0;  // a row without weight mass has no gradient
    const double sub = ru * yl[u];
    for (uint32_t e = off[u] + lane_id(); e < off[u + 1]; e += 64) {
      const double g = d * ru * lam[col[e]] - sub;
      if (g32) g32[e] = (float)g;
      if (g64) g64[e] = g;
    }
  }
}

__global__ void k_wpr_grad_p(uint32_t n, const double* __restrict__ lam, double d, float* __restrict__ g32,
                             double* __restrict__ g64) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const double g = (1.0 - d) * lam[v];
  if (g32) g32[v] = (float)g;
  if (g64) g64[v] = g;
}

template <class T>
__global__ void k_to_t(uint32_t n, const double* __restrict__ src, T* __restrict__ dst) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) dst[v] = (T)src[v];
}

template <class T>
__global__ void k_to_f64(uint32_t n, const T* __restrict__ src, double* __restrict__ dst) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) dst[v] = (double)src[v];
}

// ---------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------
// Row lists of one direction and, on large graphs, compact positions of the
// columns (by decreasing number of appearances) with the columns renamed to them,
// so a round gathers from an array of the vertices that can be gathered at all.
struct WprDir {
  DevBuf shrt, wave, heavy, counts;
  uint64_t n_short = 0, n_wave = 0, n_heavy = 0;
  bool compact = false;
  DevBuf pos, scratch, colc;
};

static void build_lists(suminc_ctx* c, WprDir& L, const DevBuf& off) {
  const uint32_t n = c->n;
  L.shrt.ensure((size_t)n * 4 + 4);
  L.wave.ensure((size_t)n * 4 + 4);
  L.heavy.ensure((c->m / kChunk + c->m / kHeavyDeg + 2) * sizeof(uint2));
  L.counts.ensure(32);
  SUMINC_HIP(hipMemsetAsync(L.counts.p, 0, 32, c->stream));
  k_wpr_lists<<<grid_for(n), kBlock, 0, c->stream>>>(n, off.as<uint32_t>(), L.shrt.as<uint32_t>(),
                                                     L.wave.as<uint32_t>(), L.heavy.as<uint2>(),
                                                     L.counts.as<unsigned long long>());
  unsigned long long h[3];
  SUMINC_HIP(hipMemcpyAsync(h, L.counts.p, 24, hipMemcpyDeviceToHost, c->stream));
  SUMINC_HIP(hipStreamSynchronize(c->stream));
  L.n_short = h[0];
  L.n_wave = h[1];
  L.n_heavy = h[2];
}

struct WprState {
  uint64_t graph_version = ~0ull;    // version the lists / edge ids were built for
  uint64_t coef_version = ~0ull;     // weights version the coefficients belong to
  int coef_prec = -1;
  uint64_t weights_version = 0;
  bool fwd_ready = false;
  uint64_t fwd_coef_version = ~0ull;
  WprDir out, in;
  DevBuf eid;                        // CSR edge id per CSC position
  DevBuf s;                          // row sums (f64)
  DevBuf a_out, a_in;                // coefficients d w / s (T) in CSR / CSC order
  DevBuf x, y, sum, comp;            // round state (T)
  DevBuf xc;                         // x at the compact positions (large graphs)
  DevBuf r64, lam64, yl64;           // FP64 results of the last forward / backward
  DevBuf l1;                         // sum |seed|
  DevBuf php_a_out, php_a_in, php64;  // PHP: coefficients (T) in CSR / CSC order, FP64 result
  // binned series rounds (large graphs, the PageRank pull's rule): the forward over
  // the CSC, the adjoint over the CSR, their coefficients in layout order; built for
  // elements of bin_es bytes
  BinLayout bin_in, bin_out;
  uint64_t bin_es = 0;
};

static WprState& wstate(suminc_ctx* c) {
  if (!c->wpr) c->wpr = new WprState();
  return *reinterpret_cast<WprState*>(c->wpr);
}

void wpr_release(suminc_ctx* c) {
  delete reinterpret_cast<WprState*>(c->wpr);
  c->wpr = nullptr;
}

static void wpr_prepare_graph(suminc_ctx* c, WprState& S) {
  const uint64_t es = c->prm.pr_precision == SUMINC_PR_FP64 ? 8 : 4;
  if (S.graph_version == c->graph_version && S.bin_es == es) return;
  const uint32_t n = c->n;
  const uint64_t m = c->m;
  build_lists(c, S.out, c->out_off);
  build_lists(c, S.in, c->in_off);
  // the forward gathers over the CSC (columns: sources, counted by out-degree), the
  // adjoint over the CSR (columns: destinations, counted by in-degree)
  const bool compact = compact_pull(c, c->prm.pr_precision == SUMINC_PR_FP64 ? 8 : 4);
  S.in.compact = S.out.compact = compact && m > 0;
  if (S.in.compact) {
    build_positions_by_degree(c, c->out_off.as<uint32_t>(), c->in_off.as<uint32_t>(), c->in_col.as<uint32_t>(), m,
                              S.in.pos, S.in.scratch, S.in.colc);
    build_positions_by_degree(c, c->in_off.as<uint32_t>(), c->out_off.as<uint32_t>(), c->out_col.as<uint32_t>(),
                              m, S.out.pos, S.out.scratch, S.out.colc);
  }
  S.eid.ensure(m * 4 + 4);
  if (m) {
    DevBuf ka, kb, ia;
    ka.ensure(m * 8); kb.ensure(m * 8); ia.ensure(m * 4);
    k_wpr_edge_keys<<<grid_for(n, kBlock / 64, kWaveGrid), kBlock, 0, c->stream>>>(c->out_off.as<uint32_t>(), n, m,
                                                                       c->out_col.as<uint32_t>(),
                                                                       ka.as<unsigned long long>(), ia.as<uint32_t>());
    // keys (dst << 32 | src): the bits above 32 + log2(n) are zero
    int hi_bits = 1;
    while (hi_bits < 32 && (1ull << hi_bits) < n) ++hi_bits;
    const unsigned end_bit = 32 + hi_bits;
    size_t bytes = 0;
    SUMINC_HIP(rocprim::radix_sort_pairs(nullptr, bytes, ka.as<uint64_t>(), kb.as<uint64_t>(), ia.as<uint32_t>(),
                                         S.eid.as<uint32_t>(), (size_t)m, 0, end_bit, c->stream));
    c->tmp.ensure(bytes);
    SUMINC_HIP(rocprim::radix_sort_pairs(c->tmp.p, bytes, ka.as<uint64_t>(), kb.as<uint64_t>(), ia.as<uint32_t>(),
                                         S.eid.as<uint32_t>(), (size_t)m, 0, end_bit, c->stream));
    SUMINC_HIP(hipStreamSynchronize(c->stream));
  }
  // binned series rounds, by the PageRank pull's rule (on above 8 MiB of elements
  // or when a window is given): the gather pass multiplies by the coefficient
  S.bin_in.n_parts = S.bin_out.n_parts = 0;
  const int kib = c->prm.pull_bin_kib;
  if (m && kib >= 0 && (kib > 0 || (uint64_t)n * es > (8ull << 20))) {
    const uint64_t win = (uint64_t)(kib ? kib : 2048) * 1024;
    const unsigned long long src_in =
        S.in.compact ? nonempty_rows(c, c->out_off.as<uint32_t>(), S.bin_in.hist) : n;
    bin_layout_build(c, S.bin_in, n, c->in_off.as<uint32_t>(),
                     S.in.compact ? S.in.colc.as<uint32_t>() : c->in_col.as<uint32_t>(), src_in, es, win,
                     c->prm.pull_bin_direct, m, true);
    const unsigned long long src_out =
        S.out.compact ? nonempty_rows(c, c->in_off.as<uint32_t>(), S.bin_out.hist) : n;
    bin_layout_build(c, S.bin_out, n, c->out_off.as<uint32_t>(),
                     S.out.compact ? S.out.colc.as<uint32_t>() : c->out_col.as<uint32_t>(), src_out, es, win,
                     c->prm.pull_bin_direct, m, true);
  }
  S.bin_es = es;
  S.graph_version = c->graph_version;
  S.coef_version = ~0ull;
  S.fwd_ready = false;
}

template <class T>
static void wpr_prepare_coef(suminc_ctx* c, WprState& S) {
  const int prec = sizeof(T) == 8 ? SUMINC_PR_FP64 : SUMINC_PR_FP32_KAHAN;
  if (S.coef_version == S.weights_version && S.coef_prec == prec) return;
  const uint32_t n = c->n;
  const uint64_t m = c->m;
  S.s.ensure((size_t)n * 8 + 8);
  S.a_out.ensure(m * sizeof(T) + 8);
  S.a_in.ensure(m * sizeof(T) + 8);
  const unsigned gw = grid_for(n, kBlock / 64, kWaveGrid);
  k_wpr_rowsum<<<gw, kBlock, 0, c->stream>>>(n, c->out_off.as<uint32_t>(), c->out_w.as<float>(), S.s.as<double>());
  k_wpr_coef_csr<T><<<gw, kBlock, 0, c->stream>>>(n, c->out_off.as<uint32_t>(), c->out_w.as<float>(),
                                                  S.s.as<double>(), c->prm.damping, S.a_out.as<T>());
  if (m) k_wpr_gather<T><<<grid_for(m), kBlock, 0, c->stream>>>(m, S.eid.as<uint32_t>(), S.a_out.as<T>(), S.a_in.as<T>());
  SUMINC_HIP(hipGetLastError());
  S.coef_version = S.weights_version;
  S.coef_prec = prec;
}

// compact: the renamed columns when the direction has them (then the gathered
// vector must be at the compact positions)
static RowLists lists_of(const WprDir& L, const DevBuf& off, const DevBuf& col, bool compact) {
  RowLists R;
  R.off = off.as<uint32_t>();
  R.col = compact && L.compact ? L.colc.as<uint32_t>() : col.as<uint32_t>();
  R.shrt = L.shrt.as<uint32_t>();
  R.wave = L.wave.as<uint32_t>();
  R.heavy = L.heavy.as<uint2>();
  R.n_short = (uint32_t)L.n_short;
  R.n_wave = (uint32_t)L.n_wave;
  R.n_heavy = L.n_heavy;
  return R;
}

template <class T>
static void spmv(suminc_ctx* c, const RowLists& R, const T* a, const T* x, T* y, const Counters* cnt) {
  const uint32_t gh = R.n_heavy ? push_grid() : 0u;
  const uint32_t gw = R.n_wave ? grid_for(R.n_wave, kBlock / 64, 8192) : 0u;
  const uint32_t gs = R.n_short ? grid_for(R.n_short, kBlock, 8192) : 0u;
  if (gh + gw + gs) k_spmv<T><<<gh + gw + gs, kBlock, 0, c->stream>>>(R, a, x, y, cnt, gh, gw);
}

// Device buffers of one series run: the round state (x, y, the compensated sum),
// the compact copy of x (large graphs), the FP64 block partials and sum |seed|.
template <class T>
struct SeriesBufs {
  T *x, *y, *sum, *xc;
  float* comp;
  double *partial, *l1;
};

// Bytes of the series buffers for n vertices (partials: one per vertex block, three
// words each, as the PageRank rounds size them).
template <class T>
static size_t series_bytes(uint32_t n, bool compact) {
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t g = grid_for(n);
  return 3 * al((size_t)n * sizeof(T) + 8) + (compact ? al((size_t)n * sizeof(T) + 8) : 0) + al((size_t)n * 4 + 4) +
         al((g + 1) * 8 * 3) + al(8);
}

// Carves the series buffers out of `base` (series_bytes<T>(n, compact) bytes).
template <class T>
static SeriesBufs<T> series_carve(unsigned char* base, uint32_t n, bool compact) {
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t g = grid_for(n);
  SeriesBufs<T> B;
  size_t o = 0;
  auto take = [&](size_t bytes) { unsigned char* p = base + o; o += al(bytes); return p; };
  B.x = reinterpret_cast<T*>(take((size_t)n * sizeof(T) + 8));
  B.y = reinterpret_cast<T*>(take((size_t)n * sizeof(T) + 8));
  B.sum = reinterpret_cast<T*>(take((size_t)n * sizeof(T) + 8));
  B.xc = compact ? reinterpret_cast<T*>(take((size_t)n * sizeof(T) + 8)) : nullptr;
  B.comp = reinterpret_cast<float*>(take((size_t)n * 4 + 4));
  B.partial = reinterpret_cast<double*>(take((g + 1) * 8 * 3));
  B.l1 = reinterpret_cast<double*>(take(8));
  return B;
}

// The series buffers in the context's weighted-PageRank state (grown on demand).
template <class T>
static SeriesBufs<T> series_state(suminc_ctx* c, WprState& S, bool compact) {
  const uint32_t n = c->n;
  S.x.ensure((size_t)n * sizeof(T) + 8);
  if (compact) S.xc.ensure((size_t)n * sizeof(T) + 8);
  S.y.ensure((size_t)n * sizeof(T) + 8);
  S.sum.ensure((size_t)n * sizeof(T) + 8);
  S.comp.ensure((size_t)n * 4 + 4);
  S.l1.ensure(8);
  c->partial.ensure((size_t)(grid_for(n) + 1) * 8 * 3);
  SeriesBufs<T> B;
  B.x = S.x.as<T>();
  B.y = S.y.as<T>();
  B.sum = S.sum.as<T>();
  B.xc = compact ? S.xc.as<T>() : nullptr;
  B.comp = S.comp.as<float>();
  B.partial = c->partial.as<double>();
  B.l1 = S.l1.as<double>();
  return B;
}

// Runs sum = sum_t x_t with x_{t+1} = A x_t from x_0 = scale * seed until
// sum|x| < tol.  Result (FP64) in out64, optionally FP32 in out32 (either may be
// null).  pos != nullptr: R's columns are compact positions, and the rounds
// gather the compact copy of x (B.xc).  Allocates nothing itself.
template <class T>
static void wpr_series(suminc_ctx* c, const SeriesBufs<T>& B, const RowLists& R, const BinLayout* L,
                       const uint32_t* pos, const T* a, const float* seed_dev, T scale, double tol_abs, bool relative,
                       float* out32, double* out64, suminc_stats* st) {
  const uint32_t n = c->n;
  // the binned layout of R's direction when it exists: its coefficients taken from a
  // in layout order, the rounds through the gather and summing passes
  const bool binned = L && L->n_parts;
  if (binned && L->m)
    k_wpr_gather<T><<<grid_for(L->m), kBlock, 0, c->stream>>>(L->m, L->eid.as<uint32_t>(), a, L->coef.as<T>());
  T* xc = pos ? B.xc : nullptr;
  const unsigned g = grid_for(n);
  const unsigned gs = grid_for(n, kBlock * kStepPer);  // the round's vertex pass and its partials
  Counters* cnt = dev_counters(c);
  k_wpr_clear<<<1, 64, 0, c->stream>>>(cnt);
  k_wpr_seed<T><<<g, kBlock, 0, c->stream>>>(n, seed_dev, scale, B.x, B.y, B.sum, B.comp, B.partial, pos, xc);
  k_sum_partials<<<1, 1024, 0, c->stream>>>(B.partial, g, B.l1);
  double tol = tol_abs;
  if (relative) {
    check_begin(c, B.l1, 8);
    tol = tol_abs * *reinterpret_cast<const double*>(check_wait(c));
  }
  const int max_rounds = c->prm.max_rounds > 0 ? c->prm.max_rounds : 100000;
  SUMINC_HIP(hipEventRecord(c->ev_start, c->stream));
  int group = 8;
  while (true) {
    for (int i = 0; i < group; ++i) {
      if (binned) bin_layout_spmv<T, true>(c, *L, pos ? xc : B.x, B.y, cnt, 2);
      else spmv<T>(c, R, a, pos ? xc : B.x, B.y, cnt);
      k_wpr_step<T><<<gs, kBlock, 0, c->stream>>>(n, B.x, B.y, B.sum, B.comp, B.partial, cnt, pos, xc);
      k_wpr_decide<<<1, 1024, 0, c->stream>>>(B.partial, gs, cnt, tol, max_rounds, c->m);
    }
    SUMINC_HIP(hipGetLastError());
    read_counters(c);
    if (c->host_counters->done) break;
    if (group < 64) group *= 2;
  }
  k_wpr_finish<T><<<g, kBlock, 0, c->stream>>>(n, B.x, B.sum, B.comp, out32, out64);
  SUMINC_HIP(hipGetLastError());
  SUMINC_HIP(hipEventRecord(c->ev_stop, c->stream));
  SUMINC_HIP(hipEventSynchronize(c->ev_stop));
  float ms = 0;
  SUMINC_HIP(hipEventElapsedTime(&ms, c->ev_start, c->ev_stop));
  const Counters* h = c->host_counters;
  SUMINC_REQUIRE(!h->aux[0], SUMINC_E_NUMERIC, "NaN/Inf in the weighted PageRank state");
  SUMINC_REQUIRE(!(h->residual >= tol), SUMINC_E_NOT_CONVERGED, "round cap reached before sum|delta| < tolerance");
  suminc_stats out{};
  out.rounds = (double)h->rounds;
  out.edges_traversed = (double)h->edges;
  out.residual = h->residual;
  out.kernel_ms = ms;
  out.pull_rounds = (double)h->rounds;
  c->stats = out;
  if (st) *st = out;
}

static void check_weighted(suminc_ctx* c) {
  SUMINC_REQUIRE(c->has_graph, SUMINC_E_STATE, "no graph");
  SUMINC_REQUIRE(c->wkind == SUMINC_W_F32, SUMINC_E_STATE, "weighted PageRank needs SUMINC_W_F32 edge weights");
}

void wpr_set_weights(suminc_ctx* c, const float* w) {
  check_weighted(c);
  SUMINC_REQUIRE(w || c->m == 0, SUMINC_E_INVALID, "null weights");
  WprState& S = wstate(c);
  if (c->m) {
    SUMINC_HIP(hipMemcpyAsync(c->out_w.p, w, c->m * 4, hipMemcpyDefault, c->stream));
    wpr_prepare_graph(c, S);
    k_wpr_gather_f<<<grid_for(c->m), kBlock, 0, c->stream>>>(c->m, S.eid.as<uint32_t>(), c->out_w.as<float>(),
                                                             c->in_w.as<uint32_t>());
    SUMINC_HIP(hipGetLastError());
    SUMINC_HIP(hipStreamSynchronize(c->stream));
  }
  S.weights_version++;
  S.fwd_ready = false;
}

template <class T>
static void wpr_forward_t(suminc_ctx* c, const float* p, float* r, suminc_stats* st) {
  WprState& S = wstate(c);
  wpr_prepare_graph(c, S);
  wpr_prepare_coef<T>(c, S);
  const uint32_t n = c->n;
  DevBuf pd, r32;
  pd.ensure((size_t)n * 4);
  SUMINC_HIP(hipMemcpyAsync(pd.p, p, (size_t)n * 4, hipMemcpyDefault, c->stream));
  S.r64.ensure((size_t)n * 8 + 8);
  r32.ensure((size_t)n * 4 + 4);
  const RowLists R = lists_of(S.in, c->in_off, c->in_col, true);
  const uint32_t* pos = S.in.compact ? S.in.pos.as<uint32_t>() : nullptr;
  wpr_series<T>(c, series_state<T>(c, S, pos != nullptr), R, &S.bin_in, pos, S.a_in.as<T>(), pd.as<float>(),
                (T)(1.0 - c->prm.damping), c->prm.termcheck_threshold,
                false, r32.as<float>(), S.r64.as<double>(), st);
  if (r) SUMINC_HIP(hipMemcpyAsync(r, r32.p, (size_t)n * 4, hipMemcpyDefault, c->stream));
  SUMINC_HIP(hipStreamSynchronize(c->stream));
  S.fwd_ready = true;
  S.fwd_coef_version = S.weights_version;
}

template <class T>
static void wpr_backward_t(suminc_ctx* c, const float* cvec, float* grad_p, float* grad_w, suminc_stats* st) {
  WprState& S = wstate(c);
  SUMINC_REQUIRE(S.fwd_ready && S.fwd_coef_version == S.weights_version && S.graph_version == c->graph_version,
                 SUMINC_E_STATE, "backward needs a forward on the current graph and weights");
  wpr_prepare_coef<T>(c, S);
  const uint32_t n = c->n;
  const uint64_t m = c->m;
  DevBuf cd, gp32, gw32;
  cd.ensure((size_t)n * 4);
  SUMINC_HIP(hipMemcpyAsync(cd.p, cvec, (size_t)n * 4, hipMemcpyDefault, c->stream));
  S.lam64.ensure((size_t)n * 8 + 8);
  S.yl64.ensure((size_t)n * 8 + 8);
  const RowLists R = lists_of(S.out, c->out_off, c->out_col, true);
  const uint32_t* pos = S.out.compact ? S.out.pos.as<uint32_t>() : nullptr;
  wpr_series<T>(c, series_state<T>(c, S, pos != nullptr), R, &S.bin_out, pos, S.a_out.as<T>(), cd.as<float>(),
                (T)1, c->prm.termcheck_threshold, true, nullptr,
                S.lam64.as<double>(), st);
  // y_u = sum_k a_uk lambda_k (one more product, in FP64)
  DevBuf lam_t, y_t;
  lam_t.ensure((size_t)n * sizeof(T) + 8);
  y_t.ensure((size_t)n * sizeof(T) + 8);
  SUMINC_HIP(hipMemsetAsync(y_t.p, 0, (size_t)n * sizeof(T), c->stream));
  k_to_t<T><<<grid_for(n), kBlock, 0, c->stream>>>(n, S.lam64.as<double>(), lam_t.as<T>());
  spmv<T>(c, lists_of(S.out, c->out_off, c->out_col, false), S.a_out.as<T>(), lam_t.as<T>(), y_t.as<T>(), nullptr);
  DevBuf yd;
  yd.ensure((size_t)n * 8 + 8);
  k_to_f64<T><<<grid_for(n), kBlock, 0, c->stream>>>(n, y_t.as<T>(), yd.as<double>());
  SUMINC_HIP(hipGetLastError());
  gp32.ensure((size_t)n * 4 + 4);
  gw32.ensure(m * 4 + 4);
  k_wpr_grad_p<<<grid_for(n), kBlock, 0, c->stream>>>(n, S.lam64.as<double>(), c->prm.damping, gp32.as<float>(),
                                                      nullptr);
  k_wpr_grad_w<<<grid_for(n, kBlock / 64, kWaveGrid), kBlock, 0, c->stream>>>(
      n, c->out_off.as<uint32_t>(), c->out_col.as<uint32_t>(), S.s.as<double>(), S.r64.as<double>(),
      S.lam64.as<double>(), yd.as<double>(), c->prm.damping, gw32.as<float>(), nullptr);
  SUMINC_HIP(hipGetLastError());
  if (grad_p) SUMINC_HIP(hipMemcpyAsync(grad_p, gp32.p, (size_t)n * 4, hipMemcpyDefault, c->stream));
  if (grad_w && m) SUMINC_HIP(hipMemcpyAsync(grad_w, gw32.p, m * 4, hipMemcpyDefault, c->stream));
  SUMINC_HIP(hipStreamSynchronize(c->stream));
}

void wpr_forward(suminc_ctx* c, const float* p, float* r, suminc_stats* st) {
  check_weighted(c);
  SUMINC_REQUIRE(p, SUMINC_E_INVALID, "null personalization");
  if (c->prm.pr_precision == SUMINC_PR_FP64) wpr_forward_t<double>(c, p, r, st);
  else wpr_forward_t<float>(c, p, r, st);
}

void wpr_backward(suminc_ctx* c, const float* cvec, float* grad_p, float* grad_w, suminc_stats* st) {
  check_weighted(c);
  SUMINC_REQUIRE(cvec, SUMINC_E_INVALID, "null loss weights");
  if (c->prm.pr_precision == SUMINC_PR_FP64) wpr_backward_t<double>(c, cvec, grad_p, grad_w, st);
  else wpr_backward_t<float>(c, cvec, grad_p, grad_w, st);
}

// ---- PHP: penalized hitting probability on the series kernels --------------------
// PHPIngress (php/php_ingress.h:15-120): value = sum of the deltas a vertex takes,
// delta_0 = 1 at the source and 0 elsewhere, and a vertex u sends
// d * w_uv / s_u of its delta along every out-edge u -> v with v != source, where
// s_u = the sum of u's out-edge weights to vertices other than the source
// (frag.weight_sum, :30-45); nothing enters the source again.  That is the weighted
// series x_{t+1} = A x_t with a_uv = d w_uv / s_u (0 into the source), x_0 = e_src,
// value = sum_t x_t: the weighted-PageRank rounds over the same CSC row classes.
template <class T>
__global__ void __launch_bounds__(kBlock) k_php_coef_csr(uint32_t n, const uint32_t* __restrict__ off,
                                                         const uint32_t* __restrict__ col,
                                                         const uint32_t* __restrict__ wu,
                                                         const float* __restrict__ wf, uint32_t src, double d,
                                                         T* __restrict__ a) {
  const uint32_t waves = gridDim.x * (kBlock / 64);
  for (uint32_t v = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); v < n; v += waves) {  // whole waves
    double acc = 0;
    for (uint32_t e = off[v] + lane_id(); e < off[v + 1]; e += 64)
      if (col[e] != src) acc += wu ? (double)wu[e] : wf ? (double)wf[e] : 1.0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    const double inv = acc > 0 ? d / acc : 0.0;
    for (uint32_t e = off[v] + lane_id(); e < off[v + 1]; e += 64) {
      const double w = wu ? (double)wu[e] : wf ? (double)wf[e] : 1.0;
      a[e] = col[e] != src ? (T)(w * inv) : T(0);
    }
  }
}

__global__ void k_php_seed(uint32_t n, uint32_t src, float* __restrict__ p) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) p[v] = v == src ? 1.f : 0.f;
}

template <class T>
static void php_run_t(suminc_ctx* c, uint32_t src, double d, double* values, suminc_stats* st) {
  WprState& S = wstate(c);
  wpr_prepare_graph(c, S);
  const uint32_t n = c->n;
  const uint64_t m = c->m;
  S.php_a_out.ensure(m * sizeof(T) + 8);
  S.php_a_in.ensure(m * sizeof(T) + 8);
  S.php64.ensure((size_t)n * 8 + 8);
  const uint32_t* wu = c->wkind == SUMINC_W_U32 ? c->out_w.as<uint32_t>() : nullptr;
  const float* wf = c->wkind == SUMINC_W_F32 ? c->out_w.as<float>() : nullptr;
  k_php_coef_csr<T><<<grid_for(n, kBlock / 64, kWaveGrid), kBlock, 0, c->stream>>>(
      n, c->out_off.as<uint32_t>(), c->out_col.as<uint32_t>(), wu, wf, src, d, S.php_a_out.as<T>());
  if (m)
    k_wpr_gather<T><<<grid_for(m), kBlock, 0, c->stream>>>(m, S.eid.as<uint32_t>(), S.php_a_out.as<T>(),
                                                         S.php_a_in.as<T>());
  DevBuf seed;
  seed.ensure((size_t)n * 4 + 4);
  k_php_seed<<<grid_for(n), kBlock, 0, c->stream>>>(n, src, seed.as<float>());
  SUMINC_HIP(hipGetLastError());
  const RowLists R = lists_of(S.in, c->in_off, c->in_col, true);
  const uint32_t* pos = S.in.compact ? S.in.pos.as<uint32_t>() : nullptr;
  wpr_series<T>(c, series_state<T>(c, S, pos != nullptr), R, &S.bin_in, pos, S.php_a_in.as<T>(),
                seed.as<float>(), (T)1, c->prm.termcheck_threshold, false, nullptr, S.php64.as<double>(), st);
  if (values) SUMINC_HIP(hipMemcpyAsync(values, S.php64.p, (size_t)n * 8, hipMemcpyDefault, c->stream));
  SUMINC_HIP(hipStreamSynchronize(c->stream));
}

void php_run(suminc_ctx* c, uint32_t source, double damping, double* values, suminc_stats* st) {
  SUMINC_REQUIRE(c->has_graph, SUMINC_E_STATE, "no graph");
  SUMINC_REQUIRE(source < c->n, SUMINC_E_INVALID, "source out of range");
  SUMINC_REQUIRE(damping > 0.0 && damping < 1.0, SUMINC_E_INVALID, "damping must be in (0, 1)");
  if (c->prm.pr_precision == SUMINC_PR_FP64) php_run_t<double>(c, source, damping, values, st);
  else php_run_t<float>(c, source, damping, values, st);
}

// ---- the allocation-free forward -------------------------------------------------
// Its round state lives in `scratch`, one fixed array per GPU that the library
// holds until the process ends.  Host code never passes the array's name to a
// kernel: wpr_scratch_setup() takes the array's device address once per device
// (with that device current) and every launch uses that address.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char* g_scratch_dev[kMaxDevices] = {};

void wpr_scratch_setup(int device) {
  SUMINC_REQUIRE(device >= 0 && device < kMaxDevices, SUMINC_E_INVALID, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_dev[device]) return;
  void* p = nullptr;
  SUMINC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  SUMINC_REQUIRE(p, SUMINC_E_HIP, "no device address for the scratch array");
  g_scratch_dev[device] = reinterpret_cast<unsigned char*>(p);
}

static unsigned char* scratch_of(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return device >= 0 && device < kMaxDevices ? g_scratch_dev[device] : nullptr;
}

static bool on_device(const void* p) {
  hipPointerAttribute_t a{};
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

template <class T>
static void wpr_forward_noalloc_t(suminc_ctx* c, const float* p, float* r, suminc_stats* st) {
  SUMINC_REQUIRE(c->wpr, SUMINC_E_STATE, "no weighted PageRank state: call suminc_wpr_set_weights and one "
                                        "suminc_wpr_forward first");
  WprState& S = *reinterpret_cast<WprState*>(c->wpr);
  const int prec = sizeof(T) == 8 ? SUMINC_PR_FP64 : SUMINC_PR_FP32_KAHAN;
  SUMINC_REQUIRE(S.graph_version == c->graph_version && S.coef_version == S.weights_version && S.coef_prec == prec,
                 SUMINC_E_STATE, "the coefficients of the current graph, weights and precision are not prepared: "
                                 "run one suminc_wpr_forward first");
  const uint32_t n = c->n;
  const bool compact = S.in.compact;
  const size_t need = series_bytes<T>(n, compact);
  if (need > kScratchBytes)
    throw Error(SUMINC_E_TOO_LARGE, "suminc_wpr_forward_noalloc: " + std::to_string(n) + " vertices need " +
                                        std::to_string(need) + " bytes of scratch, the fixed array holds " +
                                        std::to_string(kScratchBytes));
  unsigned char* base = scratch_of(c->device);
  SUMINC_REQUIRE(base, SUMINC_E_STATE, "scratch array not set up on this device");
  const SeriesBufs<T> B = series_carve<T>(base, n, compact);
  const RowLists R = lists_of(S.in, c->in_off, c->in_col, true);
  const uint32_t* pos = compact ? S.in.pos.as<uint32_t>() : nullptr;
  S.fwd_ready = false;  // the rank of this forward is not kept for a backward
  wpr_series<T>(c, B, R, &S.bin_in, pos, S.a_in.as<T>(), p, (T)(1.0 - c->prm.damping), c->prm.termcheck_threshold,
                false, r,
                nullptr, st);
  SUMINC_HIP(hipStreamSynchronize(c->stream));
}

void wpr_forward_noalloc(suminc_ctx* c, const float* p, float* r, suminc_stats* st) {
  check_weighted(c);
  SUMINC_REQUIRE(p && r, SUMINC_E_INVALID, "null personalization or rank");
  SUMINC_REQUIRE(on_device(p) && on_device(r), SUMINC_E_INVALID,
                 "suminc_wpr_forward_noalloc needs device memory for the personalization and the rank");
  if (c->prm.pr_precision == SUMINC_PR_FP64) wpr_forward_noalloc_t<double>(c, p, r, st);
  else wpr_forward_noalloc_t<float>(c, p, r, st);
}

void wpr_download(suminc_ctx* c, double* r, double* lam) {
  WprState& S = wstate(c);
  SUMINC_REQUIRE(S.fwd_ready, SUMINC_E_STATE, "no weighted PageRank forward yet");
  if (r) SUMINC_HIP(hipMemcpy(r, S.r64.p, (size_t)c->n * 8, hipMemcpyDeviceToHost));
  if (lam) {
    SUMINC_REQUIRE(S.lam64.p, SUMINC_E_STATE, "no backward yet");
    SUMINC_HIP(hipMemcpy(lam, S.lam64.p, (size_t)c->n * 8, hipMemcpyDeviceToHost));
  }
}

}  // namespace suminc
