// rsketch.hip -- the SumInc PageRank sketch phase over a sketch in the reference
// compressor's own layout: the node types 0..6 (grape/types.h:70-78), the CSR
// products is_e_ (each supernode source's bound_delta: entries and in-mirrors),
// ib_e_ (exits' and Singles' external out-edges), sync_e_ (master -> in-mirrors,
// out-mirror -> master), the mirror -> master map and each supernode's
// inner_value / inner_delta (compressor_base.h:4636-4899, 5974-5987;
// iter_compressor.h:433-491).  Mirror ids are n .. all_node_num - 1.
//
// One sketch round (sum_sync_iter_worker.h:949-1127; cudaCode/my_worker.cu:
// 235-293, 331-437), on the context stream:
//   k_rs_send    Singles: delta -> rank, d*delta/outdeg onto their ib_e_ row (push
//                items); entries (OnlyIn, BothOutIn, BothOutInMaster): delta ->
//                spnode, one shortcut item (source, delta)
//   k_rs_rows    shortcut items: bound[x] += s * w over the source's is_e_ row
//   k_rs_exits   exits (OnlyOut, BothOutIn, OutMaster, BothOutInMaster): bound ->
//                rank, d*bound/outdeg onto the ib_e_ row; a master also feeds each
//                in-mirror m on its sync_e_ row: spnode[m] += bound, one shortcut
//                item (m, bound)
//   k_rs_rows    the in-mirrors' shortcut items
//   k_pr_push    the push items over the ib_e_ CSR (hardware float atomics)
//   k_rs_omir    out-mirror -> master: delta[master] += bound[out-mirror]
// The precompute that opens the phase (sum_sync_iter_worker.h:387-435,
// iter_compressor.h:534-560, 714-765) runs the members' pending delta to
// convergence inside the clusters over the subgraph rows (k_rs_pre_take takes the
// delta into rank and the member's pass-through sum `pre`, k_rs_sub_push sends
// d/outdeg of it along the row: members into delta, out-mirror ids into bound), then
// the exit pass sends each exit's pass-through sum out (rank untouched: the rounds
// took it already).
// The passes run in this order, so mass that reaches an entry crosses its cluster
// within a round, as the reference's parallel loop may let it; the in-mirrors'
// outflow and everything pushed wait for the next round.  The correction
// (sum_sync_iter_worker.h:1414-1469) first flushes the exits' pending inflow, then
// spreads every non-zero spnode through its supernode's inner_value (rank) and
// inner_delta (delta); the plain Ingress rounds follow (pr.hip).
#include <algorithm>
#include <cstring>
#include <vector>

#include "ctx.hpp"
#include "device.hpp"
#include "pr.hpp"
#include "push.hpp"

namespace suminc {

struct RSketch {
  uint32_t n = 0, all = 0;
  uint64_t version = ~0ull;  // graph version the sketch was uploaded for
  uint32_t n_single = 0, n_entry = 0, n_exit = 0, n_omir = 0;
  uint64_t n_is = 0, n_ib = 0, n_sync = 0, n_iv = 0, n_id = 0;
  DevBuf ntype, master;                          // n (u8); all - n (u32)
  DevBuf is_off, is_col, is_w;                   // all + 1, rows of bound_delta (f64 weights)
  DevBuf ib_off, ib_col;                         // n + 1
  DevBuf sync_off, sync_col;                     // all + 1
  DevBuf iv_off, iv_col, iv_w, id_off, id_col, id_w;  // all + 1, inner_value / inner_delta (f64)
  DevBuf sub_off, sub_col;                       // all + 1, the cluster subgraph (optional)
  DevBuf members, pre, moved;                    // clustered vertices; pass-through sums (n); moved mass + work
  bool has_sub = false;
  uint32_t n_member = 0;
  uint64_t n_sub = 0;
  DevBuf singles, entries, exits, omir;          // vertex lists by role
  DevBuf bound, spnode;                          // all, in the PageRank precision (sized for f64)
  DevBuf item_v, item_s, item_n;                 // shortcut items: (source, scalar f64), count
  DevBuf partial, out;                           // residual block sums
  uint64_t item_cap = 0;
  uint64_t accum_state = ~0ull;                  // PageRank state whose sketch accumulators bound and
                                                 // spnode hold (~0: to be cleared before the next round)
  // host copies kept for suminc_sketch_download
  std::vector<uint8_t> h_ntype;
  std::vector<uint32_t> h_master;
};

static RSketch& rstate(suminc_ctx* c) {
  if (!c->rsk) c->rsk = new RSketch();
  return *reinterpret_cast<RSketch*>(c->rsk);
}

void rsk_release(suminc_ctx* c) {
  delete reinterpret_cast<RSketch*>(c->rsk);
  c->rsk = nullptr;
}

template <class T>
__device__ __forceinline__ void rs_kahan(T* rank, float* comp, uint32_t v, T x) {
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

// Push item: row v of the ib_e_ CSR with contribution cval (light list or 2048-edge
// chunks of a long row), as the plain rounds' push reads them.
template <class T>
__device__ __forceinline__ void rs_push_item(bool want, uint32_t v, T cval, uint32_t rowlen, uint32_t* wl_v, T* wl_c,
                                             uint4* heavy, Counters* cnt) {
  const bool light = want && rowlen > 0 && rowlen <= kHeavyDeg;
  if (want && rowlen > kHeavyDeg) {
    const uint32_t nch = (rowlen + kChunk - 1) / kChunk;
    const unsigned long long slot = atomicAdd(&cnt->heavy_count, (unsigned long long)nch);
    for (uint32_t k = 0; k < nch; ++k) {
      uint4 h;
      h.x = v;
      h.y = k;
      if constexpr (sizeof(T) == 4) {
        h.z = __float_as_uint((float)cval);
        h.w = 0;
      } else {
        const unsigned long long bits = __double_as_longlong((double)cval);
        h.z = (uint32_t)bits;
        h.w = (uint32_t)(bits >> 32);
      }
      heavy[slot + k] = h;
    }
  }
  const unsigned long long slot = wave_append(light, &cnt->wl_count[0]);
  if (light) {
    wl_v[slot] = v;
    wl_c[slot] = cval;
  }
}

struct RsArgs {
  uint32_t n;
  const uint8_t* ntype;
  const uint32_t* off;      // full out-CSR row pointers (out-degrees)
  const uint32_t* ib_off;   // ib_e_ row pointers
  const uint32_t* sync_off;
  const uint32_t* sync_col;
  const uint32_t* master;   // mirror id - n -> master
  uint32_t* wl_v;
  uint4* heavy;
  Counters* cnt;
  uint32_t* item_v;
  double* item_s;
  unsigned long long* item_n;
};

__global__ void k_rs_begin(Counters* cnt, unsigned long long* item_n) {
  if (threadIdx.x) return;
  cnt->wl_count[0] = 0;
  cnt->wl_count[1] = 0;
  cnt->heavy_count = 0;
  cnt->mode = 0;  // the push kernel runs in push mode only
  cnt->done = 0;
  *item_n = 0;
}

__global__ void k_rs_items_clear(unsigned long long* item_n) {
  if (threadIdx.x == 0) *item_n = 0;
}

// FP64 sum |delta| over the vertices plus sum |bound| over every id, in fixed-grid
// block sums (deterministic order)
constexpr unsigned kRsGrid = 1024;
template <class T>
__global__ void __launch_bounds__(kBlock) k_rs_residual(uint32_t n, uint32_t all, const T* __restrict__ delta,
                                                        const T* __restrict__ bound, double* __restrict__ partial) {
  __shared__ double s_d[kBlock / 64], s_b[kBlock / 64];
  double a = 0, b = 0;
  for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < all; v += kRsGrid * kBlock) {
    if (v < n) a += fabs((double)delta[v]);
    b += fabs((double)bound[v]);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o);
    b += __shfl_xor(b, o);
  }
  if (lane_id() == 0) {
    s_d[threadIdx.x >> 6] = a;
    s_b[threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = 0, y = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
      x += s_d[w];
      y += s_b[w];
    }
    partial[blockIdx.x] = x;
    partial[kRsGrid + blockIdx.x] = y;
  }
}

__global__ void __launch_bounds__(1024) k_rs_residual_end(const double* __restrict__ partial, double* __restrict__ out) {
  __shared__ double s[1024], sb[1024];
  s[threadIdx.x] = partial[threadIdx.x];
  sb[threadIdx.x] = partial[kRsGrid + threadIdx.x];
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      s[threadIdx.x] += s[threadIdx.x + o];
      sb[threadIdx.x] += sb[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = s[0];
    out[1] = sb[0];
    out[2] = (s[0] != s[0] || sb[0] != sb[0] || s[0] > 1e300 || sb[0] > 1e300) ? 1.0 : 0.0;
  }
}

// Singles (list 0) and entries (list 1) of one round.
template <class T>
__global__ void __launch_bounds__(kBlock) k_rs_send(RsArgs a, const uint32_t* __restrict__ singles, uint32_t ns,
                                                    const uint32_t* __restrict__ entries, uint32_t ne,
                                                    T* __restrict__ delta, T* __restrict__ rank,
                                                    float* __restrict__ comp, T* __restrict__ spnode,
                                                    T* __restrict__ wl_c, double eps, T d) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  bool want = false;
  T cval = T(0);
  uint32_t v = 0, rowlen = 0;
  if (i < ns) {
    v = singles[i];
    const T dv = delta[v];
    if ((double)fabs_t(dv) > eps) {
      delta[v] = T(0);
      rs_kahan(rank, comp, v, dv);
      const uint32_t deg = a.off[v + 1] - a.off[v];
      rowlen = a.ib_off[v + 1] - a.ib_off[v];
      if (deg && rowlen) {
        want = true;
        cval = d * dv / (T)deg;
      }
    }
  }
  rs_push_item<T>(want, v, cval, rowlen, a.wl_v, wl_c, a.heavy, a.cnt);
  bool sc = false;
  double sval = 0;
  uint32_t u = 0;
  if (i < ne) {
    u = entries[i];
    const T dv = delta[u];
    if ((double)fabs_t(dv) > eps) {
      delta[u] = T(0);
      spnode[u] += dv;
      sc = true;
      sval = (double)dv;
    }
  }
  const unsigned long long slot = wave_append(sc, a.item_n);
  if (sc) {
    a.item_v[slot] = u;
    a.item_s[slot] = sval;
  }
}

// Exits of one round (flush: every non-zero inflow).
template <class T>
__global__ void __launch_bounds__(kBlock) k_rs_exits(RsArgs a, const uint32_t* __restrict__ exits, uint32_t nx,
                                                     T* __restrict__ bound, T* __restrict__ rank,
                                                     float* __restrict__ comp, T* __restrict__ spnode,
                                                     T* __restrict__ wl_c, double eps, T d, int flush) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  bool want = false;
  T cval = T(0), b = T(0);
  uint32_t v = 0, rowlen = 0;
  bool sent = false;
  if (i < nx) {
    v = exits[i];
    b = bound[v];
    if (flush ? b != T(0) : (double)fabs_t(b) > eps) {
      bound[v] = T(0);
      if (rank) rs_kahan(rank, comp, v, b);
      sent = true;
      const uint32_t deg = a.off[v + 1] - a.off[v];
      rowlen = a.ib_off[v + 1] - a.ib_off[v];
      if (deg && rowlen) {
        want = true;
        cval = d * b / (T)deg;
      }
    }
  }
  rs_push_item<T>(want, v, cval, rowlen, a.wl_v, wl_c, a.heavy, a.cnt);
  // a master feeds its in-mirrors (OutMaster / BothOutInMaster): the mass that
  // arrived at it also enters the clusters where it is mirrored
  const uint8_t t = i < nx ? a.ntype[v] : 0;
  const bool master = sent && (t == 5 || t == 6);
  const uint32_t sb = master ? a.sync_off[v] : 0u, se = master ? a.sync_off[v + 1] : 0u;
  const unsigned long long cnt = se - sb;
  // one reservation per thread (masters are few; their in-mirror lists are short)
  const unsigned long long at = cnt ? atomicAdd(a.item_n, cnt) : 0ull;
  for (uint32_t k = sb; k < se; ++k) {
    const uint32_t m = a.sync_col[k];
    spnode[m] += b;
    a.item_v[at + (k - sb)] = m;
    a.item_s[at + (k - sb)] = (double)b;
  }
}

// Shortcut items: target[col[e]] += s * w[e] over the item's row (block-
// cooperative walk of up to kBlock items' rows, coalesced column reads).
template <class T>
__global__ void __launch_bounds__(kBlock) k_rs_rows(const uint32_t* __restrict__ item_v,
                                                    const double* __restrict__ item_s,
                                                    const unsigned long long* __restrict__ item_n,
                                                    const uint32_t* __restrict__ roff, const uint32_t* __restrict__ rcol,
                                                    const double* __restrict__ rw, T* __restrict__ target) {
  __shared__ uint32_t s_scan[kBlock + 1];
  __shared__ uint32_t s_start[kBlock];
  __shared__ uint32_t s_warp[kBlock / 64 + 1];
  __shared__ double s_s[kBlock];
  const unsigned long long nl = *item_n;
  const uint32_t vt = frontier_tile<kBlock>(nl);
  const unsigned long long tiles = (nl + vt - 1) / vt;
  for (unsigned long long t = blockIdx.x; t < tiles; t += gridDim.x) {
    const unsigned long long i = t * vt + threadIdx.x;
    uint32_t start = 0, len = 0;
    if (threadIdx.x < vt && i < nl) {
      const uint32_t u = item_v[i];
      start = roff[u];
      len = roff[u + 1] - start;
      s_s[threadIdx.x] = item_s[i];
    }
    block_adjacency<kBlock>(start, len, s_scan, s_start, s_warp, [&](int slot, uint32_t, uint32_t e) {
      const T x = (T)(s_s[slot] * rw[e]);
      if (x != T(0)) unsafeAtomicAdd(&target[rcol[e]], x);
    });
  }
}

// out-mirror -> master (OutMirrorSyncToMaster; sum_sync_iter_worker.h:1114-1127)
template <class T>
__global__ void k_rs_omir(const uint32_t* __restrict__ omir, uint32_t no, uint32_t n,
                          const uint32_t* __restrict__ master, T* __restrict__ bound, T* __restrict__ delta) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= no) return;
  const uint32_t o = omir[i];
  const T x = bound[o];
  if (x != T(0)) {
    bound[o] = T(0);
    unsafeAtomicAdd(&delta[master[o - n]], x);
  }
}

// correction items: every id with a non-zero spnode, which is cleared
template <class T>
__global__ void k_rs_spnode_items(uint32_t all, T* __restrict__ spnode, uint32_t* __restrict__ item_v,
                                  double* __restrict__ item_s, unsigned long long* __restrict__ item_n) {
  const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
  const T s = u < all ? spnode[u] : T(0);
  const bool take = s != T(0);
  const unsigned long long slot = wave_append(take, item_n);
  if (take) {
    spnode[u] = T(0);
    item_v[slot] = u;
    item_s[slot] = (double)s;
  }
}

// precompute round, take: every member with pending delta takes it into rank and
// its pass-through sum, and becomes an item (v, d * x / outdeg) over its subgraph
// row; *moved += |x| (FP64), *work += the row lengths
template <class T>
__global__ void __launch_bounds__(kBlock) k_rs_pre_take(const uint32_t* __restrict__ members, uint32_t nm,
                                                        const uint32_t* __restrict__ off,
                                                        const uint32_t* __restrict__ sub_off, T* __restrict__ delta,
                                                        T* __restrict__ rank, float* __restrict__ comp,
                                                        T* __restrict__ pre, uint32_t* __restrict__ item_v,
                                                        double* __restrict__ item_s,
                                                        unsigned long long* __restrict__ item_n,
                                                        double* __restrict__ moved,
                                                        unsigned long long* __restrict__ work, T d) {
  __shared__ double s_m[kBlock / 64];
  __shared__ unsigned long long s_w[kBlock / 64];
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  bool take = false;
  uint32_t v = 0;
  double xm = 0, cval = 0;
  unsigned long long w = 0;
  if (i < nm) {
    v = members[i];
    const T x = delta[v];
    if (x != T(0)) {
      delta[v] = T(0);
      rs_kahan(rank, comp, v, x);
      pre[v] += x;
      xm = fabs((double)x);
      const uint32_t deg = off[v + 1] - off[v];
      const uint32_t len = sub_off[v + 1] - sub_off[v];
      if (deg && len) {
        take = true;
        cval = (double)(d * x / (T)deg);
        w = len;
      }
    }
  }
  const unsigned long long slot = wave_append(take, item_n);
  if (take) {
    item_v[slot] = v;
    item_s[slot] = cval;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    xm += __shfl_xor(xm, o);
    w += __shfl_xor(w, o);
  }
  if (lane_id() == 0) {
    s_m[threadIdx.x >> 6] = xm;
    s_w[threadIdx.x >> 6] = w;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0;
    unsigned long long b = 0;
    for (int k = 0; k < kBlock / 64; ++k) {
      a += s_m[k];
      b += s_w[k];
    }
    if (a != 0.0) unsafeAtomicAdd(moved, a);
    if (b) atomicAdd(work, b);
  }
}

// precompute round, push: item rows of the subgraph; a member target's delta, an
// out-mirror target's bound (ids >= n)
template <class T>
__global__ void __launch_bounds__(kBlock) k_rs_sub_push(const uint32_t* __restrict__ item_v,
                                                        const double* __restrict__ item_s,
                                                        const unsigned long long* __restrict__ item_n,
                                                        const uint32_t* __restrict__ roff,
                                                        const uint32_t* __restrict__ rcol, uint32_t n,
                                                        T* __restrict__ delta, T* __restrict__ bound) {
  __shared__ uint32_t s_scan[kBlock + 1];
  __shared__ uint32_t s_start[kBlock];
  __shared__ uint32_t s_warp[kBlock / 64 + 1];
  __shared__ T s_s[kBlock];
  const unsigned long long nl = *item_n;
  const uint32_t vt = frontier_tile<kBlock>(nl);
  const unsigned long long tiles = (nl + vt - 1) / vt;
  for (unsigned long long t = blockIdx.x; t < tiles; t += gridDim.x) {
    const unsigned long long i = t * vt + threadIdx.x;
    uint32_t start = 0, len = 0;
    if (threadIdx.x < vt && i < nl) {
      const uint32_t u = item_v[i];
      start = roff[u];
      len = roff[u + 1] - start;
      s_s[threadIdx.x] = (T)item_s[i];
    }
    block_adjacency<kBlock>(start, len, s_scan, s_start, s_warp, [&](int slot, uint32_t, uint32_t e) {
      const uint32_t x = rcol[e];
      unsafeAtomicAdd(x < n ? &delta[x] : &bound[x], s_s[slot]);
    });
  }
}

__global__ void k_rs_pre_begin(unsigned long long* item_n, double* moved) {
  if (threadIdx.x == 0) {
    *item_n = 0;
    *moved = 0.0;
  }
}

__global__ void k_rs_fill(uint8_t* p, uint64_t bytes) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < bytes; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// ---------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------
template <class V>
static void to_dev(suminc_ctx* c, DevBuf& b, const V* h, size_t count) {
  b.ensure(count * sizeof(V) + sizeof(V));
  if (count) SUMINC_HIP(hipMemcpyAsync(b.p, h, count * sizeof(V), hipMemcpyHostToDevice, c->stream));
}

// a CSR of the sketch: u64 row pointers at the boundary (checked), u32 on the device
static uint64_t check_rows(const uint64_t* off, uint32_t rows, uint32_t col_bound, const uint32_t* col,
                           const char* what) {
  SUMINC_REQUIRE(off, SUMINC_E_INVALID, std::string("null row pointers: ") + what);
  SUMINC_REQUIRE(off[0] == 0, SUMINC_E_INVALID, std::string(what) + ": row pointers must start at 0");
  for (uint32_t r = 0; r < rows; ++r)
    SUMINC_REQUIRE(off[r + 1] >= off[r], SUMINC_E_INVALID, std::string(what) + ": row pointers must not decrease");
  const uint64_t nnz = off[rows];
  SUMINC_REQUIRE(nnz < 0xFFFFFFFFull, SUMINC_E_TOO_LARGE, std::string(what) + ": more than 2^32 - 1 entries");
  SUMINC_REQUIRE(nnz == 0 || col, SUMINC_E_INVALID, std::string("null columns: ") + what);
  for (uint64_t e = 0; e < nnz; ++e)
    SUMINC_REQUIRE(col[e] < col_bound, SUMINC_E_INVALID, std::string(what) + ": column id out of range");
  return nnz;
}

static void rows_to_dev(suminc_ctx* c, DevBuf& doff, DevBuf& dcol, DevBuf* dw, const uint64_t* off, uint32_t rows,
                        const uint32_t* col, const float* w, const double* w64, uint64_t nnz) {
  std::vector<uint32_t> o32(rows + 1);
  for (uint32_t r = 0; r <= rows; ++r) o32[r] = (uint32_t)off[r];
  to_dev(c, doff, o32.data(), o32.size());
  to_dev(c, dcol, col, nnz);
  if (dw) {
    std::vector<double> wd(nnz);
    for (uint64_t e = 0; e < nnz; ++e) wd[e] = w64 ? w64[e] : w ? (double)w[e] : 0.0;
    to_dev(c, *dw, wd.data(), nnz);
  }
  SUMINC_HIP(hipStreamSynchronize(c->stream));  // the host vectors go out of scope
}

// is_w64 / iv_w64 / id_w64: FP64 row weights replacing the float ones (the
// library's own compressor); null: the sketch's float weights
void rsk_upload(suminc_ctx* c, const suminc_sketch_csr* s, const double* is_w64, const double* iv_w64,
                const double* id_w64) {
  SUMINC_REQUIRE(c->has_graph, SUMINC_E_STATE, "no graph");
  SUMINC_REQUIRE(s, SUMINC_E_INVALID, "null sketch");
  const uint32_t n = c->n;
  SUMINC_REQUIRE(s->n == n, SUMINC_E_INVALID, "sketch vertex count differs from the graph's");
  SUMINC_REQUIRE(s->all_node_num >= n, SUMINC_E_INVALID, "all_node_num below the vertex count");
  SUMINC_REQUIRE(s->node_type && (s->all_node_num == n || s->mirror_master), SUMINC_E_INVALID,
                 "null node types or mirror map");
  const uint32_t all = s->all_node_num;
  RSketch& R = rstate(c);
  // node types, roles
  std::vector<uint32_t> singles, entries, exits, omir;
  for (uint32_t v = 0; v < n; ++v) {
    const uint8_t t = s->node_type[v];
    SUMINC_REQUIRE(t <= 6, SUMINC_E_INVALID, "node type above BothOutInMaster (6)");
    if (t == 0) singles.push_back(v);
    if (t == 2 || t == 3 || t == 6) entries.push_back(v);
    if (t == 1 || t == 3 || t == 5 || t == 6) exits.push_back(v);
  }
  for (uint32_t m = n; m < all; ++m)
    SUMINC_REQUIRE(s->mirror_master[m - n] < n, SUMINC_E_INVALID, "mirror master out of range");
  const uint64_t n_is = check_rows(s->is_off, all, all, s->is_col, "is_e");
  SUMINC_REQUIRE(n_is == 0 || s->is_w || is_w64, SUMINC_E_INVALID, "null is_e weights");
  const uint64_t n_ib = check_rows(s->ib_off, n, n, s->ib_col, "ib_e");
  const uint64_t n_sync = check_rows(s->sync_off, all, all, s->sync_col, "sync_e");
  const uint64_t n_iv = check_rows(s->iv_off, all, n, s->iv_col, "inner_value");
  const uint64_t n_id = check_rows(s->id_off, all, n, s->id_col, "inner_delta");
  SUMINC_REQUIRE((n_iv == 0 || s->iv_w || iv_w64) && (n_id == 0 || s->id_w || id_w64), SUMINC_E_INVALID,
                 "null inner weights");
  const bool has_sub = s->sub_off != nullptr;
  const uint64_t n_sub = has_sub ? check_rows(s->sub_off, all, all, s->sub_col, "subgraph") : 0;
  std::vector<uint32_t> members;
  for (uint32_t v = 0; v < n; ++v) {
    if (s->node_type[v] != 0) members.push_back(v);
    else if (has_sub)
      SUMINC_REQUIRE(s->sub_off[v + 1] == s->sub_off[v], SUMINC_E_INVALID, "subgraph row on a Single");
  }
  // the sync rows: a vertex row lists in-mirrors (ids >= n) of a master, a mirror
  // row the master of an out-mirror
  for (uint32_t u = 0; u < all; ++u)
    for (uint64_t k = s->sync_off[u]; k < s->sync_off[u + 1]; ++k) {
      const uint32_t x = s->sync_col[k];
      if (u < n) {
        SUMINC_REQUIRE(x >= n, SUMINC_E_INVALID, "sync_e row of a vertex must list mirrors");
        SUMINC_REQUIRE(s->node_type[u] == 5 || s->node_type[u] == 6, SUMINC_E_INVALID,
                       "sync_e row on a vertex that is not a master (OutMaster / BothOutInMaster)");
      } else {
        SUMINC_REQUIRE(x == s->mirror_master[u - n], SUMINC_E_INVALID, "out-mirror sync must point to its master");
      }
    }
  for (uint32_t m = n; m < all; ++m)
    if (s->sync_off[m + 1] > s->sync_off[m]) omir.push_back(m);
  // ib_e_ rows are subsets of the out-rows: never longer than the out-degree
  std::vector<uint32_t> hoff(n + 1);
  SUMINC_HIP(hipMemcpy(hoff.data(), c->out_off.p, (size_t)(n + 1) * 4, hipMemcpyDeviceToHost));
  for (uint32_t v = 0; v < n; ++v)
    SUMINC_REQUIRE(s->ib_off[v + 1] - s->ib_off[v] <= hoff[v + 1] - hoff[v], SUMINC_E_INVALID,
                   "ib_e row longer than the vertex's out-degree");
  R.n = n;
  R.all = all;
  R.n_single = (uint32_t)singles.size();
  R.n_entry = (uint32_t)entries.size();
  R.n_exit = (uint32_t)exits.size();
  R.n_omir = (uint32_t)omir.size();
  R.n_is = n_is; R.n_ib = n_ib; R.n_sync = n_sync; R.n_iv = n_iv; R.n_id = n_id;
  to_dev(c, R.ntype, s->node_type, n);
  to_dev(c, R.master, s->mirror_master, all - n);
  to_dev(c, R.singles, singles.data(), singles.size());
  to_dev(c, R.entries, entries.data(), entries.size());
  to_dev(c, R.exits, exits.data(), exits.size());
  to_dev(c, R.omir, omir.data(), omir.size());
  rows_to_dev(c, R.is_off, R.is_col, &R.is_w, s->is_off, all, s->is_col, s->is_w, is_w64, n_is);
  rows_to_dev(c, R.ib_off, R.ib_col, nullptr, s->ib_off, n, s->ib_col, nullptr, nullptr, n_ib);
  rows_to_dev(c, R.sync_off, R.sync_col, nullptr, s->sync_off, all, s->sync_col, nullptr, nullptr, n_sync);
  rows_to_dev(c, R.iv_off, R.iv_col, &R.iv_w, s->iv_off, all, s->iv_col, s->iv_w, iv_w64, n_iv);
  rows_to_dev(c, R.id_off, R.id_col, &R.id_w, s->id_off, all, s->id_col, s->id_w, id_w64, n_id);
  R.has_sub = has_sub;
  R.n_sub = n_sub;
  R.n_member = (uint32_t)members.size();
  if (has_sub) rows_to_dev(c, R.sub_off, R.sub_col, nullptr, s->sub_off, all, s->sub_col, nullptr, nullptr, n_sub);
  to_dev(c, R.members, members.data(), members.size());
  R.pre.ensure((size_t)n * 8 + 8);
  R.moved.ensure(16);
  R.item_cap = std::max<uint64_t>({(uint64_t)all, n_sync + R.n_entry}) + 1;
  R.item_v.ensure(R.item_cap * 4);
  R.item_s.ensure(R.item_cap * 8);
  R.item_n.ensure(8);
  R.bound.ensure((size_t)all * 8 + 8);
  R.spnode.ensure((size_t)all * 8 + 8);
  R.partial.ensure(2 * kRsGrid * 8);
  R.out.ensure(4 * 8);
  R.accum_state = ~0ull;
  R.version = c->graph_version;
  R.h_ntype.assign(s->node_type, s->node_type + n);
  R.h_master.assign(s->mirror_master ? s->mirror_master : nullptr,
                    s->mirror_master ? s->mirror_master + (all - n) : nullptr);
  SUMINC_HIP(hipStreamSynchronize(c->stream));
}

static RSketch& rsk_ready(suminc_ctx* c) {
  SUMINC_REQUIRE(c->rsk, SUMINC_E_STATE, "no sketch uploaded (suminc_upload_sketch)");
  RSketch& R = *reinterpret_cast<RSketch*>(c->rsk);
  SUMINC_REQUIRE(R.version == c->graph_version, SUMINC_E_STATE,
                 "the graph changed since the sketch was uploaded: upload the compressor's new products");
  SUMINC_REQUIRE(c->pr_ready, SUMINC_E_STATE, "PageRank state not initialised");
  return R;
}

template <class T>
static void rs_clear_accum(suminc_ctx* c, RSketch& R) {
  if (R.accum_state == c->pr_state_id) return;
  const uint64_t bytes = (uint64_t)R.all * sizeof(T);
  k_rs_fill<<<grid_for(bytes, kBlock, 4096), kBlock, 0, c->stream>>>(R.bound.as<uint8_t>(), bytes);
  k_rs_fill<<<grid_for(bytes, kBlock, 4096), kBlock, 0, c->stream>>>(R.spnode.as<uint8_t>(), bytes);
  R.accum_state = c->pr_state_id;
}

static RsArgs rs_args(suminc_ctx* c, RSketch& R) {
  RsArgs a;
  a.n = R.n;
  a.ntype = R.ntype.as<uint8_t>();
  a.off = c->out_off.as<uint32_t>();
  a.ib_off = R.ib_off.as<uint32_t>();
  a.sync_off = R.sync_off.as<uint32_t>();
  a.sync_col = R.sync_col.as<uint32_t>();
  a.master = R.master.as<uint32_t>();
  a.wl_v = c->wl0.as<uint32_t>();
  a.heavy = c->heavy.as<uint4>();
  a.cnt = dev_counters(c);
  a.item_v = R.item_v.as<uint32_t>();
  a.item_s = R.item_s.as<double>();
  a.item_n = R.item_n.as<unsigned long long>();
  return a;
}

// FP64 sum |delta| (vertices) and sum |bound| (every id); sets the numeric flag.
template <class T>
static void rs_residual(suminc_ctx* c, RSketch& R, double* rd, double* rb) {
  k_rs_residual<T><<<kRsGrid, kBlock, 0, c->stream>>>(R.n, R.all, c->pr_delta.as<T>(), R.bound.as<T>(),
                                                      R.partial.as<double>());
  k_rs_residual_end<<<1, 1024, 0, c->stream>>>(R.partial.as<double>(), R.out.as<double>());
  SUMINC_HIP(hipGetLastError());
  check_begin(c, R.out.p, 3 * sizeof(double));
  const double* h = reinterpret_cast<const double*>(check_wait(c));
  SUMINC_REQUIRE(h[2] == 0.0, SUMINC_E_NUMERIC, "NaN/Inf in the sketch state");
  if (rd) *rd = h[0];
  if (rb) *rb = h[1];
}

static const uint32_t kRowGrid = 2048;

// the exit pass, the in-mirrors' shortcut items, the push and the out-mirror sync.
// src: the exits' pending sums (bound; the precompute's pass-through sums);
// rank = null: those sums are in rank already
template <class T>
static void rs_exit_part(suminc_ctx* c, RSketch& R, const RsArgs& a, double eps, T d, int flush, T* src, T* rank) {
  k_rs_items_clear<<<1, 64, 0, c->stream>>>(a.item_n);
  if (R.n_exit)
    k_rs_exits<T><<<grid_for(R.n_exit), kBlock, 0, c->stream>>>(a, R.exits.as<uint32_t>(), R.n_exit, src, rank,
                                                               c->pr_comp.as<float>(), R.spnode.as<T>(),
                                                               c->wl1.as<T>(), eps, d, flush);
  k_rs_rows<T><<<kRowGrid, kBlock, 0, c->stream>>>(a.item_v, a.item_s, a.item_n, R.is_off.as<uint32_t>(),
                                                   R.is_col.as<uint32_t>(), R.is_w.as<double>(), R.bound.as<T>());
  k_pr_push<T><<<push_grid(), kBlock, 0, c->stream>>>(R.ib_off.as<uint32_t>(), R.ib_col.as<uint32_t>(),
                                                      c->pr_delta.as<T>(), a.wl_v, c->wl1.as<T>(), a.heavy, a.cnt);
  if (R.n_omir)
    k_rs_omir<T><<<grid_for(R.n_omir), kBlock, 0, c->stream>>>(R.omir.as<uint32_t>(), R.n_omir, R.n,
                                                              R.master.as<uint32_t>(), R.bound.as<T>(),
                                                              c->pr_delta.as<T>());
  SUMINC_HIP(hipGetLastError());
}

template <class T>
static void rs_round(suminc_ctx* c, RSketch& R, double eps) {
  const T d = (T)c->prm.damping;
  const RsArgs a = rs_args(c, R);
  k_rs_begin<<<1, 64, 0, c->stream>>>(a.cnt, a.item_n);
  const uint32_t ns = std::max(R.n_single, R.n_entry);
  if (ns)
    k_rs_send<T><<<grid_for(ns), kBlock, 0, c->stream>>>(a, R.singles.as<uint32_t>(), R.n_single,
                                                        R.entries.as<uint32_t>(), R.n_entry, c->pr_delta.as<T>(),
                                                        c->pr_rank.as<T>(), c->pr_comp.as<float>(), R.spnode.as<T>(),
                                                        c->wl1.as<T>(), eps, d);
  k_rs_rows<T><<<kRowGrid, kBlock, 0, c->stream>>>(a.item_v, a.item_s, a.item_n, R.is_off.as<uint32_t>(),
                                                   R.is_col.as<uint32_t>(), R.is_w.as<double>(), R.bound.as<T>());
  rs_exit_part<T>(c, R, a, eps, d, 0, R.bound.as<T>(), c->pr_rank.as<T>());
}

template <class T>
static void rsk_step_t(suminc_ctx* c, double* residual) {
  RSketch& R = rsk_ready(c);
  rs_clear_accum<T>(c, R);
  double rd = 0, rb = 0;
  rs_residual<T>(c, R, &rd, &rb);
  if (residual) *residual = rd + rb;
  rs_round<T>(c, R, c->prm.termcheck_threshold / (double)c->n);
  SUMINC_HIP(hipStreamSynchronize(c->stream));
}

// flush of the exits' pending inflow, then the correction.  Each flush pass moves
// at most d of the pending inflow on (through the masters' in-mirrors); the passes
// stop once it is below T / 1000 (at most 256 passes) and the rest, three orders
// below the stop tolerance, is dropped, as the reference drops all of it (its
// correction spreads only spnode, sum_sync_iter_worker.h:1411-1431)
template <class T>
static void rsk_correct_t(suminc_ctx* c) {
  RSketch& R = rsk_ready(c);
  rs_clear_accum<T>(c, R);
  const T d = (T)c->prm.damping;
  const RsArgs a = rs_args(c, R);
  for (int pass = 0; pass < 256; ++pass) {
    double rb = 0;
    rs_residual<T>(c, R, nullptr, &rb);
    if (rb <= 1e-3 * c->prm.termcheck_threshold) break;
    k_rs_begin<<<1, 64, 0, c->stream>>>(a.cnt, a.item_n);
    rs_exit_part<T>(c, R, a, 0.0, d, 1, R.bound.as<T>(), c->pr_rank.as<T>());
  }
  k_rs_items_clear<<<1, 64, 0, c->stream>>>(a.item_n);
  k_rs_spnode_items<T><<<grid_for(R.all), kBlock, 0, c->stream>>>(R.all, R.spnode.as<T>(), a.item_v, a.item_s,
                                                                  a.item_n);
  k_rs_rows<T><<<kRowGrid, kBlock, 0, c->stream>>>(a.item_v, a.item_s, a.item_n, R.iv_off.as<uint32_t>(),
                                                   R.iv_col.as<uint32_t>(), R.iv_w.as<double>(), c->pr_rank.as<T>());
  k_rs_rows<T><<<kRowGrid, kBlock, 0, c->stream>>>(a.item_v, a.item_s, a.item_n, R.id_off.as<uint32_t>(),
                                                   R.id_col.as<uint32_t>(), R.id_w.as<double>(), c->pr_delta.as<T>());
  SUMINC_HIP(hipGetLastError());
  SUMINC_HIP(hipStreamSynchronize(c->stream));
  R.accum_state = ~0ull;  // drops the flush's remainder
  rs_clear_accum<T>(c, R);
}

// The precompute: Jacobi rounds of every cluster at once over the subgraph rows,
// until the mass moved in a round is at most T / n per member (the reference's
// per-cluster rule summed over the clusters) or shortcut_max_rounds, then the
// exits' pass-through sums go out.  Returns the subgraph edges read.
template <class T>
static uint64_t rsk_precompute_t(suminc_ctx* c) {
  RSketch& R = rsk_ready(c);
  SUMINC_REQUIRE(R.has_sub, SUMINC_E_STATE, "the uploaded sketch has no subgraph rows (sub_off / sub_col)");
  rs_clear_accum<T>(c, R);
  const T d = (T)c->prm.damping;
  const RsArgs a = rs_args(c, R);
  const uint64_t pbytes = (uint64_t)R.n * sizeof(T);
  k_rs_fill<<<grid_for(pbytes, kBlock, 4096), kBlock, 0, c->stream>>>(R.pre.as<uint8_t>(), pbytes);
  k_rs_fill<<<1, 64, 0, c->stream>>>(R.moved.as<uint8_t>(), 16);
  double* moved = R.moved.as<double>();
  unsigned long long* work = reinterpret_cast<unsigned long long*>(R.moved.as<uint8_t>() + 8);
  const double thr = c->prm.termcheck_threshold / (double)c->n * (double)R.n_member;
  const int maxit = c->prm.shortcut_max_rounds > 0 ? c->prm.shortcut_max_rounds : 100;
  for (int round = 0; R.n_member && round < maxit; ++round) {
    k_rs_pre_begin<<<1, 64, 0, c->stream>>>(a.item_n, moved);
    k_rs_pre_take<T><<<grid_for(R.n_member), kBlock, 0, c->stream>>>(
        R.members.as<uint32_t>(), R.n_member, c->out_off.as<uint32_t>(), R.sub_off.as<uint32_t>(),
        c->pr_delta.as<T>(), c->pr_rank.as<T>(), c->pr_comp.as<float>(), R.pre.as<T>(), a.item_v, a.item_s,
        a.item_n, moved, work, d);
    k_rs_sub_push<T><<<kRowGrid, kBlock, 0, c->stream>>>(a.item_v, a.item_s, a.item_n, R.sub_off.as<uint32_t>(),
                                                         R.sub_col.as<uint32_t>(), R.n, c->pr_delta.as<T>(),
                                                         R.bound.as<T>());
    SUMINC_HIP(hipGetLastError());
    check_begin(c, moved, sizeof(double));
    const double mv = *reinterpret_cast<const double*>(check_wait(c));
    SUMINC_REQUIRE(mv == mv && mv < 1e300, SUMINC_E_NUMERIC, "NaN/Inf in the precompute");
    if (mv <= thr) break;
  }
  k_rs_begin<<<1, 64, 0, c->stream>>>(a.cnt, a.item_n);
  rs_exit_part<T>(c, R, a, 0.0, d, 1, R.pre.as<T>(), nullptr);
  unsigned long long w = 0;
  SUMINC_HIP(hipMemcpyAsync(&w, work, 8, hipMemcpyDeviceToHost, c->stream));
  SUMINC_HIP(hipStreamSynchronize(c->stream));
  return w;
}

template <class T>
static void rsk_run_t(suminc_ctx* c, suminc_stats* st) {
  RSketch& R = rsk_ready(c);
  rs_clear_accum<T>(c, R);
  const double T_stop = c->prm.termcheck_threshold;
  const double eps = T_stop / (double)c->n;
  const int max_rounds = c->prm.max_rounds > 0 ? c->prm.max_rounds : 1000000;
  SUMINC_HIP(hipEventRecord(c->ev_start, c->stream));
  const uint64_t pre_edges = R.has_sub ? rsk_precompute_t<T>(c) : 0;
  int rounds = 0;
  while (true) {
    double rd = 0, rb = 0;
    rs_residual<T>(c, R, &rd, &rb);
    if (!(rd + rb >= T_stop) || rounds >= max_rounds) break;
    rs_round<T>(c, R, eps);
    ++rounds;
  }
  rsk_correct_t<T>(c);
  SUMINC_HIP(hipEventRecord(c->ev_stop, c->stream));
  SUMINC_HIP(hipEventSynchronize(c->ev_stop));
  float ms_sketch = 0;
  SUMINC_HIP(hipEventElapsedTime(&ms_sketch, c->ev_start, c->ev_stop));
  suminc_stats rs{};
  pr_run(c, &rs);  // the plain residual phase
  suminc_stats out = rs;
  out.rounds = rs.rounds + rounds;
  out.kernel_ms = rs.kernel_ms + ms_sketch;
  out.sketch_rounds = rounds;
  out.sketch_ms = ms_sketch;
  out.residual_rounds = rs.rounds;
  out.precompute_edges = (double)pre_edges;
  c->stats = out;
  if (st) *st = out;
}

void rsk_step(suminc_ctx* c, double* residual) {
  if (c->pr_prec == SUMINC_PR_FP64) rsk_step_t<double>(c, residual);
  else rsk_step_t<float>(c, residual);
}

void rsk_precompute(suminc_ctx* c) {
  if (c->pr_prec == SUMINC_PR_FP64) rsk_precompute_t<double>(c);
  else rsk_precompute_t<float>(c);
}

void rsk_correct(suminc_ctx* c) {
  if (c->pr_prec == SUMINC_PR_FP64) rsk_correct_t<double>(c);
  else rsk_correct_t<float>(c);
}

void rsk_run(suminc_ctx* c, suminc_stats* st) {
  if (c->pr_prec == SUMINC_PR_FP64) rsk_run_t<double>(c, st);
  else rsk_run_t<float>(c, st);
}

}  // namespace suminc
