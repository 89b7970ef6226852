// graph.hip -- CSR/CSC fragment layout in HBM: device R-MAT generation, host
// upload, CSC construction and download.
//
// Layout (one fragment, dense ids 0..n-1):
//   out_off[n+1] u32, out_col[m] u32, out_w[m] (u32 or f32, optional)
//   in_off[n+1]  u32, in_col[m]  u32, in_w[m]  (same kind, optional)
// Both sorted by (row, col).  This replaces the reference's pointer-array CSR
// (ImmutableEdgecutFragment::Init, grape/fragment/immutable_edgecut_fragment.h:
// 129-330) and the per-worker flattening into cur_oeoff/size_oe arrays
// (sum_sync_iter_worker.h:575-597): one u32 row pointer array instead of a
// (start, size) pair per vertex, so a degree costs one 8-byte read.
#include <cstring>  // rocprim's texture_cache_iterator needs memset declared
#include <rocprim/rocprim.hpp>

#include <vector>

#include "ctx.hpp"
#include "device.hpp"

namespace suminc {

// ---- R-MAT edge draw: the synthetic-input definition shared with the oracle ----
__global__ void k_rmat_keys(uint64_t raw, int scale, uint64_t t1, uint64_t t2, uint64_t t3,
                            uint32_t k0, uint32_t k1, uint64_t* keys) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= raw) return;
  uint32_t u = 0, v = 0;
  U4 o{0, 0, 0, 0};
  for (int l = 0; l < scale; ++l) {
    if ((l & 3) == 0) o = philox4x32_10((uint32_t)i, (uint32_t)(i >> 32), (uint32_t)(l >> 2), kStreamRmat, k0, k1);
    uint32_t w = (l & 3) == 0 ? o.x : (l & 3) == 1 ? o.y : (l & 3) == 2 ? o.z : o.w;
    uint64_t r = w;
    uint32_t bu, bv;
    if (r < t1) { bu = 0; bv = 0; }
    else if (r < t2) { bu = 0; bv = 1; }
    else if (r < t3) { bu = 1; bv = 0; }
    else { bu = 1; bv = 1; }
    u |= bu << (scale - 1 - l);
    v |= bv << (scale - 1 - l);
  }
  // compact key u<<scale | v; self-loops become the all-ones sentinel (sorts last)
  keys[i] = (u == v) ? ~0ull : (((uint64_t)u << scale) | v);
}

// Sorted unique keys -> columns (+ optional hashed weights).  Key layout is given
// by `shift` (scale for compact keys, 32 for general keys).
__global__ void k_keys_to_cols(const uint64_t* keys, uint64_t m, int shift, uint32_t* col, int wkind,
                               uint32_t wmin, uint32_t wmax, uint32_t k0, uint32_t k1, void* w) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  uint64_t key = keys[i];
  uint32_t u = (uint32_t)(key >> shift);
  uint32_t v = (uint32_t)(key & ((1ull << shift) - 1ull));
  col[i] = v;
  if (wkind) {
    U4 o = philox4x32_10(u, v, 0u, kStreamWeight, k0, k1);
    if (wkind == SUMINC_W_U32) ((uint32_t*)w)[i] = wmin + mulhi32(o.x, wmax - wmin + 1u);
    else ((float*)w)[i] = (float)((o.x >> 8) + 1u) * (1.0f / 16777216.0f);
  }
}

// Row pointers from sorted row ids: thread i (0..m) fills off[prev_row+1 .. row]
// with i, where thread m uses row = n.  Coalesced, O(m + n).
__global__ void k_row_pointers(const uint64_t* keys, uint64_t m, int shift, uint32_t n, uint32_t* off) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > m) return;
  int64_t row = (i == m) ? (int64_t)n : (int64_t)(keys[i] >> shift);
  int64_t prev = (i == 0) ? -1 : (int64_t)(keys[i - 1] >> shift);
  for (int64_t r = prev + 1; r <= row; ++r) off[r] = (uint32_t)i;
}

// Row id of every edge from a u32 row-pointer array (binary search).
__global__ void k_edge_rows(const uint32_t* off, uint32_t n, uint64_t m, uint32_t* rows) {
  uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m) return;
  uint32_t lo = 0, hi = n - 1;
  while (lo < hi) {
    uint32_t mid = lo + ((hi - lo + 1) >> 1);
    if (off[mid] <= e) lo = mid; else hi = mid - 1;
  }
  rows[e] = lo;
}

__global__ void k_transpose_keys(const uint32_t* rows, const uint32_t* col, uint64_t m, uint64_t* keys) {
  uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m) return;
  keys[e] = ((uint64_t)col[e] << 32) | rows[e];
}

__global__ void k_u64_to_u32_off(const uint64_t* in, uint32_t* out, uint32_t count) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) out[i] = (uint32_t)in[i];
}
__global__ void k_u32_to_u64_off(const uint32_t* in, uint64_t* out, uint32_t count) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) out[i] = in[i];
}

__global__ void k_low32(const uint64_t* keys, uint64_t m, uint32_t* out) {
  uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < m) out[e] = (uint32_t)keys[e];
}

// ---------------------------------------------------------------------------
// CSC (in-adjacency) from the out-CSR: sort (dst<<32 | src) with the weight as
// payload, then row pointers on dst.  Both arrays end up sorted by (dst, src).
// ---------------------------------------------------------------------------
void build_csc(suminc_ctx* c) {
  const uint64_t m = c->m;
  const uint32_t n = c->n;
  c->in_off.ensure(((size_t)n + 1) * 4);
  c->in_col.ensure(m * 4 + 4);
  if (c->wkind) c->in_w.ensure(m * 4 + 4);
  if (m == 0) {
    SUMINC_HIP(hipMemsetAsync(c->in_off.p, 0, ((size_t)n + 1) * 4, c->stream));
    return;
  }
  DevBuf rows, keys_a, keys_b;
  rows.ensure(m * 4);
  keys_a.ensure(m * 8);
  keys_b.ensure(m * 8);
  k_edge_rows<<<grid_for(m), kBlock, 0, c->stream>>>(c->out_off.as<uint32_t>(), n, m, rows.as<uint32_t>());
  k_transpose_keys<<<grid_for(m), kBlock, 0, c->stream>>>(rows.as<uint32_t>(), c->out_col.as<uint32_t>(), m,
                                                          keys_a.as<uint64_t>());
  int hi_bits = 1;
  while (hi_bits < 32 && (1ull << hi_bits) < n) ++hi_bits;
  const unsigned end_bit = 32 + hi_bits;
  size_t tmp_bytes = 0;
  if (c->wkind) {
    SUMINC_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys_a.as<uint64_t>(), keys_b.as<uint64_t>(),
                                         c->out_w.as<uint32_t>(), c->in_w.as<uint32_t>(), (size_t)m, 0, end_bit,
                                         c->stream));
    c->tmp.ensure(tmp_bytes);
    SUMINC_HIP(rocprim::radix_sort_pairs(c->tmp.p, tmp_bytes, keys_a.as<uint64_t>(), keys_b.as<uint64_t>(),
                                         c->out_w.as<uint32_t>(), c->in_w.as<uint32_t>(), (size_t)m, 0, end_bit,
                                         c->stream));
  } else {
    SUMINC_HIP(rocprim::radix_sort_keys(nullptr, tmp_bytes, keys_a.as<uint64_t>(), keys_b.as<uint64_t>(), (size_t)m,
                                        0, end_bit, c->stream));
    c->tmp.ensure(tmp_bytes);
    SUMINC_HIP(rocprim::radix_sort_keys(c->tmp.p, tmp_bytes, keys_a.as<uint64_t>(), keys_b.as<uint64_t>(),
                                        (size_t)m, 0, end_bit, c->stream));
  }
  k_low32<<<grid_for(m), kBlock, 0, c->stream>>>(keys_b.as<uint64_t>(), m, c->in_col.as<uint32_t>());
  k_row_pointers<<<grid_for(m + 1), kBlock, 0, c->stream>>>(keys_b.as<uint64_t>(), m, 32, n,
                                                            c->in_off.as<uint32_t>());
  SUMINC_HIP(hipGetLastError());
  SUMINC_HIP(hipStreamSynchronize(c->stream));  // keys/rows are freed on return
}

void ensure_state_sizes(suminc_ctx* c) {
  // invalidate per-graph state that depends on n
  c->pr_ready = false;
  c->sssp_ready = false;
}

void generate_rmat(suminc_ctx* c, int scale, int edge_factor, double a, double b, double cc, uint64_t seed,
                   int wkind, uint32_t wmin, uint32_t wmax) {
  SUMINC_REQUIRE(scale >= 1 && scale <= 30, SUMINC_E_INVALID, "scale must be in [1,30]");
  SUMINC_REQUIRE(edge_factor >= 1 && edge_factor <= 1024, SUMINC_E_INVALID, "edge_factor out of range");
  SUMINC_REQUIRE(a >= 0 && b >= 0 && cc >= 0 && a + b + cc <= 1.0, SUMINC_E_INVALID, "bad R-MAT probabilities");
  SUMINC_REQUIRE(wkind >= 0 && wkind <= 2, SUMINC_E_INVALID, "bad weight kind");
  SUMINC_REQUIRE(wkind != SUMINC_W_U32 || (wmin >= 1 && wmax >= wmin), SUMINC_E_INVALID, "bad weight range");
  const uint32_t n = 1u << scale;
  const uint64_t raw = (uint64_t)edge_factor << scale;
  SUMINC_REQUIRE(raw < (1ull << 32), SUMINC_E_TOO_LARGE, "edge count must stay below 2^32");
  const uint64_t t1 = (uint64_t)(a * 4294967296.0);
  const uint64_t t2 = (uint64_t)((a + b) * 4294967296.0);
  const uint64_t t3 = (uint64_t)((a + b + cc) * 4294967296.0);
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);

  DevBuf keys_a, keys_b, count;
  keys_a.ensure(raw * 8);
  keys_b.ensure(raw * 8);
  count.ensure(8);
  k_rmat_keys<<<grid_for(raw), kBlock, 0, c->stream>>>(raw, scale, t1, t2, t3, k0, k1, keys_a.as<uint64_t>());
  SUMINC_HIP(hipGetLastError());
  const unsigned end_bit = 2 * scale + 1;  // the sentinel has every bit set
  size_t tmp_bytes = 0;
  SUMINC_HIP(rocprim::radix_sort_keys(nullptr, tmp_bytes, keys_a.as<uint64_t>(), keys_b.as<uint64_t>(), (size_t)raw,
                                      0, end_bit, c->stream));
  c->tmp.ensure(tmp_bytes);
  SUMINC_HIP(rocprim::radix_sort_keys(c->tmp.p, tmp_bytes, keys_a.as<uint64_t>(), keys_b.as<uint64_t>(), (size_t)raw,
                                      0, end_bit, c->stream));
  tmp_bytes = 0;
  SUMINC_HIP(rocprim::unique(nullptr, tmp_bytes, keys_b.as<uint64_t>(), keys_a.as<uint64_t>(),
                             count.as<unsigned long long>(), (size_t)raw, rocprim::equal_to<uint64_t>(), c->stream));
  c->tmp.ensure(tmp_bytes);
  SUMINC_HIP(rocprim::unique(c->tmp.p, tmp_bytes, keys_b.as<uint64_t>(), keys_a.as<uint64_t>(),
                             count.as<unsigned long long>(), (size_t)raw, rocprim::equal_to<uint64_t>(), c->stream));
  unsigned long long uniq = 0, last = 0;
  SUMINC_HIP(hipMemcpyAsync(&uniq, count.p, 8, hipMemcpyDeviceToHost, c->stream));
  SUMINC_HIP(hipStreamSynchronize(c->stream));
  if (uniq) {
    SUMINC_HIP(hipMemcpyAsync(&last, keys_a.as<uint64_t>() + uniq - 1, 8, hipMemcpyDeviceToHost, c->stream));
    SUMINC_HIP(hipStreamSynchronize(c->stream));
  }
  const uint64_t m = (uniq && last == ~0ull) ? uniq - 1 : uniq;

  ensure_state_sizes(c);
  c->n = n;
  c->m = m;
  c->wkind = wkind;
  c->out_off.ensure(((size_t)n + 1) * 4);
  c->out_col.ensure(m * 4 + 4);
  if (wkind) c->out_w.ensure(m * 4 + 4);
  k_keys_to_cols<<<grid_for(m), kBlock, 0, c->stream>>>(keys_a.as<uint64_t>(), m, scale, c->out_col.as<uint32_t>(),
                                                        wkind, wmin, wmax, k0, k1, c->out_w.p);
  k_row_pointers<<<grid_for(m + 1), kBlock, 0, c->stream>>>(keys_a.as<uint64_t>(), m, scale, n,
                                                            c->out_off.as<uint32_t>());
  SUMINC_HIP(hipGetLastError());
  SUMINC_HIP(hipStreamSynchronize(c->stream));
  build_csc(c);
  c->has_graph = true;
  c->pos_live = false;  // a new graph: positions are renumbered
  c->graph_version++;
}

void upload_csr(suminc_ctx* c, uint32_t n, uint64_t m, const uint64_t* off, const uint32_t* col, int wkind,
                const void* w) {
  SUMINC_REQUIRE(n >= 1 && off && (m == 0 || col), SUMINC_E_INVALID, "null CSR array");
  SUMINC_REQUIRE(wkind >= 0 && wkind <= 2 && (wkind == 0 || w || m == 0), SUMINC_E_INVALID, "bad weights");
  SUMINC_REQUIRE(m < 0xFFFFFFFFull, SUMINC_E_TOO_LARGE, "edge count must stay below 2^32");
  SUMINC_REQUIRE(off[0] == 0 && off[n] == m, SUMINC_E_INVALID, "offsets must start at 0 and end at m");
  for (uint32_t i = 0; i < n; ++i)
    SUMINC_REQUIRE(off[i] <= off[i + 1], SUMINC_E_INVALID, "offsets must be non-decreasing");
  for (uint64_t e = 0; e < m; ++e) SUMINC_REQUIRE(col[e] < n, SUMINC_E_INVALID, "column id out of range");
  // integer weights start at 1: a zero-weight cycle would let the smallest-id
  // parent rule of the SSSP state close a parent cycle away from the source
  if (wkind == SUMINC_W_U32)
    for (uint64_t e = 0; e < m; ++e)
      SUMINC_REQUIRE(((const uint32_t*)w)[e] >= 1, SUMINC_E_INVALID, "integer weights must be at least 1");
  ensure_state_sizes(c);
  c->n = n;
  c->m = m;
  c->wkind = wkind;
  c->out_off.ensure(((size_t)n + 1) * 4);
  c->out_col.ensure(m * 4 + 4);
  if (wkind) c->out_w.ensure(m * 4 + 4);
  std::vector<uint32_t> off32(n + 1);
  for (uint32_t i = 0; i <= n; ++i) off32[i] = (uint32_t)off[i];
  SUMINC_HIP(hipMemcpyAsync(c->out_off.p, off32.data(), ((size_t)n + 1) * 4, hipMemcpyHostToDevice, c->stream));
  if (m) SUMINC_HIP(hipMemcpyAsync(c->out_col.p, col, m * 4, hipMemcpyHostToDevice, c->stream));
  if (wkind && m) SUMINC_HIP(hipMemcpyAsync(c->out_w.p, w, m * 4, hipMemcpyHostToDevice, c->stream));
  SUMINC_HIP(hipStreamSynchronize(c->stream));
  // rows must be sorted by column for the merge/search kernels; check on host
  for (uint32_t u = 0; u < n; ++u)
    for (uint64_t e = off[u] + 1; e < off[u + 1]; ++e)
      SUMINC_REQUIRE(col[e - 1] < col[e], SUMINC_E_INVALID, "adjacency must be sorted and free of duplicates");
  build_csc(c);
  c->has_graph = true;
  c->pos_live = false;  // a new graph: positions are renumbered
  c->graph_version++;
}

void download_csr(suminc_ctx* c, uint64_t* off, uint32_t* col, void* w, bool in_edges) {
  SUMINC_REQUIRE(c->has_graph, SUMINC_E_STATE, "no graph");
  const DevBuf& O = in_edges ? c->in_off : c->out_off;
  const DevBuf& C = in_edges ? c->in_col : c->out_col;
  const DevBuf& W = in_edges ? c->in_w : c->out_w;
  if (off) {
    std::vector<uint32_t> off32(c->n + 1);
    SUMINC_HIP(hipMemcpyAsync(off32.data(), O.p, ((size_t)c->n + 1) * 4, hipMemcpyDeviceToHost, c->stream));
    SUMINC_HIP(hipStreamSynchronize(c->stream));
    for (uint32_t i = 0; i <= c->n; ++i) off[i] = off32[i];
  }
  if (col && c->m) SUMINC_HIP(hipMemcpyAsync(col, C.p, c->m * 4, hipMemcpyDeviceToHost, c->stream));
  if (w && c->wkind && c->m) SUMINC_HIP(hipMemcpyAsync(w, W.p, c->m * 4, hipMemcpyDeviceToHost, c->stream));
  SUMINC_HIP(hipStreamSynchronize(c->stream));
}

}  // namespace suminc
