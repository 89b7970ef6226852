// batch.hip -- edge-update batches on the device: seeded batch generation, the
// CSR/CSC delta-merge with IncFragmentBuilder semantics, the PageRank reseed
// (AmendValue -1 on old / +1 on new out-edges of touched sources) and the SSSP
// repair (reset of the dependency subtree of every deleted or re-weighted parent
// edge, re-pull of reset vertices, relaxation of inserts).
//
// Reference: grape/fragment/inc_fragment_builder.h:29-133 (batch semantics),
// grape/worker/sum_sync_iter_worker.h:139-345 (reloadGraph + AmendValue),
// grape/worker/sum_sync_traversal_worker.h:113-352 (deltaCompute).
//
// Merge: ops are sorted by (u,v) (stable LSD radix sort, so the first "a" of a
// pair in file order is the first one met), collapsed into one group per pair
// {row, col, action, weight}; each group is probed in the old adjacency (binary
// search, rows are sorted).  New row length = old + #adds - #existing-pairs.  Old
// edges of untouched rows are copied to their shifted position; edges of touched
// rows get their position from two prefix counts over the row's groups (existing
// pairs before, adds before), adds are placed symmetrically.  The same kernels
// run on the transposed groups for the CSC.  Cost O(m + G log).
//
// Reseed: each of the two AmendValue passes is a push over the touched sources'
// rows, or, when their edges exceed m / pull_threshold_div, one pull over the
// CSC with the touched sources' contributions (chosen on the device).
#include <cstring>
#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <vector>

#include "ctx.hpp"
#include "device.hpp"
#include "pr.hpp"
#include "sketch.hpp"

namespace suminc {

// ---------------------------------------------------------------------------
// batch generation (synthetic input; same definition as the oracle's)
// ---------------------------------------------------------------------------
__global__ void k_gen_dels(const uint32_t* off, const uint32_t* col, uint32_t n, uint64_t m, uint32_t k0,
                           uint32_t k1, uint32_t batch_id, uint64_t n_del, suminc_update* ops) {
  uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_del) return;
  U4 o = philox4x32_10((uint32_t)j, (uint32_t)(j >> 32), batch_id, kStreamDel, k0, k1);
  uint64_t r = ((uint64_t)o.x << 32) | o.y;
  uint64_t idx = __umul64hi(r, m);
  uint32_t lo = 0, hi = n - 1;  // last row with off[row] <= idx
  while (lo < hi) {
    uint32_t mid = lo + ((hi - lo + 1) >> 1);
    if (off[mid] <= idx) lo = mid; else hi = mid - 1;
  }
  suminc_update op;
  op.type = SUMINC_OP_DEL; op.u = lo; op.v = col[idx]; op.w_u32 = 0; op.w_f32 = 0.f; op.pad = 0;
  ops[j] = op;
}

__global__ void k_gen_ins(uint32_t n, int wkind, uint32_t wmin, uint32_t wmax, uint32_t k0, uint32_t k1,
                          uint32_t batch_id, uint64_t n_ins, suminc_update* ops) {
  uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_ins) return;
  U4 o = philox4x32_10((uint32_t)j, (uint32_t)(j >> 32), batch_id, kStreamIns, k0, k1);
  uint32_t u = mulhi32(o.x, n), v = mulhi32(o.y, n);
  if (v == u) v = (v + 1u == n) ? 0u : v + 1u;
  suminc_update op;
  op.type = SUMINC_OP_ADD; op.u = u; op.v = v; op.pad = 0;
  op.w_u32 = wkind == SUMINC_W_U32 ? wmin + mulhi32(o.z, wmax - wmin + 1u) : 0u;
  op.w_f32 = wkind == SUMINC_W_F32 ? (float)((o.z >> 8) + 1u) * (1.0f / 16777216.0f) : 0.f;
  ops[j] = op;
}

void generate_batch(suminc_ctx* c, uint64_t seed, uint32_t batch_id, uint64_t n_del, uint64_t n_ins, uint32_t wmin,
                    uint32_t wmax, suminc_update* ops) {
  SUMINC_REQUIRE(c->has_graph, SUMINC_E_STATE, "no graph");
  SUMINC_REQUIRE(ops || n_del + n_ins == 0, SUMINC_E_INVALID, "null ops");
  SUMINC_REQUIRE(n_del == 0 || c->m > 0, SUMINC_E_INVALID, "cannot delete from an empty graph");
  SUMINC_REQUIRE(c->wkind != SUMINC_W_U32 || (wmin >= 1 && wmax >= wmin), SUMINC_E_INVALID, "bad weight range");
  const uint64_t total = n_del + n_ins;
  if (!total) return;
  DevBuf d;
  d.ensure(total * sizeof(suminc_update));
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  if (n_del)
    k_gen_dels<<<grid_for(n_del), kBlock, 0, c->stream>>>(c->out_off.as<uint32_t>(), c->out_col.as<uint32_t>(), c->n,
                                                          c->m, k0, k1, batch_id, n_del, d.as<suminc_update>());
  if (n_ins)
    k_gen_ins<<<grid_for(n_ins), kBlock, 0, c->stream>>>(c->n, c->wkind, wmin, wmax, k0, k1, batch_id, n_ins,
                                                         d.as<suminc_update>() + n_del);
  SUMINC_HIP(hipGetLastError());
  SUMINC_HIP(hipMemcpyAsync(ops, d.p, total * sizeof(suminc_update), hipMemcpyDeviceToHost, c->stream));
  SUMINC_HIP(hipStreamSynchronize(c->stream));
}

// ---------------------------------------------------------------------------
// grouping (all counts stay on the device; grids are sized by the op count)
// ---------------------------------------------------------------------------
__global__ void k_op_keys(const suminc_update* ops, uint64_t n_ops, uint64_t* keys, uint32_t* idx) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_ops) return;
  keys[i] = ((uint64_t)ops[i].u << 32) | ops[i].v;
  idx[i] = (uint32_t)i;
}

// Group heads: action 0 = add (first add of the pair in batch order, weight
// bits of that add), 1 = delete.
__global__ void k_group_heads(const uint64_t* skeys, const uint32_t* sidx, const suminc_update* ops, uint64_t n_ops,
                              int wkind, uint4* rec, uint8_t* flag) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_ops) return;
  uint64_t k = skeys[i];
  bool head = (i == 0) || skeys[i - 1] != k;
  flag[i] = head;
  if (!head) return;
  uint32_t act = 1, wbits = 0;
  for (uint64_t j = i; j < n_ops && skeys[j] == k; ++j) {
    const suminc_update& op = ops[sidx[j]];
    if (op.type == SUMINC_OP_ADD) {
      act = 0;
      wbits = wkind == SUMINC_W_F32 ? __float_as_uint(op.w_f32) : op.w_u32;
      break;
    }
  }
  rec[i] = make_uint4((uint32_t)(k >> 32), (uint32_t)k, act, wbits);
}

// Group arrays in structure-of-arrays form, for the CSR (row=u, col=v) or the
// transposed view for the CSC (row=v, col=u).  Sizes live on the device.
struct Groups {
  const unsigned* countp;   // number of groups G
  const uint32_t* pos;      // position of the pair in the old edge array (or its insertion point)
  const uint32_t* row;
  const uint32_t* col;
  const uint8_t* act;       // 0 add, 1 delete
  const uint32_t* w;        // weight bits of adds
  const uint8_t* ex;        // pair exists in the old graph
  const uint32_t* ex_pre;   // exclusive prefix of ex (G+1 valid)
  const uint32_t* add_pre;  // exclusive prefix of (act == 0) (G+1 valid)
  const uint32_t* tix;      // row -> index in touched list or INF32
  const uint32_t* t_gbeg;   // touched row -> first group (T+1)
  const uint32_t* t_list;   // touched rows, ascending (T)
  const unsigned* tcount;   // T
};

__global__ void k_split_rec(const uint4* rec, const unsigned* bc, uint32_t N, uint32_t* row, uint32_t* col,
                            uint8_t* act, uint32_t* w) {
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= bc[0]) return;
  uint4 r = rec[g];
  row[g] = r.x; col[g] = r.y; act[g] = (uint8_t)r.z; w[g] = r.w;
  (void)N;
}

// existence probe in the old out-CSR; also records the old weight bits
// existence probe in the old adjacency of one direction; also records the old
// weight bits (ex / eold_w may be null) and the pair's position in the old edge
// array (its slot if it exists, else the slot it is inserted before)
__global__ void k_probe(const uint32_t* off, const uint32_t* colv, const uint32_t* wv, const uint32_t* grow,
                        const uint32_t* gcol, const unsigned* bc, uint8_t* ex, uint32_t* eold_w, uint32_t* pos) {
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= bc[0]) return;
  uint32_t u = grow[g], v = gcol[g];
  uint32_t lo = off[u], hi = off[u + 1];
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (colv[mid] < v) lo = mid + 1; else hi = mid;
  }
  bool found = lo < off[u + 1] && colv[lo] == v;
  if (ex) ex[g] = found;
  if (eold_w) eold_w[g] = (found && wv) ? wv[lo] : 0u;
  pos[g] = lo;
}

__global__ void k_flags_u32(const uint8_t* ex, const uint8_t* act, const unsigned* bc, uint32_t N, uint32_t* exu,
                            uint32_t* addu) {
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g > N) return;
  const uint32_t G = bc[0];
  exu[g] = g < G ? ex[g] : 0u;
  addu[g] = g < G ? (act[g] == 0) : 0u;
}

__global__ void k_row_heads(const uint32_t* row, const unsigned* bc, uint32_t N, uint8_t* head) {
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= N) return;
  head[g] = g < bc[0] && ((g == 0) || row[g - 1] != row[g]);
}

// touched rows: list, first group, and row -> index map (reset after the batch)
__global__ void k_touched(const uint32_t* row, const uint32_t* t_raw, const unsigned* bc, int which, uint32_t N,
                          uint32_t* t_list, uint32_t* tix, uint32_t* t_gbeg) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t T = bc[which];
  if (t > T || t > N) return;
  if (t == T) { t_gbeg[T] = bc[0]; return; }
  uint32_t g = t_raw[t];
  t_gbeg[t] = g;
  t_list[t] = row[g];
  tix[row[g]] = t;
}

// Clears both touched-row maps (CSR: count bc[1], CSC: count bc[2]) for the next batch.
__global__ void k_untouch(const uint32_t* t_list, uint32_t* tix, const uint32_t* tt_list, uint32_t* ttix,
                          const unsigned* bc, uint32_t N) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < N && t < bc[1]) tix[t_list[t]] = INF32;
  if (t < N && t < bc[2]) ttix[tt_list[t]] = INF32;
}

__global__ void k_new_m(const unsigned* bc_in, const uint32_t* ex_pre, const uint32_t* add_pre, uint32_t m,
                        unsigned* bc) {
  if (threadIdx.x) return;
  const uint32_t G = bc_in[0];
  bc[3] = m + add_pre[G] - ex_pre[G];
}

// New row pointers: a row pointer moves by the net change (adds - existing pairs)
// of the groups of all rows before it, off'[v] = off[v] + add_pre[g] - ex_pre[g]
// with g the first group of the first touched row >= v (g = G past the last).
// Vertices go in tiles of kOffTile; vts[t] is the first touched row of tile t
// (touched rows are ascending), and a workgroup stages its tile's touched rows
// and their shifts in LDS.
constexpr uint32_t kOffTile = 8 * kBlock;

__global__ void k_off_tiles(const uint32_t* __restrict__ t_list, const unsigned* tcount, uint32_t ntiles,
                            uint32_t* vts) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t T = *tcount;
  if (i > T) return;
  const int64_t cur = i == T ? (int64_t)ntiles : (int64_t)(t_list[i] / kOffTile);
  const int64_t prev = i == 0 ? -1 : (int64_t)(t_list[i - 1] / kOffTile);
  for (int64_t t = prev + 1; t <= cur; ++t) vts[t] = i;
}

__global__ void __launch_bounds__(kBlock) k_new_off(const uint32_t* __restrict__ off, uint32_t n, Groups gs,
                                                    const uint32_t* __restrict__ vts, uint32_t* __restrict__ noff) {
  __shared__ uint32_t s_row[kOffTile];
  __shared__ uint32_t s_delta[kOffTile + 1];
  const uint32_t tile = blockIdx.x;
  // the tile's old row pointers are loaded first: their latency overlaps the
  // staging of the touched rows
  constexpr int kPer = kOffTile / kBlock;
  uint32_t ov[kPer];
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const uint32_t v = tile * kOffTile + j * kBlock + threadIdx.x;
    ov[j] = v <= n ? off[v] : 0u;
  }
  const uint32_t lo = vts[tile], hi = vts[tile + 1], r = hi - lo;  // touched rows of the tile (<= kOffTile)
  for (uint32_t i = threadIdx.x; i <= r; i += kBlock) {
    const uint32_t g = gs.t_gbeg[lo + i];
    if (i < r) s_row[i] = gs.t_list[lo + i];
    s_delta[i] = gs.add_pre[g] - gs.ex_pre[g];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const uint32_t v = tile * kOffTile + j * kBlock + threadIdx.x;
    if (v > n) break;
    uint32_t a = 0, b = r;  // first touched row >= v
    while (a < b) {
      const uint32_t mid = (a + b) >> 1;
      if (s_row[mid] < v) a = mid + 1; else b = mid;
    }
    noff[v] = ov[j] + s_delta[a];
  }
}

// Copy surviving old edges to their new positions.  Between two consecutive group
// positions the destination is the source shifted by a constant: an old edge e
// moves to e - (existing pairs grouped before e) + (adds grouped at or before e),
// and is dropped when its own pair is grouped as existing (deleted, or re-added
// with a new weight, which k_place_adds writes).  Groups are sorted by (row, col),
// so their positions are non-decreasing; each workgroup copies a tile of
// kCopyTile edges, finds its first group once and stages the tile's groups in LDS.
// Tile t owns the output range [base_t, base_{t+1}) (its surviving edges and the
// add slots among them), so the tiles partition the new edge array: a tile
// reads its edges with 16-byte loads, places each survivor in LDS at its offset
// in the range, and streams the range out with aligned 16-byte stores.  The add
// slots inside a range receive stale LDS words; k_place_adds, which runs after
// this kernel, writes them.  The old arrays are not read again and the new ones
// exceed the caches, so both directions use non-temporal accesses.
constexpr uint32_t kCopyTile = 1024;   // 4 edges per thread
constexpr uint32_t kCopyStage = 256;   // LDS-staged groups per tile

// first group of every tile (groups with pos >= t * kCopyTile); entry ntiles = G
__global__ void k_tile_groups(const unsigned* countp, const uint32_t* __restrict__ pos, uint32_t ntiles,
                              uint32_t* tstart) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > ntiles) return;
  const uint32_t G = *countp;
  if (t == ntiles) { tstart[t] = G; return; }
  const uint32_t e0 = t * kCopyTile;
  uint32_t lo = 0, hi = G;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (pos[mid] < e0) lo = mid + 1; else hi = mid;
  }
  tstart[t] = lo;
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(kBlock) k_copy_old(uint32_t m, const uint32_t* __restrict__ colv,
                                                     const uint32_t* __restrict__ wv, uint32_t* __restrict__ ncol,
                                                     uint32_t* __restrict__ nw, const uint32_t* __restrict__ tstart,
                                                     Groups gs) {
  static_assert(kCopyTile == 4 * kBlock, "one 16-byte load per thread and array");
  __shared__ uint32_t s_pos[kCopyStage];
  __shared__ uint32_t s_ex[kCopyStage];   // ex_pre | (pair exists) << 31
  __shared__ uint32_t s_add[kCopyStage];  // add_pre + (group is an add)
  // the tile's output range holds at most kCopyTile survivors plus kCopyStage adds
  __shared__ __attribute__((aligned(16))) uint32_t s_c[kCopyTile + kCopyStage + 8];
  __shared__ __attribute__((aligned(16))) uint32_t s_w[kCopyTile + kCopyStage + 8];
  const uint32_t e0 = blockIdx.x * kCopyTile;
  const uint32_t e1 = (uint32_t)min((uint64_t)m, (uint64_t)e0 + kCopyTile);
  const uint32_t j0 = tstart[blockIdx.x], cnt = tstart[blockIdx.x + 1] - j0;
  if (cnt > kCopyStage) {
    // a tile with more groups than fit in LDS: one edge per thread, global reads
    for (uint32_t e = e0 + threadIdx.x; e < e1; e += kBlock) {
      uint32_t lo = 0, hi = cnt;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (gs.pos[j0 + mid] < e) lo = mid + 1; else hi = mid;
      }
      uint32_t k = lo;
      bool dropped = false;
      while (k < cnt && gs.pos[j0 + k] == e) dropped |= gs.ex[j0 + k++] != 0;
      if (dropped) continue;
      const uint32_t dst = e - gs.ex_pre[j0 + lo] + gs.add_pre[j0 + k];
      ncol[dst] = colv[e];
      if (nw) nw[dst] = wv[e];
    }
    return;
  }
  for (uint32_t i = threadIdx.x; i < cnt; i += kBlock) {
    const uint32_t g = j0 + i;
    s_pos[i] = gs.pos[g];
    s_ex[i] = gs.ex_pre[g] | (gs.ex[g] ? 0x80000000u : 0u);
    s_add[i] = gs.add_pre[g] + (gs.act[g] == 0 ? 1u : 0u);
  }
  // output range [base, end): every survivor lands at or after base (each
  // existing pair grouped before it dropped a distinct edge of the tile)
  const uint32_t ex_base = gs.ex_pre[j0], add_base = gs.add_pre[j0];
  const uint32_t base = e0 - ex_base + add_base;
  const uint32_t end = e1 - gs.ex_pre[j0 + cnt] + gs.add_pre[j0 + cnt];
  const uint32_t abase = base & ~3u;
  const uint32_t eb = e0 + 4 * threadIdx.x;
  u32x4 cv = {0u, 0u, 0u, 0u}, wq = {0u, 0u, 0u, 0u};
  if (eb + 4 <= e1) {
    cv = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(colv + eb));
    if (nw) wq = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(wv + eb));
  } else {
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (eb + j < e1) {
        cv[j] = colv[eb + j];
        if (nw) wq[j] = wv[eb + j];
      }
  }
  __syncthreads();
  uint32_t i = 0;  // first staged group with pos >= the current edge
  {
    uint32_t hi = cnt;
    while (i < hi) {
      const uint32_t mid = (i + hi) >> 1;
      if (s_pos[mid] < eb) i = mid + 1; else hi = mid;
    }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint32_t e = eb + j;
    if (e >= e1) break;
    while (i < cnt && s_pos[i] < e) ++i;
    uint32_t k = i;
    bool dropped = false;
    while (k < cnt && s_pos[k] == e) {
      dropped |= (s_ex[k] >> 31) != 0;
      ++k;
    }
    if (dropped) continue;
    const uint32_t drops = i ? (s_ex[i - 1] & 0x7FFFFFFFu) + (s_ex[i - 1] >> 31) : ex_base;
    const uint32_t adds = k ? s_add[k - 1] : add_base;
    const uint32_t l = e - drops + adds - abase;
    s_c[l] = cv[j];
    s_w[l] = wq[j];
  }
  __syncthreads();
  // aligned 16-byte chunks [a0, a1) of the range, single words before and after
  const uint32_t a0 = (base + 3) & ~3u, a1 = end & ~3u;
  const bool chunks = a0 <= a1;             // else the range lies inside one chunk
  const uint32_t h1 = chunks ? a0 : end;
  for (uint32_t d = base + threadIdx.x; d < h1; d += kBlock) {
    ncol[d] = s_c[d - abase];
    if (nw) nw[d] = s_w[d - abase];
  }
  if (chunks) {
    for (uint32_t d = a1 + threadIdx.x; d < end; d += kBlock) {
      ncol[d] = s_c[d - abase];
      if (nw) nw[d] = s_w[d - abase];
    }
    for (uint32_t q = a0 / 4 + threadIdx.x; q < a1 / 4; q += kBlock) {
      __builtin_nontemporal_store(*reinterpret_cast<const u32x4*>(&s_c[q * 4 - abase]),
                                  reinterpret_cast<u32x4*>(ncol) + q);
      if (nw)
        __builtin_nontemporal_store(*reinterpret_cast<const u32x4*>(&s_w[q * 4 - abase]),
                                    reinterpret_cast<u32x4*>(nw) + q);
    }
  }
}

// ncolc (carried compact positions): the renamed column, pos[column], as well
__global__ void k_place_adds(const uint32_t* off, const uint32_t* colv, const uint32_t* noff, uint32_t* ncol,
                             uint32_t* nw, Groups gs, uint32_t* ncolc, const uint32_t* pos) {
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= *gs.countp || gs.act[g] != 0) return;
  uint32_t u = gs.row[g], v = gs.col[g];
  uint32_t ti = gs.tix[u];
  uint32_t b = gs.t_gbeg[ti];
  uint32_t old_before = gs.pos[g] - off[u];  // the pair's insertion point in the old row
  uint32_t kept_before = old_before - (gs.ex_pre[g] - gs.ex_pre[b]);
  uint32_t adds_before = gs.add_pre[g] - gs.add_pre[b];
  uint32_t dst = noff[u] + kept_before + adds_before;
  ncol[dst] = v;
  if (nw) nw[dst] = gs.w[g];
  if (ncolc) ncolc[dst] = pos[v];
}

__global__ void k_transpose_group_keys(const uint32_t* row, const uint32_t* col, const unsigned* bc, uint32_t N,
                                       uint64_t* keys, uint32_t* idx) {
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= N) return;
  keys[g] = g < bc[0] ? (((uint64_t)col[g] << 32) | row[g]) : ~0ull;
  idx[g] = g;
}

__global__ void k_gather_groups(const uint32_t* sidx, const unsigned* bc, const uint32_t* row, const uint32_t* col,
                                const uint8_t* act, const uint32_t* w, const uint8_t* ex, uint32_t* trow,
                                uint32_t* tcol, uint8_t* tact, uint32_t* tw, uint8_t* tex) {
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= bc[0]) return;
  uint32_t s = sidx[g];
  trow[g] = col[s];
  tcol[g] = row[s];
  tact[g] = act[s];
  tw[g] = w[s];
  tex[g] = ex[s];
}

// ---------------------------------------------------------------------------
// PageRank reseed: (u, type*d*r_u/deg) for touched u over the current CSR, as a
// push list and, with pos, as the pull contribution at u's compact source
// position (contrib is zero elsewhere); the edge total picks the scatter like a
// round does (k_reseed_mode)
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(kBlock) k_reseed_list(const uint32_t* t_list, const unsigned* bc, uint32_t N,
                                                        const uint32_t* off, const T* rank, const float* comp, T d,
                                                        T type, uint32_t* wl_v, T* wl_c, uint4* heavy, T* contrib,
                                                        const uint32_t* __restrict__ pos, Counters* cnt) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  bool light = false;
  uint32_t u = 0;
  T c = T(0);
  unsigned long long edges = 0;
  if (t < N && t < bc[1]) {
    u = t_list[t];
    uint32_t deg = off[u + 1] - off[u];
    T r = rank[u];
    if constexpr (sizeof(T) == 4) r = (T)((double)r - (double)comp[u]);
    if (deg > 0 && r != T(0)) {
      c = type * d * r / (T)deg;
      contrib[pos ? pos[u] : u] = c;  // deg > 0: u has a position
      edges = deg;
      if (deg <= kHeavyDeg) {
        light = true;
      } else {
        uint32_t nch = (deg + kChunk - 1) / kChunk;
        unsigned long long slot = atomicAdd(&cnt->heavy_count, (unsigned long long)nch);
        for (uint32_t k = 0; k < nch; ++k) {
          uint4 h;
          h.x = u; h.y = k;
          if constexpr (sizeof(T) == 4) { h.z = __float_as_uint((float)c); h.w = 0; }
          else {
            unsigned long long bits = __double_as_longlong((double)c);
            h.z = (uint32_t)bits; h.w = (uint32_t)(bits >> 32);
          }
          heavy[slot + k] = h;
        }
      }
    }
  }
  unsigned long long slot = wave_append(light, &cnt->wl_count[0]);
  if (light) { wl_v[slot] = u; wl_c[slot] = c; }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) edges += __shfl_xor(edges, o);
  if (lane_id() == 0 && edges) atomicAdd(&cnt->round_edges, edges);
}

// Scatter of the reseed: pull over the CSC when its edges exceed m / div (the
// rounds' rule), else push; the edges add to the batch's reseed total (aux[5]).
__global__ void k_reseed_mode(Counters* cnt, uint64_t m, int div) {
  if (threadIdx.x) return;
  cnt->reseed_mode = div > 0 && cnt->round_edges * (unsigned long long)div > m ? 1ull : 0ull;
  cnt->aux[5] += cnt->round_edges;
  cnt->round_edges = 0;
}

// Same scatter as the PageRank push round, over the reseed list.
template <class T>
__global__ void __launch_bounds__(kBlock) k_reseed_push(const uint32_t* __restrict__ off, const uint32_t* __restrict__ col,
                                                        T* __restrict__ delta, const uint32_t* __restrict__ wl_v,
                                                        const T* __restrict__ wl_c, const uint4* __restrict__ heavy,
                                                        const Counters* cnt) {
  __shared__ uint32_t s_scan[kBlock + 1];
  __shared__ uint32_t s_start[kBlock];
  __shared__ uint32_t s_warp[kBlock / 64 + 1];
  __shared__ T s_c[kBlock];
  if (cnt->reseed_mode != 0) return;
  const unsigned long long nl = cnt->wl_count[0];
  const unsigned long long nh = cnt->heavy_count;
  const uint32_t vt = frontier_tile<kBlock>(nl);
  const unsigned long long tiles = (nl + vt - 1) / vt;
  for (unsigned long long t = blockIdx.x; t < tiles; t += gridDim.x) {
    unsigned long long i = t * vt + threadIdx.x;
    uint32_t start = 0, len = 0;
    if (threadIdx.x < vt && i < nl) {
      uint32_t v = wl_v[i];
      start = off[v];
      len = off[v + 1] - start;
      s_c[threadIdx.x] = wl_c[i];
    }
    block_adjacency<kBlock>(start, len, s_scan, s_start, s_warp,
                            [&](int slot, uint32_t, uint32_t e) { unsafeAtomicAdd(&delta[col[e]], s_c[slot]); });
  }
  for (unsigned long long h = blockIdx.x; h < nh; h += gridDim.x) {
    uint4 hv = heavy[h];
    T c;
    if constexpr (sizeof(T) == 4) c = __uint_as_float(hv.z);
    else c = __longlong_as_double((long long)(((unsigned long long)hv.w << 32) | hv.z));
    uint32_t b = off[hv.x] + hv.y * kChunk;
    uint32_t e = min(off[hv.x + 1], b + kChunk);
    for (uint32_t k = b + threadIdx.x; k < e; k += kBlock) unsafeAtomicAdd(&delta[col[k]], c);
  }
}

__global__ void k_reset_lists(Counters* cnt) {
  if (threadIdx.x) return;
  cnt->wl_count[0] = 0;
  cnt->heavy_count = 0;
  cnt->round_edges = 0;
  cnt->reseed_mode = 0;
}

// ---------------------------------------------------------------------------
// SSSP repair
// ---------------------------------------------------------------------------
// roots: deleted or re-weighted existing pair (u,v) with parent(v) == u
__global__ void k_sssp_roots(const uint32_t* grow, const uint32_t* gcol, const uint8_t* gact, const uint32_t* gw,
                             const uint8_t* ex, const uint32_t* eold_w, const unsigned* bc, uint32_t N,
                             const unsigned long long* dp, uint32_t src, uint32_t* rbits, uint32_t* rlist,
                             Counters* cnt) {
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  bool add = false;
  uint32_t v = 0;
  if (g < N && g < bc[0] && ex[g]) {
    uint32_t u = grow[g];
    v = gcol[g];
    bool changed = gact[g] == 1 || gw[g] != eold_w[g];
    unsigned long long x = dp[v];
    if (changed && v != src && (uint32_t)(x >> 32) != INF32 && (uint32_t)x == u) {
      uint32_t bit = 1u << (v & 31);
      add = !(atomicOr(&rbits[v >> 5], bit) & bit);
    }
  }
  unsigned long long slot = wave_append(add, &cnt->aux[4]);
  if (add) rlist[slot] = v;
}

// one round of the subtree walk over the OLD out-CSR: children x of u with
// parent(x) == u join the reset set (sum_sync_traversal_worker.h:170-194)
// Rows longer than kHeavyDeg are not walked by the tile that holds them: they are
// cut into kChunk-edge chunks (heavy list) that the *_heavy kernels spread over
// all workgroups, so a hub does not serialise one workgroup.
__device__ __forceinline__ void heavy_chunks(uint32_t v, uint32_t len, uint2* heavy, unsigned long long* ctr) {
  const uint32_t nch = (len + kChunk - 1) / kChunk;
  const unsigned long long slot = atomicAdd(ctr, (unsigned long long)nch);
  for (uint32_t k = 0; k < nch; ++k) heavy[slot + k] = make_uint2(v, k);
}

// x joins the walk when its parent is u.  A joining vertex with a long row is
// also cut into chunks at once, walked by the next round beside the light
// rows of the next frontier (not one round later, after the frontier has met it).
__device__ __forceinline__ void reset_child(const uint32_t* off, const unsigned long long* dp, uint32_t u, uint32_t x,
                                            uint32_t* rbits, uint32_t* rlist, unsigned long long base,
                                            unsigned long long* ctr, uint2* hnew, unsigned long long* hcount) {
  bool add = false;
  if (x != u && (uint32_t)dp[x] == u && (uint32_t)(dp[x] >> 32) != INF32) {
    uint32_t bit = 1u << (x & 31);
    add = !(atomicOr(&rbits[x >> 5], bit) & bit);
  }
  unsigned long long slot = wave_append(add, ctr);
  if (add) {
    rlist[base + slot] = x;
    const uint32_t len = off[x + 1] - off[x];
    if (len > kHeavyDeg) heavy_chunks(x, len, hnew, hcount);
  }
}

// The walk's rounds are numbered on the host, so no kernel closes a round.  rlist
// holds the roots (counted by aux[4]) followed by every round's new vertices;
// round r's frontier is rlist[S_r, S_r + C_r) with S_r in aux[2 + r % 2] (S_0 = 0)
// and C_r the count of round r - 1's appends, rc[r % 3] (C_0 = aux[4]); round r
// appends at S_r + C_r, counting in rc[(r + 1) % 3].  Rows longer than kHeavyDeg
// become kChunk-edge chunks walked by the next round (the walk is a reachability
// closure, so the order does not matter): round 0 cuts its heavy roots, later
// rounds cut each heavy vertex as it joins (reset_child) and skip it in the
// frontier; round r writes chunks to heavy half r % 2 counted by hc[r % 3] and
// walks the previous round's (hc[(r - 1) % 3]).  Its
// first workgroup records S_{r+1}, clears rc[(r + 2) % 3] and hc[(r + 1) % 3] (last
// read by round r - 1), and when nothing is left sets done, the total in aux[4]
// and the depth in aux[9].
__global__ void __launch_bounds__(kBlock) k_reset_round(const uint32_t* off, const uint32_t* colv,
                                                        const unsigned long long* dp, uint32_t* rbits,
                                                        uint32_t* rlist, uint2* heavy, uint64_t heavy_half,
                                                        Counters* cnt, uint32_t r) {
  __shared__ uint32_t s_scan[kBlock + 1];
  __shared__ uint32_t s_start[kBlock];
  __shared__ uint32_t s_warp[kBlock / 64 + 1];
  __shared__ uint32_t s_u[kBlock];
  if (cnt->done) return;
  const unsigned long long S = r == 0 ? 0ull : cnt->aux[2 + (r & 1)];
  const unsigned long long C = r == 0 ? cnt->aux[4] : cnt->rc[r % 3];
  const unsigned long long nh = r ? cnt->hc[(r - 1) % 3] : 0ull;
  unsigned long long* out = &cnt->rc[(r + 1) % 3];
  unsigned long long* hcount = &cnt->hc[r % 3];
  uint2* hnew = heavy + (r & 1) * heavy_half;
  const uint2* hold = heavy + ((r + 1) & 1) * heavy_half;
  const bool stop = C == 0 && nh == 0;  // counts of earlier launches: the whole launch agrees
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    cnt->aux[2 + ((r + 1) & 1)] = S + C;
    cnt->rc[(r + 2) % 3] = 0;
    cnt->hc[(r + 1) % 3] = 0;
    if (stop) {
      cnt->done = 1;
      cnt->aux[4] = S;  // every reset vertex
      cnt->aux[9] = r;  // walk rounds
    }
  }
  if (stop) return;
  const unsigned long long b = S, e = S + C, base = e;
  const uint32_t vt = frontier_tile<kBlock>(C);
  const unsigned long long tiles = (C + vt - 1) / vt;
  for (unsigned long long t = blockIdx.x; t < tiles; t += gridDim.x) {
    unsigned long long i = b + t * vt + threadIdx.x;
    uint32_t start = 0, len = 0;
    if (threadIdx.x < vt && i < e) {
      uint32_t u = rlist[i];
      s_u[threadIdx.x] = u;
      start = off[u];
      len = off[u + 1] - start;
      if (len > kHeavyDeg) {
        if (r == 0) heavy_chunks(u, len, hnew, hcount);  // later rounds: cut when it joined
        len = 0;
      }
    }
    block_adjacency<kBlock>(start, len, s_scan, s_start, s_warp, [&](int slot, uint32_t, uint32_t k) {
      reset_child(off, dp, s_u[slot], colv[k], rbits, rlist, base, out, hnew, hcount);
    });
  }
  // the previous round's heavy chunks
  for (unsigned long long h = blockIdx.x; h < nh; h += gridDim.x) {
    const uint2 hv = hold[h];
    const uint32_t u = hv.x;
    const uint32_t cb = off[u] + hv.y * kChunk;
    const uint32_t ce = min(off[u + 1], cb + kChunk);
    for (uint32_t k0 = cb; k0 < ce; k0 += kBlock) {  // same trip count for every lane (wave ballots)
      const uint32_t k = k0 + threadIdx.x;
      if (k < ce) {
        reset_child(off, dp, u, colv[k], rbits, rlist, base, out, hnew, hcount);
      } else {
        unsigned long long unused = wave_append(false, out);
        (void)unused;
      }
    }
  }
}

// Resets the words of the walk's vertices; also clears the chunk count the re-pull
// appends to.
__global__ void k_reset_apply(const uint32_t* rlist, Counters* cnt, unsigned long long* dp) {
  if (blockIdx.x == 0 && threadIdx.x == 0) cnt->heavy_count = 0;
  const unsigned long long nr = cnt->aux[4];
  for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < nr;
       i += (unsigned long long)gridDim.x * blockDim.x) {
    uint32_t x = rlist[i];
    dp[x] = ((unsigned long long)INF32 << 32) | x;
  }
}

// reset vertices pull from surviving (non-reset) in-neighbours over the NEW CSC
__device__ __forceinline__ unsigned long long pull_candidate(const uint32_t* icol, const uint32_t* iw,
                                                             const uint32_t* rbits, const unsigned long long* dp,
                                                             uint32_t k, int app) {
  const uint32_t u = icol[k];
  if ((rbits[u >> 5] >> (u & 31)) & 1u) return ~0ull;
  const uint32_t du = (uint32_t)(dp[u] >> 32);
  if (du == INF32) return ~0ull;
  const unsigned long long nd = trav_extend(app, du, iw ? iw[k] : 1u);
  if (nd >= INF32) return ~0ull;
  return (nd << 32) | u;
}

__global__ void __launch_bounds__(kBlock) k_reset_pull(const uint32_t* ioff, const uint32_t* icol,
                                                       const uint32_t* iw, const uint32_t* rbits,
                                                       const uint32_t* rlist, Counters* cnt,
                                                       unsigned long long* dp, uint2* heavy, int app) {
  __shared__ uint32_t s_scan[kBlock + 1];
  __shared__ uint32_t s_start[kBlock];
  __shared__ uint32_t s_warp[kBlock / 64 + 1];
  __shared__ uint32_t s_x[kBlock];
  const unsigned long long nr = cnt->aux[4];
  const uint32_t vt = frontier_tile<kBlock>(nr);
  const unsigned long long tiles = (nr + vt - 1) / vt;
  for (unsigned long long t = blockIdx.x; t < tiles; t += gridDim.x) {
    unsigned long long i = t * vt + threadIdx.x;
    uint32_t start = 0, len = 0;
    if (threadIdx.x < vt && i < nr) {
      uint32_t x = rlist[i];
      s_x[threadIdx.x] = x;
      start = ioff[x];
      len = ioff[x + 1] - start;
      if (len > kHeavyDeg) {
        heavy_chunks(x, len, heavy, &cnt->heavy_count);
        len = 0;
      }
    }
    block_adjacency<kBlock>(start, len, s_scan, s_start, s_warp, [&](int slot, uint32_t, uint32_t k) {
      const unsigned long long cand = pull_candidate(icol, iw, rbits, dp, k, app);
      // the reset words only decrease here: a candidate no smaller than a plain
      // read cannot win and skips the atomic
      if (cand != ~0ull && cand < dp[s_x[slot]]) atomicMin(&dp[s_x[slot]], cand);
    });
  }
}

// heavy in-rows of reset vertices: block minimum per chunk, one atomicMin
__global__ void __launch_bounds__(kBlock) k_reset_pull_heavy(const uint32_t* ioff, const uint32_t* icol,
                                                             const uint32_t* iw, const uint32_t* rbits,
                                                             const uint2* heavy, Counters* cnt,
                                                             unsigned long long* dp, int app) {
  __shared__ unsigned long long s_min[kBlock / 64];
  const unsigned long long nh = cnt->heavy_count;
  for (unsigned long long h = blockIdx.x; h < nh; h += gridDim.x) {
    const uint2 hv = heavy[h];
    const uint32_t x = hv.x;
    const uint32_t b = ioff[x] + hv.y * kChunk;
    const uint32_t e = min(ioff[x + 1], b + kChunk);
    unsigned long long best = ~0ull;
    for (uint32_t k = b + threadIdx.x; k < e; k += kBlock) {
      const unsigned long long cand = pull_candidate(icol, iw, rbits, dp, k, app);
      best = cand < best ? cand : best;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const unsigned long long y = __shfl_xor(best, o);
      best = y < best ? y : best;
    }
    if (lane_id() == 0) s_min[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long m = ~0ull;
      for (int w = 0; w < kBlock / 64; ++w) m = s_min[w] < m ? s_min[w] : m;
      if (m != ~0ull) atomicMin(&dp[x], m);
    }
    __syncthreads();
  }
}

// Seeds SSSP list 0 with every finite reset vertex and clears the reset bitmap
// (the next repair starts from a clear one) and the re-pull's chunk count.
__global__ void k_reset_seed(const uint32_t* rlist, uint32_t* rbits, const unsigned long long* dp, uint32_t* wl0,
                             uint32_t* bits0, Counters* cnt, unsigned long long* seeds) {
  if (blockIdx.x == 0 && threadIdx.x == 0) cnt->heavy_count = 0;
  const unsigned long long nr = cnt->aux[4];
  const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
  // whole waves iterate together (the append's ballot)
  for (unsigned long long i0 = (unsigned long long)blockIdx.x * blockDim.x; i0 < nr; i0 += stride) {
    const unsigned long long i = i0 + threadIdx.x;
    bool add = false;
    uint32_t x = 0;
    if (i < nr) {
      x = rlist[i];
      atomicAnd(&rbits[x >> 5], ~(1u << (x & 31)));
      if ((uint32_t)(dp[x] >> 32) != INF32) {
        uint32_t bit = 1u << (x & 31);
        add = !(atomicOr(&bits0[x >> 5], bit) & bit);
      }
    }
    unsigned long long slot = wave_append(add, seeds);
    if (add) wl0[slot] = x;
  }
}

// relax every add (new pair or new weight) from its current source distance
__global__ void k_insert_relax(const uint32_t* grow, const uint32_t* gcol, const uint8_t* gact, const uint32_t* gw,
                               const unsigned* bc, uint32_t N, int unit, int app, unsigned long long* dp, uint32_t* wl0,
                               uint32_t* bits0, unsigned long long* seeds) {
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  bool add = false;
  uint32_t v = 0;
  if (g < N && g < bc[0] && gact[g] == 0) {
    uint32_t u = grow[g];
    v = gcol[g];
    uint32_t du = (uint32_t)(dp[u] >> 32);
    unsigned long long nd = trav_extend(app, du, unit ? 1u : gw[g]);
    // SSWP: strictly better words only (no parent change on a tie)
    if (du != INF32 && nd < INF32 && (app != SUMINC_APP_SSWP || nd < (dp[v] >> 32))) {
      unsigned long long cand = (nd << 32) | u;
      unsigned long long old = atomicMin(&dp[v], cand);
      if ((old >> 32) > nd) {
        uint32_t bit = 1u << (v & 31);
        add = !(atomicOr(&bits0[v >> 5], bit) & bit);
      }
    }
  }
  unsigned long long slot = wave_append(add, seeds);
  if (add) wl0[slot] = v;
}

__global__ void k_sssp_repair_prepare(Counters* cnt) {
  if (threadIdx.x) return;
  cnt->wl_count[0] = 0;
  cnt->wl_count[1] = 0;
  cnt->heavy_count = 0;
  cnt->aux[1] = 0;
  cnt->aux[2] = 0;
  cnt->aux[3] = 0;
  cnt->aux[4] = 0;
  cnt->aux[9] = 0;
  cnt->rc[0] = cnt->rc[1] = cnt->rc[2] = 0;
  cnt->hc[0] = cnt->hc[1] = cnt->hc[2] = 0;
  cnt->done = 0;
  cnt->rounds = 0;
  cnt->edges = 0;
}
// End of a batch: the group counts go out with the counters (one copy to the
// host), and a finished repair walk leaves the stop flag clear for the query.
__global__ void k_batch_end(Counters* cnt, const unsigned* bc, int undone) {
  if (threadIdx.x) return;
  for (int i = 0; i < 4; ++i) cnt->bcopy[i] = bc[i];
  if (undone) cnt->done = 0;
}
__global__ void k_batch_zero(Counters* cnt, unsigned* val) {
  if (threadIdx.x) return;
  cnt->aux[5] = 0;
  val[0] = 0;            // invalid ops
  val[1] = 0xFFFFFFFFu;  // first invalid line
}
__global__ void k_fill_u32(uint32_t* p, uint32_t n, uint32_t v) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// host orchestration
// ---------------------------------------------------------------------------
template <class F>
static void cub_call(suminc_ctx* c, DevBuf& tmp, F&& f) {
  size_t bytes = 0;
  SUMINC_HIP(f((void*)nullptr, bytes));
  tmp.ensure(bytes);
  SUMINC_HIP(f(tmp.p, bytes));
}

// Rebuilds one adjacency direction into the workspace buffers, then swaps them
// with the live arrays (the old arrays become next batch's workspace).
// colc / ncolc (the CSC with carried compact positions): the renamed columns are
// merged the same way, an added edge's from pos.
static void merge_direction(suminc_ctx* c, DevBuf& off, DevBuf& col, DevBuf& w, DevBuf& noff, DevBuf& ncol,
                            DevBuf& nw, const Groups& gs, uint32_t N, uint64_t cap_m, size_t prof_slot,
                            DevBuf* colc = nullptr, DevBuf* ncolc = nullptr, const uint32_t* pos = nullptr) {
  const uint32_t n = c->n;
  BatchWs& b = c->bw;
  noff.ensure(((size_t)n + 1) * 4);
  ncol.ensure_slack(cap_m * 4 + 4);
  if (c->wkind) nw.ensure_slack(cap_m * 4 + 4);
  if (colc) ncolc->ensure_slack(cap_m * 4 + 4);
  const uint32_t vtiles = (n + 1 + kOffTile - 1) / kOffTile;
  b.vts.ensure(((size_t)vtiles + 1) * 4);
  k_off_tiles<<<grid_for((uint64_t)N + 1), kBlock, 0, c->stream>>>(gs.t_list, gs.tcount, vtiles, b.vts.as<uint32_t>());
  k_new_off<<<vtiles, kBlock, 0, c->stream>>>(off.as<uint32_t>(), n, gs, b.vts.as<uint32_t>(), noff.as<uint32_t>());
  if (c->m) {
    const uint32_t m = (uint32_t)c->m;
    const uint32_t ntiles = (m + kCopyTile - 1) / kCopyTile;
    b.tstart.ensure_slack(((size_t)ntiles + 1) * 4);
    k_tile_groups<<<grid_for((uint64_t)ntiles + 1), kBlock, 0, c->stream>>>(gs.countp, gs.pos, ntiles,
                                                                           b.tstart.as<uint32_t>());
    if (c->prof_on) SUMINC_HIP(hipEventRecord(prof_event(c, prof_slot), c->stream));
    k_copy_old<<<ntiles, kBlock, 0, c->stream>>>(m, col.as<uint32_t>(), c->wkind ? w.as<uint32_t>() : nullptr,
                                                 ncol.as<uint32_t>(), c->wkind ? nw.as<uint32_t>() : nullptr,
                                                 b.tstart.as<uint32_t>(), gs);
    if (c->prof_on) SUMINC_HIP(hipEventRecord(prof_event(c, prof_slot + 1), c->stream));
    if (colc)
      k_copy_old<<<ntiles, kBlock, 0, c->stream>>>(m, colc->as<uint32_t>(), nullptr, ncolc->as<uint32_t>(), nullptr,
                                                   b.tstart.as<uint32_t>(), gs);
  }
  k_place_adds<<<grid_for(N), kBlock, 0, c->stream>>>(off.as<uint32_t>(), col.as<uint32_t>(), noff.as<uint32_t>(),
                                                      ncol.as<uint32_t>(), c->wkind ? nw.as<uint32_t>() : nullptr,
                                                      gs, colc ? ncolc->as<uint32_t>() : nullptr, pos);
  SUMINC_HIP(hipGetLastError());
  off.swap(noff);
  col.swap(ncol);
  if (c->wkind) w.swap(nw);
  if (colc) colc->swap(*ncolc);
}

// AmendValue over the touched sources' out-edges of the current graph: push for a
// few edges, or, when they exceed m / div, one pull over the CSC (whose work
// lists must be current: `lists` false leaves the pull out).
template <class T>
static void pr_reseed(suminc_ctx* c, uint32_t N, double type, bool lists) {
  Counters* cnt = dev_counters(c);
  BatchWs& b = c->bw;
  const int div = lists ? c->prm.pull_threshold_div : 0;
  k_reset_lists<<<1, 64, 0, c->stream>>>(cnt);
  if (div > 0) SUMINC_HIP(hipMemsetAsync(c->pr_contrib.p, 0, (size_t)c->n * sizeof(T), c->stream));
  k_reseed_list<T><<<grid_for(N), kBlock, 0, c->stream>>>(b.t_list.as<uint32_t>(), b.bc.as<unsigned>(), N,
                                                          c->out_off.as<uint32_t>(), c->pr_rank.as<T>(),
                                                          c->pr_comp.as<float>(), (T)c->prm.damping, (T)type,
                                                          c->wl0.as<uint32_t>(), c->wl1.as<T>(), c->heavy.as<uint4>(),
                                                          c->pr_contrib.as<T>(),
                                                          div > 0 ? pull_pos(c) : nullptr, cnt);
  k_reseed_mode<<<1, 64, 0, c->stream>>>(cnt, c->m, div);
  k_reseed_push<T><<<push_grid(), kBlock, 0, c->stream>>>(c->out_off.as<uint32_t>(), c->out_col.as<uint32_t>(),
                                                          c->pr_delta.as<T>(), c->wl0.as<uint32_t>(), c->wl1.as<T>(),
                                                          c->heavy.as<uint4>(), cnt);
  if (div > 0) pull_scatter<T>(c, 1);
  k_reset_lists<<<1, 64, 0, c->stream>>>(cnt);
  SUMINC_HIP(hipGetLastError());
}

// Builds the per-row group view (heads, prefix counts, touched rows) for one
// direction; `which` selects the touched-row count slot in bc.
static void build_view(suminc_ctx* c, uint32_t N, const uint32_t* row, const uint8_t* ex, const uint8_t* act,
                       DevBuf& exu, DevBuf& addu, DevBuf& expre, DevBuf& addpre, DevBuf& head, DevBuf& t_raw,
                       DevBuf& t_list, DevBuf& t_gbeg, DevBuf& tix, int which) {
  BatchWs& b = c->bw;
  const unsigned* bc = b.bc.as<unsigned>();
  k_flags_u32<<<grid_for((uint64_t)N + 1), kBlock, 0, c->stream>>>(ex, act, bc, N, exu.as<uint32_t>(),
                                                                   addu.as<uint32_t>());
  cub_call(c, c->tmp, [&](void* p, size_t& bytes) {
    return rocprim::exclusive_scan(p, bytes, exu.as<uint32_t>(), expre.as<uint32_t>(), 0u, (size_t)N + 1,
                                   rocprim::plus<uint32_t>(), c->stream);
  });
  cub_call(c, c->tmp, [&](void* p, size_t& bytes) {
    return rocprim::exclusive_scan(p, bytes, addu.as<uint32_t>(), addpre.as<uint32_t>(), 0u, (size_t)N + 1,
                                   rocprim::plus<uint32_t>(), c->stream);
  });
  k_row_heads<<<grid_for(N), kBlock, 0, c->stream>>>(row, bc, N, head.as<uint8_t>());
  cub_call(c, c->tmp, [&](void* p, size_t& bytes) {
    return rocprim::select(p, bytes, rocprim::counting_iterator<uint32_t>(0), head.as<uint8_t>(),
                           t_raw.as<uint32_t>(), b.bc.as<unsigned>() + which, (size_t)N, c->stream);
  });
  k_touched<<<grid_for((uint64_t)N + 1), kBlock, 0, c->stream>>>(row, t_raw.as<uint32_t>(), bc, which, N,
                                                                 t_list.as<uint32_t>(), tix.as<uint32_t>(),
                                                                 t_gbeg.as<uint32_t>());
}

static void ensure_workspace(suminc_ctx* c, uint64_t n_ops) {
  BatchWs& b = c->bw;
  const size_t N = n_ops;
  b.ops.ensure(N * sizeof(suminc_update));
  b.keys_a.ensure(N * 8); b.keys_b.ensure(N * 8); b.idx_a.ensure(N * 4); b.idx_b.ensure(N * 4);
  b.rec.ensure(N * 16); b.grec.ensure(N * 16); b.flags.ensure(N);
  for (DevBuf* d : {&b.g_row, &b.g_col, &b.g_w, &b.g_eoldw, &b.tg_row, &b.tg_col, &b.tg_w, &b.g_pos, &b.tg_pos})
    d->ensure(N * 4 + 4);
  for (DevBuf* d : {&b.g_act, &b.g_ex, &b.head, &b.tg_act, &b.tg_ex, &b.thead}) d->ensure(N + 1);
  for (DevBuf* d : {&b.g_exu, &b.g_addu, &b.g_expre, &b.g_addpre, &b.t_raw, &b.t_list, &b.t_gbeg, &b.tg_exu,
                    &b.tg_addu, &b.tg_expre, &b.tg_addpre, &b.tt_raw, &b.tt_list, &b.tt_gbeg})
    d->ensure((N + 1) * 4);
  b.bc.ensure(64);
  b.val.ensure(8);
  if (b.tix_n != c->n || !b.tix.p) {
    b.tix.ensure((size_t)c->n * 4);
    b.ttix.ensure((size_t)c->n * 4);
    k_fill_u32<<<grid_for(c->n), kBlock, 0, c->stream>>>(b.tix.as<uint32_t>(), c->n, INF32);
    k_fill_u32<<<grid_for(c->n), kBlock, 0, c->stream>>>(b.ttix.as<uint32_t>(), c->n, INF32);
    b.tix_n = c->n;
  }
}

// Device-side check of a device-resident batch: counts invalid ops and keeps the
// smallest invalid index (out[0] = count, out[1] = first index).
__global__ void k_validate_ops(const suminc_update* ops, uint32_t N, uint32_t n, int wkind, unsigned* out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const suminc_update o = ops[i];
  bool bad = (o.type != SUMINC_OP_ADD && o.type != SUMINC_OP_DEL) || o.u >= n || o.v >= n ||
             (wkind == SUMINC_W_U32 && o.type == SUMINC_OP_ADD && o.w_u32 < 1);
  if (bad) {
    atomicAdd(&out[0], 1u);
    atomicMin(&out[1], i);
  }
}

void apply_batch(suminc_ctx* c, const suminc_update* ops, uint64_t n_ops, bool ops_on_device) {
  SUMINC_REQUIRE(c->has_graph, SUMINC_E_STATE, "no graph");
  SUMINC_REQUIRE(ops || n_ops == 0, SUMINC_E_INVALID, "null ops");
  SUMINC_REQUIRE(n_ops < 0x7FFFFFFFull, SUMINC_E_TOO_LARGE, "batch too large");
  if (!ops_on_device) {
    for (uint64_t i = 0; i < n_ops; ++i) {
      SUMINC_REQUIRE(ops[i].type == SUMINC_OP_ADD || ops[i].type == SUMINC_OP_DEL, SUMINC_E_INVALID,
                     "op type must be SUMINC_OP_ADD or SUMINC_OP_DEL (line " + std::to_string(i) + ")");
      SUMINC_REQUIRE(ops[i].u < c->n && ops[i].v < c->n, SUMINC_E_INVALID,
                     "vertex id out of range (line " + std::to_string(i) + ")");
      SUMINC_REQUIRE(c->wkind != SUMINC_W_U32 || ops[i].type == SUMINC_OP_DEL || ops[i].w_u32 >= 1,
                     SUMINC_E_INVALID, "integer weights must be at least 1 (line " + std::to_string(i) + ")");
    }
  }
  const uint64_t cap_m = c->m + n_ops;
  SUMINC_REQUIRE(cap_m < 0xFFFFFFFFull, SUMINC_E_TOO_LARGE, "edge count could exceed 2^32-1");
  c->stats = suminc_stats{};
  if (n_ops == 0) return;
  const uint32_t n = c->n;
  const uint32_t N = (uint32_t)n_ops;
  Counters* cnt = dev_counters(c);
  ensure_workspace(c, n_ops);
  BatchWs& b = c->bw;
  k_batch_zero<<<1, 64, 0, c->stream>>>(cnt, b.val.as<unsigned>());
  if (ops_on_device) {
    // validated before any kernel indexes by the ops' vertex ids; the sort and
    // grouping below (which only read the ops) run while the host waits for it
    k_validate_ops<<<grid_for(N), kBlock, 0, c->stream>>>(ops, N, n, c->wkind, b.val.as<unsigned>());
    SUMINC_HIP(hipGetLastError());
    check_begin(c, b.val.p, 8);
  }
  // heavy-chunk list sized for the largest graph this batch can produce
  if (c->pr_ready || c->sssp_ready) c->heavy.ensure_slack((cap_m / kChunk + cap_m / kHeavyDeg + 2) * sizeof(uint4));
  unsigned* bc = b.bc.as<unsigned>();
  // An SSSP state that was initialised but never run is still the initial state:
  // it is simply re-initialised on the new graph instead of repaired.
  const bool sssp_repair = c->sssp_ready && !c->sssp_pending;
  const bool sssp_reinit = c->sssp_ready && c->sssp_pending;

  // 1. upload + stable sort by (u,v) + one group per pair
  const suminc_update* dops = ops;  // device ops are read in place
  if (!ops_on_device) {
    SUMINC_HIP(hipMemcpyAsync(b.ops.p, ops, n_ops * sizeof(suminc_update), hipMemcpyHostToDevice, c->stream));
    dops = b.ops.as<suminc_update>();
  }
  k_op_keys<<<grid_for(N), kBlock, 0, c->stream>>>(dops, N, b.keys_a.as<uint64_t>(),
                                                   b.idx_a.as<uint32_t>());
  cub_call(c, c->tmp, [&](void* p, size_t& bytes) {
    return rocprim::radix_sort_pairs(p, bytes, b.keys_a.as<uint64_t>(), b.keys_b.as<uint64_t>(),
                                     b.idx_a.as<uint32_t>(), b.idx_b.as<uint32_t>(), (size_t)N, 0, 64, c->stream);
  });
  k_group_heads<<<grid_for(N), kBlock, 0, c->stream>>>(b.keys_b.as<uint64_t>(), b.idx_b.as<uint32_t>(),
                                                       dops, N, c->wkind, b.rec.as<uint4>(),
                                                       b.flags.as<uint8_t>());
  cub_call(c, c->tmp, [&](void* p, size_t& bytes) {
    return rocprim::select(p, bytes, b.rec.as<uint4>(), b.flags.as<uint8_t>(), b.grec.as<uint4>(), bc, (size_t)N,
                           c->stream);
  });

  // 2. CSR view: existence probe, prefix counts, touched rows
  k_split_rec<<<grid_for(N), kBlock, 0, c->stream>>>(b.grec.as<uint4>(), bc, N, b.g_row.as<uint32_t>(),
                                                     b.g_col.as<uint32_t>(), b.g_act.as<uint8_t>(),
                                                     b.g_w.as<uint32_t>());
  if (ops_on_device) {
    const unsigned* res = static_cast<const unsigned*>(check_wait(c));
    SUMINC_REQUIRE(res[0] == 0, SUMINC_E_INVALID,
                   std::to_string(res[0]) + " invalid ops (first at line " + std::to_string(res[1]) +
                       "): type must be add/delete, ids below n, integer weights at least 1");
  }
  k_probe<<<grid_for(N), kBlock, 0, c->stream>>>(c->out_off.as<uint32_t>(), c->out_col.as<uint32_t>(),
                                                 c->wkind ? c->out_w.as<uint32_t>() : nullptr, b.g_row.as<uint32_t>(),
                                                 b.g_col.as<uint32_t>(), bc, b.g_ex.as<uint8_t>(),
                                                 b.g_eoldw.as<uint32_t>(), b.g_pos.as<uint32_t>());
  build_view(c, N, b.g_row.as<uint32_t>(), b.g_ex.as<uint8_t>(), b.g_act.as<uint8_t>(), b.g_exu, b.g_addu,
             b.g_expre, b.g_addpre, b.head, b.t_raw, b.t_list, b.t_gbeg, b.tix, 1);
  k_new_m<<<1, 64, 0, c->stream>>>(bc, b.g_expre.as<uint32_t>(), b.g_addpre.as<uint32_t>(), (uint32_t)c->m, bc);
  SUMINC_HIP(hipGetLastError());
  // supernode sketch: dirty clusters, entry/exit edge counts, internal edge list
  sketch_batch(c, b.g_row.as<uint32_t>(), b.g_col.as<uint32_t>(), b.g_act.as<uint8_t>(), b.g_ex.as<uint8_t>(),
               b.g_w.as<uint32_t>(), bc, N);

  // 3. PageRank reseed -1 on the old out-edges of touched sources
  const bool pr_pull = c->pr_ready && c->prm.pull_threshold_div > 0;
  if (pr_pull) ensure_pull_heavy(c);
  if (c->pr_ready) {
    if (c->pr_prec == SUMINC_PR_FP64) pr_reseed<double>(c, N, -1.0, pr_pull);
    else pr_reseed<float>(c, N, -1.0, pr_pull);
  }

  // 4. SSSP: reset roots + subtree walk on the old graph
  uint32_t walk_r = 0;
  auto reset_round = [&](uint32_t r) {
    k_reset_round<<<push_grid(), kBlock, 0, c->stream>>>(
        c->out_off.as<uint32_t>(), c->out_col.as<uint32_t>(), c->sssp_dp.as<unsigned long long>(),
        b.rbits.as<uint32_t>(), b.rlist.as<uint32_t>(), c->heavy.as<uint2>(), c->heavy.cap / sizeof(uint2) / 2, cnt,
        r);
  };
  if (sssp_repair) {
    const size_t words = ((size_t)n + 31) / 32 + 1;
    if (b.rbits.cap < words * 4 || !b.rbits_clear) {  // otherwise the last repair left it clear
      b.rbits.ensure(words * 4);
      SUMINC_HIP(hipMemsetAsync(b.rbits.p, 0, words * 4, c->stream));
    }
    b.rbits_clear = false;
    b.rlist.ensure((size_t)n * 4 + 4);
    k_sssp_repair_prepare<<<1, 64, 0, c->stream>>>(cnt);
    k_sssp_roots<<<grid_for(N), kBlock, 0, c->stream>>>(b.g_row.as<uint32_t>(), b.g_col.as<uint32_t>(),
                                                        b.g_act.as<uint8_t>(), b.g_w.as<uint32_t>(),
                                                        b.g_ex.as<uint8_t>(), b.g_eoldw.as<uint32_t>(), bc, N,
                                                        c->sssp_dp.as<unsigned long long>(), c->sssp_src,
                                                        b.rbits.as<uint32_t>(), b.rlist.as<uint32_t>(), cnt);
    // rounds are enqueued in groups (the walk's depth of the previous batch + 1
    // first); the transposed grouping below keeps the device busy while the host
    // reads the stop flag of the first group
    for (; walk_r < (uint32_t)c->hint_reset; ++walk_r) reset_round(walk_r);
    SUMINC_HIP(hipGetLastError());
    counters_begin(c);
  }

  // 5a. the transposed groups (CSC view), which neither the walk nor the CSR merge
  // touch; positions of the transposed pairs in the old in-CSR
  k_transpose_group_keys<<<grid_for(N), kBlock, 0, c->stream>>>(b.g_row.as<uint32_t>(), b.g_col.as<uint32_t>(), bc,
                                                                N, b.keys_a.as<uint64_t>(), b.idx_a.as<uint32_t>());
  cub_call(c, c->tmp, [&](void* p, size_t& bytes) {
    return rocprim::radix_sort_pairs(p, bytes, b.keys_a.as<uint64_t>(), b.keys_b.as<uint64_t>(),
                                     b.idx_a.as<uint32_t>(), b.idx_b.as<uint32_t>(), (size_t)N, 0, 64, c->stream);
  });
  k_gather_groups<<<grid_for(N), kBlock, 0, c->stream>>>(b.idx_b.as<uint32_t>(), bc, b.g_row.as<uint32_t>(),
                                                         b.g_col.as<uint32_t>(), b.g_act.as<uint8_t>(),
                                                         b.g_w.as<uint32_t>(), b.g_ex.as<uint8_t>(),
                                                         b.tg_row.as<uint32_t>(), b.tg_col.as<uint32_t>(),
                                                         b.tg_act.as<uint8_t>(), b.tg_w.as<uint32_t>(),
                                                         b.tg_ex.as<uint8_t>());
  build_view(c, N, b.tg_row.as<uint32_t>(), b.tg_ex.as<uint8_t>(), b.tg_act.as<uint8_t>(), b.tg_exu, b.tg_addu,
             b.tg_expre, b.tg_addpre, b.thead, b.tt_raw, b.tt_list, b.tt_gbeg, b.ttix, 2);
  Groups gt;
  gt.countp = bc;
  gt.row = b.tg_row.as<uint32_t>(); gt.col = b.tg_col.as<uint32_t>(); gt.act = b.tg_act.as<uint8_t>();
  gt.w = b.tg_w.as<uint32_t>(); gt.ex = b.tg_ex.as<uint8_t>();
  gt.ex_pre = b.tg_expre.as<uint32_t>(); gt.add_pre = b.tg_addpre.as<uint32_t>();
  gt.tix = b.ttix.as<uint32_t>(); gt.t_gbeg = b.tt_gbeg.as<uint32_t>();
  gt.t_list = b.tt_list.as<uint32_t>(); gt.tcount = bc + 2;
  k_probe<<<grid_for(N), kBlock, 0, c->stream>>>(c->in_off.as<uint32_t>(), c->in_col.as<uint32_t>(), nullptr,
                                                 b.tg_row.as<uint32_t>(), b.tg_col.as<uint32_t>(), bc, nullptr,
                                                 nullptr, b.tg_pos.as<uint32_t>());
  gt.pos = b.tg_pos.as<uint32_t>();

  // 4b. the rest of the walk, a few more rounds at a time, widening
  if (sssp_repair) {
    counters_wait(c);
    int more = 4;
    while (!c->host_counters->done) {
      for (int i = 0; i < more; ++i, ++walk_r) reset_round(walk_r);
      SUMINC_HIP(hipGetLastError());
      read_counters(c);
      more = std::min(more * 2, 16);
    }
    c->hint_reset = std::max(2, (int)c->host_counters->aux[9] + 1);
    c->stats.reset_vertices = (double)c->host_counters->aux[4];  // reset vertex count
  }

  // 5b. topology: the CSR, then the CSC from the transposed groups
  Groups gs;
  gs.countp = bc;
  gs.row = b.g_row.as<uint32_t>(); gs.col = b.g_col.as<uint32_t>(); gs.act = b.g_act.as<uint8_t>();
  gs.w = b.g_w.as<uint32_t>(); gs.ex = b.g_ex.as<uint8_t>();
  gs.ex_pre = b.g_expre.as<uint32_t>(); gs.add_pre = b.g_addpre.as<uint32_t>();
  gs.tix = b.tix.as<uint32_t>(); gs.t_gbeg = b.t_gbeg.as<uint32_t>();
  gs.t_list = b.t_list.as<uint32_t>(); gs.tcount = bc + 1;
  gs.pos = b.g_pos.as<uint32_t>();
  const uint64_t m_old = c->m;
  // the compact positions carried over (most vertices keep theirs; the vertices
  // that gained out-edges get new ones after the CSR merge) instead of renumbered
  const bool carry = pr_pull && pr_positions_carry(c);
  merge_direction(c, c->out_off, c->out_col, c->out_w, b.noff_out, b.ncol_out, b.nw_out, gs, N, cap_m, 0);
  if (carry) pr_positions_append(c);
  merge_direction(c, c->in_off, c->in_col, c->in_w, b.noff_in, b.ncol_in, b.nw_in, gt, N, cap_m, 2,
                  carry ? &c->in_col_c : nullptr, carry ? &b.ncolc_in : nullptr,
                  carry ? c->src_pos.as<uint32_t>() : nullptr);

  // 6. PageRank reseed +1 on the new out-edges; the pull lists of the new CSC are
  // built here (they serve the next query as well)
  if (pr_pull) build_pull_lists(c, cap_m);
  if (c->pr_ready) {
    if (c->pr_prec == SUMINC_PR_FP64) pr_reseed<double>(c, N, +1.0, pr_pull);
    else pr_reseed<float>(c, N, +1.0, pr_pull);
  }
  k_untouch<<<grid_for(N), kBlock, 0, c->stream>>>(b.t_list.as<uint32_t>(), b.tix.as<uint32_t>(),
                                                   b.tt_list.as<uint32_t>(), b.ttix.as<uint32_t>(), bc, N);

  // 7. SSSP: reset, re-pull, relax inserts -> frontier list 0
  if (sssp_reinit) sssp_init(c, c->sssp_src);
  if (sssp_repair) {
    k_reset_apply<<<push_grid(), kBlock, 0, c->stream>>>(b.rlist.as<uint32_t>(), cnt,
                                                         c->sssp_dp.as<unsigned long long>());
    const uint32_t* iw = c->wkind == SUMINC_W_U32 && c->sssp_app != SUMINC_APP_BFS ? c->in_w.as<uint32_t>() : nullptr;
    k_reset_pull<<<push_grid(), kBlock, 0, c->stream>>>(c->in_off.as<uint32_t>(), c->in_col.as<uint32_t>(), iw,
                                                        b.rbits.as<uint32_t>(), b.rlist.as<uint32_t>(), cnt,
                                                        c->sssp_dp.as<unsigned long long>(), c->heavy.as<uint2>(),
                                                        c->sssp_app);
    k_reset_pull_heavy<<<push_grid(), kBlock, 0, c->stream>>>(c->in_off.as<uint32_t>(), c->in_col.as<uint32_t>(), iw,
                                                              b.rbits.as<uint32_t>(), c->heavy.as<uint2>(), cnt,
                                                              c->sssp_dp.as<unsigned long long>(), c->sssp_app);
    k_reset_seed<<<push_grid(), kBlock, 0, c->stream>>>(b.rlist.as<uint32_t>(), b.rbits.as<uint32_t>(),
                                                        c->sssp_dp.as<unsigned long long>(),
                                                        c->sssp_wl0.as<uint32_t>(), c->sssp_flag0.as<uint32_t>(), cnt,
                                                        c->sssp_seeds.as<unsigned long long>());
    k_insert_relax<<<grid_for(N), kBlock, 0, c->stream>>>(b.g_row.as<uint32_t>(), b.g_col.as<uint32_t>(),
                                                          b.g_act.as<uint8_t>(), b.g_w.as<uint32_t>(), bc, N,
                                                          c->wkind != SUMINC_W_U32 || c->sssp_app == SUMINC_APP_BFS,
                                                          c->sssp_app, c->sssp_dp.as<unsigned long long>(),
                                                          c->sssp_wl0.as<uint32_t>(), c->sssp_flag0.as<uint32_t>(),
                                                          c->sssp_seeds.as<unsigned long long>());
  }
  k_batch_end<<<1, 64, 0, c->stream>>>(cnt, bc, sssp_repair ? 1 : 0);
  SUMINC_HIP(hipGetLastError());
  read_counters(c);  // synchronises the stream
  if (sssp_repair) b.rbits_clear = true;
  unsigned hbc[4];
  for (int i = 0; i < 4; ++i) hbc[i] = (unsigned)c->host_counters->bcopy[i];
  sketch_batch_done(c);
  if (c->prof_on && m_old) {
    prof_collect(c, SUMINC_KERNEL_BATCH_COPY, 2, 0, 2);
    suminc_kernel_profile& p = c->prof[SUMINC_KERNEL_BATCH_COPY];
    // algorithmic bytes of the old-edge copy: every old edge's column (and weight)
    // read once and written once, per direction
    const double per_edge = c->wkind ? 16.0 : 8.0;
    p.edges += 2.0 * (double)m_old;
    p.alg_bytes += 2.0 * per_edge * (double)m_old;
  }
  c->m = hbc[3];
  c->graph_version++;
  if (pr_pull) c->pull_heavy_version = c->graph_version;  // built in step 6
  // per-round scratch sized by m may need to grow
  if (c->pr_ready || c->sssp_ready) c->heavy.ensure_slack((c->m / kChunk + c->m / kHeavyDeg + 2) * sizeof(uint4));
  c->stats.batch_groups = (double)hbc[0];
  c->stats.touched_rows = (double)hbc[1];
  c->stats.reseed_edges = (double)c->host_counters->aux[5];  // reseed edges
}

}  // namespace suminc
