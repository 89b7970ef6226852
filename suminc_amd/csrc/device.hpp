// device.hpp -- device-side building blocks: Philox4x32-10, wave64 helpers and
// the block-cooperative adjacency loop used by every edge-parallel kernel.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/suminc.h"

namespace suminc {

// Philox4x32-10 (Salmon et al., SC'11), same stream layout as the synthetic-input
// definition used by the generator and the test oracle.
struct U4 { uint32_t x, y, z, w; };
__device__ __forceinline__ U4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                             uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
    uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
  }
  return {c0, c1, c2, c3};
}
constexpr uint32_t kStreamRmat = 0x524D4154u;
constexpr uint32_t kStreamWeight = 0x57474854u;
constexpr uint32_t kStreamDel = 0x44454C45u;
constexpr uint32_t kStreamIns = 0x494E5345u;

__device__ __forceinline__ uint32_t mulhi32(uint32_t a, uint32_t b) { return __umulhi(a, b); }

// ---- wave64 helpers ----
__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Appends `pred` items of this wave to a global list with one atomic per wave.
// Returns this lane's slot (valid only where pred is true).
__device__ __forceinline__ unsigned long long wave_append(bool pred, unsigned long long* counter) {
  unsigned long long mask = __ballot(pred);
  if (mask == 0) return 0;
  int leader = __ffsll((long long)mask) - 1;
  unsigned long long base = 0;
  if (lane_id() == leader) base = atomicAdd(counter, (unsigned long long)__popcll(mask));
  base = __shfl(base, leader);
  unsigned long long below = mask & ((1ull << lane_id()) - 1ull);
  return base + __popcll(below);
}

// Block-wide exclusive scan of one uint32 per thread (blockDim.x == BLOCK).
template <int BLOCK>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds_warp, uint32_t* total) {
  constexpr int W = BLOCK / 64;
  const int lane = lane_id(), wid = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) lds_warp[wid] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    for (int w = 0; w < W; ++w) { uint32_t t = lds_warp[w]; lds_warp[w] = acc; acc += t; }
    lds_warp[W] = acc;
  }
  __syncthreads();
  uint32_t r = x - v + lds_warp[wid];
  *total = lds_warp[W];
  return r;
}

// Block-cooperative loop over the adjacency of up to BLOCK vertices, one per
// thread: thread t supplies (start_t, len_t); afterwards every thread runs
// fn(slot, k, start_slot + k) for a share of the edges, consecutive threads taking
// consecutive edges (coalesced column reads).  lds must hold BLOCK+1 uint32 for the
// scan plus BLOCK for starts.
template <int BLOCK, class Fn>
__device__ __forceinline__ void block_adjacency(uint32_t start, uint32_t len, uint32_t* s_scan,
                                                uint32_t* s_start, uint32_t* s_warp, Fn&& fn) {
  uint32_t total;
  uint32_t ex = block_exclusive_scan<BLOCK>(len, s_warp, &total);
  s_scan[threadIdx.x] = ex;
  s_start[threadIdx.x] = start;
  if (threadIdx.x == 0) s_scan[BLOCK] = total;
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < total; i += BLOCK) {
    // owner slot: last j with s_scan[j] <= i
    int lo = 0, hi = BLOCK - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (s_scan[mid] <= i) lo = mid; else hi = mid - 1;
    }
    uint32_t k = i - s_scan[lo];
    fn(lo, k, s_start[lo] + k);
  }
  __syncthreads();
}

// Vertices per tile when a grid of workgroups walks the adjacency of a frontier of
// nl vertices: BLOCK (one per thread) once the frontier fills the grid, fewer
// below that, down to one vertex per workgroup.  A short frontier then spreads
// its edges over the whole grid, each lane taking a few (every edge is a chain of
// dependent column, distance and atomic accesses), instead of queueing a tile's
// whole adjacency on a few workgroups.  Tile t holds vertices [t * vt, t * vt + vt);
// thread j < vt takes the j-th.
template <int BLOCK>
__device__ __forceinline__ uint32_t frontier_tile(unsigned long long nl) {
  const unsigned long long per = (nl + gridDim.x - 1) / gridDim.x;
  return per >= (unsigned long long)BLOCK ? (uint32_t)BLOCK : per ? (uint32_t)per : 1u;
}

template <class T>
__device__ __forceinline__ T fabs_t(T x) { return x < T(0) ? -x : x; }

// The traversal apps' path extension on the high word of the packed
// {word, parent} state: SSSP dist + w, BFS hops + 1, SSWP ~min(width, w) with the
// word holding ~width (so that the smaller word is the better value for all
// three).  Returns a value >= 0xFFFFFFFF when the extension does not reach.
__device__ __forceinline__ unsigned long long trav_extend(int app, uint32_t word_u, uint32_t w) {
  if (app == SUMINC_APP_SSWP) {
    const uint32_t width = ~word_u;
    return (unsigned long long)(uint32_t)~(w < width ? w : width);
  }
  return (unsigned long long)word_u + (app == SUMINC_APP_BFS ? 1u : w);
}

}  // namespace suminc
