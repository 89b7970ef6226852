// push.hpp -- the delta scatter shared by the PageRank rounds, the reseed and
// the sketch rounds: persistent blocks stride over tiles of the light list (256
// (row, contribution) entries; the block walks their adjacency cooperatively
// with coalesced column reads) and then over 2048-edge chunks of long rows;
// every edge adds the row's contribution to delta[col] with the hardware
// floating-point atomic.  Runs only in push mode (mode 0).
#pragma once
#include "ctx.hpp"
#include "device.hpp"

namespace suminc {

// Rows come from (off, col).  kSketch: a row of a clustered vertex (type[v] != 0,
// i.e. an exit of the sketch) skips the edges that stay inside its cluster.
template <class T, bool kSketch = false>
__global__ void __launch_bounds__(kBlock) k_pr_push(const uint32_t* __restrict__ off, const uint32_t* __restrict__ col,
                                                    T* __restrict__ delta, const uint32_t* __restrict__ wl_v,
                                                    const T* __restrict__ wl_c, const uint4* __restrict__ heavy,
                                                    const Counters* cnt, const uint8_t* __restrict__ type = nullptr,
                                                    const uint32_t* __restrict__ cid = nullptr) {
  __shared__ uint32_t s_scan[kBlock + 1];
  __shared__ uint32_t s_start[kBlock];
  __shared__ uint32_t s_warp[kBlock / 64 + 1];
  __shared__ T s_c[kBlock];
  __shared__ uint32_t s_k[kBlock];  // cluster of a clustered row, INF32 for a Single's row
  if (cnt->done || cnt->mode != 0) return;
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
      if constexpr (kSketch) s_k[threadIdx.x] = type[v] != 0 ? cid[v] : INF32;
    }
    block_adjacency<kBlock>(start, len, s_scan, s_start, s_warp, [&](int slot, uint32_t, uint32_t e) {
      const uint32_t x = col[e];
      if constexpr (kSketch) {
        if (s_k[slot] != INF32 && cid[x] == s_k[slot]) return;
      }
      unsafeAtomicAdd(&delta[x], s_c[slot]);
    });
  }
  for (unsigned long long h = blockIdx.x; h < nh; h += gridDim.x) {
    uint4 hv = heavy[h];
    T c;
    if constexpr (sizeof(T) == 4) c = __uint_as_float(hv.z);
    else c = __longlong_as_double((long long)(((unsigned long long)hv.w << 32) | hv.z));
    uint32_t kc = INF32;
    if constexpr (kSketch) kc = type[hv.x] != 0 ? cid[hv.x] : INF32;
    uint32_t b = off[hv.x] + hv.y * kChunk;
    uint32_t e = min(off[hv.x + 1], b + kChunk);
    for (uint32_t k = b + threadIdx.x; k < e; k += kBlock) {
      const uint32_t x = col[k];
      if (kSketch && kc != INF32 && cid[x] == kc) continue;
      unsafeAtomicAdd(&delta[x], c);
    }
  }
}

}  // namespace suminc
