// Single-pass exclusive scan (decoupled look-back) for gfx950.
//
// Tiles of kScanTile items are claimed in launch order through a ticket, so a
// tile only ever waits on tiles that are already running.  Each tile
// publishes its aggregate, then one wavefront looks back over 64 predecessor
// status words at a time; status word = (value << 2) | flag with flag 1 =
// tile aggregate, 2 = inclusive prefix.  Status words and the ticket must be
// zero before the launch; the grid must hold floor(n / kScanTile) + 1 tiles so
// that out[n] (the total) is written.
#pragma once
#include "device_utils.hpp"

namespace nsp {

constexpr int kScanBS = 256, kScanItems = 8, kScanTile = kScanBS * kScanItems;

__device__ __forceinline__ uint64_t ld_status(const uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_status(uint64_t *p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Look-back of tile `tile` with aggregate agg; returns the tile's exclusive
// prefix in every thread (call from all threads of the block).
__device__ __forceinline__ int64_t scan_lookback(int64_t tile, int64_t agg, uint64_t *status,
                                                 int64_t *s_prefix) {
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    if (tile == 0) {
      if (lane == 0) {
        st_status(&status[0], ((uint64_t)agg << 2) | 2u);
        *s_prefix = 0;
      }
    } else {
      if (lane == 0) st_status(&status[tile], ((uint64_t)agg << 2) | 1u);
      int64_t excl = 0;
      int64_t j = tile - 1;
      while (true) {
        const int64_t idx = j - lane;
        uint64_t s = (idx >= 0) ? ld_status(&status[idx]) : 2u;
        while ((s & 3u) == 0) s = ld_status(&status[idx]);
        const uint64_t incl = __ballot((s & 3u) == 2u);
        const int64_t val = (int64_t)(s >> 2);
        if (incl) {
          const int first = __ffsll((unsigned long long)incl) - 1;
          excl += group_sum<64>(lane <= first ? val : (int64_t)0);
          break;
        }
        excl += group_sum<64>(val);
        j -= 64;
      }
      if (lane == 0) {
        st_status(&status[tile], ((uint64_t)(excl + agg) << 2) | 2u);
        *s_prefix = excl;
      }
    }
  }
  __syncthreads();
  return *s_prefix;
}

// out[i] = sum_{j<i} in[j] for i <= n (out has n+1 entries).
template <typename InT, typename OutT>
__global__ __launch_bounds__(kScanBS) void k_exscan(int64_t n, const InT *__restrict__ in,
                                                   OutT *__restrict__ out, uint64_t *__restrict__ status,
                                                   uint32_t *__restrict__ ticket) {
  __shared__ int64_t s_wsum[kScanBS / 64 + 1];
  __shared__ int64_t s_prefix;
  __shared__ uint32_t s_tile;
  if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
  __syncthreads();
  const int64_t tile = s_tile;
  const int64_t base = tile * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int64_t v[kScanItems];
  int64_t tsum = 0;
#pragma unroll
  for (int q = 0; q < kScanItems; q++) {
    const int64_t i = base + q;
    v[q] = (i < n) ? (int64_t)in[i] : 0;
    tsum += v[q];
  }
  int64_t agg;
  const int64_t texcl = block_excl_scan<kScanBS, int64_t>(tsum, s_wsum, &agg);
  int64_t run = scan_lookback(tile, agg, status, &s_prefix) + texcl;
#pragma unroll
  for (int q = 0; q < kScanItems; q++) {
    const int64_t i = base + q;
    if (i <= n) out[i] = (OutT)run;
    run += v[q];
  }
}

}  // namespace nsp
