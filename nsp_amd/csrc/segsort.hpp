// Sorting of key segments (rows or columns of a CSR) on one device.
//
// Keys are uint64 and unique within a segment (callers pack the sort key in
// the high 32 bits and the entry's position in the low 32 bits, which makes
// every sort stable).  Segments of up to kSegTile keys are sorted by one
// block in LDS; longer ones are cut into kSegTile tiles sorted in LDS, then
// merged pairwise in global memory (merge path: each thread finds its
// output diagonal by binary search and merges 16 keys), doubling the run
// width each pass.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "nsp_internal.hpp"

namespace nsp {

constexpr int kSegTile = 4096;

// Bitonic sort of S[0, n) in LDS by a block of 256 threads (n <= kSegTile);
// S must hold the next power of two of n keys; the padding is set here.
__device__ __forceinline__ void block_sort_keys(uint64_t *S, int n) {
  int n2 = 1;
  while (n2 < n) n2 <<= 1;
  for (int g = n + threadIdx.x; g < n2; g += blockDim.x) S[g] = ~0ull;
  __syncthreads();
  for (int k = 2; k <= n2; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int t = threadIdx.x; t < n2 / 2; t += blockDim.x) {
        const int i = ((t / j) * 2 * j) + (t % j);
        const int q = i + j;
        const bool asc = (i & k) == 0;
        const uint64_t x = S[i], y = S[q];
        if (asc ? (x > y) : (x < y)) {
          S[i] = y;
          S[q] = x;
        }
      }
      __syncthreads();
    }
  }
}

// Sorts keys[off[s], off[s] + len[s]) for every segment s (host vectors;
// segments must not overlap).  Uses the handle's sort workspace for the
// ping-pong buffer and the tile lists.  Synchronises the host once (the tile
// lists are uploaded from host memory).
int sort_segments(Handle *h, const std::vector<int64_t> &off, const std::vector<int64_t> &len, uint64_t *keys);

// Sorts the n CSR rows listed in rows (device int64 row ids, each once) by
// column, stably (ties keep their order), moving the values along.  Keys are
// (column biased to unsigned order, position in the row); rows of 2^32 or
// more entries return NSP_ERR_UNSUPPORTED.  Synchronises the host.
template <typename RowT, typename V>
int sort_listed_rows(Handle *h, const int64_t *rows, int64_t n, const RowT *rpt, int32_t *col, V *val);

}  // namespace nsp
