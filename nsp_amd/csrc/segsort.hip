// Segment sort (see segsort.hpp): LDS tile sort plus merge-path passes.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "segsort.hpp"

namespace nsp {

// One unit of work: kSegTile keys or fewer.  Tile sorts read and write
// src[a, a + n).  Merges read runs A = src[a, a + na) and B = src[a + na,
// a + na + nb) and write outputs [d0, d0 + n) of their merge to dst[b + d0
// ...]; nb = 0 copies.
struct SortTile {
  int64_t a, b, na, nb, d0;
  int64_t n;
};

__global__ __launch_bounds__(256) void k_seg_tile_sort(const SortTile *__restrict__ tiles, uint64_t *keys) {
  __shared__ uint64_t S[kSegTile];
  const SortTile T = tiles[blockIdx.x];
  const int n = (int)T.n;
  for (int g = threadIdx.x; g < n; g += 256) S[g] = keys[T.a + g];
  block_sort_keys(S, n);
  for (int g = threadIdx.x; g < n; g += 256) keys[T.a + g] = S[g];
}

constexpr int kMergePerThread = kSegTile / 256;

__global__ __launch_bounds__(256) void k_seg_merge(const SortTile *__restrict__ tiles,
                                                   const uint64_t *__restrict__ src, uint64_t *__restrict__ dst) {
  const SortTile T = tiles[blockIdx.x];
  const int64_t d = T.d0 + (int64_t)threadIdx.x * kMergePerThread;
  const int64_t dend = std::min<int64_t>(T.d0 + T.n, d + kMergePerThread);
  if (d >= dend) return;
  const uint64_t *A = src + T.a;
  const uint64_t *B = A + T.na;
  // merge path: i = number of A keys among the d smallest of A U B
  int64_t lo = std::max<int64_t>(0, d - T.nb), hi = std::min<int64_t>(d, T.na);
  while (lo < hi) {
    const int64_t i = (lo + hi) >> 1;
    if (A[i] < B[d - i - 1]) lo = i + 1;
    else hi = i;
  }
  int64_t i = lo, j = d - lo;
  uint64_t *out = dst + T.b;
  for (int64_t o = d; o < dend; o++) {
    const bool take_a = j >= T.nb || (i < T.na && A[i] < B[j]);
    out[o] = take_a ? A[i++] : B[j++];
  }
}

static int segw_reserve(Handle *h, size_t bytes) {
  if (bytes <= h->segw.bytes) return NSP_OK;
  if (h->segw.base) NSP_HIP(h, hipFree(h->segw.base));
  h->segw.base = nullptr;
  h->segw.bytes = 0;
  if (hipMalloc(&h->segw.base, bytes + bytes / 4) != hipSuccess) {
    (void)hipGetLastError();
    return set_error(h, NSP_ERR_OOM, "segment sort workspace allocation failed");
  }
  h->segw.bytes = bytes + bytes / 4;
  return NSP_OK;
}

int sort_segments(Handle *h, const std::vector<int64_t> &off, const std::vector<int64_t> &len, uint64_t *keys) {
  // tiles of the first pass (every segment), and the long segments
  std::vector<SortTile> first;
  std::vector<size_t> longs;
  int64_t total_long = 0, maxlen = 0;
  for (size_t s = 0; s < off.size(); s++) {
    if (len[s] <= 1) continue;
    for (int64_t p = 0; p < len[s]; p += kSegTile)
      first.push_back(SortTile{off[s] + p, 0, 0, 0, 0, std::min<int64_t>(kSegTile, len[s] - p)});
    if (len[s] > kSegTile) {
      longs.push_back(s);
      total_long += len[s];
      maxlen = std::max(maxlen, len[s]);
    }
  }
  if (first.empty()) return NSP_OK;
  int passes = 0;
  for (int64_t w = kSegTile; w < maxlen; w *= 2) passes++;
  // merge tiles: pass k merges runs of width kSegTile << k; every long
  // segment takes part in every pass (a lone run is copied) so that all of
  // them alternate buffers together; an odd pass count copies back.
  std::vector<SortTile> merges;
  std::vector<int64_t> soff(longs.size());
  {
    int64_t o = 0;
    for (size_t q = 0; q < longs.size(); q++) {
      soff[q] = o;
      o += len[longs[q]];
    }
  }
  const int total_passes = passes + (passes & 1);
  for (int k = 0; k < total_passes; k++) {
    const int64_t w = (k < passes) ? ((int64_t)kSegTile << k) : INT64_MAX / 4;
    const bool from_keys = (k % 2) == 0;
    for (size_t q = 0; q < longs.size(); q++) {
      const int64_t L = len[longs[q]];
      const int64_t base_src = from_keys ? off[longs[q]] : soff[q];
      const int64_t base_dst = from_keys ? soff[q] : off[longs[q]];
      for (int64_t p = 0; p < L; p += 2 * w) {
        const int64_t na = std::min(w, L - p), nb = std::max<int64_t>(0, std::min(w, L - p - na));
        for (int64_t d0 = 0; d0 < na + nb; d0 += kSegTile)
          merges.push_back(SortTile{base_src + p, base_dst + p, na, nb, d0, std::min<int64_t>(kSegTile, na + nb - d0)});
        if (w >= L) break;
      }
    }
  }
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t tb = al((first.size() + merges.size()) * sizeof(SortTile));
  NSP_TRY(segw_reserve(h, tb + (size_t)total_long * sizeof(uint64_t)));
  SortTile *d_tiles = (SortTile *)h->segw.base;
  uint64_t *scratch = (uint64_t *)((char *)h->segw.base + tb);
  hipStream_t st = h->stream;
  std::vector<SortTile> all(first);
  all.insert(all.end(), merges.begin(), merges.end());
  NSP_HIP(h, hipMemcpyAsync(d_tiles, all.data(), all.size() * sizeof(SortTile), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_seg_tile_sort, dim3((unsigned)first.size()), dim3(256), 0, st, d_tiles, keys);
  NSP_HIP(h, hipGetLastError());
  if (total_passes > 0) {
    const size_t per = merges.size() / total_passes;
    for (int k = 0; k < total_passes; k++) {
      const bool from_keys = (k % 2) == 0;
      hipLaunchKernelGGL(k_seg_merge, dim3((unsigned)per), dim3(256), 0, st, d_tiles + first.size() + k * per,
                         from_keys ? (const uint64_t *)keys : (const uint64_t *)scratch, from_keys ? scratch : keys);
      NSP_HIP(h, hipGetLastError());
    }
  }
  // the tile list lives in host memory until the copy has been consumed
  NSP_HIP(h, hipStreamSynchronize(st));
  return NSP_OK;
}

// Extents of the listed rows: ext[2k] = row_ptr[row], ext[2k+1] = length.
template <typename RowT>
__global__ void k_rows_extents(const int64_t *__restrict__ rows, int64_t n, const RowT *__restrict__ rpt,
                               int64_t *__restrict__ ext) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = rows[k];
    ext[2 * k] = (int64_t)rpt[r];
    ext[2 * k + 1] = (int64_t)rpt[r + 1] - (int64_t)rpt[r];
  }
}

// seg[3k..3k+2] = (row start in col/val, length, start in keys)
template <typename V>
__global__ void k_rows_keys(const int64_t *__restrict__ seg, int64_t nseg, const int32_t *__restrict__ col,
                            const V *__restrict__ val, uint64_t *__restrict__ keys, V *__restrict__ oldv) {
  for (int64_t s = blockIdx.x; s < nseg; s += gridDim.x) {
    const int64_t off = seg[3 * s], len = seg[3 * s + 1], ko = seg[3 * s + 2];
    for (int64_t e = threadIdx.x; e < len; e += blockDim.x) {
      keys[ko + e] = ((uint64_t)((uint32_t)col[off + e] ^ 0x80000000u) << 32) | (uint64_t)e;
      oldv[ko + e] = val[off + e];
    }
  }
}

template <typename V>
__global__ void k_rows_gather(const int64_t *__restrict__ seg, int64_t nseg, const uint64_t *__restrict__ keys,
                              const V *__restrict__ oldv, int32_t *__restrict__ col, V *__restrict__ val) {
  for (int64_t s = blockIdx.x; s < nseg; s += gridDim.x) {
    const int64_t off = seg[3 * s], len = seg[3 * s + 1], ko = seg[3 * s + 2];
    for (int64_t e = threadIdx.x; e < len; e += blockDim.x) {
      const uint64_t k = keys[ko + e];
      col[off + e] = (int32_t)((uint32_t)(k >> 32) ^ 0x80000000u);
      val[off + e] = oldv[ko + (uint32_t)k];
    }
  }
}

static int rsort_reserve(Handle *h, size_t bytes) {
  if (bytes <= h->rsort.bytes) return NSP_OK;
  if (h->rsort.base) NSP_HIP(h, hipFree(h->rsort.base));
  h->rsort.base = nullptr;
  h->rsort.bytes = 0;
  if (hipMalloc(&h->rsort.base, bytes + bytes / 4) != hipSuccess) {
    (void)hipGetLastError();
    return set_error(h, NSP_ERR_OOM, "row sort workspace allocation failed");
  }
  h->rsort.bytes = bytes + bytes / 4;
  return NSP_OK;
}

template <typename RowT, typename V>
int sort_listed_rows(Handle *h, const int64_t *rows, int64_t n, const RowT *rpt, int32_t *col, V *val) {
  if (n <= 0) return NSP_OK;
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  hipStream_t st = h->stream;
  const int g = (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, (int64_t)h->num_cu * 8));
  NSP_TRY(rsort_reserve(h, al((size_t)n * 16)));
  int64_t *d_ext = (int64_t *)h->rsort.base;
  hipLaunchKernelGGL((k_rows_extents<RowT>), dim3(g), dim3(256), 0, st, rows, n, rpt, d_ext);
  NSP_HIP(h, hipGetLastError());
  std::vector<int64_t> ext((size_t)(2 * n));
  NSP_HIP(h, hipMemcpyAsync(ext.data(), d_ext, (size_t)n * 16, hipMemcpyDeviceToHost, st));
  NSP_HIP(h, hipStreamSynchronize(st));
  std::vector<int64_t> seg((size_t)(3 * n)), koff((size_t)n), klen((size_t)n);
  int64_t total = 0;
  for (int64_t k = 0; k < n; k++) {
    const int64_t len = ext[2 * k + 1];
    if (len >= (int64_t(1) << 32)) return set_error(h, NSP_ERR_UNSUPPORTED, "row sort: a row of 2^32 or more entries");
    seg[3 * k] = ext[2 * k];
    seg[3 * k + 1] = len;
    seg[3 * k + 2] = total;
    koff[k] = total;
    klen[k] = len;
    total += len;
  }
  const size_t b_seg = al((size_t)n * 24), b_keys = al((size_t)total * 8), b_vals = al((size_t)total * sizeof(V));
  NSP_TRY(rsort_reserve(h, b_seg + b_keys + b_vals));
  int64_t *d_seg = (int64_t *)h->rsort.base;
  uint64_t *keys = (uint64_t *)((char *)h->rsort.base + b_seg);
  V *oldv = (V *)((char *)h->rsort.base + b_seg + b_keys);
  NSP_HIP(h, hipMemcpyAsync(d_seg, seg.data(), (size_t)n * 24, hipMemcpyHostToDevice, st));
  const int gs = (int)std::min<int64_t>(n, (int64_t)h->num_cu * 8);
  hipLaunchKernelGGL((k_rows_keys<V>), dim3(gs), dim3(256), 0, st, d_seg, n, col, (const V *)val, keys, oldv);
  NSP_HIP(h, hipGetLastError());
  NSP_TRY(sort_segments(h, koff, klen, keys));
  hipLaunchKernelGGL((k_rows_gather<V>), dim3(gs), dim3(256), 0, st, d_seg, n, keys, oldv, col, val);
  NSP_HIP(h, hipGetLastError());
  NSP_HIP(h, hipStreamSynchronize(st));  // seg lives in host memory until its copy is consumed
  return NSP_OK;
}

template int sort_listed_rows<int32_t, double>(Handle *, const int64_t *, int64_t, const int32_t *, int32_t *,
                                               double *);
template int sort_listed_rows<int32_t, float>(Handle *, const int64_t *, int64_t, const int32_t *, int32_t *,
                                              float *);
template int sort_listed_rows<int64_t, double>(Handle *, const int64_t *, int64_t, const int64_t *, int32_t *,
                                               double *);
template int sort_listed_rows<int64_t, float>(Handle *, const int64_t *, int64_t, const int64_t *, int32_t *,
                                              float *);

}  // namespace nsp
