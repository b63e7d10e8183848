// Local sparse product C = A*B for gfx950: row binning, symbolic hash count,
// wave64 decoupled-look-back scan into C's row pointer, numeric hash
// accumulate with an in-row sort.
//
// Reference path replaced (nsp/src/nsp.cu): nsp_set_intprod_num 306-320,
// nsp_set_bin / nsp_set_row_perm 325-411, symbolic kernels 478-915 and their
// launcher 921-1130, the CUB exclusive scan 937/1124, numeric binning
// 1137-1300, numeric kernels 1308-2151 and their launcher 2178-2440, and the
// driver nsp_spgemm_kernel_hash 2448-2488.
//
// Layout of the work: every hash kernel walks a list of rows with a
// grid-stride loop, so a launch of any grid size covers it.  The list is a
// range of bins of the binned row list, or all rows in natural order (direct
// mode).  The host reads the bin counts back twice per product (after
// binning, and with nnz(C) after the scan) and launches only what is needed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "device_utils.hpp"
#include "nsp_internal.hpp"
#include "row_products.hpp"
#include "scan.hpp"
#include "spgemm_internal.hpp"

namespace nsp {

__host__ __device__ inline int sym_bin_of(int64_t ip) {
  int b = 0;
  while (b < kBins - 1 && ip > kSymBound[b]) b++;
  return b;
}
__host__ __device__ inline int num_bin_of(int64_t nz) {
  int b = 0;
  while (b < kBins - 1 && nz > kNumBound[b]) b++;
  return b;
}

constexpr uint8_t kNoBin = 255;

// ---------------------------------------------------------------------------
// init: zero the per-call counters and the scan's tile status words.
// (A fill kernel rather than hipMemsetAsync so the sequence stays valid inside
// a captured graph.)
__global__ void k_init(Info *info, uint64_t *tile_status, int64_t ntiles, int64_t cap) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int64_t *w = reinterpret_cast<int64_t *>(info);
  constexpr int nw = sizeof(Info) / sizeof(int64_t);
  constexpr int wcap = (int)(offsetof(Info, cap) / sizeof(int64_t));
  for (int64_t i = t; i < nw; i += (int64_t)gridDim.x * blockDim.x) w[i] = (i == wcap) ? cap : 0;
  for (int64_t i = t; i < ntiles; i += (int64_t)gridDim.x * blockDim.x) tile_status[i] = 0;
}

// ---------------------------------------------------------------------------
// Row info in two balanced passes.
//
// k_extents (entry-parallel, every thread takes U entries per round): for
// every A entry (i,k) the start and length of the B row it selects (ab_b0,
// ab_len), which the hash passes then read with one independent load
// instead of the dependent A.col -> B.row_ptr gather.  The row of an entry
// is not needed, so the work is spread evenly whatever the row lengths
// (power-law matrices put most entries in a few rows).
//
// k_rowinfo (one lane per row, 8 loads in flight; rows past kLongRow
// entries go to k_rowinfo_long): intermediate products per row as the sum of
// its entries' ab_len (nsp.cu:306-320), the symbolic bin, per-bin histogram,
// nprod, max, and an int64 copy of A's row pointer.
//
// Batched products (a_row_off != NULL): A and B hold `count` matrices as
// concatenated rows with matrix-local columns; an entry of matrix b selects
// B row b_row_off[b] + k.  A block finds the matrices its entries belong to
// with one binary search over the matrices' first entries.  This is the only
// place the batch structure enters: every later pass works on the
// concatenated rows.
constexpr int kExtU = 4;

// Last b in [from, count) with arpt[a_row_off[b]] <= p, given that b = from
// qualifies: gallop forward, then bisect (a block's tiles move forward
// through the batch, so this is usually a step or two).
template <typename RowT>
__device__ __forceinline__ int64_t matrix_of_entry(const RowT *__restrict__ arpt, const int64_t *__restrict__ a_row_off,
                                                   int64_t count, int64_t from, int64_t p) {
  int64_t lo = from, hi = from + 1, step = 1;
  while (hi < count && (int64_t)arpt[a_row_off[hi]] <= p) {
    lo = hi;
    step <<= 1;
    hi = lo + step;
  }
  if (hi > count) hi = count;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int64_t)arpt[a_row_off[mid]] <= p) lo = mid; else hi = mid;
  }
  return lo;
}

template <typename RowT>
__global__ __launch_bounds__(256) void k_extents(int64_t nnzA, const RowT *__restrict__ arpt,
                                                 const int32_t *__restrict__ acol,
                                                 const RowT *__restrict__ brpt,
                                                 const int64_t *__restrict__ a_row_off,
                                                 const int64_t *__restrict__ b_row_off, int64_t count,
                                                 int64_t *__restrict__ ab_b0, int32_t *__restrict__ ab_len,
                                                 unsigned long long *__restrict__ partial) {
  __shared__ int64_t s_mat[2];
  unsigned long long lsum = 0;  // this thread's intermediate products (for `partial`)
  const int t = threadIdx.x;
  constexpr int64_t TILE = 256 * kExtU;
  // each block takes a contiguous run of tiles
  const int64_t ntiles = (nnzA + TILE - 1) / TILE;
  const int64_t t0 = blockIdx.x * ntiles / gridDim.x, t1 = (blockIdx.x + 1) * ntiles / gridDim.x;
  int64_t from = 0;  // matrix search start (batched)
  for (int64_t tile = t0; tile < t1; tile++) {
    const int64_t tb = tile * TILE;
    if (a_row_off) {
      // matrices of the tile's first and last entries
      if (t < 2) {
        const int64_t p = t == 0 ? tb : (tb + TILE < nnzA ? tb + TILE : nnzA) - 1;
        s_mat[t] = matrix_of_entry(arpt, a_row_off, count, from, p);
      }
      __syncthreads();
    }
    int64_t p[kExtU], k[kExtU];
    bool v[kExtU];
#pragma unroll
    for (int u = 0; u < kExtU; u++) {
      p[u] = tb + u * 256 + t;
      v[u] = p[u] < nnzA;
      k[u] = v[u] ? acol[p[u]] : 0;
    }
    if (a_row_off) {
      const int64_t m0 = s_mat[0], m1 = s_mat[1];
#pragma unroll
      for (int u = 0; u < kExtU; u++) {
        int64_t lo = m0, hi = m1 + 1;
        while (hi - lo > 1) {
          const int64_t mid = (lo + hi) >> 1;
          if ((int64_t)arpt[a_row_off[mid]] <= p[u]) lo = mid; else hi = mid;
        }
        k[u] += b_row_off[lo];
      }
      from = m1;
    }
    int64_t b0[kExtU], b1[kExtU];
#pragma unroll
    for (int u = 0; u < kExtU; u++) {
      b0[u] = v[u] ? (int64_t)brpt[k[u]] : 0;
      b1[u] = v[u] ? (int64_t)brpt[k[u] + 1] : 0;
    }
#pragma unroll
    for (int u = 0; u < kExtU; u++) {
      if (v[u]) {
        ab_b0[p[u]] = b0[u];
        ab_len[p[u]] = (int32_t)(b1[u] - b0[u]);
        lsum += (unsigned long long)(b1[u] - b0[u]);
      }
    }
    if (a_row_off) __syncthreads();  // s_mat is rewritten by the next tile
  }
  if (partial) {  // the block's intermediate products, one slot per block (no atomics)
    __shared__ unsigned long long s_red[4];
    lsum = group_sum<64>(lsum);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = lsum;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = s_red[0] + s_red[1] + s_red[2] + s_red[3];
  }
}

// rows with more entries than this are summed by a whole block
constexpr int64_t kLongRow = 1024;

// Finishes a row whose intermediate-product count is s.
__device__ __forceinline__ void finish_row(int64_t i, unsigned long long s, int64_t p0, int64_t *__restrict__ intprod,
                                           int64_t *__restrict__ arpt64, int32_t *__restrict__ row_nz,
                                           uint8_t *__restrict__ cls, unsigned long long *hist) {
  intprod[i] = (int64_t)s;
  arpt64[i] = p0;
  if (s == 0) {
    cls[i] = kNoBin;
    row_nz[i] = 0;
  } else {
    const int b = sym_bin_of((int64_t)s);
    cls[i] = (uint8_t)b;
    atomicAdd(&hist[b], 1ull);
  }
}

// Entries of a block's 256 rows staged through LDS per round of k_rowinfo.
constexpr int kRowInfoStage = 4096;

template <typename RowT>
__global__ __launch_bounds__(256) void k_rowinfo(int64_t MA, const RowT *__restrict__ arpt,
                                                 const int32_t *__restrict__ ab_len,
                                                 int64_t *__restrict__ arpt64, int64_t *__restrict__ intprod,
                                                 int32_t *__restrict__ row_nz, uint8_t *__restrict__ cls,
                                                 int32_t *__restrict__ long_rows, Info *info) {
  __shared__ unsigned long long hist[kBins];
  __shared__ unsigned long long s_nprod, s_max;
  __shared__ int64_t s_e[2];
  __shared__ int32_t stage[kRowInfoStage];
  if (threadIdx.x < kBins) hist[threadIdx.x] = 0;
  if (threadIdx.x == 0) { s_nprod = 0; s_max = 0; }
  __syncthreads();
  unsigned long long my_sum = 0, my_max = 0;
  // One lane per row.  The block's 256 rows own one contiguous run of
  // entries; while it spans at most four stages of kRowInfoStage entries it
  // is read coalesced into LDS, stage by stage, and every lane sums its
  // row's part of each stage there (rows one lane apart are ~row-length
  // entries apart, so direct per-lane loads would touch a cache line each).
  // Longer runs (a long row in the block) fall back to direct loads.
  for (int64_t i0 = blockIdx.x * (int64_t)blockDim.x; i0 < MA; i0 += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = i0 + threadIdx.x;
    const int64_t ilast = (i0 + (int64_t)blockDim.x < MA ? i0 + (int64_t)blockDim.x : MA) - 1;
    const bool valid = i < MA;
    int64_t p0 = 0, p1 = 0;
    if (valid) {
      p0 = (int64_t)arpt[i];
      p1 = (int64_t)arpt[i + 1];
    }
    if (i == i0) s_e[0] = p0;
    if (i == ilast) s_e[1] = p1;
    const bool is_long = p1 - p0 > kLongRow;
    if (is_long)  // k_rowinfo_long finishes it
      long_rows[atomicAdd((unsigned long long *)&info->nlong, 1ull)] = (int32_t)i;
    const bool mine = valid && !is_long;
    __syncthreads();
    const int64_t e0 = s_e[0], e1 = s_e[1];
    __syncthreads();  // s_e is rewritten by the next round
    int64_t s = 0;
    if (e1 - e0 <= 4 * (int64_t)kRowInfoStage) {
      for (int64_t c0 = e0; c0 < e1; c0 += kRowInfoStage) {
        const int64_t c1 = c0 + kRowInfoStage < e1 ? c0 + kRowInfoStage : e1;
        constexpr int PER = kRowInfoStage / 256;
        int32_t l[PER];
#pragma unroll
        for (int u = 0; u < PER; u++) {
          const int64_t p = c0 + u * 256 + threadIdx.x;
          l[u] = p < c1 ? ab_len[p] : 0;
        }
#pragma unroll
        for (int u = 0; u < PER; u++) stage[u * 256 + threadIdx.x] = l[u];
        __syncthreads();
        if (mine) {
          const int64_t a = p0 > c0 ? p0 : c0, b = p1 < c1 ? p1 : c1;
          for (int64_t p = a; p < b; p++) s += stage[p - c0];
        }
        __syncthreads();
      }
    } else if (mine) {
      int64_t p = p0;
      for (; p + 8 <= p1; p += 8) {
        int32_t l[8];
#pragma unroll
        for (int u = 0; u < 8; u++) l[u] = ab_len[p + u];
#pragma unroll
        for (int u = 0; u < 8; u++) s += l[u];
      }
      for (; p < p1; p++) s += ab_len[p];
    }
    if (mine) {
      finish_row(i, (unsigned long long)s, p0, intprod, arpt64, row_nz, cls, hist);
      my_sum += (unsigned long long)s;
      my_max = my_max > (unsigned long long)s ? my_max : (unsigned long long)s;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) arpt64[MA] = (int64_t)arpt[MA];
  my_sum = group_sum<64>(my_sum);
  my_max = group_max<64>(my_max);
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&s_nprod, my_sum);
    atomicMax(&s_max, my_max);
  }
  __syncthreads();
  if (threadIdx.x < kBins && hist[threadIdx.x])
    atomicAdd((unsigned long long *)&info->count_sym[threadIdx.x], hist[threadIdx.x]);
  if (threadIdx.x == 0) {
    atomicAdd((unsigned long long *)&info->nprod, s_nprod);
    atomicMax((unsigned long long *)&info->max_intprod, s_max);
  }
}

// Rows listed by k_rowinfo as longer than kLongRow entries: one block per
// row sums its entries' B row lengths (8 independent loads per thread per
// round), then finishes it like k_rowinfo.
template <typename RowT>
__global__ __launch_bounds__(256) void k_rowinfo_long(const RowT *__restrict__ arpt,
                                                      const int32_t *__restrict__ ab_len,
                                                      const int32_t *__restrict__ long_rows,
                                                      int64_t *__restrict__ arpt64, int64_t *__restrict__ intprod,
                                                      int32_t *__restrict__ row_nz, uint8_t *__restrict__ cls,
                                                      Info *info) {
  __shared__ unsigned long long hist[kBins];
  __shared__ unsigned long long s_red[4];
  const int64_t nlong = info->nlong;
  if (threadIdx.x < kBins) hist[threadIdx.x] = 0;
  __syncthreads();
  unsigned long long my_sum = 0, my_max = 0;
  for (int64_t q = blockIdx.x; q < nlong; q += gridDim.x) {
    const int64_t i = long_rows[q];
    const int64_t p0 = (int64_t)arpt[i], p1 = (int64_t)arpt[i + 1];
    int64_t s = 0;
    for (int64_t pb = p0; pb < p1; pb += 8 * 256) {
      int32_t l[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int64_t p = pb + u * 256 + threadIdx.x;
        l[u] = p < p1 ? ab_len[p] : 0;
      }
#pragma unroll
      for (int u = 0; u < 8; u++) s += l[u];
    }
    s = group_sum<64>(s);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = (unsigned long long)s;
    __syncthreads();
    if (threadIdx.x == 0) {
      const unsigned long long t = s_red[0] + s_red[1] + s_red[2] + s_red[3];
      finish_row(i, t, p0, intprod, arpt64, row_nz, cls, hist);
      my_sum += t;
      my_max = my_max > t ? my_max : t;
    }
    __syncthreads();
  }
  if (threadIdx.x < kBins && hist[threadIdx.x])
    atomicAdd((unsigned long long *)&info->count_sym[threadIdx.x], hist[threadIdx.x]);
  if (threadIdx.x == 0 && my_sum) {
    atomicAdd((unsigned long long *)&info->nprod, my_sum);
    atomicMax((unsigned long long *)&info->max_intprod, my_max);
  }
}

// The part of the row info a one-call product needs when it reuses the
// previous call's slot size: A's row pointer widened to int64, and the total
// of the intermediate products from k_extents' per-block sums (block 0 adds
// them up; no per-row sums, bins or long-row list).
template <typename RowT>
__global__ __launch_bounds__(256) void k_rowinfo_light(int64_t MA, const RowT *__restrict__ arpt,
                                                       const unsigned long long *__restrict__ partial, int npart,
                                                       int64_t *__restrict__ arpt64, Info *info) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = t; i <= MA; i += stride) arpt64[i] = (int64_t)arpt[i];
  if (blockIdx.x == 0) {
    unsigned long long s = 0;
    for (int q = threadIdx.x; q < npart; q += blockDim.x) s += partial[q];
    __shared__ unsigned long long s_red[4];
    s = group_sum<64>(s);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) info->nprod = (int64_t)(s_red[0] + s_red[1] + s_red[2] + s_red[3]);
  }
}

// ---------------------------------------------------------------------------
// Scatter rows into per-bin lists, stably within a block's chunk of rows:
// pass 1 counts the chunk's rows per bin and claims one range per bin with a
// single global atomic; pass 2 places rows in order using wavefront ballots.
__global__ __launch_bounds__(256) void k_scatter(int64_t M, int64_t chunk,
                                                 const uint8_t *__restrict__ cls,
                                                 const int64_t *__restrict__ counts,
                                                 int64_t *__restrict__ cursor,
                                                 int32_t *__restrict__ perm) {
  __shared__ int64_t off[kBins];
  __shared__ int64_t run[kBins];
  __shared__ unsigned int hist[kBins];
  __shared__ int wcnt[4][kBins];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t lt = (1ull << lane) - 1;
  const int64_t c0 = blockIdx.x * chunk;
  const int64_t c1 = c0 + chunk < M ? c0 + chunk : M;
  if (threadIdx.x < kBins) hist[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t i = c0 + threadIdx.x; i < c1; i += 256) {
    const int c = cls[i];
    if (c != kNoBin) atomicAdd(&hist[c], 1u);
  }
  __syncthreads();
  if (threadIdx.x < kBins) {
    int64_t s = 0;
    for (int b = 0; b < (int)threadIdx.x; b++) s += counts[b];
    off[threadIdx.x] = s;
    const unsigned int hcount = hist[threadIdx.x];
    run[threadIdx.x] =
        hcount ? (int64_t)atomicAdd((unsigned long long *)&cursor[threadIdx.x], (unsigned long long)hcount) : 0;
  }
  __syncthreads();
  for (int64_t i0 = c0; i0 < c1; i0 += 256) {
    const int64_t i = i0 + threadIdx.x;
    const int c = (i < c1) ? (int)cls[i] : (int)kNoBin;
    int myrank = 0;
#pragma unroll
    for (int b = 0; b < kBins; b++) {
      const uint64_t m = __ballot(c == b);
      if (lane == 0) wcnt[w][b] = __popcll(m);
      if (c == b) myrank = __popcll(m & lt);
    }
    __syncthreads();
    if (c != kNoBin) {
      int64_t pos = run[c] + myrank;
      for (int q = 0; q < w; q++) pos += wcnt[q][c];
      perm[off[c] + pos] = (int32_t)i;
    }
    __syncthreads();
    if (threadIdx.x < kBins) run[threadIdx.x] += wcnt[0][threadIdx.x] + wcnt[1][threadIdx.x] +
                                                  wcnt[2][threadIdx.x] + wcnt[3][threadIdx.x];
    __syncthreads();
  }
}

// Rows handled by one launch: bins [lo, hi] of the binned row list, or, in
// direct mode (direct_rows >= 0, row list NULL), rows [0, direct_rows) in
// natural order -- the reference's DIRECT mode (nsp/src/nsp.cu:443-452),
// used when one bin's table fits every row.
__device__ __forceinline__ void bin_range(const int64_t *counts, int lo, int hi, int64_t direct_rows,
                                          int64_t &count, int64_t &off) {
  if (direct_rows >= 0) {
    count = direct_rows;
    off = 0;
    return;
  }
  int64_t s = 0;
  for (int q = 0; q < lo; q++) s += counts[q];
  off = s;
  count = 0;
  for (int q = lo; q <= hi; q++) count += counts[q];
}

// Clears GS lanes' share of an LDS table with 16-byte stores.
template <int GS, int N, typename T>
__device__ __forceinline__ void clear_lds(T *p, int glane, int32_t fill) {
  static_assert((N * sizeof(T)) % 16 == 0, "table size");
  int4 *q = reinterpret_cast<int4 *>(p);
  constexpr int NV = (int)(N * sizeof(T) / 16);
  const int4 f = make_int4(fill, fill, fill, fill);
  for (int s = glane; s < NV; s += GS) q[s] = f;
}

// ---------------------------------------------------------------------------
// Symbolic, GS lanes per row (sub-wave groups or one wave), SH-slot LDS key
// table per row; products enumerated with R loads in flight per lane.
// Groups of 8 and 16 lanes with tables of up to 256 slots have LDS for five
// or more waves per SIMD; the register budget is held to that (96 VGPRs).
template <int GS, int SH, int R, bool O32>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(GS <= 16 && SH <= 256 ? 5 : 1)))
void k_sym_group(int bin_lo, int bin_hi, int64_t direct_rows,
                                                   const Info *__restrict__ info,
                                                   const int32_t *__restrict__ perm,
                                                   const int64_t *__restrict__ arpt,
                                                   const int64_t *__restrict__ ab_b0,
                                                   const int32_t *__restrict__ ab_len,
                                                   const int32_t *__restrict__ bcol,
                                                   int32_t *__restrict__ row_nz) {
  constexpr int RPB = 256 / GS;
  constexpr int LOG = Log2<SH>::value;
  constexpr int SHD = SH + (GS < 16 ? 16 : GS);  // + one dummy slot per lane
  __shared__ int32_t keys[RPB][SHD];
  __shared__ EntryScratch<GS, float> ent[RPB];
  int64_t count, off;
  bin_range(info->count_sym, bin_lo, bin_hi, direct_rows, count, off);
  const int slot = threadIdx.x / GS, glane = threadIdx.x % GS;
  int32_t *K = keys[slot];
  const XcdTiles xt((count + RPB - 1) / RPB);
  // the group's rows, one tile apart; the next row's A extent and first
  // chunk of entries are loaded while the current row is hashed
  auto row_at = [&](int64_t tile) -> int32_t {
    const int64_t t = tile * RPB + slot;
    return (tile < xt.end && t < count) ? (perm ? perm[off + t] : (int32_t)t) : -1;
  };
  int32_t rn = row_at(xt.first);
  int64_t an0 = 0, an1 = 0;
  FirstChunk<float> pre;
  if (rn >= 0) {
    an0 = arpt[rn];
    an1 = arpt[rn + 1];
    pre = load_first_chunk<false, float>(glane, an0, an1, ab_b0, ab_len, nullptr);
  }
  for (int64_t tile = xt.first; rn >= 0; tile += xt.step) {
    const int32_t r = rn;
    const int64_t a0 = an0, a1 = an1;
    const FirstChunk<float> cur = pre;
    rn = row_at(tile + xt.step);
    if (rn >= 0) {
      an0 = arpt[rn];
      an1 = arpt[rn + 1];
    }
    clear_lds<GS, SHD>(K, glane, kEmpty);
    wave_lds_sync();
    int cnt = 0;
    auto ins = [&](const int32_t(&key)[R], const float(&)[R], const bool(&ok)[R], auto &&...) {
      cnt += hash_insert_keys<LOG, R>(K, key, ok, glane);
    };
    row_products<GS, R, false, float, false, O32>(glane, a0, a1, 0, GS, ab_b0, ab_len, nullptr, bcol, nullptr,
                                                  &ent[slot], ins, &cur);
    if (rn >= 0) pre = load_first_chunk<false, float>(glane, an0, an1, ab_b0, ab_len, nullptr);
    cnt = group_sum<GS>(cnt);
    if (glane == 0) row_nz[r] = cnt;
    wave_lds_sync();
  }
}

// Symbolic, one block of BS threads per row; the row's A entries are split
// over the block's waves in chunks of 64.
template <int BS, int SH, int R>
__global__ __launch_bounds__(BS) void k_sym_block(int bin_lo, int bin_hi, int64_t direct_rows,
                                                  const Info *__restrict__ info,
                                                  const int32_t *__restrict__ perm,
                                                  const int64_t *__restrict__ arpt,
                                                  const int64_t *__restrict__ ab_b0,
                                                  const int32_t *__restrict__ ab_len,
                                                  const int32_t *__restrict__ bcol,
                                                  int32_t *__restrict__ row_nz) {
  constexpr int LOG = Log2<SH>::value;
  constexpr int NW = BS / 64;
  __shared__ int32_t K[SH + 64];
  __shared__ EntryScratch<64, float> ent[NW];
  __shared__ int red[NW];
  int64_t count, off;
  bin_range(info->count_sym, bin_lo, bin_hi, direct_rows, count, off);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int64_t t = blockIdx.x; t < count; t += gridDim.x) {
    const int32_t r = perm ? perm[off + t] : (int32_t)t;
    clear_lds<BS, SH + 64>(K, threadIdx.x, kEmpty);
    __syncthreads();
    int cnt = 0;
    row_products<64, R, false, float>(
        lane, arpt[r], arpt[r + 1], 64 * w, 64 * NW, ab_b0, ab_len, nullptr, bcol, nullptr, &ent[w],
        [&](const int32_t(&key)[R], const float(&)[R], const bool(&ok)[R], auto &&...) {
          cnt += hash_insert_keys<LOG, R>(K, key, ok, lane);
        });
    cnt = group_sum<64>(cnt);
    if (lane == 0) red[w] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
      int s = 0;
      for (int q = 0; q < NW; q++) s += red[q];
      row_nz[r] = s;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Exclusive scan of row_nz into C's row pointer (single pass, decoupled
// look-back; scan.hpp), fused with the numeric binning of rows by nnz.
template <typename RowT>
__global__ __launch_bounds__(kScanBS) void k_scan(int64_t M, const int32_t *__restrict__ row_nz,
                                                  RowT *__restrict__ crpt,
                                                  int64_t *__restrict__ crpt64,
                                                  uint8_t *__restrict__ cls,
                                                  uint64_t *__restrict__ status, Info *info) {
  __shared__ int64_t s_wsum[kScanBS / 64 + 1];
  __shared__ int64_t s_prefix;
  __shared__ uint32_t s_tile;
  __shared__ unsigned long long hist[kBins];
  __shared__ unsigned long long s_max;
  if (threadIdx.x == 0) {
    s_tile = atomicAdd(&info->tile_ticket, 1u);
    s_max = 0;
  }
  if (threadIdx.x < kBins) hist[threadIdx.x] = 0;
  __syncthreads();
  const int64_t tile = s_tile;
  const int64_t base = tile * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int32_t v[kScanItems];
  int64_t tsum = 0;
  unsigned long long tmax = 0;
#pragma unroll
  for (int q = 0; q < kScanItems; q++) {
    const int64_t i = base + q;
    v[q] = (i < M) ? row_nz[i] : 0;
    tsum += v[q];
    if (i < M) {
      if (v[q] == 0) {
        cls[i] = kNoBin;
      } else {
        const int b = num_bin_of(v[q]);
        cls[i] = (uint8_t)b;
        atomicAdd(&hist[b], 1ull);
      }
      tmax = tmax > (unsigned long long)v[q] ? tmax : (unsigned long long)v[q];
    }
  }
  int64_t agg;
  int64_t texcl = block_excl_scan<kScanBS, int64_t>(tsum, s_wsum, &agg);
  const int64_t prefix = scan_lookback(tile, agg, status, &s_prefix);
  tmax = group_max<64>(tmax);
  if ((threadIdx.x & 63) == 0) atomicMax(&s_max, tmax);
  __syncthreads();
  // The grid has floor(M / tile) + 1 tiles, so exactly one thread holds
  // index M (the end of C's row pointer, i.e. nnz(C)).
  int64_t run = prefix + texcl;
#pragma unroll
  for (int q = 0; q < kScanItems; q++) {
    const int64_t i = base + q;
    if (i <= M) {
      crpt[i] = (RowT)run;
      crpt64[i] = run;
      if (i == M) {
        info->nnz = run;
        if (sizeof(RowT) == 4 && run > (int64_t)INT32_MAX) info->error = NSP_ERR_OVERFLOW;
      }
    }
    run += v[q];
  }
  if (threadIdx.x < kBins && hist[threadIdx.x])
    atomicAdd((unsigned long long *)&info->count_num[threadIdx.x], hist[threadIdx.x]);
  if (threadIdx.x == 0) atomicMax((unsigned long long *)&info->max_nz, s_max);
}

// ---------------------------------------------------------------------------
// Accumulator of the LDS value tables: FP64 for FP64 and FP32 values alike.
// On gfx950 an LDS FP32 atomic add (ds_add_f32) runs about 20x slower than
// the FP64 one (ds_add_f64), with or without address conflicts, so FP32
// products (formed in FP32) are summed in FP64 and rounded once on output:
// faster, and never less accurate than FP32 accumulation.
using Acc = double;

// Numeric, GS lanes per row, SH-slot LDS table (keys + values) per row, then
// compaction of the occupied slots, a register bitonic sort of the row's
// columns (GS lanes x E keys) and a lookup of each sorted column's value.
template <typename V, int GS, int NS, int SHF, int R, bool O32, bool SPEC = false>
__global__ __launch_bounds__(256) void k_num_group(int bin_lo, int bin_hi, int64_t direct_rows,
                                                   const Info *__restrict__ info,
                                                   const int32_t *__restrict__ perm,
                                                   const int64_t *__restrict__ arpt,
                                                   const int64_t *__restrict__ ab_b0,
                                                   const V *__restrict__ aval,
                                                   const int32_t *__restrict__ ab_len,
                                                   const int32_t *__restrict__ bcol,
                                                   const V *__restrict__ bval,
                                                   const int64_t *__restrict__ crpt,
                                                   int32_t *__restrict__ ccol,
                                                   V *__restrict__ cval, int pack,
                                                   int32_t *__restrict__ spec_nz) {
  constexpr int RPB = 256 / GS;
  constexpr int SH = NS * SHF;    // table slots (NS = largest row nnz of the bin)
  constexpr int LOG = Log2<SH>::value;
  constexpr int SHD = SH + GS;    // + one dummy slot per lane
  constexpr int E = NS / GS;      // keys per lane in the sort
  constexpr int SPL = SH / GS;    // table slots per lane in the compaction
  static_assert(E >= 1, "sort width below group size");
  __shared__ int32_t keys[RPB][SHD];
  __shared__ Acc vals[RPB][SHD];
  __shared__ uint32_t stage[RPB][NS];
  __shared__ EntryScratch<GS, V> ent[RPB];
  int64_t count, off;
  bin_range(info->count_num, bin_lo, bin_hi, direct_rows, count, off);
  const int slot = threadIdx.x / GS, glane = threadIdx.x % GS;
  int32_t *K = keys[slot];
  Acc *W = vals[slot];
  uint32_t *S = stage[slot];
  const XcdTiles xt((count + RPB - 1) / RPB);
  // the group's rows, one tile apart; the next row's A extent is loaded at
  // the top of the current one and its first chunk of entries after the
  // current row's products, so both arrive while this row is hashed, sorted
  // and stored
  auto row_at = [&](int64_t tile) -> int32_t {
    const int64_t t = tile * RPB + slot;
    return (tile < xt.end && t < count) ? (perm ? perm[off + t] : (int32_t)t) : -1;
  };
  int32_t rn = row_at(xt.first);
  int64_t an0 = 0, an1 = 0;
  FirstChunk<V> pre;
  if (rn >= 0) {
    an0 = arpt[rn];
    an1 = arpt[rn + 1];
    pre = load_first_chunk<true>(glane, an0, an1, ab_b0, ab_len, aval);
  }
  for (int64_t tile = xt.first; rn >= 0; tile += xt.step) {
    {
      const int32_t r = rn;
      const int64_t a0 = an0, a1 = an1;
      const FirstChunk<V> cur = pre;
      rn = row_at(tile + xt.step);
      if (rn >= 0) {
        an0 = arpt[rn];
        an1 = arpt[rn + 1];
      }
      // SPEC (one-call product, C's row pointer not known yet): the row
      // goes to slot r of NS entries of ccol/cval, its length to spec_nz
      const int64_t c0 = SPEC ? (int64_t)r * NS : crpt[r];
      int nz = SPEC ? 0 : (int)(crpt[r + 1] - c0);
      if (!SPEC && c0 + nz > info->cap) {  // C buffers too small: skip the row, flag it
        if (glane == 0) atomicOr((unsigned int *)&const_cast<Info *>(info)->error, (unsigned int)NSP_ERR_OVERFLOW);
        if (rn >= 0) pre = load_first_chunk<true>(glane, an0, an1, ab_b0, ab_len, aval);
        continue;
      }
      clear_lds<GS, SHD>(K, glane, kEmpty);
      clear_lds<GS, SHD>(W, glane, 0);
      wave_lds_sync();
      auto acc = [&](const int32_t(&key)[R], const V(&prod)[R], const bool(&ok)[R], auto &&...) {
        Acc pa[R];
#pragma unroll
        for (int u = 0; u < R; u++) pa[u] = (Acc)prod[u];
        hash_accumulate_batch<LOG, R, Acc>(K, W, key, pa, ok, glane);
      };
      row_products<GS, R, true, V, true, O32>(glane, a0, a1, 0, GS, ab_b0, ab_len, aval, bcol, bval, &ent[slot],
                                              acc, &cur);
      wave_lds_sync();
      if (rn >= 0) pre = load_first_chunk<true>(glane, an0, an1, ab_b0, ab_len, aval);
      // compact occupied slots (strided: slot q*GS + glane) into S; with
      // `pack` (C's columns below 2^(32-LOG)) the key carries its slot in
      // the low LOG bits, so the sorted entries read their values directly
      int c = 0;
#pragma unroll
      for (int q = 0; q < SPL; q++) c += (K[q * GS + glane] != kEmpty);
      int pos = group_incl_scan<GS>(c, glane) - c;
      if constexpr (SPEC) {
        nz = group_last<GS>(pos + c);
        if (glane == 0) spec_nz[r] = nz > NS ? 0 : nz;
        if (nz > NS) {  // more columns than the slot and the sort hold: the caller redoes the product
          if (glane == 0) atomicOr((unsigned int *)&const_cast<Info *>(info)->spec_over, 1u);
          wave_lds_sync();
          continue;
        }
      }
#pragma unroll
      for (int q = 0; q < SPL; q++) {
        const int32_t k = K[q * GS + glane];
        if (k != kEmpty) S[pos++] = pack ? (((uint32_t)k << LOG) | (uint32_t)(q * GS + glane)) : (uint32_t)k;
      }
      wave_lds_sync();
      uint32_t v[E];
#pragma unroll
      for (int q = 0; q < E; q++) {
        const int g = glane * E + q;
        v[q] = (g < nz) ? S[g] : kPadKey;
      }
      group_bitonic_sort<GS, E>(v, glane);
      if (pack) {
#pragma unroll
        for (int q = 0; q < E; q++) {
          const int g = glane * E + q;
          if (g < nz) {
            ccol[c0 + g] = (int32_t)(v[q] >> LOG);
            cval[c0 + g] = (V)W[v[q] & ((1u << LOG) - 1u)];
          }
        }
      } else {
#pragma unroll
        for (int q = 0; q < E; q++) {
          const int g = glane * E + q;
          if (g < nz) {
            ccol[c0 + g] = (int32_t)v[q];
            cval[c0 + g] = (V)hash_lookup<LOG, Acc>(K, W, (int32_t)v[q]);
          }
        }
      }
      wave_lds_sync();
    }
  }
}

// ---------------------------------------------------------------------------
// Narrow products (B with at most kNarrowCols columns, such as a batch of
// matrices of up to 4096 rows and columns stored as one concatenated CSR):
// a row's columns as an LDS bitmap of kNarrowCols bits instead of a hash
// table.  The symbolic pass sets one bit per product (an atomic OR: no
// compare-and-swap and no probing, however many of the products are
// distinct) and counts the bits.  The numeric pass sets the bits, turns
// them into each column's position in the sorted row (a running count of
// set bits per word), walks the products a second time adding each into
// the value slot of its position, and writes the row in column order: no
// sort.  Columns are masked to the bitmap (a column past it, which
// violates the product's preconditions, cannot write outside the row's
// bitmap).
constexpr int kNarrowCols = 4096;
constexpr int kNarrowWords = kNarrowCols / 32;

template <int GS, int R, bool O32>
__global__ __launch_bounds__(256) void k_sym_narrow(int bin_lo, int bin_hi, int64_t direct_rows,
                                                    const Info *__restrict__ info,
                                                    const int32_t *__restrict__ perm,
                                                    const int64_t *__restrict__ arpt,
                                                    const int64_t *__restrict__ ab_b0,
                                                    const int32_t *__restrict__ ab_len,
                                                    const int32_t *__restrict__ bcol,
                                                    int32_t *__restrict__ row_nz) {
  constexpr int RPB = 256 / GS;
  constexpr int WPL = kNarrowWords / GS;  // bitmap words per lane in the count
  // rows padded to 640 bytes, so that the groups of a wavefront reading
  // their bitmaps at the same offsets start half the LDS banks apart
  __shared__ __align__(16) uint32_t bits[RPB][kNarrowWords + 32];
  __shared__ EntryScratch<GS, float> ent[RPB];
  int64_t count, off;
  bin_range(info->count_sym, bin_lo, bin_hi, direct_rows, count, off);
  const int slot = threadIdx.x / GS, glane = threadIdx.x % GS;
  uint32_t *Bm = bits[slot];
  const XcdTiles xt((count + RPB - 1) / RPB);
  // the group's rows, one tile apart; the next row's A extent and first
  // chunk of entries are loaded while the current row is counted
  auto row_at = [&](int64_t tile) -> int32_t {
    const int64_t t = tile * RPB + slot;
    return (tile < xt.end && t < count) ? (perm ? perm[off + t] : (int32_t)t) : -1;
  };
  int32_t rn = row_at(xt.first);
  int64_t an0 = 0, an1 = 0;
  FirstChunk<float> pre;
  if (rn >= 0) {
    an0 = arpt[rn];
    an1 = arpt[rn + 1];
    pre = load_first_chunk<false, float>(glane, an0, an1, ab_b0, ab_len, nullptr);
  }
  for (int64_t tile = xt.first; rn >= 0; tile += xt.step) {
    {
      const int32_t r = rn;
      const int64_t a0 = an0, a1 = an1;
      const FirstChunk<float> cur = pre;
      rn = row_at(tile + xt.step);
      if (rn >= 0) {
        an0 = arpt[rn];
        an1 = arpt[rn + 1];
      }
      clear_lds<GS, kNarrowWords>(Bm, glane, 0);
      wave_lds_sync();
      row_products<GS, R, false, float, false, O32>(
          glane, a0, a1, 0, GS, ab_b0, ab_len, nullptr, bcol, nullptr, &ent[slot],
          [&](const int32_t(&key)[R], const float(&)[R], const bool(&ok)[R], auto &&...) {
#pragma unroll
            for (int u = 0; u < R; u++) {
              // lanes past the end OR nothing into a word of their own
              const uint32_t k = (uint32_t)key[u] & (kNarrowCols - 1);
              atomicOr(&Bm[ok[u] ? (k >> 5) : (uint32_t)glane], ok[u] ? (1u << (k & 31)) : 0u);
            }
          },
          &cur);
      if (rn >= 0) pre = load_first_chunk<false, float>(glane, an0, an1, ab_b0, ab_len, nullptr);
      wave_lds_sync();
      int cnt = 0;
      if constexpr (WPL >= 4) {
        // 16-byte words of the bitmap taken lane by lane (consecutive lanes on
        // consecutive banks; the count needs no order)
        const uint4 *B4 = reinterpret_cast<const uint4 *>(Bm);
#pragma unroll
        for (int q = 0; q < WPL / 4; q++) {
          const uint4 x = B4[q * GS + glane];
          cnt += __popc(x.x) + __popc(x.y) + __popc(x.z) + __popc(x.w);
        }
      } else {
#pragma unroll
        for (int q = 0; q < WPL; q++) cnt += __popc(Bm[q * GS + glane]);
      }
      cnt = group_sum<GS>(cnt);
      if (glane == 0) row_nz[r] = cnt;
      wave_lds_sync();
    }
  }
}

// LDS of a block of narrow numeric rows (RPB = 256 / GS rows).  Per bitmap
// word: .x its bits, .y the set bits of the row's words before it.  Word j
// of a row sits at wpos(j) = (j % WPL) * GS + j / WPL: the running count
// gives lane g words [g*WPL, (g+1)*WPL), and at each step the group's lanes
// read consecutive entries; rows are padded to 1152 bytes, so that the
// groups of a wavefront start half the banks apart.
template <typename V, int GS, int NS, typename AccN>
struct NarrowNumLds {
  static constexpr int RPB = 256 / GS;
  alignas(16) uint2 wb[RPB][kNarrowWords + 16];
  alignas(16) AccN vals[RPB][NS + GS];  // + one dummy slot per lane
  EntryScratch<GS, V> ent[RPB];
};

// The rows of tiles first, first + step, ... (< end) of a row list of
// `count` rows (perm[off + t], or t itself without perm), RPB = 256 / GS
// rows a tile, row `slot` of each tile taken by this group of GS lanes,
// with the group's LDS Wb / Vs / ent.
template <typename V, int GS, int NS, int R, bool O32, typename AccN>
__device__ __forceinline__ void num_narrow_rows(int tid, int64_t first, int64_t end, int64_t step, int64_t count,
                                                int64_t off, const Info *__restrict__ info,
                                                const int32_t *__restrict__ perm,
                                                const int64_t *__restrict__ arpt,
                                                const int64_t *__restrict__ ab_b0,
                                                const V *__restrict__ aval,
                                                const int32_t *__restrict__ ab_len,
                                                const int32_t *__restrict__ bcol,
                                                const V *__restrict__ bval,
                                                const int64_t *__restrict__ crpt, int32_t *__restrict__ ccol,
                                                V *__restrict__ cval, uint2 *Wb, AccN *Vs,
                                                EntryScratch<GS, V> *ent) {
  constexpr int RPB = 256 / GS;
  constexpr int WPL = kNarrowWords / GS;
  const int slot = tid / GS, glane = tid % GS;
  auto wpos = [](uint32_t j) -> uint32_t { return (j % WPL) * GS + j / WPL; };
  // adds a window's products into the value slots of their positions; with
  // `first` (the product that set its column's bit), that product also
  // writes the column into C at its position, so each column is written
  // exactly once and no pass over the bitmap is needed
  auto accumulate = [&](const int32_t(&key)[R], const V(&prod)[R], const bool(&ok)[R], const bool *first,
                        int32_t *crow) {
    bool live[R];
#pragma unroll
    for (int u = 0; u < R; u++) live[u] = __ballot(ok[u]) != 0;
#pragma unroll
    for (int u = 0; u < R; u++) {
      if (!live[u]) continue;
      const uint32_t k = (uint32_t)key[u] & (kNarrowCols - 1);
      const uint2 e = Wb[wpos(k >> 5)];
      const uint32_t rank = e.y + (uint32_t)__popc(e.x & ((1u << (k & 31)) - 1u));
      const uint32_t idx = ok[u] ? rank : (uint32_t)(NS + glane);
      __hip_atomic_fetch_add(&Vs[idx], (AccN)prod[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      if (first && first[u]) crow[rank] = (int32_t)k;
    }
  };
  // the group's rows, one tile apart; the next row's A extent and first
  // chunk of entries are loaded while the current row is accumulated
  auto row_at = [&](int64_t tile) -> int32_t {
    const int64_t t = tile * RPB + slot;
    return (tile < end && t < count) ? (perm ? perm[off + t] : (int32_t)t) : -1;
  };
  int32_t rn = row_at(first);
  int64_t an0 = 0, an1 = 0;
  FirstChunk<V> pre;
  if (rn >= 0) {
    an0 = arpt[rn];
    an1 = arpt[rn + 1];
    pre = load_first_chunk<true>(glane, an0, an1, ab_b0, ab_len, aval);
  }
  for (int64_t tile = first; rn >= 0; tile += step) {
    {
      const int32_t r = rn;
      const int64_t a0 = an0, a1 = an1;
      const FirstChunk<V> cur = pre;
      rn = row_at(tile + step);
      if (rn >= 0) {
        an0 = arpt[rn];
        an1 = arpt[rn + 1];
      }
      const int64_t c0 = crpt[r];
      const int nz = (int)(crpt[r + 1] - c0);
      if (c0 + nz > info->cap) {  // C buffers too small: skip the row, flag it
        if (glane == 0) atomicOr((unsigned int *)&const_cast<Info *>(info)->error, (unsigned int)NSP_ERR_OVERFLOW);
        if (rn >= 0) pre = load_first_chunk<true>(glane, an0, an1, ab_b0, ab_len, aval);
        continue;
      }
      clear_lds<GS, kNarrowWords>(Wb, glane, 0);
      clear_lds<GS, NS + GS>(Vs, glane, 0);
      wave_lds_sync();
      // walk 1: the row's columns into the bitmap; a row that is one window
      // long keeps its products in registers for the accumulation, with a
      // flag on the product that set its column's bit first
      int32_t sk[R];
      V sp[R];
      bool so[R], sf[R];
      int windows = 0;
      row_products<GS, R, true, V, true, O32>(
          glane, a0, a1, 0, GS, ab_b0, ab_len, aval, bcol, bval, ent,
          [&](const int32_t(&key)[R], const V(&prod)[R], const bool(&ok)[R], auto &&...) {
#pragma unroll
            for (int u = 0; u < R; u++) {
              // lanes past the end OR nothing into a word of their own
              const uint32_t k = (uint32_t)key[u] & (kNarrowCols - 1);
              const uint32_t bit = ok[u] ? (1u << (k & 31)) : 0u;
              const uint32_t old = atomicOr(&Wb[ok[u] ? wpos(k >> 5) : (uint32_t)glane].x, bit);
              sk[u] = key[u];
              sp[u] = prod[u];
              so[u] = ok[u];
              sf[u] = ok[u] && !(old & bit);
            }
            windows++;
          },
          &cur);
      if (rn >= 0) pre = load_first_chunk<true>(glane, an0, an1, ab_b0, ab_len, aval);
      wave_lds_sync();
      // running counts: lane g holds words [g*WPL, (g+1)*WPL)
      {
        uint32_t cw[WPL];
        int c = 0;
#pragma unroll
        for (int w = 0; w < WPL; w++) {
          cw[w] = (uint32_t)__popc(Wb[w * GS + glane].x);
          c += (int)cw[w];
        }
        uint32_t run = (uint32_t)(group_incl_scan<GS>(c, glane) - c);
#pragma unroll
        for (int w = 0; w < WPL; w++) {
          Wb[w * GS + glane].y = run;
          run += cw[w];
        }
      }
      wave_lds_sync();
      if (windows == 1) {
        accumulate(sk, sp, so, sf, ccol + c0);
      } else {
        // walk 2: each product into the slot of its column's position
        row_products<GS, R, true, V, true, O32>(
            glane, a0, a1, 0, GS, ab_b0, ab_len, aval, bcol, bval, ent,
            [&](const int32_t(&key)[R], const V(&prod)[R], const bool(&ok)[R], auto &&...) {
              accumulate(key, prod, ok, nullptr, nullptr);
            },
            &cur);
        // the columns from the bits (positions from the running counts)
#pragma unroll
        for (int w = 0; w < WPL; w++) {
          const uint2 e = Wb[w * GS + glane];
          uint32_t x = e.x, rk = e.y;
          while (x) {
            const int b = __ffs(x) - 1;
            x &= x - 1u;
            if ((int)rk < nz) ccol[c0 + rk] = (glane * WPL + w) * 32 + b;
            rk++;
          }
        }
      }
      wave_lds_sync();
      // the values in column order
      for (int j = glane; j < nz; j += GS) cval[c0 + j] = (V)Vs[j];
      wave_lds_sync();
    }
  }
}

template <typename V, int GS, int NS, int R, bool O32, typename AccN = Acc>
__global__ __launch_bounds__(256) void k_num_narrow(int bin_lo, int bin_hi, int64_t direct_rows,
                                                    const Info *__restrict__ info,
                                                    const int32_t *__restrict__ perm,
                                                    const int64_t *__restrict__ arpt,
                                                    const int64_t *__restrict__ ab_b0,
                                                    const V *__restrict__ aval,
                                                    const int32_t *__restrict__ ab_len,
                                                    const int32_t *__restrict__ bcol,
                                                    const V *__restrict__ bval,
                                                    const int64_t *__restrict__ crpt,
                                                    int32_t *__restrict__ ccol,
                                                    V *__restrict__ cval) {
  __shared__ NarrowNumLds<V, GS, NS, AccN> L;
  int64_t count, off;
  bin_range(info->count_num, bin_lo, bin_hi, direct_rows, count, off);
  const int slot = threadIdx.x / GS;
  const XcdTiles xt((count + L.RPB - 1) / L.RPB);
  num_narrow_rows<V, GS, NS, R, O32, AccN>(threadIdx.x, xt.first, xt.end, xt.step, count, off, info, perm, arpt, ab_b0, aval,
                                           ab_len, bcol, bval, crpt, ccol, cval, L.wb[slot], L.vals[slot],
                                           &L.ent[slot]);
}

constexpr int kNumR = 4;  // B loads in flight per lane, numeric (at most)
// Loads per lane of a numeric group kernel: a window of GS*R products is
// held near the bin's row-length bound NS (a row has at least as many
// products as entries); at least 2.
constexpr int num_r(int GS, int NS) { return NS / GS >= kNumR ? kNumR : 2; }

// Narrow numeric rows of the wave-level bins 0..5 in one persistent launch
// driven by row-bin work queues.  The tiles of each bin (RPB = 256 / GS rows
// of its row list) are split into kQueues equal ranges, one per XCD; queue
// x hands out tickets of `chunk` tiles over its ranges, the bins with the
// longest rows first.  A block draws tickets from
// the queue of its own XCD (blocks are dispatched to the XCDs in turn) and,
// once that queue is spent, from the others in turn; it ends when all are
// spent.  Each ticket's tiles are walked exactly as the per-bin kernel walks
// its share, with the bin's group size (16 lanes up to 64 entries, 32 up to
// 256, 64 up to 512), in one LDS pool that every bin's layout fits.  The
// next ticket is drawn while the current one is worked, and one barrier per
// ticket hands it to the block's waves (and frees the pool for the next
// bin's layout).  Queue heads live in Info (zeroed by k_init), so the launch
// is valid inside a captured graph.
constexpr int kQueues = 8;
constexpr int kQueueBins = 6;
constexpr int queue_gs(int b) { return b <= 2 ? 16 : b <= 4 ? 32 : 64; }
constexpr int queue_ns(int b) { return (int)kNumBound[b]; }
template <typename V, typename AccN>
constexpr size_t queue_pool_bytes() {
  const size_t a = sizeof(NarrowNumLds<V, 16, 64, AccN>), b = sizeof(NarrowNumLds<V, 32, 256, AccN>),
               c = sizeof(NarrowNumLds<V, 64, 512, AccN>);
  return a > b ? (a > c ? a : c) : (b > c ? b : c);
}

// Bin and tile range [t0, t1) of ticket k of queue q, or bin -1 when the
// queue has no ticket k.
__device__ __forceinline__ int queue_ticket(const int64_t *cnt, int q, int64_t k, int chunk, int64_t &t0,
                                            int64_t &t1) {
  int bin = -1;
#pragma unroll
  for (int b = kQueueBins - 1; b >= 0; b--) {
    const int rpb = 256 / queue_gs(b);
    const int64_t nt = (cnt[b] + rpb - 1) / rpb;
    const int64_t s = nt * q / kQueues, e = nt * (q + 1) / kQueues;
    const int64_t nk = (e - s + chunk - 1) / chunk;
    if (bin < 0) {
      if (k < nk) {
        bin = b;
        t0 = s + k * chunk;
        t1 = t0 + chunk < e ? t0 + chunk : e;
      } else {
        k -= nk;
      }
    }
  }
  return bin;
}

// FP32 values: at most 128 VGPRs (4 waves per SIMD, as the per-bin kernels
// run; 2 VGPRs spill); FP64 values take what they need (151, 3 waves).
template <typename V, bool O32, typename AccN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(V) == 4 ? 4 : 1)))
void k_num_narrow_queue(int chunk, Info *__restrict__ info, const int32_t *__restrict__ perm,
                        const int64_t *__restrict__ arpt, const int64_t *__restrict__ ab_b0,
                        const V *__restrict__ aval, const int32_t *__restrict__ ab_len,
                        const int32_t *__restrict__ bcol, const V *__restrict__ bval,
                        const int64_t *__restrict__ crpt, int32_t *__restrict__ ccol, V *__restrict__ cval) {
  __shared__ __align__(16) unsigned char pool[queue_pool_bytes<V, AccN>()];
  __shared__ uint32_t s_tk[2];
  __shared__ int64_t cnt[kQueueBins], off[kQueueBins];  // rows of each bin and their place in perm
  if (threadIdx.x == 0) {
    int64_t o = 0;
    for (int b = 0; b < kQueueBins; b++) {
      cnt[b] = info->count_num[b];
      off[b] = o;
      o += cnt[b];
    }
  }
  uint32_t *heads = info->queue_num;
  int q = (int)(blockIdx.x % kQueues), tried = 1, it = 0;
  if (threadIdx.x == 0) s_tk[0] = atomicAdd(&heads[q * kQueueStride], 1u);
  __syncthreads();
  // the ticket is the same in every lane: readfirstlane keeps it, and the
  // bin switch below, scalar
  uint32_t tk = __builtin_amdgcn_readfirstlane(s_tk[0]);
  for (;;) {
    int64_t t0 = 0, t1 = 0;
    const int bin = __builtin_amdgcn_readfirstlane(queue_ticket(cnt, q, (int64_t)tk, chunk, t0, t1));
    if (bin < 0) {  // this queue is spent: the next one, until all are
      if (tried == kQueues) break;
      tried++;
      q = (q + 1) % kQueues;
      if (threadIdx.x == 0) s_tk[(it + 1) & 1] = atomicAdd(&heads[q * kQueueStride], 1u);
      __syncthreads();
      tk = __builtin_amdgcn_readfirstlane(s_tk[(it + 1) & 1]);
      it++;
      continue;
    }
    uint32_t nxt = 0;
    if (threadIdx.x == 0) nxt = atomicAdd(&heads[q * kQueueStride], 1u);
    // each bin's path works from an opaque copy of the thread index (an
    // empty asm): otherwise the compiler hoists every path's loop invariants
    // (slot, lane, LDS addresses) out of the ticket loop, all six stay live
    // together, and the kernel needs about 200 VGPRs
#define QUEUE_BIN(B)                                                                                            \
  case B: {                                                                                                     \
    constexpr int GS = queue_gs(B), NS = queue_ns(B);                                                           \
    int tid = (int)threadIdx.x;                                                                                 \
    asm volatile("" : "+v"(tid));                                                                               \
    auto *L = reinterpret_cast<NarrowNumLds<V, GS, NS, AccN> *>(pool);                                          \
    const int slot = tid / GS;                                                                                  \
    num_narrow_rows<V, GS, NS, num_r(GS, NS), O32, AccN>(tid, t0, t1, 1, cnt[B], off[B], info, perm, arpt, ab_b0, \
                                                         aval, ab_len, bcol, bval, crpt, ccol, cval, L->wb[slot], \
                                                         L->vals[slot], &L->ent[slot]);                          \
  } break
    switch (bin) {
      QUEUE_BIN(0);
      QUEUE_BIN(1);
      QUEUE_BIN(2);
      QUEUE_BIN(3);
      QUEUE_BIN(4);
      default: QUEUE_BIN(5);
    }
#undef QUEUE_BIN
    if (threadIdx.x == 0) s_tk[(it + 1) & 1] = nxt;
    __syncthreads();
    tk = __builtin_amdgcn_readfirstlane(s_tk[(it + 1) & 1]);
    it++;
  }
}

// Numeric, one block of BS threads per row; table and sort in LDS.
template <typename V, int BS, int NS, int SHF, int R>
__global__ __launch_bounds__(BS) void k_num_block(int bin_lo, int bin_hi, int64_t direct_rows,
                                                  const Info *__restrict__ info,
                                                  const int32_t *__restrict__ perm,
                                                  const int64_t *__restrict__ arpt,
                                                  const int64_t *__restrict__ ab_b0,
                                                  const V *__restrict__ aval,
                                                  const int32_t *__restrict__ ab_len,
                                                  const int32_t *__restrict__ bcol,
                                                  const V *__restrict__ bval,
                                                  const int64_t *__restrict__ crpt,
                                                  int32_t *__restrict__ ccol,
                                                  V *__restrict__ cval) {
  constexpr int SH = NS * SHF;
  constexpr int LOG = Log2<SH>::value;
  constexpr int NW = BS / 64;
  __shared__ int32_t K[SH + 64];
  __shared__ Acc W[SH + 64];
  __shared__ uint32_t S[NS];
  __shared__ EntryScratch<64, V> ent[NW];
  __shared__ int scan_tmp[NW + 1];
  int64_t count, off;
  bin_range(info->count_num, bin_lo, bin_hi, direct_rows, count, off);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int64_t t = blockIdx.x; t < count; t += gridDim.x) {
    const int32_t r = perm ? perm[off + t] : (int32_t)t;
    if (crpt[r + 1] > info->cap) {  // C buffers too small: skip the row, flag it
      if (threadIdx.x == 0)
        atomicOr((unsigned int *)&const_cast<Info *>(info)->error, (unsigned int)NSP_ERR_OVERFLOW);
      continue;
    }
    clear_lds<BS, SH + 64>(K, threadIdx.x, kEmpty);
    clear_lds<BS, SH + 64>(W, threadIdx.x, 0);
    __syncthreads();
    row_products<64, R, true, V>(lane, arpt[r], arpt[r + 1], 64 * w, 64 * NW, ab_b0, ab_len, aval, bcol, bval,
                                 &ent[w],
                                 [&](const int32_t(&key)[R], const V(&prod)[R], const bool(&ok)[R], auto &&...) {
                                   Acc pa[R];
#pragma unroll
                                   for (int u = 0; u < R; u++) pa[u] = (Acc)prod[u];
                                   hash_accumulate_batch<LOG, R, Acc>(K, W, key, pa, ok, lane);
                                 });
    __syncthreads();
    constexpr int SPT = SH / BS;
    int c = 0;
    for (int q = 0; q < SPT; q++) c += (K[q * BS + threadIdx.x] != kEmpty);
    int pos = block_excl_scan<BS, int>(c, scan_tmp, nullptr);
    for (int q = 0; q < SPT; q++) {
      const int32_t k = K[q * BS + threadIdx.x];
      if (k != kEmpty) S[pos++] = (uint32_t)k;
    }
    const int64_t c0 = crpt[r];
    const int nz = (int)(crpt[r + 1] - c0);
    int n2 = 1;
    while (n2 < nz) n2 <<= 1;
    for (int g = nz + threadIdx.x; g < n2; g += BS) S[g] = kPadKey;
    __syncthreads();
    block_bitonic_sort_lds<BS>(S, n2);
    for (int g = threadIdx.x; g < nz; g += BS) {
      const int32_t key = (int32_t)S[g];
      ccol[c0 + g] = key;
      cval[c0 + g] = (V)hash_lookup<LOG, Acc>(K, W, key);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Heavy rows (more than 16384 intermediate products, or more than 4096
// entries in C): one block of 1024 threads per row, a bitmap of the row's
// columns in LDS instead of a hash table.  The row's column range comes from
// the first and last columns of the B rows it touches (B rows are sorted);
// ranges wider than one bitmap window are covered window by window, each
// window a new walk over the row's products.  (The reference's heavy-row
// fallback is a global-memory hash per row, nsparse kernel_spgemm_hash_d.cu
// 604-678 and 1037-1165, and 8192-column window passes in nsp.cu.)
constexpr int kHeavyBS = 1024;
constexpr int kHeavyNW = kHeavyBS / 64;
constexpr int64_t kSymHeavyWin = 1 << 20;   // columns per symbolic window (128 KB of bits)
constexpr int kHeavyStage = 256;         // C entries a wave stages in LDS per copy-out
constexpr int64_t kNumHeavyWin = 1 << 20;   // columns per numeric window (128 KB bits + 8 KB group ranks)
constexpr int kRankGroup = 8;               // bitmap words per stored rank (two 16-byte reads)
constexpr int kRankSuper = 64;              // rank groups per 32-bit base; group ranks are 16-bit offsets

// Smallest and largest column of a row's products.
template <int BS>
__device__ __forceinline__ void heavy_col_range(int64_t a0, int64_t a1, const int64_t *__restrict__ ab_b0,
                                                const int32_t *__restrict__ ab_len,
                                                const int32_t *__restrict__ bcol, int64_t *s_red,
                                                int64_t &lo, int64_t &hi) {
  int64_t mn = INT64_MAX, mx = -1;
  for (int64_t p = a0 + threadIdx.x; p < a1; p += BS) {
    const int len = ab_len[p];
    if (len > 0) {
      const int64_t b0 = ab_b0[p];
      const int64_t f = bcol[b0], l = bcol[b0 + len - 1];
      mn = f < mn ? f : mn;
      mx = l > mx ? l : mx;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const int64_t a = __shfl_xor(mn, o), b = __shfl_xor(mx, o);
    mn = a < mn ? a : mn;
    mx = b > mx ? b : mx;
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    s_red[2 * w] = mn;
    s_red[2 * w + 1] = mx;
  }
  __syncthreads();
  lo = INT64_MAX;
  hi = -1;
  for (int q = 0; q < BS / 64; q++) {
    lo = s_red[2 * q] < lo ? s_red[2 * q] : lo;
    hi = s_red[2 * q + 1] > hi ? s_red[2 * q + 1] : hi;
  }
  __syncthreads();
}

template <int R>
__global__ __launch_bounds__(kHeavyBS) void k_sym_heavy(int bin_lo, int bin_hi, int64_t direct_rows,
                                                        const Info *__restrict__ info,
                                                        const int32_t *__restrict__ perm,
                                                        const int64_t *__restrict__ arpt,
                                                        const int64_t *__restrict__ ab_b0,
                                                        const int32_t *__restrict__ ab_len,
                                                        const int32_t *__restrict__ bcol,
                                                        int32_t *__restrict__ row_nz) {
  constexpr int NWORD = (int)(kSymHeavyWin / 32);
  __shared__ uint32_t bm[NWORD];
  __shared__ EntryScratch<64, float> ent[kHeavyNW];
  __shared__ int64_t s_red[2 * kHeavyNW];
  __shared__ int s_cnt[kHeavyNW + 1];
  int64_t count, off;
  bin_range(info->count_sym, bin_lo, bin_hi, direct_rows, count, off);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int64_t t = blockIdx.x; t < count; t += gridDim.x) {
    const int32_t r = perm ? perm[off + t] : (int32_t)t;
    const int64_t a0 = arpt[r], a1 = arpt[r + 1];
    int64_t lo, hi;
    heavy_col_range<kHeavyBS>(a0, a1, ab_b0, ab_len, bcol, s_red, lo, hi);
    int64_t total = 0;
    for (int64_t w0 = lo; w0 <= hi; w0 += kSymHeavyWin) {
      for (int q = threadIdx.x; q < NWORD; q += kHeavyBS) bm[q] = 0;
      __syncthreads();
      row_products<64, R, false, float>(
          lane, a0, a1, 64 * w, 64 * kHeavyNW, ab_b0, ab_len, nullptr, bcol, nullptr, &ent[w],
          [&](const int32_t(&key)[R], const float(&)[R], const bool(&ok)[R], auto &&...) {
#pragma unroll
            for (int u = 0; u < R; u++) {
              const int64_t c = (int64_t)key[u] - w0;
              if (ok[u] && c >= 0 && c < kSymHeavyWin) atomicOr(&bm[c >> 5], 1u << (c & 31));
            }
          });
      __syncthreads();
      int c = 0;
      for (int q = threadIdx.x; q < NWORD; q += kHeavyBS) c += __popc(bm[q]);
      c = group_sum<64>(c);
      if (lane == 0) s_cnt[w] = c;
      __syncthreads();
      for (int q = 0; q < kHeavyNW; q++) total += s_cnt[q];
      __syncthreads();
    }
    if (threadIdx.x == 0) row_nz[r] = (int32_t)total;
  }
}

template <typename V, int R, int RK>
__global__ __launch_bounds__(kHeavyBS) void k_num_heavy(int bin_lo, int bin_hi, int64_t direct_rows,
                                                        const Info *__restrict__ info,
                                                        const int32_t *__restrict__ perm,
                                                        const int64_t *__restrict__ arpt,
                                                        const int64_t *__restrict__ ab_b0,
                                                        const V *__restrict__ aval,
                                                        const int32_t *__restrict__ ab_len,
                                                        const int32_t *__restrict__ bcol,
                                                        const V *__restrict__ bval,
                                                        const int64_t *__restrict__ crpt,
                                                        int32_t *__restrict__ ccol,
                                                        V *__restrict__ cval) {
  constexpr int NWORD = (int)(kNumHeavyWin / 32);
  constexpr int WPW = NWORD / kHeavyNW;  // consecutive words per wave in the rank scan
  static_assert(64 % kRankGroup == 0 && WPW % (kRankGroup * kRankSuper) == 0,
                "rank groups and super-groups start on a wave's 64-word steps");
  __shared__ __align__(16) uint32_t bm[NWORD];
  // row position of the first column of each word group: a 32-bit base per
  // super-group of kRankSuper groups plus a 16-bit offset per group (a
  // super-group covers 2^14 columns, so the offset fits)
  static_assert(kRankGroup * kRankSuper * 32 <= 65536, "group offsets fit 16 bits");
  __shared__ uint16_t gbase[NWORD / kRankGroup];
  __shared__ uint32_t sbase[NWORD / (kRankGroup * kRankSuper)];
  __shared__ EntryScratch<64, V> ent[kHeavyNW];
  static_assert(sizeof(EntryScratch<64, V>) >= kHeavyStage * sizeof(int32_t), "the output stage fits a wave's scratch");
  __shared__ int64_t s_red[2 * kHeavyNW];
  __shared__ int s_scan[kHeavyNW + 1];
  int64_t count, off;
  bin_range(info->count_num, bin_lo, bin_hi, direct_rows, count, off);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int64_t t = blockIdx.x; t < count; t += gridDim.x) {
    const int32_t r = perm ? perm[off + t] : (int32_t)t;
    if (crpt[r + 1] > info->cap) {  // C buffers too small: skip the row, flag it
      if (threadIdx.x == 0)
        atomicOr((unsigned int *)&const_cast<Info *>(info)->error, (unsigned int)NSP_ERR_OVERFLOW);
      continue;
    }
    const int64_t a0 = arpt[r], a1 = arpt[r + 1];
    const int64_t c0 = crpt[r], nzr = crpt[r + 1] - c0;
    int64_t lo, hi;
    heavy_col_range<kHeavyBS>(a0, a1, ab_b0, ab_len, bcol, s_red, lo, hi);
    int64_t base = 0;  // entries of the row placed by earlier windows
    for (int64_t w0 = lo; w0 <= hi; w0 += kNumHeavyWin) {
      for (int q = threadIdx.x; q < NWORD; q += kHeavyBS) bm[q] = 0;
      __syncthreads();
      // keys-only walk: RK loads in flight per lane
      row_products<64, RK, false, V>(
          lane, a0, a1, 64 * w, 64 * kHeavyNW, ab_b0, ab_len, nullptr, bcol, nullptr, &ent[w],
          [&](const int32_t(&key)[RK], const V(&)[RK], const bool(&ok)[RK], auto &&...) {
#pragma unroll
            for (int u = 0; u < RK; u++) {
              const int64_t c = (int64_t)key[u] - w0;
              if (ok[u] && c >= 0 && c < kNumHeavyWin) atomicOr(&bm[c >> 5], 1u << (c & 31));
            }
          });
      __syncthreads();
      // ranks: wave w owns words [w*WPW, (w+1)*WPW), taken 64 at a time with
      // lane l on word i*64 + l, so that the columns a wave writes in one
      // step are consecutive entries of C (coalesced stores)
      const int qw = w * WPW;
      int mine = 0;
#pragma unroll
      for (int i = 0; i < WPW / 64; i++) mine += __popc(bm[qw + i * 64 + lane]);
      int wtotal;
      int run = __shfl(block_excl_scan<kHeavyBS, int>(mine, s_scan, &wtotal), 0);  // entries before the wave
      int sb = 0;  // row position of the current super-group's first column
      // sorted columns of the window, values zeroed for the accumulation:
      // the lanes place a step's columns in a per-wave LDS stage (the
      // enumerator's scratch, idle between the walks), kHeavyStage at a time,
      // and the wave copies the stage out with consecutive lanes on
      // consecutive entries, whatever the density of the step's words
      int32_t *stage = reinterpret_cast<int32_t *>(&ent[w]);
#pragma unroll 1
      for (int i = 0; i < WPW / 64; i++) {
        const int q = qw + i * 64 + lane;
        uint32_t x = bm[q];
        const int p = __popc(x);
        const int incl = group_incl_scan<64>(p, lane);
        const int E = group_last<64>(incl);  // entries of the step
        int pos = incl - p;                  // lane's first entry, from the step's first
        if (i % (kRankGroup * kRankSuper / 64) == 0) {  // a super-group starts at lane 0
          sb = run;
          if (lane == 0) sbase[q / (kRankGroup * kRankSuper)] = (uint32_t)run;
        }
        if (lane % kRankGroup == 0) gbase[q / kRankGroup] = (uint16_t)(run + pos - sb);
        const int32_t colq = (int32_t)(w0 + (int64_t)q * 32);
        for (int r0 = 0; r0 < E; r0 += kHeavyStage) {
          while (x && pos < r0 + kHeavyStage) {
            const int b = __ffs(x) - 1;
            x &= x - 1;
            stage[pos - r0] = colq + b;
            pos++;
          }
          wave_lds_sync();
          const int n = E - r0 < kHeavyStage ? E - r0 : kHeavyStage;
          for (int j = lane; j < n; j += 64) {
            const int64_t k = base + run + r0 + j;
            if (k < nzr) {  // the symbolic count bounds the row
              ccol[c0 + k] = stage[j];
              cval[c0 + k] = (V)0;
            }
          }
          wave_lds_sync();
        }
        run += E;
      }
      // the zeroed values are in this XCD's L2, where the block's atomics
      // below land: the barrier's workgroup-scope release (wait for the
      // stores) orders them; an agent-scope fence would write back and
      // invalidate the whole L2 on every row
      __syncthreads();
      row_products<64, R, true, V>(
          lane, a0, a1, 64 * w, 64 * kHeavyNW, ab_b0, ab_len, aval, bcol, bval, &ent[w],
          [&](const int32_t(&key)[R], const V(&prod)[R], const bool(&ok)[R], auto &&...) {
#pragma unroll
            for (int u = 0; u < R; u++) {
              const int64_t c = (int64_t)key[u] - w0;
              if (ok[u] && c >= 0 && c < kNumHeavyWin) {
                // rank = group base + set bits of the group's words before
                // this column (two 16-byte LDS reads, masked popcounts)
                const int q = (int)(c >> 5), b = (int)(c & 31);
                const int g = q / kRankGroup, j = q % kRankGroup;
                const uint4 *G = reinterpret_cast<const uint4 *>(&bm[g * kRankGroup]);
                int cnt = (int)sbase[g / kRankSuper] + (int)gbase[g];
#pragma unroll
                for (int v4 = 0; v4 < kRankGroup / 4; v4++) {
                  const uint4 x = G[v4];
                  const uint32_t wv[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
                  for (int e = 0; e < 4; e++) {
                    const int i = v4 * 4 + e;
                    const uint32_t m = i < j ? 0xffffffffu : (i == j ? ((1u << b) - 1u) : 0u);
                    cnt += __popc(wv[e] & m);
                  }
                }
                const int64_t rk = base + cnt;
                if (rk < nzr) unsafeAtomicAdd(&cval[c0 + rk], prod[u]);
              }
            }
          });
      base += wtotal;
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------------
// One-call product in one hash pass.  k_sym_sample counts the distinct
// columns of up to kSpecSample rows, one at a pseudo-random place in each of
// kSpecSample equal blocks of rows (evenly spaced rows would all fall on one
// face of a grid whose row count per plane divides the spacing), one
// wavefront per row,
// a 2048-slot LDS table; rows past 1024 products are left out, as such a
// product does not take this path) and keeps the largest count.  When every
// row of C is then expected to fit a numeric bin's NS entries, k_num_group
// runs once over all rows with `spec` = NS, writing row r to slot r of a
// buffer of NS entries per row and its length to row_nz; k_scan turns the
// lengths into C's row pointer, and k_spec_compact copies the slots into C.
// A row with more than NS columns is flagged (Info::spec_over) and the
// product is redone in two passes.
constexpr int64_t kSpecSample = 1024;

template <int R>
__global__ __launch_bounds__(256) void k_sym_sample(int64_t MA, int64_t nsample, const int64_t *__restrict__ arpt,
                                                    const int64_t *__restrict__ intprod,
                                                    const int64_t *__restrict__ ab_b0,
                                                    const int32_t *__restrict__ ab_len,
                                                    const int32_t *__restrict__ bcol, Info *info) {
  constexpr int SH = 2048;
  constexpr int LOG = Log2<SH>::value;
  __shared__ int32_t keys[4][SH + 64];
  __shared__ EntryScratch<64, float> ent[4];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int32_t *K = keys[w];
  const int64_t stride = MA / nsample;
  for (int64_t t = blockIdx.x * 4 + w; t < nsample; t += (int64_t)gridDim.x * 4) {
    uint32_t x = (uint32_t)t * 0x9E3779B1u;
    x ^= x >> 15;
    const int64_t r = t * stride + (int64_t)(((uint64_t)x * (uint64_t)stride) >> 32);
    if (intprod[r] > SH / 2) continue;
    clear_lds<64, SH + 64>(K, lane, kEmpty);
    wave_lds_sync();
    int cnt = 0;
    row_products<64, R, false, float>(
        lane, arpt[r], arpt[r + 1], 0, 64, ab_b0, ab_len, nullptr, bcol, nullptr, &ent[w],
        [&](const int32_t(&key)[R], const float(&)[R], const bool(&ok)[R], auto &&...) {
          cnt += hash_insert_keys<LOG, R>(K, key, ok, lane);
        });
    cnt = group_sum<64>(cnt);
    if (lane == 0) atomicMax((unsigned long long *)&info->smp_max, (unsigned long long)cnt);
    wave_lds_sync();
  }
}

// Copies row r's slot (NS entries from r*NS) to its place in C, one
// wavefront per row, launched with a wavefront for every row (the row's
// extent is a dependent load before its copy, so rows are not looped).
template <typename V>
__global__ __launch_bounds__(256) void k_spec_compact(int64_t MA, int64_t NS, int64_t cap,
                                                      const int64_t *__restrict__ crpt,
                                                      const int32_t *__restrict__ scol, const V *__restrict__ sval,
                                                      int32_t *__restrict__ ccol, V *__restrict__ cval) {
  const int lane = threadIdx.x & 63;
  for (int64_t r = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6; r < MA;
       r += ((int64_t)gridDim.x * blockDim.x) >> 6) {
    const int64_t c0 = crpt[r], s0 = r * NS;
    const int64_t n = (crpt[r + 1] < cap ? crpt[r + 1] : cap) - c0;
    for (int64_t j = lane; j < n; j += 64) {
      ccol[c0 + j] = scol[s0 + j];
      cval[c0 + j] = sval[s0 + j];
    }
  }
}

// ---------------------------------------------------------------------------
// host-side launch tables
// Blocks per CU of the numeric block-per-row kernels at most: their rows are
// long and few, and more blocks lengthened the R-MAT numeric pass (0.66 ->
// 0.70 s at 32; its symbolic pass, on the other hand, went from 0.184 to
// 0.176 s, so the symbolic block kernels take the handle's grid size).
constexpr int kBlockGridCap = 8;

static int grid_for(const Handle *h, int64_t items, int per_block, int blocks_per_cu) {
  int64_t need = (items + per_block - 1) / per_block;
  int64_t cap = (int64_t)h->num_cu * blocks_per_cu;
  if (need < 1) need = 1;
  return (int)std::min(need, cap);
}

// Heavy-row kernels: one 1024-thread block per row, at most "heavy_blocks"
// blocks in flight (default one per CU; fewer keeps the C rows being
// accumulated inside L2).
static int heavy_grid(const Handle *h, int64_t rows) {
  const int64_t cap = h->heavy_blocks > 0 ? h->heavy_blocks : h->num_cu;
  return (int)std::max<int64_t>(1, std::min<int64_t>(rows, cap));
}

constexpr int kSymR = 8;  // B loads in flight per lane, symbolic (at most)

// Loads per lane of a symbolic group kernel: a window of GS*R products never
// exceeds the bin's product bound SH (tables hold SH >= the row's products),
// so the lanes of a short row do not issue dummy inserts; at least 2.
constexpr int sym_r(int GS, int SH) { return SH / GS >= kSymR ? kSymR : (SH / GS >= 4 ? 4 : 2); }
// Symbolic table slots of a bin of product bound PB with the doubled table.
constexpr int sym_slots2(int PB) { return PB <= 256 ? 2 * PB : PB; }


// Launches the kernel of bin `bin` over bins [lo, bin] (rows of lower bins fit
// its larger table), or over all rows in natural order when direct >= 0.
// Whether B's entries can be addressed with 32-bit byte offsets (8-byte
// values: at most 2^29 entries).
static bool offsets32(int64_t nnzB) { return nnzB <= ((int64_t)1 << 29); }

// Narrow kernels of the wave-level bins 0..5, with the group size of the
// bin's rows-per-wave parameter.
template <int PB>
static void launch_sym_narrow(Handle *h, int rpw, bool o32, int lo, int bin, int64_t direct, int64_t count,
                              const int32_t *perm, const int64_t *arpt, const int64_t *ab_b0, const int32_t *ab_len,
                              const int32_t *bcol) {
#define SYM_NARROW(GS)                                                                                       \
  do {                                                                                                       \
    const dim3 grid(grid_for(h, count, 256 / GS, h->grid_cap));                                              \
    if (o32)                                                                                                 \
      hipLaunchKernelGGL((k_sym_narrow<GS, sym_r(GS, PB), true>), grid, dim3(256), 0, h->stream, lo, bin, direct, \
                         h->d_info, perm, arpt, ab_b0, ab_len, bcol, h->row_nz);                             \
    else                                                                                                     \
      hipLaunchKernelGGL((k_sym_narrow<GS, sym_r(GS, PB), false>), grid, dim3(256), 0, h->stream, lo, bin,   \
                         direct, h->d_info, perm, arpt, ab_b0, ab_len, bcol, h->row_nz);                     \
  } while (0)
  if (rpw >= 8 && PB <= 128) SYM_NARROW(8);
  else if (rpw >= 4) SYM_NARROW(16);
  else if (rpw == 2) SYM_NARROW(32);
  else SYM_NARROW(64);
#undef SYM_NARROW
}

template <typename V, int GS, int NS>
static void launch_num_narrow_gs(Handle *h, bool o32, int lo, int bin, int64_t direct, int64_t count,
                                 const int32_t *perm, const int64_t *arpt, const int64_t *ab_b0, const V *aval,
                                 const int32_t *ab_len, const int32_t *bcol, const V *bval, int32_t *ccol, V *cval) {
  const dim3 grid(grid_for(h, count, 256 / GS, h->grid_cap));
  if constexpr (sizeof(V) == 4) {
    if (h->fp32_accumulate) {  // FP32 values summed in FP32
      if (o32)
        hipLaunchKernelGGL((k_num_narrow<V, GS, NS, num_r(GS, NS), true, float>), grid, dim3(256), 0, h->stream, lo,
                           bin, direct, h->d_info, perm, arpt, ab_b0, aval, ab_len, bcol, bval, h->c_rpt64, ccol, cval);
      else
        hipLaunchKernelGGL((k_num_narrow<V, GS, NS, num_r(GS, NS), false, float>), grid, dim3(256), 0, h->stream,
                           lo, bin, direct, h->d_info, perm, arpt, ab_b0, aval, ab_len, bcol, bval, h->c_rpt64, ccol,
                           cval);
      return;
    }
  }
  if (o32)
    hipLaunchKernelGGL((k_num_narrow<V, GS, NS, num_r(GS, NS), true>), grid, dim3(256), 0, h->stream, lo, bin,
                       direct, h->d_info, perm, arpt, ab_b0, aval, ab_len, bcol, bval, h->c_rpt64, ccol, cval);
  else
    hipLaunchKernelGGL((k_num_narrow<V, GS, NS, num_r(GS, NS), false>), grid, dim3(256), 0, h->stream, lo, bin,
                       direct, h->d_info, perm, arpt, ab_b0, aval, ab_len, bcol, bval, h->c_rpt64, ccol, cval);
}

// Group size from the bin's rows-per-wave parameter, within what the bin's
// value slots allow in LDS (8 lanes for rows of up to 64 entries, 16 up to
// 256) and never more lanes than its largest row has entries.
template <typename V, int NS>
static void launch_num_narrow(Handle *h, int rpw, bool o32, int lo, int bin, int64_t direct, int64_t count,
                              const int32_t *perm, const int64_t *arpt, const int64_t *ab_b0, const V *aval,
                              const int32_t *ab_len, const int32_t *bcol, const V *bval, int32_t *ccol, V *cval) {
  int gs = rpw >= 8 ? 8 : rpw >= 4 ? 16 : rpw == 2 ? 32 : 64;
  if (NS > 64 && gs < 16) gs = 16;
  if (NS > 256 && gs < 32) gs = 32;
  if (gs > NS) gs = NS < 8 ? 8 : NS;
#define NUM_NARROW(GS) \
  launch_num_narrow_gs<V, GS, NS>(h, o32, lo, bin, direct, count, perm, arpt, ab_b0, aval, ab_len, bcol, bval, ccol, cval)
  if constexpr (NS <= 64) {
    if (gs == 8) { NUM_NARROW(8); return; }
  }
  if constexpr (NS <= 256) {
    if (gs == 16) { NUM_NARROW(16); return; }
  }
  if constexpr (NS >= 32) {
    if (gs == 32) { NUM_NARROW(32); return; }
  }
  if constexpr (NS >= 64) {
    NUM_NARROW(64);
  }
#undef NUM_NARROW
}

static void launch_sym_bin(Handle *h, int bin, int lo, int64_t direct, int64_t count,
                           const int64_t *arpt, const int64_t *ab_b0, const int32_t *ab_len,
                           const int32_t *bcol, int64_t nnzB, int64_t ncols) {
  if (h->narrow_mode && ncols <= kNarrowCols && bin <= 5) {
    const int rpw = bin <= 2 ? h->sym_small_rows_per_wave : h->sym_rows_per_wave;
    const bool o32 = offsets32(nnzB);
    const int32_t *perm = direct >= 0 ? nullptr : h->perm_sym;
    switch (bin) {
      case 0: launch_sym_narrow<32>(h, rpw, o32, lo, bin, direct, count, perm, arpt, ab_b0, ab_len, bcol); break;
      case 1: launch_sym_narrow<64>(h, rpw, o32, lo, bin, direct, count, perm, arpt, ab_b0, ab_len, bcol); break;
      case 2: launch_sym_narrow<128>(h, rpw, o32, lo, bin, direct, count, perm, arpt, ab_b0, ab_len, bcol); break;
      case 3: launch_sym_narrow<256>(h, rpw, o32, lo, bin, direct, count, perm, arpt, ab_b0, ab_len, bcol); break;
      case 4: launch_sym_narrow<512>(h, rpw, o32, lo, bin, direct, count, perm, arpt, ab_b0, ab_len, bcol); break;
      default: launch_sym_narrow<1024>(h, rpw, o32, lo, bin, direct, count, perm, arpt, ab_b0, ab_len, bcol); break;
    }
    return;
  }
  hipStream_t st = h->stream;
  const Info *info = h->d_info;
  const int32_t *perm = direct >= 0 ? nullptr : h->perm_sym;
  int32_t *rnz = h->row_nz;
  const int cap = h->grid_cap;
  const bool o32 = offsets32(nnzB);
  // table slots: the bin's product bound PB, or twice that in bins 0..3
  // ("sym_table_factor" 2): a row whose products are mostly distinct fills
  // a table of PB slots almost completely, and linear probing then walks far
#define SYM_GROUP_T(GS, PB, SH)                                                                    \
  if (o32)                                                                                         \
    hipLaunchKernelGGL((k_sym_group<GS, SH, sym_r(GS, PB), true>), dim3(grid_for(h, count, 256 / GS, cap)), \
                       dim3(256), 0, st, lo, bin, direct, info, perm, arpt, ab_b0, ab_len, bcol, rnz); \
  else                                                                                             \
    hipLaunchKernelGGL((k_sym_group<GS, SH, sym_r(GS, PB), false>), dim3(grid_for(h, count, 256 / GS, cap)), \
                       dim3(256), 0, st, lo, bin, direct, info, perm, arpt, ab_b0, ab_len, bcol, rnz)
#define SYM_GROUP(GS, PB)                                                                          \
  do {                                                                                             \
    if (PB <= 256 && h->sym_table_factor >= 2) { SYM_GROUP_T(GS, PB, sym_slots2(PB)); }           \
    else { SYM_GROUP_T(GS, PB, PB); }                                                              \
  } while (0)
#define SYM_BLOCK(BS, SH)                                                                          \
  hipLaunchKernelGGL((k_sym_block<BS, SH, kSymR>), dim3(grid_for(h, count, 1, cap)), dim3(BS), 0, \
                     st, lo, bin, direct, info, perm, arpt, ab_b0, ab_len, bcol, rnz)
  // wave-level bins: 1, 2 or 4 rows per wavefront (GS = 64, 32, 16 lanes per
  // row; more rows in flight per CU against fewer lanes per row), set by
  // "sym_small_rows_per_wave" for bins 0..2 and "sym_rows_per_wave" for 3..5
  const int rpw = bin <= 2 ? h->sym_small_rows_per_wave : h->sym_rows_per_wave;
#define SYM_RPW(SH) \
  if (rpw >= 4) SYM_GROUP(16, SH); else if (rpw == 2) SYM_GROUP(32, SH); else SYM_GROUP(64, SH);
  // bins 0..2 also take 8 rows per wavefront (8-lane groups)
#define SYM_RPW_SMALL(SH) \
  if (rpw >= 8) { SYM_GROUP(8, SH); } else { SYM_RPW(SH); }
  switch (bin) {
    case 0: SYM_RPW_SMALL(32); break;
    case 1: SYM_RPW_SMALL(64); break;
    case 2: SYM_RPW_SMALL(128); break;
    case 3: SYM_RPW(256); break;
    case 4: SYM_RPW(512); break;
    case 5: SYM_RPW(1024); break;
    case 6: SYM_BLOCK(256, 2048); break;
    case 7: SYM_BLOCK(512, 16384); break;
    case 8:
      hipLaunchKernelGGL((k_sym_heavy<kSymR>), dim3(heavy_grid(h, count)), dim3(kHeavyBS), 0, st, lo, bin,
                         direct, info, perm, arpt, ab_b0, ab_len, bcol, rnz);
      break;
    default: break;
  }
#undef SYM_RPW_SMALL
#undef SYM_RPW
#undef SYM_GROUP
#undef SYM_GROUP_T
#undef SYM_BLOCK
}

// Whether C's columns leave room for a table slot of `slots` entries in the
// low bits of a 32-bit sort key (and stay below the padding key).
static int pack_ok(int64_t ncols, int slots) {
  int log = 0;
  while ((1 << log) < slots) log++;
  return ncols < (int64_t(1) << (32 - log)) ? 1 : 0;
}

// Whether the narrow numeric bins 0..5 run as one work-queue launch.
static bool queue_ok(const Handle *h, int64_t ncols) {
  return h->queue_mode && h->narrow_mode && ncols <= kNarrowCols;
}

// The work-queue launch over the narrow numeric bins 0..5 (rows listed by
// bin in perm_num); a persistent grid of queue_blocks blocks per CU.
template <typename V>
static void launch_num_queue(Handle *h, const int64_t *arpt, const int64_t *ab_b0, const V *aval,
                             const int32_t *ab_len, const int32_t *bcol, const V *bval, int32_t *ccol, V *cval,
                             int64_t nnzB) {
  const dim3 grid((unsigned)std::max(1, h->num_cu * h->queue_blocks));
  const bool o32 = offsets32(nnzB);
#define NQ(O, ACC)                                                                                            \
  hipLaunchKernelGGL((k_num_narrow_queue<V, O, ACC>), grid, dim3(256), 0, h->stream, h->queue_chunk, h->d_info, \
                     h->perm_num, arpt, ab_b0, aval, ab_len, bcol, bval, h->c_rpt64, ccol, cval)
  if constexpr (sizeof(V) == 4) {
    if (h->fp32_accumulate) {  // FP32 values summed in FP32
      if (o32) NQ(true, float);
      else NQ(false, float);
      return;
    }
  }
  if (o32) NQ(true, Acc);
  else NQ(false, Acc);
#undef NQ
}

template <typename V>
static void launch_num_bin(Handle *h, int bin, int lo, int64_t direct, int64_t count,
                           const int64_t *arpt, const int64_t *ab_b0, const V *aval,
                           const int32_t *ab_len, const int32_t *bcol, const V *bval,
                           int32_t *ccol, V *cval, int64_t ncols, int64_t nnzB, int spec = 0,
                           int32_t *spec_nz = nullptr) {
  if (h->narrow_mode && ncols <= kNarrowCols && bin <= 5) {
    const int rpw = bin <= 2 ? h->num_small_rows_per_wave : h->num_rows_per_wave;
    const bool o32 = offsets32(nnzB);
    const int32_t *perm = direct >= 0 ? nullptr : h->perm_num;
#define NN(NS) \
  launch_num_narrow<V, NS>(h, rpw, o32, lo, bin, direct, count, perm, arpt, ab_b0, aval, ab_len, bcol, bval, ccol, cval)
    switch (bin) {
      case 0: NN(16); break;
      case 1: NN(32); break;
      case 2: NN(64); break;
      case 3: NN(128); break;
      case 4: NN(256); break;
      default: NN(512); break;
    }
#undef NN
    return;
  }
  hipStream_t st = h->stream;
  const Info *info = h->d_info;
  const int32_t *perm = direct >= 0 ? nullptr : h->perm_num;
  const int64_t *crpt = h->c_rpt64;
  const int cap = h->grid_cap;
  const bool o32 = offsets32(nnzB);
  // `spec` (the one-call product's slot pass) exists for 32-bit B offsets
  // and table factor 2 only; the caller sets both
#define NUM_GROUP(GS, NS, SHF)                                                                      \
  do {                                                                                              \
    if (spec) {                                                                                     \
      if constexpr ((SHF) == 2)                                                                     \
        hipLaunchKernelGGL((k_num_group<V, GS, NS, 2, num_r(GS, NS), true, true>),                  \
                           dim3(grid_for(h, count, 256 / GS, cap)), dim3(256), 0, st, lo, bin, direct, info, \
                           perm, arpt, ab_b0, aval, ab_len, bcol, bval, crpt, ccol, cval,          \
                           pack_ok(ncols, NS * 2), spec_nz);                                        \
    } else if (o32) {                                                                               \
      hipLaunchKernelGGL((k_num_group<V, GS, NS, SHF, num_r(GS, NS), true>), dim3(grid_for(h, count, 256 / GS, cap)), \
                         dim3(256), 0, st, lo, bin, direct, info, perm, arpt, ab_b0, aval, ab_len,  \
                         bcol, bval, crpt, ccol, cval, pack_ok(ncols, NS * SHF), nullptr);          \
    } else {                                                                                        \
      hipLaunchKernelGGL((k_num_group<V, GS, NS, SHF, num_r(GS, NS), false>), dim3(grid_for(h, count, 256 / GS, cap)), \
                         dim3(256), 0, st, lo, bin, direct, info, perm, arpt, ab_b0, aval, ab_len,  \
                         bcol, bval, crpt, ccol, cval, pack_ok(ncols, NS * SHF), nullptr);          \
    }                                                                                               \
  } while (0)
#define NUM_BLOCK(BS, NS, SHF)                                                                      \
  hipLaunchKernelGGL((k_num_block<V, BS, NS, SHF, kNumR>), dim3(grid_for(h, count, 1, std::min(cap, kBlockGridCap))), dim3(BS),   \
                     0, st, lo, bin, direct, info, perm, arpt, ab_b0, aval, ab_len, bcol, bval,     \
                     crpt, ccol, cval)
  // wave-level bins: 1, 2 or 4 rows per wavefront (GS = 64, 32, 16; never
  // more lanes than the bin's largest row has entries), set by
  // "num_small_rows_per_wave" for bins 0..2 and "num_rows_per_wave" for 3..4;
  // table factor 2 or 4 slots per row entry ("num_table_factor") for 0..3
  const int rpw = bin <= 2 ? h->num_small_rows_per_wave : h->num_rows_per_wave;
  const bool f4 = h->num_table_factor >= 4;
#define NUM_F(GS, NS) \
  if (f4) NUM_GROUP(GS, NS, 4); else NUM_GROUP(GS, NS, 2);
  switch (bin) {
    case 0: if (rpw >= 8) { NUM_F(8, 16); } else { NUM_F(16, 16); } break;
    case 1: if (rpw >= 8) { NUM_F(8, 32); } else if (rpw >= 4) { NUM_F(16, 32); } else { NUM_F(32, 32); } break;
    case 2:
      if (rpw >= 8) { NUM_F(8, 64); } else if (rpw >= 4) { NUM_F(16, 64); } else if (rpw == 2) { NUM_F(32, 64); }
      else { NUM_F(64, 64); }
      break;
    case 3: if (rpw >= 4) { NUM_F(16, 128); } else if (rpw == 2) { NUM_F(32, 128); } else { NUM_F(64, 128); } break;
    case 4:
      if (rpw >= 4) NUM_GROUP(16, 256, 2); else if (rpw == 2) NUM_GROUP(32, 256, 2); else NUM_GROUP(64, 256, 2);
      break;
    case 5: NUM_GROUP(64, 512, 2); break;
    case 6: NUM_BLOCK(256, 2048, 2); break;
    case 7: NUM_BLOCK(512, 4096, 2); break;
    case 8:
      hipLaunchKernelGGL((k_num_heavy<V, kNumR, kSymR>), dim3(heavy_grid(h, count)), dim3(kHeavyBS), 0, st, lo,
                         bin, direct, info, perm, arpt, ab_b0, aval, ab_len, bcol, bval, crpt, ccol, cval);
      break;
    default: break;
  }
#undef NUM_F
#undef NUM_GROUP
#undef NUM_BLOCK
}

// ---------------------------------------------------------------------------
// Launch plan over the bins.  Wave-level bins 0..5: if the highest non-empty
// one holds at least 15% of their rows, its kernel takes all of them (its
// table fits every smaller row) -- and when no row needs a block-level bin,
// it walks all rows in natural order with no row list (the reference's
// DIRECT rule, nsp/src/nsp.cu:443-452).  Block-level bins launch one kernel
// each.
struct Launch {
  int bin, lo;
  int64_t direct, count;
};

static int plan_launches(const int64_t *counts, int64_t M, Launch *out) {
  int n = 0;
  int bmax = -1;
  int64_t group_rows = 0, block_rows = 0;
  for (int b = 0; b <= 5; b++)
    if (counts[b] > 0) { bmax = b; group_rows += counts[b]; }
  for (int b = 6; b < kBins; b++) block_rows += counts[b];
  if (bmax >= 0) {
    if (counts[bmax] * 100 >= group_rows * 15) {
      if (block_rows == 0) out[n++] = Launch{bmax, 0, M, M};
      else out[n++] = Launch{bmax, 0, -1, group_rows};
    } else {
      for (int b = bmax; b >= 0; b--)
        if (counts[b] > 0) out[n++] = Launch{b, b, -1, counts[b]};
    }
  }
  for (int b = kBins - 1; b >= 6; b--)
    if (counts[b] > 0) out[n++] = Launch{b, b, -1, counts[b]};
  return n;
}

static bool needs_row_list(const Launch *L, int n) {
  for (int q = 0; q < n; q++)
    if (L[q].direct < 0) return true;
  return false;
}

static void launch_scatter(Handle *h, int64_t M, int64_t *counts, int64_t *cursor, int32_t *perm) {
  if (M <= 0) return;
  const int64_t g = std::max<int64_t>(1, std::min<int64_t>((M + 2047) / 2048, 2048));
  int64_t chunk = (M + g - 1) / g;
  chunk = (chunk + 255) / 256 * 256;
  const int64_t blocks = (M + chunk - 1) / chunk;
  hipLaunchKernelGGL(k_scatter, dim3((unsigned)blocks), dim3(256), 0, h->stream, M, chunk, h->cls, counts,
                     cursor, perm);
}

// Optional per-phase device timing with HIP events on the handle's stream:
// 0 start, 1 after binning, 2 after symbolic, 3 after scan, 4/5 numeric.
static void timing_mark(Handle *h, int i) {
  if (!h->timing) return;
  if (!h->events[i]) (void)hipEventCreate(&h->events[i]);
  (void)hipEventRecord(h->events[i], h->stream);
}

int collect_timing(Handle *h) {
  if (!h->timing || !h->timing_pending) return NSP_OK;
  float t = 0;
  if (h->timing_fused) {  // one-call product: binning + sample, numeric into slots, scan, compaction
    NSP_HIP(h, hipEventSynchronize(h->events[7]));
    if (hipEventElapsedTime(&t, h->events[0], h->events[1]) == hipSuccess) h->stats.t_bin_ms = t;
    h->stats.t_sym_ms = 0;
    if (hipEventElapsedTime(&t, h->events[4], h->events[5]) == hipSuccess) h->stats.t_num_ms = t;
    if (hipEventElapsedTime(&t, h->events[5], h->events[3]) == hipSuccess) h->stats.t_scan_ms = t;
    if (hipEventElapsedTime(&t, h->events[6], h->events[7]) == hipSuccess) h->stats.t_compact_ms = t;
    h->timing_pending = false;
    return NSP_OK;
  }
  NSP_HIP(h, hipEventSynchronize(h->events[5]));
  if (hipEventElapsedTime(&t, h->events[0], h->events[1]) == hipSuccess) h->stats.t_bin_ms = t;
  if (hipEventElapsedTime(&t, h->events[1], h->events[2]) == hipSuccess) h->stats.t_sym_ms = t;
  if (hipEventElapsedTime(&t, h->events[2], h->events[3]) == hipSuccess) h->stats.t_scan_ms = t;
  if (hipEventElapsedTime(&t, h->events[4], h->events[5]) == hipSuccess) h->stats.t_num_ms = t;
  h->timing_pending = false;
  return NSP_OK;
}

// ---------------------------------------------------------------------------
// Bytes of the product workspace for A with MA rows and nnzA entries.
// Blocks per CU of k_extents (32 was slower: 0.094 against 0.085 ms of
// binning on the stencil share), and the per-block product sums it leaves
// for the light row info, at most.
constexpr int kExtBlocksPerCU = 8;
constexpr int kExtPartials = 4096;

static size_t workspace_bytes(int64_t MA, int64_t nnzA) {
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const int64_t ntiles = MA / kScanTile + 2;
  return al((MA + 1) * 8) * 3 + al(nnzA * 8 + 8) + al(nnzA * 4 + 4) + al(MA * 4) * 3 + al(MA + 1) +
         al(ntiles * 8) + al(kExtPartials * 8);
}

static int ensure_workspace(Handle *h, int64_t MA, int64_t nnzA) {
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const int64_t ntiles = MA / kScanTile + 2;
  const size_t need = workspace_bytes(MA, nnzA);
  if (need > h->ws.bytes) {
    h->ext_rows = -1;  // the cached B row extents go with the old buffer
    if (h->ws.base) NSP_HIP(h, hipFree(h->ws.base));
    h->ws.base = nullptr;
    h->ws.bytes = 0;
    size_t bytes = need + need / 4;
    if (hipMalloc(&h->ws.base, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return set_error(h, NSP_ERR_OOM, "workspace allocation failed");
    }
    h->ws.bytes = bytes;
  }
  char *p = (char *)h->ws.base;
  auto take = [&](size_t b) { char *q = p; p += al(b); return (void *)q; };
  h->a_rpt64 = (int64_t *)take((MA + 1) * 8);
  h->c_rpt64 = (int64_t *)take((MA + 1) * 8);
  h->intprod = (int64_t *)take((MA + 1) * 8);
  h->ab_b0 = (int64_t *)take(nnzA * 8 + 8);
  h->ab_len = (int32_t *)take(nnzA * 4 + 4);
  h->row_nz = (int32_t *)take(MA * 4);
  h->perm_sym = (int32_t *)take(MA * 4);
  h->perm_num = (int32_t *)take(MA * 4);
  h->cls = (uint8_t *)take(MA + 1);
  h->tile_status = (uint64_t *)take(ntiles * 8);
  h->ext_partial = (unsigned long long *)take(kExtPartials * 8);
  h->ntiles = ntiles;
  return NSP_OK;
}

static int check_csr_args(Handle *h, const char *name, const nsp_csr_t *X) {
  if (!X) return set_error(h, NSP_ERR_INVALID, std::string(name) + " is NULL");
  if (X->rows < 0 || X->cols < 0 || X->nnz < 0)
    return set_error(h, NSP_ERR_INVALID, std::string(name) + " has a negative size");
  if (X->rows >= INT32_MAX || X->cols > INT32_MAX)
    return set_error(h, NSP_ERR_UNSUPPORTED, std::string(name) + " has >= 2^31 rows or columns");
  if (!X->row_ptr)  // rows+1 entries, so at least one even for an empty matrix
    return set_error(h, NSP_ERR_INVALID, std::string(name) + ".row_ptr is NULL");
  if (X->nnz > 0 && (!X->col || !X->val))
    return set_error(h, NSP_ERR_INVALID, std::string(name) + ".col/.val is NULL");
  return NSP_OK;
}

// The counters the host reads back: Info up to the work-queue heads, which
// only the device uses.
constexpr size_t kInfoReadBytes = offsetof(Info, queue_num);

static int read_info(Handle *h) {
  NSP_HIP(h, hipMemcpyAsync(h->h_info, h->d_info, kInfoReadBytes, hipMemcpyDeviceToHost, h->stream));
  NSP_HIP(h, hipStreamSynchronize(h->stream));
  h->stats.host_syncs++;
  return NSP_OK;
}

// Row info + symbolic binning (no host sync).  Also used by nsp_intprod.
template <typename RowT>
static int run_rowinfo(Handle *h, const nsp_csr_t *A, const nsp_csr_t *B, const Batch &bt = Batch{},
                       int64_t cap = INT64_MAX) {
  const int64_t MA = A->rows;
  NSP_TRY(ensure_workspace(h, MA, A->nnz));
  timing_mark(h, 0);
  hipLaunchKernelGGL(k_init, dim3(std::max<int64_t>(1, std::min<int64_t>((h->ntiles + 255) / 256, 1024))),
                     dim3(256), 0, h->stream, h->d_info, h->tile_status, h->ntiles, cap);
  const int64_t nnzA = A->nnz;
  if (nnzA > 0) {
    const int ge = (int)std::max<int64_t>(
        1, std::min<int64_t>((nnzA + 256 * kExtU - 1) / (256 * kExtU), (int64_t)h->num_cu * kExtBlocksPerCU));
    hipLaunchKernelGGL(k_extents<RowT>, dim3(ge), dim3(256), 0, h->stream, nnzA, (const RowT *)A->row_ptr, A->col,
                       (const RowT *)B->row_ptr, bt.a_off, bt.b_off, bt.count, h->ab_b0, h->ab_len,
                       (unsigned long long *)nullptr);
  }
  // two blocks per CU: every block ends with a few global atomics on the
  // same counters, which serialise (8 blocks per CU: 58 us on the 524 K-row
  // stencil share, 2 per CU: 33 us)
  const int g = (int)std::max<int64_t>(1, std::min<int64_t>((MA + 255) / 256, (int64_t)h->num_cu * 2));
  // perm_num is free until the numeric binning: it holds the long-row list
  hipLaunchKernelGGL(k_rowinfo<RowT>, dim3(g), dim3(256), 0, h->stream, MA, (const RowT *)A->row_ptr, h->ab_len,
                     h->a_rpt64, h->intprod, h->row_nz, h->cls, h->perm_num, h->d_info);
  hipLaunchKernelGGL(k_rowinfo_long<RowT>, dim3(h->num_cu), dim3(256), 0, h->stream, (const RowT *)A->row_ptr,
                     h->ab_len, (const int32_t *)h->perm_num, h->a_rpt64, h->intprod, h->row_nz, h->cls, h->d_info);
  NSP_HIP(h, hipGetLastError());
  // what ab_b0 / ab_len now describe
  h->ext_key[0] = A->col;
  h->ext_key[1] = A->row_ptr;
  h->ext_key[2] = B->row_ptr;
  h->ext_key[3] = bt.a_off;
  h->ext_rows = MA;
  h->ext_nnz = A->nnz;
  h->ext_nprod = -1;
  h->poff_valid = false;
  return NSP_OK;
}

// Row info of a one-call product that reuses the previous call's slot size
// (k_rowinfo_light in place of k_rowinfo: the slot pass walks every row in
// natural order and needs neither bins nor per-row product counts).
template <typename RowT>
static int run_rowinfo_light(Handle *h, const nsp_csr_t *A, const nsp_csr_t *B) {
  const int64_t MA = A->rows;
  NSP_TRY(ensure_workspace(h, MA, A->nnz));
  timing_mark(h, 0);
  hipLaunchKernelGGL(k_init, dim3(std::max<int64_t>(1, std::min<int64_t>((h->ntiles + 255) / 256, 1024))),
                     dim3(256), 0, h->stream, h->d_info, h->tile_status, h->ntiles, INT64_MAX);
  const int64_t nnzA = A->nnz;
  unsigned long long *partial = h->ext_partial;  // one sum per k_extents block
  int ge = 0;
  if (nnzA > 0) {
    ge = (int)std::max<int64_t>(
        1, std::min<int64_t>({(nnzA + 256 * kExtU - 1) / (256 * kExtU), (int64_t)h->num_cu * kExtBlocksPerCU,
                              (int64_t)kExtPartials}));
    hipLaunchKernelGGL(k_extents<RowT>, dim3(ge), dim3(256), 0, h->stream, nnzA, (const RowT *)A->row_ptr, A->col,
                       (const RowT *)B->row_ptr, (const int64_t *)nullptr, (const int64_t *)nullptr, (int64_t)0,
                       h->ab_b0, h->ab_len, partial);
  }
  hipLaunchKernelGGL(k_rowinfo_light<RowT>,
                     dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((MA + 256) / 256, (int64_t)h->num_cu * 8))),
                     dim3(256), 0, h->stream, MA, (const RowT *)A->row_ptr, (const unsigned long long *)partial, ge,
                     h->a_rpt64, h->d_info);
  NSP_HIP(h, hipGetLastError());
  h->ext_key[0] = A->col;
  h->ext_key[1] = A->row_ptr;
  h->ext_key[2] = B->row_ptr;
  h->ext_key[3] = nullptr;
  h->ext_rows = MA;
  h->ext_nnz = A->nnz;
  h->ext_nprod = -1;
  h->poff_valid = false;
  return NSP_OK;
}

template <typename RowT>
static int symbolic_impl(Handle *h, const nsp_csr_t *A, const nsp_csr_t *B, void *c_row_ptr,
                         int64_t *c_nnz, const Batch &bt = Batch{}, bool info_ready = false) {
  const int64_t MA = A->rows;
  if (!info_ready) {  // else the row info of this product is in place and read back
    h->stats = nsp_stats_t{};
    h->have_plan = false;
    h->timing_pending = false;
    h->timing_fused = false;
    NSP_TRY(run_rowinfo<RowT>(h, A, B, bt));
    NSP_TRY(read_info(h));
  }
  const Info &hi = *h->h_info;
  {
    Launch L[kBins];
    const int n = plan_launches(hi.count_sym, MA, L);
    if (needs_row_list(L, n))
      launch_scatter(h, MA, h->d_info->count_sym, h->d_info->cursor_sym, h->perm_sym);
    timing_mark(h, 1);
    for (int q = 0; q < n; q++)
      launch_sym_bin(h, L[q].bin, L[q].lo, L[q].direct, L[q].count, h->a_rpt64, h->ab_b0, h->ab_len, B->col,
                     B->nnz, B->cols);
    h->stats.sym_launches = n;
    if (n == 1 && L[0].direct >= 0) h->stats.direct_mode |= 1;
  }
  timing_mark(h, 2);
  NSP_HIP(h, hipGetLastError());
  if (MA > 0) {
    int64_t ntile_used = MA / kScanTile + 1;
    hipLaunchKernelGGL(k_scan<RowT>, dim3((unsigned)ntile_used), dim3(kScanBS), 0, h->stream, MA,
                       h->row_nz, (RowT *)c_row_ptr, h->c_rpt64, h->cls, h->tile_status, h->d_info);
  } else {
    RowT zero = 0;
    NSP_HIP(h, hipMemcpyAsync(c_row_ptr, &zero, sizeof(RowT), hipMemcpyHostToDevice, h->stream));
    int64_t z64 = 0;
    NSP_HIP(h, hipMemcpyAsync(h->c_rpt64, &z64, 8, hipMemcpyHostToDevice, h->stream));
  }
  timing_mark(h, 3);
  NSP_HIP(h, hipGetLastError());
  NSP_TRY(read_info(h));
  if (hi.error) return set_error(h, (int)hi.error, "C has 2^31 or more entries; use NSP_IDX64");
  h->stats.nprod = hi.nprod;
  h->stats.nnz = (MA > 0) ? hi.nnz : 0;
  h->stats.max_intprod = hi.max_intprod;
  h->stats.max_row_nnz = hi.max_nz;
  h->stats.n_sym_bins = kBins;
  h->stats.n_num_bins = kBins;
  for (int b = 0; b < kBins; b++) {
    h->stats.sym_bin[b] = hi.count_sym[b];
    h->stats.num_bin[b] = hi.count_num[b];
    h->num_count[b] = hi.count_num[b];
  }
  *c_nnz = h->stats.nnz;
  h->have_plan = true;
  h->plan_rows = MA;
  h->plan_nnz = h->stats.nnz;
  h->plan_a = A->col;
  h->plan_b = B->col;
  return NSP_OK;
}

int spgemm_symbolic(Handle *h, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *A, const nsp_csr_t *B,
                    void *c_row_ptr, int64_t *c_nnz) {
  NSP_TRY(check_csr_args(h, "A", A));
  NSP_TRY(check_csr_args(h, "B", B));
  if (vt != NSP_F64 && vt != NSP_F32) return set_error(h, NSP_ERR_INVALID, "unknown value type");
  if (A->cols != B->rows) return set_error(h, NSP_ERR_INVALID, "A.cols != B.rows");
  if (!c_row_ptr || !c_nnz) return set_error(h, NSP_ERR_INVALID, "c_row_ptr / c_nnz is NULL");
  if (it == NSP_IDX32) {
    if (A->nnz > INT32_MAX || B->nnz > INT32_MAX)
      return set_error(h, NSP_ERR_OVERFLOW, "nnz >= 2^31 needs NSP_IDX64");
    return symbolic_impl<int32_t>(h, A, B, c_row_ptr, c_nnz);
  }
  if (it == NSP_IDX64) return symbolic_impl<int64_t>(h, A, B, c_row_ptr, c_nnz);
  return set_error(h, NSP_ERR_INVALID, "unknown index type");
}

template <typename V>
static int numeric_impl(Handle *h, const nsp_csr_t *A, const nsp_csr_t *B, nsp_csr_t *C) {
  int nwave = 0;
  for (int b = 0; b < kQueueBins; b++) nwave += h->num_count[b] > 0;
  if (nwave >= 2 && queue_ok(h, B->cols)) {  // the wave-level bins in one work-queue launch
    launch_scatter(h, A->rows, h->d_info->count_num, h->d_info->cursor_num, h->perm_num);
    timing_mark(h, 4);
    launch_num_queue<V>(h, h->a_rpt64, h->ab_b0, (const V *)A->val, h->ab_len, B->col, (const V *)B->val, C->col,
                        (V *)C->val, B->nnz);
    int n = 1;
    for (int b = kBins - 1; b >= kQueueBins; b--)
      if (h->num_count[b] > 0) {
        launch_num_bin<V>(h, b, b, -1, h->num_count[b], h->a_rpt64, h->ab_b0, (const V *)A->val, h->ab_len,
                          B->col, (const V *)B->val, C->col, (V *)C->val, B->cols, B->nnz);
        n++;
      }
    h->stats.num_launches = n;
    h->stats.direct_mode &= 1;
    timing_mark(h, 5);
    h->timing_fused = false;
    h->timing_pending = h->timing;
    NSP_HIP(h, hipGetLastError());
    return NSP_OK;
  }
  Launch L[kBins];
  const int n = plan_launches(h->num_count, A->rows, L);
  if (needs_row_list(L, n))
    launch_scatter(h, A->rows, h->d_info->count_num, h->d_info->cursor_num, h->perm_num);
  timing_mark(h, 4);
  for (int q = 0; q < n; q++)
    launch_num_bin<V>(h, L[q].bin, L[q].lo, L[q].direct, L[q].count, h->a_rpt64, h->ab_b0, (const V *)A->val,
                      h->ab_len, B->col, (const V *)B->val, C->col, (V *)C->val, B->cols, B->nnz);
  h->stats.num_launches = n;
  h->stats.direct_mode = (h->stats.direct_mode & 1) | ((n == 1 && L[0].direct >= 0) ? 2 : 0);
  timing_mark(h, 5);
  h->timing_fused = false;
  h->timing_pending = h->timing;
  NSP_HIP(h, hipGetLastError());
  return NSP_OK;
}

int spgemm_numeric(Handle *h, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *A, const nsp_csr_t *B,
                   nsp_csr_t *C) {
  if (!h->have_plan) return set_error(h, NSP_ERR_STATE, "nsp_spgemm_numeric without a symbolic phase");
  if (!A || !B || !C) return set_error(h, NSP_ERR_INVALID, "NULL matrix");
  if (A->rows != h->plan_rows || A->col != h->plan_a || B->col != h->plan_b)
    return set_error(h, NSP_ERR_STATE, "A/B differ from the symbolic phase");
  if (C->nnz != h->plan_nnz) return set_error(h, NSP_ERR_INVALID, "C.nnz differs from the symbolic count");
  if (C->nnz > 0 && (!C->col || !C->val)) return set_error(h, NSP_ERR_INVALID, "C.col/C.val is NULL");
  (void)it;
  if (vt == NSP_F64) return numeric_impl<double>(h, A, B, C);
  if (vt == NSP_F32) return numeric_impl<float>(h, A, B, C);
  return set_error(h, NSP_ERR_INVALID, "unknown value type");
}

// ---------------------------------------------------------------------------
// One-call product (nsp_spgemm_into, nsp_spgemm): C = A*B with the numeric
// pass as the only hash pass when a sample of rows says every row of C fits
// a wave-level numeric bin (see k_sym_sample).  spgemm_fused_count leaves
// C's row pointer and nnz, with the rows in the handle's slot buffer;
// spgemm_fused_fill copies them into C.  Products the sample rules out
// (rows past 1024 products, column-bitmap products, sampled rows of more
// than 512 columns) and products where a row outgrows its slot take the
// two-pass path, whose numeric pass spgemm_fused_fill then runs.
constexpr size_t kSpecMaxBytes = size_t(16) << 30;

// Copies the slots in the handle's buffer (rows of NS entries) into C's
// col/val (capacity cap; entries past it are not written).
static void launch_spec_compact(Handle *h, nsp_value_t vt, int64_t MA, int64_t NS, int64_t cap, int32_t *ccol,
                                void *cval) {
  if (MA <= 0 || cap <= 0) return;
  const size_t cbytes = ((size_t)MA * NS * 4 + 255) & ~(size_t)255;
  const int32_t *scol = (const int32_t *)h->spec.base;
  const void *sval = (const char *)h->spec.base + cbytes;
  const dim3 grid((unsigned)std::max<int64_t>(1, std::min<int64_t>((MA + 3) / 4, INT32_MAX)));
  if (vt == NSP_F64)
    hipLaunchKernelGGL(k_spec_compact<double>, grid, dim3(256), 0, h->stream, MA, NS, cap, h->c_rpt64, scol,
                       (const double *)sval, ccol, (double *)cval);
  else
    hipLaunchKernelGGL(k_spec_compact<float>, grid, dim3(256), 0, h->stream, MA, NS, cap, h->c_rpt64, scol,
                       (const float *)sval, ccol, (float *)cval);
}

template <typename RowT, typename V>
static int fused_count_impl(Handle *h, const nsp_csr_t *A, const nsp_csr_t *B, void *c_row_ptr, int64_t *c_nnz,
                            nsp_csr_t *into, bool *filled) {
  const int64_t MA = A->rows;
  h->fused_plan = false;
  h->timing_fused = false;
  if (filled) *filled = false;
  if (h->fused_mode == 0 || MA == 0 || A->nnz == 0 || B->nnz == 0 || !offsets32(B->nnz) ||
      (h->narrow_mode && B->cols <= kNarrowCols))
    return symbolic_impl<RowT>(h, A, B, c_row_ptr, c_nnz);
  h->stats = nsp_stats_t{};
  h->have_plan = false;
  h->timing_pending = false;
  // The slot size of the previous one-call product on the same buffers and
  // sizes is taken again without a sample and without reading the row info
  // back (no host synchronisation before the slot pass); a row that no
  // longer fits is caught like any other.
  const void *key[4] = {A->col, A->row_ptr, B->col, B->row_ptr};
  const int64_t dims[4] = {MA, A->nnz, B->nnz, B->cols};
  const bool cached = h->fc_bin >= 0 && std::equal(key, key + 4, h->fc_key) && std::equal(dims, dims + 4, h->fc_dims);
  h->fc_bin = -1;
  if (cached) NSP_TRY(run_rowinfo_light<RowT>(h, A, B));
  else NSP_TRY(run_rowinfo<RowT>(h, A, B));
  const Info &hi = *h->h_info;
  int bin = -1;
  if (cached) {
    bin = h->fc_last;
    timing_mark(h, 1);
  } else {
    const int64_t ns = std::min<int64_t>(MA, kSpecSample);
    hipLaunchKernelGGL(k_sym_sample<kSymR>, dim3((unsigned)std::min<int64_t>((ns + 3) / 4, (int64_t)h->num_cu * 4)),
                       dim3(256), 0, h->stream, MA, ns, h->a_rpt64, h->intprod, h->ab_b0, h->ab_len, B->col,
                       h->d_info);
    NSP_HIP(h, hipGetLastError());
    timing_mark(h, 1);
    NSP_TRY(read_info(h));
    // the smallest wave-level numeric bin whose NS holds the sampled rows (a
    // row past it is caught and redone in two passes), or the bin that holds
    // the largest row's product count, which no row of C can outgrow, when
    // it is at most one bin larger
    if (hi.count_sym[6] + hi.count_sym[7] + hi.count_sym[8] == 0) {
      int bs = -1, bp = -1;
      for (int b = 5; b >= 0; b--) {
        if (hi.smp_max > 0 && hi.smp_max <= kNumBound[b]) bs = b;
        if (hi.max_intprod <= kNumBound[b]) bp = b;
      }
      bin = (bp >= 0 && (bs < 0 || bp <= bs + 1)) ? bp : bs;
    }
  }
  const int64_t NS = bin >= 0 ? kNumBound[bin] : 0;
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t cbytes = al((size_t)MA * NS * 4), need = cbytes + al((size_t)MA * NS * sizeof(V));
  if (bin < 0 || need > kSpecMaxBytes) {
    if (cached) return symbolic_impl<RowT>(h, A, B, c_row_ptr, c_nnz);  // the light row info has no bins
    return symbolic_impl<RowT>(h, A, B, c_row_ptr, c_nnz, Batch{}, true);
  }
  if (need > h->spec.bytes) {
    if (h->spec.base) NSP_HIP(h, hipFree(h->spec.base));
    h->spec.base = nullptr;
    h->spec.bytes = 0;
    if (hipMalloc(&h->spec.base, need) != hipSuccess) {
      (void)hipGetLastError();
      if (cached) return symbolic_impl<RowT>(h, A, B, c_row_ptr, c_nnz);
      return symbolic_impl<RowT>(h, A, B, c_row_ptr, c_nnz, Batch{}, true);
    }
    h->spec.bytes = need;
  }
  int32_t *scol = (int32_t *)h->spec.base;
  V *sval = (V *)((char *)h->spec.base + cbytes);
  timing_mark(h, 4);
  const int ntf = h->num_table_factor;
  h->num_table_factor = 2;
  launch_num_bin<V>(h, bin, bin, MA, MA, h->a_rpt64, h->ab_b0, (const V *)A->val, h->ab_len, B->col,
                    (const V *)B->val, scol, sval, B->cols, B->nnz, (int)NS, h->row_nz);
  h->num_table_factor = ntf;
  timing_mark(h, 5);
  hipLaunchKernelGGL(k_scan<RowT>, dim3((unsigned)(MA / kScanTile + 1)), dim3(kScanBS), 0, h->stream, MA,
                     h->row_nz, (RowT *)c_row_ptr, h->c_rpt64, h->cls, h->tile_status, h->d_info);
  timing_mark(h, 3);
  if (into) {  // C's capacity is known: copy the slots now and read back once
    timing_mark(h, 6);
    launch_spec_compact(h, sizeof(V) == 8 ? NSP_F64 : NSP_F32, MA, NS, into->nnz, into->col, into->val);
    timing_mark(h, 7);
  }
  NSP_HIP(h, hipGetLastError());
  NSP_TRY(read_info(h));
  if (hi.spec_over) return symbolic_impl<RowT>(h, A, B, c_row_ptr, c_nnz);  // a row outgrew its slot
  if (hi.error) return set_error(h, (int)hi.error, "C has 2^31 or more entries; use NSP_IDX64");
  h->stats.nprod = hi.nprod;
  h->stats.nnz = hi.nnz;
  h->stats.max_intprod = cached ? -1 : hi.max_intprod;  // no per-row counts on the light row info
  h->stats.max_row_nnz = hi.max_nz;
  h->stats.n_sym_bins = kBins;
  h->stats.n_num_bins = kBins;
  for (int b = 0; b < kBins; b++) {
    h->stats.sym_bin[b] = hi.count_sym[b];
    h->stats.num_bin[b] = hi.count_num[b];
  }
  h->stats.sym_launches = 0;
  h->stats.num_launches = 1;
  h->stats.direct_mode = 2;
  h->stats.fused = 1;
  *c_nnz = hi.nnz;
  h->fc_bin = bin;
  h->fc_last = bin;
  std::copy(key, key + 4, h->fc_key);
  std::copy(dims, dims + 4, h->fc_dims);
  h->fused_ns = NS;
  h->fused_key[0] = A->col;
  h->fused_key[1] = B->col;
  h->plan_rows = MA;
  h->plan_nnz = hi.nnz;
  if (into) {
    *filled = true;
    h->timing_fused = h->timing;
    h->timing_pending = h->timing;
  } else {
    h->fused_plan = true;
  }
  return NSP_OK;
}

int spgemm_fused_count(Handle *h, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *A, const nsp_csr_t *B,
                       void *c_row_ptr, int64_t *c_nnz, nsp_csr_t *into, bool *filled) {
  NSP_TRY(check_csr_args(h, "A", A));
  NSP_TRY(check_csr_args(h, "B", B));
  if (vt != NSP_F64 && vt != NSP_F32) return set_error(h, NSP_ERR_INVALID, "unknown value type");
  if (A->cols != B->rows) return set_error(h, NSP_ERR_INVALID, "A.cols != B.rows");
  if (!c_row_ptr || !c_nnz) return set_error(h, NSP_ERR_INVALID, "c_row_ptr / c_nnz is NULL");
  if (it == NSP_IDX32) {
    if (A->nnz > INT32_MAX || B->nnz > INT32_MAX)
      return set_error(h, NSP_ERR_OVERFLOW, "nnz >= 2^31 needs NSP_IDX64");
    return vt == NSP_F64 ? fused_count_impl<int32_t, double>(h, A, B, c_row_ptr, c_nnz, into, filled)
                         : fused_count_impl<int32_t, float>(h, A, B, c_row_ptr, c_nnz, into, filled);
  }
  if (it == NSP_IDX64)
    return vt == NSP_F64 ? fused_count_impl<int64_t, double>(h, A, B, c_row_ptr, c_nnz, into, filled)
                         : fused_count_impl<int64_t, float>(h, A, B, c_row_ptr, c_nnz, into, filled);
  return set_error(h, NSP_ERR_INVALID, "unknown index type");
}

int spgemm_fused_fill(Handle *h, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *A, const nsp_csr_t *B,
                      nsp_csr_t *C) {
  if (!h->fused_plan) return spgemm_numeric(h, it, vt, A, B, C);
  h->fused_plan = false;
  if (!A || !B || !C) return set_error(h, NSP_ERR_INVALID, "NULL matrix");
  if (A->rows != h->plan_rows || A->col != h->fused_key[0] || B->col != h->fused_key[1])
    return set_error(h, NSP_ERR_STATE, "A/B differ from the counting call");
  if (C->nnz != h->plan_nnz) return set_error(h, NSP_ERR_INVALID, "C.nnz differs from the count");
  if (C->nnz > 0 && (!C->col || !C->val)) return set_error(h, NSP_ERR_INVALID, "C.col/C.val is NULL");
  timing_mark(h, 6);
  launch_spec_compact(h, vt, A->rows, h->fused_ns, C->nnz, C->col, C->val);
  timing_mark(h, 7);
  h->timing_fused = h->timing;
  h->timing_pending = h->timing;
  NSP_HIP(h, hipGetLastError());
  (void)it;
  return NSP_OK;
}

// ---------------------------------------------------------------------------
// Batched products: the plain two-phase path over concatenated rows.
static int check_batch_args(Handle *h, const nsp_csr_t *A, const int64_t *a_off, const nsp_csr_t *B,
                            const int64_t *b_off, int64_t count) {
  NSP_TRY(check_csr_args(h, "A", A));
  NSP_TRY(check_csr_args(h, "B", B));
  if (count < 0 || (count > 0 && (!a_off || !b_off)))
    return set_error(h, NSP_ERR_INVALID, "batched product: bad matrix offsets");
  if (count == 0 && A->rows > 0) return set_error(h, NSP_ERR_INVALID, "batched product: rows but no matrices");
  return NSP_OK;
}

int spgemm_batched_symbolic(Handle *h, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *A, const int64_t *a_off,
                            const nsp_csr_t *B, const int64_t *b_off, int64_t count, void *c_row_ptr,
                            int64_t *c_nnz) {
  NSP_TRY(check_batch_args(h, A, a_off, B, b_off, count));
  if (vt != NSP_F64 && vt != NSP_F32) return set_error(h, NSP_ERR_INVALID, "unknown value type");
  if (!c_row_ptr || !c_nnz) return set_error(h, NSP_ERR_INVALID, "c_row_ptr / c_nnz is NULL");
  const Batch bt{a_off, b_off, count};
  if (it == NSP_IDX32) return symbolic_impl<int32_t>(h, A, B, c_row_ptr, c_nnz, bt);
  if (it == NSP_IDX64) return symbolic_impl<int64_t>(h, A, B, c_row_ptr, c_nnz, bt);
  return set_error(h, NSP_ERR_INVALID, "unknown index type");
}

// ---------------------------------------------------------------------------
// Graph-captured product: binning, every bin's symbolic kernel, the scan, and
// every bin's numeric kernel recorded once into a hipGraph on a private
// workspace, with no host synchronisation inside.  Each kernel reads its
// bin's row count from device memory and walks it with a grid-stride loop
// sized for the worst case, so the graph stays valid for any pattern of the
// same row and nnz counts.  C's column/value buffers have a fixed capacity;
// rows that would pass it are skipped and reported.
struct GraphPlan {
  Handle hh;
  hipStream_t cap_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t launch_stream = nullptr;
  int64_t rows = 0;
};

template <typename RowT, typename V>
static int graph_record(Handle *g, const nsp_csr_t *A, const nsp_csr_t *B, const Batch &bt, nsp_csr_t *C) {
  const int64_t MA = A->rows;
  NSP_TRY(run_rowinfo<RowT>(g, A, B, bt, C->nnz));
  // grid sizes: the rows a bin may hold (every row), except for the
  // block-per-row bins, whose grids stop at the blocks a CU holds at once
  // (their LDS: symbolic 8 and 2, numeric 2 and 1 per CU), so that a bin
  // left empty by the pattern costs a short launch
  auto rows_hint = [&](int b, bool num) -> int64_t {
    const int64_t all = std::max<int64_t>(MA, 1);
    const int64_t per_cu = b == 6 ? (num ? 2 : 8) : b == 7 ? (num ? 1 : 2) : 0;
    return per_cu ? std::min<int64_t>(all, per_cu * g->num_cu) : all;
  };
  launch_scatter(g, MA, g->d_info->count_sym, g->d_info->cursor_sym, g->perm_sym);
  for (int b = 0; b < kBins; b++)
    launch_sym_bin(g, b, b, -1, rows_hint(b, false), g->a_rpt64, g->ab_b0, g->ab_len, B->col, B->nnz, B->cols);
  hipLaunchKernelGGL(k_scan<RowT>, dim3((unsigned)(MA / kScanTile + 1)), dim3(kScanBS), 0, g->stream, MA, g->row_nz,
                     (RowT *)C->row_ptr, g->c_rpt64, g->cls, g->tile_status, g->d_info);
  launch_scatter(g, MA, g->d_info->count_num, g->d_info->cursor_num, g->perm_num);
  int b0 = 0;
  if (queue_ok(g, B->cols)) {  // the wave-level bins in one work-queue launch
    launch_num_queue<V>(g, g->a_rpt64, g->ab_b0, (const V *)A->val, g->ab_len, B->col, (const V *)B->val, C->col,
                        (V *)C->val, B->nnz);
    b0 = kQueueBins;
  }
  for (int b = b0; b < kBins; b++)
    launch_num_bin<V>(g, b, b, -1, rows_hint(b, true), g->a_rpt64, g->ab_b0, (const V *)A->val, g->ab_len,
                      B->col, (const V *)B->val, C->col, (V *)C->val, B->cols, B->nnz);
  NSP_HIP(g, hipGetLastError());
  return NSP_OK;
}

int graph_create(Handle *h, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *A, const int64_t *a_off,
                 const nsp_csr_t *B, const int64_t *b_off, int64_t count, nsp_csr_t *C, GraphPlan **out) {
  if (!out || !C || !C->row_ptr) return set_error(h, NSP_ERR_INVALID, "graph: bad arguments");
  if (a_off) {
    NSP_TRY(check_batch_args(h, A, a_off, B, b_off, count));
  } else {
    NSP_TRY(check_csr_args(h, "A", A));
    NSP_TRY(check_csr_args(h, "B", B));
    if (A->cols != B->rows) return set_error(h, NSP_ERR_INVALID, "A.cols != B.rows");
  }
  if (A->rows < 1) return set_error(h, NSP_ERR_INVALID, "graph: A has no rows");
  if (C->nnz > 0 && (!C->col || !C->val)) return set_error(h, NSP_ERR_INVALID, "graph: C buffers are NULL");
  if (it == NSP_IDX32 && C->nnz > INT32_MAX) return set_error(h, NSP_ERR_OVERFLOW, "graph: capacity needs NSP_IDX64");
  GraphPlan *gp = new GraphPlan();
  Handle *g = &gp->hh;
  g->device = h->device;
  g->num_cu = h->num_cu;
  g->grid_cap = h->grid_cap;
  g->sym_small_rows_per_wave = h->sym_small_rows_per_wave;
  g->sym_rows_per_wave = h->sym_rows_per_wave;
  g->num_small_rows_per_wave = h->num_small_rows_per_wave;
  g->num_rows_per_wave = h->num_rows_per_wave;
  g->heavy_blocks = h->heavy_blocks;
  g->num_table_factor = h->num_table_factor;
  g->sym_table_factor = h->sym_table_factor;
  g->narrow_mode = h->narrow_mode;
  g->fp32_accumulate = h->fp32_accumulate;
  g->queue_mode = h->queue_mode;
  g->queue_chunk = h->queue_chunk;
  g->queue_blocks = h->queue_blocks;
  gp->rows = A->rows;
  gp->launch_stream = h->stream;
  auto fail = [&](int s) {
    h->last_error = g->last_error;
    if (gp->exec) (void)hipGraphExecDestroy(gp->exec);
    if (gp->graph) (void)hipGraphDestroy(gp->graph);
    if (g->ws.base) (void)hipFree(g->ws.base);
    if (g->d_info) (void)hipFree(g->d_info);
    if (g->h_info) (void)hipHostFree(g->h_info);
    if (gp->cap_stream) (void)hipStreamDestroy(gp->cap_stream);
    delete gp;
    return s;
  };
  if (hipMalloc((void **)&g->d_info, sizeof(Info)) != hipSuccess ||
      hipHostMalloc((void **)&g->h_info, sizeof(Info), hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    g->last_error = "graph: allocation failed";
    return fail(NSP_ERR_OOM);
  }
  if (hipStreamCreateWithFlags(&gp->cap_stream, hipStreamNonBlocking) != hipSuccess) return fail(NSP_ERR_HIP);
  g->stream = gp->cap_stream;
  // fix the workspace before capture: it must not move under the graph
  {
    int s = ensure_workspace(g, A->rows, A->nnz);
    if (s != NSP_OK) return fail(s);
  }
  if (hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return fail(NSP_ERR_HIP);
  const Batch bt{a_off, b_off, a_off ? count : 0};
  int s;
  if (it == NSP_IDX32) s = (vt == NSP_F64) ? graph_record<int32_t, double>(g, A, B, bt, C)
                                           : graph_record<int32_t, float>(g, A, B, bt, C);
  else s = (vt == NSP_F64) ? graph_record<int64_t, double>(g, A, B, bt, C)
                           : graph_record<int64_t, float>(g, A, B, bt, C);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(g->stream, &graph);
  gp->graph = graph;
  if (s != NSP_OK) return fail(s);
  if (ec != hipSuccess || !graph) {
    g->last_error = std::string("graph capture failed: ") + hipGetErrorString(ec);
    return fail(NSP_ERR_HIP);
  }
  if (hipGraphInstantiate(&gp->exec, gp->graph, nullptr, nullptr, 0) != hipSuccess) {
    g->last_error = "graph instantiation failed";
    return fail(NSP_ERR_HIP);
  }
  *out = gp;
  return NSP_OK;
}

int graph_launch(GraphPlan *gp) {
  NSP_HIP(&gp->hh, hipGraphLaunch(gp->exec, gp->launch_stream));
  return NSP_OK;
}

int graph_result(GraphPlan *gp, int64_t *nnz) {
  Handle *g = &gp->hh;
  NSP_HIP(g, hipMemcpyAsync(g->h_info, g->d_info, kInfoReadBytes, hipMemcpyDeviceToHost, gp->launch_stream));
  NSP_HIP(g, hipStreamSynchronize(gp->launch_stream));
  const Info &hi = *g->h_info;
  if (nnz) *nnz = hi.nnz;
  if (hi.error) return set_error(g, (int)hi.error, "graph: C exceeds the buffer capacity or the index width");
  return NSP_OK;
}

const char *graph_last_error(GraphPlan *gp) { return gp->hh.last_error.c_str(); }

int graph_destroy(GraphPlan *gp) {
  if (!gp) return NSP_OK;
  Handle *g = &gp->hh;
  if (gp->launch_stream) (void)hipStreamSynchronize(gp->launch_stream);
  if (gp->exec) (void)hipGraphExecDestroy(gp->exec);
  if (gp->graph) (void)hipGraphDestroy(gp->graph);
  if (g->ws.base) (void)hipFree(g->ws.base);
  if (g->d_info) (void)hipFree(g->d_info);
  if (g->h_info) (void)hipHostFree(g->h_info);
  if (gp->cap_stream) (void)hipStreamDestroy(gp->cap_stream);
  delete gp;
  return NSP_OK;
}

// ---------------------------------------------------------------------------
// Product without allocation (nsp_spgemm_noalloc): the graph path's kernel
// sequence (no host synchronisation between kernels; every bin's kernel
// reads its row count from device memory) run eagerly on the handle's
// stream, with the per-call counters and the workspace in `scratch`, a
// file-scope device array, instead of memory the library allocates.  Each
// device holds its own copy of the array (for the life of the process);
// its address on a device comes from hipGetSymbolAddress with that device
// current, looked up once per device (at nsp_create, or at the first call)
// and kept in g_scratch_ptr.  Host code never passes the array's name to a
// kernel: it is only a host-side stand-in for the symbol.  Calls on one
// device share the array: calls that overlap on one device race.
constexpr size_t kNoAllocScratch = (size_t)NSP_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratch];

constexpr int kMaxDevices = 64;
static void *g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_mu;

int noalloc_setup(Handle *h) {
  if (h->device < 0 || h->device >= kMaxDevices)
    return set_error(h, NSP_ERR_UNSUPPORTED, "nsp_spgemm_noalloc: device ids past 63 are not supported");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[h->device]) return NSP_OK;
  int cur = -1;
  NSP_HIP(h, hipGetDevice(&cur));
  if (cur != h->device) NSP_HIP(h, hipSetDevice(h->device));
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (cur != h->device) (void)hipSetDevice(cur);
  if (e != hipSuccess || !p)
    return set_error(h, NSP_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[h->device] = p;
  return NSP_OK;
}

template <typename RowT, typename V>
static int noalloc_impl(Handle *h, const nsp_csr_t *A, const nsp_csr_t *B, nsp_csr_t *C, int64_t *nnz) {
  NSP_TRY(noalloc_setup(h));
  char *base = (char *)g_scratch_ptr[h->device];
  constexpr size_t kInfoBytes = (sizeof(Info) + 255) & ~(size_t)255;
  const size_t need = kInfoBytes + workspace_bytes(A->rows, A->nnz);
  if (need > kNoAllocScratch)
    return set_error(h, NSP_ERR_UNSUPPORTED,
                     "nsp_spgemm_noalloc: A needs " + std::to_string(need) + " bytes of scratch, more than the " +
                         std::to_string(kNoAllocScratch) + " the library holds (use nsp_spgemm)");
  // a handle on the stack whose counters and workspace are the scratch
  // array (it owns no device memory and is not destroyed)
  Handle g;
  g.device = h->device;
  g.num_cu = h->num_cu;
  g.stream = h->stream;
  g.grid_cap = h->grid_cap;
  g.sym_small_rows_per_wave = h->sym_small_rows_per_wave;
  g.sym_rows_per_wave = h->sym_rows_per_wave;
  g.num_small_rows_per_wave = h->num_small_rows_per_wave;
  g.num_rows_per_wave = h->num_rows_per_wave;
  g.heavy_blocks = h->heavy_blocks;
  g.num_table_factor = h->num_table_factor;
  g.sym_table_factor = h->sym_table_factor;
  g.narrow_mode = h->narrow_mode;
  g.fp32_accumulate = h->fp32_accumulate;
  g.queue_mode = h->queue_mode;
  g.queue_chunk = h->queue_chunk;
  g.queue_blocks = h->queue_blocks;
  g.d_info = reinterpret_cast<Info *>(base);
  g.ws.base = base + kInfoBytes;
  g.ws.bytes = kNoAllocScratch - kInfoBytes;  // >= need: ensure_workspace only carves it
  const int s = graph_record<RowT, V>(&g, A, B, Batch{}, C);
  if (s != NSP_OK) {
    h->last_error = g.last_error;
    return s;
  }
  h->have_plan = false;  // the handle's own plan is untouched but no longer current
  if (!nnz) return NSP_OK;
  Info hi;
  NSP_HIP(h, hipMemcpyAsync(&hi, g.d_info, kInfoReadBytes, hipMemcpyDeviceToHost, h->stream));
  NSP_HIP(h, hipStreamSynchronize(h->stream));
  *nnz = hi.nnz;
  if (hi.error == NSP_ERR_OVERFLOW && sizeof(RowT) == 4 && hi.nnz > (int64_t)INT32_MAX)
    return set_error(h, NSP_ERR_OVERFLOW, "nsp_spgemm_noalloc: C has 2^31 or more entries; use NSP_IDX64");
  if (hi.error)
    return set_error(h, (int)hi.error,
                     "nsp_spgemm_noalloc: C has " + std::to_string(hi.nnz) + " entries, more than its capacity " +
                         std::to_string(C->nnz) + " (rows past it were skipped)");
  return NSP_OK;
}

int spgemm_noalloc(Handle *h, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *A, const nsp_csr_t *B,
                   nsp_csr_t *C, int64_t *nnz) {
  NSP_TRY(check_csr_args(h, "A", A));
  NSP_TRY(check_csr_args(h, "B", B));
  if (A->cols != B->rows) return set_error(h, NSP_ERR_INVALID, "A.cols != B.rows");
  if (!C || !C->row_ptr || C->nnz < 0 || (C->nnz > 0 && (!C->col || !C->val)))
    return set_error(h, NSP_ERR_INVALID, "nsp_spgemm_noalloc: C needs a row pointer and col/val of C->nnz entries");
  if (A->rows < 1) return set_error(h, NSP_ERR_INVALID, "nsp_spgemm_noalloc: A has no rows");
  if (vt != NSP_F64 && vt != NSP_F32) return set_error(h, NSP_ERR_INVALID, "unknown value type");
  if (it == NSP_IDX32) {
    if (A->nnz > INT32_MAX || B->nnz > INT32_MAX || C->nnz > INT32_MAX)
      return set_error(h, NSP_ERR_OVERFLOW, "nnz >= 2^31 needs NSP_IDX64");
    return vt == NSP_F64 ? noalloc_impl<int32_t, double>(h, A, B, C, nnz)
                         : noalloc_impl<int32_t, float>(h, A, B, C, nnz);
  }
  if (it == NSP_IDX64)
    return vt == NSP_F64 ? noalloc_impl<int64_t, double>(h, A, B, C, nnz)
                         : noalloc_impl<int64_t, float>(h, A, B, C, nnz);
  return set_error(h, NSP_ERR_INVALID, "unknown index type");
}

int spgemm_row_extents(Handle *h, nsp_index_t it, const nsp_csr_t *A, const nsp_csr_t *B, bool *recomputed) {
  NSP_TRY(check_csr_args(h, "A", A));
  NSP_TRY(check_csr_args(h, "B", B));
  if (A->cols != B->rows) return set_error(h, NSP_ERR_INVALID, "A.cols != B.rows");
  if (recomputed) *recomputed = false;
  if (h->reuse_extents && h->ext_key[0] == A->col && h->ext_key[1] == A->row_ptr && h->ext_key[2] == B->row_ptr &&
      h->ext_key[3] == nullptr && h->ext_rows == A->rows && h->ext_nnz == A->nnz && h->ext_nprod >= 0)
    return NSP_OK;
  h->have_plan = false;
  h->poff_valid = false;
  if (recomputed) *recomputed = true;
  if (it == NSP_IDX32) NSP_TRY(run_rowinfo<int32_t>(h, A, B));
  else NSP_TRY(run_rowinfo<int64_t>(h, A, B));
  NSP_TRY(read_info(h));
  h->ext_nprod = h->h_info->nprod;
  return NSP_OK;
}

int spgemm_intprod(Handle *h, nsp_index_t it, const nsp_csr_t *A, const nsp_csr_t *B, int64_t *intprod,
                   int64_t *nprod) {
  NSP_TRY(check_csr_args(h, "A", A));
  NSP_TRY(check_csr_args(h, "B", B));
  if (A->cols != B->rows) return set_error(h, NSP_ERR_INVALID, "A.cols != B.rows");
  h->have_plan = false;
  if (it == NSP_IDX32) NSP_TRY(run_rowinfo<int32_t>(h, A, B));
  else NSP_TRY(run_rowinfo<int64_t>(h, A, B));
  if (A->rows > 0 && intprod)
    NSP_HIP(h, hipMemcpyAsync(intprod, h->intprod, A->rows * 8, hipMemcpyDeviceToDevice, h->stream));
  NSP_TRY(read_info(h));
  if (nprod) *nprod = h->h_info->nprod;
  return NSP_OK;
}

}  // namespace nsp
