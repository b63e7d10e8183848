// Backward pair of C = A*B on fixed sparsity patterns, and the CSR transpose
// it caches.  Given dC = dL/dC on C's pattern:
//   dA = (dC * B^T) sampled on pat(A):  dA_ik = sum_{j in B_k} B_kj dC_ij
//   dB = (A^T * dC) sampled on pat(B):  dB_kj = sum_{i in A^T_k} A_ik dC_ij
// Every (i, j) a forward product reaches lies in pat(C), so both are sums over
// the forward's intermediate products; no new structure is created.
//
// dA walks A's rows: C row i goes into an LDS hash (column -> dC_ij) and the
// lane of entry (i,k) sums B_kj * dC_ij over B row k in order.  dB walks B's
// entries with the cached A^T (rows sorted by i): for every (i, A_ik) of A^T
// row k it binary-searches dC_ij in C row i; lanes split A^T's row and a fixed
// shuffle tree adds their sums.  Both are deterministic.
//
// The reference has no backward; its only transpose is cuSPARSE csr2csc
// (src/CSR.cu:739-797).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_utils.hpp"
#include "nsp_internal.hpp"
#include "scan.hpp"
#include "segsort.hpp"
#include "spgemm_internal.hpp"

namespace nsp {

// ---------------------------------------------------------------------------
// transpose: count per column -> scan -> scatter -> sort each row of X^T by
// its (original row) index so that X^T is a sorted CSR, as csr2csc gives it.
__global__ void k_t_zero(int64_t n, int64_t *cnt, uint64_t *status, int64_t ntiles, uint32_t *ticket) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  for (int64_t i = t; i < n; i += (int64_t)gridDim.x * blockDim.x) cnt[i] = 0;
  for (int64_t i = t; i < ntiles; i += (int64_t)gridDim.x * blockDim.x) status[i] = 0;
  if (t == 0) ticket[0] = 0;
}

template <typename RowT>
__global__ void k_t_count(int64_t nnz, const int32_t *__restrict__ col, int64_t *__restrict__ cnt) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x)
    atomicAdd((unsigned long long *)&cnt[col[e]], 1ull);
}

template <typename RowT, typename V>
__global__ void k_t_scatter(int64_t rows, const RowT *__restrict__ rpt, const int32_t *__restrict__ col,
                            const V *__restrict__ val, const RowT *__restrict__ trpt,
                            int64_t *__restrict__ cursor, int32_t *__restrict__ tcol, V *__restrict__ tval) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < rows; i += (int64_t)gridDim.x * blockDim.x) {
    for (int64_t p = (int64_t)rpt[i]; p < (int64_t)rpt[i + 1]; p++) {
      const int32_t c = col[p];
      const int64_t d = (int64_t)trpt[c] + (int64_t)atomicAdd((unsigned long long *)&cursor[c], 1ull);
      tcol[d] = (int32_t)i;
      tval[d] = val[p];
    }
  }
}

// Rows of X^T are sorted by their (original row) index in three tiers:
// up to 64 entries by one wavefront in registers (k_t_sort_small), up to
// kTSortMax by one block in LDS (k_t_sort, rows listed by the first tier),
// longer ones by the merge sort of sort_listed_rows (listed as well).
constexpr int kTSortMax = 4096;

// One wavefront per row of up to 64 entries: a bitonic network over the 64
// lanes on keys (row index << 32 | position); values follow through a
// shuffle from the lane that held them.  Rows past 64 entries are listed.
template <typename RowT, typename V>
__global__ __launch_bounds__(256) void k_t_sort_small(int64_t trows, const RowT *__restrict__ trpt,
                                                      int32_t *__restrict__ tcol, V *__restrict__ tval,
                                                      unsigned long long *__restrict__ counts,
                                                      int64_t *__restrict__ mid_rows,
                                                      int64_t *__restrict__ long_rows) {
  const int lane = lane_id();
  const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t r = wave; r < trows; r += nwaves) {
    const int64_t p0 = (int64_t)trpt[r], n = (int64_t)trpt[r + 1] - p0;
    if (n <= 1) continue;
    if (n > 64) {
      if (lane == 0) {
        if (n > kTSortMax) long_rows[atomicAdd(&counts[1], 1ull)] = r;
        else mid_rows[atomicAdd(&counts[0], 1ull)] = r;
      }
      continue;
    }
    uint64_t key = ~0ull;
    V v = (V)0;
    if (lane < n) {
      key = ((uint64_t)(uint32_t)tcol[p0 + lane] << 32) | (uint32_t)lane;
      v = tval[p0 + lane];
    }
#pragma unroll
    for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
      for (int j = k >> 1; j > 0; j >>= 1) {
        const uint64_t o = __shfl_xor(key, j);
        const bool up = (lane & k) == 0, lower = (lane & j) == 0;
        key = (lower == up) ? (key < o ? key : o) : (key < o ? o : key);
      }
    }
    const V w = __shfl(v, (int)((uint32_t)key & 63));
    if (lane < n) {
      tcol[p0 + lane] = (int32_t)(key >> 32);
      tval[p0 + lane] = w;
    }
  }
}

// One block of 256 threads per listed row of 65..kTSortMax entries: bitonic
// sort of (row index, position) keys in LDS, values gathered by position.
template <typename RowT, typename V>
__global__ __launch_bounds__(256) void k_t_sort(const int64_t *__restrict__ rows,
                                                const unsigned long long *__restrict__ nrows,
                                                const RowT *__restrict__ trpt, int32_t *__restrict__ tcol,
                                                V *__restrict__ tval) {
  __shared__ uint64_t S[kTSortMax];  // (row index << 32) | position in the row
  __shared__ V Vt[kTSortMax];
  const int64_t count = (int64_t)*nrows;
  for (int64_t qi = blockIdx.x; qi < count; qi += gridDim.x) {
    const int64_t r = rows[qi];
    const int64_t p0 = (int64_t)trpt[r], n = (int64_t)trpt[r + 1] - p0;
    int n2 = 1;
    while (n2 < n) n2 <<= 1;
    for (int g = threadIdx.x; g < n2; g += 256) {
      if (g < n) {
        S[g] = ((uint64_t)(uint32_t)tcol[p0 + g] << 32) | (uint32_t)g;
        Vt[g] = tval[p0 + g];
      } else {
        S[g] = ~0ull;
      }
    }
    __syncthreads();
    for (int k = 2; k <= n2; k <<= 1) {
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int t = threadIdx.x; t < n2 / 2; t += 256) {
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
    for (int g = threadIdx.x; g < n; g += 256) {
      tcol[p0 + g] = (int32_t)(S[g] >> 32);
      tval[p0 + g] = Vt[(uint32_t)S[g]];
    }
    __syncthreads();
  }
}

static int aux2_reserve(Handle *h, size_t bytes) {
  if (bytes <= h->aux2.bytes) return NSP_OK;
  if (h->aux2.base) NSP_HIP(h, hipFree(h->aux2.base));
  h->aux2.base = nullptr;
  h->aux2.bytes = 0;
  if (hipMalloc(&h->aux2.base, bytes + bytes / 4) != hipSuccess) {
    (void)hipGetLastError();
    return set_error(h, NSP_ERR_OOM, "transpose workspace allocation failed");
  }
  h->aux2.bytes = bytes + bytes / 4;
  return NSP_OK;
}

template <typename RowT, typename V>
static int transpose_impl(Handle *h, const nsp_csr_t *X, nsp_csr_t *XT) {
  const int64_t n = X->cols, nnz = X->nnz;
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const int64_t ntiles = n / kScanTile + 2;
  const int64_t long_cap = nnz / (kTSortMax + 1) + 1;  // rows of X^T past kTSortMax
  const int64_t mid_cap = nnz / 65 + 1;                 // rows of X^T past 64 entries
  NSP_TRY(aux2_reserve(h, al((n + 1) * 8) + al(ntiles * 8) + 256 + al(long_cap * 8) + al(mid_cap * 8)));
  char *p = (char *)h->aux2.base;
  int64_t *cnt = (int64_t *)p;
  uint64_t *status = (uint64_t *)(p + al((n + 1) * 8));
  uint32_t *ticket = (uint32_t *)(p + al((n + 1) * 8) + al(ntiles * 8));
  unsigned long long *counts = (unsigned long long *)(ticket + 2);  // [0] mid rows, [1] long rows
  int64_t *long_rows = (int64_t *)(p + al((n + 1) * 8) + al(ntiles * 8) + 256);
  int64_t *mid_rows = (int64_t *)(p + al((n + 1) * 8) + al(ntiles * 8) + 256 + al(long_cap * 8));
  hipStream_t st = h->stream;
  const int g =(int)std::max<int64_t>(1, std::min<int64_t>((std::max(n, nnz) + 255) / 256, (int64_t)h->num_cu * 16));
  hipLaunchKernelGGL(k_t_zero, dim3(g), dim3(256), 0, st, n + 1, cnt, status, ntiles, ticket);
  hipLaunchKernelGGL((k_t_count<RowT>), dim3(g), dim3(256), 0, st, nnz, X->col, cnt);
  hipLaunchKernelGGL((k_exscan<int64_t, RowT>), dim3((unsigned)(n / kScanTile + 1)), dim3(kScanBS), 0, st, n, cnt,
                     (RowT *)XT->row_ptr, status, ticket);
  // reuse cnt as the scatter cursor; both row-list counts = 0
  hipLaunchKernelGGL(k_t_zero, dim3(g), dim3(256), 0, st, n + 1, cnt, (uint64_t *)counts, 2, ticket + 1);
  hipLaunchKernelGGL((k_t_scatter<RowT, V>), dim3(g), dim3(256), 0, st, X->rows, (const RowT *)X->row_ptr, X->col,
                     (const V *)X->val, (const RowT *)XT->row_ptr, cnt, XT->col, (V *)XT->val);
  const int gw = (int)std::max<int64_t>(1, std::min<int64_t>((n + 3) / 4, (int64_t)h->num_cu * 16));
  hipLaunchKernelGGL((k_t_sort_small<RowT, V>), dim3(gw), dim3(256), 0, st, n, (const RowT *)XT->row_ptr, XT->col,
                     (V *)XT->val, counts, mid_rows, long_rows);
  hipLaunchKernelGGL((k_t_sort<RowT, V>), dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>(mid_cap, (int64_t)h->num_cu * 8))),
                     dim3(256), 0, st, (const int64_t *)mid_rows, (const unsigned long long *)counts,
                     (const RowT *)XT->row_ptr, XT->col, (V *)XT->val);
  NSP_HIP(h, hipGetLastError());
  unsigned long long nl = 0;
  NSP_HIP(h, hipMemcpyAsync(&nl, counts + 1, 8, hipMemcpyDeviceToHost, st));
  NSP_HIP(h, hipStreamSynchronize(st));
  if (nl > 0) return sort_listed_rows<RowT, V>(h, long_rows, (int64_t)nl, (const RowT *)XT->row_ptr, XT->col, (V *)XT->val);
  return NSP_OK;
}

int transpose(Handle *h, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *X, nsp_csr_t *XT) {
  if (!X || !XT || !X->row_ptr || !XT->row_ptr || X->nnz != XT->nnz || XT->rows != X->cols ||
      (X->nnz > 0 && (!X->col || !X->val || !XT->col || !XT->val)))
    return set_error(h, NSP_ERR_INVALID, "nsp_transpose: bad arguments");
  if (it == NSP_IDX32 && vt == NSP_F64) return transpose_impl<int32_t, double>(h, X, XT);
  if (it == NSP_IDX32) return transpose_impl<int32_t, float>(h, X, XT);
  if (vt == NSP_F64) return transpose_impl<int64_t, double>(h, X, XT);
  return transpose_impl<int64_t, float>(h, X, XT);
}

// ---------------------------------------------------------------------------
// LDS hash of one C row with a table size chosen per row (the smallest power
// of two >= 2 nz, at least 16 slots), so short rows clear and probe small
// tables.  Same 24-bit multiplicative hash as the forward tables.
__device__ __forceinline__ uint32_t hash_slot_rt(int32_t key, int log) {
  return hash_mul24(key) >> (24 - log);
}

template <typename V>
__device__ __forceinline__ V hash_lookup_rt(const int32_t *keys, const V *vals, int32_t key, int log) {
  const uint32_t mask = (1u << log) - 1;
  uint32_t h = hash_slot_rt(key, log);
  for (uint32_t step = 0; step <= mask; step++) {  // bounded: a full table ends the walk
    const int32_t k = keys[h];
    if (k == key) return vals[h];
    if (k == kEmpty) return (V)0;
    h = (h + 1) & mask;
  }
  return (V)0;
}

__device__ __forceinline__ int table_log(int64_t nz) {
  int log = 4;
  while ((1ll << log) < 2 * nz) log++;
  return log;
}

// dA: GS lanes per A row, one lane per A entry.  C row i goes into an LDS
// table column -> dC sized for the row; lane (i,k) walks B row k in order and
// sums B_kj * dC_ij (deterministic).  A launch takes rows [0, MA) (row list
// NULL) or the rows of a device list; rows whose C row passes NS entries are
// appended to long_rows for the next tier.
template <typename V, typename RowT, int GS, int NS>
__global__ __launch_bounds__(256) void k_grad_a(int64_t MA, const int32_t *__restrict__ rows,
                                                const unsigned int *__restrict__ nrows,
                                                const RowT *__restrict__ arpt,
                                                const int64_t *__restrict__ ab_b0,
                                                const int32_t *__restrict__ ab_len,
                                                const int32_t *__restrict__ bcol, const V *__restrict__ bval,
                                                const RowT *__restrict__ crpt, const int32_t *__restrict__ ccol,
                                                const V *__restrict__ dc, V *__restrict__ da,
                                                const V *__restrict__ aval, const int64_t *__restrict__ poff,
                                                V *__restrict__ w,
                                                int32_t *__restrict__ long_rows, unsigned int *__restrict__ nlong) {
  constexpr int RPB = 256 / GS;
  constexpr int SH = 2 * NS;
  __shared__ int32_t keys[RPB][SH];
  __shared__ V vals[RPB][SH];
  const int slot = threadIdx.x / GS, glane = threadIdx.x % GS;
  int32_t *K = keys[slot];
  V *W = vals[slot];
  const int64_t count = rows ? (int64_t)*nrows : MA;
  for (int64_t t = blockIdx.x * (int64_t)RPB + slot; t < count; t += (int64_t)gridDim.x * RPB) {
    const int64_t i = rows ? rows[t] : t;
    const int64_t c0 = (int64_t)crpt[i], nz = (int64_t)crpt[i + 1] - c0;
    if (nz > NS) {
      if (glane == 0) long_rows[atomicAdd(nlong, 1u)] = (int32_t)i;
      continue;
    }
    const int log = table_log(nz);
    const uint32_t mask = (1u << log) - 1;
    for (int s = glane; s <= (int)mask; s += GS) K[s] = kEmpty;
    wave_lds_sync();
    for (int64_t q = glane; q < nz; q += GS) {
      const int32_t key = ccol[c0 + q];
      uint32_t hs = hash_slot_rt(key, log);
      for (uint32_t step = 0; step <= mask; step++) {  // nz < table size: a free slot exists
        if (atomicCAS(&K[hs], kEmpty, key) == kEmpty) break;
        hs = (hs + 1) & mask;
      }
      W[hs] = dc[c0 + q];
    }
    wave_lds_sync();
    const int64_t a0 = (int64_t)arpt[i], a1 = (int64_t)arpt[i + 1];
    // product offsets of the row's entries: the row's first one read once,
    // the rest by a group scan of the B row lengths (no per-entry gather)
    int64_t run = w ? poff[a0] : 0;
    for (int64_t c0 = a0; c0 < a1; c0 += GS) {
      const int64_t p = c0 + glane;
      const bool act = p < a1;
      const int len = act ? ab_len[p] : 0;
      const int incl = group_incl_scan<GS>(len, glane);
      const int64_t pofs = run + incl - len;
      run += group_last<GS>(incl);
      if (!act) continue;
      const int64_t b0 = ab_b0[p];
      V s = (V)0;
      if (w) {
        // also the product's share of dB: w[product] = A_ik dC_ij
        const V a = aval[p];
        V *wp = w + pofs;
        for (int j = 0; j < len; j++) {
          const V d = hash_lookup_rt<V>(K, W, bcol[b0 + j], log);
          s += bval[b0 + j] * d;
          wp[j] = a * d;
        }
      } else {
        for (int j = 0; j < len; j++) s += bval[b0 + j] * hash_lookup_rt<V>(K, W, bcol[b0 + j], log);
      }
      da[p] = s;
    }
    wave_lds_sync();
  }
}

// dA for rows whose C row passes the last table tier (more than 1024
// entries): one wavefront per listed row, one lane per A entry (i,k), and
// dC_ij found by binary search in the sorted C row i (no table, so no bound
// on the row length).  Same sums in the same order as k_grad_a.
template <typename V, typename RowT>
__global__ __launch_bounds__(256) void k_grad_a_search(const int32_t *__restrict__ rows,
                                                       const unsigned int *__restrict__ nrows,
                                                       const RowT *__restrict__ arpt,
                                                       const int64_t *__restrict__ ab_b0,
                                                       const int32_t *__restrict__ ab_len,
                                                       const int32_t *__restrict__ bcol, const V *__restrict__ bval,
                                                       const RowT *__restrict__ crpt, const int32_t *__restrict__ ccol,
                                                       const V *__restrict__ dc, V *__restrict__ da,
                                                       const V *__restrict__ aval, const int64_t *__restrict__ poff,
                                                       V *__restrict__ w) {
  const int lane = threadIdx.x & 63;
  const int64_t count = (int64_t)*nrows;
  for (int64_t t = blockIdx.x * 4 + (threadIdx.x >> 6); t < count; t += (int64_t)gridDim.x * 4) {
    const int64_t i = rows[t];
    const int64_t c0 = (int64_t)crpt[i], c1 = (int64_t)crpt[i + 1];
    const int64_t a0 = (int64_t)arpt[i], a1 = (int64_t)arpt[i + 1];
    for (int64_t p = a0 + lane; p < a1; p += 64) {
      const int64_t b0 = ab_b0[p];
      const int len = ab_len[p];
      const V a = w ? aval[p] : (V)0;
      V s = (V)0;
      int64_t lo = c0;  // B row k is sorted: each search starts at the previous hit
      for (int j = 0; j < len; j++) {
        const int32_t key = bcol[b0 + j];
        int64_t l = lo, r = c1;
        while (l < r) {
          const int64_t mid = (l + r) >> 1;
          if (ccol[mid] < key) l = mid + 1; else r = mid;
        }
        // key is in C row i whenever A_ik and B_kj are stored
        const V d = (l < c1 && ccol[l] == key) ? dc[l] : (V)0;
        lo = l;
        s += bval[b0 + j] * d;
        if (w) w[poff[p] + j] = a * d;
      }
      da[p] = s;
    }
  }
}

// dB: entry-parallel over B.  A block takes kGradBRows consecutive B rows and
// stages their row pointers in LDS; sub-groups of SG lanes take the chunk's B
// entries (k, j) in turn, and the SG lanes of a sub-group split A^T row k's
// entries t (lane l takes t = l, l+SG, ...), each binary-searching dC_ij in
// the sorted C row i.  Per-lane sums in t order, then a fixed shuffle tree:
// deterministic.  A's value of A^T entry t is AT.val[t], or A.val[perm[t]]
// when the permutation is given.
constexpr int kGradBRows = 64;

// Largest per-product contribution buffer the backward allocates (bytes);
// past it dB searches C's rows instead.
constexpr size_t kBackwardContribMax = (size_t)8 << 30;

template <typename V, typename RowT, int SG>
__global__ __launch_bounds__(256) void k_grad_b(int64_t MB, const RowT *__restrict__ brpt,
                                                const int32_t *__restrict__ bcol, const RowT *__restrict__ atrpt,
                                                const int32_t *__restrict__ atcol, const V *__restrict__ atval,
                                                const V *__restrict__ aval, const RowT *__restrict__ at_perm,
                                                const RowT *__restrict__ crpt, const int32_t *__restrict__ ccol,
                                                const V *__restrict__ dc, V *__restrict__ db) {
  constexpr int NSG = 256 / SG;
  __shared__ int64_t s_rpt[kGradBRows + 1];
  const int sg = threadIdx.x / SG, sl = threadIdx.x % SG;
  for (int64_t k0 = blockIdx.x * (int64_t)kGradBRows; k0 < MB; k0 += (int64_t)gridDim.x * kGradBRows) {
    const int nr = (int)(MB - k0 < kGradBRows ? MB - k0 : kGradBRows);
    for (int q = threadIdx.x; q <= nr; q += 256) s_rpt[q] = (int64_t)brpt[k0 + q];
    __syncthreads();
    const int64_t e0 = s_rpt[0], e1 = s_rpt[nr];
    // every sub-group runs the same number of rounds (the shuffle tree needs
    // all lanes of the wavefront)
    for (int64_t eb = e0; eb < e1; eb += NSG) {
      const int64_t p = eb + sg;
      const bool active = p < e1;
      V s = (V)0;
      if (active) {
        int lo = 0, hi = nr;
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (s_rpt[mid] <= p) lo = mid; else hi = mid;
        }
        const int64_t k = k0 + lo;
        const int32_t j = bcol[p];
        const int64_t t0 = (int64_t)atrpt[k], t1 = (int64_t)atrpt[k + 1];
        for (int64_t t = t0 + sl; t < t1; t += SG) {
          const int32_t i = atcol[t];
          const V a = at_perm ? aval[(int64_t)at_perm[t]] : atval[t];
          const int64_t end = (int64_t)crpt[i + 1];
          int64_t lo2 = (int64_t)crpt[i], hi2 = end;
          while (lo2 < hi2) {
            const int64_t mid = (lo2 + hi2) >> 1;
            if (ccol[mid] < j) lo2 = mid + 1; else hi2 = mid;
          }
          // j is in C row i whenever A_ik and B_kj are stored (C = A*B's pattern)
          if (lo2 < end && ccol[lo2] == j) s += a * dc[lo2];
        }
      }
      s = group_sum<SG>(s);
      if (active && sl == 0) db[p] = s;
    }
    __syncthreads();  // the next chunk reuses s_rpt
  }
}

// First product of the A entry behind each A^T entry, wpos[t] =
// poff[perm[t]]: computed once per pattern, so that the dB pass reads one
// coalesced array instead of gathering through the permutation.
template <typename RowT>
__global__ void k_wpos(int64_t n, const RowT *__restrict__ at_perm, const int64_t *__restrict__ poff,
                       int64_t *__restrict__ wpos) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    wpos[t] = poff[(int64_t)at_perm[t]];
}

// dB from the products' contributions w (written by k_grad_a, numbered in
// the forward's product order: A entry p owns products poff[p] ..
// poff[p] + nnz(B_k) - 1): dB_kq = sum over A^T row k of w[wpos[t] + q].
// Same chunking, lane split and fixed shuffle tree as k_grad_b.
template <typename V, typename RowT, int SG>
__global__ __launch_bounds__(256) void k_grad_b_sum(int64_t MB, const RowT *__restrict__ brpt,
                                                    const RowT *__restrict__ atrpt, const int64_t *__restrict__ wpos,
                                                    const V *__restrict__ w, V *__restrict__ db) {
  constexpr int NSG = 256 / SG;
  __shared__ int64_t s_rpt[kGradBRows + 1];
  const int sg = threadIdx.x / SG, sl = threadIdx.x % SG;
  for (int64_t k0 = blockIdx.x * (int64_t)kGradBRows; k0 < MB; k0 += (int64_t)gridDim.x * kGradBRows) {
    const int nr = (int)(MB - k0 < kGradBRows ? MB - k0 : kGradBRows);
    for (int q = threadIdx.x; q <= nr; q += 256) s_rpt[q] = (int64_t)brpt[k0 + q];
    __syncthreads();
    const int64_t e0 = s_rpt[0], e1 = s_rpt[nr];
    for (int64_t eb = e0; eb < e1; eb += NSG) {
      const int64_t p = eb + sg;
      const bool active = p < e1;
      V s = (V)0;
      if (active) {
        int lo = 0, hi = nr;
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (s_rpt[mid] <= p) lo = mid; else hi = mid;
        }
        const int64_t k = k0 + lo, q = p - s_rpt[lo];
        const int64_t t0 = (int64_t)atrpt[k], t1 = (int64_t)atrpt[k + 1];
        for (int64_t t = t0 + sl; t < t1; t += SG) s += w[wpos[t] + q];
      }
      s = group_sum<SG>(s);
      if (active && sl == 0) db[p] = s;
    }
    __syncthreads();
  }
}

__global__ void k_zero_words(uint32_t *p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

// Longest row of a CSR (its row pointer) into *out (zeroed before).
template <typename RowT>
__global__ __launch_bounds__(256) void k_row_maxlen(int64_t rows, const RowT *__restrict__ rpt,
                                                    unsigned long long *out) {
  unsigned long long m = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < rows; i += (int64_t)gridDim.x * blockDim.x) {
    const unsigned long long l = (unsigned long long)((int64_t)rpt[i + 1] - (int64_t)rpt[i]);
    m = l > m ? l : m;
  }
  m = group_max<64>(m);
  if ((threadIdx.x & 63) == 0 && m) atomicMax(out, m);
}

template <typename V, typename RowT>
static int backward_impl(Handle *h, const nsp_csr_t *A, const nsp_csr_t *B, const nsp_csr_t *AT,
                         const RowT *at_perm, const nsp_csr_t *C, const V *dc, V *da, V *db) {
  hipStream_t st = h->stream;
  // per-entry B row extents: the binning pass of the forward (reused when
  // the handle allows it and the buffers are the same)
  NSP_TRY(spgemm_row_extents(h, sizeof(RowT) == 4 ? NSP_IDX32 : NSP_IDX64, A, B));
  const int64_t MA = A->rows, nnzA = A->nnz, nprod = h->ext_nprod;
  // dB from per-product contributions when A^T's permutation is given and
  // the contributions fit the budget, else by search in C's rows
  const bool contrib_ok = at_perm && nprod > 0 && (size_t)nprod * sizeof(V) <= kBackwardContribMax;
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const int64_t ntiles = nnzA / kScanTile + 2;
  const size_t o_rows = 256, o_status = o_rows + 3 * al((size_t)std::max<int64_t>(MA, 1) * 4),
               o_poff = o_status + al((size_t)ntiles * 8), o_wpos = o_poff + al((size_t)(nnzA + 1) * 8),
               o_w = o_wpos + al((size_t)std::max<int64_t>(nnzA, 1) * 8),
               need = contrib_ok ? o_w + al((size_t)nprod * sizeof(V)) : o_poff;
  if (need > h->bwd.bytes) {
    h->poff_valid = false;
    if (h->bwd.base) NSP_HIP(h, hipFree(h->bwd.base));
    h->bwd.base = nullptr;
    h->bwd.bytes = 0;
    if (hipMalloc(&h->bwd.base, need + need / 8) != hipSuccess) {
      (void)hipGetLastError();
      return set_error(h, NSP_ERR_OOM, "backward workspace allocation failed");
    }
    h->bwd.bytes = need + need / 8;
  }
  char *base = (char *)h->bwd.base;
  // [1] rows for the second tier, [2] scan ticket, [3] rows for the third
  // tier, [4] rows for the search tier, [6..7] longest C row
  uint32_t *flags = (uint32_t *)base;
  const size_t list_bytes = al((size_t)std::max<int64_t>(MA, 1) * 4);
  int32_t *mid_rows = (int32_t *)(base + o_rows);
  int32_t *long_rows = (int32_t *)(base + o_rows + list_bytes);
  int32_t *search_rows = (int32_t *)(base + o_rows + 2 * list_bytes);
  // The longest C row decides which tiers run.  It is found once per C
  // pattern (one host synchronisation); with reuse_extents and the same C
  // buffers the cached value is used and the call does not synchronise.
  const bool c_cached = h->reuse_extents && h->bwd_c_key[0] == C->row_ptr && h->bwd_c_key[1] == C->col &&
                        h->bwd_c_rows == C->rows && h->bwd_c_nnz == C->nnz;
  if (!c_cached) {
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, st, flags, 8);
    if (C->rows > 0)
      hipLaunchKernelGGL(k_row_maxlen<RowT>, dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((C->rows + 255) / 256,
                                                                                           (int64_t)h->num_cu * 4))),
                         dim3(256), 0, st, C->rows, (const RowT *)C->row_ptr, (unsigned long long *)(flags + 6));
    unsigned long long cm = 0;
    NSP_HIP(h, hipMemcpyAsync(&cm, flags + 6, 8, hipMemcpyDeviceToHost, st));
    NSP_HIP(h, hipStreamSynchronize(st));
    h->bwd_cmax = (int64_t)cm;
    h->bwd_c_key[0] = C->row_ptr;
    h->bwd_c_key[1] = C->col;
    h->bwd_c_rows = C->rows;
    h->bwd_c_nnz = C->nnz;
  }
  const int64_t cmax = h->bwd_cmax;
  // dB from the products' contributions (written by dA's pass; "bwd_contrib"
  // 1, the default) or by search in C's rows (0), or contributions only when
  // C has rows past 32 entries (2).  On the 64^3 stencil times the
  // aggregation prolongator (C rows of at most 18 entries) the contributions
  // measured faster: fwd+bwd 0.282 ms against 0.313 ms by search.
  const bool contrib = contrib_ok && (h->bwd_contrib == 1 || (h->bwd_contrib == 2 && cmax > 32));
  int64_t *poff = contrib ? (int64_t *)(base + o_poff) : nullptr;
  V *w = contrib ? (V *)(base + o_w) : nullptr;
  // the row lists of tiers 2-4 and the scan ticket start at zero
  if (cmax > 32 || (contrib && !h->poff_valid)) hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, st, flags, 5);
  if (contrib && !h->poff_valid) {
    // product offsets of the A entries: exclusive scan of the B row lengths
    uint64_t *status = (uint64_t *)(base + o_status);
    hipLaunchKernelGGL(k_t_zero, dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((ntiles + 255) / 256, 1024))),
                       dim3(256), 0, st, (int64_t)0, (int64_t *)nullptr, status, ntiles, flags + 2);
    hipLaunchKernelGGL((k_exscan<int32_t, int64_t>), dim3((unsigned)(nnzA / kScanTile + 1)), dim3(kScanBS), 0, st,
                       nnzA, h->ab_len, poff, status, flags + 2);
    h->poff_valid = true;
    h->wpos_key = nullptr;
  }
  int64_t *wpos = contrib ? (int64_t *)(base + o_wpos) : nullptr;
  if (contrib && h->wpos_key != (const void *)at_perm && nnzA > 0) {
    hipLaunchKernelGGL(k_wpos<RowT>, dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((nnzA + 255) / 256,
                                                                                          (int64_t)h->num_cu * 16))),
                       dim3(256), 0, st, nnzA, at_perm, (const int64_t *)poff, wpos);
    h->wpos_key = at_perm;
  }
  if (MA > 0) {
    // tier 1: C rows up to 32 entries (16 lanes per row, 64-slot tables:
    // twice the rows in flight of 32 lanes; on the 64^3 stencil times the
    // aggregation prolongator fwd+bwd 0.228 -> 0.216 ms in FP32, 8 lanes
    // slower); longer rows go to a device list for tier 2 (up to 128
    // entries), whose longer rows go to one for tier 3 (up to 1024 entries,
    // 64 lanes), whose longer rows go to the search tier (any length).  The
    // small first tier keeps the LDS per block low (occupancy) for the
    // common short C rows.
    const int ga = (int)std::max<int64_t>(1, std::min<int64_t>((MA + 15) / 16, (int64_t)h->num_cu * 32));
    hipLaunchKernelGGL((k_grad_a<V, RowT, 16, 32>), dim3(ga), dim3(256), 0, st, MA, (const int32_t *)nullptr,
                       (const unsigned int *)nullptr, (const RowT *)A->row_ptr, h->ab_b0, h->ab_len, B->col,
                       (const V *)B->val, (const RowT *)C->row_ptr, C->col, dc, da, (const V *)A->val,
                       (const int64_t *)poff, w, mid_rows, flags + 1);
    if (cmax > 32)
      hipLaunchKernelGGL((k_grad_a<V, RowT, 32, 128>), dim3(h->num_cu * 2), dim3(256), 0, st, MA,
                         (const int32_t *)mid_rows, (const unsigned int *)(flags + 1), (const RowT *)A->row_ptr,
                         h->ab_b0, h->ab_len, B->col, (const V *)B->val, (const RowT *)C->row_ptr, C->col, dc, da,
                         (const V *)A->val, (const int64_t *)poff, w, long_rows, flags + 3);
    if (cmax > 128)
      hipLaunchKernelGGL((k_grad_a<V, RowT, 64, 1024>), dim3(h->num_cu), dim3(256), 0, st, MA,
                         (const int32_t *)long_rows, (const unsigned int *)(flags + 3), (const RowT *)A->row_ptr,
                         h->ab_b0, h->ab_len, B->col, (const V *)B->val, (const RowT *)C->row_ptr, C->col, dc, da,
                         (const V *)A->val, (const int64_t *)poff, w, search_rows, flags + 4);
    if (cmax > 1024)
      hipLaunchKernelGGL((k_grad_a_search<V, RowT>), dim3(h->num_cu), dim3(256), 0, st, (const int32_t *)search_rows,
                         (const unsigned int *)(flags + 4), (const RowT *)A->row_ptr, h->ab_b0, h->ab_len, B->col,
                         (const V *)B->val, (const RowT *)C->row_ptr, C->col, dc, da, (const V *)A->val,
                         (const int64_t *)poff, w);
  }
  if (B->rows > 0 && B->nnz > 0) {
    // lanes per B entry from the mean A^T row length
    const int64_t avg = A->cols > 0 ? (A->nnz + A->cols - 1) / A->cols : 1;
    const int g = (int)std::max<int64_t>(
        1, std::min<int64_t>((B->rows + kGradBRows - 1) / kGradBRows, (int64_t)h->num_cu * 32));
#define GRAD_B(SG)                                                                                            \
  if (contrib)                                                                                                \
    hipLaunchKernelGGL((k_grad_b_sum<V, RowT, SG>), dim3(g), dim3(256), 0, st, B->rows,                       \
                       (const RowT *)B->row_ptr, (const RowT *)AT->row_ptr, (const int64_t *)wpos, (const V *)w, \
                       db);                                                                                   \
  else                                                                                                        \
    hipLaunchKernelGGL((k_grad_b<V, RowT, SG>), dim3(g), dim3(256), 0, st, B->rows, (const RowT *)B->row_ptr, \
                       B->col, (const RowT *)AT->row_ptr, AT->col, (const V *)AT->val, (const V *)A->val,     \
                       at_perm, (const RowT *)C->row_ptr, C->col, dc, db)
    if (avg <= 4) { GRAD_B(4); }
    else if (avg <= 8) { GRAD_B(8); }
    else if (avg <= 16) { GRAD_B(16); }
    else if (avg <= 32) { GRAD_B(32); }
    else { GRAD_B(64); }
#undef GRAD_B
  }
  NSP_HIP(h, hipGetLastError());
  return NSP_OK;
}

int spgemm_backward(Handle *h, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *A, const nsp_csr_t *B,
                    const nsp_csr_t *AT, const void *at_perm, const nsp_csr_t *C, const void *dc, void *da,
                    void *db) {
  if (!A || !B || !AT || !C || !A->row_ptr || !B->row_ptr || !AT->row_ptr || !C->row_ptr ||
      A->cols != B->rows || AT->rows != A->cols || AT->nnz != A->nnz || C->rows != A->rows)
    return set_error(h, NSP_ERR_INVALID, "nsp_spgemm_backward: bad arguments");
  if ((A->nnz && !da) || (B->nnz && !db) || (C->nnz && !dc))
    return set_error(h, NSP_ERR_INVALID, "nsp_spgemm_backward: NULL gradient buffer");
  if (AT->nnz && (!AT->col || (!AT->val && !at_perm)))
    return set_error(h, NSP_ERR_INVALID, "nsp_spgemm_backward: A^T needs its columns and values or a permutation");
  if (it == NSP_IDX32 && vt == NSP_F64)
    return backward_impl<double, int32_t>(h, A, B, AT, (const int32_t *)at_perm, C, (const double *)dc,
                                          (double *)da, (double *)db);
  if (it == NSP_IDX32)
    return backward_impl<float, int32_t>(h, A, B, AT, (const int32_t *)at_perm, C, (const float *)dc, (float *)da,
                                         (float *)db);
  if (vt == NSP_F64)
    return backward_impl<double, int64_t>(h, A, B, AT, (const int64_t *)at_perm, C, (const double *)dc,
                                          (double *)da, (double *)db);
  return backward_impl<float, int64_t>(h, A, B, AT, (const int64_t *)at_perm, C, (const float *)dc, (float *)da,
                                       (float *)db);
}

}  // namespace nsp
