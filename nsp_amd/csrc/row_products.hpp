// Enumeration of a CSR row's intermediate products a_ik * b_kj by a group of
// GS lanes (GS = 16, 32 or 64), built for memory-level parallelism on gfx950.
//
// A entries are taken GS at a time, one per lane (column k, value a_ik, B row
// start and length), so the dependent loads A.col -> B.row_ptr cost one round
// trip per GS entries.  The group's products are then numbered in entry order
// and walked in windows of R*GS: in window w lane g handles products
// w + u*GS + g, u < R, and issues all R of its B loads before any of the R
// callbacks runs, keeping R B-row loads in flight per lane with consecutive
// lanes reading consecutive B entries.  A product's entry is found without a
// search: a per-window bit mask marks where entries start, one 64-bit word per
// GS products, and
//     entry = #(entries starting before the word) + popc(word & lanes<=g) - 1
// (entries with empty B rows are compacted away first).
#pragma once
#include "device_utils.hpp"

namespace nsp {

// Per-group LDS scratch of the enumerator.
template <int GS, int R, typename V>
struct EntryScratch {
  int64_t base[GS];          // B index of product 0 of entry e, minus e's first product number
  V av[GS];                  // a_ik of entry e
  unsigned long long mask[R];
};

template <int GS>
__device__ __forceinline__ uint64_t group_ballot(bool p) {
  const uint64_t b = __ballot(p);
  if constexpr (GS == 64) {
    return b;
  } else {
    return (b >> ((lane_id() / GS) * GS)) & ((1ull << GS) - 1);
  }
}

// f(key[R], product[R], ok[R]) is called once per window: lane g's R products
// of the window, ok[u] false past the end of the chunk.  VAL=false skips the
// values (symbolic pass; products are 0).
// Within a group all GS lanes must be active; different groups of a wave may
// be in different iterations.
// Entries are taken from chunks [c0, c0+GS), c0 = a0 + first, first + step, ...
// (step = GS for a group that owns the row; a block splits a row's chunks
// over its waves with first = wave*64, step = 64*waves).
template <int GS, int R, bool VAL, typename V, typename F>
__device__ __forceinline__ void row_products(const int glane, const int64_t a0, const int64_t a1,
                                             const int64_t first, const int64_t step,
                                             const int32_t *__restrict__ acol,
                                             const V *__restrict__ aval,
                                             const int64_t *__restrict__ brpt,
                                             const int32_t *__restrict__ bcol,
                                             const V *__restrict__ bval,
                                             EntryScratch<GS, R, V> *__restrict__ es, F &&f) {
  const uint64_t le = (glane == 63) ? ~0ull : ((2ull << glane) - 1);
  const uint64_t lt = (1ull << glane) - 1;
  for (int64_t c0 = a0 + first; c0 < a1; c0 += step) {
    const int64_t ja = c0 + glane;
    int64_t b0 = 0;
    int len = 0;
    V av = (V)0;
    if (ja < a1) {
      const int32_t k = acol[ja];
      if (VAL) av = aval[ja];
      b0 = brpt[k];
      len = (int)(brpt[k + 1] - b0);
    }
    const uint64_t nzm = group_ballot<GS>(len > 0);
    const int end = group_incl_scan<GS>(len, glane);
    const int start = end - len;
    const int T = __shfl(end, GS - 1, GS);
    if (len > 0) {
      const int cidx = __popcll(nzm & lt);
      es->base[cidx] = b0 - start;
      if (VAL) es->av[cidx] = av;
    }
    for (int w = 0; w < T; w += GS * R) {
      if (glane < R) es->mask[glane] = 0;
      wave_lds_sync();
      if (len > 0 && start >= w && start < w + GS * R)
        atomicOr(&es->mask[(start - w) / GS], 1ull << ((start - w) % GS));
      wave_lds_sync();
      int32_t key[R];
      V bv[R], a[R], prod[R];
      bool ok[R];
#pragma unroll
      for (int u = 0; u < R; u++) {
        const int p = w + u * GS + glane;
        const int before = __popcll(group_ballot<GS>(len > 0 && start < w + u * GS));
        ok[u] = p < T;
        key[u] = kEmpty;
        bv[u] = (V)0;
        a[u] = (V)0;
        if (ok[u]) {
          const int e = before + __popcll(es->mask[u] & le) - 1;
          const int64_t jb = es->base[e] + p;
          key[u] = bcol[jb];
          if (VAL) {
            bv[u] = bval[jb];
            a[u] = es->av[e];
          }
        }
      }
#pragma unroll
      for (int u = 0; u < R; u++) prod[u] = VAL ? a[u] * bv[u] : (V)0;
      f(key, prod, ok);
    }
    wave_lds_sync();
  }
}

}  // namespace nsp
