// Enumeration of a CSR row's intermediate products a_ik * b_kj by a group of
// GS lanes (GS = 8, 16, 32 or 64), built for memory-level parallelism on gfx950.
//
// A entries are taken GS at a time, one per lane.  Each entry's B row start
// and length come from arrays the binning pass already gathered (ab_b0,
// ab_len: independent loads, no dependent A.col -> B.row_ptr chain).  The
// group's products are numbered in entry order and walked in windows of R*GS:
// in a window lane g handles products w + u*GS + g, u < R.  Windows are
// software-pipelined: the B loads of window i+1 are issued before the
// callback of window i runs, so loads stay in flight while the LDS hash work
// of the previous window executes (for groups of 32 and 64 lanes the two
// windows alternate roles, so nothing is copied between them).  A product's
// entry is found without a search: a bit mask marks where entries start, one
// word per GS products, and
//     entry = pre[word] + popc(word & lanes<=g) - 1
// where pre[word] counts the entries starting before the word (entries with
// empty B rows are compacted away first).  Masks and counts cover a
// super-window of kMaskWords*GS products at a time; for GS <= 32 the count
// sits in the upper half of the mask word, so one 64-bit LDS read gives both.
#pragma once
#include "device_utils.hpp"

namespace nsp {

constexpr int kMaskWords = 32;

// Per-group LDS scratch of the enumerator.
template <int GS, typename V>
struct EntryScratch {
  int64_t base[GS];          // B index of entry e's first product minus that product's number
  V av[GS];                  // a_ik of entry e
  unsigned long long mask[kMaskWords];  // entry starts; GS <= 32: count before the word in bits 32..63
  int pre[GS == 64 ? kMaskWords : 1];   // GS = 64: count of entries starting before the word
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

// One window's worth of loaded products for one lane.
template <int R, typename V>
struct Window {
  int32_t key[R];
  V bv[R];
  int e[R];
  bool ok[R];
};

// A lane's A entry of a row's first chunk of GS entries (its B row start and
// length, and its value), loaded ahead of the row: the group kernels load
// the next row's first chunk while the current row is hashed and sorted, so
// a row starts on its B loads instead of on two dependent global round
// trips (row extent, then entry extents).
template <typename V>
struct FirstChunk {
  int64_t b0 = 0;
  int len = 0;
  V av = (V)0;
};

template <bool VAL, typename V>
__device__ __forceinline__ FirstChunk<V> load_first_chunk(int glane, int64_t a0, int64_t a1,
                                                         const int64_t *__restrict__ ab_b0,
                                                         const int32_t *__restrict__ ab_len,
                                                         const V *__restrict__ aval) {
  FirstChunk<V> f;
  const int64_t ja = a0 + glane;
  if (ja < a1) {
    f.b0 = ab_b0[ja];
    f.len = ab_len[ja];
    if (VAL) f.av = aval[ja];
  }
  return f;
}

// f(key[R], product[R], ok[R], bval[R], entry[R]) is called once per window
// with lane g's R products of the window; ok[u] is false past the end of the
// entries, entry[u] is the product's compacted entry index in the chunk.
// VAL = true loads B values and forms a_ik*b_kj; LOADB = true loads B values
// without A values (products are then 0); both false skips values (symbolic
// pass).  O32 = true (B with at most 2^29 entries) addresses B's entries
// with 32-bit byte offsets from the arrays' bases (saddr loads, no 64-bit
// address arithmetic per product).  Within a group all GS lanes must be
// active; different groups of a wave may be in different iterations.  Entries are taken from chunks
// [c0, c0+GS), c0 = a0 + first, first + step, ... (step = GS for a group that
// owns the row; a block splits a row's chunks over its waves with
// first = wave*64, step = 64*waves).
// With `pre` (first == 0), the row's first chunk comes from *pre instead of
// being loaded.
template <int GS, int R, bool VAL, typename V, bool LOADB = VAL, bool O32 = false, typename F>
__device__ __forceinline__ void row_products(const int glane, const int64_t a0, const int64_t a1,
                                             const int64_t first, const int64_t step,
                                             const int64_t *__restrict__ ab_b0,
                                             const int32_t *__restrict__ ab_len,
                                             const V *__restrict__ aval,
                                             const int32_t *__restrict__ bcol,
                                             const V *__restrict__ bval,
                                             EntryScratch<GS, V> *__restrict__ es, F &&f,
                                             const FirstChunk<V> *pre = nullptr) {
  static_assert(kMaskWords % R == 0, "a window must not straddle the mask words");
  const uint64_t le = (glane == 63) ? ~0ull : ((2ull << glane) - 1);
  const uint64_t lt = (1ull << glane) - 1;
  for (int64_t c0 = a0 + first; c0 < a1; c0 += step) {
    const int64_t ja = c0 + glane;
    int64_t b0 = 0;
    int len = 0;
    V av = (V)0;
    if (pre && c0 == a0) {
      b0 = pre->b0;
      len = pre->len;
      av = pre->av;
    } else if (ja < a1) {
      b0 = ab_b0[ja];
      len = ab_len[ja];
      if (VAL) av = aval[ja];
    }
    const uint64_t nzm = group_ballot<GS>(len > 0);
    const int end = group_incl_scan<GS>(len, glane);
    const int start = end - len;
    const int T = group_last<GS>(end);
    if (len > 0) {
      const int cidx = __popcll(nzm & lt);
      es->base[cidx] = b0 - start;
      if (VAL) es->av[cidx] = av;
    }
    for (int sw = 0; sw < T; sw += GS * kMaskWords) {
      // entry-start masks of this super-window
      for (int w = glane; w < kMaskWords; w += GS) es->mask[w] = 0;
      wave_lds_sync();
      if (len > 0 && start >= sw && start < sw + GS * kMaskWords)
        atomicOr(&es->mask[(start - sw) / GS], 1ull << ((start - sw) % GS));
      wave_lds_sync();
      // count of entries starting before each word: entries before the
      // super-window plus a group scan of the words' popcounts (WPL words
      // per lane, in order)
      {
        constexpr int WPL = kMaskWords > GS ? kMaskWords / GS : 1;
        const bool has = glane * WPL < kMaskWords;
        int cw[WPL], cnt = 0;
#pragma unroll
        for (int w = 0; w < WPL; w++) {
          cw[w] = has ? __popcll(es->mask[glane * WPL + w]) : 0;
          cnt += cw[w];
        }
        int run = __popcll(group_ballot<GS>(len > 0 && start < sw)) + group_incl_scan<GS>(cnt, glane) - cnt;
        if (has) {
#pragma unroll
          for (int w = 0; w < WPL; w++) {
            if constexpr (GS == 64)
              es->pre[glane * WPL + w] = run;
            else
              es->mask[glane * WPL + w] |= (unsigned long long)(uint32_t)run << 32;
            run += cw[w];
          }
        }
      }
      wave_lds_sync();
      const int tend = (T - sw < GS * kMaskWords) ? T - sw : GS * kMaskWords;
      const int nwin = (tend + GS * R - 1) / (GS * R);
      // Loads of window i: products sw + (i*R + u)*GS + glane.  Lanes past the
      // end load the chunk's first product instead, so every load is
      // unconditional and stays countable in flight.
      auto load = [&](int i, Window<R, V> &W) {
#pragma unroll
        for (int u = 0; u < R; u++) {
          const int q = i * R + u;
          const unsigned long long m = es->mask[q];
          int e;
          if constexpr (GS == 64)
            e = es->pre[q] + __popcll(m & le) - 1;
          else
            e = (int)(m >> 32) + __popc((uint32_t)m & (uint32_t)le) - 1;
          W.ok[u] = q * GS + glane < tend;
          e = W.ok[u] ? e : 0;
          if constexpr (O32) {
            const uint32_t jb = (uint32_t)es->base[e] + (uint32_t)(W.ok[u] ? sw + q * GS + glane : 0);
            W.key[u] = *reinterpret_cast<const int32_t *>(reinterpret_cast<const char *>(bcol) + (uint64_t)(jb * 4u));
            if (VAL || LOADB)
              W.bv[u] = *reinterpret_cast<const V *>(reinterpret_cast<const char *>(bval) +
                                                     (uint64_t)(jb * (uint32_t)sizeof(V)));
          } else {
            const int64_t jb = es->base[e] + (W.ok[u] ? sw + q * GS + glane : 0);
            W.key[u] = bcol[jb];
            if (VAL || LOADB) W.bv[u] = bval[jb];
          }
          W.e[u] = e;
        }
      };
      auto consume = [&](Window<R, V> &W) {
        V prod[R];
#pragma unroll
        for (int u = 0; u < R; u++) prod[u] = VAL ? es->av[W.e[u]] * W.bv[u] : (V)0;
        f(W.key, prod, W.ok, W.bv, W.e);
      };
      // Inside the loops the loads of the next window are issued with no
      // branch around them, so that the compiler waits for the current
      // window's loads alone (vmcnt of one window) instead of every load in
      // flight; only the last window or two, after the loop, wait for all.
      if constexpr (GS >= 32) {
        // long rows: the two windows alternate roles (no copies)
        Window<R, V> wa, wb;
        load(0, wa);
        int i = 0;
        for (; i + 2 < nwin; i += 2) {
          load(i + 1, wb);
          consume(wa);
          load(i + 2, wa);
          consume(wb);
        }
        if (i + 1 < nwin) load(i + 1, wb);
        consume(wa);
        if (i + 1 < nwin) consume(wb);
      } else {
        // short rows (mostly one window): fewer live registers, one more
        // wave per SIMD in the small bins
        Window<R, V> cur, nxt;
        load(0, cur);
        for (int i = 0; i + 1 < nwin; i++) {
          load(i + 1, nxt);
          consume(cur);
          cur = nxt;
        }
        consume(cur);
      }
    }
    wave_lds_sync();
  }
}

}  // namespace nsp

