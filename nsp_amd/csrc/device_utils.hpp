// Device-side building blocks for the gfx950 SpGEMM kernels: wave64 group
// reductions and scans, LDS open-addressing hash tables, and register
// bitonic sorts.  Everything here is written for 64-wide wavefronts.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace nsp {

constexpr int kWave = 64;
constexpr int32_t kEmpty = -1;           // empty hash slot (columns are >= 0)
constexpr uint32_t kPadKey = 0xFFFFFFFFu; // sorts after every column

template <int N> struct Log2 { static constexpr int value = 1 + Log2<N / 2>::value; };
template <> struct Log2<1> { static constexpr int value = 0; };

// Relaxed LDS read of a hash slot that other lanes may CAS concurrently (an
// atomic load keeps it a ds_read; a volatile access becomes a flat load that
// waits on every outstanding global load).
__device__ __forceinline__ int32_t lds_ld(const int32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint32_t lds_ld(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Orders LDS accesses of one wavefront.  A wave's LDS instructions execute in
// issue order, so a compiler barrier is all that is needed between a write by
// one lane and a read by another lane of the same wave.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Tiles of a work list in XCD-contiguous order.  The hardware deals
// workgroups to the XCDs round-robin (block b runs on XCD b % 8, and each XCD
// has its own L2), so block b takes tiles from XCD (b % 8)'s contiguous share
// of the list, every (blocks on that XCD)-th one.  The tiles in flight on one
// XCD are then neighbours, and so are the B rows they read (neighbouring rows
// of a stencil or FEM matrix share most of theirs), which keeps those reads
// in that XCD's L2.  Every tile is visited once for any grid size.
struct XcdTiles {
  int64_t first, end, step;
  __device__ __forceinline__ explicit XcdTiles(int64_t ntiles) {
    const int nx = gridDim.x < 8 ? (int)gridDim.x : 8;
    const int x = (int)(blockIdx.x % nx);
    const int64_t j = blockIdx.x / nx;
    step = ((int64_t)gridDim.x - x + nx - 1) / nx;  // blocks on this XCD
    first = ntiles * x / nx + j;
    end = ntiles * (x + 1) / nx;
  }
};

// Multiplicative (Fibonacci) hashing in 24-bit arithmetic: slot = the top LOG
// bits of (key * floor(2^24/phi)) mod 2^24, with the full-rate 24-bit multiply
// (v_mul_u32_u24) instead of the quarter-rate 32-bit one and one bit-field
// extract.  Arithmetic progressions of columns (stencil rows: runs at strides
// 1, m, m^2) spread over the table instead of aliasing as under key & (SH-1).
// Only the low 24 bits of a column enter the hash; columns 2^24 apart collide,
// which costs probes, never correctness.
// The multiply is written as the instruction itself: with only the low 24
// bits of the product used, hipcc otherwise drops __umul24's operand mask and
// emits the quarter-rate v_mul_lo_u32 (same low bits, four times the cycles).
// (volatile: the compiler would otherwise sink it into the branches of its
// users' selects instead of computing it unconditionally)
__device__ __forceinline__ uint32_t hash_mul24(int32_t key) {
  uint32_t p;
  asm volatile("v_mul_u32_u24 %0, 0x9e3779, %1" : "=v"(p) : "v"(key));
  return p & 0xFFFFFFu;
}

template <int LOG>
__device__ __forceinline__ uint32_t hash_slot(int32_t key) {
  return hash_mul24(key) >> (24 - LOG);
}

constexpr int32_t kDummy = -2;  // key of a lane's private dummy slot

// Continues the linear probe of `key` after slot h (its home slot or a slot
// already found taken by another key); returns the key's slot and sets
// inserted = 1 when this call claimed the slot.
template <int LOG>
__device__ __forceinline__ uint32_t hash_probe_from(int32_t *keys, int32_t key, uint32_t h, int &inserted) {
  constexpr uint32_t MASK = (1u << LOG) - 1;
  // bounded: a full table (impossible when the bin bounds hold) ends the walk
  for (uint32_t step = 0; step <= MASK; step++) {
    h = (h + 1) & MASK;
    const int32_t k = __hip_atomic_load(&keys[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (k == key) return h;
    if (k == kEmpty) {
      const int32_t old = atomicCAS(&keys[h], kEmpty, key);
      if (old == kEmpty) {
        inserted = 1;
        return h;
      }
      if (old == key) return h;
    }
  }
  return h;
}

// Batched inserts of R keys per lane (one window of products).  The first
// probe of every key is an unconditional CAS on its home slot, all R issued
// back to back; only keys that find their home slot taken by another key walk
// on.  Lanes with ok[u] false use their private dummy slot SH + glane (tables
// hold SH + GS slots), so no branch is needed on the common path.
// Returns the number of keys this lane inserted for the first time.
template <int LOG, int R>
__device__ __forceinline__ int hash_insert_keys(int32_t *keys, const int32_t (&key)[R], const bool (&ok)[R],
                                                int glane) {
  constexpr uint32_t SH = 1u << LOG;
  int32_t kk[R], old[R];
  uint32_t h[R];
#pragma unroll
  for (int u = 0; u < R; u++) {
    const uint32_t hs = hash_slot<LOG>(key[u]);
    kk[u] = ok[u] ? key[u] : kDummy;
    h[u] = ok[u] ? hs : SH + glane;
  }
#pragma unroll
  for (int u = 0; u < R; u++) old[u] = atomicCAS(&keys[h[u]], kEmpty, kk[u]);
  int cnt = 0;
#pragma unroll
  for (int u = 0; u < R; u++) {
    int ins = (old[u] == kEmpty) & ok[u];
    if (old[u] != kEmpty && old[u] != kk[u]) hash_probe_from<LOG>(keys, kk[u], h[u], ins);
    cnt += ins;
  }
  return cnt;
}

// Batched accumulate: finds or inserts every key as above, then adds the R
// products into the value slots (LDS atomic adds).
template <int LOG, int R, typename V>
__device__ __forceinline__ void hash_accumulate_batch(int32_t *keys, V *vals, const int32_t (&key)[R],
                                                      const V (&prod)[R], const bool (&ok)[R], int glane) {
  constexpr uint32_t SH = 1u << LOG;
  int32_t kk[R], old[R];
  uint32_t h[R];
#pragma unroll
  for (int u = 0; u < R; u++) {
    const uint32_t hs = hash_slot<LOG>(key[u]);
    kk[u] = ok[u] ? key[u] : kDummy;
    h[u] = ok[u] ? hs : SH + glane;
  }
  // a product slot that no lane of the wavefront fills (the tail of a row's
  // last window) issues neither the swap nor the add
  bool live[R];
#pragma unroll
  for (int u = 0; u < R; u++) live[u] = __ballot(ok[u]) != 0;
#pragma unroll
  for (int u = 0; u < R; u++) old[u] = live[u] ? atomicCAS(&keys[h[u]], kEmpty, kk[u]) : kk[u];
#pragma unroll
  for (int u = 0; u < R; u++) {
    int ins = 0;
    if (old[u] != kEmpty && old[u] != kk[u]) h[u] = hash_probe_from<LOG>(keys, kk[u], h[u], ins);
  }
  // lanes with ok[u] false add into their dummy slot, which no reader visits
#pragma unroll
  for (int u = 0; u < R; u++)
    if (live[u]) __hip_atomic_fetch_add(&vals[h[u]], prod[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// ---- group (sub-wave) collectives; GS lanes, GS a power of two <= 64 ------
// The 32-bit value of the group's last lane in every lane of the group,
// without an LDS round trip: DPP row_share (row_newbcast) for groups within
// a 16-lane row, v_readlane for groups of 32 and 64.
template <int GS>
__device__ __forceinline__ int group_last(int x) {
  if constexpr (GS == 8) {
    const int lo = __builtin_amdgcn_update_dpp(0, x, 0x157, 0xF, 0xF, false);  // lane 7 of the row
    const int hi = __builtin_amdgcn_update_dpp(0, x, 0x15F, 0xF, 0xF, false);  // lane 15 of the row
    return (__lane_id() & 8) ? hi : lo;
  } else if constexpr (GS == 16) {
    return __builtin_amdgcn_update_dpp(0, x, 0x15F, 0xF, 0xF, false);
  } else if constexpr (GS == 32) {
    const int a = __builtin_amdgcn_readlane(x, 31), b = __builtin_amdgcn_readlane(x, 63);
    return __lane_id() < 32 ? a : b;
  } else {
    return __builtin_amdgcn_readlane(x, 63);
  }
}

template <int GS, int RL = 0>
__device__ __forceinline__ int group_incl_scan_dpp(int v, int glane);

template <int GS, typename T>
__device__ __forceinline__ T group_sum(T x) {
  if constexpr (sizeof(T) == 4 && std::is_integral<T>::value && GS >= 8) {
    // an inclusive scan (DPP) and the last lane's total
    const int glane = (int)(__lane_id() % GS);
    return (T)group_last<GS>(group_incl_scan_dpp<GS>((int)x, glane));
  } else {
#pragma unroll
    for (int o = GS / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, GS);
    return x;
  }
}

template <int GS, typename T>
__device__ __forceinline__ T group_max(T x) {
#pragma unroll
  for (int o = GS / 2; o > 0; o >>= 1) {
    T y = __shfl_xor(x, o, GS);
    x = x > y ? x : y;
  }
  return x;
}

// inclusive scan within groups of GS lanes (GS >= 2).  32-bit integers use DPP lane
// moves (VALU, no LDS round trip): row_shr:1,2,4,8 within each 16-lane row,
// then row_bcast:15 (lane 15 of a row added into the next row, rows 1 and 3)
// and row_bcast:31 (lane 31 into rows 2 and 3) to join rows into groups of 32
// and 64.  Other types go through shuffles.
template <int GS, int RL>
__device__ __forceinline__ int group_incl_scan_dpp(int v, int glane) {
  // rows of 16 lanes hold several groups when GS < 16
  auto step = [&](int y, int o) { v += (GS < 16 && glane < o) ? 0 : y; };
  step(__builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, false), 1);  // row_shr:1
  if constexpr (GS > 2) step(__builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, false), 2);
  if constexpr (GS > 4) step(__builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, false), 4);
  if constexpr (GS > 8) step(__builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, false), 8);
  if constexpr (GS >= 32) v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, false);  // row_bcast:15
  if constexpr (GS == 64) v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, false);  // row_bcast:31
  return v;
}

template <int GS, typename T>
__device__ __forceinline__ T group_incl_scan(T x, int glane) {
  if constexpr (sizeof(T) == 4 && std::is_integral<T>::value) {
    return (T)group_incl_scan_dpp<GS>((int)x, glane);
  } else {
#pragma unroll
    for (int o = 1; o < GS; o <<= 1) {
      T y = __shfl_up(x, o, GS);
      if (glane >= o) x += y;
    }
    return x;
  }
}

// ---- LDS hash tables -------------------------------------------------------
// Value of a key known to be present.
template <int LOG, typename V>
__device__ __forceinline__ V hash_lookup(const int32_t* keys, const V* vals, int32_t key) {
  constexpr uint32_t MASK = (1u << LOG) - 1;
  uint32_t h = hash_slot<LOG>(key);
  for (uint32_t step = 0; step <= MASK && keys[h] != key; step++) h = (h + 1) & MASK;
  return vals[h];
}

// Runtime-sized variants (window kernels pick the table size per launch).
__device__ __forceinline__ int hash_insert_key_rt(int32_t* keys, int32_t key, int log2sh) {
  const uint32_t mask = (1u << log2sh) - 1;
  uint32_t h = ((uint32_t)key * 0x9E3779B1u) >> (32 - log2sh);
  while (true) {
    int32_t k = lds_ld(&keys[h]);
    if (k == key) return 0;
    if (k == kEmpty) {
      int32_t old = atomicCAS(&keys[h], kEmpty, key);
      if (old == kEmpty) return 1;
      if (old == key) return 0;
    }
    h = (h + 1) & mask;
  }
}

// Lane exchange x -> lane ^ X for a constant X: DPP quad permutations for 1 and
// 2, ds_swizzle (xor mode, within 32 lanes, no memory access) for 4..16,
// ds_bpermute for 32.
__device__ __forceinline__ int shfl_xor_const(int X, int v) {
  switch (X) {
    case 1: return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
    case 2: return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false);   // quad_perm [2,3,0,1]
    case 4: return __builtin_amdgcn_ds_swizzle(v, (4 << 10) | 0x1F);
    case 8: return __builtin_amdgcn_ds_swizzle(v, (8 << 10) | 0x1F);
    case 16: return __builtin_amdgcn_ds_swizzle(v, (16 << 10) | 0x1F);
    default: return __shfl_xor(v, X, 64);
  }
}

// ---- register bitonic sort over a group ------------------------------------
// GS lanes hold E keys each (blocked layout: element g = glane*E + r); sorts
// the GS*E keys ascending.  GS and E powers of two; all GS lanes active.
// Phase k merges runs of k keys, ascending where (g & k) == 0 and descending
// elsewhere.  Keys of descending runs are held complemented (~x reverses the
// order of uint32) for the whole phase, so every compare-exchange puts the
// minimum at the lower index: an in-lane pair is one min and one max, a
// cross-lane one a min, a max and a select, with no per-element direction.
// Leaving phase k and entering phase 2k is one xor per key.
template <int GS, int E>
__device__ __forceinline__ void group_bitonic_sort(uint32_t (&v)[E], int glane) {
  constexpr int N = GS * E;
  // The lane masks below depend on the lane only; an opaque copy of the lane
  // index keeps the compiler from hoisting them out of the caller's row loop,
  // where they would stay live (registers) across the whole kernel.
  int gl = glane;
  asm volatile("" : "+v"(gl));
  // complement mask of element gl*E + r in phase kk (none in the last)
  auto flip = [&](int kk, int r) -> uint32_t {
    return (kk >= 2 && kk < N && ((gl * E + r) & kk)) ? ~0u : 0u;
  };
#pragma unroll
  for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
    for (int r = 0; r < E; r++) v[r] ^= flip(k, r) ^ flip(k >> 1, r);
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (j < E) {
#pragma unroll
        for (int r = 0; r < E; r++) {
          if ((r & j) == 0) {
            const uint32_t a = v[r], b = v[r | j];
            v[r] = a < b ? a : b;
            v[r | j] = a < b ? b : a;
          }
        }
      } else {
        const bool lower = ((gl * E) & j) == 0;
#pragma unroll
        for (int r = 0; r < E; r++) {
          const uint32_t o = (uint32_t)shfl_xor_const(j / E, (int)v[r]);
          const uint32_t mn = v[r] < o ? v[r] : o;
          const uint32_t mx = v[r] < o ? o : v[r];
          v[r] = lower ? mn : mx;
        }
      }
    }
  }
}

// ---- block bitonic sort in LDS -------------------------------------------
// Sorts n2 (a power of two) uint32 keys in LDS with all BS threads of the block.
template <int BS>
__device__ __forceinline__ void block_bitonic_sort_lds(uint32_t* a, int n2) {
  for (int k = 2; k <= n2; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int t = threadIdx.x; t < n2 / 2; t += BS) {
        // t-th pair: lower index i has bit j clear
        int i = ((t / j) * 2 * j) + (t % j);
        int p = i + j;
        bool asc = (i & k) == 0;
        uint32_t x = a[i], y = a[p];
        if (asc ? (x > y) : (x < y)) { a[i] = y; a[p] = x; }
      }
      __syncthreads();
    }
  }
}

// ---- block-wide exclusive scan (BS threads, one value each) ---------------
template <int BS, typename T>
__device__ __forceinline__ T block_excl_scan(T x, T* smem /* BS/64 + 1 */, T* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  T inc = group_incl_scan<64>(x, lane);
  if (lane == 63) smem[w] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    T s = 0;
    for (int i = 0; i < BS / 64; i++) { T t = smem[i]; smem[i] = s; s += t; }
    smem[BS / 64] = s;
  }
  __syncthreads();
  T r = inc - x + smem[w];
  if (total) *total = smem[BS / 64];
  __syncthreads();
  return r;
}

}  // namespace nsp
