// Device-side building blocks for gfx950 (CDNA4, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace oc {

constexpr int kWave = 64;
constexpr uint32_t kEmpty = 0xFFFFFFFFu;  // HASH_UNUSED (include/Common.h:71)

// --------------------------------------------------------------- wave ops
__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Make this lane's LDS writes visible to the other lanes of its wave.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint64_t lanemask_lt() {
  int l = lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Inclusive sum over groups of `width` consecutive lanes (width = 2^k <= 64).
template <typename T>
__device__ __forceinline__ T group_incl_sum(T v, int width) {
  int lig = lane_id() & (width - 1);
  for (int d = 1; d < width; d <<= 1) {
    T o = __shfl_up(v, d, width);
    if (lig >= d) v += o;
  }
  return v;
}

template <typename T>
__device__ __forceinline__ T group_sum(T v, int width) {
  for (int d = width >> 1; d > 0; d >>= 1) v += __shfl_xor(v, d, width);
  return v;
}

template <typename T>
__device__ __forceinline__ T group_max(T v, int width) {
  for (int d = width >> 1; d > 0; d >>= 1) {
    T o = __shfl_xor(v, d, width);
    v = o > v ? o : v;
  }
  return v;
}

template <typename T>
__device__ __forceinline__ T group_min(T v, int width) {
  for (int d = width >> 1; d > 0; d >>= 1) {
    T o = __shfl_xor(v, d, width);
    v = o < v ? o : v;
  }
  return v;
}

__device__ __forceinline__ double shfl_d(double v, int src, int width = 64) {
  return __shfl(v, src, width);
}

// Ballot restricted to the lane's group of `width` lanes, shifted to bit 0.
__device__ __forceinline__ uint64_t group_ballot(bool p, int width) {
  uint64_t b = __ballot(p);
  if (width == 64) return b;
  int base = lane_id() & ~(width - 1);
  return (b >> base) & ((1ull << width) - 1);
}

// Block-wide exclusive sum for blockDim.x <= 1024; `scratch` holds 16
// + 1 entries.  Returns the exclusive prefix; *total receives the block sum.
template <typename T>
__device__ __forceinline__ T block_excl_sum(T v, T *scratch, T *total) {
  const int lane = lane_id(), wid = threadIdx.x / kWave;
  const int nw = (blockDim.x + kWave - 1) / kWave;
  T incl = group_incl_sum(v, kWave);
  if (lane == kWave - 1) scratch[wid] = incl;
  __syncthreads();
  if (wid == 0) {
    T s = lane < nw ? scratch[lane] : T(0);
    T si = group_incl_sum(s, kWave);
    if (lane < nw) scratch[lane] = si - s;
    if (lane == nw - 1) scratch[16] = si;
  }
  __syncthreads();
  T r = scratch[wid] + incl - v;
  *total = scratch[16];
  __syncthreads();
  return r;
}

// Inclusive scan over the 64 lanes of a wave with DPP moves: row shifts by
// 1, 2, 4, 8 within each row of 16 lanes, then lane 15 of rows 0 and 2
// into rows 1 and 3 and lane 31 into rows 2 and 3.  Lanes whose source is
// outside the row take `id`.  All 64 lanes must be active.
template <typename Op>
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x, uint32_t id, Op op) {
  x = op(x, (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)x, 0x111, 0xf, 0xf, false));
  x = op(x, (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)x, 0x112, 0xf, 0xf, false));
  x = op(x, (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)x, 0x114, 0xf, 0xf, false));
  x = op(x, (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)x, 0x118, 0xf, 0xf, false));
  x = op(x, (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)x, 0x142, 0xa, 0xf, false));
  x = op(x, (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)x, 0x143, 0xc, 0xf, false));
  return x;
}

// The value of the lane below (lane 0 takes `id`): DPP wave shift right.
__device__ __forceinline__ uint32_t wave_shr1(uint32_t x, uint32_t id) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)x, 0x138, 0xf, 0xf, false);
}

// 32-bit block-wide exclusive sum (blockDim.x a multiple of 64, at most
// 1024): DPP wave scans, one LDS slot per wave; every wave then scans the
// slots across its lanes and reads its own prefix with readlane (no
// per-wave masks, which the compiler would keep live in SGPRs).  `scratch`
// holds 16 entries.
__device__ __forceinline__ uint32_t block_excl_sum(uint32_t v, uint32_t *scratch,
                                                   uint32_t *total) {
  const int lane = lane_id();
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int nw = blockDim.x / kWave;
  auto add = [](uint32_t a, uint32_t b) { return a + b; };
  const uint32_t incl = wave_incl_scan(v, 0u, add);
  if (lane == kWave - 1) scratch[wid] = incl;
  __syncthreads();
  const uint32_t s = lane < nw ? scratch[lane] : 0u;
  const uint32_t si = wave_incl_scan(s, 0u, add);
  *total = (uint32_t)__builtin_amdgcn_readlane((int)si, kWave - 1);
  const uint32_t before = (uint32_t)__builtin_amdgcn_readlane((int)(si - s), wid);
  __syncthreads();
  return before + incl - v;
}

// Block-wide exclusive max of 32-bit values (identity 0), same shape.
__device__ __forceinline__ uint32_t block_excl_max(uint32_t v, uint32_t *scratch) {
  const int lane = lane_id();
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int nw = blockDim.x / kWave;
  auto mx = [](uint32_t a, uint32_t b) { return a > b ? a : b; };
  const uint32_t incl = wave_incl_scan(v, 0u, mx);
  if (lane == kWave - 1) scratch[wid] = incl;
  __syncthreads();
  const uint32_t s = lane < nw ? scratch[lane] : 0u;
  const uint32_t sx = wave_shr1(wave_incl_scan(s, 0u, mx), 0u);
  const uint32_t waves_before = (uint32_t)__builtin_amdgcn_readlane((int)sx, wid);
  __syncthreads();
  return mx(waves_before, wave_shr1(incl, 0u));
}

// Block-wide exclusive max (identity 0) for blockDim.x <= 1024; `scratch`
// holds 16 entries.  Returns the max over the threads before this one.
template <typename T>
__device__ __forceinline__ T block_excl_max(T v, T *scratch) {
  const int lane = lane_id(), wid = threadIdx.x / kWave;
  T incl = v;
  for (int d = 1; d < kWave; d <<= 1) {
    const T o = __shfl_up(incl, d, kWave);
    if (lane >= d) incl = o > incl ? o : incl;
  }
  T excl = __shfl_up(incl, 1, kWave);
  if (lane == 0) excl = T(0);
  if (lane == kWave - 1) scratch[wid] = incl;
  __syncthreads();
  for (int w = 0; w < wid; ++w) excl = scratch[w] > excl ? scratch[w] : excl;
  __syncthreads();
  return excl;
}

// ------------------------------------------------------------- hashing
// Linear-probing slot of a column: Fibonacci hash, then Lemire range
// reduction onto a table of any size (no integer division).
__device__ __forceinline__ uint32_t hash_slot(uint32_t key, uint32_t size) {
  uint32_t h = key * 0x9E3779B1u;
  return (uint32_t)(((uint64_t)h * size) >> 32);
}

// MurmurHash3_x86_32 of one 4-byte key with seed 1234 (the public
// algorithm; kernels/MurmurHash.cuh:44-98 of the reference).
__device__ __forceinline__ uint32_t murmur3_col(uint32_t k1) {
  uint32_t h1 = 1234u;
  k1 *= 0xcc9e2d51u;
  k1 = (k1 << 15) | (k1 >> 17);
  k1 *= 0x1b873593u;
  h1 ^= k1;
  h1 = (h1 << 13) | (h1 >> 19);
  h1 = h1 * 5u + 0xe6546b64u;
  h1 ^= 4u;
  h1 ^= h1 >> 16;
  h1 *= 0x85ebca6bu;
  h1 ^= h1 >> 13;
  h1 *= 0xc2b2ae35u;
  h1 ^= h1 >> 16;
  return h1;
}

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ULL + 1ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}

// --------------------------------------------------------- value access
template <typename T>
__device__ __forceinline__ double ldv(const T *p, uint64_t i) {
  return (double)p[i];
}

// ---------------------------------------------------- LDS radix sort
// Stable LSD sort of n (key, idx) pairs held in LDS, 8-bit digits, over the
// low key_bits bits.  Items are split into contiguous per-wave chunks; each
// 64-item round ranks equal digits with eight ballots; per-(wave,digit)
// counters (u16, nwaves*256) give the scatter positions; dsum holds 256
// digit totals.  The result ends in
// (k0,i0) if the returned parity is 0, else in (k1,i1).  All threads of the
// block must call it.
__device__ __forceinline__ uint64_t peer_mask(uint32_t digit, bool valid) {
  uint64_t m = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    uint64_t bal = __ballot((digit >> b) & 1u);
    m &= ((digit >> b) & 1u) ? bal : ~bal;
  }
  return m;
}

__device__ inline int block_radix_sort(uint32_t *k0, uint16_t *i0,
                                       uint32_t *k1, uint16_t *i1,
                                       uint16_t *cnt, uint32_t *dsum, int n,
                                       int key_bits) {
  const int nw = blockDim.x / kWave;
  const int wid = threadIdx.x / kWave, lane = lane_id();
  const int chunk = ((n + nw - 1) / nw + kWave - 1) / kWave * kWave;
  const int beg = wid * chunk;
  const int end = min(n, beg + chunk);
  uint16_t *mycnt = cnt + wid * 256;
  int parity = 0;
  for (int shift = 0; shift < key_bits; shift += 8) {
    uint32_t *ks = parity ? k1 : k0;
    uint16_t *is = parity ? i1 : i0;
    uint32_t *kd = parity ? k0 : k1;
    uint16_t *id = parity ? i0 : i1;
    for (int d = lane; d < 256; d += kWave) mycnt[d] = 0;
    wave_sync();
    // pass 1: per-wave digit counts
    for (int base = beg; base < end; base += kWave) {
      int it = base + lane;
      bool valid = it < end;
      uint32_t dg = valid ? (ks[it] >> shift) & 255u : 0u;
      uint64_t m = peer_mask(dg, valid);
      if (valid && (m & lanemask_lt()) == 0) mycnt[dg] += (uint16_t)__popcll(m);
      wave_sync();
    }
    __syncthreads();
    // digit-major, wave-minor exclusive offsets
    for (int d = threadIdx.x; d < 256; d += blockDim.x) {
      uint32_t run = 0;
      for (int w = 0; w < nw; ++w) {
        uint32_t c = cnt[w * 256 + d];
        cnt[w * 256 + d] = (uint16_t)run;
        run += c;
      }
      dsum[d] = run;
    }
    __syncthreads();
    if (threadIdx.x < kWave) {
      // exclusive scan of the 256 digit totals by one wave (4 per lane)
      uint32_t t0 = dsum[4 * lane], t1 = dsum[4 * lane + 1],
               t2 = dsum[4 * lane + 2], t3 = dsum[4 * lane + 3];
      uint32_t s = t0 + t1 + t2 + t3;
      uint32_t incl = group_incl_sum(s, kWave);
      uint32_t ex = incl - s;
      dsum[4 * lane] = ex;
      dsum[4 * lane + 1] = ex + t0;
      dsum[4 * lane + 2] = ex + t0 + t1;
      dsum[4 * lane + 3] = ex + t0 + t1 + t2;
    }
    __syncthreads();
    for (int d = threadIdx.x; d < 256; d += blockDim.x) {
      uint32_t db = dsum[d];
      for (int w = 0; w < nw; ++w) cnt[w * 256 + d] = (uint16_t)(cnt[w * 256 + d] + db);
    }
    __syncthreads();
    // pass 2: scatter
    for (int base = beg; base < end; base += kWave) {
      int it = base + lane;
      bool valid = it < end;
      uint32_t key = valid ? ks[it] : 0u;
      uint16_t ix = valid ? is[it] : (uint16_t)0;
      uint32_t dg = (key >> shift) & 255u;
      uint64_t m = peer_mask(dg, valid);
      uint32_t pos = 0;
      if (valid) pos = (uint32_t)mycnt[dg] + (uint32_t)__popcll(m & lanemask_lt());
      wave_sync();
      if (valid && (m & lanemask_lt()) == 0) mycnt[dg] += (uint16_t)__popcll(m);
      if (valid) {
        kd[pos] = key;
        id[pos] = ix;
      }
      wave_sync();
    }
    __syncthreads();
    parity ^= 1;
  }
  return parity;
}

}  // namespace oc
