// gfx950 accumulator kernels of the SpGEMM hot path:
//   ESC            rows with <= 16/32/64 products, one product per lane
//   LDS hash       symbolic (set) and numeric (col -> FP64 sum) tables with
//                  linear probing on LDS compare-and-swap, numeric rows are
//                  sorted in LDS (radix) and streamed to their final place
//   LDS dense      bitmap (symbolic) / FP64 array + bitmap (numeric) over
//                  the row's column span; output is produced in order
//   windowed       persistent kernels for rows that fit neither: the column
//                  span is swept in LDS-sized windows with per-A-entry
//                  cursors into the (sorted) B rows
// All accumulation is in FP64; FP32 inputs/outputs are converted at the
// load/store.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "k_common.hpp"

namespace oc {

static int key_bits_for(uint64_t cols) {
  int b = 0;
  uint64_t x = cols > 0 ? cols - 1 : 0;
  while (x) {
    ++b;
    x >>= 1;
  }
  return b;
}

// Lanes per A entry for a row (power of two <= 64): about one lane per two
// products of the average referenced B row, widened while the A row leaves
// whole groups idle.  Same spirit as localLoadBalance
// (kernels/AccumulatorCommon.cuh:67-110).  (One lane per product left half
// the lanes idle on B rows just over a power of two: banded rows of 33
// entries took 64-lane groups; half as wide ran the dense bins of a
// 512-product batch 11.9 -> 11.1 ms (numeric) and 12.2 -> 11.7 ms
// (symbolic); a quarter as wide, 12.3 and 12.0 ms.)  Lanes of one group take
// consecutive entries of one B row, so the LDS atomics of a wave land on
// distinct columns; narrower groups (below) put several B rows in a wave,
// which for rows whose B rows overlap (banded products) makes lanes of one
// instruction add into the same word, and ran the dense numeric bin
// 11.9 -> 13.9 ms on a 512-product batch.
__device__ __forceinline__ int pick_group(uint32_t nA, uint32_t prod,
                                          uint32_t maxb, int bs) {
  if (nA == 0) return 1;
  uint32_t avg = prod / nA;
  if (avg < 1) avg = 1;
  int g = 1;
  while (g < 64 && 2u * (uint32_t)g < avg) g <<= 1;
  while (g < 64 && (uint32_t)(bs / g) > nA && (uint32_t)g < maxb) g <<= 1;
  return g;
}

// Lanes per A entry chosen by latency, for the symbolic hash kernel (5.6 ->
// 5.0 ms on a 512-product batch; no gain in the numeric kernels, whose
// FP64 adds suffer more from narrow groups).  A group walks its A entries
// one after another, and each entry is a chain of dependent
// loads (the A column, the B row's bounds, then its entries, U per lane
// in flight), so a row costs about
//   ceil(nA / groups) * (2 + ceil(avg / (U g))) + ceil(maxb / (U g))
// memory round trips: the group's entries, plus the longest B row.  The
// cheapest G wins (ties: the wider, better coalesced one), among the G
// whose heavy-row threshold (list_heavy) is not below the longest B row;
// G = 64 is always allowed.
template <int U>
__device__ __forceinline__ int pick_group_lat(uint32_t nA, uint32_t prod,
                                              uint32_t maxb, int bs) {
  if (nA == 0) return 1;
  const uint32_t avg = max(1u, prod / nA);
  int best = 64;
  uint32_t best_cost = 0xFFFFFFFFu;
  for (int g = 1; g <= 64; g <<= 1) {
    const uint32_t groups = max(1u, (uint32_t)(bs / g));
    const uint32_t share = (prod + groups - 1) / groups;
    if (g < 64 && maxb > max(8u * (uint32_t)g, 2u * share)) continue;
    const uint32_t w = (uint32_t)(U * g);
    const uint32_t iters = (nA + groups - 1) / groups;
    const uint32_t cost = iters * (2u + (avg + w - 1) / w) + (maxb + w - 1) / w;
    if (cost <= best_cost) {
      best_cost = cost;
      best = g;
    }
  }
  return best;
}

__device__ __forceinline__ uint32_t lds_load(uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Walk one B row segment [s, e) with a group of G lanes (q = s + lig, step
// G), U iterations at a time: the U column loads are issued before any of
// them is used, so each lane keeps U gathers in flight.  TAIL: the last
// (at most U - 1) loads of a lane are issued together too, which pays on
// the longest rows and costs the small-workgroup kernels a few percent.
template <int U, bool TAIL = false, typename F>
__device__ __forceinline__ void seg_cols(uint32_t s, uint32_t e, int lig, int G,
                                         const uint32_t *__restrict__ b_ci, F &&f) {
  uint32_t q = s + lig;
  const uint32_t step = (uint32_t)G;
  for (; q + (U - 1) * step < e; q += U * step) {
    uint32_t c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) c[u] = b_ci[q + u * step];
#pragma unroll
    for (int u = 0; u < U; ++u) f(c[u]);
  }
  if constexpr (TAIL) {
    uint32_t c[U - 1];
#pragma unroll
    for (int u = 0; u < U - 1; ++u) c[u] = q + u * step < e ? b_ci[q + u * step] : 0u;
#pragma unroll
    for (int u = 0; u < U - 1; ++u)
      if (q + u * step < e) f(c[u]);
  } else {
    for (; q < e; q += step) f(b_ci[q]);
  }
}

// Same with the B values; f(col, av * b_value).
template <int U, typename T, typename F>
__device__ __forceinline__ void seg_vals(uint32_t s, uint32_t e, int lig, int G, double av,
                                         const uint32_t *__restrict__ b_ci,
                                         const T *__restrict__ b_v, F &&f) {
  uint32_t q = s + lig;
  const uint32_t step = (uint32_t)G;
  for (; q + (U - 1) * step < e; q += U * step) {
    uint32_t c[U];
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      c[u] = b_ci[q + u * step];
      v[u] = b_v[q + u * step];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) f(c[u], av * (double)v[u]);
  }
  for (; q < e; q += step) f(b_ci[q], av * (double)b_v[q]);
}

// seg_cols / seg_vals with the U entries of a full step handed over at
// once (fb), so the callee can issue their LDS operations together; the
// tail goes one entry at a time (f1).
template <int U, typename FB, typename F1>
__device__ __forceinline__ void seg_cols_batch(uint32_t s, uint32_t e, int lig, int G,
                                               const uint32_t *__restrict__ b_ci, FB &&fb,
                                               F1 &&f1) {
  uint32_t q = s + lig;
  const uint32_t step = (uint32_t)G;
  for (; q + (U - 1) * step < e; q += U * step) {
    uint32_t c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) c[u] = b_ci[q + u * step];
    fb(c);
  }
  for (; q < e; q += step) f1(b_ci[q]);
}

template <int U, typename T, typename FB, typename F1>
__device__ __forceinline__ void seg_vals_batch(uint32_t s, uint32_t e, int lig, int G, double av,
                                               const uint32_t *__restrict__ b_ci,
                                               const T *__restrict__ b_v, FB &&fb, F1 &&f1) {
  uint32_t q = s + lig;
  const uint32_t step = (uint32_t)G;
  for (; q + (U - 1) * step < e; q += U * step) {
    uint32_t c[U];
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      c[u] = b_ci[q + u * step];
      v[u] = b_v[q + u * step];
    }
    double pv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) pv[u] = av * (double)v[u];
    fb(c, pv);
  }
  for (; q < e; q += step) f1(b_ci[q], av * (double)b_v[q]);
}

// =================================================================== ESC
// One product per lane of an N-lane group; the group ranks its N
// (col, product) pairs by compare-counting and sums equal columns in
// product order, exactly as countSortCompactFused / ESCKernel
// (kernels/AccumulatorESC.cuh:12-127).  Owners of products are found by a
// prefix sum of the B row lengths and a binary search with shuffles.
template <int N, typename T>
__global__ __launch_bounds__(256) void k_esc(KArgs<T> a, uint32_t rows,
                                             KOut<T> o) {
  __shared__ uint32_t sk[256];
  __shared__ double sv[256];
  const int sub = threadIdx.x / N, lig = threadIdx.x % N;
  uint32_t *mk = sk + sub * N;
  double *mv = sv + sub * N;
  const uint32_t per_block = 256 / N;
  for (uint32_t r = blockIdx.x * per_block + sub; r < rows;
       r += gridDim.x * per_block) {
    const uint32_t p = a.products[r];
    if (p > (uint32_t)N) continue;
    const uint32_t a0 = a.a_rp[r], a1 = a.a_rp[r + 1];
    uint32_t col = kEmpty;
    double val = 0.0;
    uint32_t base = 0;
    for (uint32_t c0 = a0; c0 < a1 && base < p; c0 += N) {
      const uint32_t j = c0 + lig;
      uint32_t L = 0, bs = 0;
      double av = 0.0;
      if (j < a1) {
        const uint32_t k = a.a_ci[j];
        bs = a.b_rp[k];
        L = a.b_rp[k + 1] - bs;
        av = (double)a.a_v[j];
      }
      const uint32_t P = group_incl_sum(L, N);
      const uint32_t tot = __shfl(P, N - 1, N);
      const int qq = lig - (int)base;
      int lo = 0, hi = N - 1;
#pragma unroll
      for (int s = N; s > 1; s >>= 1) {
        int mid = (lo + hi) >> 1;
        int pm = (int)__shfl(P, mid, N);
        if (pm > qq) hi = mid;
        else lo = mid + 1;
      }
      const int owner = lo < N ? lo : N - 1;
      const uint32_t prev = __shfl(P, owner > 0 ? owner - 1 : 0, N);
      const uint32_t obs = __shfl(bs, owner, N);
      const double oav = __shfl(av, owner, N);
      if (qq >= 0 && (uint32_t)qq < tot) {
        const uint32_t idx = obs + ((uint32_t)qq - (owner > 0 ? prev : 0u));
        col = a.b_ci[idx];
        val = oav * (double)a.b_v[idx];
      }
      base += tot;
    }
    mk[lig] = col;
    mv[lig] = val;
    wave_sync();
    double sum = 0.0;
    bool lead = col != kEmpty;
    uint64_t less = 0;
#pragma unroll 4
    for (int k = 0; k < N; ++k) {
      const uint32_t kk = mk[k];
      if (kk == col) {
        sum += mv[k];
        if (k < lig) lead = false;
      }
      if (kk < col) less |= 1ull << k;
    }
    const uint64_t lm = group_ballot(lead, N);
    const uint32_t pos = (uint32_t)__popcll(less & lm);
    const uint32_t n = (uint32_t)__popcll(lm);
    const uint64_t off = o.off[r];
    if (lead) {
      o.cols[off + pos] = col;
      o.vals[off + pos] = (T)sum;
    }
    if (lig == 0) o.nnz[r] = n;
    wave_sync();
  }
}

int launch_esc(int esc_n, const AccArgs &a, const NumOut &o, hipStream_t s) {
  const uint32_t rows = (uint32_t)a.A.rows;
  if (rows == 0) return OC_OK;
  const uint32_t per_block = 256 / esc_n;
  dim3 grid(std::min<uint32_t>(cdiv(rows, per_block), 1u << 18));
#define GO(N, T) \
  hipLaunchKernelGGL((k_esc<N, T>), grid, dim3(256), 0, s, kargs<T>(a), rows, kout<T>(o))
  if (a.A.vtype == OC_F64) {
    if (esc_n == 16) GO(16, double);
    else if (esc_n == 32) GO(32, double);
    else if (esc_n == 64) GO(64, double);
    else goto bad;
  } else {
    if (esc_n == 16) GO(16, float);
    else if (esc_n == 32) GO(32, float);
    else if (esc_n == 64) GO(64, float);
    else goto bad;
  }
#undef GO
  LAUNCH_CHECK();
  return OC_OK;
bad:
  set_error("ESC kernel size %d unsupported", esc_n);
  return OC_ERR_INVALID;
}

// ====================================================== A entries of a row
// The entries j = a0 + gid, a0 + gid + ngroups, ... of a row with their B
// row bounds, software-pipelined: the next entry's bounds and the column of
// the entry after it are loaded while f sweeps entry j, so a group waits
// about one memory latency per entry instead of three (A column, B row
// bounds, B entries).  f(j, bs, be) returns false to stop.  (Dense
// symbolic rows: R-MAT 2^20 25.8 -> 25.1 ms, a 512-product batch 11.2 ->
// 10.7 ms; the other row kernels unchanged or slightly faster.)
template <typename KA, typename F>
__device__ __forceinline__ void for_entries(const KA &a, uint32_t a0, uint32_t a1, uint32_t gid,
                                            uint32_t ngroups, F &&f) {
  uint32_t j = a0 + gid;
  if (j >= a1) return;
  const uint32_t k = a.a_ci[j];
  uint32_t bs = a.b_rp[k], be = a.b_rp[k + 1];
  uint32_t kn = j + ngroups < a1 ? a.a_ci[j + ngroups] : 0u;
  for (;;) {
    const uint32_t cj = j, cbs = bs, cbe = be;
    j += ngroups;
    const bool more = j < a1;
    if (more) {
      bs = a.b_rp[kn];
      be = a.b_rp[kn + 1];
      if (j + ngroups < a1) kn = a.a_ci[j + ngroups];
    }
    if (!f(cj, cbs, cbe) || !more) break;
  }
}

// =========================================================== row scalars
// The per-row scalars every row kernel needs, loaded together when the row
// starts: independent loads share one memory round trip instead of each
// waiting behind the previous barrier.
struct RowHead {
  uint32_t a0, a1, prod, maxb, left, right;
};
template <typename KA>
__device__ __forceinline__ RowHead row_head(const KA &a, uint32_t r) {
  RowHead h;
  h.a0 = a.a_rp[r];
  h.a1 = a.a_rp[r + 1];
  h.prod = a.products[r];
  h.maxb = a.max_b[r];
  h.left = a.left[r];
  h.right = a.right[r];
  return h;
}

// ============================================================ heavy rows
// A B row far longer than a lane group's share of the row's products (an
// R-MAT hub) keeps its group busy long after the others finish.  The
// groups list such rows first; when there are at most `cap` of them the
// whole workgroup sweeps each listed row after the groups have done the
// rest.  Returns the number listed (0: none, or too many to list, and the
// groups take every row); the listing pass, which ends with a workgroup
// barrier, is skipped when the row's longest B row is not heavy.
template <typename KA>
__device__ __forceinline__ uint32_t list_heavy(const KA &a, const RowHead &rh, uint32_t a0,
                                               uint32_t a1, int G, int gid, int ngroups,
                                               int lig, uint32_t *count, uint32_t *list,
                                               uint32_t cap, uint32_t *heavy) {
  const uint32_t share = (rh.prod + ngroups - 1) / ngroups;
  *heavy = max(8u * (uint32_t)G, 2u * share);
  if (rh.maxb <= *heavy) return 0u;  // no heavy row: no listing pass
  for (uint32_t j = a0 + gid; j < a1; j += ngroups) {
    const uint32_t k = a.a_ci[j];
    if (lig == 0 && a.b_rp[k + 1] - a.b_rp[k] > *heavy) {
      const uint32_t idx = atomicAdd(count, 1u);
      if (idx < cap) list[idx] = j;
    }
  }
  __syncthreads();
  const uint32_t n = *count;
  return n <= cap ? n : 0u;
}

// Persistent row loop over the rows [bounds[0], bounds[1]) of a bin list:
// chunks of kClaimRows rows claimed from *claim (zeroed before the launch).
// A workgroup that starts late or draws costly rows takes fewer chunks;
// each workgroup striding over a fixed share of the rows ran the bins
// 10-40% slower than launching one workgroup per row.  One loop latch,
// after the barriers that publish the claim.
constexpr uint32_t kClaimRows = 16;
template <typename F>
__device__ __forceinline__ void claim_rows(const uint32_t *__restrict__ bounds, uint32_t *claim,
                                           F &&row) {
  // The claim is published in the first word of the dynamic LDS, which no
  // row keeps between rows (each row initialises its LDS).  A static slot
  // would add LDS to every launch of the kernel: 4 bytes moved the dynamic
  // LDS off its 16-byte alignment (k_sym_dense 32.5 -> 37.2 ms), and 16
  // bytes took a workgroup per CU from bins sized to the CU's LDS
  // (k_sym_hash 4.6 -> 6.4 ms on R-MAT 2^20).
  // Rows per claim: kClaimRows, fewer (down to one) when the bin has
  // fewer than 16 claims per workgroup, so that a bin of few, heavy rows
  // spreads them over the grid instead of queueing them on a handful of
  // workgroups, and leaves a short tail (a 38,424-row panel of R-MAT 2^21
  // degree 32: windowed symbolic rows 33.1 -> 9.6 ms; the 48 panels of
  // that product 3.09 -> 2.17 s).
  extern __shared__ uint32_t s_claim[];
  const uint32_t b0 = bounds[0], b1 = bounds[1];
  const uint32_t n = b1 - min(b0, b1);
  const uint32_t chunk = max(1u, min(kClaimRows, n / (16u * gridDim.x)));
  for (;;) {
    if (threadIdx.x == 0) s_claim[0] = atomicAdd(claim, chunk);
    __syncthreads();
    const uint32_t c = s_claim[0];
    __syncthreads();
    if (c >= n) break;
    const uint32_t e = min(b1, b0 + c + chunk);
    for (uint32_t i = b0 + c; i < e; ++i) {
      row(i);
      __syncthreads();  // the next row re-initialises the LDS
    }
  }
}

// ========================================================== symbolic hash
constexpr uint32_t kSymHeavy = 56;  // listed heavy B rows per row (hdr[8..63])
constexpr uint32_t kDenseHeavy = 32;  // dense kernels: hdr[32..63], count hdr[4]

// Exact row nnz from an LDS hash set (kernels/AccumulatorHash.cuh:11-152).
// With limit < S (estimated-compaction tables) a row whose occupancy
// reaches the limit, or whose probe sequence wraps, is listed as overflow
// and recounted by the windowed kernel.
template <int BS>
__device__ __forceinline__ void sym_hash_row(KArgs<float> a,
                                                 uint32_t r,
                                                 uint32_t S, uint32_t limit,
                                                 uint32_t *__restrict__ nnz,
                                                 uint32_t *ovf_rows,
                                                 uint32_t *ovf_count) {
  extern __shared__ uint32_t lds_u32[];
  uint32_t *hdr = lds_u32;
  uint32_t *tab = lds_u32 + 64;
  const bool track = limit < S;
  const RowHead rh = row_head(a, r);
  // With full-size tables (no load limit), a row whose column span fits the
  // table's S words as a bitmap is counted with ORs and popcounts instead
  // of probing; both give the exact count.  Shrunken tables keep hashing,
  // so the rows that outgrow them are listed as before.
  const uint32_t blo = rh.left & ~31u, bhi = rh.right;
  const uint32_t bw = bhi >= rh.left ? ((bhi - blo) >> 5) + 1 : 0u;
  const bool use_bm = !track && bw && bw <= S;
  if (use_bm) {
    for (uint32_t i = threadIdx.x; i < bw; i += BS) tab[i] = 0u;
  } else {
    for (uint32_t i = threadIdx.x; i < S; i += BS) tab[i] = kEmpty;
  }
  if (threadIdx.x < 4) hdr[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t a0 = rh.a0, a1 = rh.a1;
  const int G = pick_group_lat<8>(a1 - a0, rh.prod, rh.maxb, BS);
  const int ngroups = BS / G, gid = threadIdx.x / G, lig = threadIdx.x % G;
  if (use_bm) {
    auto mark = [&](uint32_t col) {
      const uint32_t c = col - blo;
      atomicOr(&tab[c >> 5], 1u << (c & 31));
    };
    uint32_t heavy;
    const uint32_t nheavy =
        list_heavy(a, rh, a0, a1, G, gid, ngroups, lig, &hdr[3], hdr + 8, kSymHeavy, &heavy);
    for_entries(a, a0, a1, (uint32_t)gid, (uint32_t)ngroups,
                [&](uint32_t, uint32_t bs, uint32_t be) {
                  if (!(nheavy && be - bs > heavy))  // listed: swept below
                    seg_cols<8, false>(bs, be, lig, G, a.b_ci, mark);
                  return true;
                });
    for (uint32_t h = 0; h < nheavy; ++h) {
      const uint32_t k = a.a_ci[hdr[8 + h]];
      seg_cols<8, false>(a.b_rp[k], a.b_rp[k + 1], (int)threadIdx.x, BS, a.b_ci, mark);
    }
    __syncthreads();
    uint32_t cnt = 0;
    for (uint32_t i = threadIdx.x; i < bw; i += BS) cnt += __popc(tab[i]);
    cnt = group_sum(cnt, kWave);
    if (lane_id() == 0) atomicAdd(&hdr[2], cnt);
    __syncthreads();
    if (threadIdx.x == 0) nnz[r] = hdr[2];
    return;
  }
  uint32_t mine = 0;
  bool ovf = false;
  // linear probing from `slot` (`probes` slots already tried), each probe
  // a compare-and-swap that either claims an empty slot or finds the key
  auto probe = [&](uint32_t c, uint32_t slot, uint32_t probes) {
    for (; probes < S; ++probes) {
      const uint32_t old = atomicCAS(&tab[slot], kEmpty, c);
      if (old == kEmpty) {
        ++mine;
        if (track) atomicAdd(&hdr[0], 1u);
        return;
      }
      if (old == c) return;
      if (++slot == S) slot = 0;
    }
    ovf = true;
  };
  auto insert = [&](uint32_t c) {
    if (!ovf) probe(c, hash_slot(c, S), 0);
  };
  // eight keys: first probes issued together, collisions probed on
  auto insert8 = [&](const uint32_t (&c)[8]) {
    if (ovf) return;
    uint32_t sl[8], old[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) sl[u] = hash_slot(c[u], S);
#pragma unroll
    for (int u = 0; u < 8; ++u) old[u] = atomicCAS(&tab[sl[u]], kEmpty, c[u]);
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (old[u] == kEmpty) {
        ++mine;
        if (track) atomicAdd(&hdr[0], 1u);
      } else if (old[u] != c[u] && !ovf) {
        probe(c[u], sl[u] + 1 == S ? 0u : sl[u] + 1, 1);
      }
    }
  };
  uint32_t heavy;
  const uint32_t nheavy =
      list_heavy(a, rh, a0, a1, G, gid, ngroups, lig, &hdr[3], hdr + 8, kSymHeavy, &heavy);
  for_entries(a, a0, a1, (uint32_t)gid, (uint32_t)ngroups,
              [&](uint32_t, uint32_t bs, uint32_t be) {
                if (track && lds_load(&hdr[0]) > limit) {  // past the load limit
                  ovf = true;
                  return false;
                }
                if (!(nheavy && be - bs > heavy))  // listed: swept below
                  seg_cols_batch<8>(bs, be, lig, G, a.b_ci, insert8, insert);
                return !ovf;
              });
  __syncthreads();
  for (uint32_t h = 0; h < nheavy && !ovf; ++h) {
    if (track && lds_load(&hdr[0]) > limit) {
      ovf = true;
      break;
    }
    const uint32_t k = a.a_ci[hdr[8 + h]];
    seg_cols_batch<8>(a.b_rp[k], a.b_rp[k + 1], (int)threadIdx.x, BS, a.b_ci, insert8, insert);
  }
  if (ovf) hdr[1] = 1;
  atomicAdd(&hdr[2], mine);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (hdr[1]) ovf_rows[atomicAdd(ovf_count, 1u)] = r;
    else nnz[r] = hdr[2];
  }
}

template <int BS>
__global__ __launch_bounds__(BS) void k_sym_hash(KArgs<float> a, const uint32_t *__restrict__ rows,
    const uint32_t *__restrict__ bounds, uint32_t S, uint32_t limit, uint32_t *__restrict__ nnz, uint32_t *ovf_rows, uint32_t *ovf_count,
    uint32_t *claim) {
  if (!bounds) {
    sym_hash_row<BS>(a, rows[blockIdx.x], S, limit, nnz, ovf_rows, ovf_count);
    return;
  }
  claim_rows(bounds, claim, [&](uint32_t i) {
    sym_hash_row<BS>(a, rows[i], S, limit, nnz, ovf_rows, ovf_count);
  });
}

// ========================================================= symbolic dense
// Exact row nnz from an LDS bitmap over [left, right]
// (kernels/AccumulatorDense.cuh:43-126).
// bmoff/bm (optional): rows with bmoff[r] != ~0 also store their bitmap,
// n4 uint4 groups from column left & ~127, at bm + bmoff[r], for the
// one-gather rank tiles of the numeric pass.
// WIN: the row's span may exceed the LDS bitmap; it is counted in windows
// of wbits columns (a multiple of 128 the LDS holds), each window sweeping
// the row's B rows again and keeping the columns inside it (the role of the
// windowed symbolic pass for rows wider than any dense bin).
template <int BS, bool WIN = false>
__device__ __forceinline__ void sym_dense_row(KArgs<float> a,
                                                  uint32_t r,
                                                  uint32_t *__restrict__ nnz,
                                                  const uint64_t *__restrict__ bmoff,
                                                  uint4 *__restrict__ bm, uint32_t wbits = 0) {
  extern __shared__ uint32_t lds_u32[];
  uint32_t *hdr = lds_u32;
  uint32_t *bits = lds_u32 + 64;
  const RowHead rh = row_head(a, r);
  const uint32_t a0 = rh.a0, a1 = rh.a1;
  const uint32_t lo = rh.left & ~127u, hi = rh.right;
  if (a1 == a0 || rh.prod == 0 || hi < rh.left) {
    if (threadIdx.x == 0) nnz[r] = 0;
    return;
  }
  const uint64_t bo = bmoff ? bmoff[r] : ~0ull;
  const uint32_t nwin = WIN ? (uint32_t)(((uint64_t)hi - lo) / wbits) + 1 : 1u;
  uint4 *bits4 = reinterpret_cast<uint4 *>(bits);
  if (threadIdx.x == 0) hdr[0] = hdr[4] = 0;
  __syncthreads();
  const int G = pick_group(a1 - a0, rh.prod, rh.maxb, BS);
  const int ngroups = BS / G, gid = threadIdx.x / G, lig = threadIdx.x % G;
  uint32_t heavy;  // heavy B rows listed and swept by the whole workgroup
  const uint32_t nheavy =
      list_heavy(a, rh, a0, a1, G, gid, ngroups, lig, &hdr[4], hdr + 32, kDenseHeavy, &heavy);
  for (uint32_t w = 0; w < nwin; ++w) {
    // window [w0, w1] (the whole span without WIN); the bitmap is cleared
    // and counted 128 bits at a time (the launch allocates whole 16-byte
    // groups)
    const uint32_t w0 = lo + w * wbits;
    const uint32_t w1 = WIN ? (uint32_t)min((uint64_t)hi, (uint64_t)w0 + wbits - 1) : hi;
    const uint32_t nwords = ((w1 - w0) >> 5) + 1, n4 = (nwords + 3) / 4;
    for (uint32_t i = threadIdx.x; i < n4; i += BS) bits4[i] = uint4{0, 0, 0, 0};
    __syncthreads();
    auto mark = [&](uint32_t col) {
      const uint32_t c = col - w0;
      if (!WIN || (col >= w0 && col <= w1)) atomicOr(&bits[c >> 5], 1u << (c & 31));
    };
    for_entries(a, a0, a1, (uint32_t)gid, (uint32_t)ngroups,
                [&](uint32_t, uint32_t bs, uint32_t be) {
                  if (!(nheavy && be - bs > heavy))
                    seg_cols<8, BS == 1024>(bs, be, lig, G, a.b_ci, mark);
                  return true;
                });
    for (uint32_t h = 0; h < nheavy; ++h) {
      const uint32_t k = a.a_ci[hdr[32 + h]];
      seg_cols<8, BS == 1024>(a.b_rp[k], a.b_rp[k + 1], (int)threadIdx.x, BS, a.b_ci, mark);
    }
    __syncthreads();
    uint32_t cnt = 0;
    const uint64_t go = bo + (w0 - lo) / 128;  // the window's first stored group
    for (uint32_t i = threadIdx.x; i < n4; i += BS) {
      const uint4 v = bits4[i];
      cnt += __popc(v.x) + __popc(v.y) + __popc(v.z) + __popc(v.w);
      if (bo != ~0ull) bm[go + i] = v;
    }
    cnt = group_sum(cnt, kWave);
    if (lane_id() == 0) atomicAdd(&hdr[0], cnt);
    __syncthreads();  // counted before the next window clears the bitmap
  }
  if (threadIdx.x == 0) nnz[r] = hdr[0];
}

template <int BS, bool WIN = false>
__global__ __launch_bounds__(BS) void k_sym_dense(KArgs<float> a, const uint32_t *__restrict__ rows,
    const uint32_t *__restrict__ bounds, uint32_t *__restrict__ nnz, const uint64_t *__restrict__ bmoff, uint4 *__restrict__ bm,
    uint32_t wbits, uint32_t *claim) {
  if (!bounds) {
    sym_dense_row<BS, WIN>(a, rows[blockIdx.x], nnz, bmoff, bm, wbits);
    return;
  }
  claim_rows(bounds, claim, [&](uint32_t i) {
    sym_dense_row<BS, WIN>(a, rows[i], nnz, bmoff, bm, wbits);
  });
}

// ========================================================== numeric hash
// C row values in an LDS hash table, FP64 sums via LDS atomics
// (kernels/AccumulatorHash.cuh:509-651), then compaction, LDS radix sort
// of (col, slot) and a streamed, ordered write of the row.  EST: the row's
// output capacity is cap[r]; rows whose table fills past 2/3 or beyond
// their capacity are listed as overflow (kernels/AccumulatorHash.cuh:
// 605-623) and recomputed by the windowed kernel.
struct HashLayout {
  uint32_t S, nmax;
  uint32_t off_keys, off_vals, off_ck, off_ci, off_cnt, bytes;
};

__host__ __device__ inline HashLayout hash_layout(uint32_t S, uint32_t nwaves) {
  HashLayout L;
  L.S = S;
  L.nmax = 2 * S / 3 + 1;
  L.off_keys = 1280;
  uint32_t p = L.off_keys + 4 * S + 8;
  p = (p + 7) & ~7u;
  L.off_vals = p;
  p += 8 * S;
  L.off_ck = p;
  p += 4 * L.nmax;
  L.off_ci = p;
  p += 2 * L.nmax;
  p = (p + 3) & ~3u;
  L.off_cnt = p;
  p += nwaves * 512;
  L.bytes = p;
  return L;
}

template <int BS, typename T, bool EST>
__device__ __forceinline__ void num_hash_row(KArgs<T> a,
                                                 uint32_t r,
                                                 HashLayout L, KOut<T> o,
                                                 int key_bits) {
  extern __shared__ __align__(16) uint8_t smem[];
  uint32_t *hdr = reinterpret_cast<uint32_t *>(smem);
  uint32_t *dsum = reinterpret_cast<uint32_t *>(smem + 256);
  uint32_t *keys = reinterpret_cast<uint32_t *>(smem + L.off_keys);
  double *vals = reinterpret_cast<double *>(smem + L.off_vals);
  uint32_t *ck = reinterpret_cast<uint32_t *>(smem + L.off_ck);
  uint16_t *ci = reinterpret_cast<uint16_t *>(smem + L.off_ci);
  uint16_t *cnt = reinterpret_cast<uint16_t *>(smem + L.off_cnt);
  const uint32_t S = L.S;
  const RowHead rh = row_head(a, r);
  const uint64_t off = o.off[r];
  const uint32_t cap = EST ? o.cap[r] : 0u;
  for (uint32_t i = threadIdx.x; i < S; i += BS) {
    keys[i] = kEmpty;
    vals[i] = 0.0;
  }
  if (threadIdx.x < 4) hdr[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t a0 = rh.a0, a1 = rh.a1;
  const int G = pick_group(a1 - a0, rh.prod, rh.maxb, BS);
  const int ngroups = BS / G, gid = threadIdx.x / G, lig = threadIdx.x % G;
  uint32_t limit = S;
  if (EST) limit = min(L.nmax - 1, cap);
  bool ovf = false;
  // linear probing from `slot` (`probes` slots already tried), each probe
  // a compare-and-swap that either claims an empty slot or finds the key;
  // the product is added to the slot found
  auto probe = [&](uint32_t c, double pv, uint32_t slot, uint32_t probes) {
    for (; probes < S; ++probes) {
      const uint32_t old = atomicCAS(&keys[slot], kEmpty, c);
      if (old == kEmpty || old == c) {
        if (EST && old == kEmpty) atomicAdd(&hdr[0], 1u);
        lds_add_f64(&vals[slot], pv);
        return;
      }
      if (++slot == S) slot = 0;
    }
    ovf = true;
  };
  auto insert = [&](uint32_t c, double pv) {
    if (!ovf) probe(c, pv, hash_slot(c, S), 0);
  };
  // four products: first probes issued together, collisions probed on
  auto insert4 = [&](const uint32_t (&c)[4], const double (&pv)[4]) {
    if (ovf) return;
    uint32_t sl[4], old[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) sl[u] = hash_slot(c[u], S);
#pragma unroll
    for (int u = 0; u < 4; ++u) old[u] = atomicCAS(&keys[sl[u]], kEmpty, c[u]);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (old[u] == kEmpty || old[u] == c[u]) {
        if (EST && old[u] == kEmpty) atomicAdd(&hdr[0], 1u);
        lds_add_f64(&vals[sl[u]], pv[u]);
      } else if (!ovf) {
        probe(c[u], pv[u], sl[u] + 1 == S ? 0u : sl[u] + 1, 1);
      }
    }
  };
  uint32_t heavy;  // heavy B rows listed and swept by the whole workgroup
  const uint32_t nheavy =
      list_heavy(a, rh, a0, a1, G, gid, ngroups, lig, &hdr[3], hdr + 8, kSymHeavy, &heavy);
  for_entries(a, a0, a1, (uint32_t)gid, (uint32_t)ngroups,
              [&](uint32_t j, uint32_t bs, uint32_t be) {
                if (EST && lds_load(&hdr[0]) > limit) {  // more keys than the capacity
                  ovf = true;
                  return false;
                }
                if (!(nheavy && be - bs > heavy))  // listed: swept below
                  seg_vals_batch<4>(bs, be, lig, G, (double)a.a_v[j], a.b_ci, a.b_v, insert4,
                                    insert);
                return !ovf;
              });
  __syncthreads();
  for (uint32_t h = 0; h < nheavy && !ovf; ++h) {
    if (EST && lds_load(&hdr[0]) > limit) {
      ovf = true;
      break;
    }
    const uint32_t j = hdr[8 + h];
    const uint32_t k = a.a_ci[j];
    seg_vals_batch<4>(a.b_rp[k], a.b_rp[k + 1], (int)threadIdx.x, BS, (double)a.a_v[j], a.b_ci,
                      a.b_v, insert4, insert);
  }
  if (ovf) hdr[1] = 1;
  __syncthreads();
  const uint32_t lo = rh.left;
  // Output order by rank, no sort: when a bitmap of the row's column span
  // and one prefix per word (8 bytes per 32 columns) fit in the sort's
  // scratch, each key's place in the output is the popcount of the set
  // bits before it.  Wider rows compact the occupied slots and sort.
  const uint32_t spanw = rh.right >= lo ? (rh.right - lo) / 32 + 1 : 0u;
  if (!hdr[1] && spanw && 8ull * spanw <= L.bytes - L.off_ck) {
    uint32_t *bits = reinterpret_cast<uint32_t *>(smem + L.off_ck);
    uint32_t *pref = bits + spanw;
    for (uint32_t i = threadIdx.x; i < spanw; i += BS) bits[i] = 0u;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < S; i += BS) {
      const uint32_t kk = keys[i];
      if (kk != kEmpty) {
        const uint32_t c = kk - lo;
        atomicOr(&bits[c >> 5], 1u << (c & 31));
      }
    }
    __syncthreads();
    // word prefixes: each thread a contiguous run of words
    const uint32_t per = (spanw + BS - 1) / BS;
    const uint32_t w0 = min(spanw, threadIdx.x * per), w1 = min(spanw, w0 + per);
    uint32_t mine = 0;
    for (uint32_t w = w0; w < w1; ++w) mine += __popc(bits[w]);
    uint32_t n;
    uint32_t run = block_excl_sum(mine, dsum, &n);
    for (uint32_t w = w0; w < w1; ++w) {
      pref[w] = run;
      run += __popc(bits[w]);
    }
    if (n > L.nmax || (EST && n > cap)) {  // as the sorting path
      if (threadIdx.x == 0) o.ovf_rows[atomicAdd(o.ovf_count, 1u)] = r;
      return;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < S; i += BS) {
      const uint32_t kk = keys[i];
      if (kk != kEmpty) {
        const uint32_t c = kk - lo, w = c >> 5;
        const uint32_t pos = pref[w] + __popc(bits[w] & ((1u << (c & 31)) - 1u));
        o.cols[off + pos] = kk;
        o.vals[off + pos] = (T)vals[i];
      }
    }
    if (o.nnz && threadIdx.x == 0) o.nnz[r] = n;
    return;
  }
  // compaction of occupied slots; keys relative to the row's first column,
  // so the sort needs only the bits of the row's column span
  if (!hdr[1]) {
    for (uint32_t i = threadIdx.x; i < S; i += BS) {
      const uint32_t kk = keys[i];
      if (kk != kEmpty) {
        const uint32_t pos = atomicAdd(&hdr[2], 1u);
        if (pos < L.nmax) {
          ck[pos] = kk - lo;
          ci[pos] = (uint16_t)i;
        }
      }
    }
  }
  __syncthreads();
  const uint32_t n = hdr[2];
  const bool overflow = hdr[1] || n > L.nmax || (EST && n > cap);
  if (overflow) {
    if (threadIdx.x == 0) o.ovf_rows[atomicAdd(o.ovf_count, 1u)] = r;
    return;
  }
  uint32_t *k1 = keys;
  uint16_t *i1 = reinterpret_cast<uint16_t *>(keys + L.nmax);
  // every output column lies in [left, right]; an empty row has n == 0
  const uint32_t span = n ? rh.right - lo : 0u;
  const int kb = min(key_bits, span ? 32 - __clz(span) : 0);
  const int parity = block_radix_sort(ck, ci, k1, i1, cnt, dsum, (int)n, kb);
  const uint32_t *sk = parity ? k1 : ck;
  const uint16_t *si = parity ? i1 : ci;
  for (uint32_t i = threadIdx.x; i < n; i += BS) {
    o.cols[off + i] = sk[i] + lo;
    o.vals[off + i] = (T)vals[si[i]];
  }
  if (o.nnz && threadIdx.x == 0) o.nnz[r] = n;
}

template <int BS, typename T, bool EST>
__global__ __launch_bounds__(BS) void k_num_hash(KArgs<T> a, const uint32_t *__restrict__ rows,
    const uint32_t *__restrict__ bounds, const uint32_t *__restrict__ err, HashLayout L, KOut<T> o, int key_bits,
    uint32_t *claim) {
  if (err && *err) return;  // the capacity check failed: write nothing
  if (!bounds) {
    num_hash_row<BS, T, EST>(a, rows[blockIdx.x], L, o, key_bits);
    return;
  }
  claim_rows(bounds, claim, [&](uint32_t i) {
    num_hash_row<BS, T, EST>(a, rows[i], L, o, key_bits);
  });
}

// ========================================================= numeric dense
// C row values over the LDS span [left, right] (FP64 sums + bitmap), then
// an ordered compaction (kernels/AccumulatorDense.cuh:132-346).  EST: rows
// whose count exceeds cap[r] are listed as overflow
// (kernels/AccumulatorDense.cuh:327-339).
template <int BS, typename T, bool EST>
__device__ __forceinline__ void num_dense_row(KArgs<T> a,
                                                  uint32_t r,
                                                  KOut<T> o) {
  extern __shared__ __align__(16) uint8_t smem[];
  uint32_t *hdr = reinterpret_cast<uint32_t *>(smem);
  const RowHead rh = row_head(a, r);
  const uint64_t off = o.off[r];
  const uint32_t cap = EST ? o.cap[r] : 0u;
  const uint32_t a0 = rh.a0, a1 = rh.a1;
  const uint32_t lo = rh.left & ~31u, hi = rh.right;
  if (a1 == a0 || rh.prod == 0 || hi < rh.left) {
    if (threadIdx.x == 0 && o.nnz) o.nnz[r] = 0;
    return;
  }
  const uint32_t nwords = ((hi - lo) >> 5) + 1;
  double *vals = reinterpret_cast<double *>(smem + 256);
  uint32_t *bits = reinterpret_cast<uint32_t *>(smem + 256 + (size_t)nwords * 256);
  for (uint32_t i = threadIdx.x; i < nwords * 32; i += BS) vals[i] = 0.0;
  for (uint32_t i = threadIdx.x; i < nwords; i += BS) bits[i] = 0;
  if (threadIdx.x == 0) hdr[4] = 0;
  __syncthreads();
  const int G = pick_group(a1 - a0, rh.prod, rh.maxb, BS);
  const int ngroups = BS / G, gid = threadIdx.x / G, lig = threadIdx.x % G;
  auto add = [&](uint32_t col, double pv) {
    const uint32_t c = col - lo;
    lds_add_f64(&vals[c], pv);
    atomicOr(&bits[c >> 5], 1u << (c & 31));
  };
  uint32_t heavy;  // heavy B rows listed and swept by the whole workgroup
  const uint32_t nheavy =
      list_heavy(a, rh, a0, a1, G, gid, ngroups, lig, &hdr[4], hdr + 32, kDenseHeavy, &heavy);
  for_entries(a, a0, a1, (uint32_t)gid, (uint32_t)ngroups,
              [&](uint32_t j, uint32_t bs, uint32_t be) {
                if (!(nheavy && be - bs > heavy))
                  seg_vals<4>(bs, be, lig, G, (double)a.a_v[j], a.b_ci, a.b_v, add);
                return true;
              });
  for (uint32_t h = 0; h < nheavy; ++h) {
    const uint32_t j = hdr[32 + h];
    const uint32_t k = a.a_ci[j];
    seg_vals<4>(a.b_rp[k], a.b_rp[k + 1], (int)threadIdx.x, BS, (double)a.a_v[j], a.b_ci,
                a.b_v, add);
  }
  __syncthreads();
  const uint32_t wpt = (nwords + BS - 1) / BS;
  const uint32_t w0 = min(nwords, threadIdx.x * wpt), w1 = min(nwords, w0 + wpt);
  uint32_t mycnt = 0;
  for (uint32_t w = w0; w < w1; ++w) mycnt += __popc(bits[w]);
  uint32_t total;
  const uint32_t base = block_excl_sum(mycnt, hdr + 8, &total);
  if (EST && total > cap) {
    if (threadIdx.x == 0) o.ovf_rows[atomicAdd(o.ovf_count, 1u)] = r;
    return;
  }
  uint64_t pos = off + base;
  for (uint32_t w = w0; w < w1; ++w) {
    uint32_t m = bits[w];
    while (m) {
      const int b = __ffs(m) - 1;
      m &= m - 1;
      const uint32_t c = w * 32 + b;
      o.cols[pos] = lo + c;
      o.vals[pos] = (T)vals[c];
      ++pos;
    }
  }
  if (o.nnz && threadIdx.x == 0) o.nnz[r] = total;
}

template <int BS, typename T, bool EST>
__global__ __launch_bounds__(BS) void k_num_dense(KArgs<T> a, const uint32_t *__restrict__ rows,
    const uint32_t *__restrict__ bounds, const uint32_t *__restrict__ err, KOut<T> o,
    uint32_t *claim) {
  if (err && *err) return;  // the capacity check failed: write nothing
  if (!bounds) {
    num_dense_row<BS, T, EST>(a, rows[blockIdx.x], o);
    return;
  }
  claim_rows(bounds, claim, [&](uint32_t i) { num_dense_row<BS, T, EST>(a, rows[i], o); });
}

// ============================================================== windowed
// The numeric fallback for long rows (use_tiled_numeric 0, or no tile
// table): persistent, one 1024-thread workgroup per CU, rows taken from an
// atomic counter.  The row's column span is swept in windows of W columns;
// each A entry keeps a cursor into its B row, whose columns ascend, so a
// window consumes a prefix of what is left of every B row.  The cursor
// lives with the lane group that owns the A entry (LDS for rows up to 4096
// entries, else a per-workgroup global scratch slot).  Windows hold FP64
// sums + a bitmap and are written out in order
// (kernels/AccumulatorDense.cuh:610-806).  (Its symbolic form was replaced
// by the windowed form of k_sym_dense.)
constexpr int kWinBlock = 1024;
constexpr uint32_t kWinLdsCursors = 4096;

template <typename T>
__global__ __launch_bounds__(kWinBlock) void k_window(
    KArgs<T> a, const uint32_t *__restrict__ rows, uint32_t nrows,
    uint32_t *counter, uint32_t W, KOut<T> o,
    uint32_t *__restrict__ scratch, uint32_t max_nA, const uint32_t *__restrict__ bounds,
    const uint32_t *__restrict__ err) {
  if (err && *err) return;  // a capacity check of the run failed
  if (bounds) {             // rows [bounds[0], bounds[1]) of the list
    rows += bounds[0];
    nrows = bounds[1] - bounds[0];
  }
  extern __shared__ __align__(16) uint8_t smem[];
  uint32_t *hdr = reinterpret_cast<uint32_t *>(smem);
  uint32_t *lcur = reinterpret_cast<uint32_t *>(smem + 256);
  uint32_t *bits = reinterpret_cast<uint32_t *>(smem + 256 + kWinLdsCursors * 4);
  double *vals = reinterpret_cast<double *>(smem + 256 + kWinLdsCursors * 4 + (W / 32) * 4);
  for (uint32_t i = threadIdx.x; i < W / 32; i += kWinBlock) bits[i] = 0;
  for (uint32_t i = threadIdx.x; i < W; i += kWinBlock) vals[i] = 0.0;
  __syncthreads();
  for (;;) {
    if (threadIdx.x == 0) hdr[0] = atomicAdd(counter, 1u);
    __syncthreads();
    const uint32_t idx = hdr[0];
    __syncthreads();
    if (idx >= nrows) break;
    const uint32_t r = rows[idx];
    const uint32_t a0 = a.a_rp[r], a1 = a.a_rp[r + 1], nA = a1 - a0;
    const uint32_t left = a.left[r], hi = a.right[r];
    // No `continue` and one latch after a barrier: with a second back edge
    // (an early `continue`) the compiler split this loop in two, leaving
    // thread 0's row claim in an outer loop and the barriers that publish it
    // in an inner loop the other 1023 threads never leave, so the workgroup
    // spun on its first row for ever.
    uint64_t total = 0;
    if (nA != 0 && a.products[r] != 0 && hi >= left) {
    const uint32_t lo = left & ~31u;
    uint32_t *cur = nA <= kWinLdsCursors ? lcur : scratch + (uint64_t)blockIdx.x * max_nA;
    const uint32_t nwin = (hi - lo) / W + 1;
    const uint32_t per = (uint32_t)(((uint64_t)a.products[r] + (uint64_t)nA * nwin - 1) /
                                    ((uint64_t)nA * nwin));
    int G = 1;
    while (G < 64 && (uint32_t)G < per) G <<= 1;
    const int ngroups = kWinBlock / G, gid = threadIdx.x / G, lig = threadIdx.x % G;
    const uint64_t off = o.off[r];
    for (uint32_t w = 0; w < nwin; ++w) {
      const uint64_t w0 = (uint64_t)lo + (uint64_t)w * W;
      const uint64_t w1 = min((uint64_t)hi + 1, w0 + W);
      for (uint32_t jj = gid; jj < nA; jj += ngroups) {
        const uint32_t k = a.a_ci[a0 + jj];
        const uint32_t e = a.b_rp[k + 1];
        uint32_t c = w == 0 ? a.b_rp[k] : cur[jj];
        const double av = (double)a.a_v[a0 + jj];
        for (;;) {
          const uint32_t q = c + lig;
          const uint32_t col = q < e ? a.b_ci[q] : kEmpty;
          const bool valid = q < e && (uint64_t)col < w1;
          if (valid) {
            const uint32_t x = (uint32_t)(col - w0);
            atomicOr(&bits[x >> 5], 1u << (x & 31));
            lds_add_f64(&vals[x], av * (double)a.b_v[q]);
          }
          const uint32_t nv = (uint32_t)__popcll(group_ballot(valid, G));
          c += nv;
          if (nv < (uint32_t)G) break;
        }
        if (lig == 0) cur[jj] = c;
      }
      __syncthreads();
      const uint32_t nwords = (uint32_t)((w1 - w0 + 31) / 32);
      const uint32_t wpt = (nwords + kWinBlock - 1) / kWinBlock;
      const uint32_t b0 = min(nwords, threadIdx.x * wpt), b1 = min(nwords, b0 + wpt);
      uint32_t mycnt = 0;
      for (uint32_t i = b0; i < b1; ++i) mycnt += __popc(bits[i]);
      uint32_t wtotal;
      const uint32_t base = block_excl_sum(mycnt, hdr + 8, &wtotal);
      uint64_t pos = off + total + base;
      for (uint32_t i = b0; i < b1; ++i) {
        uint32_t m = bits[i];
        bits[i] = 0;
        while (m) {
          const int b = __ffs(m) - 1;
          m &= m - 1;
          const uint32_t x = i * 32 + b;
          o.cols[pos] = (uint32_t)(w0 + x);
          o.vals[pos] = (T)vals[x];
          vals[x] = 0.0;
          ++pos;
        }
      }
      total += wtotal;
      __syncthreads();
    }
    }
    if (threadIdx.x == 0) {
      if (o.nnz) o.nnz[r] = (uint32_t)total;
    }
    __syncthreads();
  }
}

// ================================================================= tiled
// Column tiles of W = 2^logW columns.  ts[k * T1 + t] is the index of the
// first entry of B row k whose column is >= t * W (T1 = tiles + 1), so the
// part of any B row inside any tile is known without walking the row.
// A wave takes 64 / P rows at once, P = 2^logP lanes per row (T1 rounded up
// to a power of two, at most 64), so narrow tables keep every lane busy.
__global__ __launch_bounds__(256) void k_tile_starts(const uint32_t *__restrict__ b_rp,
                                                     const uint32_t *__restrict__ b_ci,
                                                     uint32_t rows, uint32_t logW,
                                                     uint32_t T1, const uint32_t *__restrict__ base,
                                                     uint32_t *__restrict__ ts, uint32_t logP) {
  const uint32_t lane = (uint32_t)lane_id(), P = 1u << logP, sub = lane >> logP;
  const uint32_t per_wave = kWave >> logP;
  const uint32_t waves = gridDim.x * 4;
  for (uint32_t k0 = (blockIdx.x * 4 + threadIdx.x / kWave) * per_wave; k0 < rows;
       k0 += waves * per_wave) {
    const uint32_t k = k0 + sub;
    if (k >= rows) continue;
    const uint32_t s = b_rp[k], e = b_rp[k + 1];
    const uint64_t b0 = base ? base[k] : 0;  // tiles counted from the row's block
    for (uint32_t t = lane & (P - 1); t < T1; t += P) {
      const uint64_t bound = b0 + ((uint64_t)t << logW);
      uint32_t lo = s, hi = e;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((uint64_t)b_ci[mid] < bound) lo = mid + 1;
        else hi = mid;
      }
      ts[(uint64_t)k * T1 + t] = lo;
    }
  }
}

// Numeric rows swept tile by tile (one 1024-thread workgroup per row,
// rows taken from an atomic counter).  For each tile the segments of the
// referenced B rows come from the tile-start table; their products are
// spread evenly over all lanes (prefix sum of segment lengths + binary
// search for the owning A entry), so every lane issues independent
// gathers.  FP64 sums + bitmap over the tile in LDS; the tile is written
// out in column order (the role of kernels/AccumulatorDense.cuh:610-806).
static size_t tile_lds(uint32_t W, uint32_t bs) {
  return 256 + (size_t)bs * (4 + 4 + 4 + 8) + (W / 32) * 4 + (size_t)W * 8;
}

template <typename T, int kTileBlock>
__global__ __launch_bounds__(kTileBlock) void k_num_tiles(
    KArgs<T> a, const uint32_t *__restrict__ rows, uint32_t nrows, uint32_t *counter,
    uint32_t logW, const uint32_t *__restrict__ ts, uint32_t T1, KOut<T> o,
    const uint32_t *nrows_dev) {
  constexpr uint32_t kTileA = kTileBlock;  // A entries staged per pass
  if (nrows_dev) nrows = *nrows_dev;
  extern __shared__ __align__(16) uint8_t smem[];
  const uint32_t W = 1u << logW;
  uint32_t *hdr = reinterpret_cast<uint32_t *>(smem);
  uint32_t *sk = reinterpret_cast<uint32_t *>(smem + 256);
  uint32_t *sst = sk + kTileA;
  uint32_t *spre = sst + kTileA;
  double *sav = reinterpret_cast<double *>(spre + kTileA);
  uint32_t *bits = reinterpret_cast<uint32_t *>(sav + kTileA);
  double *vals = reinterpret_cast<double *>(bits + W / 32);
  const uint32_t tid = threadIdx.x;
  for (uint32_t i = tid; i < W / 32; i += kTileBlock) bits[i] = 0;
  for (uint32_t i = tid; i < W; i += kTileBlock) vals[i] = 0.0;
  __syncthreads();
  for (;;) {
    if (tid == 0) hdr[0] = atomicAdd(counter, 1u);
    __syncthreads();
    const uint32_t idx = hdr[0];
    __syncthreads();
    if (idx >= nrows) break;
    const uint32_t r = rows[idx];
    const uint32_t a0 = a.a_rp[r], a1 = a.a_rp[r + 1], nA = a1 - a0;
    const uint32_t left = a.left[r], hi = a.right[r];
    if (nA == 0 || a.products[r] == 0 || hi < left) {
      if (tid == 0 && o.nnz) o.nnz[r] = 0;
      continue;
    }
    const bool staged = nA <= kTileA;
    if (staged) {
      for (uint32_t j = tid; j < nA; j += kTileBlock) {
        sk[j] = a.a_ci[a0 + j];
        sav[j] = (double)a.a_v[a0 + j];
      }
      __syncthreads();
    }
    const uint64_t off = o.off[r];
    uint64_t total = 0;
    const uint32_t t_begin = left >> logW, t_end = hi >> logW;
    for (uint32_t t = t_begin; t <= t_end; ++t) {
      const uint64_t wbase = (uint64_t)t << logW;
      uint32_t any = 0;
      for (uint32_t c0 = a0; c0 < a1; c0 += kTileA) {
        const uint32_t nc = min(kTileA, a1 - c0);
        if (!staged) {
          if (tid < nc) {
            sk[tid] = a.a_ci[c0 + tid];
            sav[tid] = (double)a.a_v[c0 + tid];
          }
          __syncthreads();
        }
        uint32_t len = 0, st = 0;
        if (tid < nc) {
          const uint64_t base = (uint64_t)sk[tid] * T1 + t;
          st = ts[base];
          len = ts[base + 1] - st;
        }
        uint32_t P;
        const uint32_t ex = block_excl_sum(len, hdr + 8, &P);
        if (tid < nc) {
          sst[tid] = st;
          spre[tid] = ex + len;
        }
        __syncthreads();
        any |= P;
        // flattened gathers, 4 products per lane in flight
        for (uint32_t p0 = 0; p0 < P; p0 += 4 * kTileBlock) {
          uint32_t col[4];
          T bv[4];
          double av[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const uint32_t p = p0 + u * kTileBlock + tid;
            col[u] = kEmpty;
            if (p < P) {
              uint32_t lo = 0, hb = nc - 1;
              while (lo < hb) {
                const uint32_t mid = (lo + hb) >> 1;
                if (spre[mid] > p) hb = mid;
                else lo = mid + 1;
              }
              const uint32_t prev = lo ? spre[lo - 1] : 0u;
              const uint32_t q = sst[lo] + (p - prev);
              col[u] = a.b_ci[q];
              bv[u] = a.b_v[q];
              av[u] = sav[lo];
            }
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            if (col[u] != kEmpty) {
              const uint32_t x = (uint32_t)(col[u] - wbase);
              lds_add_f64(&vals[x], av[u] * (double)bv[u]);
              atomicOr(&bits[x >> 5], 1u << (x & 31));
            }
          }
        }
        __syncthreads();
      }
      if (!any) continue;
      // ordered write-out of the tile (W/32 words, one or more per thread)
      const uint32_t nwords = W / 32;
      const uint32_t wpt = (nwords + kTileBlock - 1) / kTileBlock;
      const uint32_t b0 = min(nwords, tid * wpt), b1 = min(nwords, b0 + wpt);
      uint32_t mycnt = 0;
      for (uint32_t i = b0; i < b1; ++i) mycnt += __popc(bits[i]);
      uint32_t wtotal;
      const uint32_t base = block_excl_sum(mycnt, hdr + 8, &wtotal);
      uint64_t pos = off + total + base;
      for (uint32_t i = b0; i < b1; ++i) {
        uint32_t m = bits[i];
        bits[i] = 0;
        while (m) {
          const int b = __ffs(m) - 1;
          m &= m - 1;
          const uint32_t x = i * 32 + b;
          o.cols[pos] = (uint32_t)(wbase + x);
          o.vals[pos] = (T)vals[x];
          vals[x] = 0.0;
          ++pos;
        }
      }
      total += wtotal;
      __syncthreads();
    }
    if (tid == 0 && o.nnz) o.nnz[r] = (uint32_t)total;
  }
}

// Rank-compressed tiles for rows beyond the LDS hash/dense kernels.  A
// tile of W = 2^lw columns (lw from the row's density, 2^14..2^17) is done
// in two passes over its products: the first marks the output columns in
// an LDS bitmap, a popcount prefix over the bitmap words then gives every
// output column its rank, and the second pass adds each product into an
// FP64 slot indexed by that rank.  Sums need one slot per output entry, not
// per column, so tiles are wide and few; ranks are the output order, so the
// tile is written with coalesced stores.  Tiles with more outputs than
// slots are finished in rank ranges.  Segment bounds come from the
// tile-start table (granularity 2^logT columns).
// Shape of a rank-tile workgroup of BS threads: A entries staged per chunk
// (E per thread), products per owner map (U per thread), workgroups per
// CU, widest tile (2^MAXLOG columns) and its bitmap words.  1024 threads
// take a whole CU's LDS; 512 threads (two A entries each) take half and
// 256 a quarter, so two or four rows are in flight per CU and one row's
// LDS phases overlap another's gathers.
template <int BS>
struct RankShape {
  static constexpr uint32_t E = BS == 512 ? 2 : 1;
  static constexpr uint32_t U = 8;
  static constexpr uint32_t PER_CU = 1024 / BS;
  static constexpr uint32_t A = E * BS;
  static constexpr uint32_t ROUND = U * BS;
  static constexpr uint32_t MAXLOG = BS >= 1024 ? 17 : (BS >= 512 ? 16 : 15);
  static constexpr uint32_t WORDS = (1u << MAXLOG) / 32;
  static size_t lds_fixed() {
    return 256 + (size_t)A * (4 + 4 + 8) + (size_t)ROUND * 2 + (size_t)WORDS * 8;
  }
  static_assert(A <= 1024, "owner-map entries hold 10-bit A entry ids");
};

// One workgroup per long row (rows taken from an atomic counter).  A
// chunk's products are flattened (A entry i owns positions [ex_i,
// ex_i + len_i)) and enumerated in rounds of U * BS: an LDS owner map,
// built by marking each segment's first position (and the first position
// of every wave block it covers) and a wave-level max-scan, gives every
// position its A entry with one LDS read, and lane-strided positions keep
// the B gathers coalesced.  The next round's map is built while a round's
// gathers are in flight.
// Owner-map entries: A entry id in the low bits, map tag above it.
constexpr uint32_t kRankIdBits = 10, kRankIdMask = (1u << kRankIdBits) - 1;
constexpr uint32_t kRankTags = 1u << (16 - kRankIdBits);

template <typename T, int BS>
__global__ __launch_bounds__(BS, 4) void k_num_rank(
    KArgs<T> a, const uint32_t *__restrict__ rows, uint32_t nrows, uint32_t *counter,
    uint32_t logT, const uint32_t *__restrict__ ts, uint32_t T1,
    const uint32_t *__restrict__ tsw, uint32_t wper, const uint32_t *__restrict__ bbase,
    uint32_t cap,
    uint32_t pcap, KOut<T> o, const uint32_t *__restrict__ bounds,
    const uint32_t *__restrict__ err) {
  if (err && *err) return;  // a capacity check of the run failed
  if (bounds) {             // rows [bounds[0], bounds[1]) of the list
    rows += bounds[0];
    nrows = bounds[1] - bounds[0];
  }
  using S = RankShape<BS>;
  constexpr uint32_t E = S::E, U = S::U, V = U / 8;
  extern __shared__ __align__(16) uint8_t smem[];
  uint32_t *hdr = reinterpret_cast<uint32_t *>(smem);
  uint32_t *sk = reinterpret_cast<uint32_t *>(smem + 256);
  uint32_t *sbase = sk + S::A;  // B position of product p is sbase[owner] + p
  double *sav = reinterpret_cast<double *>(sbase + S::A);
  uint16_t *own = reinterpret_cast<uint16_t *>(sav + S::A);
  uint32_t *bits = reinterpret_cast<uint32_t *>(own + S::ROUND);
  uint32_t *wpre = bits + S::WORDS;
  double *acc = reinterpret_cast<double *>(wpre + S::WORDS);
  uint32_t *cmap = reinterpret_cast<uint32_t *>(acc + cap);
  typedef unsigned short us8 __attribute__((ext_vector_type(8)));
  const uint32_t tid = threadIdx.x;
  for (uint32_t i = tid; i < S::WORDS; i += BS) bits[i] = 0;
  for (uint32_t i = tid; i < cap; i += BS) acc[i] = 0.0;
  // owner entries hold A entry ids (< S::A) from the start, so the
  // unmasked reads past a round's end stay inside sbase
  for (uint32_t i = tid; i < S::ROUND; i += BS) own[i] = 0;
  uint32_t tag = 0;  // tag of the current owner map
  __syncthreads();
  for (;;) {
    if (tid == 0) hdr[0] = atomicAdd(counter, 1u);
    __syncthreads();
    const uint32_t idx = hdr[0];
    __syncthreads();
    if (idx >= nrows) break;
    const uint32_t r = rows[idx];
    const uint32_t a0 = a.a_rp[r], a1 = a.a_rp[r + 1], nA = a1 - a0;
    const uint32_t left = a.left[r], hi = a.right[r];
    if (nA == 0 || a.products[r] == 0 || hi < left) {
      if (tid == 0 && o.nnz) o.nnz[r] = 0;
      continue;
    }
    // tile width from the row density: about 3/4 of the slots per tile
    const uint64_t span = (uint64_t)hi - left + 1;
    uint64_t est = a.products[r] < span ? a.products[r] : span;
    if (o.nnz_hint) est = o.nnz_hint[r];
    const uint64_t prod = a.products[r];
    uint32_t lw = logT;
    while (lw < S::MAXLOG && (uint64_t)(span >> lw) >= 1 &&
           est * (1ull << (lw + 1)) <= (uint64_t)cap * 3 / 4 * span &&
           (pcap == 0 || prod * (1ull << (lw + 1)) <= (uint64_t)pcap * span))
      ++lw;
    const uint32_t step = lw - logT;
    const bool one_chunk = nA <= S::A;
    const uint64_t off = o.off[r];
    // estimated capacity (estimation workflow): a row whose outputs exceed
    // it stops before writing past its region and joins the overflow list
    const uint64_t rcap = o.cap ? (uint64_t)o.cap[r] : ~0ull;
    bool overflow = false;
    uint64_t total = 0;
    // tiles are counted from the row's column block (block-diagonal batches:
    // every B row a row references starts the same block), else from 0
    const uint32_t base_r = bbase ? bbase[a.a_ci[a0]] : 0u;
    const uint32_t t_begin = (left - base_r) >> lw, t_end = (hi - base_r) >> lw;
    const uint32_t nwords = 1u << (lw - 5);
    const uint32_t wpt = (nwords + BS - 1) / BS;
    const uint32_t b0 = min(nwords, tid * wpt), b1 = min(nwords, b0 + wpt);
    const bool quad = nwords == 4 * BS;  // words [4 tid, 4 tid + 4) per thread
    auto clear_bits = [&] {
      if (quad) reinterpret_cast<uint4 *>(bits)[tid] = uint4{0, 0, 0, 0};
      else for (uint32_t i = b0; i < b1; ++i) bits[i] = 0;
    };
    // segments of this thread's A entries tid * E + j
    uint32_t my_ex[E], my_len[E];

    // stage chunk c0.. of the A row
    auto stage = [&](uint32_t c0, uint32_t nc) {
      for (uint32_t j = tid; j < nc; j += BS) {
        sk[j] = a.a_ci[c0 + j];
        sav[j] = (double)a.a_v[c0 + j];
      }
      __syncthreads();
    };
    // Tile starts of an A entry.  Single-chunk rows walk their tiles in
    // order, so a tile's end is the next tile's start and the end after it
    // is loaded one tile ahead.  Rows at the widest tile read the compact
    // table (all of a B row's starts in one 64-byte slot, so the line a
    // row's tiles share stays cached between them).
    const bool wide = tsw != nullptr && lw == S::MAXLOG;
    auto tsv = [&](uint32_t k, uint32_t tt) {
      if (wide) return tsw[(uint64_t)k * wper + min(wper - 1, tt)];
      return ts[(uint64_t)k * T1 + min(T1 - 1, tt << step)];
    };
    uint32_t ts_cur[E], ts_end[E];
    // segment bounds of tile t for the staged chunk of nc entries
    auto bounds = [&](uint32_t t, uint32_t nc, uint32_t &P) {
      uint32_t st[E], len[E], sum = 0;
#pragma unroll
      for (uint32_t j = 0; j < E; ++j) {
        const uint32_t e = tid * E + j;
        st[j] = 0;
        len[j] = 0;
        if (e < nc) {
          if (one_chunk) {
            st[j] = ts_cur[j];
            len[j] = ts_end[j] - ts_cur[j];
            ts_cur[j] = ts_end[j];
            if (t + 1 <= t_end) ts_end[j] = tsv(sk[e], t + 2);
          } else {
            st[j] = tsv(sk[e], t);
            len[j] = tsv(sk[e], t + 1) - st[j];
          }
        }
        sum += len[j];
      }
      uint32_t ex = block_excl_sum(sum, hdr + 8, &P);
#pragma unroll
      for (uint32_t j = 0; j < E; ++j) {
        my_ex[j] = ex;
        my_len[j] = len[j];
        if (tid * E + j < nc) sbase[tid * E + j] = st[j] - ex;
        ex += len[j];
      }
      __syncthreads();
    };
    // Owner map of positions [p0, p0 + n).  Thread tid holds positions
    // [U tid, U tid + U), so a wave holds one block of U * 64.  A segment
    // marks its first position in the map and the first position of every
    // later block it covers; every block's first position is then marked
    // and a wave-level max-scan completes the map without crossing waves.
    // Marks carry the map's tag above the entry id, so the max picks this
    // map's marks over stale ones and nothing is cleared between maps (the
    // whole map is cleared when the tag wraps).
    auto build_map = [&](uint32_t p0, uint32_t n) {
      constexpr uint32_t BLK = U * kWave;
      __syncthreads();  // the previous map is no longer read
      if (++tag == kRankTags) {
        for (uint32_t i = tid; i < S::ROUND; i += BS) own[i] = 0;
        tag = 1;
        __syncthreads();
      }
      const uint32_t tg = tag << kRankIdBits;
#pragma unroll
      for (uint32_t j = 0; j < E; ++j)
        if (my_len[j] > 0 && my_ex[j] < p0 + n && my_ex[j] + my_len[j] > p0) {
          const uint32_t st = my_ex[j] > p0 ? my_ex[j] - p0 : 0u;
          const uint32_t en = min(my_ex[j] + my_len[j] - p0, n);
          const uint16_t mk = (uint16_t)(tg | (tid * E + j));
          own[st] = mk;
          for (uint32_t b = (st / BLK + 1) * BLK; b < en; b += BLK) own[b] = mk;
        }
      __syncthreads();
      us8 vv[V];
      uint32_t run = 0;
#pragma unroll
      for (uint32_t v = 0; v < V; ++v) {
        vv[v] = (U * tid + 8 * v < n) ? *reinterpret_cast<const us8 *>(own + U * tid + 8 * v)
                                      : us8{0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          run = max(run, (uint32_t)vv[v][k]);
          vv[v][k] = (unsigned short)run;
        }
      }
      auto mx = [](uint32_t x, uint32_t y) { return x > y ? x : y; };
      const uint32_t pre = wave_shr1(wave_incl_scan(run, 0u, mx), 0u);
#pragma unroll
      for (uint32_t v = 0; v < V; ++v) {
#pragma unroll
        for (int k = 0; k < 8; ++k) vv[v][k] = (unsigned short)max(pre, (uint32_t)vv[v][k]);
        if (U * tid + 8 * v < n) *reinterpret_cast<us8 *>(own + U * tid + 8 * v) = vv[v];
      }
      __syncthreads();
    };
    // products p0 + tid + BS u of the current map: B entry q[u] (kEmpty
    // past n), A entry w[u].  All owner reads are issued before any sbase
    // read, with no lane branches.
    auto locate = [&](uint32_t p0, uint32_t n, uint32_t (&q)[U], uint32_t (&w)[U]) {
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) w[u] = own[tid + u * BS] & kRankIdMask;
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) {
        const uint32_t i = tid + u * BS;
        const uint32_t sb = sbase[w[u]];
        q[u] = i < n ? sb + p0 + i : kEmpty;
      }
    };
    // the column bits of the products into the tile bitmap
    auto mark = [&](const uint32_t (&col)[U], uint64_t wbase) {
#pragma unroll
      for (uint32_t u = 0; u < U; ++u)
        if (col[u] != kEmpty) {
          const uint32_t x = (uint32_t)(col[u] - wbase);
          atomicOr(&bits[x >> 5], 1u << (x & 31));
        }
    };
    // products (column, value) summed into ranks [r0, r1): all bitmap and
    // prefix reads first, then the adds
    auto accumulate = [&](const uint32_t (&col)[U], const double (&pv)[U], uint64_t wbase,
                          uint32_t r0, uint32_t r1) {
      uint32_t wp[U], bw[U];
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) {
        const uint32_t wd = col[u] != kEmpty ? (uint32_t)(col[u] - wbase) >> 5 : 0u;
        wp[u] = wpre[wd];
        bw[u] = bits[wd];
      }
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) {
        const uint32_t x = (uint32_t)(col[u] - wbase);
        const uint32_t rk = wp[u] + __popc(bw[u] & ((1u << (x & 31)) - 1u));
        if (col[u] != kEmpty && rk >= r0 && rk < r1) {
          lds_add_f64(&acc[rk - r0], pv[u]);
          cmap[rk - r0] = col[u];
        }
      }
    };
    if (one_chunk) {
      stage(a0, nA);
#pragma unroll
      for (uint32_t j = 0; j < E; ++j) {
        ts_cur[j] = ts_end[j] = 0;
        if (tid * E + j < nA) {
          ts_cur[j] = tsv(sk[tid * E + j], t_begin);
          ts_end[j] = tsv(sk[tid * E + j], t_begin + 1);
        }
      }
    }
    for (uint32_t t = t_begin; t <= t_end; ++t) {
      const uint64_t wbase = (uint64_t)base_r + ((uint64_t)t << lw);
      // ---- pass 1: mark the output columns
      uint32_t any = 0, Plast = 0;
      bool keep = false;
      uint32_t kc[U];  // kept columns and products of round 0 (single-chunk rows)
      double kv[U];
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) {  // defined each tile: nothing carried over
        kc[u] = kEmpty;
        kv[u] = 0.0;
      }
      for (uint32_t c0 = a0; c0 < a1; c0 += S::A) {
        const uint32_t nc = min(S::A, a1 - c0);
        if (!one_chunk) stage(c0, nc);
        uint32_t P;
        bounds(t, nc, P);
        any |= P;
        Plast = P;
        keep = one_chunk;
        // single-chunk rows take the rounds last to first, so round 0 (the
        // only full one when there are two) ends pass 1 with its columns and
        // products in registers, and pass 2 sums it without a gather
        // the next round's owner map is built while this round's gathers
        // are in flight (the map is read before them, the gathers are used
        // after it)
        const uint32_t nr = (P + S::ROUND - 1) / S::ROUND;
        auto rp0 = [&](uint32_t ri) { return (one_chunk ? nr - 1 - ri : ri) * S::ROUND; };
        if (nr) build_map(rp0(0), min(S::ROUND, P - rp0(0)));
        for (uint32_t ri = 0; ri < nr; ++ri) {
          const uint32_t p0 = rp0(ri);
          const uint32_t n = min(S::ROUND, P - p0);
          uint32_t q[U], w[U];
          locate(p0, n, q, w);
#pragma unroll
          for (uint32_t u = 0; u < U; ++u) kc[u] = q[u] != kEmpty ? a.b_ci[q[u]] : kEmpty;
          if (one_chunk && p0 == 0) {  // the last round: nothing to build
#pragma unroll
            for (uint32_t u = 0; u < U; ++u)
              kv[u] = q[u] != kEmpty ? sav[w[u]] * (double)a.b_v[q[u]] : 0.0;
          }
          if (ri + 1 < nr) build_map(rp0(ri + 1), min(S::ROUND, P - rp0(ri + 1)));
          mark(kc, wbase);
        }
        if (!one_chunk) __syncthreads();
      }
      if (!any) continue;
      __syncthreads();
      // ---- ranks: exclusive popcount prefix per bitmap word (the widest
      // tile gives each thread four aligned words: one 128-bit access)
      uint32_t m, run;
      if (quad) {
        const uint4 w4 = reinterpret_cast<const uint4 *>(bits)[tid];
        const uint32_t c0 = __popc(w4.x), c1 = __popc(w4.y), c2 = __popc(w4.z);
        run = block_excl_sum(c0 + c1 + c2 + __popc(w4.w), hdr + 8, &m);
        reinterpret_cast<uint4 *>(wpre)[tid] =
            uint4{run, run + c0, run + c0 + c1, run + c0 + c1 + c2};
      } else {
        uint32_t mycnt = 0;
        for (uint32_t i = b0; i < b1; ++i) mycnt += __popc(bits[i]);
        run = block_excl_sum(mycnt, hdr + 8, &m);
        for (uint32_t i = b0; i < b1; ++i) {
          wpre[i] = run;
          run += __popc(bits[i]);
        }
      }
      __syncthreads();
      if (total + m > rcap) {  // m and total are uniform over the block
        clear_bits();
        __syncthreads();
        overflow = true;
        break;
      }
      // ---- pass 2: sums by rank (in rank ranges when m > cap)
      // products [first, P) of the current chunk, round by round: owner
      // map, columns and values gathered, summed into ranks [r0, r1); the
      // next round's map is built while this round's gathers are in flight
      auto sum_from = [&](uint32_t first, uint32_t P, uint32_t r0, uint32_t r1) {
        if (first >= P) return;
        build_map(first, min(S::ROUND, P - first));
        for (uint32_t p0 = first; p0 < P; p0 += S::ROUND) {
          uint32_t q[U], w[U];
          locate(p0, min(S::ROUND, P - p0), q, w);
          uint32_t col[U];
          T bv[U];
#pragma unroll
          for (uint32_t u = 0; u < U; ++u) {
            col[u] = q[u] != kEmpty ? a.b_ci[q[u]] : kEmpty;
            bv[u] = q[u] != kEmpty ? a.b_v[q[u]] : T(0);
          }
          if (p0 + S::ROUND < P) build_map(p0 + S::ROUND, min(S::ROUND, P - p0 - S::ROUND));
          double pv[U];
#pragma unroll
          for (uint32_t u = 0; u < U; ++u) pv[u] = q[u] != kEmpty ? sav[w[u]] * (double)bv[u] : 0.0;
          accumulate(col, pv, wbase, r0, r1);
        }
      };
      // every round of the tile summed into ranks [r0, r1)
      auto sum_rounds = [&](uint32_t r0, uint32_t r1) {
        for (uint32_t c0 = a0; c0 < a1; c0 += S::A) {
          const uint32_t nc = min(S::A, a1 - c0);
          uint32_t P = Plast;
          if (!one_chunk) {
            stage(c0, nc);
            bounds(t, nc, P);
          }
          sum_from(0, P, r0, r1);
          if (!one_chunk) __syncthreads();
        }
      };
      // coalesced write-out of ranks [r0, r1)
      auto write_ranks = [&](uint32_t r0, uint32_t r1) {
        __syncthreads();
        for (uint32_t i = tid; i < r1 - r0; i += BS) {
          o.cols[off + total + r0 + i] = cmap[i];
          o.vals[off + total + r0 + i] = (T)acc[i];
          acc[i] = 0.0;
        }
        __syncthreads();
      };
      // the first rank range takes round 0 from registers; later ranges
      // (more outputs than slots) gather every round again
      const uint32_t r1a = min(m, cap);
      if (keep) {
        accumulate(kc, kv, wbase, 0, r1a);
        sum_from(S::ROUND, Plast, 0, r1a);
      } else {
        sum_rounds(0, r1a);
      }
      write_ranks(0, r1a);
      for (uint32_t r0 = cap; r0 < m; r0 += cap) {
        const uint32_t r1 = min(m, r0 + cap);
        sum_rounds(r0, r1);
        write_ranks(r0, r1);
      }
      clear_bits();
      total += m;
      __syncthreads();
    }
    if (tid == 0) {
      if (overflow) o.ovf_rows[atomicAdd(o.ovf_count, 1u)] = r;
      if (o.nnz) o.nnz[r] = overflow ? 0u : (uint32_t)total;
    }
  }
}

uint32_t rank_wide_log(uint32_t bs) {
  if (bs == 256) return RankShape<256>::MAXLOG;
  if (bs == 512) return RankShape<512>::MAXLOG;
  return RankShape<1024>::MAXLOG;
}

int wide_starts_plan(uint64_t b_cols, uint32_t bs, size_t *entries_per_row) {
  return wide_starts_for_log(b_cols, rank_wide_log(bs), entries_per_row);
}

int wide_starts_for_log(uint64_t b_cols, uint32_t lw, size_t *entries_per_row) {
  const uint64_t tiles = b_cols ? ((b_cols - 1) >> lw) + 1 : 1;
  if (tiles + 1 > kMaxWideStarts) return 0;
  *entries_per_row = (size_t)tiles + 1;
  return 1;
}

int launch_num_rank(const AccArgs &a, const uint32_t *rows, uint32_t nrows, uint32_t logT,
                    const uint32_t *ts, uint32_t T1, const uint32_t *tsw, uint32_t wper,
                    const uint32_t *bbase,
                    uint32_t pcap, uint32_t bs, const NumOut &o, void *scratch, hipStream_t s,
                    const RowSel &sel) {
  if (nrows == 0 && !sel.bounds) return OC_OK;
  const Tables &t = tables();
  uint32_t *counter = (uint32_t *)scratch;
  OC_TRY(zero_u32(counter, 1, s));
  if (bs != 256 && bs != 512) bs = 1024;
#define GO(T, B)                                                                    \
  do {                                                                              \
    const uint32_t per_cu = RankShape<B>::PER_CU;                                   \
    const size_t budget =                                                           \
        std::min<size_t>(t.lds_max_block, t.lds_per_cu / per_cu) & ~255ull;         \
    const uint32_t full = std::max<uint32_t>(1, t.num_cu) * per_cu;                 \
    const uint32_t grid = sel.bounds ? full : std::min<uint32_t>(nrows, full);      \
    const size_t fixed = RankShape<B>::lds_fixed();                                 \
    if (fixed + 64 * 12 > budget) {                                                 \
      set_error("rank-tile kernel does not fit the LDS of this device");            \
      return OC_ERR_INVALID;                                                        \
    }                                                                               \
    const uint32_t cap = (uint32_t)(((budget - fixed) / 12) & ~63ull);              \
    const size_t lds = fixed + (size_t)cap * 12;                                    \
    OC_TRY(set_lds(k_num_rank<T, B>, lds));                                         \
    hipLaunchKernelGGL((k_num_rank<T, B>), dim3(grid), dim3(B), lds, s, kargs<T>(a), \
                       rows, nrows, counter, logT, ts, T1, tsw, wper, bbase, cap, pcap,   \
                       kout<T>(o), sel.bounds, sel.err);                                     \
  } while (0)
  if (a.A.vtype == OC_F64) {
    if (bs == 256) GO(double, 256);
    else if (bs == 512) GO(double, 512);
    else GO(double, 1024);
  } else {
    if (bs == 256) GO(float, 256);
    else if (bs == 512) GO(float, 512);
    else GO(float, 1024);
  }
#undef GO
  LAUNCH_CHECK();
  return OC_OK;
}

int tile_plan(uint64_t b_rows, uint64_t b_cols, uint32_t log2_cols, uint32_t bs,
              uint32_t *logW, uint32_t *T1, size_t *ts_bytes) {
  const Tables &t = tables();
  uint32_t lw = log2_cols;
  if (lw < 10 || lw > 16 || (bs != 256 && bs != 512 && bs != 1024)) return 0;
  if (tile_lds(1u << lw, bs) > t.lds_max_block) return 0;
  const uint64_t tiles = b_cols ? ((b_cols - 1) >> lw) + 1 : 1;
  const uint64_t bytes = b_rows * (tiles + 1) * 4;
  if (tiles > 4096 || bytes > (8ull << 30)) return 0;
  *logW = lw;
  *T1 = (uint32_t)(tiles + 1);
  *ts_bytes = (size_t)bytes;
  return 1;
}

int launch_tile_starts(const DevCsr &B, uint32_t logW, uint32_t T1, uint32_t *ts,
                       hipStream_t s, const uint32_t *base) {
  if (B.rows == 0) return OC_OK;
  uint32_t logP = 0;
  while (logP < 6 && (1u << logP) < T1) ++logP;
  const uint32_t rows_per_wg = 4u * (64u >> logP);
  hipLaunchKernelGGL(k_tile_starts,
                     dim3(std::min<uint32_t>(cdiv(B.rows, rows_per_wg), 1u << 20)), dim3(256),
                     0, s, B.rp, B.ci, (uint32_t)B.rows, logW, T1, base, ts, logP);
  LAUNCH_CHECK();
  return OC_OK;
}

int launch_num_tiles(const AccArgs &a, const uint32_t *rows, uint32_t nrows, uint32_t logW,
                     uint32_t bs, const uint32_t *ts, uint32_t T1, const NumOut &o,
                     void *scratch, hipStream_t s, const uint32_t *nrows_dev) {
  if (nrows == 0) return OC_OK;
  const Tables &t = tables();
  const size_t lds = tile_lds(1u << logW, bs);
  uint32_t *counter = (uint32_t *)scratch;
  OC_TRY(zero_u32(counter, 1, s));
  // persistent: as many workgroups as fit on the chip at once
  const uint32_t per_cu = std::max<uint32_t>(
      1, std::min<uint32_t>(2048 / bs, (uint32_t)(t.lds_per_cu / lds)));
  const uint32_t grid = std::min<uint32_t>(nrows, std::max<uint32_t>(1, t.num_cu) * per_cu);
#define GO(T, BS)                                                                   \
  do {                                                                              \
    OC_TRY(set_lds(k_num_tiles<T, BS>, lds));                                       \
    hipLaunchKernelGGL((k_num_tiles<T, BS>), dim3(grid), dim3(BS), lds, s,          \
                       kargs<T>(a), rows, nrows, counter, logW, ts, T1, kout<T>(o),  \
                       nrows_dev);                                                   \
  } while (0)
  if (a.A.vtype == OC_F64) {
    if (bs == 256) GO(double, 256);
    else if (bs == 512) GO(double, 512);
    else GO(double, 1024);
  } else {
    if (bs == 256) GO(float, 256);
    else if (bs == 512) GO(float, 512);
    else GO(float, 1024);
  }
#undef GO
  LAUNCH_CHECK();
  return OC_OK;
}

size_t window_scratch_bytes(uint32_t max_a_nnz) {
  const Tables &t = tables();
  uint64_t grid = std::max<uint32_t>(1, t.num_cu);
  uint64_t per = max_a_nnz > kWinLdsCursors ? max_a_nnz : 0;
  return (size_t)(grid * per * 4 + 256);
}

static size_t window_lds(uint32_t W) {
  return 256 + kWinLdsCursors * 4 + (W / 32) * 4 + (size_t)W * 8;
}

template <typename T>
static int window_go(const AccArgs &a, const uint32_t *rows, uint32_t nrows,
                     uint32_t max_a_nnz, const NumOut &o, void *scratch, hipStream_t s,
                     const RowSel &sel) {
  if (nrows == 0 && !sel.bounds) return OC_OK;
  const Tables &t = tables();
  const uint32_t W = t.win_num_cols;
  const size_t lds = window_lds(W);
  if (lds > t.lds_max_block) {
    set_error("window kernel needs %zu B of LDS, device allows %u", lds, t.lds_max_block);
    return OC_ERR_INVALID;
  }
  OC_TRY(set_lds(k_window<T>, lds));
  uint32_t *counter = (uint32_t *)scratch;
  uint32_t *cur = counter + 64;
  OC_TRY(zero_u32(counter, 1, s));
  const uint32_t grid = sel.bounds ? std::max<uint32_t>(1, t.num_cu)
                                  : std::min<uint32_t>(nrows, std::max<uint32_t>(1, t.num_cu));
  KOut<T> ko = kout<T>(o);
  hipLaunchKernelGGL((k_window<T>), dim3(grid), dim3(kWinBlock), lds, s,
                     kargs<T>(a), rows, nrows, counter, W, ko, cur, max_a_nnz, sel.bounds,
                     sel.err);
  LAUNCH_CHECK();
  return OC_OK;
}

int launch_num_window(const AccArgs &a, const uint32_t *rows, uint32_t nrows,
                      uint32_t max_a_nnz, const NumOut &o, void *scratch,
                      hipStream_t s, const RowSel &sel) {
  if (a.A.vtype == OC_F64)
    return window_go<double>(a, rows, nrows, max_a_nnz, o, scratch, s, sel);
  return window_go<float>(a, rows, nrows, max_a_nnz, o, scratch, s, sel);
}

// ------------------------------------------------------------ launchers
// Block sizes per bin (gfx950): 64, 64, 128, 256, 512, 512, 1024.
// the bin's workgroup size (tables().block, default {64,64,128,256,512,
// 512,1024}) picks the kernel instance
#define BIN_DISPATCH(bin, MACRO)                                   \
  switch ((bin) >= 0 && (bin) < 7 ? tables().block[(bin)] : 0u) { \
    case 64: MACRO(64); break;                                     \
    case 128: MACRO(128); break;                                   \
    case 256: MACRO(256); break;                                   \
    case 512: MACRO(512); break;                                   \
    case 1024: MACRO(1024); break;                                 \
    default:                                                       \
      set_error("bad bin %d", bin);                                \
      return OC_ERR_INVALID;                                       \
  }

// Workgroups of a persistent launch: as many as fit on the chip at once.
// Persistent grids hold as many workgroups as the device keeps resident at
// once for the kernel (its registers, LDS and waves), so that every
// workgroup of the row-striding loop starts at launch.
template <typename K>
static uint32_t persistent_grid(K kernel, uint32_t bs, size_t lds) {
  const Tables &t = tables();
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, (int)bs, lds) !=
          hipSuccess ||
      per_cu < 1) {
    (void)hipGetLastError();
    per_cu = 1;
  }
  return std::max<uint32_t>(1, t.num_cu) * (uint32_t)per_cu;
}

// The grid of a persistent launch over a bin: every resident workgroup, or
// at most one per row when the host knows the bin's size (nrows > 0).
template <typename K>
static uint32_t claim_grid(K kernel, uint32_t bs, size_t lds, uint32_t nrows) {
  const uint32_t full = persistent_grid(kernel, bs, lds);
  return nrows ? std::min<uint32_t>(full, nrows) : full;
}

// A bin launch whose host knows the bin (sel.first, nrows) runs one
// workgroup per row when more than two workgroups fit a CU: the hardware
// then hides the gap between one row's workgroup and the next, and the
// claims cost more than they gain (R-MAT 2^20: k_num_dense 0.11 -> 0.46 ms
// persistent).  With one or two per CU the CU idles in that gap, and the
// claims win (k_sym_dense<1024> 32.6 -> 26.1 ms).
template <typename K>
static bool run_persistent(const RowSel &sel, uint32_t nrows, K kernel, uint32_t bs, size_t lds) {
  if (!sel.bounds) return false;
  if (!nrows || sel.first == ~0u) return true;
  return persistent_grid(kernel, bs, lds) <= 2 * std::max<uint32_t>(1, tables().num_cu);
}

int launch_sym_hash(int bin, const AccArgs &a, const uint32_t *rows,
                    uint32_t nrows, uint32_t limit, uint32_t *nnz,
                    uint32_t *ovf_rows, uint32_t *ovf_count, hipStream_t s, const RowSel &sel) {
  if (nrows == 0 && !sel.bounds) return OC_OK;
  const Tables &t = tables();
  const uint32_t S = t.hash_sym[bin];
  const size_t lds = 256 + (size_t)S * 4;
  KArgs<float> k = kargs<float>(a);
#define GO(BS)                                                                \
  do {                                                                        \
    OC_TRY(set_lds(k_sym_hash<BS>, lds));                                     \
    const bool per = run_persistent(sel, nrows, k_sym_hash<BS>, BS, lds);    \
    const uint32_t grid = per ? claim_grid(k_sym_hash<BS>, BS, lds, nrows) : nrows; \
    if (per) OC_TRY(zero_u32(sel.claim, 1, s));                               \
    hipLaunchKernelGGL(k_sym_hash<BS>, dim3(grid), dim3(BS), lds, s, k,        \
                       per || !sel.bounds ? rows : rows + sel.first,          \
                       per ? sel.bounds : nullptr, S, std::min(limit, S), nnz, \
                       ovf_rows, ovf_count, sel.claim);                       \
  } while (0)
  BIN_DISPATCH(bin, GO)
#undef GO
  LAUNCH_CHECK();
  return OC_OK;
}

int launch_sym_dense(int dense_bin, const AccArgs &a, const uint32_t *rows,
                     uint32_t nrows, uint32_t *nnz, hipStream_t s, const uint64_t *bmoff,
                     void *bm, const RowSel &sel) {
  if (nrows == 0 && !sel.bounds) return OC_OK;
  const Tables &t = tables();
  // whole 16-byte groups from column left & ~127 over the widest span
  const size_t lds = 256 + ((size_t)(t.dense_sym[dense_bin] + 126) / 128 + 1) * 16;
  KArgs<float> k = kargs<float>(a);
#define GO(BS)                                                                 \
  do {                                                                         \
    OC_TRY(set_lds(k_sym_dense<BS>, lds));                                     \
    const bool per = run_persistent(sel, nrows, k_sym_dense<BS>, BS, lds);     \
    const uint32_t grid = per ? claim_grid(k_sym_dense<BS>, BS, lds, nrows) : nrows; \
    if (per) OC_TRY(zero_u32(sel.claim, 1, s));                                \
    hipLaunchKernelGGL(k_sym_dense<BS>, dim3(grid), dim3(BS), lds, s, k,        \
                       per || !sel.bounds ? rows : rows + sel.first,           \
                       per ? sel.bounds : nullptr, nnz, bmoff, (uint4 *)bm, 0u, \
                       sel.claim);                                             \
  } while (0)
  BIN_DISPATCH(dense_bin, GO)
#undef GO
  LAUNCH_CHECK();
  return OC_OK;
}

// Rows wider than every dense bin (and rows whose hash table filled up):
// the windowed form of the dense symbolic kernel, 1024 threads and the
// largest LDS bitmap of one workgroup, one window of wbits columns at a
// time.  max_a_nnz and scratch are unused (the cursor kernel needed them).
int launch_sym_window(const AccArgs &a, const uint32_t *rows, uint32_t nrows,
                      uint32_t max_a_nnz, uint32_t *nnz, void *scratch,
                      hipStream_t s, const RowSel &sel, const uint64_t *bmoff, void *bm) {
  (void)max_a_nnz;
  (void)scratch;
  if (nrows == 0 && !sel.bounds) return OC_OK;
  const Tables &t = tables();
  const size_t lds = std::min<size_t>(t.lds_max_block, t.lds_per_cu) & ~(size_t)15;
  if (lds < 256 + 3 * 16) {
    set_error("windowed symbolic kernel: %zu B of LDS is too little", lds);
    return OC_ERR_INVALID;
  }
  const uint32_t wbits = (uint32_t)(((lds - 256) / 16 - 1) * 128);
  OC_TRY(set_lds(k_sym_dense<1024, true>, lds));
  const bool per = run_persistent(sel, nrows, k_sym_dense<1024, true>, 1024, lds);
  const uint32_t grid = per ? claim_grid(k_sym_dense<1024, true>, 1024, lds, nrows) : nrows;
  if (per) OC_TRY(zero_u32(sel.claim, 1, s));
  hipLaunchKernelGGL((k_sym_dense<1024, true>), dim3(grid), dim3(1024), lds, s,
                     kargs<float>(a), per || !sel.bounds ? rows : rows + sel.first,
                     per ? sel.bounds : nullptr, nnz, bmoff, (uint4 *)bm, wbits, sel.claim);
  LAUNCH_CHECK();
  return OC_OK;
}

int launch_num_hash(int bin, const AccArgs &a, const uint32_t *rows,
                    uint32_t nrows, const NumOut &o, hipStream_t s, const RowSel &sel) {
  if (nrows == 0 && !sel.bounds) return OC_OK;
  const Tables &t = tables();
  const HashLayout L = hash_layout(t.hash_num[bin], t.block[bin] / 64);
  if (L.bytes > t.lds[bin]) {
    set_error("numeric hash bin %d layout %u B exceeds %u B", bin, L.bytes, t.lds[bin]);
    return OC_ERR_INVALID;
  }
  const int kb = key_bits_for(a.B.cols);
  const bool est = o.cap != nullptr;
#define GO3(BS, T, E)                                                          \
  do {                                                                         \
    OC_TRY(set_lds(k_num_hash<BS, T, E>, L.bytes));                            \
    const bool per = run_persistent(sel, nrows, k_num_hash<BS, T, E>, BS, L.bytes); \
    const uint32_t grid = per ? claim_grid(k_num_hash<BS, T, E>, BS, L.bytes, nrows) : nrows; \
    if (per) OC_TRY(zero_u32(sel.claim, 1, s));                                \
    hipLaunchKernelGGL((k_num_hash<BS, T, E>), dim3(grid), dim3(BS), L.bytes,  \
                       s, kargs<T>(a), per || !sel.bounds ? rows : rows + sel.first, \
                       per ? sel.bounds : nullptr, sel.err, L, kout<T>(o), kb, \
                       sel.claim);                                             \
  } while (0)
#define GO(BS)                                            \
  do {                                                    \
    if (a.A.vtype == OC_F64) {                            \
      if (est) GO3(BS, double, true);                     \
      else GO3(BS, double, false);                        \
    } else {                                              \
      if (est) GO3(BS, float, true);                      \
      else GO3(BS, float, false);                         \
    }                                                     \
  } while (0)
  BIN_DISPATCH(bin, GO)
#undef GO
#undef GO3
  LAUNCH_CHECK();
  return OC_OK;
}

int launch_num_dense(int dense_bin, const AccArgs &a, const uint32_t *rows,
                     uint32_t nrows, const NumOut &o, hipStream_t s, const RowSel &sel) {
  if (nrows == 0 && !sel.bounds) return OC_OK;
  const Tables &t = tables();
  const uint32_t cap = t.dense_num[dense_bin];
  const uint32_t nwords = cap / 32 + 2;
  const size_t lds = 256 + (size_t)nwords * 256 + (size_t)nwords * 4;
  if (lds > t.lds[dense_bin] + 8) {
    set_error("dense numeric bin %d needs %zu B", dense_bin, lds);
    return OC_ERR_INVALID;
  }
  const bool est = o.cap != nullptr;
#define GO3(BS, T, E)                                                           \
  do {                                                                          \
    OC_TRY(set_lds(k_num_dense<BS, T, E>, lds));                                \
    const bool per = run_persistent(sel, nrows, k_num_dense<BS, T, E>, BS, lds); \
    const uint32_t grid = per ? claim_grid(k_num_dense<BS, T, E>, BS, lds, nrows) : nrows; \
    if (per) OC_TRY(zero_u32(sel.claim, 1, s));                                 \
    hipLaunchKernelGGL((k_num_dense<BS, T, E>), dim3(grid), dim3(BS), lds, s,   \
                       kargs<T>(a), per || !sel.bounds ? rows : rows + sel.first, \
                       per ? sel.bounds : nullptr, sel.err, kout<T>(o), sel.claim); \
  } while (0)
#define GO(BS)                                            \
  do {                                                    \
    if (a.A.vtype == OC_F64) {                            \
      if (est) GO3(BS, double, true);                     \
      else GO3(BS, double, false);                        \
    } else {                                              \
      if (est) GO3(BS, float, true);                      \
      else GO3(BS, float, false);                         \
    }                                                     \
  } while (0)
  BIN_DISPATCH(dense_bin, GO)
#undef GO
#undef GO3
  LAUNCH_CHECK();
  return OC_OK;
}

// ============================================================ row sorting
// Segmented sort of the (col, value) rows of a device CSR in place
// (kernels/Epilogue.cuh:67-222).  Rows up to kSortLdsMax entries: LDS
// radix sort of (col, position) with values staged in LDS.  Longer rows:
// the same stable LSD radix sort streamed through global scratch in
// 1024-item tiles.
constexpr int kSortBlock = 1024;
constexpr uint32_t kSortLdsMax = 6144;

template <typename O, typename T>
__global__ __launch_bounds__(kSortBlock) void k_sort_rows_lds(
    const void *rp, uint32_t *__restrict__ cols, T *__restrict__ vals,
    const uint32_t *__restrict__ rows, uint64_t nrows, int key_bits,
    uint32_t *long_rows, uint32_t *long_count) {
  extern __shared__ __align__(16) uint8_t smem[];
  uint32_t *dsum = reinterpret_cast<uint32_t *>(smem);
  uint16_t *cnt = reinterpret_cast<uint16_t *>(smem + 1024);
  uint32_t *k0 = reinterpret_cast<uint32_t *>(smem + 1024 + 16 * 512);
  uint32_t *k1 = k0 + kSortLdsMax;
  uint16_t *i0 = reinterpret_cast<uint16_t *>(k1 + kSortLdsMax);
  uint16_t *i1 = i0 + kSortLdsMax;
  T *v = reinterpret_cast<T *>(i1 + kSortLdsMax);
  for (uint64_t t = blockIdx.x; t < nrows; t += gridDim.x) {
    const uint32_t r = rows ? rows[t] : (uint32_t)t;
    const uint64_t s = (uint64_t)reinterpret_cast<const O *>(rp)[r];
    const uint64_t e = (uint64_t)reinterpret_cast<const O *>(rp)[r + 1];
    const uint64_t n = e - s;
    if (n > kSortLdsMax) {
      if (threadIdx.x == 0) long_rows[atomicAdd(long_count, 1u)] = r;
      continue;
    }
    if (n <= 1) continue;
    for (uint32_t i = threadIdx.x; i < n; i += kSortBlock) {
      k0[i] = cols[s + i];
      i0[i] = (uint16_t)i;
      v[i] = vals[s + i];
    }
    __syncthreads();
    const int par = block_radix_sort(k0, i0, k1, i1, cnt, dsum, (int)n, key_bits);
    const uint32_t *sk = par ? k1 : k0;
    const uint16_t *si = par ? i1 : i0;
    for (uint32_t i = threadIdx.x; i < n; i += kSortBlock) {
      cols[s + i] = sk[i];
      vals[s + i] = v[si[i]];
    }
    __syncthreads();
  }
}

// One block per long row; keys/positions ping-pong in global scratch.
template <typename O, typename T>
__global__ __launch_bounds__(kSortBlock) void k_sort_rows_global(
    const void *rp, uint32_t *__restrict__ cols, T *__restrict__ vals,
    const uint32_t *__restrict__ long_rows, const uint32_t *long_count,
    int key_bits, uint8_t *scratch, uint64_t scratch_per_block) {
  __shared__ uint32_t base[256];
  __shared__ uint32_t wc[kSortBlock / kWave][256];
  const int wid = threadIdx.x / kWave, lane = lane_id();
  const uint32_t nlong = *long_count;
  uint8_t *my = scratch + (uint64_t)blockIdx.x * scratch_per_block;
  for (uint32_t t = blockIdx.x; t < nlong; t += gridDim.x) {
    const uint32_t r = long_rows[t];
    const uint64_t s = (uint64_t)reinterpret_cast<const O *>(rp)[r];
    const uint64_t n = (uint64_t)reinterpret_cast<const O *>(rp)[r + 1] - s;
    uint32_t *ka = reinterpret_cast<uint32_t *>(my);
    uint32_t *kb = ka + n;
    uint32_t *pa = kb + n;
    uint32_t *pb = pa + n;
    T *vt = reinterpret_cast<T *>(pb + n);
    for (uint64_t i = threadIdx.x; i < n; i += kSortBlock) {
      ka[i] = cols[s + i];
      pa[i] = (uint32_t)i;
      vt[i] = vals[s + i];
    }
    __syncthreads();
    for (int shift = 0; shift < key_bits; shift += 8) {
      if (threadIdx.x < 256) base[threadIdx.x] = 0;
      __syncthreads();
      for (uint64_t i = threadIdx.x; i < n; i += kSortBlock)
        atomicAdd(&base[(ka[i] >> shift) & 255u], 1u);
      __syncthreads();
      if (threadIdx.x < kWave) {
        uint32_t t0 = base[4 * lane], t1 = base[4 * lane + 1], t2 = base[4 * lane + 2],
                 t3 = base[4 * lane + 3];
        uint32_t sum = t0 + t1 + t2 + t3;
        uint32_t ex = group_incl_sum(sum, kWave) - sum;
        base[4 * lane] = ex;
        base[4 * lane + 1] = ex + t0;
        base[4 * lane + 2] = ex + t0 + t1;
        base[4 * lane + 3] = ex + t0 + t1 + t2;
      }
      __syncthreads();
      for (uint64_t t0 = 0; t0 < n; t0 += kSortBlock) {
        const uint64_t i = t0 + threadIdx.x;
        const bool valid = i < n;
        const uint32_t key = valid ? ka[i] : 0u;
        const uint32_t pos0 = valid ? pa[i] : 0u;
        const uint32_t dg = (key >> shift) & 255u;
        for (int d = lane; d < 256; d += kWave) wc[wid][d] = 0;
        __syncthreads();
        const uint64_t m = peer_mask(dg, valid);
        if (valid && (m & lanemask_lt()) == 0) wc[wid][dg] = (uint32_t)__popcll(m);
        __syncthreads();
        if (valid) {
          uint32_t dst = base[dg] + (uint32_t)__popcll(m & lanemask_lt());
          for (int w = 0; w < wid; ++w) dst += wc[w][dg];
          kb[dst] = key;
          pb[dst] = pos0;
        }
        __syncthreads();
        if (threadIdx.x < 256) {
          uint32_t add = 0;
          for (int w = 0; w < kSortBlock / kWave; ++w) add += wc[w][threadIdx.x];
          base[threadIdx.x] += add;
        }
        __syncthreads();
      }
      uint32_t *tk = ka;
      ka = kb;
      kb = tk;
      uint32_t *tp = pa;
      pa = pb;
      pb = tp;
      __syncthreads();
    }
    for (uint64_t i = threadIdx.x; i < n; i += kSortBlock) {
      cols[s + i] = ka[i];
      vals[s + i] = vt[pa[i]];
    }
    __syncthreads();
  }
}

template <typename O, typename T>
static int sort_rows_t(oc_csr &M, const uint32_t *rows, uint64_t n,
                       hipStream_t s) {
  const int kb = key_bits_for(M.cols);
  const size_t lds = 1024 + 16 * 512 + (size_t)kSortLdsMax * (4 + 4 + 2 + 2 + sizeof(T));
  OC_TRY(set_lds(k_sort_rows_lds<O, T>, lds));
  // host-side: find the longest row to size the global scratch
  uint32_t *long_info = nullptr;
  HIP_TRY(hipMallocAsync((void **)&long_info, (n + 64) * 4, s));
  HIP_TRY(hipMemsetAsync(long_info, 0, 4, s));
  const uint32_t grid = (uint32_t)std::min<uint64_t>(n, 4096);
  hipLaunchKernelGGL((k_sort_rows_lds<O, T>), dim3(grid), dim3(kSortBlock), lds, s,
                     (const void *)M.row_offsets, M.col_ids, (T *)M.values, rows, n,
                     kb, long_info + 64, long_info);
  LAUNCH_CHECK();
  uint32_t nlong = 0;
  HIP_TRY(hipMemcpyAsync(&nlong, long_info, 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (nlong) {
    // longest listed row
    std::vector<uint32_t> lr(nlong);
    HIP_TRY(hipMemcpy(lr.data(), long_info + 64, 4 * nlong, hipMemcpyDeviceToHost));
    uint64_t maxn = 0;
    for (uint32_t r : lr) {
      O o2[2];
      HIP_TRY(hipMemcpy(o2, (const O *)M.row_offsets + r, 2 * sizeof(O), hipMemcpyDeviceToHost));
      maxn = std::max<uint64_t>(maxn, (uint64_t)(o2[1] - o2[0]));
    }
    const uint32_t g = std::min<uint32_t>(nlong, 64);
    const uint64_t per = (maxn * (16 + sizeof(T)) + 255) & ~255ull;
    uint8_t *scr = nullptr;
    HIP_TRY(hipMallocAsync((void **)&scr, per * g, s));
    hipLaunchKernelGGL((k_sort_rows_global<O, T>), dim3(g), dim3(kSortBlock), 0, s,
                       (const void *)M.row_offsets, M.col_ids, (T *)M.values,
                       (const uint32_t *)(long_info + 64), (const uint32_t *)long_info,
                       kb, scr, per);
    LAUNCH_CHECK();
    HIP_TRY(hipFreeAsync(scr, s));
  }
  HIP_TRY(hipFreeAsync(long_info, s));
  return OC_OK;
}

int launch_sort_rows(oc_csr &M, const uint32_t *rows, uint64_t n, hipStream_t s) {
  if (n == 0) return OC_OK;
  if (M.off_bytes == 8) {
    if (M.vtype == OC_F64) return sort_rows_t<uint64_t, double>(M, rows, n, s);
    return sort_rows_t<uint64_t, float>(M, rows, n, s);
  }
  if (M.vtype == OC_F64) return sort_rows_t<uint32_t, double>(M, rows, n, s);
  return sort_rows_t<uint32_t, float>(M, rows, n, s);
}

}  // namespace oc
