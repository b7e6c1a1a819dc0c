// One-gather rank tiles: the numeric kernel for rows beyond the LDS hash
// and dense kernels (the role of denseNumericIterKernel,
// kernels/AccumulatorDense.cuh:610-806 of the reference).
//
// A row is swept in column tiles of W = 2^lw columns.  Inside a tile the
// output columns are known as a bitmap, a popcount prefix over its words
// gives every output column its rank (its place in the output), and each
// product is added into the FP64 slot of its rank; the slots are written
// out in rank order, so no sort is needed.
//
// The bitmap of a tile comes from the symbolic pass when it stored the
// row's bitmap (k_sym_dense writes the LDS bitmap it counted; those rows
// gather every product once, here), else from a first pass over the
// tile's products that marks their columns.
//
// Products are enumerated without an owner map: the tile's non-empty B
// segments are compacted (entry j owns positions [cex[j], cex[j+1])), each
// wave takes a contiguous range of positions in blocks of 64, and finds
// the owners of a block from a register window of the next 64 segment
// starts with two ballots.  Waves never wait for each other inside a pass.
#include "k_common.hpp"

namespace oc {

constexpr uint32_t kR1Block = 1024;
constexpr uint32_t kR1U = 8;  // 64-position blocks per step

// Shape of a one-gather workgroup of BS threads: one A entry per thread
// per chunk, widest tile 2^MAXLOG columns (1024 threads take a whole CU's
// LDS; 512 threads half, two workgroups per CU, so one's gathers overlap
// the other's scans and write-outs).
template <int BS>
struct R1Shape {
  static constexpr uint32_t NW = BS / kWave;
  static constexpr uint32_t MAXLOG = BS >= 1024 ? 17 : 16;
  static constexpr uint32_t WORDS = (1u << MAXLOG) / 32;
  static constexpr uint32_t PER_CU = 1024 / BS;
  // LDS bytes besides the cap slots (12 B each: FP64 sum + column)
  static size_t lds_fixed() {
    return 1024 + (size_t)BS * 8 + (size_t)WORDS * 8 + (size_t)BS * 4 + (size_t)(BS + 4) * 4;
  }
  static_assert(BS == 512 || BS == 1024, "one-gather workgroups of 512 or 1024 threads");
};

// Exclusive block sums of three 32-bit values at once (NW waves, at most
// 16): DPP wave scans, one LDS slot per wave and value, one barrier.  `scr`
// holds 64 entries; consecutive calls must alternate between two scratch
// areas (the barrier of the next call orders this call's reads before the
// call after it writes the same area again).
template <uint32_t NW>
__device__ __forceinline__ void block_excl_sum3(uint32_t x, uint32_t y, uint32_t z,
                                                uint32_t *scr, uint32_t &ex, uint32_t &ey,
                                                uint32_t &ez, uint32_t &tx, uint32_t &ty,
                                                uint32_t &tz) {
  const int lane = lane_id();
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  auto add = [](uint32_t a, uint32_t b) { return a + b; };
  const uint32_t ix = wave_incl_scan(x, 0u, add);
  const uint32_t iy = wave_incl_scan(y, 0u, add);
  const uint32_t iz = wave_incl_scan(z, 0u, add);
  if (lane == kWave - 1) {
    scr[wid] = ix;
    scr[16 + wid] = iy;
    scr[32 + wid] = iz;
  }
  __syncthreads();
  static_assert(NW <= 16, "at most 16 waves");
  const bool in = lane < (int)NW;
  const uint32_t sx = in ? scr[lane] : 0u, sy = in ? scr[16 + lane] : 0u,
                 sz = in ? scr[32 + lane] : 0u;
  const uint32_t jx = wave_incl_scan(sx, 0u, add);
  const uint32_t jy = wave_incl_scan(sy, 0u, add);
  const uint32_t jz = wave_incl_scan(sz, 0u, add);
  tx = (uint32_t)__builtin_amdgcn_readlane((int)jx, kWave - 1);
  ty = (uint32_t)__builtin_amdgcn_readlane((int)jy, kWave - 1);
  tz = (uint32_t)__builtin_amdgcn_readlane((int)jz, kWave - 1);
  ex = (uint32_t)__builtin_amdgcn_readlane((int)(jx - sx), wid) + ix - x;
  ey = (uint32_t)__builtin_amdgcn_readlane((int)(jy - sy), wid) + iy - y;
  ez = (uint32_t)__builtin_amdgcn_readlane((int)(jz - sz), wid) + iz - z;
}

template <bool V>
struct BoolC {
  static constexpr bool value = V;
};

// Tile width of a row: the widest 2^lw (from 2^logT up to 2^maxlog) whose
// tiles are expected to fill at most about 3/4 of the cap slots.
__device__ __forceinline__ uint32_t r1_tile_log(uint32_t logT, uint32_t maxlog, uint64_t span,
                                                uint64_t est, uint32_t cap) {
  uint32_t lw = logT;
  while (lw < maxlog && (span >> lw) >= 1 &&
         est * (1ull << (lw + 1)) <= (uint64_t)cap * 3 / 4 * span)
    ++lw;
  return lw;
}

// Pieces of long rows.  A row with more than per_piece products is cut into
// K pieces of consecutive tiles (K at most its tiles and kR1MaxPieces),
// which different workgroups take.  A piece finds the outputs of the row
// before its first tile from the stored bitmap, or, for a row without one
// (only when `counted`), from the per-piece counts of a count-only pass of
// the kernel; no piece waits for another.  kp[i] = K of list entry i; all K
// pieces are appended to `extra` as (i, piece) pairs, consecutive, at most
// max_extra of them (a row that does not fit stays whole).  The kernel takes the pieces before the
// list, so the heaviest work starts first (taken after the list, the
// pieces left a tail: R-MAT 2^20 went from 110.7 to 114.3 ms).
constexpr uint32_t kR1MaxPieces = 64;

template <int BS>
__global__ __launch_bounds__(256) void k_r1_pieces(
    const uint32_t *__restrict__ rows, uint32_t nrows, const uint32_t *__restrict__ products,
    const uint32_t *__restrict__ left, const uint32_t *__restrict__ right,
    const uint32_t *__restrict__ nnz_hint, const uint64_t *__restrict__ bmoff, uint32_t logT,
    uint32_t cap, uint32_t per_piece, uint8_t *__restrict__ kp, uint32_t *__restrict__ extra,
    uint32_t *n_extra, uint32_t max_extra, uint32_t counted) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nrows) return;
  const uint32_t r = rows[i];
  const uint32_t l = left[r], h = right[r];
  const uint64_t prod = products[r];
  uint32_t K = 1;
  const bool stored = bmoff && bmoff[r] != ~0ull;
  if ((stored || counted) && h >= l && prod > per_piece) {
    const uint64_t span = (uint64_t)h - l + 1;
    const uint64_t est = nnz_hint ? nnz_hint[r] : (prod < span ? prod : span);
    const uint32_t lw = r1_tile_log(logT, R1Shape<BS>::MAXLOG, span, est, cap);
    const uint64_t nt = (uint64_t)(h >> lw) - (l >> lw) + 1;
    const uint64_t want = (prod + per_piece - 1) / per_piece;
    K = (uint32_t)min(min(want, nt), (uint64_t)kR1MaxPieces);
    if (K > 1) {
      // reserve K extra entries, or none (lock-free: a failed swap means
      // another row's reservation succeeded)
      uint32_t cur = __hip_atomic_load(n_extra, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      uint32_t base = ~0u;
      while (cur + K <= max_extra) {
        const uint32_t prev = atomicCAS(n_extra, cur, cur + K);
        if (prev == cur) {
          base = cur;
          break;
        }
        cur = prev;
      }
      if (base == ~0u) {
        K = 1;
      } else {
        for (uint32_t j = 0; j < K; ++j) {
          extra[2 * (base + j)] = i;
          extra[2 * (base + j) + 1] = j;
        }
      }
    }
  }
  kp[i] = (uint8_t)K;
}

// One 1024-thread workgroup per row or row piece, taken from an atomic
// counter: the pieces of k_r1_pieces first (when kp is given), then list
// entries [0, nrows), skipping those of split rows.  count_only: only the
// pieces of rows without a stored bitmap, whose outputs are counted into
// pcnt (indexed like `extra`) and nothing is written.  Requires fewer than 2^31
// products per row (the analysis stores them as u32).  bm/bmoff: stored
// row bitmaps (uint4 groups from column left & ~127; bmoff[r] = ~0 for rows
// without one); only used without block bases.
template <typename T, int BS>
__global__ __launch_bounds__(BS) void k_num_rank1(
    KArgs<T> a, const uint32_t *__restrict__ rows, uint32_t nrows, uint32_t *counter,
    uint32_t logT, const uint32_t *__restrict__ ts, uint32_t T1,
    const uint32_t *__restrict__ tsw, uint32_t wper, const uint32_t *__restrict__ bbase,
    const uint4 *__restrict__ bm, const uint64_t *__restrict__ bmoff, uint32_t cap,
    KOut<T> o, const uint32_t *__restrict__ bounds, const uint32_t *__restrict__ err,
    const uint8_t *__restrict__ kp, const uint32_t *__restrict__ extra,
    const uint32_t *__restrict__ n_extra, uint32_t *__restrict__ pcnt, uint32_t count_only) {
  if (err && *err) return;  // a capacity check of the run failed
  if (bounds) {             // rows [bounds[0], bounds[1]) of the list
    rows += bounds[0];
    nrows = bounds[1] - bounds[0];
  }
  using SH = R1Shape<BS>;
  constexpr uint32_t U = kR1U, NW = SH::NW, MAXLOG = SH::MAXLOG, WORDS = SH::WORDS;
  extern __shared__ __align__(16) uint8_t smem[];
  uint32_t *scr = reinterpret_cast<uint32_t *>(smem);  // two 64-entry scan areas
  uint32_t *hdr = reinterpret_cast<uint32_t *>(smem + 512);
  double *cav = reinterpret_cast<double *>(smem + 1024);  // A value of segment j
  uint2 *bw = reinterpret_cast<uint2 *>(cav + BS);        // {bits, rank of bit 0}
  double *acc = reinterpret_cast<double *>(bw + WORDS);
  uint32_t *cq = reinterpret_cast<uint32_t *>(acc + cap);  // B position = cq[j] + p
  uint32_t *cex = cq + BS;                                 // first position of segment j
  uint32_t *cmap = cex + BS + 4;                           // column of rank slot
  const uint32_t tid = threadIdx.x, lane = lane_id();
  const uint32_t wid = __builtin_amdgcn_readfirstlane(tid / kWave);
  for (uint32_t i = tid; i < WORDS; i += BS) bw[i] = uint2{0u, 0u};
  for (uint32_t i = tid; i < cap; i += BS) acc[i] = 0.0;
  uint32_t sp = 0;  // scan scratch area of the next block scan
  auto scan3 = [&](uint32_t x, uint32_t y, uint32_t z, uint32_t &ex, uint32_t &ey,
                   uint32_t &ez, uint32_t &tx, uint32_t &ty, uint32_t &tz) {
    block_excl_sum3<NW>(x, y, z, scr + 64 * sp, ex, ey, ez, tx, ty, tz);
    sp ^= 1u;
  };
  const uint32_t nex = kp ? *n_extra : 0u;
  __syncthreads();
  for (;;) {
    if (tid == 0) {
      const uint32_t c = atomicAdd(counter, 1u);
      // piece ~0: skip the entry; counted pieces: the outputs before it and
      // the row's total from the count pass
      uint32_t i = ~0u, piece = 0, before = 0, rtot = 0, counted = 0;
      if (c < nex) {
        i = extra[2 * c];
        piece = extra[2 * c + 1];
        const bool st = bbase == nullptr && bmoff && bmoff[rows[i]] != ~0ull;
        if (count_only) {
          if (st) piece = ~0u;  // a stored bitmap needs no count
        } else if (!st) {
          counted = 1;
          const uint32_t first = c - piece, K = kp[i];
          for (uint32_t q = 0; q < K; ++q) {
            const uint32_t v = pcnt[first + q];
            before += q < piece ? v : 0u;
            rtot += v;
          }
        }
      } else if (!count_only && c - nex < nrows) {
        i = c - nex;
        piece = kp && kp[i] > 1 ? ~0u : 0u;  // a split row: its pieces did it
      }
      hdr[0] = i;
      hdr[1] = piece;
      hdr[2] = before;
      hdr[3] = rtot;
      hdr[4] = counted;
      hdr[5] = c;
    }
    __syncthreads();
    const uint32_t idx = hdr[0], hp = hdr[1], h_before = hdr[2], h_rtot = hdr[3];
    const bool counted = hdr[4] != 0u;
    const uint32_t claim = hdr[5];
    __syncthreads();
    if (idx >= nrows) break;
    const bool skip = hp == ~0u;
    const uint32_t piece = skip ? 0u : hp;
    const uint32_t K = kp ? kp[idx] : 1u;
    const uint32_t r = rows[idx];
    const uint32_t a0 = a.a_rp[r], a1 = a.a_rp[r + 1], nA = a1 - a0;
    const uint32_t left = a.left[r], hi = a.right[r];
    uint64_t total = counted ? h_before : 0u;
    // a counted row over its estimated capacity: no piece works on it, and
    // piece 0 lists it for the recount
    const bool row_over = counted && o.cap && h_rtot > o.cap[r];
    bool overflow = false;
    if (!skip && !row_over && nA != 0 && a.products[r] != 0 && hi >= left) {
      // tile width from the row density: about 3/4 of the slots per tile
      const uint64_t span = (uint64_t)hi - left + 1;
      const uint64_t prod = a.products[r];
      const uint64_t est = o.nnz_hint ? o.nnz_hint[r] : (prod < span ? prod : span);
      const uint32_t lw = r1_tile_log(logT, MAXLOG, span, est, cap);
      const uint32_t step = lw - logT;
      const uint64_t off = o.off[r];
      // estimated capacity (estimation workflow): a row whose outputs exceed
      // it stops before writing past its region and joins the overflow list
      const uint64_t rcap = o.cap ? (uint64_t)o.cap[r] : ~0ull;
      const uint64_t bo = bmoff ? bmoff[r] : ~0ull;
      const bool stored = bbase == nullptr && bo != ~0ull;
      const uint32_t lo4 = left & ~127u;
      const uint32_t n4row = stored ? ((hi - lo4) >> 7) + 1 : 0u;
      // tiles are counted from the row's column block (block-diagonal
      // batches), else from 0
      const uint32_t base_r = bbase ? bbase[a.a_ci[a0]] : 0u;
      uint32_t t_begin = (left - base_r) >> lw, t_end = (hi - base_r) >> lw;
      // this thread's bitmap words [w0, w0 + wpt) of the tile
      const uint32_t nwords = 1u << (lw - 5);
      const uint32_t wpt = nwords >= BS ? nwords / BS : 1u;
      const uint32_t w0 = tid * wpt;
      const bool has_w = w0 < nwords;
      const bool one_chunk = nA <= BS;
      const bool wide = tsw != nullptr && lw == MAXLOG;
      auto tsv = [&](uint32_t k, uint32_t tt) -> uint32_t {
        if (wide) return tsw[(uint64_t)k * wper + min(wper - 1, tt)];
        return ts[(uint64_t)k * T1 + min(T1 - 1, tt << step)];
      };
      // stored bitmap words of the tile at wbase (zero outside the row)
      auto load_words = [&](uint64_t wbase, uint32_t (&wv)[4]) {
        wv[0] = wv[1] = wv[2] = wv[3] = 0u;
        if (stored && has_w) {
          const int64_t g0 = ((int64_t)wbase - (int64_t)lo4) / 32 + (int64_t)w0;
          const int64_t gn = 4 * (int64_t)n4row;
          if (wpt == 4) {
            if (g0 >= 0 && g0 < gn) {  // g0 is a multiple of 4
              const uint4 v = bm[bo + (uint64_t)(g0 >> 2)];
              wv[0] = v.x;
              wv[1] = v.y;
              wv[2] = v.z;
              wv[3] = v.w;
            }
          } else {
            const uint32_t *b32 = reinterpret_cast<const uint32_t *>(bm + bo);
            for (uint32_t j = 0; j < wpt; ++j) {
              const int64_t g = g0 + j;
              if (g >= 0 && g < gn) wv[j] = b32[g];
            }
          }
        }
      };
      // words and their rank prefixes into LDS (ew: ranks before word w0)
      auto write_words = [&](const uint32_t (&wv)[4], uint32_t ew) {
        if (!has_w) return;
        for (uint32_t j = 0; j < wpt; ++j) {
          bw[w0 + j] = uint2{wv[j], ew};
          ew += __popc(wv[j]);
        }
      };
      // rank prefixes of the marked words; returns the tile's outputs
      auto prefix_marks = [&]() -> uint32_t {
        uint32_t wv[4] = {0u, 0u, 0u, 0u}, pc = 0;
        if (has_w)
          for (uint32_t j = 0; j < wpt; ++j) {
            wv[j] = bw[w0 + j].x;
            pc += __popc(wv[j]);
          }
        uint32_t ew, d1, d2, m, t1, t2;
        scan3(pc, 0u, 0u, ew, d1, d2, m, t1, t2);
        write_words(wv, ew);
        __syncthreads();
        return m;
      };
      auto clear_words = [&]() {
        if (has_w)
          for (uint32_t j = 0; j < wpt; ++j) bw[w0 + j].x = 0u;
      };
      // segments of A entries [c0, c0 + nc) in tile t, compacted; *P
      // products, *ne non-empty segments
      auto setup_chunk = [&](uint32_t c0, uint32_t nc, uint32_t t, uint32_t &P, uint32_t &ne) {
        uint32_t len = 0, st = 0;
        double av = 0.0;
        if (tid < nc) {
          const uint32_t k = a.a_ci[c0 + tid];
          av = (double)a.a_v[c0 + tid];
          st = tsv(k, t);
          len = tsv(k, t + 1) - st;
        }
        uint32_t ex, ej, d, t3;
        scan3(len, len != 0u ? 1u : 0u, 0u, ex, ej, d, P, ne, t3);
        if (len) {
          cex[ej] = ex;
          cq[ej] = st - ex;
          cav[ej] = av;
        }
        if (tid == 0) cex[ne] = P;
        __syncthreads();
      };
      // Products [0, P) of the compacted segments: MARK sets their column
      // bits; else each product is added into its rank slot if the rank is
      // in [r0, r1).
      auto enumerate = [&](auto markc, uint64_t wbase, uint32_t P, uint32_t ne, uint32_t r0,
                           uint32_t r1) {
        constexpr bool MARK = decltype(markc)::value;
        const uint32_t per = ((P + NW - 1) / NW + 63) & ~63u;
        const uint32_t beg = min(P, wid * per), end = min(P, beg + per);
        if (beg >= end) return;
        // owner of position beg: the last segment starting at or before it
        uint32_t e_base;
        {
          const uint32_t S = (ne + 63) / 64;
          const uint32_t i1 = lane * S;
          const uint64_t b1 = __ballot(i1 < ne && cex[min(i1, ne - 1)] <= beg);
          const uint32_t L1 = 63 - __clzll((long long)b1);
          const uint32_t i2 = L1 * S + lane;
          const uint64_t b2 = __ballot(lane < S && i2 < ne && cex[min(i2, ne - 1)] <= beg);
          e_base = L1 * S + (63 - __clzll((long long)b2));
        }
        // window: starts of segments e_base + 1 .. e_base + 64
        uint32_t nxw = e_base + 1 + lane < ne ? cex[e_base + 1 + lane] : kEmpty;
        for (uint32_t pb = beg; pb < end; pb += 64 * U) {
          uint32_t own[U];
#pragma unroll
          for (uint32_t u = 0; u < U; ++u) {
            const uint32_t p0 = pb + 64 * u;
            own[u] = 0;
            if (p0 < end) {
              const uint32_t p = p0 + lane;
              uint64_t mlo = __ballot(nxw <= p0);
              // a block holds at most 63 segment starts after its first
              // position; reload the window when it may not reach past it
              if ((uint32_t)__builtin_amdgcn_readlane((int)nxw, kWave - 1) <= p0 + 63) {
                e_base += (uint32_t)__popcll(mlo);
                nxw = e_base + 1 + lane < ne ? cex[e_base + 1 + lane] : kEmpty;
                mlo = 0;
              }
              uint64_t cand = __ballot(nxw <= p0 + 63) & ~mlo;
              uint32_t ow = e_base + (uint32_t)__popcll(mlo);
              while (cand) {
                const int i = __builtin_ctzll(cand);
                cand &= cand - 1;
                ow += (uint32_t)__builtin_amdgcn_readlane((int)nxw, i) <= p ? 1u : 0u;
              }
              own[u] = ow;
            }
          }
          uint32_t col[U];
          T bv[U];
          double av[U];
#pragma unroll
          for (uint32_t u = 0; u < U; ++u) {
            const uint32_t p = pb + 64 * u + lane;
            col[u] = kEmpty;
            bv[u] = T(0);
            av[u] = 0.0;
            if (p < end) {
              const uint32_t q = cq[own[u]] + p;
              col[u] = a.b_ci[q];
              if (!MARK) {
                bv[u] = a.b_v[q];
                av[u] = cav[own[u]];
              }
            }
          }
#pragma unroll
          for (uint32_t u = 0; u < U; ++u) {
            if (col[u] == kEmpty) continue;
            const uint32_t x = (uint32_t)(col[u] - wbase);
            if (MARK) {
              atomicOr(&bw[x >> 5].x, 1u << (x & 31));
            } else {
              const uint2 e = bw[x >> 5];
              const uint32_t rk = e.y + __popc(e.x & ((1u << (x & 31)) - 1u));
              if (rk >= r0 && rk < r1) {
                lds_add_f64(&acc[rk - r0], av[u] * (double)bv[u]);
                cmap[rk - r0] = col[u];
              }
            }
          }
        }
      };
      // coalesced write-out of ranks [r0, r1) of the tile
      auto write_out = [&](uint32_t r0, uint32_t r1) {
        for (uint32_t i = tid; i < r1 - r0; i += BS) {
          o.cols[off + total + r0 + i] = cmap[i];
          o.vals[off + total + r0 + i] = (T)acc[i];
          acc[i] = 0.0;
        }
      };

      // a piece: tiles [t_begin, t_end] of its share, after the row outputs
      // counted in the stored words before it (or by the count pass)
      if (K > 1) {
        const uint32_t nt = t_end - t_begin + 1, t0 = t_begin;
        t_begin = t0 + (uint32_t)((uint64_t)piece * nt / K);
        t_end = t0 + (uint32_t)((uint64_t)(piece + 1) * nt / K) - 1;
        if (piece && stored) {
          const uint32_t g4 = (uint32_t)((((uint64_t)t_begin << lw) - lo4) >> 7);
          uint32_t c = 0;
          for (uint32_t g = tid; g < g4; g += BS) {
            const uint4 v = bm[bo + g];
            c += __popc(v.x) + __popc(v.y) + __popc(v.z) + __popc(v.w);
          }
          uint32_t d0, d1, d2, before, t1, t2;
          scan3(c, 0u, 0u, d0, d1, d2, before, t1, t2);
          total = before;
        }
      }
      // single-chunk rows: each thread keeps its A entry and the bounds of
      // its segment; the next tile's end is loaded one tile ahead
      uint32_t my_k = 0, ts_cur = 0, ts_end = 0;
      double my_av = 0.0;
      if (one_chunk && tid < nA) {
        my_k = a.a_ci[a0 + tid];
        my_av = (double)a.a_v[a0 + tid];
        ts_cur = tsv(my_k, t_begin);
        ts_end = tsv(my_k, t_begin + 1);
      }
      // (loading the next tile's stored words during the current tile's sums
      // measured slower: the block barriers wait for the loads)
      uint32_t wv[4];
      for (uint32_t t = t_begin; t <= t_end; ++t) {
        const uint64_t wbase = (uint64_t)base_r + ((uint64_t)t << lw);
        load_words(wbase, wv);
        const uint32_t pc = __popc(wv[0]) + __popc(wv[1]) + __popc(wv[2]) + __popc(wv[3]);
        uint32_t P = 0, ne = 0, m = 0;
        if (one_chunk) {
          const uint32_t len = tid < nA ? ts_end - ts_cur : 0u;
          const uint32_t nxt = tid < nA && t + 1 <= t_end ? tsv(my_k, t + 2) : 0u;
          uint32_t ex, ej, ew;
          scan3(len, len != 0u ? 1u : 0u, pc, ex, ej, ew, P, ne, m);
          if (len) {
            cex[ej] = ex;
            cq[ej] = ts_cur - ex;
            cav[ej] = my_av;
          }
          if (tid == 0) cex[ne] = P;
          if (stored) {
            write_words(wv, ew);
          }
          ts_cur = ts_end;
          ts_end = nxt;
          __syncthreads();
          if (!stored) {
            if (P) enumerate(BoolC<true>(), wbase, P, ne, 0u, 0u);
            __syncthreads();
            m = prefix_marks();
          }
        } else if (stored) {
          uint32_t ew, d1, d2, t1, t2;
          scan3(pc, 0u, 0u, ew, d1, d2, m, t1, t2);
          write_words(wv, ew);  // published by the first chunk's barrier
        } else {
          for (uint32_t c0 = a0; c0 < a1; c0 += BS) {
            setup_chunk(c0, min(BS, a1 - c0), t, P, ne);
            if (P) enumerate(BoolC<true>(), wbase, P, ne, 0u, 0u);
            __syncthreads();
          }
          m = prefix_marks();
        }
        // (the count pass only counts: no capacity check, sums or writes)
        if (!count_only && total + m > rcap) {  // m and total are uniform over the block
          clear_words();
          overflow = true;
          break;
        }
        for (uint32_t r0 = 0; !count_only && r0 < m; r0 += cap) {
          const uint32_t r1 = min(m, r0 + cap);
          if (r0) __syncthreads();  // the previous range's slots are cleared
          if (one_chunk) {
            if (P) enumerate(BoolC<false>(), wbase, P, ne, r0, r1);
            __syncthreads();
          } else {
            for (uint32_t c0 = a0; c0 < a1; c0 += BS) {
              setup_chunk(c0, min(BS, a1 - c0), t, P, ne);
              if (P) enumerate(BoolC<false>(), wbase, P, ne, r0, r1);
              __syncthreads();
            }
          }
          write_out(r0, r1);
        }
        clear_words();
        total += m;
      }
    }
    if (tid == 0 && !skip) {
      if (count_only) {
        pcnt[claim] = (uint32_t)total;
      } else {
        if (overflow || (row_over && piece == 0)) o.ovf_rows[atomicAdd(o.ovf_count, 1u)] = r;
        // the last piece's total counts the whole row
        if (o.nnz && piece == K - 1) o.nnz[r] = overflow || row_over ? 0u : (uint32_t)total;
      }
    }
  }
}

int launch_num_rank1(const AccArgs &a, const uint32_t *rows, uint32_t nrows, uint32_t logT,
                     const uint32_t *ts, uint32_t T1, const uint32_t *tsw, uint32_t wper,
                     const uint32_t *bbase, const void *bm, const uint64_t *bmoff,
                     const NumOut &o, void *scratch, hipStream_t s, const RowSel &sel,
                     uint32_t bs, const R1Pieces &pc) {
  if (nrows == 0 && !sel.bounds) return OC_OK;
  const Tables &t = tables();
  if (bs != 512) bs = 1024;
  if (logT > rank1_wide_log(bs) || logT < 10) {
    set_error("one-gather rank tiles of %u threads need a tile-start table of 2^10..2^%u "
              "columns", bs, rank1_wide_log(bs));
    return OC_ERR_INVALID;
  }
  uint32_t *counter = (uint32_t *)scratch;
  OC_TRY(zero_u32(counter, 1, s));
#define GO(T, B)                                                                          \
  do {                                                                                    \
    const size_t budget =                                                                 \
        std::min<size_t>(t.lds_max_block, t.lds_per_cu / R1Shape<B>::PER_CU) & ~255ull;   \
    const size_t fixed = R1Shape<B>::lds_fixed();                                         \
    if (fixed + 64 * 12 > budget) {                                                       \
      set_error("one-gather rank-tile kernel does not fit the LDS of this device");       \
      return OC_ERR_INVALID;                                                              \
    }                                                                                     \
    const uint32_t cap = (uint32_t)(((budget - fixed) / 12) & ~63ull);                    \
    const size_t lds = fixed + (size_t)cap * 12;                                          \
    const uint32_t full = std::max<uint32_t>(1, t.num_cu) * R1Shape<B>::PER_CU;           \
    uint32_t grid = sel.bounds ? full : std::min<uint32_t>(nrows, full);                  \
    OC_TRY(set_lds(k_num_rank1<T, B>, lds));                                              \
    const bool split = pc.kp && !sel.bounds && !bbase && !(o.cap && bmoff);               \
    if (split) {                                                                          \
      OC_TRY(zero_u32(pc.n_extra, 1, s));                                                 \
      hipLaunchKernelGGL(k_r1_pieces<B>, dim3(cdiv(nrows, 256)), dim3(256), 0, s, rows,   \
                         nrows, a.products, a.left, a.right, o.nnz_hint, bmoff, logT,     \
                         cap, pc.per_piece, pc.kp, pc.extra, pc.n_extra, pc.max_extra,    \
                         pc.pcnt ? 1u : 0u);                                              \
      LAUNCH_CHECK();                                                                     \
      grid = full;                                                                        \
      if (pc.pcnt) { /* count the pieces of rows without a stored bitmap */               \
        hipLaunchKernelGGL((k_num_rank1<T, B>), dim3(grid), dim3(B), lds, s, kargs<T>(a), \
                           rows, nrows, counter, logT, ts, T1, tsw, wper, bbase,          \
                           (const uint4 *)bm, bmoff, cap, kout<T>(o), sel.bounds,         \
                           sel.err, pc.kp, pc.extra, pc.n_extra, pc.pcnt, 1u);            \
        LAUNCH_CHECK();                                                                   \
        OC_TRY(zero_u32(counter, 1, s));                                                  \
      }                                                                                   \
    }                                                                                     \
    hipLaunchKernelGGL((k_num_rank1<T, B>), dim3(grid), dim3(B), lds, s, kargs<T>(a),     \
                       rows, nrows, counter, logT, ts, T1, tsw, wper, bbase,              \
                       (const uint4 *)bm, bmoff, cap, kout<T>(o), sel.bounds, sel.err,    \
                       split ? pc.kp : nullptr, pc.extra, pc.n_extra, pc.pcnt, 0u);       \
  } while (0)
  if (a.A.vtype == OC_F64) {
    if (bs == 512) GO(double, 512);
    else GO(double, 1024);
  } else {
    if (bs == 512) GO(float, 512);
    else GO(float, 1024);
  }
#undef GO
  LAUNCH_CHECK();
  return OC_OK;
}

uint32_t rank1_wide_log(uint32_t bs) {
  return bs == 512 ? R1Shape<512>::MAXLOG : R1Shape<1024>::MAXLOG;
}

// ------------------------------------------------------ stored row bitmaps
// Rows of the dense symbolic bins with at least min_products products get
// a bitmap region: n4 = ((right - (left & ~127)) >> 7) + 1 uint4 groups.
__global__ __launch_bounds__(256) void k_bm_sizes(const uint32_t *__restrict__ rows, uint32_t n,
                                                  const uint32_t *__restrict__ products,
                                                  const uint32_t *__restrict__ left,
                                                  const uint32_t *__restrict__ right,
                                                  uint32_t min_products, uint32_t *__restrict__ sz) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t r = rows[i];
  const uint32_t l = left[r], h = right[r];
  sz[i] = (products[r] >= min_products && h >= l) ? ((h - (l & ~127u)) >> 7) + 1 : 0u;
}

__global__ __launch_bounds__(256) void k_bm_fill(uint64_t *__restrict__ bmoff, uint64_t m) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * blockDim.x)
    bmoff[i] = ~0ull;
}

__global__ __launch_bounds__(256) void k_bm_scatter(const uint32_t *__restrict__ rows, uint32_t n,
                                                    const uint32_t *__restrict__ sz,
                                                    const uint64_t *__restrict__ offs,
                                                    uint64_t *__restrict__ bmoff) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (sz[i]) bmoff[rows[i]] = offs[i];
}

int launch_bitmap_offsets(const uint32_t *rows, uint32_t n, uint64_t m,
                          const uint32_t *products, const uint32_t *left,
                          const uint32_t *right, uint32_t min_products, uint32_t *sz,
                          uint64_t *offs, void *scan_tmp, uint64_t *bmoff, hipStream_t s) {
  hipLaunchKernelGGL(k_bm_fill, dim3((uint32_t)std::min<uint64_t>(cdiv(m, 256), 4096)), dim3(256),
                     0, s, bmoff, m);
  LAUNCH_CHECK();
  if (n == 0) return OC_OK;
  hipLaunchKernelGGL(k_bm_sizes, dim3(cdiv(n, 256)), dim3(256), 0, s, rows, n, products, left,
                     right, min_products, sz);
  LAUNCH_CHECK();
  OC_TRY(launch_scan_u32(sz, offs, n, scan_tmp, s));
  hipLaunchKernelGGL(k_bm_scatter, dim3(cdiv(n, 256)), dim3(256), 0, s, rows, n, sz, offs, bmoff);
  LAUNCH_CHECK();
  return OC_OK;
}

}  // namespace oc
