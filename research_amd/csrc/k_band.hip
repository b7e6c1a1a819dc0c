// Long rows swept band by band across all rows (band-major order).
//
// The columns are cut into bands of 2^logW.  Work items are (band t, group
// of kBandGroup consecutive long rows), taken from an atomic counter in
// band-major order, so at any time the GPU works on one band: the slice of
// B that falls into that band (about B / bands bytes) stays in each XCD's
// L2 while every long row gathers from it.  Row-major sweeps touch all of B
// at once and miss L2 on most gathers.
//
// Per (row, band): the A row is staged in chunks, the B segments of the
// band come from a band-major start table, the segments' products are
// flattened and enumerated through an LDS owner map (one LDS read per
// product), their columns are ORed into a band bitmap, and a popcount
// prefix gives every output column its rank.  COUNT mode stores the
// number of outputs of each (row, band); numeric mode adds the products
// into FP64 slots indexed by rank (in rank ranges when the band has more
// outputs than slots) and writes them, in column order, at the row's
// offset plus the outputs of its earlier bands.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "dev_common.hpp"
#include "launch.hpp"

namespace oc {

#define HIP_TRY(call)                                                         \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess) {                                                   \
      set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,                    \
                hipGetErrorString(e_));                                       \
      return e_ == hipErrorOutOfMemory ? OC_ERR_NOMEM : OC_ERR_HIP;           \
    }                                                                         \
  } while (0)
#define LAUNCH_CHECK() HIP_TRY(hipGetLastError())

constexpr int kBandBlock = 256;
constexpr uint32_t kBandA = kBandBlock;              // A entries per chunk
constexpr uint32_t kBandRound = 8 * kBandBlock;      // products per owner map
constexpr uint32_t kBandGroup = 8;                   // rows per work item
constexpr uint32_t kBandMaxLog = 15;                 // widest band 2^15 columns
constexpr uint32_t kBandSlots = 1024;                // FP64 rank slots (numeric)

template <typename T>
struct BandArgs {
  const uint32_t *a_rp, *a_ci;
  const T *a_v;
  const uint32_t *b_ci;
  const T *b_v;
  const uint32_t *left, *right;
  const uint32_t *rows;  // long rows, ascending
  uint32_t nrows, nbands, logW, b_rows;
  const uint32_t *ts;    // band-major starts: ts[t * b_rows + k], t = 0..nbands
  uint32_t *counter;
  uint32_t *cnt;         // COUNT: outputs of (row j, band t) at j * nbands + t
  const uint64_t *pref;  // numeric: exclusive prefix of cnt (u64)
  const uint64_t *c_off; // numeric: C row offsets
  uint32_t *c_cols;
  T *c_vals;
};

static size_t band_lds(bool num, uint32_t logW) {
  const size_t words = (size_t)1 << (logW - 5);
  size_t b = 256 + (size_t)kBandA * 8 + (size_t)kBandRound * 2 + words * 4;
  if (num) b += (size_t)kBandA * 8 + words * 4 + (size_t)kBandSlots * 12;
  return b;
}

template <typename T, bool NUM>
__global__ __launch_bounds__(kBandBlock) void k_band(BandArgs<T> g) {
  extern __shared__ __align__(16) uint8_t smem[];
  typedef unsigned short us8 __attribute__((ext_vector_type(8)));
  const uint32_t nwords = 1u << (g.logW - 5);
  uint32_t *hdr = reinterpret_cast<uint32_t *>(smem);
  uint32_t *sk = reinterpret_cast<uint32_t *>(smem + 256);
  uint32_t *sbase = sk + kBandA;
  uint16_t *own = reinterpret_cast<uint16_t *>(sbase + kBandA);
  uint32_t *bits = reinterpret_cast<uint32_t *>(own + kBandRound);
  uint32_t *wpre = bits + nwords;                 // numeric only
  double *sav = reinterpret_cast<double *>(wpre + nwords);
  double *acc = sav + kBandA;
  uint32_t *cmap = reinterpret_cast<uint32_t *>(acc + kBandSlots);
  const uint32_t tid = threadIdx.x;
  const uint32_t wpt = (nwords + kBandBlock - 1) / kBandBlock;
  const uint32_t b0 = min(nwords, tid * wpt), b1 = min(nwords, b0 + wpt);
  for (uint32_t i = tid; i < nwords; i += kBandBlock) bits[i] = 0;
  if (NUM)
    for (uint32_t i = tid; i < kBandSlots; i += kBandBlock) acc[i] = 0.0;
  const uint32_t ngroups = (g.nrows + kBandGroup - 1) / kBandGroup;
  const uint64_t nitems = (uint64_t)ngroups * g.nbands;
  __syncthreads();
  uint32_t my_ex = 0, my_len = 0;

  auto build_map = [&](uint32_t p0, uint32_t n) {
    __syncthreads();  // the previous map is no longer read
    if (8 * tid < n) *reinterpret_cast<us8 *>(own + 8 * tid) = us8{0, 0, 0, 0, 0, 0, 0, 0};
    __syncthreads();
    if (my_len > 0 && my_ex < p0 + n && my_ex + my_len > p0)
      own[my_ex > p0 ? my_ex - p0 : 0] = (uint16_t)tid;
    __syncthreads();
    us8 v = (8 * tid < n) ? *reinterpret_cast<const us8 *>(own + 8 * tid)
                          : us8{0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t run = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      run = max(run, (uint32_t)v[j]);
      v[j] = (unsigned short)run;
    }
    const uint32_t pre = block_excl_max(run, hdr + 32);
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = (unsigned short)max(pre, (uint32_t)v[j]);
    if (8 * tid < n) *reinterpret_cast<us8 *>(own + 8 * tid) = v;
    __syncthreads();
  };
  auto locate = [&](uint32_t p0, uint32_t n, uint32_t (&q)[8], uint32_t (&w)[8]) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const uint32_t i = tid + (uint32_t)u * kBandBlock;
      q[u] = kEmpty;
      w[u] = 0;
      if (i < n) {
        const uint32_t ow = own[i];
        q[u] = sbase[ow] + p0 + i;
        w[u] = ow;
      }
    }
  };

  for (;;) {
    if (tid == 0) hdr[0] = atomicAdd(g.counter, 1u);
    __syncthreads();
    const uint32_t item = hdr[0];
    __syncthreads();
    if ((uint64_t)item >= nitems) break;
    const uint32_t t = item / ngroups, grp = item % ngroups;
    const uint32_t j1 = min(g.nrows, (grp + 1) * kBandGroup);
    const uint64_t tcol0 = (uint64_t)t << g.logW;
    for (uint32_t j = grp * kBandGroup; j < j1; ++j) {
      const uint32_t r = g.rows[j];
      const uint32_t a0 = g.a_rp[r], a1 = g.a_rp[r + 1];
      const uint32_t lo = g.left[r], hi = g.right[r];
      const bool inside = a1 > a0 && hi >= lo && t >= (lo >> g.logW) && t <= (hi >> g.logW);
      if (!inside) {
        if (!NUM && tid == 0) g.cnt[(uint64_t)j * g.nbands + t] = 0;
        continue;
      }
      const bool one_chunk = a1 - a0 <= kBandA;
      auto stage_bounds = [&](uint32_t c0, uint32_t nc, uint32_t &P) {
        uint32_t len = 0, st = 0;
        if (tid < nc) {
          const uint32_t k = g.a_ci[c0 + tid];
          sk[tid] = k;
          if (NUM) sav[tid] = (double)g.a_v[c0 + tid];
          st = g.ts[(uint64_t)t * g.b_rows + k];
          len = g.ts[(uint64_t)(t + 1) * g.b_rows + k] - st;
        }
        const uint32_t ex = block_excl_sum(len, hdr + 8, &P);
        my_ex = ex;
        my_len = len;
        if (tid < nc) sbase[tid] = st - ex;
        __syncthreads();
      };
      // ---- pass 1: mark the output columns of the band
      uint32_t any = 0, Plast = 0;
      bool keep = false;
      uint32_t kc[8];
      for (uint32_t c0 = a0; c0 < a1; c0 += kBandA) {
        const uint32_t nc = min(kBandA, a1 - c0);
        uint32_t P;
        stage_bounds(c0, nc, P);
        any |= P;
        Plast = P;
        keep = one_chunk && P <= kBandRound;
        for (uint32_t p0 = 0; p0 < P; p0 += kBandRound) {
          const uint32_t n = min(kBandRound, P - p0);
          build_map(p0, n);
          uint32_t q[8], w[8];
          locate(p0, n, q, w);
#pragma unroll
          for (int u = 0; u < 8; ++u) kc[u] = q[u] != kEmpty ? g.b_ci[q[u]] : kEmpty;
          uint32_t cw = kEmpty, cm = 0;
#pragma unroll
          for (int u = 0; u < 8; ++u)
            if (kc[u] != kEmpty) {
              const uint32_t x = (uint32_t)(kc[u] - tcol0);
              if ((x >> 5) != cw) {
                if (cm) atomicOr(&bits[cw], cm);
                cw = x >> 5;
                cm = 0;
              }
              cm |= 1u << (x & 31);
            }
          if (cm) atomicOr(&bits[cw], cm);
        }
        __syncthreads();  // the chunk's staging is reused
      }
      if (!any) {
        if (!NUM && tid == 0) g.cnt[(uint64_t)j * g.nbands + t] = 0;
        continue;
      }
      // ---- outputs of the band: popcount (and its prefix for the ranks)
      uint32_t mycnt = 0;
      for (uint32_t i = b0; i < b1; ++i) mycnt += __popc(bits[i]);
      uint32_t m;
      uint32_t run = block_excl_sum(mycnt, hdr + 8, &m);
      if (!NUM) {
        if (tid == 0) g.cnt[(uint64_t)j * g.nbands + t] = m;
        for (uint32_t i = b0; i < b1; ++i) bits[i] = 0;
        __syncthreads();
        continue;
      }
      for (uint32_t i = b0; i < b1; ++i) {
        wpre[i] = run;
        run += __popc(bits[i]);
      }
      __syncthreads();
      const uint64_t base = g.c_off[r] + (g.pref[(uint64_t)j * g.nbands + t] -
                                          g.pref[(uint64_t)j * g.nbands]);
      // ---- pass 2: sums by rank, in rank ranges of kBandSlots
      for (uint32_t r0 = 0; r0 < m; r0 += kBandSlots) {
        const uint32_t r1 = min(m, r0 + kBandSlots);
        if (keep) {
          // the single round's owner map and staging are still in LDS
          uint32_t q[8], w[8];
          locate(0, Plast, q, w);
          double pv[8];
#pragma unroll
          for (int u = 0; u < 8; ++u)
            pv[u] = q[u] != kEmpty ? sav[w[u]] * (double)g.b_v[q[u]] : 0.0;
#pragma unroll
          for (int u = 0; u < 8; ++u)
            if (kc[u] != kEmpty) {
              const uint32_t x = (uint32_t)(kc[u] - tcol0);
              const uint32_t wd = x >> 5;
              const uint32_t rk = wpre[wd] + __popc(bits[wd] & ((1u << (x & 31)) - 1u));
              if (rk >= r0 && rk < r1) {
                atomicAdd(&acc[rk - r0], pv[u]);
                cmap[rk - r0] = kc[u];
              }
            }
        } else {
          for (uint32_t c0 = a0; c0 < a1; c0 += kBandA) {
            const uint32_t nc = min(kBandA, a1 - c0);
            uint32_t P;
            stage_bounds(c0, nc, P);
            for (uint32_t p0 = 0; p0 < P; p0 += kBandRound) {
              const uint32_t n = min(kBandRound, P - p0);
              build_map(p0, n);
              uint32_t q[8], w[8];
              locate(p0, n, q, w);
              uint32_t col[8];
              double pv[8];
#pragma unroll
              for (int u = 0; u < 8; ++u) {
                col[u] = q[u] != kEmpty ? g.b_ci[q[u]] : kEmpty;
                pv[u] = q[u] != kEmpty ? sav[w[u]] * (double)g.b_v[q[u]] : 0.0;
              }
#pragma unroll
              for (int u = 0; u < 8; ++u)
                if (col[u] != kEmpty) {
                  const uint32_t x = (uint32_t)(col[u] - tcol0);
                  const uint32_t wd = x >> 5;
                  const uint32_t rk = wpre[wd] + __popc(bits[wd] & ((1u << (x & 31)) - 1u));
                  if (rk >= r0 && rk < r1) {
                    atomicAdd(&acc[rk - r0], pv[u]);
                    cmap[rk - r0] = col[u];
                  }
                }
            }
            __syncthreads();
          }
        }
        __syncthreads();
        for (uint32_t i = tid; i < r1 - r0; i += kBandBlock) {
          g.c_cols[base + r0 + i] = cmap[i];
          g.c_vals[base + r0 + i] = (T)acc[i];
          acc[i] = 0.0;
        }
        __syncthreads();
      }
      for (uint32_t i = b0; i < b1; ++i) bits[i] = 0;
      __syncthreads();
    }
  }
}

// Band-major start table: ts[t * rows + k] = first entry of B row k with
// column >= t * 2^logW, for t = 0 .. nbands.
__global__ __launch_bounds__(256) void k_band_starts(const uint32_t *__restrict__ b_rp,
                                                     const uint32_t *__restrict__ b_ci,
                                                     uint32_t rows, uint32_t logW,
                                                     uint32_t nbands, uint32_t *__restrict__ ts) {
  const int lane = lane_id();
  for (uint32_t k = blockIdx.x * 4 + threadIdx.x / kWave; k < rows; k += gridDim.x * 4) {
    const uint32_t s = b_rp[k], e = b_rp[k + 1];
    for (uint32_t t = lane; t <= nbands; t += kWave) {
      const uint64_t bound = (uint64_t)t << logW;
      uint32_t lo = s, hi = e;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((uint64_t)b_ci[mid] < bound) lo = mid + 1;
        else hi = mid;
      }
      ts[(uint64_t)t * rows + k] = lo;
    }
  }
}

// Long-row selection: flag[r] = products[r] >= min_products and the row's
// column span is wider than one band.
__global__ void k_band_flags(const uint32_t *__restrict__ products,
                             const uint32_t *__restrict__ left,
                             const uint32_t *__restrict__ right, uint64_t rows,
                             uint32_t min_products, uint32_t band_cols,
                             uint32_t *__restrict__ flag) {
  for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < rows;
       r += (uint64_t)gridDim.x * blockDim.x) {
    const bool wide = right[r] >= left[r] && right[r] - left[r] >= band_cols;
    flag[r] = (products[r] >= min_products && wide) ? 1u : 0u;
  }
}

// Rows with flag set, in ascending order, from the exclusive prefix.
__global__ void k_band_compact(const uint32_t *__restrict__ flag,
                               const uint64_t *__restrict__ pos, uint64_t rows,
                               uint32_t *__restrict__ out) {
  for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < rows;
       r += (uint64_t)gridDim.x * blockDim.x)
    if (flag[r]) out[pos[r]] = (uint32_t)r;
}

// nnz of each long row from the prefix of its band counts.
__global__ void k_band_row_nnz(const uint32_t *__restrict__ rows, uint32_t nrows,
                               uint32_t nbands, const uint64_t *__restrict__ pref,
                               uint32_t *__restrict__ nnz) {
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nrows;
       j += gridDim.x * blockDim.x)
    nnz[rows[j]] = (uint32_t)(pref[(uint64_t)(j + 1) * nbands] - pref[(uint64_t)j * nbands]);
}

static uint32_t grid1(uint64_t n, uint32_t bs) {
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + bs - 1) / bs, 1u << 16));
}

int band_plan(uint64_t b_cols, uint32_t log2_cols, uint32_t *nbands) {
  if (log2_cols < 10 || log2_cols > kBandMaxLog || b_cols == 0) return 0;
  const uint64_t nb = ((b_cols - 1) >> log2_cols) + 1;
  if (nb > 65536) return 0;
  *nbands = (uint32_t)nb;
  return 1;
}

int launch_band_starts(const DevCsr &B, uint32_t logW, uint32_t nbands, uint32_t *ts,
                       hipStream_t s) {
  if (B.rows == 0) return OC_OK;
  hipLaunchKernelGGL(k_band_starts, dim3(std::min<uint64_t>((B.rows + 3) / 4, 1u << 20)),
                     dim3(256), 0, s, B.rp, B.ci, (uint32_t)B.rows, logW, nbands, ts);
  LAUNCH_CHECK();
  return OC_OK;
}

int launch_band_select(const RowInfo &ri, uint64_t rows, uint32_t min_products,
                       uint32_t band_cols, uint32_t *flag, uint64_t *pos, void *scan_tmp,
                       uint32_t *list, hipStream_t s) {
  hipLaunchKernelGGL(k_band_flags, dim3(grid1(rows, 256)), dim3(256), 0, s, ri.products,
                     ri.left, ri.right, rows, min_products, band_cols, flag);
  LAUNCH_CHECK();
  OC_TRY(launch_scan_u32(flag, pos, rows, scan_tmp, s));
  hipLaunchKernelGGL(k_band_compact, dim3(grid1(rows, 256)), dim3(256), 0, s, flag, pos, rows,
                     list);
  LAUNCH_CHECK();
  return OC_OK;
}

int launch_band_row_nnz(const uint32_t *rows, uint32_t nrows, uint32_t nbands,
                        const uint64_t *pref, uint32_t *nnz, hipStream_t s) {
  if (!nrows) return OC_OK;
  hipLaunchKernelGGL(k_band_row_nnz, dim3(grid1(nrows, 256)), dim3(256), 0, s, rows, nrows,
                     nbands, pref, nnz);
  LAUNCH_CHECK();
  return OC_OK;
}

static int zero_counter(uint32_t *c, hipStream_t s);

template <typename T, bool NUM>
static int band_go(const AccArgs &a, const uint32_t *rows, uint32_t nrows, uint32_t logW,
                   uint32_t nbands, const uint32_t *ts, uint32_t *cnt, const uint64_t *pref,
                   const NumOut *o, uint32_t *counter, hipStream_t s) {
  BandArgs<T> g;
  g.a_rp = a.A.rp;
  g.a_ci = a.A.ci;
  g.a_v = (const T *)a.A.v;
  g.b_ci = a.B.ci;
  g.b_v = (const T *)a.B.v;
  g.left = a.left;
  g.right = a.right;
  g.rows = rows;
  g.nrows = nrows;
  g.nbands = nbands;
  g.logW = logW;
  g.b_rows = (uint32_t)a.B.rows;
  g.ts = ts;
  g.counter = counter;
  g.cnt = cnt;
  g.pref = pref;
  g.c_off = o ? o->off : nullptr;
  g.c_cols = o ? o->cols : nullptr;
  g.c_vals = o ? (T *)o->vals : nullptr;
  const size_t lds = band_lds(NUM, logW);
  const Tables &tb = tables();
  if (lds > tb.lds_max_block) {
    set_error("band kernel needs %zu B of LDS", lds);
    return OC_ERR_INVALID;
  }
  if (lds > 65536)
    HIP_TRY(hipFuncSetAttribute((const void *)k_band<T, NUM>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  // resident workgroups per CU: VGPRs allow 8 (count) or 5 (numeric)
  const uint32_t per_cu = std::max<uint32_t>(1, std::min<uint32_t>(NUM ? 5 : 8, (uint32_t)(tb.lds_per_cu / lds)));
  const uint64_t items = (uint64_t)((nrows + kBandGroup - 1) / kBandGroup) * nbands;
  const uint32_t grid = (uint32_t)std::min<uint64_t>(items, (uint64_t)std::max<uint32_t>(1, tb.num_cu) * per_cu);
  OC_REQUIRE(items < (1ull << 31), "too many band work items");
  OC_TRY(zero_counter(counter, s));
  hipLaunchKernelGGL((k_band<T, NUM>), dim3(grid), dim3(kBandBlock), lds, s, g);
  LAUNCH_CHECK();
  return OC_OK;
}

__global__ void k_band_zero(uint32_t *p) {
  if (threadIdx.x == 0) p[0] = 0;
}
static int zero_counter(uint32_t *c, hipStream_t s) {
  hipLaunchKernelGGL(k_band_zero, dim3(1), dim3(64), 0, s, c);
  LAUNCH_CHECK();
  return OC_OK;
}

int launch_band_count(const AccArgs &a, const uint32_t *rows, uint32_t nrows, uint32_t logW,
                      uint32_t nbands, const uint32_t *ts, uint32_t *cnt, uint32_t *counter,
                      hipStream_t s) {
  if (!nrows) return OC_OK;
  if (a.A.vtype == OC_F64)
    return band_go<double, false>(a, rows, nrows, logW, nbands, ts, cnt, nullptr, nullptr,
                                  counter, s);
  return band_go<float, false>(a, rows, nrows, logW, nbands, ts, cnt, nullptr, nullptr,
                               counter, s);
}

int launch_band_numeric(const AccArgs &a, const uint32_t *rows, uint32_t nrows, uint32_t logW,
                        uint32_t nbands, const uint32_t *ts, const uint64_t *pref,
                        const NumOut &o, uint32_t *counter, hipStream_t s) {
  if (!nrows) return OC_OK;
  if (a.A.vtype == OC_F64)
    return band_go<double, true>(a, rows, nrows, logW, nbands, ts, nullptr, pref, &o, counter,
                                 s);
  return band_go<float, true>(a, rows, nrows, logW, nbands, ts, nullptr, pref, &o, counter, s);
}

}  // namespace oc
