// Spy-plot segmentation of A (analysis only: C does not depend on it), the
// role of Segmentation.cuh:374-673.
//
//   density   A's nonzeros counted per tile of a grid_h x grid_w grid.  Each
//             workgroup takes a run of rows inside one grid row and counts
//             into grid_w LDS bins, then adds its nonzero bins to the grid.
//   grid      two single-workgroup launches: the mean and deviation of the
//             tile counts and hot tiles (count >= mean + k * deviation,
//             count > 0) labelled by their flat index; then synchronous
//             4-neighbour min-label sweeps (at most cc_max_iters) and
//             pointer-jumping rounds, so a hot tile ends with the smallest
//             index of its connected component once the sweeps converge
//             (no host round trip per sweep).
//   dominant  per grid row, the label held by most hot tiles (ties to the
//             smaller label) and the row's hot-tile count.
//   classify  per row segment: sums and extremes over its rows of the row
//             analysis, then the reference's class rule.
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

constexpr int kSegBlock = 256;
constexpr int kSegGridBlock = 1024;
constexpr uint32_t kSegMaxW = 4096;  // LDS bins of the density pass

__global__ __launch_bounds__(kSegBlock) void k_seg_density(
    const uint32_t *__restrict__ a_rp, const uint32_t *__restrict__ a_ci, uint32_t a_rows,
    uint32_t grid_h, uint32_t grid_w, uint32_t tile_rows, uint32_t tile_cols,
    uint32_t chunks_per_grid_row, uint32_t chunk_rows, uint32_t *__restrict__ density) {
  __shared__ uint32_t bins[kSegMaxW];
  const uint32_t g = blockIdx.x / chunks_per_grid_row, ch = blockIdx.x % chunks_per_grid_row;
  for (uint32_t i = threadIdx.x; i < grid_w; i += kSegBlock) bins[i] = 0;
  __syncthreads();
  const uint64_t g0 = (uint64_t)g * tile_rows;
  const uint64_t g1 = g + 1 == grid_h ? a_rows : std::min<uint64_t>(a_rows, g0 + tile_rows);
  const uint64_t r0 = g0 + (uint64_t)ch * chunk_rows;
  const uint64_t r1 = std::min<uint64_t>(g1, r0 + chunk_rows);
  // one wave per row, lanes over the row's entries
  for (uint64_t r = r0 + threadIdx.x / kWave; r < r1; r += kSegBlock / kWave) {
    const uint32_t e = a_rp[r + 1];
    for (uint32_t j = a_rp[r] + lane_id(); j < e; j += kWave) {
      uint32_t tc = a_ci[j] / tile_cols;
      if (tc >= grid_w) tc = grid_w - 1;
      atomicAdd(&bins[tc], 1u);
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < grid_w; i += kSegBlock)
    if (bins[i]) atomicAdd(&density[(uint64_t)g * grid_w + i], bins[i]);
}

// out[0..1] = sum and sum of squares of the tile counts (u64); labels[] =
// flat index of each hot tile, ~0 for the others
__global__ __launch_bounds__(kSegGridBlock) void k_seg_threshold(
    const uint32_t *__restrict__ density, uint32_t n, double k, uint32_t *__restrict__ labels,
    unsigned long long *__restrict__ out) {
  __shared__ unsigned long long red[2][kSegGridBlock / kWave];
  __shared__ double stats[2];
  unsigned long long s = 0, sq = 0;
  for (uint32_t i = threadIdx.x; i < n; i += kSegGridBlock) {
    const unsigned long long v = density[i];
    s += v;
    sq += v * v;
  }
  s = group_sum(s, kWave);
  sq = group_sum(sq, kWave);
  if (lane_id() == 0) {
    red[0][threadIdx.x / kWave] = s;
    red[1][threadIdx.x / kWave] = sq;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long ts = 0, tq = 0;
    for (int w = 0; w < kSegGridBlock / kWave; ++w) {
      ts += red[0][w];
      tq += red[1][w];
    }
    out[0] = ts;
    out[1] = tq;
    const double mean = (double)ts / (double)n;
    double var = (double)tq / (double)n - mean * mean;
    if (var < 0.0) var = 0.0;
    stats[0] = mean;
    stats[1] = sqrt(var);
  }
  __syncthreads();
  const double thresh = stats[0] + k * stats[1];
  for (uint32_t i = threadIdx.x; i < n; i += kSegGridBlock) {
    const uint32_t d = density[i];
    labels[i] = ((double)d >= thresh && d > 0u) ? i : kEmpty;
  }
}

// Synchronous sweeps (next = min over the hot 4-neighbourhood), at most
// max_sweeps of them, then up to 8 synchronous pointer-jumping rounds
// (label = label[label]), each stage stopping early once nothing changes
// (the bounded propagation and flattening of Segmentation.cuh:495-537).
// Labels are always flat indices of hot tiles.  Converged sweeps leave each
// hot tile with the smallest index of its 4-connected component.
// out[2] = sweeps that changed a label.
__global__ __launch_bounds__(kSegGridBlock) void k_seg_components(
    uint32_t grid_h, uint32_t grid_w, uint32_t max_sweeps, uint32_t *__restrict__ labels,
    uint32_t *__restrict__ next, unsigned long long *__restrict__ out) {
  __shared__ int changed_any;
  const uint32_t n = grid_h * grid_w;
  uint32_t sweeps = 0;
  for (uint32_t it = 0; it < max_sweeps; ++it) {
    if (threadIdx.x == 0) changed_any = 0;
    __syncthreads();
    int mine = 0;
    for (uint32_t i = threadIdx.x; i < n; i += kSegGridBlock) {
      uint32_t best = labels[i];
      if (best != kEmpty) {
        const uint32_t x = i % grid_w, y = i / grid_w;
        if (x > 0) best = min(best, labels[i - 1]);
        if (x + 1 < grid_w) best = min(best, labels[i + 1]);
        if (y > 0) best = min(best, labels[i - grid_w]);
        if (y + 1 < grid_h) best = min(best, labels[i + grid_w]);
        mine |= best != labels[i];
      }
      next[i] = best;
    }
    if (mine) changed_any = 1;
    __syncthreads();
    const int ch = changed_any;
    for (uint32_t i = threadIdx.x; i < n; i += kSegGridBlock) labels[i] = next[i];
    __syncthreads();
    if (!ch) break;
    ++sweeps;
  }
  for (int round = 0; round < 8; ++round) {
    if (threadIdx.x == 0) changed_any = 0;
    __syncthreads();
    int mine = 0;
    for (uint32_t i = threadIdx.x; i < n; i += kSegGridBlock) {
      const uint32_t l = labels[i];
      const uint32_t p = l == kEmpty ? kEmpty : labels[l];
      mine |= p != l;
      next[i] = p;
    }
    if (mine) changed_any = 1;
    __syncthreads();
    const int ch = changed_any;
    for (uint32_t i = threadIdx.x; i < n; i += kSegGridBlock) labels[i] = next[i];
    __syncthreads();
    if (!ch) break;
  }
  if (threadIdx.x == 0) out[2] = sweeps;
}

// per grid row: label held by most hot tiles (ties: smaller label), hot count
__global__ __launch_bounds__(kSegBlock) void k_seg_dominant(const uint32_t *__restrict__ labels,
                                                            uint32_t grid_w,
                                                            uint32_t *__restrict__ row_label,
                                                            uint32_t *__restrict__ row_hot) {
  __shared__ unsigned long long best[kSegBlock / kWave];
  __shared__ uint32_t hot[kSegBlock / kWave];
  const uint32_t *row = labels + (uint64_t)blockIdx.x * grid_w;
  // key = count << 32 | ~label: the max key is the most common label,
  // the smaller label on ties
  unsigned long long key = 0;
  uint32_t nhot = 0;
  for (uint32_t x = threadIdx.x; x < grid_w; x += kSegBlock) {
    const uint32_t l = row[x];
    if (l == kEmpty) continue;
    ++nhot;
    uint32_t c = 0;
    for (uint32_t y = 0; y < grid_w; ++y) c += row[y] == l;
    const unsigned long long kk = ((unsigned long long)c << 32) | (uint32_t)~l;
    key = kk > key ? kk : key;
  }
  key = group_max(key, kWave);
  nhot = group_sum(nhot, kWave);
  if (lane_id() == 0) {
    best[threadIdx.x / kWave] = key;
    hot[threadIdx.x / kWave] = nhot;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long b = 0;
    uint32_t h = 0;
    for (int w = 0; w < kSegBlock / kWave; ++w) {
      b = best[w] > b ? best[w] : b;
      h += hot[w];
    }
    row_label[blockIdx.x] = b ? ~(uint32_t)(b & 0xFFFFFFFFull) : kEmpty;
    row_hot[blockIdx.x] = h;
  }
}

// per segment [start, end): u64 sums of products, A entries and squared
// products; max B row, max A row, min left, max right (one workgroup each)
__global__ __launch_bounds__(kSegBlock) void k_seg_reduce(
    const uint32_t *__restrict__ seg_start, const uint32_t *__restrict__ seg_end,
    const uint32_t *__restrict__ a_rp, const uint32_t *__restrict__ products,
    const uint32_t *__restrict__ max_b, const uint32_t *__restrict__ left,
    const uint32_t *__restrict__ right, SegSums *__restrict__ out) {
  __shared__ unsigned long long r64[3][kSegBlock / kWave];
  __shared__ uint32_t r32[4][kSegBlock / kWave];
  const uint32_t s0 = seg_start[blockIdx.x], s1 = seg_end[blockIdx.x];
  unsigned long long sp = 0, sa = 0, sq = 0;
  uint32_t mb = 0, ma = 0, lo = kEmpty, hi = 0;
  for (uint32_t r = s0 + threadIdx.x; r < s1; r += kSegBlock) {
    const unsigned long long p = products[r];
    sp += p;
    sq += p * p;
    const uint32_t na = a_rp[r + 1] - a_rp[r];
    sa += na;
    ma = max(ma, na);
    mb = max(mb, max_b[r]);
    lo = min(lo, left[r]);
    hi = max(hi, right[r]);
  }
  sp = group_sum(sp, kWave);
  sa = group_sum(sa, kWave);
  sq = group_sum(sq, kWave);
  mb = group_max(mb, kWave);
  ma = group_max(ma, kWave);
  lo = group_min(lo, kWave);
  hi = group_max(hi, kWave);
  const int w = threadIdx.x / kWave;
  if (lane_id() == 0) {
    r64[0][w] = sp;
    r64[1][w] = sa;
    r64[2][w] = sq;
    r32[0][w] = mb;
    r32[1][w] = ma;
    r32[2][w] = lo;
    r32[3][w] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    SegSums t{};
    t.min_left = kEmpty;
    for (int i = 0; i < kSegBlock / kWave; ++i) {
      t.sum_products += r64[0][i];
      t.sum_a_nnz += r64[1][i];
      t.sum_sq_products += r64[2][i];
      t.max_b_row = max(t.max_b_row, r32[0][i]);
      t.max_a_row = max(t.max_a_row, r32[1][i]);
      t.min_left = min(t.min_left, r32[2][i]);
      t.max_right = max(t.max_right, r32[3][i]);
    }
    out[blockIdx.x] = t;
  }
}

int launch_seg_density(const DevCsr &A, uint32_t grid_h, uint32_t grid_w, uint32_t tile_rows,
                       uint32_t tile_cols, uint32_t *density, hipStream_t s) {
  if (A.rows == 0 || grid_h == 0 || grid_w == 0) return OC_OK;
  if (grid_w > kSegMaxW) {
    set_error("segmentation: grid width %u beyond %u", grid_w, kSegMaxW);
    return OC_ERR_INVALID;
  }
  const uint32_t chunk_rows = 512;
  const uint32_t chunks = (tile_rows + chunk_rows - 1) / chunk_rows;
  const uint64_t grid = (uint64_t)grid_h * chunks;
  hipLaunchKernelGGL(k_seg_density, dim3((uint32_t)grid), dim3(kSegBlock), 0, s, A.rp, A.ci,
                     (uint32_t)A.rows, grid_h, grid_w, tile_rows, tile_cols, chunks, chunk_rows,
                     density);
  LAUNCH_CHECK();
  return OC_OK;
}

int launch_seg_threshold(const uint32_t *density, uint32_t n, double k, uint32_t *labels,
                         unsigned long long *out, hipStream_t s) {
  hipLaunchKernelGGL(k_seg_threshold, dim3(1), dim3(kSegGridBlock), 0, s, density, n, k, labels,
                     out);
  LAUNCH_CHECK();
  return OC_OK;
}

int launch_seg_components(uint32_t grid_h, uint32_t grid_w, uint32_t max_sweeps,
                          uint32_t *labels, uint32_t *next, unsigned long long *out,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_seg_components, dim3(1), dim3(kSegGridBlock), 0, s, grid_h, grid_w,
                     max_sweeps, labels, next, out);
  LAUNCH_CHECK();
  return OC_OK;
}

int launch_seg_dominant(const uint32_t *labels, uint32_t grid_h, uint32_t grid_w,
                        uint32_t *row_label, uint32_t *row_hot, hipStream_t s) {
  hipLaunchKernelGGL(k_seg_dominant, dim3(grid_h), dim3(kSegBlock), 0, s, labels, grid_w,
                     row_label, row_hot);
  LAUNCH_CHECK();
  return OC_OK;
}

int launch_seg_reduce(const uint32_t *seg_start, const uint32_t *seg_end, uint32_t nseg,
                      const DevCsr &A, const RowInfo &ri, SegSums *out, hipStream_t s) {
  if (nseg == 0) return OC_OK;
  hipLaunchKernelGGL(k_seg_reduce, dim3(nseg), dim3(kSegBlock), 0, s, seg_start, seg_end, A.rp,
                     ri.products, ri.max_b, ri.left, ri.right, out);
  LAUNCH_CHECK();
  return OC_OK;
}

}  // namespace oc
