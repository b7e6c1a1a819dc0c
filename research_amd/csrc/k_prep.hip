// gfx950 kernels ahead of the accumulators: row analysis, device scan,
// MurmurHash3 / HyperLogLog, binning, row copy and validation helpers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "dev_common.hpp"
#include "launch.hpp"

namespace oc {

#define HIP_TRY(call)                                                         \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess) {                                                   \
      set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,                    \
                hipGetErrorString(e_));                                       \
      return OC_ERR_HIP;                                                      \
    }                                                                         \
  } while (0)

#define LAUNCH_CHECK() HIP_TRY(hipGetLastError())

static inline uint32_t cdiv(uint64_t a, uint64_t b) {
  return (uint32_t)((a + b - 1) / b);
}

// ------------------------------------------------------------------ tables
static Tables g_tab;
static bool g_tab_ready = false;
static uint32_t g_lds_cu = 163840, g_lds_blk = 163840, g_ncu = 256;
// tables() on this thread while a TablesScope lives (per-handle shapes)
static thread_local const Tables *g_override = nullptr;

// Bin shapes (workgroups per CU and threads per workgroup) -> LDS per
// workgroup, table sizes and dense spans.
static Tables derive_tables(const int32_t *conc_in, const int32_t *block_in) {
  static const uint32_t conc0[7] = {32, 32, 16, 8, 4, 2, 1};
  static const uint32_t block0[7] = {64, 64, 128, 256, 512, 512, 1024};
  Tables t;
  std::memset(&t, 0, sizeof(t));
  t.lds_per_cu = g_lds_cu;
  t.lds_max_block = g_lds_blk;
  t.num_cu = g_ncu;
  for (int b = 0; b < 7; ++b) {
    const uint32_t conc = conc_in && conc_in[b] > 0 ? (uint32_t)conc_in[b] : conc0[b];
    const uint32_t block = block_in && block_in[b] > 0 ? (uint32_t)block_in[b] : block0[b];
    uint32_t l = std::min(g_lds_cu / conc, g_lds_blk) & ~255u;
    uint32_t nw = block / 64;
    t.block[b] = block;
    t.lds[b] = l;
    t.hash_sym[b] = ((l - 256) / 4) & ~7u;
    // dense bitmap over a span (plus one word of alignment slack)
    t.dense_sym[b] = b == 0 ? 0 : (l - 512) * 8;
    // numeric hash: keys (4 B) + FP64 sums (8 B) + sort buffers (4 B per
    // slot) + 1280 B header/digit totals + 512 B of radix counters per wave
    t.hash_num[b] = b == 0 ? 0 : ((l - 1312 - nw * 512) / 16) & ~7u;
    // dense numeric: FP64 sum + 1 bit per column
    t.dense_num[b] = b == 6 ? 0xFFFFFFFFu : (((l - 1024) * 8) / 65) & ~31u;
  }
  t.max_dense_sym_bin = 6;
  // numeric window: FP64 accumulators + bitmap + 4096 LDS cursors + 1 KB
  uint32_t w = 1u << 16;
  while (w > 1024 && (size_t)w * 8 + w / 8 + 4096 * 4 + 1024 > t.lds[6]) w >>= 1;
  t.win_num_cols = w;
  return t;
}

int tables_init() {
  if (g_tab_ready) return OC_OK;
  int dev = 0;
  hipDeviceProp_t p;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipGetDeviceProperties(&p, dev) == hipSuccess) {
    if (p.maxSharedMemoryPerMultiProcessor > 0)
      g_lds_cu = (uint32_t)std::min<size_t>(p.maxSharedMemoryPerMultiProcessor, 163840);
    if (p.sharedMemPerBlock > 0)
      g_lds_blk = (uint32_t)std::min<size_t>(p.sharedMemPerBlock, 163840);
    if (p.multiProcessorCount > 0) g_ncu = (uint32_t)p.multiProcessorCount;
  } else {
    (void)hipGetLastError();
  }
  g_tab = derive_tables(nullptr, nullptr);
  g_tab_ready = true;
  return OC_OK;
}

const Tables &tables() {
  if (g_override) return *g_override;
  tables_init();
  return g_tab;
}

int tables_for(const int32_t *conc, const int32_t *block, Tables *out) {
  tables_init();
  Tables t = derive_tables(conc, block);
  for (int b = 0; b < 7; ++b) {
    // every bin's numeric hash layout and dense span must fit its LDS
    const uint32_t nw = t.block[b] / 64;
    OC_REQUIRE(t.lds[b] >= 1312 + nw * 512 + 64 * 16 && t.lds[b] >= 2048,
               "bin %d: %u bytes of LDS per workgroup cannot hold the tables of a "
               "%u-thread workgroup (lower bin_concurrency or bin_block)",
               b, t.lds[b], t.block[b]);
  }
  *out = t;
  return OC_OK;
}

TablesScope::TablesScope(const Tables *t) : prev(g_override) {
  if (t) g_override = t;
}
TablesScope::~TablesScope() { g_override = prev; }

// ---------------------------------------------------------------- analysis
// Per A row: products = sum of referenced B row lengths, longest B row, the
// average, and the min first / max last column of the referenced non-empty
// B rows; grid-wide max A row length and max products
// (kernels/Analysis.cuh:31-105).
//
// A first pass writes one 16-byte record per B row {length, first column,
// last column}, so each A entry costs one gather of a compact table (B.rows
// x 16 B, cache-resident for the block-local column ranges of batched
// inputs) instead of three dependent gathers into B's offsets and columns.
// Rows of A up to kLongA entries take G lanes each; longer rows are listed
// and taken by a whole workgroup each in a second pass, so one hub row does
// not serialise its group for thousands of dependent loads.
constexpr uint32_t kLongAIters = 32;  // iterations per lane before a row is "long"
constexpr int kLongBlock = 256;

__global__ __launch_bounds__(256) void k_brow_info(const uint32_t *__restrict__ b_rp,
                                                   const uint32_t *__restrict__ b_ci,
                                                   uint32_t b_rows, uint4 *__restrict__ info) {
  for (uint32_t k = blockIdx.x * 256 + threadIdx.x; k < b_rows; k += gridDim.x * 256) {
    const uint32_t bs = b_rp[k], be = b_rp[k + 1];
    uint4 v = make_uint4(0u, 0xFFFFFFFFu, 0u, 0u);
    if (be > bs) v = make_uint4(be - bs, b_ci[bs], b_ci[be - 1], 0u);
    info[k] = v;
  }
}

struct AnaAcc {
  uint32_t prod = 0, mb = 0, lo = 0xFFFFFFFFu, hi = 0;
  __device__ __forceinline__ void add(const uint4 v) {
    // an empty B row is {0, ~0, 0}: neutral for every field
    prod += v.x;
    mb = max(mb, v.x);
    lo = min(lo, v.y);
    hi = max(hi, v.z);
  }
};

template <int G>
__global__ __launch_bounds__(256) void k_analysis(
    const uint32_t *__restrict__ a_rp, const uint32_t *__restrict__ a_ci,
    uint32_t a_rows, const uint4 *__restrict__ info, uint32_t *__restrict__ products,
    uint32_t *__restrict__ max_b, uint32_t *__restrict__ avg_b,
    uint32_t *__restrict__ left, uint32_t *__restrict__ right,
    uint32_t *max_a_nnz, uint32_t *max_products, uint32_t *__restrict__ long_rows,
    uint32_t *long_count) {
  constexpr int RPB = 256 / G;
  const int lig = threadIdx.x % G;
  uint32_t my_max_nnz = 0, my_max_prod = 0;
  for (uint32_t r = blockIdx.x * RPB + threadIdx.x / G; r < a_rows;
       r += gridDim.x * RPB) {
    const uint32_t s = a_rp[r], e = a_rp[r + 1];
    const uint32_t n = e - s;
    if (n > kLongAIters * G) {  // uniform over the group
      if (lig == 0) {
        long_rows[atomicAdd(long_count, 1u)] = r;
        my_max_nnz = max(my_max_nnz, n);
      }
      continue;
    }
    AnaAcc acc;
    uint32_t j = s + lig;
    for (; j + G < e; j += 2 * G) {  // two independent gathers in flight
      const uint32_t k0 = a_ci[j], k1 = a_ci[j + G];
      const uint4 v0 = info[k0], v1 = info[k1];
      acc.add(v0);
      acc.add(v1);
    }
    if (j < e) acc.add(info[a_ci[j]]);
    const uint32_t prod = group_sum(acc.prod, G);
    const uint32_t mb = group_max(acc.mb, G);
    const uint32_t lo = group_min(acc.lo, G);
    const uint32_t hi = group_max(acc.hi, G);
    if (lig == 0) {
      products[r] = prod;
      max_b[r] = mb;
      avg_b[r] = n == 0 ? 0u : prod / n;
      left[r] = lo;
      right[r] = hi;
      my_max_nnz = max(my_max_nnz, n);
      my_max_prod = max(my_max_prod, prod);
    }
  }
  // one atomic per workgroup: same-address atomics serialise in L2, so
  // one per wave over millions of rows would dominate the kernel
  __shared__ uint32_t red[2][256 / kWave];
  my_max_nnz = group_max(my_max_nnz, kWave);
  my_max_prod = group_max(my_max_prod, kWave);
  if (lane_id() == 0) {
    red[0][threadIdx.x / kWave] = my_max_nnz;
    red[1][threadIdx.x / kWave] = my_max_prod;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 256 / kWave; ++w) {
      my_max_nnz = max(my_max_nnz, red[0][w]);
      my_max_prod = max(my_max_prod, red[1][w]);
    }
    if (max_a_nnz) atomicMax(max_a_nnz, my_max_nnz);
    if (max_products) atomicMax(max_products, my_max_prod);
  }
}

// The listed long rows, one workgroup per row (grid-stride over the list).
__global__ __launch_bounds__(kLongBlock) void k_analysis_long(
    const uint32_t *__restrict__ a_rp, const uint32_t *__restrict__ a_ci,
    const uint4 *__restrict__ info, uint32_t *__restrict__ products,
    uint32_t *__restrict__ max_b, uint32_t *__restrict__ avg_b,
    uint32_t *__restrict__ left, uint32_t *__restrict__ right, uint32_t *max_products,
    const uint32_t *__restrict__ long_rows, const uint32_t *long_count) {
  constexpr int NW = kLongBlock / kWave;
  __shared__ uint4 part[NW];
  const uint32_t nlong = *long_count;
  const int wid = threadIdx.x / kWave;
  for (uint32_t i = blockIdx.x; i < nlong; i += gridDim.x) {
    const uint32_t r = long_rows[i];
    const uint32_t s = a_rp[r], e = a_rp[r + 1];
    AnaAcc acc;
    uint32_t j = s + threadIdx.x;
    for (; j + 3 * kLongBlock < e; j += 4 * kLongBlock) {
      const uint32_t k0 = a_ci[j], k1 = a_ci[j + kLongBlock], k2 = a_ci[j + 2 * kLongBlock],
                     k3 = a_ci[j + 3 * kLongBlock];
      const uint4 v0 = info[k0], v1 = info[k1], v2 = info[k2], v3 = info[k3];
      acc.add(v0);
      acc.add(v1);
      acc.add(v2);
      acc.add(v3);
    }
    for (; j < e; j += kLongBlock) acc.add(info[a_ci[j]]);
    const uint32_t prod = group_sum(acc.prod, kWave);
    const uint32_t mb = group_max(acc.mb, kWave);
    const uint32_t lo = group_min(acc.lo, kWave);
    const uint32_t hi = group_max(acc.hi, kWave);
    if (lane_id() == 0) part[wid] = make_uint4(prod, mb, lo, hi);
    __syncthreads();
    if (threadIdx.x == 0) {
      uint4 t = part[0];
      for (int w = 1; w < NW; ++w) {
        t.x += part[w].x;
        t.y = max(t.y, part[w].y);
        t.z = min(t.z, part[w].z);
        t.w = max(t.w, part[w].w);
      }
      products[r] = t.x;
      max_b[r] = t.y;
      avg_b[r] = t.x / (e - s);
      left[r] = t.z;
      right[r] = t.w;
      if (max_products) atomicMax(max_products, t.x);
    }
    __syncthreads();  // part[] is reused by the next row
  }
}

__global__ void k_zero3(uint32_t *a, uint32_t *b, uint32_t *c) {
  if (threadIdx.x == 0) {
    if (a) *a = 0;
    if (b) *b = 0;
    if (c) *c = 0;
  }
}

// Checks of a whole-path replay on the device: err |= bit when *x > limit.
__global__ void k_check_gt(const void *x, int is64, uint64_t limit, uint32_t bit,
                           uint32_t *err) {
  if (threadIdx.x == 0) {
    const uint64_t v = is64 ? *(const uint64_t *)x : (uint64_t) * (const uint32_t *)x;
    if (v > limit) atomicOr(err, bit);
  }
}
__global__ void k_fill_u32(uint32_t *p, uint32_t n, uint32_t v) {
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = v;
}

int launch_check_gt(const void *x, int is64, uint64_t limit, uint32_t bit, uint32_t *err,
                    hipStream_t s) {
  hipLaunchKernelGGL(k_check_gt, dim3(1), dim3(64), 0, s, x, is64, limit, bit, err);
  LAUNCH_CHECK();
  return OC_OK;
}
__global__ void k_count_bounds(const uint32_t *count, uint32_t *bounds) {
  if (threadIdx.x == 0) {
    bounds[0] = 0;
    bounds[1] = *count;
  }
}
int launch_count_bounds(const uint32_t *count, uint32_t *bounds, hipStream_t s) {
  hipLaunchKernelGGL(k_count_bounds, dim3(1), dim3(64), 0, s, count, bounds);
  LAUNCH_CHECK();
  return OC_OK;
}
int launch_fill_u32(uint32_t *p, uint32_t n, uint32_t v, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, s, p, n, v);
  LAUNCH_CHECK();
  return OC_OK;
}

size_t analysis_temp_bytes(uint64_t a_rows, uint64_t a_nnz, uint64_t b_rows) {
  const uint64_t nlong = std::min<uint64_t>(a_rows, a_nnz / (kLongAIters * 4) + 1);
  return (b_rows + 1) * 16 + 256 + nlong * 4;
}

int launch_analysis(const DevCsr &A, const DevCsr &B, const RowInfo &ri,
                    uint32_t *max_a_nnz, uint32_t *max_products, void *temp,
                    hipStream_t s) {
  if (A.rows == 0) {
    if (max_a_nnz || max_products) {
      hipLaunchKernelGGL(k_zero3, dim3(1), dim3(64), 0, s, nullptr, max_a_nnz, max_products);
      LAUNCH_CHECK();
    }
    return OC_OK;
  }
  OC_REQUIRE(A.rows < (1ull << 32) && B.rows < (1ull << 32),
             "analysis: more than 2^32 - 1 rows");
  void *own = nullptr;
  if (!temp) {
    HIP_TRY(hipMallocAsync(&own, analysis_temp_bytes(A.rows, A.nnz, B.rows), s));
    temp = own;
  }
  uint4 *info = (uint4 *)temp;
  uint32_t *long_count = (uint32_t *)(info + B.rows + 1);
  uint32_t *long_rows = long_count + 64;
  // counters cleared by a kernel, not memset nodes (capture-safe)
  hipLaunchKernelGGL(k_zero3, dim3(1), dim3(64), 0, s, long_count, max_a_nnz, max_products);
  LAUNCH_CHECK();
  if (B.rows)
    hipLaunchKernelGGL(k_brow_info, dim3(std::min<uint64_t>(cdiv(B.rows, 256), 1u << 20)),
                       dim3(256), 0, s, B.rp, B.ci, (uint32_t)B.rows, info);
  const double avg = (double)A.nnz / (double)A.rows;
  const uint32_t rows = (uint32_t)A.rows;
  // grid-stride over at most 64 workgroups per CU
  const uint32_t max_grid = 64 * std::max<uint32_t>(tables().num_cu, 1);
#define GO(G)                                                                 \
  hipLaunchKernelGGL(k_analysis<G>,                                           \
                     dim3(std::min<uint32_t>(cdiv(rows, 256 / G), max_grid)), \
                     dim3(256), 0, s, A.rp, A.ci, rows, info,                 \
                     ri.products, ri.max_b, ri.avg_b, ri.left, ri.right,      \
                     max_a_nnz, max_products, long_rows, long_count)
  if (avg <= 4) GO(4);
  else if (avg <= 8) GO(8);
  else if (avg <= 16) GO(16);
  else if (avg <= 32) GO(32);
  else GO(64);
#undef GO
  // a row is long only past kLongAIters * 4 entries, so the list holds at
  // most nnz / (kLongAIters * 4) rows, the size analysis_temp_bytes gives it
  hipLaunchKernelGGL(k_analysis_long, dim3(1024), dim3(kLongBlock), 0, s, A.rp, A.ci, info,
                     ri.products, ri.max_b, ri.avg_b, ri.left, ri.right, max_products,
                     long_rows, long_count);
  LAUNCH_CHECK();
  if (own) HIP_TRY(hipFreeAsync(own, s));
  return OC_OK;
}

// -------------------------------------------------------------------- scan
// Reduce-then-scan exclusive prefix sum of u32 into u64 (out[n] = total).
// 1024 threads x 8 items per block; deterministic.
constexpr int kScanItems = 8;
constexpr int kScanBlock = 1024;
constexpr uint32_t kScanTile = kScanItems * kScanBlock;

__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(
    const uint32_t *__restrict__ in, uint64_t n, uint64_t *__restrict__ part) {
  __shared__ uint64_t sc[17];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  uint64_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i)
    if (base + i < n) s += in[base + i];
  uint64_t tot;
  block_excl_sum(s, sc, &tot);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_partials(
    uint64_t *part, uint32_t nparts, uint64_t *total_out) {
  __shared__ uint64_t sc[17];
  uint64_t carry = 0;
  for (uint32_t b0 = 0; b0 < nparts; b0 += kScanBlock) {
    uint32_t i = b0 + threadIdx.x;
    uint64_t v = i < nparts ? part[i] : 0;
    uint64_t tot;
    uint64_t ex = block_excl_sum(v, sc, &tot);
    if (i < nparts) part[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) *total_out = carry;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_final(
    const uint32_t *__restrict__ in, uint64_t *__restrict__ out, uint64_t n,
    const uint64_t *__restrict__ part, int write_total) {
  __shared__ uint64_t sc[17];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint64_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    v[i] = base + i < n ? in[base + i] : 0u;
    s += v[i];
  }
  uint64_t tot;
  uint64_t run = block_excl_sum(s, sc, &tot) + (part ? part[blockIdx.x] : 0ull);
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    if (base + i < n) out[base + i] = run;
    run += v[i];
  }
  if (write_total && threadIdx.x == 0) out[n] = tot;
}

size_t scan_temp_bytes(uint64_t n) { return (size_t)(cdiv(n, kScanTile) + 1) * 8; }

int launch_scan_u32(const uint32_t *in, uint64_t *out, uint64_t n, void *temp,
                    hipStream_t s) {
  if (n == 0) {  // out[0] = 0 by a kernel, not a memset node (capture-safe)
    hipLaunchKernelGGL(k_scan_final, dim3(1), dim3(kScanBlock), 0, s, in, out,
                       (uint64_t)0, (const uint64_t *)nullptr, 1);
    LAUNCH_CHECK();
    return OC_OK;
  }
  uint32_t nb = cdiv(n, kScanTile);
  if (nb == 1) {
    hipLaunchKernelGGL(k_scan_final, dim3(1), dim3(kScanBlock), 0, s, in, out,
                       n, (const uint64_t *)nullptr, 1);
    LAUNCH_CHECK();
    return OC_OK;
  }
  uint64_t *part = (uint64_t *)temp;
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kScanBlock), 0, s, in, n, part);
  LAUNCH_CHECK();
  hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(kScanBlock), 0, s, part, nb,
                     out + n);
  LAUNCH_CHECK();
  hipLaunchKernelGGL(k_scan_final, dim3(nb), dim3(kScanBlock), 0, s, in, out, n,
                     (const uint64_t *)part, 0);
  LAUNCH_CHECK();
  return OC_OK;
}

// --------------------------------------------------------------- murmur3
__global__ void k_murmur(const uint32_t *__restrict__ c, uint32_t *__restrict__ h,
                         uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    h[i] = murmur3_col(c[i]);
}

int launch_murmur(const uint32_t *cols, uint32_t *out, uint64_t n, hipStream_t s) {
  if (n == 0) return OC_OK;
  hipLaunchKernelGGL(k_murmur, dim3(std::min<uint32_t>(cdiv(n, 256), 65536)),
                     dim3(256), 0, s, cols, out, n);
  LAUNCH_CHECK();
  return OC_OK;
}

// ------------------------------------------------------------------- HLL
// Sketch of each B row: 16 lanes per row, registers in LDS, merged with
// LDS atomicMax (kernels/HLL.cuh:12-24, 29-103).
template <int LOGP>
__global__ __launch_bounds__(256) void k_hll_build(const uint32_t *__restrict__ rp,
                                                   const uint32_t *__restrict__ ci,
                                                   uint32_t rows,
                                                   uint8_t *__restrict__ out) {
  constexpr int P = 1 << LOGP;
  constexpr uint32_t CAP = 33 - LOGP;
  __shared__ uint32_t regs[16][P];
  const int g = threadIdx.x / 16, lig = threadIdx.x % 16;
  for (uint32_t r = blockIdx.x * 16 + g; r < rows; r += gridDim.x * 16) {
    for (int i = lig; i < P; i += 16) regs[g][i] = 0;
    wave_sync();
    for (uint32_t j = rp[r] + lig; j < rp[r + 1]; j += 16) {
      const uint32_t h = murmur3_col(ci[j]);
      const uint32_t item = min((uint32_t)__clz(h) + 1u, CAP);
      atomicMax(&regs[g][h & (P - 1)], item);
    }
    wave_sync();
    constexpr int BPL = P / 16;  // bytes per lane
    uint32_t packed = 0;
#pragma unroll
    for (int i = 0; i < BPL; ++i) packed |= regs[g][lig * BPL + i] << (8 * i);
    if (BPL == 4)
      reinterpret_cast<uint32_t *>(out + (uint64_t)r * P)[lig] = packed;
    else
      reinterpret_cast<uint16_t *>(out + (uint64_t)r * P)[lig] = (uint16_t)packed;
    wave_sync();
  }
}

int launch_hll_build(int log_p, const DevCsr &B, uint8_t *hll, hipStream_t s) {
  if (B.rows == 0) return OC_OK;
  uint32_t rows = (uint32_t)B.rows;
  dim3 grid(std::min<uint32_t>(cdiv(rows, 16), 1u << 20));
  if (log_p == 5)
    hipLaunchKernelGGL(k_hll_build<5>, grid, dim3(256), 0, s, B.rp, B.ci, rows, hll);
  else if (log_p == 6)
    hipLaunchKernelGGL(k_hll_build<6>, grid, dim3(256), 0, s, B.rp, B.ci, rows, hll);
  else {
    set_error("unsupported HLL precision 2^%d", log_p);
    return OC_ERR_INVALID;
  }
  LAUNCH_CHECK();
  return OC_OK;
}

__device__ __forceinline__ uint32_t bytemax4(uint32_t a, uint32_t b) {
  uint32_t r = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    uint32_t x = (a >> (8 * i)) & 255u, y = (b >> (8 * i)) & 255u;
    r |= (x > y ? x : y) << (8 * i);
  }
  return r;
}

struct HllConst {
  int32_t lc[65];          // linear-counting table, lc[v] for v = 1..P
  uint32_t num_sizes[7];   // numeric hash table sizes (estimate rounding)
};

// Raw estimate from the merged words held by the first P/4 lanes
// (kernels/HLL.cuh:277-323); valid on lane 0.
template <int LOGP>
__device__ __forceinline__ int32_t hll_estimate_wave(uint32_t m, const HllConst &hc) {
  constexpr int P = 1 << LOGP;
  constexpr int W = P / 4;
  double z = 0.0;
  int empty = 0;
  if (lane_id() < W) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      uint32_t b = (m >> (8 * i)) & 255u;
      z += 1.0 / (double)(1ull << b);
      empty += b == 0;
    }
  }
  z = group_sum(z, kWave);
  empty = group_sum(empty, kWave);
  constexpr double alpha = LOGP == 5 ? 0.697 : 0.709;  // HLL_CONSTANT
  int32_t result = (int32_t)(alpha * P * P / z);
  if (empty != 0 && result < 2.5 * P) result = hc.lc[empty];
  return result;
}

// One wave per A row (rows with at most `heavy` entries) or one block per
// listed heavy row.  MODE 0: estimate * expansion rounded up to the numeric
// table sizes (kernels/HLL.cuh:108-236).  MODE 1: sampled rows, raw
// estimate (kernels/HLL.cuh:240-357).
template <int LOGP>
__device__ __forceinline__ uint32_t hll_merge_range(const uint32_t *a_ci,
                                                    const uint8_t *hll,
                                                    uint32_t s, uint32_t e,
                                                    int lane) {
  constexpr int P = 1 << LOGP;
  constexpr int W = P / 4;     // words per sketch
  constexpr int S = kWave / W; // sketches per step
  const int sk = lane / W, w = lane % W;
  uint32_t m = 0;
  for (uint32_t j = s + sk; j < e; j += S) {
    const uint32_t k = a_ci[j];
    m = bytemax4(m, reinterpret_cast<const uint32_t *>(hll + (uint64_t)k * P)[w]);
  }
  for (int d = W; d < kWave; d <<= 1) m = bytemax4(m, __shfl_xor(m, d));
  return m;
}

template <int LOGP, int MODE>
__global__ __launch_bounds__(256) void k_hll_merge(
    const uint32_t *__restrict__ a_rp, const uint32_t *__restrict__ a_ci,
    uint32_t rows, const uint8_t *__restrict__ hll, double expansion,
    HllConst hc, uint32_t *__restrict__ est, uint32_t nsamples,
    uint32_t *__restrict__ sample_rows, int32_t *__restrict__ sample_est,
    const uint32_t *__restrict__ products, uint32_t *__restrict__ sample_prod,
    uint32_t heavy, uint32_t *heavy_list, uint32_t *heavy_count) {
  const int wid = threadIdx.x / kWave, lane = lane_id();
  const uint32_t ntasks = MODE == 0 ? rows : nsamples;
  for (uint32_t t = blockIdx.x * 4 + wid; t < ntasks; t += gridDim.x * 4) {
    uint32_t r = t;
    if (MODE == 1) {
      uint64_t x = 112233ull + (uint64_t)t * 2654435761ull;
      x = (1664525ull * x + 1013904223ull) & 0xFFFFFFFFull;
      r = (uint32_t)(x % rows);
    }
    const uint32_t s = a_rp[r], e = a_rp[r + 1];
    if (e - s > heavy) {
      if (lane == 0) heavy_list[atomicAdd(heavy_count, 1u)] = t;
      continue;
    }
    const uint32_t m = hll_merge_range<LOGP>(a_ci, hll, s, e, lane);
    const int32_t result = hll_estimate_wave<LOGP>(m, hc);
    if (lane == 0) {
      if (MODE == 0) {
        uint32_t temp = (uint32_t)(result * expansion);
        for (int b = 0; b < 6; ++b)
          if (temp <= hc.num_sizes[b]) {
            temp = hc.num_sizes[b];
            break;
          }
        est[r] = temp;
      } else {
        sample_rows[t] = r;
        sample_est[t] = result;
        if (sample_prod) sample_prod[t] = products[r];
      }
    }
  }
}

template <int LOGP, int MODE>
__global__ __launch_bounds__(1024) void k_hll_merge_heavy(
    const uint32_t *__restrict__ a_rp, const uint32_t *__restrict__ a_ci,
    uint32_t rows, const uint8_t *__restrict__ hll, double expansion,
    HllConst hc, uint32_t *__restrict__ est, uint32_t *__restrict__ sample_rows,
    int32_t *__restrict__ sample_est, const uint32_t *__restrict__ products,
    uint32_t *__restrict__ sample_prod, const uint32_t *heavy_list,
    const uint32_t *heavy_count) {
  constexpr int P = 1 << LOGP;
  constexpr int W = P / 4;
  __shared__ uint32_t part[16][W];
  const int wid = threadIdx.x / kWave, lane = lane_id();
  const uint32_t nheavy = *heavy_count;
  for (uint32_t h = blockIdx.x; h < nheavy; h += gridDim.x) {
    const uint32_t t = heavy_list[h];
    uint32_t r = t;
    if (MODE == 1) {
      uint64_t x = 112233ull + (uint64_t)t * 2654435761ull;
      x = (1664525ull * x + 1013904223ull) & 0xFFFFFFFFull;
      r = (uint32_t)(x % rows);
    }
    const uint32_t s = a_rp[r], e = a_rp[r + 1];
    const uint32_t per = (e - s + 15) / 16;
    const uint32_t ws = s + wid * per, we = min(e, ws + per);
    uint32_t m = hll_merge_range<LOGP>(a_ci, hll, ws < e ? ws : e, ws < e ? we : e, lane);
    if (lane < W) part[wid][lane] = m;
    __syncthreads();
    if (wid == 0) {
      m = 0;
      if (lane < W)
        for (int i = 0; i < 16; ++i) m = bytemax4(m, part[i][lane]);
      const int32_t result = hll_estimate_wave<LOGP>(m, hc);
      if (lane == 0) {
        if (MODE == 0) {
          uint32_t temp = (uint32_t)(result * expansion);
          for (int b = 0; b < 6; ++b)
            if (temp <= hc.num_sizes[b]) {
              temp = hc.num_sizes[b];
              break;
            }
          est[r] = temp;
        } else {
          sample_rows[t] = r;
          sample_est[t] = result;
          if (sample_prod) sample_prod[t] = products[r];
        }
      }
    }
    __syncthreads();
  }
}

size_t hll_merge_temp_bytes(uint64_t rows) { return (rows + 16) * 4; }

static void make_lc(int log_p, int32_t *lc) {
  const int p = 1 << log_p;
  lc[0] = 0;
  for (int v = 1; v <= p; ++v) {
    float q = (float)p / (float)v;
    lc[v] = (int32_t)((float)p * logf(q));
  }
}

int launch_hll_merge(int log_p, const DevCsr &A, const uint8_t *hll,
                     double expansion, uint32_t *est, uint32_t nsamples,
                     uint32_t *sample_rows, int32_t *sample_est,
                     const uint32_t *products, uint32_t *sample_prod, int mode,
                     void *temp, hipStream_t s) {
  if (A.rows == 0) return OC_OK;
  if (log_p != 5 && log_p != 6) {
    set_error("unsupported HLL precision 2^%d", log_p);
    return OC_ERR_INVALID;
  }
  HllConst hc;
  std::memset(&hc, 0, sizeof(hc));
  make_lc(log_p, hc.lc);
  std::memcpy(hc.num_sizes, tables().hash_num, sizeof(hc.num_sizes));
  uint32_t *heavy_count = (uint32_t *)temp;
  uint32_t *heavy_list = heavy_count + 16;
  HIP_TRY(hipMemsetAsync(heavy_count, 0, 4, s));
  const uint32_t rows = (uint32_t)A.rows;
  const uint32_t heavy = 2048;
  const uint32_t ntasks = mode == 0 ? rows : nsamples;
  dim3 grid(std::max<uint32_t>(1, std::min<uint32_t>(cdiv(ntasks, 4), 1u << 18)));
  dim3 hgrid(std::max<uint32_t>(1, tables().num_cu));
#define GO(L, M)                                                               \
  do {                                                                         \
    hipLaunchKernelGGL((k_hll_merge<L, M>), grid, dim3(256), 0, s, A.rp, A.ci, \
                       rows, hll, expansion, hc, est, nsamples, sample_rows,   \
                       sample_est, products, sample_prod, heavy, heavy_list,   \
                       heavy_count);                                           \
    LAUNCH_CHECK();                                                            \
    hipLaunchKernelGGL((k_hll_merge_heavy<L, M>), hgrid, dim3(1024), 0, s,     \
                       A.rp, A.ci, rows, hll, expansion, hc, est, sample_rows, \
                       sample_est, products, sample_prod,                      \
                       (const uint32_t *)heavy_list,                           \
                       (const uint32_t *)heavy_count);                         \
    LAUNCH_CHECK();                                                            \
  } while (0)
  if (log_p == 5) {
    if (mode == 0) GO(5, 0); else GO(5, 1);
  } else {
    if (mode == 0) GO(6, 0); else GO(6, 1);
  }
#undef GO
  return OC_OK;
}

// ---------------------------------------------------------------- binning
// Deterministic stable binning: classify (bin per row + per-block
// histogram), scan the histograms, scatter rows in ascending row order.
// The rules are those of binning / numericBinning
// (kernels/Analysis.cuh:110-303) with the gfx950 tables.
struct BinTables {
  uint32_t hash_sym[7], dense_sym[7], hash_num[7], dense_num[7];
  int max_dense_sym_bin;
};

constexpr int kBinBlock = 256;

__device__ __forceinline__ int sym_bin(uint32_t products, uint32_t lo,
                                       uint32_t hi, int use_compaction,
                                       double expand, double comp,
                                       const BinTables &bt) {
  const uint32_t range = hi - lo + 1u;
  int dense_bin = 7;
  for (int b = 0; b <= bt.max_dense_sym_bin; ++b)
    if (range <= bt.dense_sym[b]) {
      dense_bin = b;
      break;
    }
  uint32_t val = products;
  if (use_compaction) val = (uint32_t)(val / comp);
  val = (uint32_t)fmin((double)val * expand, 4294967295.0);  // saturating
  int bin = 6;
  for (int b = 0; b < 6; ++b)
    if (val <= bt.hash_sym[b]) {
      bin = b;
      break;
    }
  if (dense_bin <= bin) bin = min(bin, bt.max_dense_sym_bin) + 7;
  return bin;
}

__device__ __forceinline__ int num_bin_one(uint32_t val, int dense_bin,
                                           int max_bin, const BinTables &bt) {
  int bin = 7;
  for (int b = 1; b < max_bin; ++b)
    if (val <= bt.hash_num[b]) {
      bin = b;
      break;
    }
  if (dense_bin <= bin && !(bin == 6 && dense_bin == 6)) bin = dense_bin + 7;
  return bin;
}

struct BinArgs {
  int kind;
  const uint32_t *input;
  uint32_t *input_mut;
  const uint32_t *left, *right;
  int use_compaction, expand, modify_dense_est, use_largest;
  double expand_factor, avg_compaction;
  uint32_t hybrid_max;
  const uint32_t *filter;
  uint32_t skip_below;
  const uint32_t *skip_flag;    // rows handled elsewhere (long-row bands)
  const uint64_t *hybrid_rank;  // exclusive rank among hybrid candidates
};

__device__ __forceinline__ bool skipped(const BinArgs &ba, uint32_t r) {
  return (ba.filter && ba.filter[r] <= ba.skip_below) || (ba.skip_flag && ba.skip_flag[r]);
}

// numeric: compute dense bin and the (possibly modified) value
__device__ __forceinline__ void num_prepare(const BinArgs &ba, const BinTables &bt,
                                            uint32_t r, int &dense_bin,
                                            uint32_t &val) {
  const uint32_t range = ba.right[r] - ba.left[r] + 1u;
  dense_bin = 6;
  uint32_t in = ba.input[r];
  for (int b = 0; b < 6; ++b)
    if (range <= bt.dense_num[b]) {
      dense_bin = b;
      if (b == 0 && ba.modify_dense_est) in = bt.dense_num[0];
      break;
    }
  val = in;
  if (ba.expand) val = (uint32_t)fmin((double)val * ba.expand_factor, 4294967295.0);
}

// Hybrid candidates (numeric bin 6 when allowed) for the deterministic cap.
__global__ void k_bin_hybrid_flags(BinArgs ba, BinTables bt, uint32_t rows,
                                   uint32_t *flags) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < rows;
       r += gridDim.x * blockDim.x) {
    int db;
    uint32_t val;
    num_prepare(ba, bt, r, db, val);
    flags[r] = (!skipped(ba, r) && num_bin_one(val, db, 7, bt) == 6) ? 1u : 0u;
  }
}

__device__ __forceinline__ int classify(const BinArgs &ba, const BinTables &bt,
                                        uint32_t r) {
  if (skipped(ba, r)) return kSkipBin;
  if (ba.kind == 0)
    return sym_bin(ba.input[r], ba.left[r], ba.right[r], ba.use_compaction,
                   ba.expand_factor, ba.avg_compaction, bt);
  int db;
  uint32_t val;
  num_prepare(ba, bt, r, db, val);
  int bin = num_bin_one(val, db, 6, bt);
  if (ba.use_largest && ba.hybrid_rank) {
    if (num_bin_one(val, db, 7, bt) == 6 && ba.hybrid_rank[r] < ba.hybrid_max) bin = 6;
  }
  return bin;
}

__global__ __launch_bounds__(kBinBlock) void k_bin_classify(
    BinArgs ba, BinTables bt, uint32_t rows, uint8_t *__restrict__ bin_of_row,
    uint32_t *__restrict__ hist /* [nblocks][kBinSlots] */) {
  __shared__ uint32_t h[kBinSlots];
  if (threadIdx.x < kBinSlots) h[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t r = blockIdx.x * kBinBlock + threadIdx.x;
  if (r < rows) {
    const int bin = classify(ba, bt, r);
    if (ba.kind == 1 && ba.modify_dense_est && ba.input_mut) {
      const uint32_t range = ba.right[r] - ba.left[r] + 1u;
      if (range <= bt.dense_num[0] && !skipped(ba, r))
        ba.input_mut[r] = bt.dense_num[0];
    }
    bin_of_row[r] = (uint8_t)bin;
    atomicAdd(&h[bin], 1u);
  }
  __syncthreads();
  if (threadIdx.x < kBinSlots) hist[(uint64_t)blockIdx.x * kBinSlots + threadIdx.x] = h[threadIdx.x];
}

// One block: per bin, exclusive scan over blocks plus bin bases.
__global__ __launch_bounds__(1024) void k_bin_scan(uint32_t *hist, uint32_t nblocks,
                                                   uint32_t *bin_offsets) {
  __shared__ uint32_t tot[kBinSlots];
  __shared__ uint64_t sc[17];
  for (int b = 0; b < kBinSlots; ++b) {
    uint64_t carry = 0;
    for (uint32_t b0 = 0; b0 < nblocks; b0 += 1024) {
      uint32_t i = b0 + threadIdx.x;
      uint64_t v = i < nblocks ? hist[(uint64_t)i * kBinSlots + b] : 0;
      uint64_t t;
      uint64_t ex = block_excl_sum(v, sc, &t);
      if (i < nblocks) hist[(uint64_t)i * kBinSlots + b] = (uint32_t)(carry + ex);
      carry += t;
    }
    if (threadIdx.x == 0) tot[b] = (uint32_t)carry;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int b = 0; b < kBinSlots; ++b) {
      bin_offsets[b] = run;
      run += tot[b];
    }
    bin_offsets[kBinSlots] = run;
  }
  __syncthreads();
  // add bin bases
  for (uint32_t i = threadIdx.x; i < nblocks; i += 1024)
    for (int b = 0; b < kBinSlots; ++b) {
      uint32_t base = 0;
      for (int q = 0; q < b; ++q) base += tot[q];
      hist[(uint64_t)i * kBinSlots + b] += base;
    }
}

__global__ __launch_bounds__(kBinBlock) void k_bin_scatter(
    const uint8_t *__restrict__ bin_of_row, uint32_t rows,
    const uint32_t *__restrict__ hist, uint32_t *__restrict__ bin_rows) {
  __shared__ uint32_t wc[kBinBlock / kWave][16];
  const int wid = threadIdx.x / kWave, lane = lane_id();
  const uint32_t r = blockIdx.x * kBinBlock + threadIdx.x;
  const bool valid = r < rows;
  const uint32_t bin = valid ? bin_of_row[r] : 0u;
  uint64_t m = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    uint64_t bal = __ballot((bin >> b) & 1u);
    m &= ((bin >> b) & 1u) ? bal : ~bal;
  }
  if (lane < 16) wc[wid][lane] = 0;
  __syncthreads();
  if (valid && (m & lanemask_lt()) == 0) wc[wid][bin] = (uint32_t)__popcll(m);
  __syncthreads();
  if (valid) {
    uint32_t pos = hist[(uint64_t)blockIdx.x * kBinSlots + bin];
    for (int w = 0; w < wid; ++w) pos += wc[w][bin];
    pos += (uint32_t)__popcll(m & lanemask_lt());
    bin_rows[pos] = r;
  }
}

size_t bin_temp_bytes(uint64_t rows) {
  uint64_t nb = (rows + kBinBlock - 1) / kBinBlock;
  return nb * kBinSlots * 4 + (rows + 1) * 4 + (rows + 1) * 8 + scan_temp_bytes(rows) + 64;
}

int launch_binning(const BinParams &bp, uint64_t rows, uint8_t *bin_of_row,
                   uint32_t *bin_rows, uint32_t *bin_offsets, void *temp,
                   hipStream_t s) {
  const Tables &t = tables();
  BinTables bt;
  std::memcpy(bt.hash_sym, t.hash_sym, sizeof(bt.hash_sym));
  std::memcpy(bt.dense_sym, t.dense_sym, sizeof(bt.dense_sym));
  std::memcpy(bt.hash_num, t.hash_num, sizeof(bt.hash_num));
  std::memcpy(bt.dense_num, t.dense_num, sizeof(bt.dense_num));
  bt.max_dense_sym_bin = (int)t.max_dense_sym_bin;
  BinArgs ba;
  std::memset(&ba, 0, sizeof(ba));
  ba.kind = bp.kind;
  ba.input = bp.input;
  ba.input_mut = bp.input_mut;
  ba.left = bp.left;
  ba.right = bp.right;
  ba.use_compaction = bp.use_compaction;
  ba.expand = bp.expand;
  ba.modify_dense_est = bp.modify_dense_est;
  ba.use_largest = bp.use_largest;
  ba.expand_factor = bp.expand_factor;
  ba.avg_compaction = bp.avg_compaction;
  ba.hybrid_max = bp.hybrid_max;
  ba.filter = bp.filter;
  ba.skip_below = bp.skip_below;
  ba.skip_flag = bp.skip_flag;
  if (rows == 0) {
    HIP_TRY(hipMemsetAsync(bin_offsets, 0, 4 * (kBinSlots + 1), s));
    return OC_OK;
  }
  const uint32_t nb = cdiv(rows, kBinBlock);
  uint32_t *hist = (uint32_t *)temp;
  uint32_t *flags = hist + (uint64_t)nb * kBinSlots;
  uint64_t *rank = (uint64_t *)(((uintptr_t)(flags + rows + 1) + 7) & ~(uintptr_t)7);
  void *scan_tmp = rank + rows + 1;
  if (bp.kind == 1 && bp.use_largest) {
    hipLaunchKernelGGL(k_bin_hybrid_flags, dim3(std::min<uint32_t>(nb, 65536)),
                       dim3(kBinBlock), 0, s, ba, bt, (uint32_t)rows, flags);
    LAUNCH_CHECK();
    OC_TRY(launch_scan_u32(flags, rank, rows, scan_tmp, s));
    ba.hybrid_rank = rank;
  }
  hipLaunchKernelGGL(k_bin_classify, dim3(nb), dim3(kBinBlock), 0, s, ba, bt,
                     (uint32_t)rows, bin_of_row, hist);
  LAUNCH_CHECK();
  hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(1024), 0, s, hist, nb, bin_offsets);
  LAUNCH_CHECK();
  hipLaunchKernelGGL(k_bin_scatter, dim3(nb), dim3(kBinBlock), 0, s,
                     (const uint8_t *)bin_of_row, (uint32_t)rows,
                     (const uint32_t *)hist, bin_rows);
  LAUNCH_CHECK();
  return OC_OK;
}

// --------------------------------------------------------------- row copy
// Temp -> C copy of whole rows (kernels/Epilogue.cuh:8-65): one wave per
// row, coalesced.
template <typename T>
__global__ __launch_bounds__(256) void k_copy_rows(
    uint64_t nrows, const uint32_t *__restrict__ rows,
    const uint32_t *__restrict__ nnz, const uint64_t *__restrict__ src_off,
    const uint32_t *__restrict__ cap, const uint32_t *__restrict__ sc,
    const T *__restrict__ sv, const uint64_t *__restrict__ dst_off,
    uint32_t *__restrict__ dc, T *__restrict__ dv) {
  const int lane = lane_id();
  for (uint64_t i = blockIdx.x * 4ull + threadIdx.x / kWave; i < nrows;
       i += gridDim.x * 4ull) {
    const uint32_t r = rows ? rows[i] : (uint32_t)i;
    const uint32_t n = nnz[r];
    if (cap && n > cap[r]) continue;  // outgrew its slot: computed in place
    const uint64_t so = src_off[r], d = dst_off[r];
    // eight loads in flight per lane before the stores: a long row is
    // copied by one wave, so its copy time is latency x n / (64 x 8)
    uint32_t j = lane;
    for (; j + 7 * kWave < n; j += 8 * kWave) {
      uint32_t c[8];
      T v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        c[u] = sc[so + j + u * kWave];
        v[u] = sv[so + j + u * kWave];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        dc[d + j + u * kWave] = c[u];
        dv[d + j + u * kWave] = v[u];
      }
    }
    for (; j < n; j += kWave) {
      dc[d + j] = sc[so + j];
      dv[d + j] = sv[so + j];
    }
  }
}

int launch_copy_rows(int vtype, uint64_t nrows, const uint32_t *rows,
                     const uint32_t *nnz, const uint64_t *src_off, const uint32_t *cap,
                     const uint32_t *src_cols, const void *src_vals,
                     const uint64_t *dst_off, uint32_t *dst_cols,
                     void *dst_vals, hipStream_t s) {
  if (nrows == 0) return OC_OK;
  dim3 grid(std::min<uint32_t>(cdiv(nrows, 4), 1u << 20));
  if (vtype == OC_F64)
    hipLaunchKernelGGL(k_copy_rows<double>, grid, dim3(256), 0, s, nrows, rows,
                       nnz, src_off, cap, src_cols, (const double *)src_vals, dst_off,
                       dst_cols, (double *)dst_vals);
  else
    hipLaunchKernelGGL(k_copy_rows<float>, grid, dim3(256), 0, s, nrows, rows,
                       nnz, src_off, cap, src_cols, (const float *)src_vals, dst_off,
                       dst_cols, (float *)dst_vals);
  LAUNCH_CHECK();
  return OC_OK;
}

// ------------------------------------------------- overflow spare space
// Each overflow row reserves min(products, cols) entries of the spare
// region behind the estimated temp space (kernels/AccumulatorHash.cuh:
// 612-617); its output offset becomes base + reservation.
__global__ void k_assign_spare(const uint32_t *__restrict__ ovf_rows, uint32_t n,
                               const uint32_t *__restrict__ products,
                               const uint32_t *__restrict__ left,
                               const uint32_t *__restrict__ right, uint64_t base,
                               unsigned long long *spare_used,
                               uint64_t *__restrict__ off) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    const uint32_t r = ovf_rows[i];
    // a row has at most min(products, column span) outputs
    const uint64_t span = right[r] >= left[r] ? (uint64_t)right[r] - left[r] + 1 : 0;
    const unsigned long long need =
        (unsigned long long)min((uint64_t)products[r], span);
    off[r] = base + atomicAdd(spare_used, need);
  }
}

int launch_assign_spare(const uint32_t *ovf_rows, uint32_t n,
                        const uint32_t *products, const uint32_t *left,
                        const uint32_t *right, uint64_t base,
                        unsigned long long *spare_used, uint64_t *off,
                        hipStream_t s) {
  if (n == 0) return OC_OK;
  hipLaunchKernelGGL(k_assign_spare, dim3(std::min<uint32_t>(cdiv(n, 256), 4096)),
                     dim3(256), 0, s, ovf_rows, n, products, left, right, base,
                     spare_used, off);
  LAUNCH_CHECK();
  return OC_OK;
}

// --------------------------------------------------- work accounting
// Sums over a row list (or all rows): A entries, products and output
// entries (from u64 offsets or a count array); out[0..3] += {rows, a_nnz,
// products, out_nnz}.
__global__ __launch_bounds__(256) void k_row_sums(const uint32_t *__restrict__ rows, uint32_t n,
                           const uint32_t *__restrict__ a_rp,
                           const uint32_t *__restrict__ products,
                           const uint64_t *__restrict__ c_rp,
                           const uint32_t *__restrict__ cnt,
                           unsigned long long *out) {
  unsigned long long an = 0, pr = 0, nz = 0, k = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    const uint32_t r = rows ? rows[i] : i;
    an += a_rp[r + 1] - a_rp[r];
    pr += products[r];
    nz += c_rp ? c_rp[r + 1] - c_rp[r] : (cnt ? cnt[r] : 0u);
    ++k;
  }
  an = group_sum(an, kWave);
  pr = group_sum(pr, kWave);
  nz = group_sum(nz, kWave);
  k = group_sum(k, kWave);
  // the block's four sums, then one atomic each per block (every wave
  // adding into the same four words contended: 0.2 ms per launch)
  __shared__ unsigned long long part[4][4];
  const uint32_t w = threadIdx.x / kWave;
  if (lane_id() == 0) {
    part[w][0] = k;
    part[w][1] = an;
    part[w][2] = pr;
    part[w][3] = nz;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    unsigned long long s = 0;
    for (uint32_t i = 0; i < blockDim.x / kWave; ++i) s += part[i][threadIdx.x];
    atomicAdd(&out[threadIdx.x], s);
  }
}

int launch_row_sums(const uint32_t *rows, uint32_t n, const uint32_t *a_rp,
                    const uint32_t *products, const uint64_t *c_rp,
                    const uint32_t *cnt, unsigned long long *out, hipStream_t s) {
  if (n == 0) return OC_OK;
  hipLaunchKernelGGL(k_row_sums, dim3(std::min<uint32_t>(cdiv(n, 256), 256)), dim3(256),
                     0, s, rows, n, a_rp, products, c_rp, cnt, out);
  LAUNCH_CHECK();
  return OC_OK;
}

// ------------------------------------------------------------ validation
template <typename O>
__device__ __forceinline__ uint64_t off_at(const void *p, uint64_t i) {
  return (uint64_t)reinterpret_cast<const O *>(p)[i];
}

template <typename O>
__global__ __launch_bounds__(256) void k_validate(const void *rp,
                                                  const uint32_t *__restrict__ ci,
                                                  uint64_t rows, uint64_t cols,
                                                  unsigned long long *bad) {
  const int lane = lane_id();
  unsigned long long nbad = 0;
  for (uint64_t r = blockIdx.x * 4ull + threadIdx.x / kWave; r < rows;
       r += gridDim.x * 4ull) {
    const uint64_t s = off_at<O>(rp, r), e = off_at<O>(rp, r + 1);
    bool ok = e >= s;
    for (uint64_t j = s + lane; ok && j < e; j += kWave) {
      uint32_t c = ci[j];
      if (c >= cols) ok = false;
      if (j + 1 < e && ci[j + 1] <= c) ok = false;
    }
    bool any_bad = __ballot(!ok) != 0;
    if (lane == 0 && any_bad) ++nbad;
  }
  if (lane == 0 && nbad) atomicAdd(bad, nbad);
}

template <typename O>
__global__ __launch_bounds__(256) void k_set_hash(const void *rp,
                                                  const uint32_t *__restrict__ ci,
                                                  uint64_t rows, uint64_t *h) {
  const int lane = lane_id();
  for (uint64_t r = blockIdx.x * 4ull + threadIdx.x / kWave; r < rows;
       r += gridDim.x * 4ull) {
    const uint64_t s = off_at<O>(rp, r), e = off_at<O>(rp, r + 1);
    uint64_t acc = 0;
    for (uint64_t j = s + lane; j < e; j += kWave) acc += mix64(ci[j]);
    acc = group_sum(acc, kWave);
    if (lane == 0) h[r] = acc;
  }
}

template <typename O, typename T>
__global__ __launch_bounds__(256) void k_spmv(const void *rp,
                                              const uint32_t *__restrict__ ci,
                                              const T *__restrict__ v,
                                              uint64_t rows,
                                              const double *__restrict__ x,
                                              double *__restrict__ y) {
  const int lane = lane_id();
  for (uint64_t r = blockIdx.x * 4ull + threadIdx.x / kWave; r < rows;
       r += gridDim.x * 4ull) {
    const uint64_t s = off_at<O>(rp, r), e = off_at<O>(rp, r + 1);
    double acc = 0.0;
    for (uint64_t j = s + lane; j < e; j += kWave) acc += (double)v[j] * x[ci[j]];
    acc = group_sum(acc, kWave);
    if (lane == 0) y[r] = acc;
  }
}

static dim3 row_grid(uint64_t rows) {
  return dim3(std::max<uint32_t>(1, std::min<uint32_t>(cdiv(rows, 4), 1u << 18)));
}

int launch_validate(const oc_csr &M, unsigned long long *bad, hipStream_t s) {
  if (M.rows == 0) return OC_OK;
  if (M.off_bytes == 8)
    hipLaunchKernelGGL(k_validate<uint64_t>, row_grid(M.rows), dim3(256), 0, s,
                       (const void *)M.row_offsets, M.col_ids, M.rows, M.cols, bad);
  else
    hipLaunchKernelGGL(k_validate<uint32_t>, row_grid(M.rows), dim3(256), 0, s,
                       (const void *)M.row_offsets, M.col_ids, M.rows, M.cols, bad);
  LAUNCH_CHECK();
  return OC_OK;
}

int launch_row_set_hash(const oc_csr &M, uint64_t *hash, hipStream_t s) {
  if (M.rows == 0) return OC_OK;
  if (M.off_bytes == 8)
    hipLaunchKernelGGL(k_set_hash<uint64_t>, row_grid(M.rows), dim3(256), 0, s,
                       (const void *)M.row_offsets, M.col_ids, M.rows, hash);
  else
    hipLaunchKernelGGL(k_set_hash<uint32_t>, row_grid(M.rows), dim3(256), 0, s,
                       (const void *)M.row_offsets, M.col_ids, M.rows, hash);
  LAUNCH_CHECK();
  return OC_OK;
}

int launch_spmv(const oc_csr &M, const double *x, double *y, hipStream_t s) {
  if (M.rows == 0) return OC_OK;
  const void *rp = M.row_offsets;
  if (M.off_bytes == 8) {
    if (M.vtype == OC_F64)
      hipLaunchKernelGGL((k_spmv<uint64_t, double>), row_grid(M.rows), dim3(256), 0,
                         s, rp, M.col_ids, (const double *)M.values, M.rows, x, y);
    else
      hipLaunchKernelGGL((k_spmv<uint64_t, float>), row_grid(M.rows), dim3(256), 0,
                         s, rp, M.col_ids, (const float *)M.values, M.rows, x, y);
  } else {
    if (M.vtype == OC_F64)
      hipLaunchKernelGGL((k_spmv<uint32_t, double>), row_grid(M.rows), dim3(256), 0,
                         s, rp, M.col_ids, (const double *)M.values, M.rows, x, y);
    else
      hipLaunchKernelGGL((k_spmv<uint32_t, float>), row_grid(M.rows), dim3(256), 0,
                         s, rp, M.col_ids, (const float *)M.values, M.rows, x, y);
  }
  LAUNCH_CHECK();
  return OC_OK;
}

}  // namespace oc
