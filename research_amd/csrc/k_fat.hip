// FAT products: A short and wide, B tall and narrow, so C (A.rows x B.cols)
// is small and dense-ish while each A row holds very many entries (the role
// of denseNumericStaticKernel + numeric_kernel_fat, AccumulatorDense.cuh:
// 809-893 and SpGEMM.cuh:925-1050).
//
// The reference gives each A row one block; with a handful of rows that
// leaves most of the GPU idle.  Here each row's A entries are split over
// `splits` workgroups.  A workgroup sums its share into a dense FP64 LDS row
// over all B.cols columns plus a bitmap, then flushes the set columns into
// a global dense row (FP64 atomics) and ORs its bitmap into the row's global
// bitmap.  A count pass and a scan give C's offsets; the compaction writes
// each row in column order from the dense row.  No symbolic pass, no sort.
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

constexpr int kFatBlock = 512;

template <typename T>
struct FatArgs {
  const uint32_t *a_rp, *a_ci;
  const T *a_v;
  const uint32_t *b_rp, *b_ci;
  const T *b_v;
  uint32_t ncols, nwords, splits;
  int G;         // lanes per A entry
  double *dense;  // [rows][ncols], zeroed
  uint32_t *bits; // [rows][nwords], zeroed
};

template <typename T>
__global__ __launch_bounds__(kFatBlock) void k_fat_accumulate(FatArgs<T> f) {
  extern __shared__ __align__(16) uint8_t smem[];
  double *acc = reinterpret_cast<double *>(smem);
  uint32_t *lbits = reinterpret_cast<uint32_t *>(smem + (size_t)f.ncols * 8);
  const uint32_t r = blockIdx.x / f.splits, s = blockIdx.x % f.splits;
  for (uint32_t i = threadIdx.x; i < f.ncols; i += kFatBlock) acc[i] = 0.0;
  for (uint32_t i = threadIdx.x; i < f.nwords; i += kFatBlock) lbits[i] = 0;
  __syncthreads();
  const uint32_t a0 = f.a_rp[r], a1 = f.a_rp[r + 1];
  const uint64_t len = a1 - a0;
  const uint32_t j0 = a0 + (uint32_t)(len * s / f.splits);
  const uint32_t j1 = a0 + (uint32_t)(len * (s + 1) / f.splits);
  const int G = f.G, ngroups = kFatBlock / G;
  const int gid = threadIdx.x / G, lig = threadIdx.x % G;
  for (uint32_t j = j0 + gid; j < j1; j += ngroups) {
    const uint32_t k = f.a_ci[j];
    const double av = (double)f.a_v[j];
    const uint32_t e = f.b_rp[k + 1];
    for (uint32_t q = f.b_rp[k] + lig; q < e; q += G) {
      const uint32_t c = f.b_ci[q];
      if (c >= f.ncols) continue;  // not a column of B
      __hip_atomic_fetch_add(&acc[c], av * (double)f.b_v[q], __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WORKGROUP);
      atomicOr(&lbits[c >> 5], 1u << (c & 31));
    }
  }
  __syncthreads();
  double *drow = f.dense + (uint64_t)r * f.ncols;
  uint32_t *brow = f.bits + (uint64_t)r * f.nwords;
  for (uint32_t w = threadIdx.x; w < f.nwords; w += kFatBlock) {
    uint32_t m = lbits[w];
    if (!m) continue;
    atomicOr(&brow[w], m);
    while (m) {
      const uint32_t c = w * 32 + (uint32_t)(__ffs(m) - 1);
      m &= m - 1;
      atomicAdd(&drow[c], acc[c]);
    }
  }
}

// nnz of each row: popcount of its bitmap (one wave per row)
__global__ void k_fat_count(const uint32_t *__restrict__ bits, uint32_t rows, uint32_t nwords,
                            uint32_t *__restrict__ nnz) {
  const uint32_t r = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (r >= rows) return;
  uint32_t cnt = 0;
  for (uint32_t w = lane_id(); w < nwords; w += kWave) cnt += __popc(bits[(uint64_t)r * nwords + w]);
  cnt = group_sum(cnt, kWave);
  if (lane_id() == 0) nnz[r] = cnt;
}

// row r written in column order: each thread takes a contiguous run of
// bitmap words, a workgroup prefix of their popcounts gives its positions
template <typename T>
__global__ __launch_bounds__(kFatBlock) void k_fat_compact(
    const double *__restrict__ dense, const uint32_t *__restrict__ bits, uint32_t ncols,
    uint32_t nwords, const uint64_t *__restrict__ c_off, uint32_t *__restrict__ c_cols,
    T *__restrict__ c_vals) {
  __shared__ uint32_t scratch[17];
  const uint32_t r = blockIdx.x;
  const uint32_t *brow = bits + (uint64_t)r * nwords;
  const double *drow = dense + (uint64_t)r * ncols;
  const uint32_t wpt = (nwords + kFatBlock - 1) / kFatBlock;
  const uint32_t w0 = min(nwords, threadIdx.x * wpt), w1 = min(nwords, w0 + wpt);
  uint32_t mine = 0;
  for (uint32_t w = w0; w < w1; ++w) mine += __popc(brow[w]);
  uint32_t total;
  uint64_t pos = c_off[r] + block_excl_sum(mine, scratch, &total);
  for (uint32_t w = w0; w < w1; ++w) {
    uint32_t m = brow[w];
    while (m) {
      const uint32_t c = w * 32 + (uint32_t)(__ffs(m) - 1);
      m &= m - 1;
      c_cols[pos] = c;
      c_vals[pos] = (T)drow[c];
      ++pos;
    }
  }
}

size_t fat_lds_bytes(uint64_t b_cols) {
  return (size_t)b_cols * 8 + ((b_cols + 31) / 32) * 4;
}

int launch_fat_accumulate(const DevCsr &A, const DevCsr &B, uint32_t splits, int G,
                          double *dense, uint32_t *bits, hipStream_t s) {
  if (A.rows == 0 || B.cols == 0) return OC_OK;
  const Tables &t = tables();
  const size_t lds = fat_lds_bytes(B.cols);
  if (lds > t.lds_max_block) {
    set_error("FAT accumulator needs %zu B of LDS, device allows %u", lds, t.lds_max_block);
    return OC_ERR_INVALID;
  }
  if (G < 1 || G > 64 || (G & (G - 1))) {
    set_error("FAT: lanes per A entry must be a power of two up to 64");
    return OC_ERR_INVALID;
  }
  const uint64_t grid = A.rows * (uint64_t)splits;
  if (grid >= (1ull << 31)) {
    set_error("FAT: %llu workgroups", (unsigned long long)grid);
    return OC_ERR_INVALID;
  }
#define GO(T)                                                                          \
  do {                                                                                 \
    FatArgs<T> f{A.rp, A.ci, (const T *)A.v, B.rp, B.ci, (const T *)B.v, (uint32_t)B.cols, \
                 (uint32_t)((B.cols + 31) / 32), splits, G, dense, bits};               \
    if (lds > 65536)                                                                   \
      HIP_TRY(hipFuncSetAttribute((const void *)k_fat_accumulate<T>,                   \
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds)); \
    hipLaunchKernelGGL(k_fat_accumulate<T>, dim3((uint32_t)grid), dim3(kFatBlock), lds, s, f); \
  } while (0)
  if (A.vtype == OC_F64) GO(double);
  else GO(float);
#undef GO
  LAUNCH_CHECK();
  return OC_OK;
}

int launch_fat_count(const uint32_t *bits, uint64_t rows, uint64_t b_cols, uint32_t *nnz,
                     hipStream_t s) {
  if (rows == 0) return OC_OK;
  const uint32_t nwords = (uint32_t)((b_cols + 31) / 32);
  hipLaunchKernelGGL(k_fat_count, dim3((uint32_t)((rows + 3) / 4)), dim3(256), 0, s, bits,
                     (uint32_t)rows, nwords, nnz);
  LAUNCH_CHECK();
  return OC_OK;
}

int launch_fat_compact(int vtype, const double *dense, const uint32_t *bits, uint64_t rows,
                       uint64_t b_cols, const uint64_t *c_off, uint32_t *c_cols, void *c_vals,
                       hipStream_t s) {
  if (rows == 0) return OC_OK;
  const uint32_t nwords = (uint32_t)((b_cols + 31) / 32);
  if (vtype == OC_F64)
    hipLaunchKernelGGL(k_fat_compact<double>, dim3((uint32_t)rows), dim3(kFatBlock), 0, s, dense,
                       bits, (uint32_t)b_cols, nwords, c_off, c_cols, (double *)c_vals);
  else
    hipLaunchKernelGGL(k_fat_compact<float>, dim3((uint32_t)rows), dim3(kFatBlock), 0, s, dense,
                       bits, (uint32_t)b_cols, nwords, c_off, c_cols, (float *)c_vals);
  LAUNCH_CHECK();
  return OC_OK;
}

}  // namespace oc
