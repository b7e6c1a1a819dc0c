// Shared pieces of the accumulator kernel files: error macros, kernel
// argument structs (CSR inputs, numeric outputs) and small LDS helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

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

template <typename T>
struct KArgs {
  const uint32_t *__restrict__ a_rp;
  const uint32_t *__restrict__ a_ci;
  const T *__restrict__ a_v;
  const uint32_t *__restrict__ b_rp;
  const uint32_t *__restrict__ b_ci;
  const T *__restrict__ b_v;
  uint32_t b_cols;
  const uint32_t *__restrict__ products;
  const uint32_t *__restrict__ max_b;
  const uint32_t *__restrict__ left;
  const uint32_t *__restrict__ right;
};

template <typename T>
struct KOut {
  const uint64_t *off;
  uint32_t *cols;
  T *vals;
  uint32_t *nnz;
  const uint32_t *cap;
  uint32_t *ovf_rows;
  uint32_t *ovf_count;
  const uint32_t *nnz_hint;
};

template <typename T>
static KArgs<T> kargs(const AccArgs &a) {
  KArgs<T> k;
  k.a_rp = a.A.rp;
  k.a_ci = a.A.ci;
  k.a_v = (const T *)a.A.v;
  k.b_rp = a.B.rp;
  k.b_ci = a.B.ci;
  k.b_v = (const T *)a.B.v;
  k.b_cols = (uint32_t)a.B.cols;
  k.products = a.products;
  k.max_b = a.max_b;
  k.left = a.left;
  k.right = a.right;
  return k;
}

template <typename T>
static KOut<T> kout(const NumOut &o) {
  KOut<T> k;
  k.off = o.off;
  k.cols = o.cols;
  k.vals = (T *)o.vals;
  k.nnz = o.nnz;
  k.cap = o.cap;
  k.ovf_rows = o.ovf_rows;
  k.ovf_count = o.ovf_count;
  k.nnz_hint = o.nnz_hint;
  return k;
}

__device__ __forceinline__ void lds_add_f64(double *p, double v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}


template <typename K>
static int set_lds(K kernel, size_t bytes) {
  if (bytes > 65536)
    HIP_TRY(hipFuncSetAttribute((const void *)kernel,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)bytes));
  return OC_OK;
}


// Clears work counters with a kernel rather than a memset node, so a
// launch sequence can be captured into a HIP graph and replayed.
static __global__ void k_zero_u32(uint32_t *p, uint32_t n) {
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}
static inline int zero_u32(uint32_t *p, uint32_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, p, n);
  LAUNCH_CHECK();
  return OC_OK;
}

// Most entries per B row of the compact tile-start table of the widest
// tile (its tiles + 1: 63 tiles of 2^17 columns reach 8.2M columns).
constexpr uint32_t kMaxWideStarts = 64;

}  // namespace oc
