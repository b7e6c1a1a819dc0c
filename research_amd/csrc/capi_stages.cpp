// Per-stage C-ABI entry points (device pointers in, device pointers out).
// They run one stage of the hot path on their own, for tests and for
// callers that compose the stages themselves.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "launch.hpp"

using namespace oc;

#define HIP_TRY(call)                                                         \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess) {                                                   \
      set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,                    \
                hipGetErrorString(e_));                                       \
      return OC_ERR_HIP;                                                      \
    }                                                                         \
  } while (0)

namespace {

DevCsr dev_view(const oc_csr *m) {
  return DevCsr{m->rows, m->cols, m->nnz, (const uint32_t *)m->row_offsets,
                m->col_ids, m->values, m->vtype};
}

int need_dev32(const oc_csr *m, const char *what) {
  OC_REQUIRE(m && m->on_device && m->off_bytes == 4,
             "%s: need a device matrix with 32-bit offsets", what);
  return OC_OK;
}

// Scoped stream-ordered temporary.
struct Tmp {
  void *p = nullptr;
  hipStream_t s;
  explicit Tmp(hipStream_t st) : s(st) {}
  ~Tmp() {
    if (p) (void)hipFreeAsync(p, s);
  }
  int get(size_t bytes) {
    hipError_t e = hipMallocAsync(&p, std::max<size_t>(bytes, 256), s);
    if (e != hipSuccess) {
      set_error("temporary allocation of %zu bytes failed", bytes);
      return OC_ERR_NOMEM;
    }
    return OC_OK;
  }
};

int sync(hipStream_t s) {
  HIP_TRY(hipStreamSynchronize(s));
  return OC_OK;
}

}  // namespace

extern "C" int oc_get_tables(int vtype, oc_tables *out) {
  OC_REQUIRE(out, "oc_get_tables: null output");
  OC_REQUIRE(vtype == OC_F64 || vtype == OC_F32, "bad vtype");
  OC_TRY(tables_init());
  const Tables &t = tables();
  std::memcpy(out->hash_sym, t.hash_sym, sizeof(out->hash_sym));
  std::memcpy(out->dense_sym, t.dense_sym, sizeof(out->dense_sym));
  std::memcpy(out->hash_num, t.hash_num, sizeof(out->hash_num));
  std::memcpy(out->dense_num, t.dense_num, sizeof(out->dense_num));
  std::memcpy(out->block_size, t.block, sizeof(out->block_size));
  std::memcpy(out->lds_bytes, t.lds, sizeof(out->lds_bytes));
  out->max_dense_sym_bin = t.max_dense_sym_bin;
  out->window_cols = t.win_num_cols;
  out->lds_per_cu = t.lds_per_cu;
  out->num_cu = t.num_cu;
  return OC_OK;
}

extern "C" int oc_analysis(const oc_csr *dA, const oc_csr *dB, uint32_t *num_products,
                           uint32_t *max_b_len, uint32_t *avg_b_len, uint32_t *b_left,
                           uint32_t *b_right, uint64_t *products_psum,
                           uint32_t *max_a_nnz, uint32_t *max_products, void *stream) {
  OC_TRY(need_dev32(dA, "oc_analysis A"));
  OC_TRY(need_dev32(dB, "oc_analysis B"));
  OC_REQUIRE(dA->cols == dB->rows, "Incompatible matrix dimensions!");
  hipStream_t s = (hipStream_t)stream;
  RowInfo ri{num_products, max_b_len, avg_b_len, b_left, b_right};
  Tmp sc(s);
  OC_TRY(sc.get(16 + scan_temp_bytes(dA->rows + 1)));
  uint32_t *scal = (uint32_t *)sc.p;
  OC_TRY(launch_analysis(dev_view(dA), dev_view(dB), ri, scal, scal + 1, nullptr, s));
  if (products_psum)
    OC_TRY(launch_scan_u32(num_products, products_psum, dA->rows, (uint8_t *)sc.p + 16, s));
  if (max_a_nnz) HIP_TRY(hipMemcpyAsync(max_a_nnz, scal, 4, hipMemcpyDeviceToDevice, s));
  if (max_products) HIP_TRY(hipMemcpyAsync(max_products, scal + 1, 4, hipMemcpyDeviceToDevice, s));
  return sync(s);
}

extern "C" int oc_scan_u32(const uint32_t *in, uint64_t *out, uint64_t n, void *stream) {
  OC_REQUIRE(in || n == 0, "oc_scan_u32: null input");
  OC_REQUIRE(out, "oc_scan_u32: null output");
  hipStream_t s = (hipStream_t)stream;
  Tmp t(s);
  OC_TRY(t.get(scan_temp_bytes(n)));
  OC_TRY(launch_scan_u32(in, out, n, t.p, s));
  return sync(s);
}

extern "C" int oc_murmur3_cols(const uint32_t *cols, uint32_t *hashes, uint64_t n,
                               void *stream) {
  hipStream_t s = (hipStream_t)stream;
  OC_TRY(launch_murmur(cols, hashes, n, s));
  return sync(s);
}

extern "C" int oc_hll_build(int log_p, const oc_csr *dB, uint8_t *hll_b, void *stream) {
  OC_TRY(need_dev32(dB, "oc_hll_build"));
  hipStream_t s = (hipStream_t)stream;
  OC_TRY(launch_hll_build(log_p, dev_view(dB), hll_b, s));
  return sync(s);
}

extern "C" int oc_hll_estimate(int log_p, const oc_csr *dA, const uint8_t *hll_b,
                               double expansion, uint32_t *est, void *stream) {
  OC_TRY(need_dev32(dA, "oc_hll_estimate"));
  hipStream_t s = (hipStream_t)stream;
  Tmp t(s);
  OC_TRY(t.get(hll_merge_temp_bytes(dA->rows)));
  OC_TRY(launch_hll_merge(log_p, dev_view(dA), hll_b, expansion, est, 0, nullptr, nullptr,
                          nullptr, nullptr, 0, t.p, s));
  return sync(s);
}

extern "C" int oc_hll_sample(int log_p, const oc_csr *dA, const uint8_t *hll_b,
                             uint32_t nsamples, uint32_t *sample_rows, int32_t *sample_est,
                             void *stream) {
  OC_TRY(need_dev32(dA, "oc_hll_sample"));
  OC_REQUIRE(dA->rows > 0, "oc_hll_sample: empty matrix");
  hipStream_t s = (hipStream_t)stream;
  Tmp t(s);
  OC_TRY(t.get(hll_merge_temp_bytes(nsamples)));
  OC_TRY(launch_hll_merge(log_p, dev_view(dA), hll_b, 1.0, nullptr, nsamples, sample_rows,
                          sample_est, nullptr, nullptr, 1, t.p, s));
  return sync(s);
}

static int bin_common(BinParams &bp, uint64_t rows, uint8_t *bin_of_row,
                      uint32_t *bin_rows, uint32_t *bin_offsets, void *stream) {
  hipStream_t s = (hipStream_t)stream;
  Tmp t(s);
  OC_TRY(t.get(bin_temp_bytes(rows)));
  OC_TRY(launch_binning(bp, rows, bin_of_row, bin_rows, bin_offsets, t.p, s));
  return sync(s);
}

extern "C" int oc_bin_symbolic(int vtype, uint64_t rows, const uint32_t *num_products,
                               const uint32_t *b_left, const uint32_t *b_right,
                               int use_compaction, double expand_factor,
                               double avg_compaction, uint8_t *bin_of_row,
                               uint32_t *bin_rows, uint32_t *bin_offsets, void *stream) {
  OC_REQUIRE(vtype == OC_F64 || vtype == OC_F32, "bad vtype");
  BinParams bp;
  std::memset(&bp, 0, sizeof(bp));
  bp.kind = 0;
  bp.input = num_products;
  bp.left = b_left;
  bp.right = b_right;
  bp.use_compaction = use_compaction;
  bp.expand_factor = expand_factor;
  bp.avg_compaction = avg_compaction;
  return bin_common(bp, rows, bin_of_row, bin_rows, bin_offsets, stream);
}

extern "C" int oc_bin_numeric(int vtype, uint64_t rows, uint32_t *input,
                              const uint32_t *b_left, const uint32_t *b_right, int expand,
                              double expand_factor, int modify_dense_est, int use_largest,
                              uint32_t hybrid_max, uint8_t *bin_of_row, uint32_t *bin_rows,
                              uint32_t *bin_offsets, void *stream) {
  OC_REQUIRE(vtype == OC_F64 || vtype == OC_F32, "bad vtype");
  BinParams bp;
  std::memset(&bp, 0, sizeof(bp));
  bp.kind = 1;
  bp.input = input;
  bp.input_mut = modify_dense_est ? input : nullptr;
  bp.left = b_left;
  bp.right = b_right;
  bp.expand = expand;
  bp.expand_factor = expand_factor;
  bp.modify_dense_est = modify_dense_est;
  bp.use_largest = use_largest;
  bp.hybrid_max = hybrid_max;
  return bin_common(bp, rows, bin_of_row, bin_rows, bin_offsets, stream);
}

extern "C" int oc_symbolic(const oc_config *cfg, const oc_csr *dA, const oc_csr *dB,
                           uint32_t *nnz_row, void *stream) {
  OC_TRY(need_dev32(dA, "oc_symbolic A"));
  OC_TRY(need_dev32(dB, "oc_symbolic B"));
  OC_REQUIRE(dA->cols == dB->rows, "Incompatible matrix dimensions!");
  oc_config c;
  if (cfg) c = *cfg;
  else oc_config_default(&c);
  hipStream_t s = (hipStream_t)stream;
  const uint64_t m = dA->rows;
  if (m == 0) return OC_OK;
  // row analysis, symbolic binning and the symbolic kernels of every bin
  Tmp t(s);
  const size_t off_ri = 0, ri_bytes = 5 * (m + 1) * 4;
  const size_t off_bins = (off_ri + ri_bytes + 255) & ~(size_t)255;
  const size_t bins_bytes = (m + 1) + (m + 1) * 4 + 64 * 4 + bin_temp_bytes(m) + 1024;
  const size_t off_win = (off_bins + bins_bytes + 255) & ~(size_t)255;
  OC_TRY(t.get(off_win + 1024));
  uint8_t *base = (uint8_t *)t.p;
  RowInfo ri;
  ri.products = (uint32_t *)(base + off_ri);
  ri.max_b = ri.products + (m + 1);
  ri.avg_b = ri.max_b + (m + 1);
  ri.left = ri.avg_b + (m + 1);
  ri.right = ri.left + (m + 1);
  uint32_t *scal = nullptr;
  HIP_TRY(hipMallocAsync((void **)&scal, 256, s));
  const DevCsr A = dev_view(dA), B = dev_view(dB);
  int rc = launch_analysis(A, B, ri, scal, scal + 1, nullptr, s);
  uint32_t hs[2] = {0, 0};
  if (rc == OC_OK) {
    HIP_TRY(hipMemcpyAsync(hs, scal, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
  }
  uint8_t *bor = base + off_bins;
  uint32_t *brows = (uint32_t *)(((uintptr_t)(bor + m + 1) + 3) & ~(uintptr_t)3);
  uint32_t *boff = brows + (m + 1);
  uint8_t *btmp = (uint8_t *)(boff + 64);
  BinParams bp;
  std::memset(&bp, 0, sizeof(bp));
  bp.kind = 0;
  bp.input = ri.products;
  bp.left = ri.left;
  bp.right = ri.right;
  bp.expand_factor = c.symbolic_expansion_coe;
  if (rc == OC_OK) rc = launch_binning(bp, m, bor, brows, boff, btmp, s);
  uint32_t hoff[kBinSlots + 1];
  if (rc == OC_OK) {
    HIP_TRY(hipMemcpyAsync(hoff, boff, sizeof(hoff), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
  }
  void *win = nullptr;
  if (rc == OC_OK) {
    hipError_t e = hipMallocAsync(&win, window_scratch_bytes(hs[0]) + 256, s);
    if (e != hipSuccess) rc = OC_ERR_NOMEM;
  }
  AccArgs a;
  a.A = A;
  a.B = B;
  a.products = ri.products;
  a.max_b = ri.max_b;
  a.left = ri.left;
  a.right = ri.right;
  const Tables &tb = tables();
  for (int b = 0; rc == OC_OK && b <= 5; ++b) {
    const uint32_t n = hoff[b + 1] - hoff[b];
    if (n) rc = launch_sym_hash(b, a, brows + hoff[b], n, tb.hash_sym[b], nnz_row,
                                scal + 8, scal + 4, s);
  }
  if (rc == OC_OK && hoff[7] > hoff[6])
    rc = launch_sym_window(a, brows + hoff[6], hoff[7] - hoff[6], hs[0], nnz_row, win, s);
  for (int b = 8; rc == OC_OK && b <= 13; ++b) {
    const uint32_t n = hoff[b + 1] - hoff[b];
    if (n) rc = launch_sym_dense(b - 7, a, brows + hoff[b], n, nnz_row, s);
  }
  if (win) (void)hipFreeAsync(win, s);
  (void)hipFreeAsync(scal, s);
  if (rc != OC_OK) return rc;
  return sync(s);
}

extern "C" int oc_sort_rows(oc_csr *dM, const uint32_t *rows, uint64_t n, void *stream) {
  OC_REQUIRE(dM && dM->on_device, "oc_sort_rows: need a device matrix");
  hipStream_t s = (hipStream_t)stream;
  const uint64_t cnt = rows ? n : dM->rows;
  OC_TRY(launch_sort_rows(*dM, rows, cnt, s));
  return sync(s);
}

extern "C" int oc_csr_validate(const oc_csr *dM, uint64_t *bad_rows, void *stream) {
  OC_REQUIRE(dM && dM->on_device && bad_rows, "oc_csr_validate: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  unsigned long long *d = nullptr;
  HIP_TRY(hipMallocAsync((void **)&d, 8, s));
  HIP_TRY(hipMemsetAsync(d, 0, 8, s));
  int rc = launch_validate(*dM, d, s);
  unsigned long long h = 0;
  if (rc == OC_OK) HIP_TRY(hipMemcpyAsync(&h, d, 8, hipMemcpyDeviceToHost, s));
  (void)hipFreeAsync(d, s);
  OC_TRY(rc);
  OC_TRY(sync(s));
  *bad_rows = h;
  return OC_OK;
}

extern "C" int oc_csr_row_set_hash(const oc_csr *dM, uint64_t *hash, void *stream) {
  OC_REQUIRE(dM && dM->on_device && hash, "oc_csr_row_set_hash: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  OC_TRY(launch_row_set_hash(*dM, hash, s));
  return sync(s);
}

extern "C" int oc_csr_spmv(const oc_csr *dM, const double *x, double *y, void *stream) {
  OC_REQUIRE(dM && dM->on_device && x && y, "oc_csr_spmv: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  OC_TRY(launch_spmv(*dM, x, y, s));
  return sync(s);
}

// ------------------------------------------------------ device memory
extern "C" int oc_device_count(int *n) {
  OC_REQUIRE(n, "null output");
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) {
    (void)hipGetLastError();
    c = 0;
  }
  *n = c;
  return OC_OK;
}

extern "C" int oc_device_malloc(void **p, size_t bytes) {
  OC_REQUIRE(p, "null output");
  hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 1));
  if (e != hipSuccess) {
    set_error("hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    return OC_ERR_NOMEM;
  }
  return OC_OK;
}

extern "C" int oc_device_free(void *p) {
  if (p) HIP_TRY(hipFree(p));
  return OC_OK;
}

extern "C" int oc_memcpy_htod(void *dst, const void *src, size_t bytes) {
  if (bytes) HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return OC_OK;
}

extern "C" int oc_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
  if (bytes) HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return OC_OK;
}

extern "C" int oc_memcpy_dtod(void *dst, const void *src, size_t bytes) {
  if (bytes) HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  return OC_OK;
}

extern "C" int oc_memset_device(void *p, int value, size_t bytes) {
  if (bytes) HIP_TRY(hipMemset(p, value, bytes));
  return OC_OK;
}

extern "C" int oc_device_synchronize(void) {
  HIP_TRY(hipDeviceSynchronize());
  return OC_OK;
}

extern "C" int oc_device_mem_info(size_t *free_bytes, size_t *total_bytes) {
  OC_REQUIRE(free_bytes && total_bytes, "oc_device_mem_info: null argument");
  HIP_TRY(hipMemGetInfo(free_bytes, total_bytes));
  return OC_OK;
}
