// Allocation-free row analysis: oc_analysis with its temporaries in a
// file-scope device array instead of stream-ordered allocations, for
// callers that cannot allocate (code under HIP graph capture, serving
// loops, harnesses that forbid hipMalloc).
//
// The array holds kNoAllocScratchBytes per device for the life of the
// process.  Every call on a device shares that device's copy: calls that
// overlap on one device (other streams, other host threads) race.  Each
// call clears the counters it needs before use.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

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

constexpr size_t kNoAllocScratchBytes = 32u << 20;

// Host code never uses this name as an address: the device address of each
// device's copy comes from hipGetSymbolAddress (scratch_ptr below).
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices] = {nullptr};

// Device address of the current device's scratch, looked up once per
// device with that device current.
int scratch_ptr(void **out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  OC_REQUIRE(dev >= 0 && dev < kMaxDevices, "device %d beyond the scratch table", dev);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void *p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    OC_REQUIRE(p != nullptr, "hipGetSymbolAddress gave no address for the scratch array");
    g_ptr[dev] = p;
  }
  *out = g_ptr[dev];
  return OC_OK;
}

DevCsr view(const oc_csr *m) {
  DevCsr d;
  d.rows = m->rows;
  d.cols = m->cols;
  d.nnz = m->nnz;
  d.rp = (const uint32_t *)m->row_offsets;
  d.ci = m->col_ids;
  d.v = m->values;
  d.vtype = m->vtype;
  return d;
}

}  // namespace

size_t analysis_noalloc_limit() { return kNoAllocScratchBytes; }

}  // namespace oc

using namespace oc;

extern "C" int oc_analysis_noalloc(const oc_csr *dA, const oc_csr *dB, uint32_t *num_products,
                                   uint32_t *max_b_len, uint32_t *avg_b_len, uint32_t *b_left,
                                   uint32_t *b_right, uint64_t *products_psum,
                                   uint32_t *max_a_nnz, uint32_t *max_products, void *stream) {
  OC_REQUIRE(dA && dA->on_device && dA->off_bytes == 4,
             "oc_analysis_noalloc A: need a device matrix with 32-bit offsets");
  OC_REQUIRE(dB && dB->on_device && dB->off_bytes == 4,
             "oc_analysis_noalloc B: need a device matrix with 32-bit offsets");
  OC_REQUIRE(dA->cols == dB->rows, "Incompatible matrix dimensions!");
  OC_REQUIRE(num_products && max_b_len && avg_b_len && b_left && b_right,
             "oc_analysis_noalloc: null per-row output");
  hipStream_t s = (hipStream_t)stream;
  // layout: 256 B of counters, the scan partials, the analysis temporaries
  const size_t scan_b = (scan_temp_bytes(dA->rows + 1) + 255) & ~(size_t)255;
  const size_t ana_b = analysis_temp_bytes(dA->rows, dA->nnz, dB->rows);
  const size_t need = 256 + scan_b + ana_b;
  OC_REQUIRE(need <= kNoAllocScratchBytes,
             "oc_analysis_noalloc: these operands need %zu bytes of scratch, beyond the "
             "%zu-byte limit (use oc_analysis)",
             need, kNoAllocScratchBytes);
  void *base = nullptr;
  OC_TRY(scratch_ptr(&base));
  uint8_t *p = (uint8_t *)base;
  uint32_t *scal = (uint32_t *)p;
  RowInfo ri{num_products, max_b_len, avg_b_len, b_left, b_right};
  OC_TRY(launch_analysis(view(dA), view(dB), ri, scal, scal + 1, p + 256 + scan_b, s));
  if (products_psum) OC_TRY(launch_scan_u32(num_products, products_psum, dA->rows, p + 256, s));
  if (max_a_nnz) HIP_TRY(hipMemcpyAsync(max_a_nnz, scal, 4, hipMemcpyDeviceToDevice, s));
  if (max_products) HIP_TRY(hipMemcpyAsync(max_products, scal + 1, 4, hipMemcpyDeviceToDevice, s));
  return OC_OK;
}
