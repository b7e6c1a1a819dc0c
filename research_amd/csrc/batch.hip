// Batched SpGEMM: many independent products C_i = A_i * B_i as one product
// of block-diagonal operands.
//
// The inputs are concatenated on the device: A_i's rows follow A_{i-1}'s,
// its column ids are shifted by the sum of the earlier A_j.cols (which is
// B_i's row offset in the concatenated B), and B_i's column ids by the sum
// of the earlier B_j.cols.  The product of the concatenated operands is
// block diagonal with C_i in its blocks, so one run of the whole path (one
// analysis, one binning over the rows of every product, one launch per bin)
// computes every C_i; launch counts do not grow with the batch size.
//
// For repeated batches with a fixed structure the batch keeps a numeric
// plan, and one captured HIP graph refreshes the concatenated values from
// the inputs and recomputes every product's values.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

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

struct BatchDesc {
  const uint32_t *rp, *ci;
  const void *v;
  uint64_t rows, nnz, row_base, nnz_base;
  uint32_t col_base, pad;
};

// The batch's operands as one block-diagonal CSR.  The rows and the entries
// of all products are cut into chunks of kConcatChunk and each workgroup
// copies one chunk, so a large product is spread over many workgroups (one
// workgroup per product left the largest ones to a single CU).  A thread
// finds the product of its first element by a binary search over the
// products' first rows (entries), then steps forward over the short products
// a chunk may hold.  VALUES_ONLY: refresh the values.
constexpr uint32_t kConcatChunk = 8192;

template <bool ENTRIES>
__device__ __forceinline__ uint64_t desc_base(const BatchDesc &x) {
  return ENTRIES ? x.nnz_base : x.row_base;
}

// the last product whose first row (entry) is at or below e; d[0] starts at 0
template <bool ENTRIES>
__device__ uint32_t product_of(const BatchDesc *__restrict__ d, uint32_t count, uint64_t e) {
  uint32_t lo = 0, hi = count - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) / 2;
    if (desc_base<ENTRIES>(d[mid]) <= e) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <bool ENTRIES, bool VALUES_ONLY, typename V>
__device__ void concat_chunk(const BatchDesc *__restrict__ d, uint32_t count, uint64_t c,
                             uint64_t n, uint32_t *__restrict__ rp_out,
                             uint32_t *__restrict__ ci_out, V *__restrict__ v_out) {
  const uint64_t x0 = c * kConcatChunk + threadIdx.x;
  const uint64_t x1 = min(n, (c + 1) * (uint64_t)kConcatChunk);
  if (x0 >= x1) return;
  uint32_t i = product_of<ENTRIES>(d, count, x0);
  BatchDesc D = d[i];
  uint64_t next = i + 1 < count ? desc_base<ENTRIES>(d[i + 1]) : ~0ull;
  uint32_t r0 = D.rp[0];
  for (uint64_t x = x0; x < x1; x += 256) {
    if (x >= next) {
      do {
        ++i;
        next = i + 1 < count ? desc_base<ENTRIES>(d[i + 1]) : ~0ull;
      } while (x >= next);
      D = d[i];
      r0 = D.rp[0];
    }
    if (!ENTRIES) {
      rp_out[x] = (uint32_t)(D.nnz_base + (D.rp[x - D.row_base] - r0));
    } else {
      const uint64_t src = r0 + (x - D.nnz_base);
      if (!VALUES_ONLY) ci_out[x] = D.ci[src] + D.col_base;
      v_out[x] = reinterpret_cast<const V *>(D.v)[src];
    }
  }
}

// blocks [0, row_chunks): rows (unless VALUES_ONLY); the rest: entries
template <bool VALUES_ONLY, typename V>
__global__ __launch_bounds__(256) void k_batch_concat(const BatchDesc *__restrict__ d,
                                                      uint32_t count, uint32_t row_chunks,
                                                      uint32_t *__restrict__ rp_out,
                                                      uint32_t *__restrict__ ci_out,
                                                      V *__restrict__ v_out,
                                                      uint64_t total_rows, uint64_t total_nnz) {
  if (!VALUES_ONLY && blockIdx.x == 0 && threadIdx.x == 0)
    rp_out[total_rows] = (uint32_t)total_nnz;
  if (!VALUES_ONLY && blockIdx.x < row_chunks)
    concat_chunk<false, VALUES_ONLY, V>(d, count, blockIdx.x, total_rows, rp_out, ci_out, v_out);
  else
    concat_chunk<true, VALUES_ONLY, V>(d, count, blockIdx.x - (VALUES_ONLY ? 0u : row_chunks),
                                       total_nnz, rp_out, ci_out, v_out);
}

// Product i of a batched result as its own CSR (u64 offsets from 0, column
// ids rebased).
template <typename T>
__global__ void k_batch_extract(const uint64_t *__restrict__ rp, const uint32_t *__restrict__ ci,
                                const T *__restrict__ v, uint64_t row_base, uint64_t rows,
                                uint32_t col_base, uint64_t *__restrict__ rp_out,
                                uint32_t *__restrict__ ci_out, T *__restrict__ v_out) {
  const uint64_t base = rp[row_base];
  const uint64_t nnz = rp[row_base + rows] - base;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t t0 = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  for (uint64_t j = t0; j <= rows; j += stride) rp_out[j] = rp[row_base + j] - base;
  for (uint64_t e = t0; e < nnz; e += stride) {
    ci_out[e] = ci[base + e] - col_base;
    v_out[e] = v[base + e];
  }
}

int spgemm_replay_async(oc_spgemm *h, hipStream_t s);
int spgemm_full_replay_async(oc_spgemm *h, hipStream_t s);
int spgemm_full_status(oc_spgemm *h, uint32_t *status);
void spgemm_drop_plan(oc_spgemm *h);
void spgemm_set_blocks(oc_spgemm *h, const uint32_t *b_block_base, uint64_t span,
                       const void *b_row_offsets);

}  // namespace oc

using namespace oc;

struct oc_batch {
  oc_spgemm *h = nullptr;
  uint32_t count = 0;
  int vtype = OC_F64;
  bool b_is_a = false;
  oc_csr A{}, B{}, C{};
  std::vector<uint64_t> row_base, col_base;  // count + 1 each
  uint64_t a_total_nnz = 0, b_total_nnz = 0;
  BatchDesc *dA = nullptr, *dB = nullptr;
  uint32_t *d_bbase = nullptr;  // per row of the concatenated B: its block's first column
  bool planned = false;
  hipStream_t own = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph_full = nullptr;  // the whole path (oc_batch_capture_full)
  hipGraphExec_t exec_full = nullptr;

  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (exec_full) (void)hipGraphExecDestroy(exec_full);
    if (graph_full) (void)hipGraphDestroy(graph_full);
    exec = nullptr;
    graph = nullptr;
    exec_full = nullptr;
    graph_full = nullptr;
  }
  void drop_result() {
    drop_graph();
    spgemm_drop_plan(h);  // the plan refers to C
    oc_csr_free(&C);
    planned = false;
  }
  void drop_inputs() {
    drop_result();
    if (!b_is_a) oc_csr_free(&B);
    std::memset(&B, 0, sizeof(B));
    oc_csr_free(&A);
    spgemm_set_blocks(h, nullptr, 0, nullptr);
    if (dA) (void)hipFree(dA);
    if (dB) (void)hipFree(dB);
    if (d_bbase) (void)hipFree(d_bbase);
    dA = dB = nullptr;
    d_bbase = nullptr;
    count = 0;
  }
};

static int concat(const oc_batch *b, const BatchDesc *d, const oc_csr &M, uint64_t total_nnz,
                  bool values_only, hipStream_t s) {
  const uint64_t rc = values_only ? 0 : (M.rows + kConcatChunk - 1) / kConcatChunk;
  const uint64_t ec = (total_nnz + kConcatChunk - 1) / kConcatChunk;
  if (rc + ec > 0x7fffffffull) {
    set_error("batched operand of %llu rows and %llu entries is too large",
              (unsigned long long)M.rows, (unsigned long long)total_nnz);
    return OC_ERR_INVALID;
  }
  const dim3 grid((uint32_t)std::max<uint64_t>(1, rc + ec));
  uint32_t *rp = (uint32_t *)M.row_offsets;
#define GO(VO, V)                                                                       \
  hipLaunchKernelGGL((k_batch_concat<VO, V>), grid, dim3(256), 0, s, d, b->count,        \
                     (uint32_t)rc, rp, M.col_ids, (V *)M.values, (uint64_t)M.rows, total_nnz)
  if (vsize(b->vtype) == 8) {
    if (values_only) GO(true, double);
    else GO(false, double);
  } else {
    if (values_only) GO(true, float);
    else GO(false, float);
  }
#undef GO
  LAUNCH_CHECK();
  return OC_OK;
}

extern "C" oc_batch *oc_batch_create(const oc_config *cfg) {
  oc_batch *b = new (std::nothrow) oc_batch();
  if (!b) {
    set_error("out of host memory");
    return nullptr;
  }
  b->h = oc_spgemm_create(cfg);
  if (!b->h) {
    delete b;
    return nullptr;
  }
  return b;
}

extern "C" void oc_batch_destroy(oc_batch *b) {
  if (!b) return;
  b->drop_inputs();
  if (b->own) (void)hipStreamDestroy(b->own);
  oc_spgemm_destroy(b->h);
  delete b;
}

extern "C" int oc_batch_set_inputs(oc_batch *b, const oc_csr *As, const oc_csr *Bs,
                                   uint32_t count, void *stream) {
  OC_REQUIRE(b && As && count > 0, "oc_batch_set_inputs: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  b->drop_inputs();
  const int vt = As[0].vtype;
  OC_REQUIRE(vt == OC_F64 || vt == OC_F32, "bad value type");
  std::vector<BatchDesc> da(count), db(count);
  uint64_t rows = 0, acols = 0, bcols = 0, annz = 0, bnnz = 0;
  std::vector<uint64_t> rb(count + 1), cb(count + 1);
  for (uint32_t i = 0; i < count; ++i) {
    const oc_csr &a = As[i];
    const oc_csr &bb = Bs ? Bs[i] : As[i];
    OC_REQUIRE(a.on_device && bb.on_device && a.off_bytes == 4 && bb.off_bytes == 4,
               "batch input %u: need device matrices with 32-bit offsets", i);
    OC_REQUIRE(a.vtype == vt && bb.vtype == vt, "batch input %u: value types differ", i);
    OC_REQUIRE(a.cols == bb.rows, "batch input %u: Incompatible matrix dimensions!", i);
    da[i] = BatchDesc{(const uint32_t *)a.row_offsets, a.col_ids, a.values, a.rows, a.nnz,
                      rows, annz, (uint32_t)acols, 0};
    db[i] = BatchDesc{(const uint32_t *)bb.row_offsets, bb.col_ids, bb.values, bb.rows, bb.nnz,
                      acols, bnnz, (uint32_t)bcols, 0};
    rb[i] = rows;
    cb[i] = bcols;
    rows += a.rows;
    acols += a.cols;
    bcols += bb.cols;
    annz += a.nnz;
    bnnz += bb.nnz;
    OC_REQUIRE(rows < (1ull << 32) && acols < (1ull << 32) && bcols < (1ull << 32) &&
                   annz < (1ull << 32) && bnnz < (1ull << 32),
               "batch exceeds 32-bit rows, columns or entries at input %u", i);
  }
  rb[count] = rows;
  cb[count] = bcols;
  b->count = count;
  b->vtype = vt;
  b->b_is_a = Bs == nullptr;
  b->row_base = rb;
  b->col_base = cb;
  b->a_total_nnz = annz;
  b->b_total_nnz = bnnz;
  OC_TRY(oc_csr_alloc(&b->A, rows, acols, annz, vt, 4, 1));
  HIP_TRY(hipMalloc((void **)&b->dA, sizeof(BatchDesc) * count));
  HIP_TRY(hipMemcpy(b->dA, da.data(), sizeof(BatchDesc) * count, hipMemcpyHostToDevice));
  OC_TRY(concat(b, b->dA, b->A, annz, false, s));
  if (b->b_is_a) {
    b->B = b->A;
    b->B.owned = 0;
  } else {
    OC_TRY(oc_csr_alloc(&b->B, acols, bcols, bnnz, vt, 4, 1));
    HIP_TRY(hipMalloc((void **)&b->dB, sizeof(BatchDesc) * count));
    HIP_TRY(hipMemcpy(b->dB, db.data(), sizeof(BatchDesc) * count, hipMemcpyHostToDevice));
    OC_TRY(concat(b, b->dB, b->B, bnnz, false, s));
  }
  // B_i's rows hold only columns of block i: tiles of the long-row kernel
  // are counted from each block's first column (tile tables then span the
  // widest block, not all of the concatenated columns)
  {
    std::vector<uint32_t> hb(acols ? acols : 1);
    uint64_t span = 0, r = 0;
    for (uint32_t i = 0; i < count; ++i) {
      for (uint64_t j = 0; j < As[i].cols; ++j) hb[r + j] = (uint32_t)cb[i];
      r += As[i].cols;
      span = std::max<uint64_t>(span, cb[i + 1] - cb[i]);
    }
    HIP_TRY(hipMalloc((void **)&b->d_bbase, hb.size() * 4));
    HIP_TRY(hipMemcpy(b->d_bbase, hb.data(), hb.size() * 4, hipMemcpyHostToDevice));
    spgemm_set_blocks(b->h, b->d_bbase, span, b->B.row_offsets);
  }
  HIP_TRY(hipStreamSynchronize(s));
  return OC_OK;
}

extern "C" int oc_batch_run(oc_batch *b, void *stream) {
  OC_REQUIRE(b && b->count, "oc_batch_run: no inputs");
  b->drop_result();
  return oc_spgemm_run(b->h, &b->A, &b->B, &b->C, stream);
}

extern "C" int oc_batch_plan(oc_batch *b, void *stream) {
  OC_REQUIRE(b && b->count, "oc_batch_plan: no inputs");
  b->drop_result();
  OC_TRY(oc_spgemm_plan(b->h, &b->A, &b->B, &b->C, stream));
  b->planned = true;
  return OC_OK;
}

extern "C" int oc_batch_capture(oc_batch *b) {
  OC_REQUIRE(b && b->planned, "oc_batch_capture: run oc_batch_plan first");
  b->drop_graph();
  if (!b->own) HIP_TRY(hipStreamCreate(&b->own));
  HIP_TRY(hipStreamBeginCapture(b->own, hipStreamCaptureModeRelaxed));
  int rc = concat(b, b->dA, b->A, b->a_total_nnz, true, b->own);
  if (rc == OC_OK && !b->b_is_a) rc = concat(b, b->dB, b->B, b->b_total_nnz, true, b->own);
  if (rc == OC_OK) rc = spgemm_replay_async(b->h, b->own);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(b->own, &g);
  if (rc != OC_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  HIP_TRY(e);
  b->graph = g;
  HIP_TRY(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  return OC_OK;
}

// The whole path in one graph: the inputs' structure and values are
// concatenated again, then analysis, binning, symbolic and numeric run for
// every product with bin sizes kept on the device (see
// oc_spgemm_capture_full).  The inputs may change structure between
// launches as long as each keeps its rows, columns and entry count.
extern "C" int oc_batch_capture_full(oc_batch *b) {
  OC_REQUIRE(b && b->planned, "oc_batch_capture_full: run oc_batch_plan first");
  if (b->exec_full) (void)hipGraphExecDestroy(b->exec_full);
  if (b->graph_full) (void)hipGraphDestroy(b->graph_full);
  b->exec_full = nullptr;
  b->graph_full = nullptr;
  if (!b->own) HIP_TRY(hipStreamCreate(&b->own));
  HIP_TRY(hipStreamBeginCapture(b->own, hipStreamCaptureModeRelaxed));
  int rc = concat(b, b->dA, b->A, b->a_total_nnz, false, b->own);
  if (rc == OC_OK && !b->b_is_a) rc = concat(b, b->dB, b->B, b->b_total_nnz, false, b->own);
  if (rc == OC_OK) rc = spgemm_full_replay_async(b->h, b->own);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(b->own, &g);
  if (rc != OC_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  HIP_TRY(e);
  b->graph_full = g;
  HIP_TRY(hipGraphInstantiate(&b->exec_full, g, nullptr, nullptr, 0));
  return OC_OK;
}

extern "C" int oc_batch_launch_full(oc_batch *b, void *stream) {
  OC_REQUIRE(b && b->exec_full, "oc_batch_launch_full: run oc_batch_capture_full first");
  if (stream) {
    HIP_TRY(hipGraphLaunch(b->exec_full, (hipStream_t)stream));
  } else {
    HIP_TRY(hipGraphLaunch(b->exec_full, b->own));
    HIP_TRY(hipStreamSynchronize(b->own));
  }
  return OC_OK;
}

extern "C" int oc_batch_full_status(oc_batch *b, uint32_t *status) {
  OC_REQUIRE(b && status, "oc_batch_full_status: bad arguments");
  return spgemm_full_status(b->h, status);
}

extern "C" int oc_batch_launch(oc_batch *b, void *stream) {
  OC_REQUIRE(b && b->exec, "oc_batch_launch: run oc_batch_capture first");
  if (stream) {
    HIP_TRY(hipGraphLaunch(b->exec, (hipStream_t)stream));
  } else {
    HIP_TRY(hipGraphLaunch(b->exec, b->own));
    HIP_TRY(hipStreamSynchronize(b->own));
  }
  return OC_OK;
}

extern "C" int oc_batch_result(const oc_batch *b, oc_csr *C, uint64_t *row_base,
                               uint64_t *col_base) {
  OC_REQUIRE(b && C && b->C.row_offsets, "oc_batch_result: no result");
  *C = b->C;
  C->owned = 0;  // a view: the batch keeps the memory
  if (row_base) std::memcpy(row_base, b->row_base.data(), 8 * (b->count + 1));
  if (col_base) std::memcpy(col_base, b->col_base.data(), 8 * (b->count + 1));
  return OC_OK;
}

extern "C" int oc_batch_extract(const oc_batch *b, uint32_t i, oc_csr *Ci, void *stream) {
  OC_REQUIRE(b && Ci && b->C.row_offsets, "oc_batch_extract: no result");
  OC_REQUIRE(i < b->count, "oc_batch_extract: product %u of %u", i, b->count);
  hipStream_t s = (hipStream_t)stream;
  const uint64_t r0 = b->row_base[i], rows = b->row_base[i + 1] - r0;
  uint64_t off[2];
  HIP_TRY(hipMemcpyAsync(&off[0], (const uint64_t *)b->C.row_offsets + r0, 8,
                         hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(&off[1], (const uint64_t *)b->C.row_offsets + r0 + rows, 8,
                         hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  const uint64_t nnz = off[1] - off[0];
  OC_TRY(oc_csr_alloc(Ci, rows, b->col_base[i + 1] - b->col_base[i], nnz, b->vtype, 8, 1));
  const uint64_t work = std::max<uint64_t>(rows + 1, nnz);
  const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((work + 255) / 256, 4096));
  if (b->vtype == OC_F64)
    hipLaunchKernelGGL(k_batch_extract<double>, dim3(grid), dim3(256), 0, s,
                       (const uint64_t *)b->C.row_offsets, b->C.col_ids,
                       (const double *)b->C.values, r0, rows, (uint32_t)b->col_base[i],
                       (uint64_t *)Ci->row_offsets, Ci->col_ids, (double *)Ci->values);
  else
    hipLaunchKernelGGL(k_batch_extract<float>, dim3(grid), dim3(256), 0, s,
                       (const uint64_t *)b->C.row_offsets, b->C.col_ids,
                       (const float *)b->C.values, r0, rows, (uint32_t)b->col_base[i],
                       (uint64_t *)Ci->row_offsets, Ci->col_ids, (float *)Ci->values);
  LAUNCH_CHECK();
  HIP_TRY(hipStreamSynchronize(s));
  return OC_OK;
}

extern "C" oc_spgemm *oc_batch_handle(oc_batch *b) { return b ? b->h : nullptr; }
