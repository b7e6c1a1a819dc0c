/*
 * ocean_capi.h -- C ABI of the MI355X (gfx950) Ocean SpGEMM library
 * (research_amd/libocean.so).
 *
 * Plain pointers and sizes only.  Status codes are returned from every call,
 * the message of the last failure is kept per thread (oc_last_error), and the
 * library never exits the process.  Device pointers are HIP device memory;
 * `stream` is a hipStream_t (NULL = the null stream).
 *
 * Each declaration names the reference interface it replaces
 * (lingling36109/research, file:line).
 */
#ifndef OCEAN_CAPI_H
#define OCEAN_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define OC_OK 0
#define OC_ERR_INVALID 1   /* bad argument or configuration             */
#define OC_ERR_HIP 2       /* HIP runtime error                         */
#define OC_ERR_CAPACITY 3  /* buffer too small (overflow spare exhausted) */
#define OC_ERR_IO 4        /* file open / read / write error            */
#define OC_ERR_MISMATCH 5  /* compare found differing structure/values  */
#define OC_ERR_NOMEM 6     /* device or host allocation failed          */

/* Message of the last failing call on this thread.
 * Replaces: exit(EXIT_FAILURE) in CHECK_CUDA (include/Common.h:296-310) and
 * the runtime_error texts of kernels/SpGEMM.cuh:362,1586. */
const char *oc_last_error(void);

/* ------------------------------------------------------------------- CSR */
#define OC_F64 0
#define OC_F32 1

/* Replaces: struct CSR / struct cuCSR (include/CSR.h:16-51).  One struct
 * serves host and device matrices (`on_device`).  row_offsets entries are
 * off_bytes wide: inputs A and B use 4 (as the .csr format stores them);
 * C produced by the library uses 8 so that nnz(C) may exceed 2^32. */
typedef struct oc_csr {
  uint64_t rows, cols, nnz;
  int32_t vtype;      /* OC_F64 or OC_F32 (data_t, include/Common.h:8) */
  int32_t off_bytes;  /* 4 or 8 */
  void *row_offsets;  /* rows+1 entries                                */
  uint32_t *col_ids;  /* nnz entries, ascending within a row           */
  void *values;       /* nnz entries of vtype                          */
  int32_t on_device;  /* 1: HIP device memory, 0: host memory          */
  int32_t owned;      /* nonzero: allocated by this library (oc_csr_free) */
} oc_csr;

/* Allocate a host or device matrix (owned).
 * Replaces: CSR::alloc (src/CSR.cpp:45-55), cuCSR::alloc (src/CSR.cpp:185-209). */
int oc_csr_alloc(oc_csr *m, uint64_t rows, uint64_t cols, uint64_t nnz,
                 int vtype, int off_bytes, int on_device);
/* Free an owned matrix and zero the struct; borrowed matrices are only
 * zeroed.  Replaces: cuCSR::~cuCSR / reset (src/CSR.cpp:211-224). */
void oc_csr_free(oc_csr *m);
/* Host -> device copy (allocates dst).
 * Replaces: convert(cuCSR&, const CSR&) (src/CSR.cpp:226-241). */
int oc_csr_upload(const oc_csr *host, oc_csr *dev, void *stream);
/* Device -> host copy (allocates dst).
 * Replaces: convert(CSR&, const cuCSR&) (src/CSR.cpp:243-250). */
int oc_csr_download(const oc_csr *dev, oc_csr *host, void *stream);

/* .csr binary file (80-byte header, State<T>, values, col_ids, row_offsets).
 * Replaces: loadCSR (src/CSR.cpp:119-152); the writer is storeCSR of
 * utils/convert.cpp:296-310 (with zeroed padding).  vtype_expected < 0
 * accepts either value type; otherwise a different typesize is an error
 * (src/CSR.cpp:136-139). */
int oc_csr_load(const char *path, int vtype_expected, oc_csr *host);
int oc_csr_store(const char *path, const oc_csr *host);

/* MatrixMarket coordinate file -> host CSR (value type vtype, columns
 * ascending in each row).  Replaces: loadMTX + convert(CSR&, const COO&)
 * (utils/convert.cpp:83-232): the same accepted headers (real / integer /
 * double / pattern / complex's real part; general / symmetric / Hermitian,
 * off-diagonal entries mirrored), skipped comment and blank lines, the same
 * index checks and messages.  Entries with the same (row, column) keep file
 * order, or are summed when merge_duplicates is nonzero (no counterpart:
 * the reference keeps them in an unspecified order).  A 0 index or more
 * entries than the size line declares is an error. */
int oc_mtx_load(const char *path, int vtype, int merge_duplicates, oc_csr *host);

/* Correctness check of two host matrices.
 * Replaces: compare (src/CSR.cpp:252-335): equal dims and nnz, equal
 * row_offsets[0..rows), equal col_ids, and value relative error
 * |a-b|/max(|a|,|b|) counted above rel_err and fatal above autofail.
 * Returns OC_OK or OC_ERR_MISMATCH. */
typedef struct oc_compare_result {
  uint64_t num_errors;     /* entries above rel_err            */
  uint64_t first_bad;      /* first offending index (or ~0)     */
  double max_rel_err;
  int32_t kind;            /* 0 ok, 1 dims, 2 offsets, 3 cols, 4 values */
} oc_compare_result;
int oc_csr_compare(const oc_csr *a, const oc_csr *b, double rel_err,
                   double autofail_rel_err, oc_compare_result *res);

/* ---------------------------------------------------------------- config */
/* Replaces: struct Config (include/Utils.h:17-108). */
typedef struct oc_config {
  int32_t use_estimation_workflow, use_ultrasparse_workflow, use_fat_workflow,
      use_assist_information, use_hybrid_accumulator, use_indirect_sort;
  int32_t use_segmented_workflow, segmentation_grid_max;
  double segmentation_hot_threshold_k;
  int32_t segmentation_min_segment_rows, segmentation_cc_max_iters;
  int32_t usparse_avg_product_threshold, estimation_input_expansion_threshold,
      estimation_output_compaction_threshold,
      dense_query_bitmap_compaction_threshold,
      numeric_hybrid_hash_kernel_row_length_threshold, fat_ratio;
  double hll_sample_rate;
  int32_t hll_sample_min_rows, hll_sample_max_rows;
  double hll_expansion_low, hll_expansion_high; /* hll_expansion_coes 5,6 */
  double hll_numeric_malloc_expansion_coe;
  int32_t hll_higher_precision_expansion_threshold, hll_precision_low,
      hll_precision_high;
  double estimate_compaction_target_confidence_interval,
      estimate_compaction_target_z_value;
  double symbolic_expansion_coe, numeric_expansion_coe;
  int32_t track_stage_times, check_correctness, output_stats;
  double compare_relative_error, compare_autofail_relative_error;
  char ground_truth_file[512];
  char stats_output_file[512];
  int32_t warmup_iters, bench_iters, print_info, print_memory;
  /* Additions (no reference counterpart): learned row-size predictor. */
  int32_t use_learned_binning;
  char learned_model_file[512];
  /* gfx950 kernel choice for rows beyond the LDS hash/dense kernels:
   * 3 one-gather rank tiles (default; the symbolic pass stores the long
   * rows' bitmaps), 2 rank-compressed tiles, 1 dense tiles, 0 cursor
   * windows. */
  int32_t use_tiled_numeric;
  int32_t tile_log2_cols;  /* tile width 2^k columns (10..16), default 14 */
  int32_t tile_block;      /* workgroup size of the tiled kernel: 256/512/1024 */
  int32_t tile_products_cap; /* rank tiles: products per tile target (0 = none) */
  double learned_quantile; /* predictor: cumulative class probability to cover */
  /* Long rows of the precise workflow swept in column bands of
   * 2^band_log2_cols (10..15; 0 = off, the default) across all rows at once, for rows
   * with at least band_min_products products and a wider span. */
  int32_t band_log2_cols;
  int32_t band_min_products;
  /* Learned predictor: take the estimation workflow when the products are
   * at least this many times the sum of the predicted capacities, else the
   * precise workflow (default 1.5). */
  double learned_min_compression;
  /* Per-bin shapes of the hash/dense kernels (bins 0..6; 0 = the gfx950
   * default).  bin_concurrency: workgroups per CU the bin's tables are
   * sized for (LDS per workgroup = 160 KB / concurrency), default
   * {32,32,16,8,4,2,1}; bin_block: threads per workgroup (64..1024, a power
   * of two), default {64,64,128,256,512,512,1024}.  The hash-table sizes,
   * dense spans and bin thresholds follow from them, as BLOCK_SIZES and
   * SHARED_MEM_PER_BLOCK fix them in the reference (include/Common.h:78-133);
   * the autotuner searches them. */
  int32_t bin_concurrency[7];
  int32_t bin_block[7];
} oc_config;

/* Replaces: Config's default member initialisers (include/Utils.h:22-107). */
void oc_config_default(oc_config *c);
/* Replaces: from_json (src/Utils.cpp:8-37): the same keys (including the
 * `track_stage_time` spelling); unknown keys are ignored. */
int oc_config_from_json(oc_config *c, const char *json_text);
int oc_config_load(oc_config *c, const char *path);
/* Range checks of a configuration (also run by oc_config_from_json and
 * oc_spgemm_create): expansion coefficients >= 1, fat_ratio >= 1,
 * HyperLogLog precisions 5 or 6, sample bounds, and the gfx950 kernel
 * knobs.  The reference applies no checks (src/Utils.cpp:8-37); a value it
 * would accept and then misbehave on is an OC_ERR_INVALID here. */
int oc_config_check(const oc_config *c);

/* ---------------------------------------------------------------- SpGEMM */
/* Replaces: class SpGEMM (kernels/SpGEMM.cuh:209-343):
 *   SpGEMM(const Config&)            -> oc_spgemm_create
 *   shared_ptr<cuCSR> run(A, B)      -> oc_spgemm_run (one product; the
 *                                       warmup/bench loop lives in the CLI)
 *   json get_stats() const           -> oc_spgemm_stats_json
 * A and B are borrowed device matrices (off_bytes 4, same vtype); C is
 * allocated by the library (off_bytes 8) and freed with oc_csr_free. */
typedef struct oc_spgemm oc_spgemm;
oc_spgemm *oc_spgemm_create(const oc_config *cfg);
void oc_spgemm_destroy(oc_spgemm *h);
int oc_spgemm_run(oc_spgemm *h, const oc_csr *dA, const oc_csr *dB,
                  oc_csr *dC, void *stream);
/* Stats of the last run as JSON (keys of SpGEMM.cuh:1788-1890 plus
 * "timing", "total_time", "gflops", "roofline").  Writes at most cap bytes
 * (NUL-terminated); *needed receives the full length + 1. */
int oc_spgemm_stats_json(const oc_spgemm *h, char *buf, size_t cap,
                         size_t *needed);
/* Milliseconds of the last run ("total") or of a stage timer
 * ("analysis_product_calc", "symbolic_hash", "numeric_window", ...; the
 * stage names of Timing, include/Utils.h:175-205); -1 when not recorded.
 * Replaces: Timing / cuTimer (include/Utils.h:120-320). */
double oc_spgemm_timer_ms(const oc_spgemm *h, const char *name);
/* Workflow override for the HLL-sampled path: -1 automatic (the decision
 * of kernels/SpGEMM.cuh:610-637), 0 precise, 1 estimation.  No reference
 * counterpart: lets tests and the learned binning pick a workflow. */
int oc_spgemm_set_workflow(oc_spgemm *h, int workflow);
/* Per-row output estimates (device, A.rows entries) used by the
 * estimation workflow instead of the HLL merge; NULL restores HLL.  No
 * reference counterpart: the learned row-size predictor and forced-overflow
 * tests feed the estimation workflow through it.  Rows whose estimate is
 * too small overflow and are recomputed exactly. */
int oc_spgemm_set_estimates(oc_spgemm *h, const uint32_t *d_est);

/* Value-only replays (no reference counterpart): for repeated products
 * with a fixed structure (batches, iterative methods).  oc_spgemm_plan runs
 * the precise workflow into dC and keeps its numeric plan (row information,
 * numeric bins, scratch) in the handle; dC stays the caller's and must stay
 * allocated while the plan is used.  oc_spgemm_replay recomputes dC's values
 * from the current values of the same A and B buffers; oc_spgemm_capture
 * records that replay into a HIP graph and oc_spgemm_graph_launch launches
 * it (stream NULL: the handle's own stream, waited for). */
int oc_spgemm_plan(oc_spgemm *h, const oc_csr *dA, const oc_csr *dB, oc_csr *dC,
                   void *stream);
int oc_spgemm_replay(oc_spgemm *h, void *stream);
int oc_spgemm_capture(oc_spgemm *h);
int oc_spgemm_graph_launch(oc_spgemm *h, void *stream);
/* The whole precise path of a kept plan, from the analysis to the numeric
 * kernels, recomputing structure and values from the current contents of
 * the planned A and B (same dimensions and entry counts; the structure may
 * change).  No reference counterpart: the reference reads bin sizes and
 * totals back to the host between stages (kernels/SpGEMM.cuh:509-514,
 * 611-615, 809-810, 1578-1580); here every bin kernel runs persistent over
 * its device-side row range, so nothing is read back, allocated or set with
 * a memset node.  oc_spgemm_full_replay runs it eagerly (synchronizing);
 * oc_spgemm_capture_full records it into a HIP graph and
 * oc_spgemm_graph_launch_full launches that graph.  dC keeps the planned
 * capacity (the planned nnz).  After a run, oc_spgemm_full_status gives its
 * check bits: 1 nnz(C) beyond the capacity, 2 an A row longer than the
 * planned longest, 8 a numeric recount; with any bit set, dC's contents
 * are not the product.  Rows whose (estimated-compaction) symbolic table
 * fills up are recounted on the device, inside the same launch. */
int oc_spgemm_full_replay(oc_spgemm *h, void *stream);
int oc_spgemm_capture_full(oc_spgemm *h);
int oc_spgemm_graph_launch_full(oc_spgemm *h, void *stream);
int oc_spgemm_full_status(oc_spgemm *h, uint32_t *status);

/* ------------------------------------------------------ batched SpGEMM */
/* No reference counterpart.  count independent products C_i = A_i * B_i
 * (B_i = A_i when Bs is NULL) run as one product of block-diagonal
 * operands: the device inputs are concatenated (A_i's columns shifted by
 * the earlier A_j.cols, B_i's by the earlier B_j.cols), so every launch of
 * the path covers the rows of all products.  The concatenation needs fewer
 * than 2^32 rows, columns and entries in each operand; A_i and B_i must
 * share one value type and have 32-bit offsets. */
typedef struct oc_batch oc_batch;
oc_batch *oc_batch_create(const oc_config *cfg);
void oc_batch_destroy(oc_batch *b);
/* Copies the inputs into the block-diagonal operands (the inputs may be
 * freed afterwards unless value refreshes via oc_batch_capture are used). */
int oc_batch_set_inputs(oc_batch *b, const oc_csr *dAs, const oc_csr *dBs, uint32_t count,
                        void *stream);
/* Whole path (automatic workflow) on the concatenated operands. */
int oc_batch_run(oc_batch *b, void *stream);
/* Precise run that keeps the numeric plan, then a HIP graph that refreshes
 * the concatenated values from the inputs' current values and recomputes
 * every product (one graph launch per batch). */
int oc_batch_plan(oc_batch *b, void *stream);
int oc_batch_capture(oc_batch *b);
int oc_batch_launch(oc_batch *b, void *stream);
/* The whole path of every product in one graph (see oc_spgemm_capture_full):
 * the inputs' structure and values are concatenated again and every
 * product is recomputed from analysis to numeric; each input may change
 * structure if it keeps its rows, columns and entry count.  Status bits as
 * oc_spgemm_full_status. */
int oc_batch_capture_full(oc_batch *b);
int oc_batch_launch_full(oc_batch *b, void *stream);
int oc_batch_full_status(oc_batch *b, uint32_t *status);
/* The batched result: a view of C (u64 offsets; product i is rows
 * [row_base[i], row_base[i+1]) with columns shifted by col_base[i]) and
 * host copies of the count+1 bases (may be NULL). */
int oc_batch_result(const oc_batch *b, oc_csr *C, uint64_t *row_base, uint64_t *col_base);
/* Product i as its own device CSR (u64 offsets, owned; oc_csr_free). */
int oc_batch_extract(const oc_batch *b, uint32_t i, oc_csr *Ci, void *stream);
/* The SpGEMM handle behind the batch (stats, timers, model). */
oc_spgemm *oc_batch_handle(oc_batch *b);

/* ---------------------------------------------------------- stage calls */
/* gfx950 bin tables (re-derived for 160 KB LDS; see include/Common.h:78-291
 * of the reference for the A100 tables they replace). */
typedef struct oc_tables {
  uint32_t hash_sym[7];    /* HASH_BIN_SIZES          (Common.h:115-123) */
  uint32_t dense_sym[7];   /* DENSE_BIN_SIZES bits    (Common.h:249-257) */
  uint32_t hash_num[7];    /* HASH_NUMERIC_BIN_SIZES  (Common.h:125-133) */
  uint32_t dense_num[7];   /* DENSE_NUMERIC_BIN_SIZES (Common.h:261-269) */
  uint32_t block_size[7];  /* BLOCK_SIZES             (Common.h:79-87)   */
  uint32_t lds_bytes[7];   /* SHARED_MEM_PER_BLOCK    (Common.h:102-110) */
  uint32_t max_dense_sym_bin;
  uint32_t window_cols;    /* column window of the windowed accumulator */
  uint32_t lds_per_cu;
  uint32_t num_cu;
} oc_tables;
int oc_get_tables(int vtype, oc_tables *t);

/* Per-row analysis.  Replaces: analysisKernel + analysisKernelWrapper
 * (kernels/Analysis.cuh:31-105, kernels/Wrappers.cuh:12-54) and the
 * ExclusiveSum/Max of num_products (kernels/SpGEMM.cuh:488-514).
 * Outputs (device, A.rows entries each): num_products, max_b_len, avg_b_len,
 * b_left, b_right; products_psum has A.rows+1 entries (u64); scalars
 * (device, may be NULL): max_a_nnz, max_products. */
int oc_analysis(const oc_csr *dA, const oc_csr *dB, uint32_t *num_products,
                uint32_t *max_b_len, uint32_t *avg_b_len, uint32_t *b_left,
                uint32_t *b_right, uint64_t *products_psum,
                uint32_t *max_a_nnz, uint32_t *max_products, void *stream);

/* oc_analysis without any allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, outputs
 * and errors as oc_analysis, but the temporaries live in a 32 MiB
 * per-device array that the library holds until the process ends (no
 * hipMalloc/hipFree, no stream-ordered allocation, no memset nodes) and the
 * call does not synchronize: results are ready when the stream reaches them.
 * Operands that need more than 32 MiB of temporaries (about 16 B per row of
 * B plus 8 B per 8192 rows of A) are rejected with OC_ERR_INVALID.  Calls
 * that overlap on one device race on its array. */
int oc_analysis_noalloc(const oc_csr *dA, const oc_csr *dB, uint32_t *num_products,
                        uint32_t *max_b_len, uint32_t *avg_b_len, uint32_t *b_left,
                        uint32_t *b_right, uint64_t *products_psum,
                        uint32_t *max_a_nnz, uint32_t *max_products, void *stream);

/* Exclusive prefix sum of n u32 values into n+1 u64 values (out[n] is the
 * total).  Replaces: cub::DeviceScan::ExclusiveSum (kernels/SpGEMM.cuh:402,
 * 488, 978, 1147, 1221, 1441, 1622). */
int oc_scan_u32(const uint32_t *in, uint64_t *out, uint64_t n, void *stream);

/* MurmurHash3_x86_32 of col ids, seed 1234, on the device (n values).
 * Replaces: murmurHash3_32 (kernels/MurmurHash.cuh:44-98). */
int oc_murmur3_cols(const uint32_t *cols, uint32_t *hashes, uint64_t n,
                    void *stream);
/* HLL sketches of B's rows (B.rows * 2^log_p bytes).
 * Replaces: hllConstruct (kernels/HLL.cuh:29-103). */
int oc_hll_build(int log_p, const oc_csr *dB, uint8_t *hll_b, void *stream);
/* Per-A-row estimate (rounded up to the numeric hash sizes of vtype).
 * Replaces: hllMerge (kernels/HLL.cuh:108-236) via est_kernel
 * (kernels/SpGEMM.cuh:691-757). */
int oc_hll_estimate(int log_p, const oc_csr *dA, const uint8_t *hll_b,
                    double expansion, uint32_t *est, void *stream);
/* Sampled rows and their raw estimates.  Replaces: hllSamplingMerge
 * (kernels/HLL.cuh:240-357). */
int oc_hll_sample(int log_p, const oc_csr *dA, const uint8_t *hll_b,
                  uint32_t nsamples, uint32_t *sample_rows,
                  int32_t *sample_est, void *stream);

/* Binning into 14 bins (0-6 hash, 7-13 dense; slot 14 holds rows left
 * out): bin_of_row[rows] (u8) and rows grouped by bin in ascending row
 * order: bin_rows[rows], bin_offsets[16] (device).
 * Replaces: binning (kernels/Analysis.cuh:110-211). */
int oc_bin_symbolic(int vtype, uint64_t rows, const uint32_t *num_products,
                    const uint32_t *b_left, const uint32_t *b_right,
                    int use_compaction, double expand_factor,
                    double avg_compaction, uint8_t *bin_of_row,
                    uint32_t *bin_rows, uint32_t *bin_offsets, void *stream);
/* Replaces: numericBinning (kernels/Analysis.cuh:214-303); `input` is
 * modified for dense-bin-0 rows when modify_dense_est is set. */
int oc_bin_numeric(int vtype, uint64_t rows, uint32_t *input,
                   const uint32_t *b_left, const uint32_t *b_right,
                   int expand, double expand_factor, int modify_dense_est,
                   int use_largest, uint32_t hybrid_max, uint8_t *bin_of_row,
                   uint32_t *bin_rows, uint32_t *bin_offsets, void *stream);

/* Exact nnz of every row of C = A*B through the binned symbolic kernels
 * (LDS hash sets, LDS bitmaps, windowed bitmaps).  Replaces:
 * symbolic_kernel (kernels/SpGEMM.cuh:759-923) with hashSymbolicLauncher /
 * denseSymbolicLauncher (kernels/Wrappers.cuh:56-280). */
int oc_symbolic(const oc_config *cfg, const oc_csr *dA, const oc_csr *dB,
                uint32_t *nnz_row, void *stream);

/* Per-row ascending sort of (col, value) segments: rows listed in
 * `rows` (n entries, NULL = all rows) of a device matrix with off_bytes
 * offsets are sorted in place with an LDS radix sort (global ping-pong for
 * rows beyond LDS).  Replaces: sortOutputDyn / sortOutputFused
 * (kernels/Epilogue.cuh:67-222) and outputSortingDynLauncher
 * (kernels/Wrappers.cuh:853-1068). */
int oc_sort_rows(oc_csr *dM, const uint32_t *rows, uint64_t n, void *stream);

/* --------------------------------------------------- matrix generators */
/* Host matrices with seeded values (no reference counterpart: synthetic
 * inputs of the benchmark configurations).  vtype picks f64/f32 values;
 * value_dist 0 = uniform [-1,1), 1 = N(0,1). */
int oc_gen_laplacian2d(uint32_t nx, uint32_t ny, int vtype, uint64_t seed,
                       oc_csr *host);
int oc_gen_rmat(uint32_t log_n, uint64_t edges, double a, double b, double c,
                int vtype, uint64_t seed, oc_csr *host);
int oc_gen_erdos_renyi(uint32_t n, double avg_degree, int vtype,
                       uint64_t seed, oc_csr *host);
int oc_gen_banded(uint32_t n, uint32_t bandwidth, int vtype, uint64_t seed,
                  oc_csr *host);
int oc_gen_rect_band(uint32_t rows, uint32_t cols, double mean_per_row,
                     double band_frac, uint32_t half_band, uint32_t stride,
                     int vtype, int value_dist, uint64_t seed, oc_csr *host);
/* Host transpose (B = A^T for the rectangular configuration). */
int oc_csr_transpose(const oc_csr *host_in, oc_csr *host_out);

/* ------------------------------------------------- device validation */
/* No reference counterpart: structural checks of a device CSR used by the
 * full-size parity tests.  bad_rows (host out) counts rows whose columns
 * are not strictly ascending or not < cols. */
int oc_csr_validate(const oc_csr *dM, uint64_t *bad_rows, void *stream);
/* Order-free per-row column-set hash sum_c mix64(c) (device out, rows). */
int oc_csr_row_set_hash(const oc_csr *dM, uint64_t *hash, void *stream);
/* y = M x in FP64 (x, y device, cols / rows entries). */
int oc_csr_spmv(const oc_csr *dM, const double *x, double *y, void *stream);

/* ------------------------------------------ learned row-size predictor */
/* No reference counterpart.  A 16 -> 64 -> 64 -> 8 MLP (BF16 weights on
 * MFMA, FP32 accumulation) that predicts the class of each row's output
 * count of C = A*B; classes end at the numeric hash capacities, 65536 and
 * unbounded (oc_mlp_class_thresholds).  Parameters (OC_MLP_NPARAM floats)
 * are laid out W1[16][64], b1[64], W2[64][64], b2[64], W3[64][8], b3[8].
 * The predicted capacities replace the HyperLogLog estimates of the
 * estimation workflow (kernels/SpGEMM.cuh:691-757); rows that exceed them
 * overflow and are recomputed exactly. */
#define OC_MLP_FEATURES 16
#define OC_MLP_HIDDEN 64
#define OC_MLP_CLASSES 8
#define OC_MLP_NPARAM 5768
typedef struct oc_mlp oc_mlp;
/* He-normal weights from a seed, zero biases, Adam state cleared. */
oc_mlp *oc_mlp_create(uint64_t seed);
void oc_mlp_destroy(oc_mlp *m);
int oc_mlp_get_params(const oc_mlp *m, float *host_params);
int oc_mlp_set_params(oc_mlp *m, const float *host_params);
/* Model file: 8-byte magic "OCMLP1", u32 parameter count, FP32 params. */
int oc_mlp_save(const oc_mlp *m, const char *path);
int oc_mlp_load(oc_mlp *m, const char *path);
/* Class upper bounds (OC_MLP_CLASSES entries, the last 0xFFFFFFFF). */
int oc_mlp_class_thresholds(uint32_t *thr);
/* Features (device out, A.rows x 16 floats) of the rows of C = A*B. */
int oc_mlp_features(const oc_csr *dA, const oc_csr *dB, float *feat, void *stream);
/* Class labels (device u8) of exact per-row counts (device u32). */
int oc_mlp_labels(const uint32_t *nnz, uint64_t n, uint8_t *labels, void *stream);
/* Logits (device out, n x 8 floats); feat 16-byte aligned. */
int oc_mlp_forward(const oc_mlp *m, const float *feat, uint64_t n, float *logits,
                   void *stream);
/* Gradients of the mean cross-entropy of one minibatch (n <= 65536) into
 * the model's gradient buffer (read with oc_mlp_get_grads). */
int oc_mlp_backward(oc_mlp *m, const float *feat, const uint8_t *labels, uint64_t n,
                    void *stream);
int oc_mlp_get_grads(const oc_mlp *m, float *host_grads, float *host_loss);
/* One Adam step (beta 0.9/0.999, eps 1e-8) with the last gradients. */
int oc_mlp_step(oc_mlp *m, float lr, void *stream);
/* Minibatch training over device data in order; mean loss per epoch
 * (host out, may be NULL). */
int oc_mlp_train(oc_mlp *m, const float *feat, const uint8_t *labels, uint64_t n,
                 uint32_t batch, uint32_t epochs, float lr, float *epoch_loss,
                 void *stream);
/* Per-row capacities (device out, A.rows) for C = A*B: the smallest class
 * whose cumulative probability reaches `quantile`, raised to the longest
 * referenced B row and clipped to min(products, column span). */
int oc_mlp_estimates(const oc_mlp *m, const oc_csr *dA, const oc_csr *dB,
                     double quantile, uint32_t *est, void *stream);
/* Use a trained predictor in oc_spgemm_run (NULL: back to HyperLogLog):
 * on the HyperLogLog-sampled path the estimation workflow then runs on the
 * predicted capacities.  The model must outlive the handle's runs. */
int oc_spgemm_set_model(oc_spgemm *h, const oc_mlp *m, double quantile);

/* ------------------------------------------------------ device memory */
/* Thin HIP helpers for bindings without their own device allocator (no
 * reference counterpart; the reference calls cudaMalloc/cudaMemcpy
 * directly, e.g. src/CSR.cpp:185-250). */
int oc_device_count(int *n);
int oc_device_malloc(void **p, size_t bytes);
int oc_device_free(void *p);
int oc_memcpy_htod(void *dst, const void *src, size_t bytes);
int oc_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int oc_memcpy_dtod(void *dst, const void *src, size_t bytes);
int oc_memset_device(void *p, int value, size_t bytes);
int oc_device_synchronize(void);
/* Free and total bytes of the current device (hipMemGetInfo; the
 * reference's printMemInfo, src/Utils.cpp:59-66, reads cudaMemGetInfo). */
int oc_device_mem_info(size_t *free_bytes, size_t *total_bytes);

#ifdef __cplusplus
}
#endif
#endif /* OCEAN_CAPI_H */
