/*
 * CPU oracle for the Ocean SpGEMM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (research_amd/, the
 * spgemm CLI, the C-ABI library) links, loads or calls this code.  Only the
 * tests under tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py use it, and only as the checker / CPU baseline.
 *
 * Each function restates one piece of the reference algorithm and cites the
 * reference file:line it follows.  How each function is pinned (known-answer
 * vectors, closed forms, brute force, library routines) is stated beside it.
 */
#ifndef OCEAN_ORACLE_H
#define OCEAN_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- MurmurHash3_x86_32 ------------------------------------------------
 * General form of the public algorithm; the reference hashes the 4 LE bytes
 * of a column id with seed 1234 (kernels/MurmurHash.cuh:44-98,
 * include/Common.h:73).  Pinned by the published MurmurHash3_x86_32 known
 * answers (tests/test_oracle_hll.py). */
uint32_t ora_murmur3_x86_32(const void *data, size_t len, uint32_t seed);
uint32_t ora_murmur3_col(uint32_t col); /* 4-byte key, seed 1234 */

/* ---- HyperLogLog ------------------------------------------------------- */
/* Linear-counting table: lc[v] = (int)(P * logf(P / (float)v)), v = 1..P
 * (kernels/HLL.cuh:321-323).  lc[0] is unused. */
void ora_hll_lc_table(int log_p, int32_t *lc /* P+1 */);
/* Sketch per B row: reg[h & (P-1)] = max(min(clz(h)+1, 33-logP))
 * (kernels/HLL.cuh:12-24, 29-103). out is rows*P bytes. */
void ora_hll_construct(int log_p, uint64_t rows, const uint32_t *rowptr,
                       const uint32_t *cols, uint8_t *out);
/* Raw HLL estimate of a merged register set (kernels/HLL.cuh:277-323). */
int32_t ora_hll_estimate(int log_p, const uint8_t *regs, const int32_t *lc);
/* Per-A-row estimate: byte-max merge of the referenced B sketches, raw
 * estimate, times expansion, rounded up to the first numeric hash table size
 * that holds it (kernels/HLL.cuh:108-236).  num_sizes has 7 entries. */
void ora_hll_merge(int log_p, uint64_t rows, const uint32_t *a_rowptr,
                   const uint32_t *a_cols, const uint8_t *hll_b,
                   double expansion, const uint32_t *num_sizes,
                   uint32_t *est);
/* LCG row choice of sample i (kernels/HLL.cuh:355-361). */
uint32_t ora_sample_row(uint32_t i, uint64_t rows);
/* Sampling pass: per sample its row and raw estimate
 * (kernels/HLL.cuh:240-357). */
void ora_hll_sample(int log_p, uint64_t rows, const uint32_t *a_rowptr,
                    const uint32_t *a_cols, const uint8_t *hll_b,
                    uint32_t nsamples, uint32_t *sample_row,
                    int32_t *sample_est);

/* ---- Row analysis (kernels/Analysis.cuh:31-105) ------------------------ */
void ora_analysis(uint64_t a_rows, const uint32_t *a_rowptr,
                  const uint32_t *a_cols, uint64_t b_rows,
                  const uint32_t *b_rowptr, const uint32_t *b_cols,
                  uint32_t *num_products, uint32_t *max_b_len,
                  uint32_t *avg_b_len, uint32_t *b_left, uint32_t *b_right,
                  uint32_t *max_a_nnz);

/* ---- Binning ------------------------------------------------------------
 * Symbolic: kernels/Analysis.cuh:110-211.  bin_out[row] in 0..13.
 * max_dense_bin is the largest dense symbolic kernel (5 in the reference,
 * whose rule sends hash bin 6 to dense bin 5; 6 where a dense bin-6 kernel
 * exists). */
void ora_bin_symbolic(uint64_t rows, const uint32_t *products,
                      const uint32_t *b_left, const uint32_t *b_right,
                      const uint32_t *hash_sizes, const uint32_t *dense_sizes,
                      int max_dense_bin, int use_compaction,
                      double expand_factor, double avg_compaction,
                      uint8_t *bin_out);
/* Numeric: kernels/Analysis.cuh:214-303.  `input` is modified in place for
 * dense-bin-0 rows when modify_dense_est is set (Analysis.cuh:250-252).  The
 * reference caps hybrid (bin 6) rows with a racy counter; here the first
 * hybrid_max candidate rows in row order get bin 6. */
void ora_bin_numeric(uint64_t rows, uint32_t *input, const uint32_t *b_left,
                     const uint32_t *b_right, const uint32_t *hash_sizes,
                     const uint32_t *dense_sizes, int expand,
                     double expand_factor, int modify_dense_est,
                     int use_largest, uint32_t hybrid_max, uint8_t *bin_out);

/* ---- Workflow decisions -------------------------------------------------
 * kernels/SpGEMM.cuh:445-547 (analysis), 549-689 (sampling),
 * kernels/Analysis.cuh:10-29 (decideMatrixType).  Inputs are the scalars the
 * reference reads back; outputs mirror its Workspace fields. */
typedef struct ora_decision {
  int ana1_type;            /* 0 ultrasparse, 1 hll, 2 fat */
  int spark_kernel_size;
  int spark_need_outlier;
  int spark_outlier_kernel_size;
  int hll_log2_precision;
  double hll_expansion_factor;
  int num_sample_rows;
  double compaction_1, compaction_2; /* ana2_avg_sampled_compaction(2) */
  int ana2_type;            /* 0 precise, 1 estimation */
  int dense_query_bitmap;
  double compaction_avg, compaction_var, safe_compaction;
  int use_estimated_symbolic;
} ora_decision;

typedef struct ora_decide_cfg {
  int use_estimation_workflow, use_ultrasparse_workflow, use_fat_workflow,
      use_assist_information;
  int usparse_avg_product_threshold, estimation_input_expansion_threshold,
      estimation_output_compaction_threshold,
      dense_query_bitmap_compaction_threshold, fat_ratio,
      hll_higher_precision_expansion_threshold, hll_precision_low,
      hll_precision_high, hll_sample_min_rows, hll_sample_max_rows;
  double hll_sample_rate, hll_expansion_low, hll_expansion_high,
      estimate_compaction_target_z_value;
  uint32_t fat_dense_limit; /* DENSE_NUMERIC_BIN_SIZES[5] */
} ora_decide_cfg;

void ora_decide_cfg_default(ora_decide_cfg *c);
/* First half: SpGEMM.cuh:455-540. */
void ora_decide_analysis(const ora_decide_cfg *c, uint64_t a_rows,
                         uint64_t a_cols, uint64_t a_nnz, uint64_t b_rows,
                         uint64_t b_cols, uint64_t total_products,
                         uint32_t max_products, ora_decision *d);
/* Number of sampled rows (SpGEMM.cuh:587-590). */
int ora_num_samples(const ora_decide_cfg *c, uint64_t a_rows);
/* Second half, given the per-sample (row products, raw estimate) pairs, in
 * sample order (SpGEMM.cuh:610-686; Analysis.cuh:10-29). */
void ora_decide_sampling(const ora_decide_cfg *c, uint64_t a_nnz,
                         uint64_t total_products, int nsamples,
                         const uint32_t *sample_products,
                         const int32_t *sample_est, ora_decision *d);

/* ---- Gustavson row-wise SpGEMM (FP64) -----------------------------------
 * Output structure is the full symbolic pattern: every column reached by an
 * intermediate product is kept, including exact cancellations
 * (kernels/AccumulatorDense.cuh:208-219, AccumulatorHash.cuh:462-466), with
 * ascending columns per row.  vtype 0 = double, 1 = float inputs.
 *
 * ora_symbolic: exact nnz of rows [r0, r1) and an order-free set hash
 * sum_{c in row} ora_mix64(c) (mod 2^64).  OpenMP over rows.
 */
uint64_t ora_mix64(uint64_t x);
void ora_symbolic(uint64_t r0, uint64_t r1, const uint32_t *a_rowptr,
                  const uint32_t *a_cols, uint64_t b_cols_n,
                  const uint32_t *b_rowptr, const uint32_t *b_cols,
                  uint32_t *nnz_row, uint64_t *set_hash, int nthreads);
/* ora_numeric: rows [r0, r1) into caller buffers sized by ora_symbolic.
 * out_rowptr has (r1-r0+1) entries starting at 0.  vals are accumulated in
 * product order (A entries in order, then B entries in order) starting from
 * 0.0 (plain, the ESC order of AccumulatorESC.cuh:36-46) when compensated=0,
 * or with Neumaier compensation when compensated=1.  cond (may be NULL)
 * receives sum |a_ik * b_kj| per output entry.  Dense SPA + touched list +
 * per-row sort, OpenMP dynamic over rows; also the CPU baseline. */
void ora_numeric(uint64_t r0, uint64_t r1, int vtype,
                 const uint32_t *a_rowptr, const uint32_t *a_cols,
                 const void *a_vals, uint64_t b_cols_n,
                 const uint32_t *b_rowptr, const uint32_t *b_cols,
                 const void *b_vals, uint64_t *out_rowptr, uint32_t *out_cols,
                 double *out_vals, double *cond, int compensated,
                 int nthreads);
/* Rows listed in `rows` (any order, may repeat); out_rowptr relative. */
void ora_numeric_rows(uint64_t nrows, const uint32_t *rows, int vtype,
                      const uint32_t *a_rowptr, const uint32_t *a_cols,
                      const void *a_vals, uint64_t b_cols_n,
                      const uint32_t *b_rowptr, const uint32_t *b_cols,
                      const void *b_vals, uint64_t *out_rowptr,
                      uint32_t *out_cols, double *out_vals, double *cond,
                      int nthreads);
uint64_t ora_symbolic_rows(uint64_t nrows, const uint32_t *rows,
                           const uint32_t *a_rowptr, const uint32_t *a_cols,
                           uint64_t b_cols_n, const uint32_t *b_rowptr,
                           const uint32_t *b_cols, uint32_t *nnz_row,
                           int nthreads);

/* CPU baseline (not a checker): plain row-wise Gustavson, dense FP64
 * accumulator + touched list, rows in column order, per-thread output
 * buffers.  Returns nnz of rows [r0, r1); *checksum = sum of values. */
uint64_t ora_baseline_spgemm(uint64_t r0, uint64_t r1, int vtype,
                             const uint32_t *a_rowptr, const uint32_t *a_cols,
                             const void *a_vals, uint64_t b_cols_n,
                             const uint32_t *b_rowptr, const uint32_t *b_cols,
                             const void *b_vals, double *checksum, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
