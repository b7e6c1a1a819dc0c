/*
 * CPU oracle for the Ocean SpGEMM hot path -- TEST INFRASTRUCTURE ONLY.
 * See oracle.h for the contract and the reference citations.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg use this code.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* MurmurHash3_x86_32 (public algorithm; kernels/MurmurHash.cuh:15-98).     */
/* ------------------------------------------------------------------------ */
static uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

static uint32_t fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

uint32_t ora_murmur3_x86_32(const void *data_, size_t len, uint32_t seed) {
  const uint8_t *data = (const uint8_t *)data_;
  const size_t nblocks = len / 4;
  const uint32_t c1 = 0xcc9e2d51u, c2 = 0x1b873593u;
  uint32_t h1 = seed;
  for (size_t i = 0; i < nblocks; ++i) {
    uint32_t k1 = (uint32_t)data[4 * i] | ((uint32_t)data[4 * i + 1] << 8) |
                  ((uint32_t)data[4 * i + 2] << 16) |
                  ((uint32_t)data[4 * i + 3] << 24);
    k1 *= c1;
    k1 = rotl32(k1, 15);
    k1 *= c2;
    h1 ^= k1;
    h1 = rotl32(h1, 13);
    h1 = h1 * 5 + 0xe6546b64u;
  }
  const uint8_t *tail = data + nblocks * 4;
  uint32_t k1 = 0;
  switch (len & 3) {
  case 3:
    k1 ^= (uint32_t)tail[2] << 16; /* fall through */
  case 2:
    k1 ^= (uint32_t)tail[1] << 8; /* fall through */
  case 1:
    k1 ^= tail[0];
    k1 *= c1;
    k1 = rotl32(k1, 15);
    k1 *= c2;
    h1 ^= k1;
  }
  h1 ^= (uint32_t)len;
  return fmix32(h1);
}

uint32_t ora_murmur3_col(uint32_t col) {
  uint8_t b[4] = {(uint8_t)col, (uint8_t)(col >> 8), (uint8_t)(col >> 16),
                  (uint8_t)(col >> 24)};
  return ora_murmur3_x86_32(b, 4, 1234u); /* MURMUR3_SEED, Common.h:73 */
}

/* ------------------------------------------------------------------------ */
/* HyperLogLog                                                               */
/* ------------------------------------------------------------------------ */
static double hll_alpha(int log_p) {
  /* HLL_CONSTANT, include/Common.h:51 */
  static const double a[8] = {0.0, 0.0, 0.0, 0.0, 0.673, 0.697, 0.709, 0.715};
  return a[log_p];
}

void ora_hll_lc_table(int log_p, int32_t *lc) {
  const int p = 1 << log_p;
  lc[0] = 0;
  for (int v = 1; v <= p; ++v) {
    /* precision * log(precision / (float)empty_reg), HLL.cuh:322 */
    float q = (float)p / (float)v;
    float l = logf(q);
    lc[v] = (int32_t)((float)p * l);
  }
}

static uint32_t clz32(uint32_t x) { return x ? (uint32_t)__builtin_clz(x) : 32u; }

void ora_hll_construct(int log_p, uint64_t rows, const uint32_t *rowptr,
                       const uint32_t *cols, uint8_t *out) {
  const int p = 1 << log_p;
  const uint32_t cap = (uint32_t)(32 - log_p) + 1; /* HLL.cuh:119-121 */
  for (uint64_t r = 0; r < rows; ++r) {
    uint8_t *reg = out + r * (uint64_t)p;
    memset(reg, 0, (size_t)p);
    for (uint32_t j = rowptr[r]; j < rowptr[r + 1]; ++j) {
      uint32_t h = ora_murmur3_col(cols[j]);
      uint32_t idx = h & (uint32_t)(p - 1);
      uint32_t item = clz32(h) + 1;
      if (item > cap) item = cap;
      if (item > reg[idx]) reg[idx] = (uint8_t)item;
    }
  }
}

int32_t ora_hll_estimate(int log_p, const uint8_t *regs, const int32_t *lc) {
  const int p = 1 << log_p;
  double z = 0.0;
  int empty = 0;
  /* HLL.cuh:278-294: register sum of 2^-reg and count of zero registers.
   * Every term is a dyadic rational with at most 34 significant bits, so
   * the sum is exact in any order. */
  for (int i = 0; i < p; ++i) {
    z += 1.0 / (double)(1ULL << regs[i]);
    empty += regs[i] == 0;
  }
  int32_t result = (int32_t)(hll_alpha(log_p) * p * p / z); /* HLL.cuh:320 */
  if (empty != 0 && result < 2.5 * p) result = lc[empty];    /* :321-323 */
  return result;
}

void ora_hll_merge(int log_p, uint64_t rows, const uint32_t *a_rowptr,
                   const uint32_t *a_cols, const uint8_t *hll_b,
                   double expansion, const uint32_t *num_sizes,
                   uint32_t *est) {
  const int p = 1 << log_p;
  int32_t lc[65];
  ora_hll_lc_table(log_p, lc);
  uint8_t merged[64];
  for (uint64_t r = 0; r < rows; ++r) {
    memset(merged, 0, sizeof(merged));
    for (uint32_t j = a_rowptr[r]; j < a_rowptr[r + 1]; ++j) {
      const uint8_t *s = hll_b + (uint64_t)a_cols[j] * (uint64_t)p;
      for (int i = 0; i < p; ++i)
        if (s[i] > merged[i]) merged[i] = s[i];
    }
    int32_t result = ora_hll_estimate(log_p, merged, lc);
    uint32_t temp = (uint32_t)(result * expansion); /* HLL.cuh:325 */
    for (int b = 0; b < 6; ++b) {                   /* HLL.cuh:326-331 */
      if (temp <= num_sizes[b]) {
        temp = num_sizes[b];
        break;
      }
    }
    est[r] = temp;
  }
}

uint32_t ora_sample_row(uint32_t i, uint64_t rows) {
  /* HLL.cuh:356-361: x = 112233 + i*2654435761; x = (1664525x+1013904223)
   * mod 2^32; row = x mod rows. */
  uint64_t x = 112233ULL + (uint64_t)i * 2654435761ULL;
  x = (1664525ULL * x + 1013904223ULL) & 0xFFFFFFFFULL;
  return (uint32_t)(x % rows);
}

void ora_hll_sample(int log_p, uint64_t rows, const uint32_t *a_rowptr,
                    const uint32_t *a_cols, const uint8_t *hll_b,
                    uint32_t nsamples, uint32_t *sample_row,
                    int32_t *sample_est) {
  const int p = 1 << log_p;
  int32_t lc[65];
  ora_hll_lc_table(log_p, lc);
  uint8_t merged[64];
  for (uint32_t i = 0; i < nsamples; ++i) {
    uint32_t r = ora_sample_row(i, rows);
    memset(merged, 0, sizeof(merged));
    for (uint32_t j = a_rowptr[r]; j < a_rowptr[r + 1]; ++j) {
      const uint8_t *s = hll_b + (uint64_t)a_cols[j] * (uint64_t)p;
      for (int k = 0; k < p; ++k)
        if (s[k] > merged[k]) merged[k] = s[k];
    }
    sample_row[i] = r;
    sample_est[i] = ora_hll_estimate(log_p, merged, lc);
  }
}

/* ------------------------------------------------------------------------ */
/* Row analysis (kernels/Analysis.cuh:31-105)                                */
/* ------------------------------------------------------------------------ */
void ora_analysis(uint64_t a_rows, const uint32_t *a_rowptr,
                  const uint32_t *a_cols, uint64_t b_rows,
                  const uint32_t *b_rowptr, const uint32_t *b_cols,
                  uint32_t *num_products, uint32_t *max_b_len,
                  uint32_t *avg_b_len, uint32_t *b_left, uint32_t *b_right,
                  uint32_t *max_a_nnz) {
  uint32_t mx = 0;
  (void)b_rows;
  for (uint64_t r = 0; r < a_rows; ++r) {
    uint32_t s = a_rowptr[r], e = a_rowptr[r + 1];
    uint32_t nnz = e - s;
    if (nnz > mx) mx = nnz;
    uint32_t prod = 0, mb = 0, left = 0xFFFFFFFFu, right = 0;
    for (uint32_t j = s; j < e; ++j) {
      uint32_t k = a_cols[j];
      uint32_t bs = b_rowptr[k], be = b_rowptr[k + 1];
      if (be <= bs) continue; /* empty B row, Analysis.cuh:73 */
      uint32_t cnt = be - bs;
      if (cnt > mb) mb = cnt;
      if (b_cols[bs] < left) left = b_cols[bs];
      if (b_cols[be - 1] > right) right = b_cols[be - 1];
      prod += cnt;
    }
    num_products[r] = prod;
    max_b_len[r] = mb;
    avg_b_len[r] = nnz == 0 ? 0 : prod / nnz; /* Analysis.cuh:94 */
    b_left[r] = left;
    b_right[r] = right;
  }
  *max_a_nnz = mx;
}

/* ------------------------------------------------------------------------ */
/* Binning                                                                   */
/* ------------------------------------------------------------------------ */
/* Expanded sizes saturate at 2^32-1 (the reference's cast of an
 * out-of-range value is undefined; the GPU binning saturates the same way). */
static uint32_t sat_u32(double x) { return x >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)x; }

void ora_bin_symbolic(uint64_t rows, const uint32_t *products,
                      const uint32_t *b_left, const uint32_t *b_right,
                      const uint32_t *hash_sizes, const uint32_t *dense_sizes,
                      int max_dense_bin, int use_compaction,
                      double expand_factor, double avg_compaction,
                      uint8_t *bin_out) {
  for (uint64_t r = 0; r < rows; ++r) {
    uint32_t range = b_right[r] - b_left[r] + 1u; /* u32 wrap, :140 */
    int dense_bin = 7;                            /* BIN_NUM */
    for (int b = 0; b <= max_dense_bin; ++b) {
      if (range <= dense_sizes[b]) {
        dense_bin = b;
        break;
      }
    }
    uint32_t val = products[r];
    if (use_compaction) val = (uint32_t)(val / avg_compaction); /* :150-152 */
    val = sat_u32((double)val * expand_factor);                 /* :153-155 */
    int bin = 6;
    for (int b = 0; b < 6; ++b) {
      if (val <= hash_sizes[b]) {
        bin = b;
        break;
      }
    }
    /* speck rule (:166-173): dense when the dense kernel needs no larger bin
     * than the hash kernel would; the workload-sized dense kernel is used,
     * capped at the largest dense kernel. */
    if (dense_bin <= bin) bin = (bin < max_dense_bin ? bin : max_dense_bin) + 7;
    bin_out[r] = (uint8_t)bin;
  }
}

static int numeric_bin_one(uint32_t val, int dense_bin, int max_bin,
                           const uint32_t *hash_sizes) {
  int bin = 7;
  for (int b = 1; b < max_bin; ++b) { /* Analysis.cuh:265-270 */
    if (val <= hash_sizes[b]) {
      bin = b;
      break;
    }
  }
  if (dense_bin <= bin && !(bin == 6 && dense_bin == 6)) bin = dense_bin + 7;
  return bin;
}

void ora_bin_numeric(uint64_t rows, uint32_t *input, const uint32_t *b_left,
                     const uint32_t *b_right, const uint32_t *hash_sizes,
                     const uint32_t *dense_sizes, int expand,
                     double expand_factor, int modify_dense_est,
                     int use_largest, uint32_t hybrid_max, uint8_t *bin_out) {
  uint32_t hybrid_cnt = 0;
  for (uint64_t r = 0; r < rows; ++r) {
    uint32_t range = b_right[r] - b_left[r] + 1u;
    int dense_bin = 6;
    for (int b = 0; b < 6; ++b) { /* Analysis.cuh:247-255 */
      if (range <= dense_sizes[b]) {
        dense_bin = b;
        if (b == 0 && modify_dense_est) input[r] = dense_sizes[0];
        break;
      }
    }
    uint32_t val = input[r];
    if (expand) val = sat_u32((double)val * expand_factor);
    int bin = numeric_bin_one(val, dense_bin, 6, hash_sizes);
    if (use_largest) {
      int cand = numeric_bin_one(val, dense_bin, 7, hash_sizes);
      if (cand == 6) {
        if (hybrid_cnt < hybrid_max) bin = 6;
        ++hybrid_cnt;
      }
    }
    bin_out[r] = (uint8_t)bin;
  }
}

/* ------------------------------------------------------------------------ */
/* Workflow decisions                                                        */
/* ------------------------------------------------------------------------ */
void ora_decide_cfg_default(ora_decide_cfg *c) {
  /* Config defaults, include/Utils.h:22-88 */
  c->use_estimation_workflow = 1;
  c->use_ultrasparse_workflow = 1;
  c->use_fat_workflow = 1;
  c->use_assist_information = 1;
  c->usparse_avg_product_threshold = 64;
  c->estimation_input_expansion_threshold = 8;
  c->estimation_output_compaction_threshold = 8;
  c->dense_query_bitmap_compaction_threshold = 2;
  c->fat_ratio = 100;
  c->hll_higher_precision_expansion_threshold = 48;
  c->hll_precision_low = 5;
  c->hll_precision_high = 6;
  c->hll_sample_min_rows = 648;
  c->hll_sample_max_rows = 10800;
  c->hll_sample_rate = 0.03;
  c->hll_expansion_low = 2.0;
  c->hll_expansion_high = 1.5;
  c->estimate_compaction_target_z_value = 2.53;
  c->fat_dense_limit = 7277;
}

void ora_decide_analysis(const ora_decide_cfg *c, uint64_t a_rows,
                         uint64_t a_cols, uint64_t a_nnz, uint64_t b_rows,
                         uint64_t b_cols, uint64_t total_products,
                         uint32_t max_products, ora_decision *d) {
  memset(d, 0, sizeof(*d));
  /* FAT shortcut, SpGEMM.cuh:456-465 (integer divisions) */
  if (a_cols / (uint64_t)c->fat_ratio > a_rows &&
      b_rows / (uint64_t)c->fat_ratio > b_cols && b_cols < c->fat_dense_limit &&
      c->use_fat_workflow) {
    d->ana1_type = 2;
    return;
  }
  double avg_product = (double)total_products / (double)a_rows; /* :518 */
  if (avg_product <= (double)c->usparse_avg_product_threshold &&
      c->use_ultrasparse_workflow) { /* :521-530 */
    static const int sizes[4] = {16, 32, 64, 128};
    d->ana1_type = 0;
    for (int i = 0; i < 4; ++i)
      if (sizes[i] >= avg_product) {
        d->spark_kernel_size = sizes[i];
        break;
      }
    if ((int64_t)max_products > d->spark_kernel_size) {
      d->spark_need_outlier = 1;
      d->spark_outlier_kernel_size = (int)(((max_products) + 31) / 32 * 32);
    }
  } else { /* :531-540 */
    uint64_t ratio = a_nnz ? total_products / a_nnz : 0;
    d->ana1_type = 1;
    if (ratio < (uint64_t)c->hll_higher_precision_expansion_threshold) {
      d->hll_log2_precision = c->hll_precision_low;
      d->hll_expansion_factor = c->hll_expansion_low;
    } else {
      d->hll_log2_precision = c->hll_precision_high;
      d->hll_expansion_factor = c->hll_expansion_high;
    }
  }
}

int ora_num_samples(const ora_decide_cfg *c, uint64_t a_rows) {
  int n = (int)(a_rows * c->hll_sample_rate); /* SpGEMM.cuh:587-590 */
  if (n < c->hll_sample_min_rows) n = c->hll_sample_min_rows;
  if (n > c->hll_sample_max_rows) n = c->hll_sample_max_rows;
  return n;
}

void ora_decide_sampling(const ora_decide_cfg *c, uint64_t a_nnz,
                         uint64_t total_products, int nsamples,
                         const uint32_t *sample_products,
                         const int32_t *sample_est, ora_decision *d) {
  /* Device sums of HLL.cuh:450-453, taken here in sample order. */
  uint64_t sum_est = 0, sum_prod = 0;
  double sum_comp = 0.0;
  for (int i = 0; i < nsamples; ++i) {
    sum_est += (uint64_t)(int64_t)sample_est[i];
    sum_prod += sample_products[i];
    sum_comp += (double)sample_products[i] / (double)sample_est[i];
  }
  d->num_sample_rows = nsamples;
  d->compaction_1 = (double)sum_prod / (double)sum_est; /* SpGEMM.cuh:618 */
  d->compaction_2 = sum_comp / nsamples;                /* :619 */
  /* decideMatrixType (Analysis.cuh:10-29).  Its parameters are index_t, so
   * total_products and A.nnz arrive truncated to 32 bits. */
  uint32_t tp32 = (uint32_t)total_products, an32 = (uint32_t)a_nnz;
  int type = 1;
  if (an32 == 0 || tp32 / an32 < (uint32_t)c->estimation_input_expansion_threshold)
    type = 0;
  else if (d->compaction_1 < c->estimation_output_compaction_threshold)
    type = 0;
  else if (d->compaction_2 < c->estimation_output_compaction_threshold)
    type = 0;
  if (!c->use_estimation_workflow) type = 0; /* :635-637 */
  d->ana2_type = type;
  /* :641-648 (NaN compares false, so a NaN compaction keeps the bitmap) */
  if ((d->compaction_1 < c->dense_query_bitmap_compaction_threshold &&
       d->compaction_2 < c->dense_query_bitmap_compaction_threshold) ||
      !c->use_assist_information)
    d->dense_query_bitmap = 0;
  else
    d->dense_query_bitmap = 1;
  if (type == 1) return;
  /* Host variance, SpGEMM.cuh:657-685 */
  double s = 0.0;
  for (int i = 0; i < nsamples; ++i)
    s += (double)sample_products[i] / (double)sample_est[i];
  d->compaction_avg = s / nsamples;
  double v = 0.0;
  for (int i = 0; i < nsamples; ++i) {
    double x = (double)sample_products[i] / (double)sample_est[i] - d->compaction_avg;
    v += x * x;
  }
  d->compaction_var = v / (nsamples - 1);
  d->safe_compaction =
      (d->compaction_avg - c->estimate_compaction_target_z_value * sqrt(d->compaction_var)) / 2;
  if (d->safe_compaction > d->compaction_avg * 0.8)
    d->safe_compaction = d->compaction_avg * 0.8;
  d->use_estimated_symbolic =
      d->safe_compaction >= 1.5 && c->use_assist_information;
}

/* ------------------------------------------------------------------------ */
/* Gustavson                                                                 */
/* ------------------------------------------------------------------------ */
uint64_t ora_mix64(uint64_t x) {
  /* splitmix64 finaliser of (x + 1) */
  x += 0x9E3779B97F4A7C15ULL + 1ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}

static int nthreads_or_default(int n) {
#ifdef _OPENMP
  return n > 0 ? n : omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

void ora_symbolic(uint64_t r0, uint64_t r1, const uint32_t *a_rowptr,
                  const uint32_t *a_cols, uint64_t b_cols_n,
                  const uint32_t *b_rowptr, const uint32_t *b_cols,
                  uint32_t *nnz_row, uint64_t *set_hash, int nthreads) {
  int nt = nthreads_or_default(nthreads);
  size_t words = (size_t)((b_cols_n + 63) / 64);
#pragma omp parallel num_threads(nt)
  {
    uint64_t *mark = (uint64_t *)calloc(words ? words : 1, sizeof(uint64_t));
    uint32_t *touched = (uint32_t *)malloc(sizeof(uint32_t) * (b_cols_n ? b_cols_n : 1));
#pragma omp for schedule(dynamic, 64)
    for (int64_t r = (int64_t)r0; r < (int64_t)r1; ++r) {
      uint32_t n = 0;
      uint64_t h = 0;
      for (uint32_t j = a_rowptr[r]; j < a_rowptr[r + 1]; ++j) {
        uint32_t k = a_cols[j];
        for (uint32_t q = b_rowptr[k]; q < b_rowptr[k + 1]; ++q) {
          uint32_t c = b_cols[q];
          uint64_t bit = 1ULL << (c & 63);
          if (!(mark[c >> 6] & bit)) {
            mark[c >> 6] |= bit;
            touched[n++] = c;
            h += ora_mix64(c);
          }
        }
      }
      for (uint32_t t = 0; t < n; ++t) mark[touched[t] >> 6] = 0;
      nnz_row[r - r0] = n;
      if (set_hash) set_hash[r - r0] = h;
    }
    free(mark);
    free(touched);
  }
}

static int cmp_u32(const void *a, const void *b) {
  uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
  return x < y ? -1 : x > y;
}

static inline double ld_val(int vtype, const void *v, uint64_t i) {
  return vtype == 0 ? ((const double *)v)[i] : (double)((const float *)v)[i];
}

/* One row of Gustavson with dense SPA (sum, Neumaier compensation, |.|
 * sum) and a touched list; writes sorted output at out_cols/out_vals. */
static uint32_t gustavson_row(uint64_t r, int vtype, const uint32_t *a_rowptr,
                              const uint32_t *a_cols, const void *a_vals,
                              const uint32_t *b_rowptr, const uint32_t *b_cols,
                              const void *b_vals, double *acc, double *comp,
                              double *absacc, uint8_t *flag, uint32_t *touched,
                              uint32_t *out_cols, double *out_vals,
                              double *cond, int compensated) {
  uint32_t n = 0;
  for (uint32_t j = a_rowptr[r]; j < a_rowptr[r + 1]; ++j) {
    uint32_t k = a_cols[j];
    double av = ld_val(vtype, a_vals, j);
    for (uint32_t q = b_rowptr[k]; q < b_rowptr[k + 1]; ++q) {
      uint32_t c = b_cols[q];
      double p = av * ld_val(vtype, b_vals, q);
      if (!flag[c]) {
        flag[c] = 1;
        touched[n++] = c;
        acc[c] = 0.0;
        comp[c] = 0.0;
        absacc[c] = 0.0;
      }
      if (compensated) {
        double s = acc[c], t = s + p;
        if (fabs(s) >= fabs(p))
          comp[c] += (s - t) + p;
        else
          comp[c] += (p - t) + s;
        acc[c] = t;
      } else {
        acc[c] = acc[c] + p;
      }
      absacc[c] += fabs(p);
    }
  }
  qsort(touched, n, sizeof(uint32_t), cmp_u32);
  for (uint32_t t = 0; t < n; ++t) {
    uint32_t c = touched[t];
    out_cols[t] = c;
    out_vals[t] = compensated ? acc[c] + comp[c] : acc[c];
    if (cond) cond[t] = absacc[c];
    flag[c] = 0;
  }
  return n;
}

void ora_numeric(uint64_t r0, uint64_t r1, int vtype,
                 const uint32_t *a_rowptr, const uint32_t *a_cols,
                 const void *a_vals, uint64_t b_cols_n,
                 const uint32_t *b_rowptr, const uint32_t *b_cols,
                 const void *b_vals, uint64_t *out_rowptr, uint32_t *out_cols,
                 double *out_vals, double *cond, int compensated,
                 int nthreads) {
  int nt = nthreads_or_default(nthreads);
  size_t nc = (size_t)(b_cols_n ? b_cols_n : 1);
#pragma omp parallel num_threads(nt)
  {
    double *acc = (double *)malloc(sizeof(double) * nc);
    double *comp = (double *)malloc(sizeof(double) * nc);
    double *absacc = (double *)malloc(sizeof(double) * nc);
    uint8_t *flag = (uint8_t *)calloc(nc, 1);
    uint32_t *touched = (uint32_t *)malloc(sizeof(uint32_t) * nc);
#pragma omp for schedule(dynamic, 64)
    for (int64_t r = (int64_t)r0; r < (int64_t)r1; ++r) {
      uint64_t o = out_rowptr[r - r0];
      gustavson_row((uint64_t)r, vtype, a_rowptr, a_cols, a_vals, b_rowptr,
                    b_cols, b_vals, acc, comp, absacc, flag, touched,
                    out_cols + o, out_vals + o, cond ? cond + o : NULL,
                    compensated);
    }
    free(acc);
    free(comp);
    free(absacc);
    free(flag);
    free(touched);
  }
}

uint64_t ora_symbolic_rows(uint64_t nrows, const uint32_t *rows,
                           const uint32_t *a_rowptr, const uint32_t *a_cols,
                           uint64_t b_cols_n, const uint32_t *b_rowptr,
                           const uint32_t *b_cols, uint32_t *nnz_row,
                           int nthreads) {
  int nt = nthreads_or_default(nthreads);
  size_t words = (size_t)((b_cols_n + 63) / 64);
  uint64_t total = 0;
#pragma omp parallel num_threads(nt) reduction(+ : total)
  {
    uint64_t *mark = (uint64_t *)calloc(words ? words : 1, sizeof(uint64_t));
    uint32_t *touched = (uint32_t *)malloc(sizeof(uint32_t) * (b_cols_n ? b_cols_n : 1));
#pragma omp for schedule(dynamic, 16)
    for (int64_t i = 0; i < (int64_t)nrows; ++i) {
      uint32_t r = rows[i], n = 0;
      for (uint32_t j = a_rowptr[r]; j < a_rowptr[r + 1]; ++j) {
        uint32_t k = a_cols[j];
        for (uint32_t q = b_rowptr[k]; q < b_rowptr[k + 1]; ++q) {
          uint32_t c = b_cols[q];
          uint64_t bit = 1ULL << (c & 63);
          if (!(mark[c >> 6] & bit)) {
            mark[c >> 6] |= bit;
            touched[n++] = c;
          }
        }
      }
      for (uint32_t t = 0; t < n; ++t) mark[touched[t] >> 6] = 0;
      nnz_row[i] = n;
      total += n;
    }
    free(mark);
    free(touched);
  }
  return total;
}

void ora_numeric_rows(uint64_t nrows, const uint32_t *rows, int vtype,
                      const uint32_t *a_rowptr, const uint32_t *a_cols,
                      const void *a_vals, uint64_t b_cols_n,
                      const uint32_t *b_rowptr, const uint32_t *b_cols,
                      const void *b_vals, uint64_t *out_rowptr,
                      uint32_t *out_cols, double *out_vals, double *cond,
                      int nthreads) {
  int nt = nthreads_or_default(nthreads);
  size_t nc = (size_t)(b_cols_n ? b_cols_n : 1);
#pragma omp parallel num_threads(nt)
  {
    double *acc = (double *)malloc(sizeof(double) * nc);
    double *comp = (double *)malloc(sizeof(double) * nc);
    double *absacc = (double *)malloc(sizeof(double) * nc);
    uint8_t *flag = (uint8_t *)calloc(nc, 1);
    uint32_t *touched = (uint32_t *)malloc(sizeof(uint32_t) * nc);
#pragma omp for schedule(dynamic, 16)
    for (int64_t i = 0; i < (int64_t)nrows; ++i) {
      uint64_t o = out_rowptr[i];
      gustavson_row(rows[i], vtype, a_rowptr, a_cols, a_vals, b_rowptr, b_cols,
                    b_vals, acc, comp, absacc, flag, touched, out_cols + o,
                    out_vals + o, cond ? cond + o : NULL, 1);
    }
    free(acc);
    free(comp);
    free(absacc);
    free(flag);
    free(touched);
  }
}

/* CPU baseline, not a checker: the plain row-wise Gustavson a CPU
 * library runs.  Dense FP64 accumulator with a touched list (no
 * compensation, no |a*b| sums); a row's columns come out in order by a
 * scan of the touched range when it is dense (at least 1/16 of it
 * touched), else by sorting the touched list; each thread appends its
 * rows to its own growing output buffer.  Returns nnz of rows [r0, r1);
 * *checksum receives the sum of all values (keeps the work observable). */
uint64_t ora_baseline_spgemm(uint64_t r0, uint64_t r1, int vtype,
                             const uint32_t *a_rowptr, const uint32_t *a_cols,
                             const void *a_vals, uint64_t b_cols_n,
                             const uint32_t *b_rowptr, const uint32_t *b_cols,
                             const void *b_vals, double *checksum, int nthreads) {
  int nt = nthreads_or_default(nthreads);
  size_t nc = (size_t)(b_cols_n ? b_cols_n : 1);
  uint64_t total = 0;
  double sum = 0.0;
#pragma omp parallel num_threads(nt) reduction(+ : total, sum)
  {
    double *acc = (double *)malloc(sizeof(double) * nc);
    uint8_t *flag = (uint8_t *)calloc(nc, 1);
    uint32_t *touched = (uint32_t *)malloc(sizeof(uint32_t) * nc);
    size_t cap = 1 << 20, used = 0;
    uint32_t *oc = (uint32_t *)malloc(sizeof(uint32_t) * cap);
    double *ov = (double *)malloc(sizeof(double) * cap);
#pragma omp for schedule(dynamic, 64)
    for (int64_t r = (int64_t)r0; r < (int64_t)r1; ++r) {
      uint32_t n = 0, lo = UINT32_MAX, hi = 0;
      for (uint32_t j = a_rowptr[r]; j < a_rowptr[r + 1]; ++j) {
        uint32_t k = a_cols[j];
        double av = ld_val(vtype, a_vals, j);
        for (uint32_t q = b_rowptr[k]; q < b_rowptr[k + 1]; ++q) {
          uint32_t c = b_cols[q];
          double p = av * ld_val(vtype, b_vals, q);
          if (!flag[c]) {
            flag[c] = 1;
            touched[n++] = c;
            acc[c] = p;
            if (c < lo) lo = c;
            if (c > hi) hi = c;
          } else {
            acc[c] += p;
          }
        }
      }
      if (used + n > cap) {
        while (used + n > cap) cap *= 2;
        oc = (uint32_t *)realloc(oc, sizeof(uint32_t) * cap);
        ov = (double *)realloc(ov, sizeof(double) * cap);
      }
      if (n && (uint64_t)n * 16 >= (uint64_t)(hi - lo + 1)) {
        uint32_t t = 0;
        for (uint32_t c = lo; c <= hi; ++c)
          if (flag[c]) touched[t++] = c;
      } else {
        qsort(touched, n, sizeof(uint32_t), cmp_u32);
      }
      for (uint32_t t = 0; t < n; ++t) {
        uint32_t c = touched[t];
        oc[used] = c;
        ov[used] = acc[c];
        sum += acc[c];
        ++used;
        flag[c] = 0;
      }
      total += n;
      /* the buffer is reused: a caller would hand each row range on */
      if (used > (1u << 24)) used = 0;
    }
    free(acc);
    free(flag);
    free(touched);
    free(oc);
    free(ov);
  }
  if (checksum) *checksum = sum;
  return total;
}
