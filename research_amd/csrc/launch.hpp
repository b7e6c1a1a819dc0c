// Host-side launchers of the gfx950 kernels (internal to libocean.so).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "oc_internal.hpp"

namespace oc {

constexpr int kNumBins = 14;   // 7 hash + 7 dense bins (include/Common.h:53)
constexpr int kSkipBin = 14;   // rows left out of a binned launch
constexpr int kBinSlots = 15;  // bins 0..13 plus the skip bin

// Device view of a CSR with 32-bit offsets (the A and B inputs).
struct DevCsr {
  uint64_t rows, cols, nnz;
  const uint32_t *rp;
  const uint32_t *ci;
  const void *v;
  int vtype;
};

// gfx950 tables (oc_tables plus derived launch shapes).
struct Tables {
  uint32_t hash_sym[7], dense_sym[7], hash_num[7], dense_num[7];
  uint32_t block[7], lds[7];
  uint32_t max_dense_sym_bin;
  uint32_t win_num_cols;   // numeric window (columns)
  uint32_t lds_per_cu, num_cu, lds_max_block;
};
// Tables for the current device (nominal 160 KB / 256 CUs without a GPU),
// or the ones a TablesScope on this thread set.
const Tables &tables();
int tables_init();
// Tables for per-bin shapes (workgroups per CU, threads per workgroup;
// null or 0 entries: the defaults); an error when a bin's tables do not fit.
int tables_for(const int32_t *concurrency, const int32_t *block, Tables *out);
// While it lives, tables() on this thread returns *t (t null: no change).
struct TablesScope {
  const Tables *prev;
  explicit TablesScope(const Tables *t);
  ~TablesScope();
  TablesScope(const TablesScope &) = delete;
  TablesScope &operator=(const TablesScope &) = delete;
};

// Per-row analysis outputs.
struct RowInfo {
  uint32_t *products, *max_b, *avg_b, *left, *right;
};

// Learned row-size predictor (k_mlp.hip): features of the rows of C = A*B
// and per-row capacities from the model.
int launch_mlp_features(const DevCsr &A, const DevCsr &B, const RowInfo &ri, float *feat,
                        hipStream_t s);
int launch_mlp_estimates(const oc_mlp *m, const float *feat, uint64_t n, const RowInfo &ri,
                         double q, uint32_t *est, hipStream_t s);

// temp: analysis_temp_bytes(A.rows, A.nnz, B.rows) bytes, or null to have
// the launcher allocate (stream-ordered) and free it
size_t analysis_temp_bytes(uint64_t a_rows, uint64_t a_nnz, uint64_t b_rows);
int launch_analysis(const DevCsr &A, const DevCsr &B, const RowInfo &ri,
                    uint32_t *max_a_nnz, uint32_t *max_products, void *temp,
                    hipStream_t s);
// Device-side checks: *err |= bit when the u32 (is64 0) or u64 at x
// exceeds limit; and a fill of n u32 (a kernel, not a memset node).
int launch_check_gt(const void *x, int is64, uint64_t limit, uint32_t bit, uint32_t *err,
                    hipStream_t s);
int launch_fill_u32(uint32_t *p, uint32_t n, uint32_t v, hipStream_t s);
// bounds = {0, *count}: the row range of a device-counted list
int launch_count_bounds(const uint32_t *count, uint32_t *bounds, hipStream_t s);
size_t scan_temp_bytes(uint64_t n);
int launch_scan_u32(const uint32_t *in, uint64_t *out, uint64_t n, void *temp,
                    hipStream_t s);
int launch_murmur(const uint32_t *cols, uint32_t *out, uint64_t n,
                  hipStream_t s);
int launch_hll_build(int log_p, const DevCsr &B, uint8_t *hll, hipStream_t s);
// mode 0: per-row estimate with expansion and rounding into est[];
// mode 1: nsamples sampled rows -> sample_rows[], sample_est[] (raw)
int launch_hll_merge(int log_p, const DevCsr &A, const uint8_t *hll,
                     double expansion, uint32_t *est, uint32_t nsamples,
                     uint32_t *sample_rows, int32_t *sample_est,
                     const uint32_t *products, uint32_t *sample_prod, int mode,
                     void *temp, hipStream_t s);
size_t hll_merge_temp_bytes(uint64_t rows);

// Binning.  kind 0 = symbolic rule, 1 = numeric rule.  Rows whose
// `filter` value is <= skip_below go to the skip bin (filter may be null).
struct BinParams {
  int kind;
  const uint32_t *input;  // products (symbolic) or nnz / estimate (numeric)
  uint32_t *input_mut;    // numeric: modified for dense-bin-0 rows (est)
  const uint32_t *left, *right;
  int use_compaction;
  double expand_factor, avg_compaction;
  int expand, modify_dense_est, use_largest;
  uint32_t hybrid_max;
  const uint32_t *filter;
  uint32_t skip_below;
  const uint32_t *skip_flag;  // rows with a non-zero flag go to the skip bin
};
size_t bin_temp_bytes(uint64_t rows);
int launch_binning(const BinParams &bp, uint64_t rows, uint8_t *bin_of_row,
                   uint32_t *bin_rows, uint32_t *bin_offsets, void *temp,
                   hipStream_t s);

// Copy rows (all rows, or those listed) from a temp layout into C.
// A row whose count exceeds cap[row] (when cap is given) is skipped.
int launch_copy_rows(int vtype, uint64_t nrows, const uint32_t *rows,
                     const uint32_t *nnz, const uint64_t *src_off, const uint32_t *cap,
                     const uint32_t *src_cols, const void *src_vals,
                     const uint64_t *dst_off, uint32_t *dst_cols,
                     void *dst_vals, hipStream_t s);

// Reserve spare space for overflow rows; off[row] = base + reservation.
int launch_assign_spare(const uint32_t *ovf_rows, uint32_t n,
                        const uint32_t *products, const uint32_t *left,
                        const uint32_t *right, uint64_t base,
                        unsigned long long *spare_used, uint64_t *off,
                        hipStream_t s);

// Work accounting over a row list: out[0..3] += {rows, a_nnz, products,
// out_nnz} (out_nnz from u64 offsets c_rp, else from cnt).
int launch_row_sums(const uint32_t *rows, uint32_t n, const uint32_t *a_rp,
                    const uint32_t *products, const uint64_t *c_rp,
                    const uint32_t *cnt, unsigned long long *out, hipStream_t s);

// Validation helpers over a device CSR (off_bytes 4 or 8).
int launch_validate(const oc_csr &M, unsigned long long *bad, hipStream_t s);
int launch_row_set_hash(const oc_csr &M, uint64_t *hash, hipStream_t s);
int launch_spmv(const oc_csr &M, const double *x, double *y, hipStream_t s);

// ------------------------------------------------------------ accumulators
struct AccArgs {
  DevCsr A, B;
  const uint32_t *products, *max_b, *left, *right;
};

// Output description of a numeric launch.
struct NumOut {
  const uint64_t *off;     // per-row output offset
  uint32_t *cols;
  void *vals;
  uint32_t *nnz;           // per-row count written (may be null)
  const uint32_t *cap;     // per-row capacity (est mode) or null
  uint32_t *ovf_rows;      // est mode: overflow row list
  uint32_t *ovf_count;
  const uint32_t *nnz_hint;  // exact per-row counts when known (or null)
};

// Rows of a launch taken from the device instead of the host: with bounds
// set, the kernel is persistent (as many workgroups as fit the chip) over
// rows[bounds[0] .. bounds[1]) of its list, so the launch needs no host
// read-back of the bin sizes (a whole run can be captured into a graph).
// err (numeric kernels): when *err != 0 at launch, nothing is written (the
// output's capacity check failed).
struct RowSel {
  const uint32_t *bounds = nullptr;
  const uint32_t *err = nullptr;
  uint32_t *claim = nullptr;  // row-claim counter of the hash/dense bins (zeroed per launch)
  // the bin's first list entry when the host knows it (eager runs, which
  // pass the bin's row count): bins with many workgroups per CU then run
  // one workgroup per row from rows + first
  uint32_t first = ~0u;
};

// Symbolic: rows of one bin.  Overflowing rows (estimated-compaction hash
// tables that fill up) are appended to ovf_rows.
int launch_sym_hash(int bin, const AccArgs &a, const uint32_t *rows,
                    uint32_t nrows, uint32_t limit, uint32_t *nnz,
                    uint32_t *ovf_rows, uint32_t *ovf_count, hipStream_t s,
                    const RowSel &sel = RowSel{});
// bmoff/bm (optional): rows with bmoff[r] != ~0 store their bitmap (uint4
// groups from column left & ~127) at bm + bmoff[r] (launch_bitmap_offsets)
int launch_sym_dense(int dense_bin, const AccArgs &a, const uint32_t *rows,
                     uint32_t nrows, uint32_t *nnz, hipStream_t s,
                     const uint64_t *bmoff = nullptr, void *bm = nullptr,
                     const RowSel &sel = RowSel{});
// Windowed kernels take the bin row list (device) and its host-known
// length; cursor scratch is grid * max_a_nnz u32.  bmoff/bm as for
// launch_sym_dense: rows with a stored bitmap get their windows' words.
int launch_sym_window(const AccArgs &a, const uint32_t *rows, uint32_t nrows,
                      uint32_t max_a_nnz, uint32_t *nnz, void *scratch,
                      hipStream_t s, const RowSel &sel = RowSel{},
                      const uint64_t *bmoff = nullptr, void *bm = nullptr);
size_t window_scratch_bytes(uint32_t max_a_nnz);

int launch_num_hash(int bin, const AccArgs &a, const uint32_t *rows,
                    uint32_t nrows, const NumOut &o, hipStream_t s,
                    const RowSel &sel = RowSel{});
int launch_num_dense(int dense_bin, const AccArgs &a, const uint32_t *rows,
                     uint32_t nrows, const NumOut &o, hipStream_t s,
                     const RowSel &sel = RowSel{});
int launch_num_window(const AccArgs &a, const uint32_t *rows, uint32_t nrows,
                      uint32_t max_a_nnz, const NumOut &o, void *scratch,
                      hipStream_t s, const RowSel &sel = RowSel{});
// Tiled numeric path for rows too large for LDS hash / dense kernels:
// tile_plan picks the tile width (2^logW columns) and the size of the
// tile-start table (B.rows x T1 u32); returns 0 when tiles do not apply.
int tile_plan(uint64_t b_rows, uint64_t b_cols, uint32_t log2_cols, uint32_t bs,
              uint32_t *logW, uint32_t *T1, size_t *ts_bytes);
// base (optional, per B row): tiles are counted from base[k] instead of 0
int launch_tile_starts(const DevCsr &B, uint32_t logW, uint32_t T1, uint32_t *ts,
                       hipStream_t s, const uint32_t *base = nullptr);
int launch_num_tiles(const AccArgs &a, const uint32_t *rows, uint32_t nrows, uint32_t logW,
                     uint32_t bs, const uint32_t *ts, uint32_t T1, const NumOut &o,
                     void *scratch, hipStream_t s, const uint32_t *nrows_dev = nullptr);
// Rank-compressed tiles (tile-start table of granularity 2^logT columns;
// tsw, when given, the compact table of the widest tile: wper =
// wide_starts_plan entries per B row (its tiles + 1), built by
// launch_tile_starts at rank_wide_log(bs)).
uint32_t rank_wide_log(uint32_t bs);
int wide_starts_plan(uint64_t b_cols, uint32_t bs, size_t *entries_per_row);
// the same for a table of 2^lw-column tiles
int wide_starts_for_log(uint64_t b_cols, uint32_t lw, size_t *entries_per_row);
// bbase (optional, per B row): the first column of the row's block in a
// block-diagonal product; the tile tables are then built from the same bases.
int launch_num_rank(const AccArgs &a, const uint32_t *rows, uint32_t nrows, uint32_t logT,
                    const uint32_t *ts, uint32_t T1, const uint32_t *tsw, uint32_t wper,
                    const uint32_t *bbase,
                    uint32_t pcap, uint32_t bs, const NumOut &o, void *scratch, hipStream_t s,
                    const RowSel &sel = RowSel{});
// One-gather rank tiles (k_rank.hip), 1024 or 512 threads (bs), tiles up
// to 2^rank1_wide_log(bs) columns; tsw only when built at that width.  Rows
// with a stored bitmap (bmoff[r] != ~0, no block bases) gather each
// product once; the others mark their tiles first.
// pc (optional): rows with more than pc.per_piece products are split into
// pieces of consecutive tiles taken by different workgroups (eager runs
// without block bases): kp holds nrows bytes, extra 2 * max_extra u32,
// n_extra one u32.  Rows without a stored bitmap are split only with pcnt
// (max_extra u32), which a count pass fills with each piece's outputs.
struct R1Pieces {
  uint8_t *kp = nullptr;
  uint32_t *extra = nullptr, *n_extra = nullptr, *pcnt = nullptr;
  uint32_t max_extra = 0, per_piece = 0;
};
uint32_t rank1_wide_log(uint32_t bs);
int launch_num_rank1(const AccArgs &a, const uint32_t *rows, uint32_t nrows, uint32_t logT,
                     const uint32_t *ts, uint32_t T1, const uint32_t *tsw, uint32_t wper,
                     const uint32_t *bbase, const void *bm, const uint64_t *bmoff,
                     const NumOut &o, void *scratch, hipStream_t s,
                     const RowSel &sel = RowSel{}, uint32_t bs = 1024,
                     const R1Pieces &pc = R1Pieces{});
// Stored-bitmap offsets (uint4 units) for the listed rows with at least
// min_products products; bmoff[0..m) = ~0 elsewhere; offs[n] = total.
// sz: n u32, offs: n + 1 u64, scan_tmp: scan_temp_bytes(n).
int launch_bitmap_offsets(const uint32_t *rows, uint32_t n, uint64_t m,
                          const uint32_t *products, const uint32_t *left,
                          const uint32_t *right, uint32_t min_products, uint32_t *sz,
                          uint64_t *offs, void *scan_tmp, uint64_t *bmoff, hipStream_t s);
// Ultra-sparse expand-sort-compress for all rows with products <= esc_n.
int launch_esc(int esc_n, const AccArgs &a, const NumOut &o, hipStream_t s);

// Per-row sort of (col,val) segments of a device CSR.
int launch_sort_rows(oc_csr &M, const uint32_t *rows, uint64_t n,
                     hipStream_t s);

// Long rows in column bands, band-major (k_band.hip).
int band_plan(uint64_t b_cols, uint32_t log2_cols, uint32_t *nbands);
int launch_band_starts(const DevCsr &B, uint32_t logW, uint32_t nbands, uint32_t *ts,
                       hipStream_t s);
// flag[r] = long row; list = flagged rows ascending; pos[rows] = count
int launch_band_select(const RowInfo &ri, uint64_t rows, uint32_t min_products,
                       uint32_t band_cols, uint32_t *flag, uint64_t *pos, void *scan_tmp,
                       uint32_t *list, hipStream_t s);
int launch_band_count(const AccArgs &a, const uint32_t *rows, uint32_t nrows, uint32_t logW,
                      uint32_t nbands, const uint32_t *ts, uint32_t *cnt, uint32_t *counter,
                      hipStream_t s);
int launch_band_row_nnz(const uint32_t *rows, uint32_t nrows, uint32_t nbands,
                        const uint64_t *pref, uint32_t *nnz, hipStream_t s);
int launch_band_numeric(const AccArgs &a, const uint32_t *rows, uint32_t nrows, uint32_t logW,
                        uint32_t nbands, const uint32_t *ts, const uint64_t *pref,
                        const NumOut &o, uint32_t *counter, hipStream_t s);

// FAT products (k_fat.hip): dense rows over all of B's columns, each A row
// split over `splits` workgroups; dense [rows][B.cols] FP64 and bits
// [rows][(B.cols+31)/32] zeroed by the caller.
size_t fat_lds_bytes(uint64_t b_cols);
int launch_fat_accumulate(const DevCsr &A, const DevCsr &B, uint32_t splits, int G,
                          double *dense, uint32_t *bits, hipStream_t s);
int launch_fat_count(const uint32_t *bits, uint64_t rows, uint64_t b_cols, uint32_t *nnz,
                     hipStream_t s);
int launch_fat_compact(int vtype, const double *dense, const uint32_t *bits, uint64_t rows,
                       uint64_t b_cols, const uint64_t *c_off, uint32_t *c_cols, void *c_vals,
                       hipStream_t s);

// Spy-plot segmentation (k_seg.hip).
struct SegSums {
  unsigned long long sum_products, sum_a_nnz, sum_sq_products;
  uint32_t max_b_row, max_a_row, min_left, max_right;
};
int launch_seg_density(const DevCsr &A, uint32_t grid_h, uint32_t grid_w, uint32_t tile_rows,
                       uint32_t tile_cols, uint32_t *density, hipStream_t s);
int launch_seg_threshold(const uint32_t *density, uint32_t n, double k, uint32_t *labels,
                         unsigned long long *out, hipStream_t s);
int launch_seg_components(uint32_t grid_h, uint32_t grid_w, uint32_t max_sweeps,
                          uint32_t *labels, uint32_t *next, unsigned long long *out,
                          hipStream_t s);
int launch_seg_dominant(const uint32_t *labels, uint32_t grid_h, uint32_t grid_w,
                        uint32_t *row_label, uint32_t *row_hot, hipStream_t s);
int launch_seg_reduce(const uint32_t *seg_start, const uint32_t *seg_end, uint32_t nseg,
                      const DevCsr &A, const RowInfo &ri, SegSums *out, hipStream_t s);

}  // namespace oc
