// SpGEMM orchestrator: workflow decisions, the ultra-sparse, precise and
// estimation workflows, per-stage timers and the stats JSON.
//
// Follows SpGEMM::run and its stage functions (kernels/SpGEMM.cuh:213-1736)
// with one HIP stream, stream-ordered allocation from the default memory
// pool (release threshold raised as SpGEMM.cuh:219-223 does) and the gfx950
// kernels of k_prep.hip / k_accum.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "launch.hpp"

using namespace oc;

#define HIP_TRY(call)                                                         \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess) {                                                   \
      set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,                    \
                hipGetErrorString(e_));                                       \
      return e_ == hipErrorOutOfMemory ? OC_ERR_NOMEM : OC_ERR_HIP;           \
    }                                                                         \
  } while (0)

namespace {

int bits_needed(uint64_t x) {
  int b = 0;
  while (x) {
    ++b;
    x >>= 1;
  }
  return b;
}

// Device memory for the temporaries of a handle's runs: chunks from
// hipMalloc, handed out by a bump pointer and all reset when a run ends
// (every run ends with a stream synchronisation).  Repeated runs reuse the
// same memory; handing every temporary back to the stream-ordered pool
// instead fragments it under products of varying size until allocations
// stall while the pool trims; and device memory freed and allocated again
// is cleared by the driver first, at tens of GB/s.  The chunks stay with the
// handle until it is destroyed, so a repeated product allocates nothing
// after its first run.
struct ChunkStore {
  struct Chunk {
    uint8_t *base;
    size_t size, used;
  };
  std::vector<Chunk> chunks;
  size_t total() const {
    size_t t = 0;
    for (const Chunk &c : chunks) t += c.size;
    return t;
  }
  void drop() {
    for (Chunk &c : chunks) (void)hipFree(c.base);
    chunks.clear();
  }
  // Frees the chunks a run left untouched (call once its stream has
  // drained, before the run's marks are reset), so the store holds what the
  // latest run needed rather than the union of every run's sizes.  Freed
  // memory is cleared by the driver before its next use, so the runs call
  // this only when the store holds more than a quarter of the device.
  void trim() {
    size_t k = 0;
    for (Chunk &c : chunks) {
      if (c.used) chunks[k++] = c;
      else (void)hipFree(c.base);
    }
    chunks.resize(k);
  }
};

// One run's allocations from the handle's chunk store.
struct Arena {
  ChunkStore *cs;
  struct Rec {
    size_t chunk, off;
    void *p;
  };
  std::vector<Rec> recs;
  explicit Arena(ChunkStore *c) : cs(c) {}
  ~Arena() { release(); }
  template <typename T>
  int get(T **p, uint64_t count) {
    const size_t bytes = ((size_t)std::max<uint64_t>(count, 1) * sizeof(T) + 255) & ~(size_t)255;
    // best fit: the chunk with the least room that holds the request, so
    // small requests do not keep large chunks in use
    size_t best = cs->chunks.size();
    for (size_t i = 0; i < cs->chunks.size(); ++i) {
      const ChunkStore::Chunk &c = cs->chunks[i];
      if (c.size - c.used >= bytes &&
          (best == cs->chunks.size() ||
           c.size - c.used < cs->chunks[best].size - cs->chunks[best].used))
        best = i;
    }
    if (best < cs->chunks.size()) {
      ChunkStore::Chunk &c = cs->chunks[best];
      *p = (T *)(c.base + c.used);
      recs.push_back(Rec{best, c.used, (void *)*p});
      c.used += bytes;
      return OC_OK;
    }
    uint8_t *base = nullptr;
    const size_t size = std::max<size_t>(bytes, 64u << 20);
    hipError_t e = hipMalloc((void **)&base, size);
    if (e == hipErrorOutOfMemory) {  // the cached product outputs back, once
      (void)hipGetLastError();
      out_release();
      e = hipMalloc((void **)&base, size);
    }
    if (e != hipSuccess) {
      (void)hipGetLastError();
      set_error("device allocation of %zu bytes failed: %s", bytes, hipGetErrorString(e));
      return OC_ERR_NOMEM;
    }
    cs->chunks.push_back(ChunkStore::Chunk{base, size, bytes});
    *p = (T *)base;
    recs.push_back(Rec{cs->chunks.size() - 1, 0, (void *)base});
    return OC_OK;
  }
  // Only the latest allocation is given back (stack order); the others
  // stay until the run ends.
  void free(void *p) {
    if (!recs.empty() && recs.back().p == p) {
      cs->chunks[recs.back().chunk].used = recs.back().off;
      recs.pop_back();
    }
  }
  void release() {
    recs.clear();
    for (ChunkStore::Chunk &c : cs->chunks) c.used = 0;
  }
};

// Stream-ordered timestamps; the time between two marks is charged to the
// later mark's stage.  Resolved once the stream has drained.
struct Marks {
  hipStream_t s;
  std::vector<std::pair<std::string, hipEvent_t>> ev;
  bool enabled = true;
  explicit Marks(hipStream_t st) : s(st) {}
  ~Marks() {
    for (auto &p : ev) (void)hipEventDestroy(p.second);
  }
  void mark(const char *name) {
    if (!enabled) return;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    (void)hipEventRecord(e, s);
    ev.emplace_back(name, e);
  }
  void resolve(std::map<std::string, double> &out, double &total) {
    total = 0;
    for (size_t i = 1; i < ev.size(); ++i) {
      float ms = 0;
      if (hipEventElapsedTime(&ms, ev[i - 1].second, ev[i].second) == hipSuccess) {
        out[ev[i].first] += ms;
        total += ms;
      }
    }
  }
};

template <typename T>
int readback(T *host, const T *dev, size_t n, hipStream_t s) {
  HIP_TRY(hipMemcpyAsync(host, dev, n * sizeof(T), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return OC_OK;
}

struct Decision {
  int ana1_type = 0;  // 0 ultrasparse, 1 HLL-sampled, 2 fat
  int spark_kernel_size = 0, spark_need_outlier = 0, spark_outlier_kernel_size = 0;
  int log_p = 0;
  double hll_expansion = 0;
  int nsamples = 0;
  double comp1 = 0, comp2 = 0;
  int ana2_type = 0;  // 0 precise, 1 estimation
  int dense_query_bitmap = 0;
  double comp_avg = 0, comp_var = 0, safe_comp = 0;
  int use_est_symbolic = 0;
  int hash_largest = 0;
};

}  // namespace

struct oc_spgemm {
  oc_config cfg;
  // last run
  Decision d;
  uint64_t a_rows = 0, a_cols = 0, a_nnz = 0, b_rows = 0, b_cols = 0, b_nnz = 0;
  uint64_t c_rows = 0, c_cols = 0, c_nnz = 0;
  uint64_t total_products = 0;
  uint32_t max_products = 0, max_a_nnz = 0;
  uint32_t bin1[kBinSlots + 1] = {0}, bin2[kBinSlots + 1] = {0};
  bool have_bin1 = false, have_bin2 = false;
  // block-diagonal B (set by the batched layer for its runs): first column
  // of each B row's block and the widest block
  const uint32_t *b_block_base = nullptr;
  uint64_t b_block_span = 0;
  const void *b_block_rp = nullptr;  // the B the bases belong to
  // spy-plot segmentation of the last run (SpGEMM.cuh:1853-1886)
  struct SegRecord {
    uint32_t row_start, row_end, col_left, col_right, cls;
    uint64_t sum_products, sum_a_nnz;
    uint32_t max_b_row_len, max_a_row_len;
    double mean_row_len, var_row_len, local_er;
  };
  struct Segmentation {
    bool ran = false;
    uint32_t grid_h = 0, grid_w = 0, tile_rows = 0, tile_cols = 0, num_components = 0;
    uint64_t cc_iterations = 0;
    double density_mean = 0, density_stddev = 0;
    float t_density = 0, t_threshold = 0, t_cc = 0, t_classify = 0;
    std::vector<SegRecord> segments;
  } seg;
  uint64_t total_est = 0;
  uint64_t est_spare_need = 0;       // entries the overflow rows reserved
  bool est_separate_spare = false;   // ... in a buffer of their own
  bool est_memory_fallback = false;
  uint64_t bitmap_bytes = 0;         // row bitmaps stored by the symbolic pass  // estimated region + C beyond the device: precise instead
  uint32_t overflow_rows = 0, sym_overflow_rows = 0, precise_recount = 0;
  int vtype = OC_F64;
  double total_ms = 0;
  std::map<std::string, double> stage_ms;
  struct FamStat {
    std::string name, stage;
    bool numeric;
    unsigned long long rows, a_nnz, products, nnz;
  };
  std::vector<FamStat> fams;
  // estimation override (learned predictor / forced fallback tests)
  const uint32_t *est_override = nullptr;
  int force_workflow = -1;  // -1 none, 0 precise, 1 estimation
  // learned row-size predictor (set by the caller or loaded from the config)
  const oc_mlp *model = nullptr;
  oc_mlp *owned_model = nullptr;
  double model_q = 0.99;
  bool learned_run = false;        // estimation workflow on learned capacities
  bool learned_consulted = false;  // the predictor ran and picked the workflow
  uint64_t learned_total = 0;      // sum of the learned capacities
  // Numeric plan of a precise run kept for value-only replays: the row
  // information, numeric bins and scratch stay allocated; C is the caller's.
  struct Plan {
    bool valid = false;
    std::vector<void *> mem;
    DevCsr A{}, B{};
    RowInfo ri{};
    const uint32_t *brows = nullptr;
    uint32_t hoff[kBinSlots + 1] = {0};
    NumOut o{};
    uint32_t *tile_ts = nullptr, *tile_tsw = nullptr;
    const uint32_t *block_base = nullptr;
    void *row_bm = nullptr;
    const uint64_t *row_bmoff = nullptr;
    // what the whole-path replay (analysis to numeric, on the device) needs
    uint32_t *scal = nullptr;           // analysis maxima
    void *ana_tmp = nullptr, *scan_tmp = nullptr, *bin_tmp = nullptr;
    uint8_t *bor1 = nullptr, *bor2 = nullptr;
    uint32_t *brows1 = nullptr, *boff1 = nullptr, *boff2 = nullptr;
    uint32_t *nnz = nullptr, *ovf = nullptr;
    uint32_t *err = nullptr;            // check bits of the last whole-path run
    uint64_t cap_nnz = 0;               // C's capacity (the planned nnz)
    uint32_t tsw_log = 0, tsw_per = 0;  // compact tile-start table shape
    BinParams sym_bp{}, num_bp{};
    bool est_symbolic = false;          // shrunken symbolic tables (recount overflow rows)
    hipGraph_t graph_full = nullptr;
    hipGraphExec_t exec_full = nullptr;
    uint32_t tile_logw = 0, tile_t1 = 0, max_a_nnz = 0;
    int tiles_ok = -1;
    void *win_scratch = nullptr;
    uint8_t *pc_kp = nullptr;  // piece lists of the long-row kernel
    uint32_t *pc_extra = nullptr, *pc_cnt = nullptr, *pc_next = nullptr, pc_cap = 0;
    uint32_t band_n = 0, band_nbands = 0, band_logw = 0;
    uint32_t *band_rows = nullptr, *band_ts = nullptr, *band_counter = nullptr;
    uint64_t *band_pref = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
  } plan;
  bool keep_plan = false;
  // per-bin shapes from the config (bin_concurrency / bin_block)
  Tables shapes{};
  bool custom_shapes = false;
  hipStream_t own_stream = nullptr;  // capture needs a non-default stream
  void drop_plan() {
    if (plan.exec) (void)hipGraphExecDestroy(plan.exec);
    if (plan.graph) (void)hipGraphDestroy(plan.graph);
    if (plan.exec_full) (void)hipGraphExecDestroy(plan.exec_full);
    if (plan.graph_full) (void)hipGraphDestroy(plan.graph_full);
    if (!plan.mem.empty()) (void)hipDeviceSynchronize();
    for (void *p : plan.mem) (void)hipFree(p);
    plan = Plan();
  }
  ChunkStore mem;  // temporaries of the runs (see ChunkStore)
  ~oc_spgemm() {
    if (!mem.chunks.empty()) (void)hipDeviceSynchronize();
    mem.drop();
    drop_plan();
    if (own_stream) (void)hipStreamDestroy(own_stream);
    oc_mlp_destroy(owned_model);
  }
};

namespace {

// ------------------------------------------------------------ decisions
// SpGEMM.cuh:455-540
void decide_analysis(const oc_config &c, const oc_spgemm &h, Decision &d) {
  d = Decision();
  if (h.a_cols / (uint64_t)c.fat_ratio > h.a_rows &&
      h.b_rows / (uint64_t)c.fat_ratio > h.b_cols &&
      h.b_cols < tables().dense_num[5] && c.use_fat_workflow) {
    d.ana1_type = 2;
    return;
  }
  const double avg = (double)h.total_products / (double)h.a_rows;
  if (avg <= (double)c.usparse_avg_product_threshold && c.use_ultrasparse_workflow) {
    static const int sizes[4] = {16, 32, 64, 128};
    d.ana1_type = 0;
    for (int k : sizes)
      if (k >= avg) {
        d.spark_kernel_size = k;
        break;
      }
    if ((int64_t)h.max_products > d.spark_kernel_size) {
      d.spark_need_outlier = 1;
      d.spark_outlier_kernel_size = (int)((h.max_products + 31) / 32 * 32);
    }
  } else {
    const uint64_t ratio = h.a_nnz ? h.total_products / h.a_nnz : 0;
    d.ana1_type = 1;
    if (ratio < (uint64_t)c.hll_higher_precision_expansion_threshold) {
      d.log_p = c.hll_precision_low;
      d.hll_expansion = c.hll_expansion_low;
    } else {
      d.log_p = c.hll_precision_high;
      d.hll_expansion = c.hll_expansion_high;
    }
  }
}

int num_samples(const oc_config &c, uint64_t rows) {
  int n = (int)(rows * c.hll_sample_rate);  // SpGEMM.cuh:587-590
  n = std::max(n, c.hll_sample_min_rows);
  n = std::min(n, c.hll_sample_max_rows);
  return n;
}

// SpGEMM.cuh:610-686 and decideMatrixType (Analysis.cuh:10-29).  Device
// sums of the reference are taken here in sample order.
void decide_sampling(const oc_config &c, const oc_spgemm &h,
                     const std::vector<uint32_t> &prod,
                     const std::vector<int32_t> &est, Decision &d) {
  const int n = (int)prod.size();
  uint64_t sum_est = 0, sum_prod = 0;
  double sum_comp = 0.0;
  for (int i = 0; i < n; ++i) {
    sum_est += (uint64_t)(int64_t)est[i];
    sum_prod += prod[i];
    sum_comp += (double)prod[i] / (double)est[i];
  }
  d.nsamples = n;
  d.comp1 = (double)sum_prod / (double)sum_est;
  d.comp2 = sum_comp / n;
  // index_t parameters: total_products and A.nnz arrive truncated to 32 bits
  const uint32_t tp32 = (uint32_t)h.total_products, an32 = (uint32_t)h.a_nnz;
  int type = 1;
  if (an32 == 0 || tp32 / an32 < (uint32_t)c.estimation_input_expansion_threshold) type = 0;
  else if (d.comp1 < c.estimation_output_compaction_threshold) type = 0;
  else if (d.comp2 < c.estimation_output_compaction_threshold) type = 0;
  if (!c.use_estimation_workflow) type = 0;
  d.ana2_type = type;
  d.dense_query_bitmap = !((d.comp1 < c.dense_query_bitmap_compaction_threshold &&
                            d.comp2 < c.dense_query_bitmap_compaction_threshold) ||
                           !c.use_assist_information);
  if (type == 1) return;
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += (double)prod[i] / (double)est[i];
  d.comp_avg = s / n;
  double v = 0.0;
  for (int i = 0; i < n; ++i) {
    const double x = (double)prod[i] / (double)est[i] - d.comp_avg;
    v += x * x;
  }
  d.comp_var = v / (n - 1);
  d.safe_comp = (d.comp_avg - c.estimate_compaction_target_z_value * std::sqrt(d.comp_var)) / 2;
  if (d.safe_comp > d.comp_avg * 0.8) d.safe_comp = d.comp_avg * 0.8;
  d.use_est_symbolic = d.safe_comp >= 1.5 && c.use_assist_information;
}

// ------------------------------------------------------------- run state
struct Run {
  oc_spgemm *h;
  TablesScope tscope;  // the handle's bin shapes for every launch of this run
  const oc_config &c;
  hipStream_t s;
  Arena ar;
  Marks mk;
  DevCsr A, B;
  RowInfo ri{};
  uint64_t *psum_prod = nullptr;
  uint32_t *scal = nullptr;
  void *scan_tmp = nullptr;
  uint8_t *hll_b = nullptr;
  uint32_t *learned_est = nullptr;  // learned capacities (learned_capacities)
  void *win_scratch = nullptr;
  // piece lists of the long-row kernel: kp for pc_cap list entries
  uint8_t *pc_kp = nullptr;
  uint32_t *pc_extra = nullptr, *pc_cnt = nullptr, *pc_next = nullptr, pc_cap = 0;
  uint32_t *tile_ts = nullptr, *tile_tsw = nullptr;
  const uint32_t *block_base = nullptr;  // per-B-row block bases of the tables
  uint32_t tile_logw = 0, tile_t1 = 0;
  int tiles_ok = -1;
  // row bitmaps stored by the dense symbolic pass (one-gather rank tiles)
  void *row_bm = nullptr;
  const uint64_t *row_bmoff = nullptr;
  void *ana_tmp = nullptr;  // analysis temporaries (kept for a plan)
  void *bin_tmp = nullptr;  // temporaries of the last binning
  uint32_t tsw_log = 0, tsw_per = 0;
  RowSel sel_big{};         // device row range of the long rows (whole-path replay)
  // long rows in column bands (precise workflow)
  struct Band {
    uint32_t n = 0, nbands = 0, logw = 0;
    uint32_t *flag = nullptr, *rows = nullptr, *ts = nullptr, *cnt = nullptr, *counter = nullptr;
    uint64_t *pref = nullptr;
  } band;
  oc_csr *C;
  struct Fam {
    const char *name, *stage;
    bool numeric;
    const uint32_t *rows;  // null: all rows
    uint32_t n;
  };
  std::vector<Fam> fams;
  bool replay = false;  // value-only replay of a kept plan: no marks, no accounting
  void fam(const char *name, const char *stage, bool numeric, const uint32_t *rows,
           uint32_t n) {
    if (c.track_stage_times && n && !replay) fams.push_back(Fam{name, stage, numeric, rows, n});
  }
  Run(oc_spgemm *hh, hipStream_t st, oc_csr *cc)
      : h(hh), tscope(hh->custom_shapes ? &hh->shapes : nullptr), c(hh->cfg), s(st),
        ar(&hh->mem), mk(st), C(cc) {}

  AccArgs acc() const {
    AccArgs a;
    a.A = A;
    a.B = B;
    a.products = ri.products;
    a.max_b = ri.max_b;
    a.left = ri.left;
    a.right = ri.right;
    return a;
  }

  // the row-claim counter of persistent bin launches (zeroed per launch)
  uint32_t *claim_ctr = nullptr;
  int claims() {
    if (!claim_ctr) OC_TRY(ar.get(&claim_ctr, 64));
    return OC_OK;
  }

  int window_scratch() {
    if (!win_scratch) {
      uint8_t *p;
      OC_TRY(ar.get(&p, window_scratch_bytes(h->max_a_nnz)));
      win_scratch = p;
    }
    return OC_OK;
  }

  // Tile-start table of B for the tiled numeric kernel (built once).  For
  // a block-diagonal product (set by the batched layer) the tiles of each
  // B row are counted from its block's first column, so the table spans the
  // widest block instead of all of B's columns; only the rank-tile kernel
  // reads such a table.
  int tiles() {
    if (tiles_ok >= 0) return OC_OK;
    size_t bytes = 0;
    // the bases apply only to the B they were set for
    const bool blocks = h->b_block_base != nullptr && h->b_block_rp == (const void *)B.rp;
    block_base = blocks ? h->b_block_base : nullptr;
    const uint64_t width = blocks ? h->b_block_span : B.cols;
    const int mode = h->cfg.use_tiled_numeric;
    tiles_ok = (mode == 2 || mode == 3 || (mode == 1 && !blocks)) &&
               tile_plan(B.rows, width, (uint32_t)h->cfg.tile_log2_cols,
                         (uint32_t)h->cfg.tile_block, &tile_logw, &tile_t1, &bytes);
    if (tiles_ok) {
      OC_TRY(ar.get(&tile_ts, bytes / 4));
      OC_TRY(launch_tile_starts(B, tile_logw, tile_t1, tile_ts, s, block_base));
      size_t per = 0;
      // the one-gather kernel runs 1024- or 512-thread workgroups
      const uint32_t bs = mode == 3 ? (h->cfg.tile_block == 512 ? 512u : 1024u)
                                    : (uint32_t)h->cfg.tile_block;
      const uint32_t wlog = mode == 3 ? rank1_wide_log(bs) : rank_wide_log(bs);
      if (mode >= 2 && wide_starts_for_log(width, wlog, &per)) {
        OC_TRY(ar.get(&tile_tsw, B.rows * per));
        tsw_log = wlog;
        tsw_per = (uint32_t)per;
        OC_TRY(launch_tile_starts(B, tsw_log, tsw_per, tile_tsw, s, block_base));
      }
      mk.mark("numeric_tile_starts");
    }
    return OC_OK;
  }

  // Row bitmaps of the windowed and dense symbolic bins (rows with at
  // least kBitmapMinProducts products), stored by the symbolic pass for the
  // one-gather rank tiles of the numeric pass.  The bins' rows are
  // [hoff[6], hoff[14]) (bin 7 is never used).  Skipped for block-diagonal
  // batches and when the bitmaps would take more than a fifth of the device.
  int row_bitmaps(const uint32_t *brows, const uint32_t *hoff) {
    constexpr uint32_t kBitmapMinProducts = 2048;
    h->bitmap_bytes = 0;
    const bool blocks = h->b_block_base != nullptr && h->b_block_rp == (const void *)B.rp;
    if (h->cfg.use_tiled_numeric != 3 || blocks || hoff[14] <= hoff[6]) return OC_OK;
    const uint32_t nd = hoff[14] - hoff[6];
    uint64_t *bmoff, *offs;
    uint32_t *sz;
    OC_TRY(ar.get(&bmoff, A.rows));
    OC_TRY(ar.get(&sz, nd));
    OC_TRY(ar.get(&offs, (uint64_t)nd + 1));
    OC_TRY(launch_bitmap_offsets(brows + hoff[6], nd, A.rows, ri.products, ri.left, ri.right,
                                 kBitmapMinProducts, sz, offs, scan_tmp, bmoff, s));
    uint64_t n4 = 0;
    OC_TRY(readback(&n4, offs + nd, 1, s));
    size_t fr = 0, tot = 0;
    HIP_TRY(hipMemGetInfo(&fr, &tot));
    if (n4 == 0 || (double)n4 * 16 > 0.2 * (double)tot) return OC_OK;
    uint8_t *bm;
    OC_TRY(ar.get(&bm, n4 * 16));
    row_bm = bm;
    row_bmoff = bmoff;
    h->bitmap_bytes = n4 * 16;
    return OC_OK;
  }

  // sel: the device row range of a whole-path replay (n is then unused)
  int big_rows(const uint32_t *rows, uint32_t n, const NumOut &o, const RowSel &sel = RowSel{}) {
    if (!n && !sel.bounds) return OC_OK;
    OC_TRY(window_scratch());
    OC_TRY(tiles());
    if (tiles_ok && h->cfg.use_tiled_numeric == 3) {
      const uint32_t bs = h->cfg.tile_block == 512 ? 512u : 1024u;
      // rows of more than kPieceProducts products are split over several
      // workgroups (eager runs without block bases)
      constexpr uint32_t kPieceProducts = 1u << 21, kMaxPieces = 1u << 16;
      R1Pieces pc;
      if (!block_base && !sel.bounds && n) {
        if (n > pc_cap && !replay) {  // a replay reuses the plan's lists
          OC_TRY(ar.get(&pc_kp, n));
          if (!pc_extra) {
            OC_TRY(ar.get(&pc_extra, 2 * kMaxPieces));
            OC_TRY(ar.get(&pc_cnt, kMaxPieces));
            OC_TRY(ar.get(&pc_next, 1));
          }
          pc_cap = n;
        }
        if (n <= pc_cap) {
          pc.kp = pc_kp;
          pc.extra = pc_extra;
          pc.n_extra = pc_next;
          pc.pcnt = pc_cnt;
          pc.max_extra = kMaxPieces;
          pc.per_piece = kPieceProducts;
        }
      }
      return launch_num_rank1(acc(), rows, n, tile_logw, tile_ts, tile_t1,
                              tsw_log == rank1_wide_log(bs) ? tile_tsw : nullptr, tsw_per,
                              block_base, row_bm, row_bmoff, o, win_scratch, s, sel, bs, pc);
    }
    if (tiles_ok && h->cfg.use_tiled_numeric == 2)
      return launch_num_rank(acc(), rows, n, tile_logw, tile_ts, tile_t1, tile_tsw, tsw_per,
                             block_base, (uint32_t)h->cfg.tile_products_cap,
                             (uint32_t)h->cfg.tile_block, o,
                             win_scratch, s, sel);
    if (tiles_ok) {
      if (sel.bounds) {
        set_error("the dense-tile long-row kernel (use_tiled_numeric 1) has no "
                  "whole-path replay");
        return OC_ERR_INVALID;
      }
      return launch_num_tiles(acc(), rows, n, tile_logw, (uint32_t)h->cfg.tile_block, tile_ts,
                              tile_t1, o, win_scratch, s);
    }
    return launch_num_window(acc(), rows, n, h->max_a_nnz, o, win_scratch, s, sel);
  }

  // C: u64 offsets (rows+1) from per-row counts, then cols/values.
  int alloc_c_from_counts(const uint32_t *nnz) {
    const uint64_t m = A.rows;
    std::memset(C, 0, sizeof(*C));
    C->rows = m;
    C->cols = B.cols;
    C->vtype = A.vtype;
    C->off_bytes = 8;
    C->on_device = 1;
    C->owned = 3;  // output cache (out_alloc)
    OC_TRY(out_alloc(&C->row_offsets, (m + 1) * 8));
    OC_TRY(launch_scan_u32(nnz, (uint64_t *)C->row_offsets, m, scan_tmp, s));
    mk.mark("epilogue_scan");
    uint64_t nnzc = 0;
    OC_TRY(readback(&nnzc, (const uint64_t *)C->row_offsets + m, 1, s));
    C->nnz = nnzc;
    OC_TRY(out_alloc((void **)&C->col_ids, std::max<uint64_t>(nnzc, 1) * 4));
    OC_TRY(out_alloc(&C->values, std::max<uint64_t>(nnzc, 1) * vsize(A.vtype)));
    mk.mark("numeric_malloc");
    return OC_OK;
  }

  int binning(const BinParams &bp, uint8_t **bor, uint32_t **brows,
              uint32_t **boff, uint32_t *hoff) {
    const uint64_t m = A.rows;
    if (!*bor) {
      OC_TRY(ar.get(bor, m));
      OC_TRY(ar.get(brows, m));
      OC_TRY(ar.get(boff, kBinSlots + 1));
    }
    uint8_t *tmp;
    OC_TRY(ar.get(&tmp, bin_temp_bytes(m)));
    bin_tmp = tmp;
    OC_TRY(launch_binning(bp, m, *bor, *brows, *boff, tmp, s));
    OC_TRY(readback(hoff, *boff, kBinSlots + 1, s));
    return OC_OK;
  }

  // Select the long rows for the band kernels and build their tables.
  int band_setup() {
    const uint64_t m = A.rows;
    band = Band();
    const uint32_t lw = (uint32_t)std::max(0, c.band_log2_cols);
    uint32_t nb = 0;
    if (lw == 0 || !band_plan(B.cols, lw, &nb)) return OC_OK;
    uint32_t *flag, *list;
    uint64_t *pos;
    OC_TRY(ar.get(&flag, m));
    OC_TRY(ar.get(&pos, m + 1));
    OC_TRY(ar.get(&list, m));
    OC_TRY(launch_band_select(ri, m, (uint32_t)std::max(1, c.band_min_products), 1u << lw, flag,
                              pos, scan_tmp, list, s));
    uint64_t n = 0;
    OC_TRY(readback(&n, pos + m, 1, s));
    if (n == 0) return OC_OK;
    band.n = (uint32_t)n;
    band.nbands = nb;
    band.logw = lw;
    band.flag = flag;
    band.rows = list;
    OC_TRY(ar.get(&band.ts, (uint64_t)(nb + 1) * B.rows));
    OC_TRY(launch_band_starts(B, lw, nb, band.ts, s));
    OC_TRY(ar.get(&band.cnt, n * nb));
    OC_TRY(ar.get(&band.pref, n * nb + 1));
    OC_TRY(ar.get(&band.counter, 64));
    mk.mark("band_setup");
    return OC_OK;
  }

  // Exact counts of the long rows (per band, then per row) into nnz.
  int band_symbolic(uint32_t *nnz) {
    if (!band.n) return OC_OK;
    const uint64_t cells = (uint64_t)band.n * band.nbands;
    OC_TRY(launch_band_count(acc(), band.rows, band.n, band.logw, band.nbands, band.ts,
                             band.cnt, band.counter, s));
    uint8_t *tmp;
    OC_TRY(ar.get(&tmp, scan_temp_bytes(cells)));
    OC_TRY(launch_scan_u32(band.cnt, band.pref, cells, tmp, s));
    OC_TRY(launch_band_row_nnz(band.rows, band.n, band.nbands, band.pref, nnz, s));
    fam("band_count", "symbolic_band", false, band.rows, band.n);
    mk.mark("symbolic_band");
    return OC_OK;
  }

  int band_numeric(const NumOut &o) {
    if (!band.n) return OC_OK;
    OC_TRY(launch_band_numeric(acc(), band.rows, band.n, band.logw, band.nbands, band.ts,
                               band.pref, o, band.counter, s));
    fam("band_numeric", "numeric_band", true, band.rows, band.n);
    mk.mark("numeric_band");
    return OC_OK;
  }

  // Numeric kernels over numeric bins 1..13 (bin 13 = windowed).
  // boff (optional): the bin offsets on the device; the hash and dense
  // bins then run persistent, claiming rows in chunks (claim_ctr)
  int numeric_bins(const uint32_t *brows, const uint32_t *hoff, const NumOut &o,
                   bool window_bin13, const uint32_t *boff = nullptr) {
    const AccArgs a = acc();
    if (boff) OC_TRY(claims());
    for (int b = 0; b <= 6; ++b) {
      const uint32_t n = hoff[b + 1] - hoff[b];
      if (!n) continue;
      if (b == 0) {
        set_error("internal: numeric hash bin 0 is not used");
        return OC_ERR_INVALID;
      }
      if (boff)
        OC_TRY(launch_num_hash(b, a, brows, n, o, s, RowSel{boff + b, nullptr, claim_ctr, hoff[b]}));
      else
        OC_TRY(launch_num_hash(b, a, brows + hoff[b], n, o, s));
    }
    fam("num_hash", "numeric_hash", true, brows + hoff[1], hoff[7] - hoff[1]);
    mk.mark("numeric_hash");
    for (int b = 7; b <= 12; ++b) {
      const uint32_t n = hoff[b + 1] - hoff[b];
      if (!n) continue;
      if (boff)
        OC_TRY(launch_num_dense(b - 7, a, brows, n, o, s,
                                RowSel{boff + b, nullptr, claim_ctr, hoff[b]}));
      else
        OC_TRY(launch_num_dense(b - 7, a, brows + hoff[b], n, o, s));
    }
    fam("num_dense", "numeric_dense", true, brows + hoff[7], hoff[13] - hoff[7]);
    mk.mark("numeric_dense");
    if (window_bin13) {
      const uint32_t n = hoff[14] - hoff[13];
      if (n) OC_TRY(tiles());
      mk.mark("numeric_tile_starts");
      OC_TRY(big_rows(brows + hoff[13], n, o));
      fam("num_window", "numeric_window", true, brows + hoff[13], n);
      mk.mark("numeric_window");
    }
    return OC_OK;
  }

  int prologue_and_analysis();
  int sampling();
  int learned_capacities(bool *use_est);
  int segmentation();
  int run_fat();
  int run_usparse();
  int run_precise();
  int run_est();
};

int Run::prologue_and_analysis() {
  const uint64_t m = A.rows;
  // SpGEMM.cuh:360-363
  if (c.use_hybrid_accumulator && !c.use_indirect_sort) {
    set_error("Hybrid accumulator is enabled, but indirect sort is disabled. Invalid configuration");
    return OC_ERR_INVALID;
  }
  OC_TRY(ar.get(&ri.products, m));
  OC_TRY(ar.get(&ri.max_b, m));
  OC_TRY(ar.get(&ri.avg_b, m));
  OC_TRY(ar.get(&ri.left, m));
  OC_TRY(ar.get(&ri.right, m));
  OC_TRY(ar.get(&psum_prod, m + 1));
  OC_TRY(ar.get(&scal, 16));
  uint8_t *st;
  OC_TRY(ar.get(&st, scan_temp_bytes(m + 1)));
  scan_tmp = st;
  // hybrid (largest) numeric hash table rule, SpGEMM.cuh:434-441
  const uint32_t idx_bits6 = 32 - bits_needed(tables().hash_num[6]);
  h->d.hash_largest = 0;
  const bool largest = !((uint32_t)bits_needed(B.cols ? B.cols - 1 : 0) > idx_bits6 ||
                         !c.use_hybrid_accumulator) &&
                       B.cols > (uint64_t)c.numeric_hybrid_hash_kernel_row_length_threshold;
  mk.mark("prologue");
  uint8_t *atmp;
  OC_TRY(ar.get(&atmp, analysis_temp_bytes(m, A.nnz, B.rows)));
  OC_TRY(launch_analysis(A, B, ri, scal, scal + 1, atmp, s));
  ana_tmp = atmp;
  // stream-ordered: released once the analysis has run, unless a plan
  // keeps it for whole-path replays
  if (!h->keep_plan) ar.free(atmp);
  mk.mark("analysis_product_calc");
  OC_TRY(launch_scan_u32(ri.products, psum_prod, m, scan_tmp, s));
  mk.mark("analysis_reduce");
  uint32_t sc[2];
  OC_TRY(readback(sc, scal, 2, s));
  OC_TRY(readback(&h->total_products, psum_prod + m, 1, s));
  h->max_a_nnz = sc[0];
  h->max_products = sc[1];
  mk.mark("analysis_mem_cpy");
  decide_analysis(c, *h, h->d);
  h->d.hash_largest = largest;
  return OC_OK;
}

int Run::sampling() {
  Decision &d = h->d;
  const int P = 1 << d.log_p;
  OC_TRY(ar.get(&hll_b, B.rows * (uint64_t)P));
  mk.mark("est_malloc");
  OC_TRY(launch_hll_build(d.log_p, B, hll_b, s));
  mk.mark("est_hll_construct");
  const int ns = num_samples(c, A.rows);
  uint32_t *srows, *sprod;
  int32_t *sest;
  uint8_t *tmp;
  OC_TRY(ar.get(&srows, ns));
  OC_TRY(ar.get(&sprod, ns));
  OC_TRY(ar.get(&sest, ns));
  OC_TRY(ar.get(&tmp, hll_merge_temp_bytes(ns)));
  OC_TRY(launch_hll_merge(d.log_p, A, hll_b, 1.0, nullptr, (uint32_t)ns, srows, sest,
                          ri.products, sprod, 1, tmp, s));
  std::vector<uint32_t> hp(ns);
  std::vector<int32_t> he(ns);
  OC_TRY(readback(hp.data(), sprod, ns, s));
  OC_TRY(readback(he.data(), sest, ns, s));
  mk.mark("est_sample");
  const int hl = d.hash_largest;
  decide_sampling(c, *h, hp, he, d);
  d.hash_largest = hl;
  mk.mark("sampling_var_calc_host");
  return OC_OK;
}

// Learned capacities of every row, and the workflow they pick.  The
// estimation workflow saves the symbolic pass (a second sweep over all
// products) but adds a copy of C from the estimated slots to its final
// place; it pays when the products outnumber the predicted output by at
// least learned_min_compression.
int Run::learned_capacities(bool *use_est) {
  const uint64_t m = A.rows;
  OC_TRY(ar.get(&learned_est, m + 1));
  float *feat;
  OC_TRY(ar.get(&feat, m * OC_MLP_FEATURES));
  OC_TRY(launch_mlp_features(A, B, ri, feat, s));
  OC_TRY(launch_mlp_estimates(h->model, feat, m, ri, h->model_q, learned_est, s));
  ar.free(feat);
  uint64_t *ps;
  OC_TRY(ar.get(&ps, m + 1));
  OC_TRY(launch_scan_u32(learned_est, ps, m, scan_tmp, s));
  OC_TRY(readback(&h->learned_total, ps + m, 1, s));
  ar.free(ps);
  h->learned_consulted = true;
  *use_est = (double)h->total_products >=
             c.learned_min_compression * (double)std::max<uint64_t>(h->learned_total, 1);
  mk.mark("est_hll_merge");
  return OC_OK;
}

// Spy-plot segmentation of A (Segmentation.cuh:374-673): analysis only,
// reported in the stats JSON.  Grid sizes, thresholds, the bounded label
// propagation and flattening, the run-length encoding of the per-grid-row
// labels, the merge of short bands and the class rule follow the reference;
// the propagation sweeps are synchronous (deterministic) and dominant-label
// ties go to the smaller label.
int Run::segmentation() {
  oc_spgemm::Segmentation &sg = h->seg;
  sg = oc_spgemm::Segmentation();
  sg.ran = true;
  const uint64_t m = A.rows;
  if (m == 0 || B.cols == 0) return OC_OK;
  const uint64_t gmax = (uint64_t)std::max(8, c.segmentation_grid_max);
  OC_REQUIRE(gmax <= 4096, "segmentation_grid_max above 4096");
  sg.grid_h = (uint32_t)std::min<uint64_t>(gmax, m);
  sg.grid_w = (uint32_t)std::min<uint64_t>(gmax, B.cols);
  sg.tile_rows = (uint32_t)((m + sg.grid_h - 1) / sg.grid_h);
  sg.tile_cols = (uint32_t)((B.cols + sg.grid_w - 1) / sg.grid_w);
  const uint32_t n = sg.grid_h * sg.grid_w;
  hipEvent_t ev[5];
  for (hipEvent_t &e : ev) HIP_TRY(hipEventCreate(&e));
  struct Ev {
    hipEvent_t *e;
    ~Ev() {
      for (int i = 0; i < 5; ++i) (void)hipEventDestroy(e[i]);
    }
  } ev_guard{ev};
  uint32_t *density, *labels, *next, *row_label, *row_hot;
  unsigned long long *out;
  OC_TRY(ar.get(&density, n));
  OC_TRY(ar.get(&labels, n));
  OC_TRY(ar.get(&next, n));
  OC_TRY(ar.get(&row_label, sg.grid_h));
  OC_TRY(ar.get(&row_hot, sg.grid_h));
  OC_TRY(ar.get(&out, 4));
  HIP_TRY(hipMemsetAsync(density, 0, (size_t)n * 4, s));
  HIP_TRY(hipEventRecord(ev[0], s));
  OC_TRY(launch_seg_density(A, sg.grid_h, sg.grid_w, sg.tile_rows, sg.tile_cols, density, s));
  HIP_TRY(hipEventRecord(ev[1], s));
  OC_TRY(launch_seg_threshold(density, n, c.segmentation_hot_threshold_k, labels, out, s));
  HIP_TRY(hipEventRecord(ev[2], s));
  OC_TRY(launch_seg_components(sg.grid_h, sg.grid_w,
                               (uint32_t)std::max(0, c.segmentation_cc_max_iters), labels, next,
                               out, s));
  HIP_TRY(hipEventRecord(ev[3], s));
  OC_TRY(launch_seg_dominant(labels, sg.grid_h, sg.grid_w, row_label, row_hot, s));
  unsigned long long ho[3] = {0, 0, 0};
  std::vector<uint32_t> band_label(sg.grid_h);
  OC_TRY(readback(ho, out, 3, s));
  OC_TRY(readback(band_label.data(), row_label, sg.grid_h, s));
  const double mean = (double)ho[0] / (double)n;
  double var = (double)ho[1] / (double)n - mean * mean;
  if (var < 0.0) var = 0.0;
  sg.density_mean = mean;
  sg.density_stddev = std::sqrt(var);
  sg.cc_iterations = ho[2];
  // run-length encoding of the grid-row labels into row bands
  struct Band {
    uint32_t start, end, label;
  };
  std::vector<Band> bands;
  uint32_t cur = band_label[0], cur_g = 0;
  auto push = [&](uint32_t g_begin, uint32_t g_end, uint32_t label) {
    const uint64_t b0 = (uint64_t)g_begin * sg.tile_rows;
    const uint64_t b1 = std::min<uint64_t>(m, (uint64_t)g_end * sg.tile_rows);
    if (b1 > b0) bands.push_back(Band{(uint32_t)b0, (uint32_t)b1, label});
  };
  for (uint32_t g = 1; g < sg.grid_h; ++g)
    if (band_label[g] != cur) {
      push(cur_g, g, cur);
      cur = band_label[g];
      cur_g = g;
    }
  {
    const uint64_t b0 = (uint64_t)cur_g * sg.tile_rows;
    if (m > b0) bands.push_back(Band{(uint32_t)b0, (uint32_t)m, cur});
  }
  // bands shorter than segmentation_min_segment_rows join the previous one
  const uint32_t min_rows = (uint32_t)std::max(1, c.segmentation_min_segment_rows);
  std::vector<Band> merged;
  for (const Band &b : bands) {
    if (!merged.empty() && b.end - b.start < min_rows) merged.back().end = b.end;
    else merged.push_back(b);
  }
  std::vector<uint32_t> uniq;
  for (const Band &b : merged)
    if (b.label != 0xFFFFFFFFu && std::find(uniq.begin(), uniq.end(), b.label) == uniq.end())
      uniq.push_back(b.label);
  sg.num_components = (uint32_t)uniq.size();
  if (merged.empty()) merged.push_back(Band{0, (uint32_t)m, 0xFFFFFFFFu});
  const uint32_t nseg = (uint32_t)merged.size();
  std::vector<uint32_t> hs(2 * (size_t)nseg);
  for (uint32_t i = 0; i < nseg; ++i) {
    hs[i] = merged[i].start;
    hs[nseg + i] = merged[i].end;
  }
  uint32_t *dseg;
  SegSums *dsum;
  OC_TRY(ar.get(&dseg, 2 * (uint64_t)nseg));
  OC_TRY(ar.get(&dsum, nseg));
  HIP_TRY(hipMemcpyAsync(dseg, hs.data(), hs.size() * 4, hipMemcpyHostToDevice, s));
  OC_TRY(launch_seg_reduce(dseg, dseg + nseg, nseg, A, ri, dsum, s));
  std::vector<SegSums> sums(nseg);
  OC_TRY(readback(sums.data(), dsum, nseg, s));
  HIP_TRY(hipEventRecord(ev[4], s));
  HIP_TRY(hipEventSynchronize(ev[4]));
  (void)hipEventElapsedTime(&sg.t_density, ev[0], ev[1]);
  (void)hipEventElapsedTime(&sg.t_threshold, ev[1], ev[2]);
  (void)hipEventElapsedTime(&sg.t_cc, ev[2], ev[3]);
  (void)hipEventElapsedTime(&sg.t_classify, ev[3], ev[4]);
  // class rule (Segmentation.cuh:597-640): banded > dense > power law
  for (uint32_t i = 0; i < nseg; ++i) {
    const SegSums &t = sums[i];
    oc_spgemm::SegRecord rec{};
    rec.row_start = merged[i].start;
    rec.row_end = merged[i].end;
    rec.col_left = t.min_left == 0xFFFFFFFFu ? 0u : t.min_left;
    rec.col_right = t.max_right;
    rec.sum_products = t.sum_products;
    rec.sum_a_nnz = t.sum_a_nnz;
    rec.max_b_row_len = t.max_b_row;
    rec.max_a_row_len = t.max_a_row;
    const uint32_t rows = rec.row_end - rec.row_start;
    const double mean_p = rows ? (double)t.sum_products / (double)rows : 0.0;
    double var_p = 0.0;
    if (rows) {
      var_p = (double)t.sum_sq_products / (double)rows - mean_p * mean_p;
      if (var_p < 0.0) var_p = 0.0;
    }
    const double cv = mean_p > 0.0 ? std::sqrt(var_p) / mean_p : 0.0;
    const double er = t.sum_a_nnz ? (double)t.sum_products / (double)t.sum_a_nnz : 0.0;
    const double span_ratio =
        (double)(uint32_t)(rec.col_right + 1u - rec.col_left) / (double)B.cols;
    const double row_density = span_ratio > 0.0 ? mean_p / ((double)B.cols * span_ratio) : 0.0;
    rec.mean_row_len = mean_p;
    rec.var_row_len = var_p;
    rec.local_er = er;
    uint32_t cls = 0;  // sparse uniform
    if (span_ratio > 0.0 && span_ratio < 0.05 && mean_p > 0.0 &&
        er >= (double)c.estimation_input_expansion_threshold)
      cls = 3;  // banded
    else if (row_density >= 0.25)
      cls = 1;  // dense
    else if (cv >= 1.5)
      cls = 2;  // power law
    rec.cls = cls;
    sg.segments.push_back(rec);
  }
  mk.mark("segmentation");
  return OC_OK;
}

// FAT workflow (SpGEMM.cuh:925-1050): one dense pass over all of B's
// columns per A row, split over several workgroups per row (k_fat.hip).
int Run::run_fat() {
  const uint64_t m = A.rows;
  const uint64_t nwords = (B.cols + 31) / 32;
  double *dense;
  uint32_t *bits, *nnz;
  OC_TRY(ar.get(&dense, m * B.cols));
  OC_TRY(ar.get(&bits, m * nwords));
  OC_TRY(ar.get(&nnz, m + 1));
  HIP_TRY(hipMemsetAsync(dense, 0, m * B.cols * 8, s));
  HIP_TRY(hipMemsetAsync(bits, 0, m * nwords * 4, s));
  mk.mark("numeric_malloc");
  // about four workgroups per CU in all, each with at least 256 A entries
  const uint64_t ncu = std::max<uint32_t>(1, tables().num_cu);
  uint64_t splits = (4 * ncu + m - 1) / m;
  splits = std::max<uint64_t>(1, std::min<uint64_t>(splits, h->max_a_nnz / 256));
  // lanes per A entry from B's mean row length (SpGEMM.cuh:963-967)
  const double avg = B.rows ? (double)B.nnz / (double)B.rows : 0.0;
  int G = 1;
  while (G < 64 && (double)G < avg) G <<= 1;
  OC_TRY(launch_fat_accumulate(A, B, (uint32_t)splits, G, dense, bits, s));
  OC_TRY(launch_fat_count(bits, m, B.cols, nnz, s));
  fam("fat", "numeric_fat", true, nullptr, (uint32_t)m);
  mk.mark("numeric_fat");
  OC_TRY(alloc_c_from_counts(nnz));
  OC_TRY(launch_fat_compact(A.vtype, dense, bits, m, B.cols, (const uint64_t *)C->row_offsets,
                            C->col_ids, C->values, s));
  mk.mark("epilogue_copy");
  return OC_OK;
}

int Run::run_usparse() {
  const Decision &d = h->d;
  const uint64_t m = A.rows;
  uint32_t *tc, *nnz;
  uint8_t *tv;
  OC_TRY(ar.get(&tc, h->total_products));
  OC_TRY(ar.get(&tv, h->total_products * vsize(A.vtype)));
  OC_TRY(ar.get(&nnz, m + 1));
  mk.mark("numeric_malloc");
  const int esc_n = std::min(d.spark_kernel_size, 64);
  NumOut o;
  std::memset(&o, 0, sizeof(o));
  o.off = psum_prod;
  o.cols = tc;
  o.vals = tv;
  o.nnz = nnz;
  OC_TRY(launch_esc(esc_n, acc(), o, s));
  fam("esc_all_rows", "numeric_usparse_main", true, nullptr, (uint32_t)m);
  mk.mark("numeric_usparse_main");
  if (h->max_products > (uint32_t)esc_n) {
    uint32_t *ovf;
    OC_TRY(ar.get(&ovf, m + 4));
    o.ovf_rows = ovf + 4;
    o.ovf_count = ovf;
    HIP_TRY(hipMemsetAsync(ovf, 0, 16, s));
    BinParams bp;
    std::memset(&bp, 0, sizeof(bp));
    bp.kind = 1;
    bp.input = ri.products;
    bp.left = ri.left;
    bp.right = ri.right;
    bp.expand = 1;
    bp.expand_factor = c.numeric_expansion_coe;
    bp.filter = ri.products;
    bp.skip_below = (uint32_t)esc_n;
    uint8_t *bor = nullptr;
    uint32_t *brows = nullptr, *boff = nullptr;
    uint32_t hoff[kBinSlots + 1];
    OC_TRY(binning(bp, &bor, &brows, &boff, hoff));
    std::memcpy(h->bin2, hoff, sizeof(hoff));
    h->have_bin2 = true;
    OC_TRY(numeric_bins(brows, hoff, o, true, boff));
    uint32_t novf = 0;
    OC_TRY(readback(&novf, ovf, 1, s));
    if (novf) {
      set_error("internal: %u ultra-sparse outlier rows overflowed", novf);
      return OC_ERR_CAPACITY;
    }
    mk.mark("numeric_usparse_outlier");
  }
  OC_TRY(alloc_c_from_counts(nnz));
  OC_TRY(launch_copy_rows(A.vtype, m, nullptr, nnz, psum_prod, nullptr, tc, tv,
                          (const uint64_t *)C->row_offsets, C->col_ids, C->values, s));
  mk.mark("epilogue_copy");
  return OC_OK;
}

int Run::run_precise() {
  const Decision &d = h->d;
  const uint64_t m = A.rows;
  uint32_t *nnz, *ovf;
  OC_TRY(ar.get(&nnz, m + 1));
  OC_TRY(ar.get(&ovf, m + 8));
  HIP_TRY(hipMemsetAsync(ovf, 0, 32, s));
  uint32_t *ovf_rows = ovf + 8;
  BinParams bp;
  std::memset(&bp, 0, sizeof(bp));
  bp.kind = 0;
  bp.input = ri.products;
  bp.left = ri.left;
  bp.right = ri.right;
  bp.use_compaction = d.use_est_symbolic;
  bp.expand_factor = c.symbolic_expansion_coe;
  bp.avg_compaction = d.safe_comp;
  OC_TRY(band_setup());
  bp.skip_flag = band.flag;
  uint8_t *bor = nullptr;
  uint32_t *brows = nullptr, *boff = nullptr;
  uint32_t hoff[kBinSlots + 1];
  OC_TRY(binning(bp, &bor, &brows, &boff, hoff));
  std::memcpy(h->bin1, hoff, sizeof(hoff));
  h->have_bin1 = true;
  mk.mark("symbolic_binning");
  const AccArgs a = acc();
  const Tables &t = tables();
  OC_TRY(row_bitmaps(brows, hoff));
  OC_TRY(claims());
  for (int b = 0; b <= 5; ++b) {
    const uint32_t n = hoff[b + 1] - hoff[b];
    if (!n) continue;
    const uint32_t S = t.hash_sym[b];
    const uint32_t limit = d.use_est_symbolic ? S - S / 8 : S;
    OC_TRY(launch_sym_hash(b, a, brows, n, limit, nnz, ovf_rows, ovf, s,
                           RowSel{boff + b, nullptr, claim_ctr, hoff[b]}));
  }
  fam("sym_hash", "symbolic_hash", false, brows + hoff[0], hoff[6] - hoff[0]);
  mk.mark("symbolic_hash");
  if (hoff[7] > hoff[6]) {
    OC_TRY(window_scratch());
    OC_TRY(launch_sym_window(a, brows, hoff[7] - hoff[6], h->max_a_nnz, nnz, win_scratch, s,
                             RowSel{boff + 6, nullptr, claim_ctr, hoff[6]}, row_bmoff, row_bm));
  }
  fam("sym_window", "symbolic_window", false, brows + hoff[6], hoff[7] - hoff[6]);
  mk.mark("symbolic_window");
  for (int b = 7; b <= 13; ++b) {
    const uint32_t n = hoff[b + 1] - hoff[b];
    if (!n) continue;
    if (b == 7) {
      set_error("internal: dense symbolic bin 0 is not used");
      return OC_ERR_INVALID;
    }
    OC_TRY(launch_sym_dense(b - 7, a, brows, n, nnz, s, row_bmoff, row_bm,
                            RowSel{boff + b, nullptr, claim_ctr, hoff[b]}));
  }
  fam("sym_dense", "symbolic_dense", false, brows + hoff[7], hoff[14] - hoff[7]);
  mk.mark("symbolic_dense");
  {
    // Rows whose hash table filled up (the shrunken tables of the estimated
    // compaction, or a symbolic expansion that undersized a bin) are
    // recounted with the windowed bitmap, which has no size limit.
    uint32_t n = 0;
    OC_TRY(readback(&n, ovf, 1, s));
    h->sym_overflow_rows = n;
    if (n) {
      OC_TRY(window_scratch());
      OC_TRY(launch_sym_window(a, ovf_rows, n, h->max_a_nnz, nnz, win_scratch, s));
    }
    mk.mark("symbolic_overflow");
  }
  OC_TRY(band_symbolic(nnz));
  OC_TRY(alloc_c_from_counts(nnz));
  // numeric binning on exact counts, SpGEMM.cuh:1247-1283
  BinParams nb;
  std::memset(&nb, 0, sizeof(nb));
  nb.kind = 1;
  nb.input = nnz;
  nb.left = ri.left;
  nb.right = ri.right;
  nb.expand = 1;
  nb.expand_factor = c.numeric_expansion_coe;
  nb.use_largest = d.hash_largest;
  nb.hybrid_max = (uint32_t)(m / 10);
  nb.skip_flag = band.flag;
  uint8_t *bor2 = nullptr;
  uint32_t *brows2 = nullptr, *boff2 = nullptr;
  OC_TRY(binning(nb, &bor2, &brows2, &boff2, hoff));
  std::memcpy(h->bin2, hoff, sizeof(hoff));
  h->have_bin2 = true;
  mk.mark("numeric_binning");
  NumOut o;
  std::memset(&o, 0, sizeof(o));
  o.off = (const uint64_t *)C->row_offsets;
  o.cols = C->col_ids;
  o.vals = C->values;
  o.ovf_rows = ovf_rows;
  o.ovf_count = ovf + 1;
  o.nnz_hint = nnz;
  OC_TRY(numeric_bins(brows2, hoff, o, true, boff2));
  OC_TRY(band_numeric(o));
  // exact capacities cannot overflow; recount defensively if one did
  uint32_t n = 0;
  OC_TRY(readback(&n, ovf + 1, 1, s));
  h->precise_recount = n;
  OC_TRY(big_rows(ovf_rows, n, o));
  if (h->keep_plan) {
    // a whole-path replay needs the window scratch and the tile tables
    // even when this run had no rows for them
    OC_TRY(window_scratch());
    OC_TRY(tiles());
    oc_spgemm::Plan &p = h->plan;
    p.A = A;
    p.B = B;
    p.ri = ri;
    p.brows = brows2;
    std::memcpy(p.hoff, hoff, sizeof(hoff));
    p.o = o;
    p.tile_ts = tile_ts;
    p.tile_tsw = tile_tsw;
    p.block_base = block_base;
    p.row_bm = row_bm;
    p.row_bmoff = row_bmoff;
    p.tsw_log = tsw_log;
    p.tsw_per = tsw_per;
    p.scal = scal;
    p.ana_tmp = ana_tmp;
    p.scan_tmp = scan_tmp;
    p.bin_tmp = bin_tmp;
    p.bor1 = bor;
    p.brows1 = brows;
    p.boff1 = boff;
    p.bor2 = bor2;
    p.boff2 = boff2;
    p.nnz = nnz;
    p.ovf = ovf;
    p.cap_nnz = C->nnz;
    p.sym_bp = bp;  // with the sampled compaction when the run used it
    p.est_symbolic = d.use_est_symbolic;
    p.num_bp = nb;
    if (!p.err) OC_TRY(ar.get(&p.err, 64));
    p.tile_logw = tile_logw;
    p.tile_t1 = tile_t1;
    p.tiles_ok = tiles_ok;
    p.win_scratch = win_scratch;
    p.pc_kp = pc_kp;
    p.pc_extra = pc_extra;
    p.pc_cnt = pc_cnt;
    p.pc_next = pc_next;
    p.pc_cap = pc_cap;
    p.max_a_nnz = h->max_a_nnz;
    p.band_n = band.n;
    p.band_nbands = band.nbands;
    p.band_logw = band.logw;
    p.band_rows = band.rows;
    p.band_ts = band.ts;
    p.band_counter = band.counter;
    p.band_pref = band.pref;
    p.valid = n == 0;  // a recount would make the numeric phase two-step
  }
  return OC_OK;
}

int Run::run_est() {
  const Decision &d = h->d;
  const uint64_t m = A.rows;
  uint32_t *est, *nnz, *ovf;
  OC_TRY(ar.get(&est, m + 1));
  OC_TRY(ar.get(&nnz, m + 1));
  OC_TRY(ar.get(&ovf, m + 8));
  HIP_TRY(hipMemsetAsync(ovf, 0, 32, s));
  uint32_t *ovf_rows = ovf + 8;
  mk.mark("est_malloc");
  if (h->est_override) {
    HIP_TRY(hipMemcpyAsync(est, h->est_override, m * 4, hipMemcpyDeviceToDevice, s));
  } else if (h->learned_run) {
    est = learned_est;  // computed when the predictor picked this workflow
  } else {
    uint8_t *tmp;
    OC_TRY(ar.get(&tmp, hll_merge_temp_bytes(m)));
    OC_TRY(launch_hll_merge(d.log_p, A, hll_b, d.hll_expansion, est, 0, nullptr, nullptr,
                            nullptr, nullptr, 0, tmp, s));
  }
  mk.mark("est_hll_merge");
  BinParams bp;
  std::memset(&bp, 0, sizeof(bp));
  bp.kind = 1;
  bp.input = est;
  bp.input_mut = est;
  bp.left = ri.left;
  bp.right = ri.right;
  bp.modify_dense_est = 1;
  bp.use_largest = d.hash_largest;
  bp.hybrid_max = (uint32_t)(m / 10);
  // HyperLogLog estimates carry their expansion already; capacities given by
  // the caller or the predictor are counts, binned with the numeric
  // expansion like exact counts (a hash table of S slots takes 2S/3 keys)
  if (h->est_override || h->learned_run) {
    bp.expand = 1;
    bp.expand_factor = c.numeric_expansion_coe;
  }
  uint8_t *bor = nullptr;
  uint32_t *brows = nullptr, *boff = nullptr;
  uint32_t hoff[kBinSlots + 1];
  OC_TRY(binning(bp, &bor, &brows, &boff, hoff));
  std::memcpy(h->bin2, hoff, sizeof(hoff));
  h->have_bin2 = true;
  uint64_t *psum_est;
  OC_TRY(ar.get(&psum_est, m + 1));
  OC_TRY(launch_scan_u32(est, psum_est, m, scan_tmp, s));
  OC_TRY(readback(&h->total_est, psum_est + m, 1, s));
  mk.mark("numeric_binning");
  // estimated slots plus a spare region for the rows that outgrow theirs
  // (SpGEMM.cuh:1560-1587)
  const double coe = c.hll_numeric_malloc_expansion_coe;
  const uint64_t cap_total = std::max<uint64_t>((uint64_t)((double)h->total_est * coe), h->total_est);
  {
    // The estimated region and C are held at once (C is at least the
    // estimate's size).  When both would take more than 60% of the device,
    // the precise workflow, which holds only C, runs instead.
    size_t fr = 0, tot = 0;
    HIP_TRY(hipMemGetInfo(&fr, &tot));
    const double bytes = (double)(cap_total + h->total_est) * (double)(4 + vsize(A.vtype));
    if (bytes > 0.6 * (double)tot) {
      h->est_memory_fallback = true;
      h->d.ana2_type = 0;
      h->learned_run = false;
      return run_precise();
    }
  }
  uint32_t *tc;
  uint8_t *tv;
  OC_TRY(ar.get(&tc, cap_total));
  OC_TRY(ar.get(&tv, cap_total * vsize(A.vtype)));
  mk.mark("numeric_malloc");
  NumOut o;
  std::memset(&o, 0, sizeof(o));
  o.off = psum_est;
  o.cols = tc;
  o.vals = tv;
  o.nnz = nnz;
  o.cap = est;
  o.ovf_rows = ovf_rows;
  o.ovf_count = ovf;
  o.nnz_hint = est;
  OC_TRY(numeric_bins(brows, hoff, o, false, boff));
  // Long rows (bin 13): the rank-tile kernel checks the estimate and works
  // in place; the other long-row kernels have no capacity check, so their
  // rows go to the spare region with the overflow rows.
  const uint32_t n13 = hoff[14] - hoff[13];
  if (n13) OC_TRY(tiles());
  const bool inplace13 = n13 && tiles_ok && c.use_tiled_numeric >= 2;
  if (inplace13) {
    OC_TRY(big_rows(brows + hoff[13], n13, o));
    fam("num_window", "numeric_window", true, brows + hoff[13], n13);
    mk.mark("numeric_window");
  }
  uint32_t n1 = 0;
  OC_TRY(readback(&n1, ovf, 1, s));
  const uint32_t nmove = inplace13 ? 0u : n13;
  if (nmove)
    HIP_TRY(hipMemcpyAsync(ovf_rows + n1, brows + hoff[13], (size_t)nmove * 4,
                           hipMemcpyDeviceToDevice, s));
  const uint32_t novf = n1 + nmove;
  h->overflow_rows = n1;
  uint64_t *src_off;
  OC_TRY(ar.get(&src_off, m + 1));
  HIP_TRY(hipMemcpyAsync(src_off, psum_est, (m + 1) * 8, hipMemcpyDeviceToDevice, s));
  bool separate = false;  // overflow rows in a buffer of their own
  uint32_t *sc = nullptr;
  uint8_t *sv = nullptr;
  if (novf) {
    unsigned long long *used;
    OC_TRY(ar.get(&used, 1));
    HIP_TRY(hipMemsetAsync(used, 0, 8, s));
    OC_TRY(launch_assign_spare(ovf_rows, novf, ri.products, ri.left, ri.right,
                               h->total_est, used, src_off, s));
    unsigned long long hu = 0;
    OC_TRY(readback(&hu, used, 1, s));
    h->est_spare_need = hu;
    NumOut o2 = o;
    o2.cap = nullptr;
    o2.nnz_hint = nullptr;  // the estimate was too small for these rows
    o2.off = src_off;
    if (h->total_est + hu > cap_total) {
      // The spare region behind the estimated slots (SpGEMM.cuh:1583-1587
      // raises here) cannot hold the rows the estimates missed: they get a
      // buffer of exactly the room they reserve, offsets from 0.
      separate = true;
      HIP_TRY(hipMemsetAsync(used, 0, 8, s));
      OC_TRY(launch_assign_spare(ovf_rows, novf, ri.products, ri.left, ri.right, 0, used,
                                 src_off, s));
      OC_TRY(ar.get(&sc, hu));
      OC_TRY(ar.get(&sv, hu * vsize(A.vtype)));
      o2.cols = sc;
      o2.vals = sv;
    }
    OC_TRY(big_rows(ovf_rows, novf, o2));
  }
  h->est_separate_spare = separate;
  fam("num_overflow", "numeric_overflow", true, ovf_rows, novf);
  mk.mark("numeric_overflow");
  OC_TRY(alloc_c_from_counts(nnz));
  if (!separate) {
    OC_TRY(launch_copy_rows(A.vtype, m, nullptr, nnz, src_off, nullptr, tc, tv,
                            (const uint64_t *)C->row_offsets, C->col_ids, C->values, s));
  } else {
    // rows within their estimate from the estimated slots (a row that
    // outgrew its slot is skipped; listed rows that fit are overwritten
    // next), then the listed rows from their own buffer
    OC_TRY(launch_copy_rows(A.vtype, m, nullptr, nnz, psum_est, est, tc, tv,
                            (const uint64_t *)C->row_offsets, C->col_ids, C->values, s));
    OC_TRY(launch_copy_rows(A.vtype, novf, ovf_rows, nnz, src_off, nullptr, sc, sv,
                            (const uint64_t *)C->row_offsets, C->col_ids, C->values, s));
  }
  mk.mark("epilogue_copy");
  return OC_OK;
}

// ------------------------------------------------------------------ JSON
struct Json {
  std::ostringstream o;
  bool first = true;
  void key(const char *k) {
    if (!first) o << ",";
    first = false;
    o << "\"" << k << "\":";
  }
  void num(const char *k, double v) {
    key(k);
    if (std::isfinite(v)) {
      o.precision(17);
      o << v;
    } else {
      o << "null";
    }
  }
  void u64(const char *k, uint64_t v) {
    key(k);
    o << v;
  }
  void boolean(const char *k, bool v) {
    key(k);
    o << (v ? "true" : "false");
  }
  void open(const char *k) {
    key(k);
    o << "{";
    first = true;
  }
  void close() {
    o << "}";
    first = false;
  }
  void str(const char *k, const char *v) {
    key(k);
    o << "\"" << v << "\"";
  }
  // a list of objects: list_open, then item_open ... item_close per item
  bool first_item = true;
  void list_open(const char *k) {
    key(k);
    o << "[";
    first_item = true;
  }
  void item_open() {
    if (!first_item) o << ",";
    first_item = false;
    o << "{";
    first = true;
  }
  void item_close() { o << "}"; }
  void list_close() {
    o << "]";
    first = false;
  }
  void arr(const char *k, const uint32_t *v, int n) {
    key(k);
    o << "[";
    for (int i = 0; i < n; ++i) o << (i ? "," : "") << v[i];
    o << "]";
  }
};

double stage(const oc_spgemm *h, const char *k) {
  auto it = h->stage_ms.find(k);
  return it == h->stage_ms.end() ? 0.0 : it->second;
}

std::string stats_json(const oc_spgemm *h) {
  Json j;
  j.o << "{";
  j.open("A");
  j.u64("rows", h->a_rows);
  j.u64("cols", h->a_cols);
  j.u64("nnz", h->a_nnz);
  j.close();
  j.open("B");
  j.u64("rows", h->b_rows);
  j.u64("cols", h->b_cols);
  j.u64("nnz", h->b_nnz);
  j.close();
  j.open("C");
  j.u64("rows", h->c_rows);
  j.u64("cols", h->c_cols);
  j.u64("nnz", h->c_nnz);
  j.close();
  j.open("stats");
  j.u64("intermediate_product", h->total_products);
  j.u64("intermediate_product_max", h->max_products);
  j.num("intermediate_product_avg", (double)h->total_products / (double)h->a_rows);
  j.num("input_expansion", (double)h->total_products / (double)h->a_nnz);
  j.num("output_compression", (double)h->total_products / (double)h->c_nnz);
  j.close();
  const Decision &d = h->d;
  int workflow = 0;  // JSON ids of SpGEMM.cuh:1807-1817
  if (d.ana1_type == 0) workflow = 2;
  else if (d.ana1_type == 2) workflow = 3;
  else if (d.ana2_type == 0) workflow = 0;
  else workflow = 1;
  j.open("decision");
  j.u64("workflow", workflow);
  if (workflow <= 0) {
    j.boolean("estimated_symbolic", d.use_est_symbolic);
    j.u64("symbolic_overflow_rows", h->sym_overflow_rows);
    j.u64("stored_row_bitmap_bytes", h->bitmap_bytes);
  }
  if (workflow <= 1) {
    j.boolean("hash_largest", d.hash_largest);
    j.boolean("dense_bitmap_query", d.dense_query_bitmap);
  }
  j.boolean("learned_estimates", h->learned_run);
  j.boolean("estimation_memory_fallback", h->est_memory_fallback);
  if (h->learned_consulted) {
    j.u64("learned_total_capacity", h->learned_total);
    j.num("learned_products_per_capacity",
          (double)h->total_products / (double)std::max<uint64_t>(h->learned_total, 1));
  }
  j.close();
  if (workflow <= 1) {
    j.open("sampling");
    j.u64("log2_precision", d.log_p);
    j.u64("num_samples", d.nsamples);
    j.num("compaction_1", d.comp1);
    j.num("compaction_2", d.comp2);
    if (workflow <= 0) {
      j.num("compaction_2_var", d.comp_var);
      j.num("safe_estimated_compaction", d.safe_comp);
    }
    j.close();
  }
  if (h->have_bin1) j.arr("binning_1", h->bin1, kBinSlots + 1);
  if (h->have_bin2) j.arr("binning_2", h->bin2, kBinSlots + 1);
  if (workflow == 1) {
    j.open("estimation");
    j.u64("overflow_rows", h->overflow_rows);
    j.u64("log2_precision", d.log_p);
    j.u64("total_est", h->total_est);
    j.u64("spare_needed", h->est_spare_need);
    j.boolean("spare_separate", h->est_separate_spare);
    j.num("expansion", h->cfg.hll_numeric_malloc_expansion_coe);
    j.close();
  }
  if (workflow == 2) {
    j.open("ultra_sparse");
    j.u64("kernel_size", d.spark_kernel_size);
    j.u64("outlier_size", d.spark_outlier_kernel_size);
    j.close();
  }
  if (h->seg.ran) {  // SpGEMM.cuh:1853-1886
    static const char *kClass[4] = {"SPARSE_UNIFORM", "DENSE", "POWER_LAW", "BANDED"};
    const oc_spgemm::Segmentation &sg = h->seg;
    j.open("segmentation");
    j.u64("grid_h", sg.grid_h);
    j.u64("grid_w", sg.grid_w);
    j.u64("tile_rows", sg.tile_rows);
    j.u64("tile_cols", sg.tile_cols);
    j.num("density_mean", sg.density_mean);
    j.num("density_stddev", sg.density_stddev);
    j.u64("num_components", sg.num_components);
    j.u64("cc_iterations", sg.cc_iterations);
    j.num("time_density_ms", sg.t_density);
    j.num("time_threshold_ms", sg.t_threshold);
    j.num("time_cc_ms", sg.t_cc);
    j.num("time_classify_ms", sg.t_classify);
    j.list_open("segments");
    for (const oc_spgemm::SegRecord &r : sg.segments) {
      j.item_open();
      j.u64("row_start", r.row_start);
      j.u64("row_end", r.row_end);
      j.u64("row_count", r.row_end - r.row_start);
      j.u64("col_left", r.col_left);
      j.u64("col_right", r.col_right);
      j.str("class", kClass[r.cls & 3]);
      j.u64("sum_products", r.sum_products);
      j.u64("sum_a_nnz", r.sum_a_nnz);
      j.num("mean_row_len", r.mean_row_len);
      j.num("var_row_len", r.var_row_len);
      j.num("local_er", r.local_er);
      j.u64("max_b_row_len", r.max_b_row_len);
      j.u64("max_a_row_len", r.max_a_row_len);
      j.item_close();
    }
    j.list_close();
    j.close();
  }
  if (h->cfg.track_stage_times) {
    j.open("timing");
    j.num("prologue", stage(h, "prologue"));
    j.open("analysis");
    j.num("product_calc", stage(h, "analysis_product_calc"));
    j.num("reduce", stage(h, "analysis_reduce"));
    j.num("mem_cpy", stage(h, "analysis_mem_cpy"));
    j.close();
    j.open("sampling");
    j.num("var_calc_host", stage(h, "sampling_var_calc_host"));
    j.close();
    j.open("estimation");
    j.num("sampling", stage(h, "est_sample"));
    j.num("malloc", stage(h, "est_malloc"));
    j.num("hll_construct", stage(h, "est_hll_construct"));
    j.num("hll_merge", stage(h, "est_hll_merge"));
    j.num("binning", 0.0);
    j.close();
    j.open("symbolic");
    j.num("binning", stage(h, "symbolic_binning"));
    j.num("hash", stage(h, "symbolic_hash"));
    j.num("dense", stage(h, "symbolic_dense"));
    j.num("overflow", stage(h, "symbolic_overflow"));
    j.close();
    j.open("numeric");
    j.num("malloc", stage(h, "numeric_malloc"));
    j.num("binning", stage(h, "numeric_binning"));
    j.num("hash", stage(h, "numeric_hash"));
    j.num("usparse_outlier", stage(h, "numeric_usparse_outlier"));
    j.num("usparse_main", stage(h, "numeric_usparse_main"));
    j.num("dense", stage(h, "numeric_dense"));
    j.num("window", stage(h, "numeric_window"));
    j.num("overflow", stage(h, "numeric_overflow"));
    j.num("fat", stage(h, "numeric_fat"));
    j.close();
    j.open("epilogue");
    j.num("scan", stage(h, "epilogue_scan"));
    j.num("sort", 0.0);
    j.num("copy", stage(h, "epilogue_copy"));
    j.close();
    j.close();
  }
  if (!h->fams.empty()) {
    // per launch family: GPU ms and algorithmic bytes (gather model: A
    // entries + B row pointers, one B entry per product, one C entry per
    // output, row offsets)
    const double sv = (double)vsize(h->vtype);
    j.open("kernels");
    std::map<std::string, oc_spgemm::FamStat> merged;
    for (const auto &f : h->fams) {
      auto it = merged.find(f.name);
      if (it == merged.end()) merged[f.name] = f;
      else {
        it->second.rows += f.rows;
        it->second.a_nnz += f.a_nnz;
        it->second.products += f.products;
        it->second.nnz += f.nnz;
      }
    }
    for (const auto &kv : merged) {
      const auto &f = kv.second;
      double bytes;
      if (f.numeric)
        bytes = 16.0 * f.rows + (4 + sv + 8) * f.a_nnz + (4 + sv) * (double)f.products +
                (4 + sv) * (double)f.nnz;
      else
        bytes = 12.0 * f.rows + 12.0 * f.a_nnz + 4.0 * (double)f.products;
      j.open(f.name.c_str());
      j.num("ms", stage(h, f.stage.c_str()));
      j.u64("rows", f.rows);
      j.u64("a_nnz", f.a_nnz);
      j.u64("products", f.products);
      j.u64("nnz", f.nnz);
      j.num("bytes_alg", bytes);
      j.close();
    }
    j.close();
  }
  j.open("timing_flat");
  for (const auto &kv : h->stage_ms) j.num(kv.first.c_str(), kv.second);
  j.close();
  j.num("total_time", h->total_ms);
  j.num("gflops", (double)(h->total_products * 2) / h->total_ms / 1e6);
  // HBM roofline of the whole product (gather model)
  const double s = (double)vsize(h->vtype);
  const double oc_bytes = h->c_nnz >= 0xFFFFFFFFull ? 8.0 : 4.0;
  const double bytes = 4.0 * (h->a_rows + 1) + (4 + s) * h->a_nnz +
                       (4 + s) * (double)h->total_products + oc_bytes * (h->a_rows + 1) +
                       (4 + s) * (double)h->c_nnz;
  j.open("roofline");
  j.num("bytes_alg", bytes);
  j.num("achieved_gbps", bytes / (h->total_ms * 1e-3) / 1e9);
  j.num("fraction_of_8tbps", bytes / (h->total_ms * 1e-3) / 8.0e12);
  j.close();
  j.o << "}";
  return j.o.str();
}

int validate_inputs(const oc_csr *A, const oc_csr *B) {
  OC_REQUIRE(A && B, "null matrix");
  OC_REQUIRE(A->on_device && B->on_device, "A and B must be device matrices");
  OC_REQUIRE(A->off_bytes == 4 && B->off_bytes == 4,
             "A and B need 32-bit row offsets (off_bytes 4)");
  OC_REQUIRE(A->vtype == B->vtype, "A and B must have the same value type");
  OC_REQUIRE(A->vtype == OC_F64 || A->vtype == OC_F32, "bad value type");
  OC_REQUIRE(A->cols == B->rows, "Incompatible matrix dimensions! (A.cols %llu != B.rows %llu)",
             (unsigned long long)A->cols, (unsigned long long)B->rows);
  OC_REQUIRE(A->rows < (1ull << 32) && B->cols <= (1ull << 32),
             "row and column counts must fit 32-bit indices");
  return OC_OK;
}

int spgemm_impl(oc_spgemm *h, const oc_csr *dA, const oc_csr *dB, oc_csr *dC,
                hipStream_t s) {
  OC_TRY(validate_inputs(dA, dB));
  OC_REQUIRE(dC != nullptr, "null output matrix");
  OC_TRY(tables_init());
  h->a_rows = dA->rows;
  h->a_cols = dA->cols;
  h->a_nnz = dA->nnz;
  h->b_rows = dB->rows;
  h->b_cols = dB->cols;
  h->b_nnz = dB->nnz;
  h->vtype = dA->vtype;
  h->have_bin1 = h->have_bin2 = false;
  h->total_est = 0;
  h->overflow_rows = h->sym_overflow_rows = h->precise_recount = 0;
  h->learned_run = h->learned_consulted = false;
  h->est_spare_need = 0;
  h->est_separate_spare = false;
  h->est_memory_fallback = false;
  h->bitmap_bytes = 0;
  h->learned_total = 0;
  h->stage_ms.clear();
  std::memset(dC, 0, sizeof(*dC));
  Run r(h, s, dC);
  r.A = DevCsr{dA->rows, dA->cols, dA->nnz, (const uint32_t *)dA->row_offsets,
               dA->col_ids, dA->values, dA->vtype};
  r.B = DevCsr{dB->rows, dB->cols, dB->nnz, (const uint32_t *)dB->row_offsets,
               dB->col_ids, dB->values, dB->vtype};
  r.mk.mark("start");
  int rc = OC_OK;
  if (dA->rows == 0) {
    dC->rows = 0;
    dC->cols = dB->cols;
    dC->vtype = dA->vtype;
    dC->off_bytes = 8;
    dC->on_device = 1;
    dC->owned = 3;
    OC_TRY(out_alloc(&dC->row_offsets, 8));
    HIP_TRY(hipMemsetAsync(dC->row_offsets, 0, 8, s));
    OC_TRY(out_alloc((void **)&dC->col_ids, 4));
    OC_TRY(out_alloc(&dC->values, 8));
    h->total_products = 0;
    h->d = Decision();
  } else {
    rc = r.prologue_and_analysis();
    h->seg = oc_spgemm::Segmentation();
    if (rc == OC_OK && h->cfg.use_segmented_workflow && h->d.ana1_type != 2)
      rc = r.segmentation();
    if (rc == OC_OK && h->keep_plan) {
      // the sampled compaction decides the symbolic tables, as in a plain
      // run of the HyperLogLog-sampled path; the plan itself is always the
      // precise workflow
      if (h->d.ana1_type == 1) rc = r.sampling();
      h->d.ana1_type = 1;  // reported as the precise workflow
      h->d.ana2_type = 0;
      if (rc == OC_OK) rc = r.run_precise();
    } else if (rc == OC_OK) {
      const bool learned = h->model != nullptr && h->d.ana1_type == 1 && h->force_workflow != 0;
      bool use_est = false;
      if (learned) rc = r.learned_capacities(&use_est);
      if (rc == OC_OK && learned && (use_est || h->force_workflow == 1)) {
        h->learned_run = true;
        h->d.ana2_type = 1;  // predicted capacities
      } else if (rc == OC_OK && learned) {
        // the predictor picked exact counts: the precise workflow runs with
        // full-size symbolic tables, without the HyperLogLog sketches and
        // sampling, which would only choose the workflow again and maybe
        // shrink those tables
        h->d.ana2_type = 0;
      } else if (rc == OC_OK && h->d.ana1_type == 1) {
        rc = r.sampling();
      }
      if (rc == OC_OK && h->force_workflow >= 0 && h->d.ana1_type == 1)
        h->d.ana2_type = h->force_workflow;
      if (rc == OC_OK) {
        const Decision &d = h->d;
        if (d.ana1_type == 0) rc = r.run_usparse();
        else if (d.ana1_type == 2) rc = r.run_fat();
        else if (d.ana2_type == 0) rc = r.run_precise();
        else rc = r.run_est();
      }
    }
  }
  r.mk.mark("end");
  h->fams.clear();
  if (rc == OC_OK && !r.fams.empty() && dC->row_offsets) {
    unsigned long long *acc = nullptr;
    const size_t nf = r.fams.size();
    if (hipMallocAsync((void **)&acc, nf * 32, s) == hipSuccess) {
      (void)hipMemsetAsync(acc, 0, nf * 32, s);
      for (size_t i = 0; i < nf && rc == OC_OK; ++i)
        rc = launch_row_sums(r.fams[i].rows, r.fams[i].n, r.A.rp, r.ri.products,
                             (const uint64_t *)dC->row_offsets, nullptr, acc + 4 * i, s);
      std::vector<unsigned long long> hs(4 * nf);
      if (rc == OC_OK &&
          hipMemcpyAsync(hs.data(), acc, nf * 32, hipMemcpyDeviceToHost, s) == hipSuccess &&
          hipStreamSynchronize(s) == hipSuccess) {
        for (size_t i = 0; i < nf; ++i)
          h->fams.push_back(oc_spgemm::FamStat{r.fams[i].name, r.fams[i].stage,
                                               r.fams[i].numeric, hs[4 * i], hs[4 * i + 1],
                                               hs[4 * i + 2], hs[4 * i + 3]});
      }
      (void)hipFreeAsync(acc, s);
    }
  }
  hipError_t e = hipStreamSynchronize(s);
  if (rc == OC_OK && e != hipSuccess) {
    set_error("stream error: %s", hipGetErrorString(e));
    rc = OC_ERR_HIP;
  }
  r.mk.resolve(h->stage_ms, h->total_ms);
  if (rc == OC_OK && h->keep_plan && h->plan.valid) {
    // the plan keeps this run's memory for replays (freed by drop_plan);
    // the next run starts a new chunk store
    for (const ChunkStore::Chunk &ck : h->mem.chunks) h->plan.mem.push_back(ck.base);
    h->mem.chunks.clear();
    r.ar.recs.clear();
  } else {
    if (h->keep_plan) h->plan.valid = false;
    size_t fr = 0, tot = 0;
    if (e == hipSuccess && hipMemGetInfo(&fr, &tot) == hipSuccess && h->mem.total() > tot / 4)
      h->mem.trim();
    (void)hipGetLastError();
    r.ar.recs.clear();
  }
  r.ar.release();
  if (rc != OC_OK) {
    oc_csr_free(dC);
    return rc;
  }
  h->c_rows = dC->rows;
  h->c_cols = dC->cols;
  h->c_nnz = dC->nnz;
  return OC_OK;
}

// Numeric phase of a kept plan, issued on stream s (no allocation, no host
// synchronisation, no memset nodes: safe to capture into a graph).
int replay_impl(oc_spgemm *h, hipStream_t s) {
  oc_spgemm::Plan &p = h->plan;
  OC_REQUIRE(p.valid, "no numeric plan: run oc_spgemm_plan first");
  Run r(h, s, nullptr);
  r.replay = true;
  r.mk.enabled = false;
  r.A = p.A;
  r.B = p.B;
  r.ri = p.ri;
  r.tile_ts = p.tile_ts;
  r.tile_tsw = p.tile_tsw;
  r.block_base = p.block_base;
  r.row_bm = p.row_bm;
  r.row_bmoff = p.row_bmoff;
  r.tsw_log = p.tsw_log;
  r.tsw_per = p.tsw_per;
  r.tile_logw = p.tile_logw;
  r.tile_t1 = p.tile_t1;
  r.tiles_ok = p.tiles_ok;
  r.win_scratch = p.win_scratch;
  r.pc_kp = p.pc_kp;
  r.pc_extra = p.pc_extra;
  r.pc_cnt = p.pc_cnt;
  r.pc_next = p.pc_next;
  r.pc_cap = p.pc_cap;
  r.band.n = p.band_n;
  r.band.nbands = p.band_nbands;
  r.band.logw = p.band_logw;
  r.band.rows = p.band_rows;
  r.band.ts = p.band_ts;
  r.band.counter = p.band_counter;
  r.band.pref = p.band_pref;
  const uint32_t keep_max = h->max_a_nnz;
  h->max_a_nnz = p.max_a_nnz;
  r.claim_ctr = p.err + 32;  // the plan's claim counter (no allocation here)
  int rc = r.numeric_bins(p.brows, p.hoff, p.o, true, p.boff2);
  if (rc == OC_OK) rc = r.band_numeric(p.o);
  h->max_a_nnz = keep_max;
  if (rc == OC_OK && !r.ar.recs.empty()) {
    set_error("internal: a replay allocated memory");
    rc = OC_ERR_INVALID;
  }
  return rc;
}

// Check bits of a whole-path replay (oc_spgemm_full_status).
constexpr uint32_t kFullCapacity = 1;   // nnz(C) beyond the planned capacity
constexpr uint32_t kFullLongA = 2;      // an A row longer than the planned longest
constexpr uint32_t kFullSymOvf = 4;     // (unused: symbolic overflow rows are recounted)
constexpr uint32_t kFullNumOvf = 8;     // a numeric row needed a recount

// The whole precise path of a kept plan on stream s, from the analysis to
// the numeric kernels, for the current contents of the planned A and B
// (same dimensions and entry counts; structure and values may change).  Bin
// sizes stay on the device: every bin kernel runs persistent over its
// device-side row range, so nothing is read back, allocated or set with a
// memset node and the sequence can be captured into one graph.  C keeps the
// planned capacity; when a product needs more (or a row outgrows a planned
// buffer) the check bits are set and the numeric kernels write nothing.
int full_replay_impl(oc_spgemm *h, hipStream_t s) {
  oc_spgemm::Plan &p = h->plan;
  OC_REQUIRE(p.valid, "no numeric plan: run oc_spgemm_plan first");
  OC_REQUIRE(p.band_n == 0 && h->cfg.band_log2_cols == 0,
             "the whole-path replay does not support band-major long rows");
  OC_REQUIRE(p.ana_tmp && p.bin_tmp && p.err && p.boff1 && p.boff2,
             "internal: the plan holds no whole-path state");
  Run r(h, s, nullptr);
  r.replay = true;
  r.mk.enabled = false;
  r.A = p.A;
  r.B = p.B;
  r.ri = p.ri;
  r.scal = p.scal;
  r.scan_tmp = p.scan_tmp;
  r.tile_ts = p.tile_ts;
  r.tile_tsw = p.tile_tsw;
  r.block_base = p.block_base;
  r.row_bm = nullptr;  // the stored bitmaps belong to the planned structure
  r.row_bmoff = nullptr;
  r.tsw_log = p.tsw_log;
  r.tsw_per = p.tsw_per;
  r.tile_logw = p.tile_logw;
  r.tile_t1 = p.tile_t1;
  r.tiles_ok = p.tiles_ok;
  r.win_scratch = p.win_scratch;
  const uint64_t m = p.A.rows;
  uint32_t *err = p.err;
  // err[0]: check bits; err[8..9]: bounds of the symbolic overflow list;
  // err[32]: the row-claim counter of the persistent bin kernels
  uint32_t *claim = err + 32;
  const AccArgs a = r.acc();
  const Tables &t = tables();
  OC_TRY(launch_fill_u32(err, 1, 0u, s));
  OC_TRY(launch_fill_u32(p.ovf, 8, 0u, s));
  // analysis (fixed grid) and the longest A row against the planned one
  OC_TRY(launch_analysis(p.A, p.B, p.ri, p.scal, p.scal + 1, p.ana_tmp, s));
  OC_TRY(launch_check_gt(p.scal, 0, p.max_a_nnz, kFullLongA, err, s));
  // tile-start tables of the long-row kernels follow B's structure
  if (p.tiles_ok > 0) {
    OC_TRY(launch_tile_starts(p.B, p.tile_logw, p.tile_t1, p.tile_ts, s, p.block_base));
    if (p.tile_tsw)
      OC_TRY(launch_tile_starts(p.B, p.tsw_log, p.tsw_per, p.tile_tsw, s, p.block_base));
  }
  // symbolic: binning, then every bin persistent over its device range
  OC_TRY(launch_binning(p.sym_bp, m, p.bor1, p.brows1, p.boff1, p.bin_tmp, s));
  uint32_t *ovf_rows = p.ovf + 8;
  for (int b = 0; b <= 5; ++b) {
    const uint32_t S = t.hash_sym[b];
    OC_TRY(launch_sym_hash(b, a, p.brows1, 0, p.est_symbolic ? S - S / 8 : S, p.nnz, ovf_rows,
                           p.ovf, s, RowSel{p.boff1 + b, nullptr, claim}));
  }
  OC_TRY(launch_sym_window(a, p.brows1, 0, p.max_a_nnz, p.nnz, p.win_scratch, s,
                           RowSel{p.boff1 + 6, err, claim}));
  for (int b = 8; b <= 13; ++b)
    OC_TRY(launch_sym_dense(b - 7, a, p.brows1, 0, p.nnz, s, nullptr, nullptr,
                            RowSel{p.boff1 + b, nullptr, claim}));
  // rows whose table filled up are recounted by the windowed kernel over
  // the device-side overflow list [0, ovf[0])
  uint32_t *ovf_bounds = err + 8;
  OC_TRY(launch_count_bounds(p.ovf, ovf_bounds, s));
  OC_TRY(launch_sym_window(a, ovf_rows, 0, p.max_a_nnz, p.nnz, p.win_scratch, s,
                           RowSel{ovf_bounds, err, claim}));
  // C's offsets, and its capacity against the new nnz
  OC_TRY(launch_scan_u32(p.nnz, (uint64_t *)p.o.off, m, p.scan_tmp, s));
  OC_TRY(launch_check_gt(p.o.off + m, 1, p.cap_nnz, kFullCapacity, err, s));
  // numeric: binning on the exact counts, then every bin persistent
  OC_TRY(launch_binning(p.num_bp, m, p.bor2, const_cast<uint32_t *>(p.brows), p.boff2, p.bin_tmp, s));
  for (int b = 1; b <= 6; ++b)
    OC_TRY(launch_num_hash(b, a, p.brows, 0, p.o, s, RowSel{p.boff2 + b, err, claim}));
  for (int b = 7; b <= 12; ++b)
    OC_TRY(launch_num_dense(b - 7, a, p.brows, 0, p.o, s, RowSel{p.boff2 + b, err, claim}));
  const uint32_t keep_max = h->max_a_nnz;
  h->max_a_nnz = p.max_a_nnz;
  int rc = r.big_rows(p.brows, 0, p.o, RowSel{p.boff2 + 13, err});
  h->max_a_nnz = keep_max;
  OC_TRY(rc);
  OC_TRY(launch_check_gt(p.ovf + 1, 0, 0, kFullNumOvf, err, s));
  if (!r.ar.recs.empty()) {
    set_error("internal: a whole-path replay allocated memory");
    return OC_ERR_INVALID;
  }
  return OC_OK;
}

}  // namespace

namespace oc {
// Used by the batched layer (batch.hip) to record the numeric phase into
// its own graph after refreshing the concatenated values.
int spgemm_replay_async(oc_spgemm *h, hipStream_t s) { return replay_impl(h, s); }
int spgemm_full_replay_async(oc_spgemm *h, hipStream_t s) { return full_replay_impl(h, s); }
int spgemm_full_status(oc_spgemm *h, uint32_t *status) {
  OC_REQUIRE(h && h->plan.valid && h->plan.err, "no whole-path plan");
  HIP_TRY(hipMemcpy(status, h->plan.err, 4, hipMemcpyDeviceToHost));
  return OC_OK;
}
void spgemm_set_blocks(oc_spgemm *h, const uint32_t *b_block_base, uint64_t span,
                       const void *b_row_offsets) {
  if (!h) return;
  h->drop_plan();  // a kept plan's tables were built for the previous B
  h->b_block_base = b_block_base;
  h->b_block_span = b_block_base ? span : 0;
  h->b_block_rp = b_block_base ? b_row_offsets : nullptr;
}
void spgemm_drop_plan(oc_spgemm *h) {
  if (h) h->drop_plan();
}
}  // namespace oc

// ================================================================== C ABI
extern "C" oc_spgemm *oc_spgemm_create(const oc_config *cfg) {
  oc_spgemm *h = new (std::nothrow) oc_spgemm();
  if (!h) {
    set_error("out of host memory");
    return nullptr;
  }
  if (cfg) h->cfg = *cfg;
  else oc_config_default(&h->cfg);
  if (oc_config_check(&h->cfg) != OC_OK) {
    delete h;  // the check's message stays set
    return nullptr;
  }
  for (int b = 0; b < 7; ++b)
    if (h->cfg.bin_concurrency[b] || h->cfg.bin_block[b]) h->custom_shapes = true;
  if (h->custom_shapes &&
      tables_for(h->cfg.bin_concurrency, h->cfg.bin_block, &h->shapes) != OC_OK) {
    delete h;  // the message names the bin
    return nullptr;
  }
  if (h->cfg.use_learned_binning && h->cfg.learned_model_file[0]) {
    h->owned_model = oc_mlp_create(0);
    if (!h->owned_model || oc_mlp_load(h->owned_model, h->cfg.learned_model_file) != OC_OK) {
      delete h;  // the error message of the failing call stays set
      return nullptr;
    }
    h->model = h->owned_model;
    const double q = h->cfg.learned_quantile;
    h->model_q = (q > 0.0 && q <= 1.0) ? q : 0.99;
  }
  // Keep freed pool memory for reuse across runs (SpGEMM.cuh:219-223).
  int dev = 0;
  hipMemPool_t pool;
  if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
    uint64_t thr = UINT64_MAX;
    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
  }
  (void)hipGetLastError();
  return h;
}

extern "C" void oc_spgemm_destroy(oc_spgemm *h) { delete h; }

extern "C" int oc_spgemm_run(oc_spgemm *h, const oc_csr *dA, const oc_csr *dB,
                             oc_csr *dC, void *stream) {
  OC_REQUIRE(h != nullptr, "null handle");
  return spgemm_impl(h, dA, dB, dC, (hipStream_t)stream);
}

extern "C" int oc_spgemm_stats_json(const oc_spgemm *h, char *buf, size_t cap,
                                    size_t *needed) {
  OC_REQUIRE(h != nullptr, "null handle");
  const std::string js = stats_json(h);
  if (needed) *needed = js.size() + 1;
  if (buf && cap) {
    size_t n = std::min(cap - 1, js.size());
    std::memcpy(buf, js.data(), n);
    buf[n] = '\0';
  }
  return OC_OK;
}

extern "C" int oc_spgemm_set_workflow(oc_spgemm *h, int workflow) {
  OC_REQUIRE(h != nullptr, "null handle");
  OC_REQUIRE(workflow >= -1 && workflow <= 1, "workflow must be -1, 0 or 1");
  h->force_workflow = workflow;
  return OC_OK;
}

extern "C" int oc_spgemm_plan(oc_spgemm *h, const oc_csr *dA, const oc_csr *dB, oc_csr *dC,
                              void *stream) {
  OC_REQUIRE(h != nullptr, "null handle");
  h->drop_plan();
  h->keep_plan = true;
  int rc = spgemm_impl(h, dA, dB, dC, (hipStream_t)stream);
  h->keep_plan = false;
  if (rc == OC_OK && !h->plan.valid) {
    h->drop_plan();
    set_error("oc_spgemm_plan: the product needed an overflow recount; no plan was kept");
    return OC_ERR_INVALID;
  }
  return rc;
}

extern "C" int oc_spgemm_replay(oc_spgemm *h, void *stream) {
  OC_REQUIRE(h != nullptr, "null handle");
  hipStream_t s = (hipStream_t)stream;
  OC_TRY(replay_impl(h, s));
  HIP_TRY(hipStreamSynchronize(s));
  return OC_OK;
}

extern "C" int oc_spgemm_capture(oc_spgemm *h) {
  OC_REQUIRE(h != nullptr, "null handle");
  OC_REQUIRE(h->plan.valid, "no numeric plan: run oc_spgemm_plan first");
  oc_spgemm::Plan &p = h->plan;
  if (p.exec) (void)hipGraphExecDestroy(p.exec);
  if (p.graph) (void)hipGraphDestroy(p.graph);
  p.exec = nullptr;
  p.graph = nullptr;
  if (!h->own_stream) HIP_TRY(hipStreamCreate(&h->own_stream));
  HIP_TRY(hipStreamBeginCapture(h->own_stream, hipStreamCaptureModeRelaxed));
  const int rc = replay_impl(h, h->own_stream);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(h->own_stream, &g);
  if (rc != OC_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  HIP_TRY(e);
  p.graph = g;
  HIP_TRY(hipGraphInstantiate(&p.exec, g, nullptr, nullptr, 0));
  return OC_OK;
}

extern "C" int oc_spgemm_capture_full(oc_spgemm *h) {
  OC_REQUIRE(h != nullptr, "null handle");
  OC_REQUIRE(h->plan.valid, "no numeric plan: run oc_spgemm_plan first");
  oc_spgemm::Plan &p = h->plan;
  if (p.exec_full) (void)hipGraphExecDestroy(p.exec_full);
  if (p.graph_full) (void)hipGraphDestroy(p.graph_full);
  p.exec_full = nullptr;
  p.graph_full = nullptr;
  if (!h->own_stream) HIP_TRY(hipStreamCreate(&h->own_stream));
  HIP_TRY(hipStreamBeginCapture(h->own_stream, hipStreamCaptureModeRelaxed));
  const int rc = full_replay_impl(h, h->own_stream);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(h->own_stream, &g);
  if (rc != OC_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  HIP_TRY(e);
  p.graph_full = g;
  HIP_TRY(hipGraphInstantiate(&p.exec_full, g, nullptr, nullptr, 0));
  return OC_OK;
}

extern "C" int oc_spgemm_graph_launch_full(oc_spgemm *h, void *stream) {
  OC_REQUIRE(h != nullptr, "null handle");
  OC_REQUIRE(h->plan.valid && h->plan.exec_full,
             "no whole-path graph: run oc_spgemm_capture_full first");
  if (stream) {
    HIP_TRY(hipGraphLaunch(h->plan.exec_full, (hipStream_t)stream));
  } else {
    HIP_TRY(hipGraphLaunch(h->plan.exec_full, h->own_stream));
    HIP_TRY(hipStreamSynchronize(h->own_stream));
  }
  return OC_OK;
}

extern "C" int oc_spgemm_full_replay(oc_spgemm *h, void *stream) {
  OC_REQUIRE(h != nullptr, "null handle");
  hipStream_t s = (hipStream_t)stream;
  OC_TRY(full_replay_impl(h, s));
  HIP_TRY(hipStreamSynchronize(s));
  return OC_OK;
}

extern "C" int oc_spgemm_full_status(oc_spgemm *h, uint32_t *status) {
  OC_REQUIRE(status != nullptr, "null status");
  return spgemm_full_status(h, status);
}

extern "C" int oc_spgemm_graph_launch(oc_spgemm *h, void *stream) {
  OC_REQUIRE(h != nullptr, "null handle");
  OC_REQUIRE(h->plan.valid && h->plan.exec, "no captured graph: run oc_spgemm_capture first");
  if (stream) {
    HIP_TRY(hipGraphLaunch(h->plan.exec, (hipStream_t)stream));
  } else {
    HIP_TRY(hipGraphLaunch(h->plan.exec, h->own_stream));
    HIP_TRY(hipStreamSynchronize(h->own_stream));
  }
  return OC_OK;
}

extern "C" int oc_spgemm_set_model(oc_spgemm *h, const oc_mlp *m, double quantile) {
  OC_REQUIRE(h != nullptr, "null handle");
  OC_REQUIRE(quantile > 0.0 && quantile <= 1.0, "quantile must be in (0, 1]");
  h->model = m ? m : h->owned_model;
  h->model_q = quantile;
  return OC_OK;
}

extern "C" int oc_spgemm_set_estimates(oc_spgemm *h, const uint32_t *d_est) {
  OC_REQUIRE(h != nullptr, "null handle");
  h->est_override = d_est;
  return OC_OK;
}

extern "C" double oc_spgemm_timer_ms(const oc_spgemm *h, const char *name) {
  if (!h || !name) return -1.0;
  if (!std::strcmp(name, "total")) return h->total_ms;
  auto it = h->stage_ms.find(name);
  return it == h->stage_ms.end() ? -1.0 : it->second;
}
