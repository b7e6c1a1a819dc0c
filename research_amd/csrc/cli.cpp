// spgemm -- command-line driver, same argv grammar as the reference
// (src/main.cu:11-84): spgemm A [B] [config.json].  With three arguments a
// ".json" suffix selects a config, anything else a second matrix.  Runs
// warmup_iters + bench_iters products (kernels/SpGEMM.cuh:226-316), prints
// per-iteration and average times, the stage report when track_stage_time
// is set, compares against ground_truth_file when check_correctness is set
// and writes the stats JSON when output_stats is set.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/ocean.hpp"

static bool ends_with(const std::string &s, const std::string &suf) {
  return s.size() >= suf.size() && s.compare(s.size() - suf.size(), suf.size(), suf) == 0;
}

static const char *kBanner =
    "================================================================";

int main(int argc, char **argv) {
  if (argc < 2 || argc > 4) {
    std::cerr << "Usage: " << argv[0] << " <matrix_file> [matrix_b_file] [config_file]\n";
    std::cerr << "Matrices are .csr files, and config should end with .json.\n";
    return 1;
  }
  std::string a_file = argv[1], b_file = argv[1], cfg_file;
  if (argc == 3) {
    std::string arg2 = argv[2];
    if (ends_with(arg2, ".mtxd_.hicsr")) b_file = arg2;
    else if (ends_with(arg2, ".json")) cfg_file = arg2;
    else b_file = arg2;
  } else if (argc == 4) {
    b_file = argv[2];
    cfg_file = argv[3];
  }
  try {
    auto A = somespgemm::loadMatrix(a_file);
    auto B = b_file == a_file ? A : somespgemm::loadMatrix(b_file);
    if (A->cols() != B->rows()) {
      std::cerr << "Incompatible matrix dimensions!" << std::endl;
      return 1;
    }
    if (A->m.vtype != B->m.vtype) {
      std::cerr << "A and B have different value types" << std::endl;
      return 1;
    }
    somespgemm::Config cfg = somespgemm::default_config();
    if (!cfg_file.empty()) {
      std::ifstream in(cfg_file);
      if (!in.is_open()) {
        std::cerr << "could not open config " << cfg_file << std::endl;
        return 1;
      }
      std::stringstream ss;
      ss << in.rdbuf();
      somespgemm::from_json(ss.str(), cfg);
    }
    somespgemm::SpGEMM spgemm(cfg);
    somespgemm::cuCSR dA(*A), dB(*B);
    std::shared_ptr<somespgemm::cuCSR> C;
    double time = 0.0;
    std::map<std::string, double> stages;
    const int total = cfg.warmup_iters + cfg.bench_iters;
    std::string stats;
    for (int i = 0; i < total; ++i) {
      C.reset();
      C = spgemm.run(dA, dB);
      const double ms = spgemm.timer_ms("total");
      std::cout << "Iteration " << i << " done. Time: " << ms << " ms\n";
      if (i >= cfg.warmup_iters) {
        time += ms;
        for (const char *k : {"prologue", "analysis_product_calc", "analysis_reduce",
                              "analysis_mem_cpy", "est_sample", "est_malloc",
                              "est_hll_construct", "sampling_var_calc_host", "est_hll_merge",
                              "symbolic_binning", "symbolic_hash", "symbolic_dense",
                              "symbolic_overflow", "numeric_malloc", "numeric_binning",
                              "numeric_hash", "numeric_usparse_outlier",
                              "numeric_usparse_main", "numeric_dense", "numeric_window",
                              "numeric_overflow", "epilogue_scan", "epilogue_copy"}) {
          double v = spgemm.timer_ms(k);
          if (v > 0) stages[k] += v;
        }
      }
      if (cfg.print_memory) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) == hipSuccess)
          std::cout << kBanner << "\nGPU memory usage: used = " << (tot - fr) / 1048576.0
                    << " MB, free = " << fr / 1048576.0 << " MB, total = " << tot / 1048576.0
                    << " MB\n" << kBanner << "\n";
      }
    }
    const int iters = cfg.bench_iters > 0 ? cfg.bench_iters : 1;
    std::cout << "Average SpGEMM time over " << cfg.bench_iters << " iters: " << time / iters
              << " ms\n";
    stats = spgemm.get_stats();
    if (cfg.track_stage_times) {
      std::cout << kBanner << "\nTiming Report (ms):\n";
      for (auto &kv : stages) std::cout << kv.first << ": " << kv.second / iters << "\n";
      std::cout << kBanner << "\n";
    }
    somespgemm::CSR Ch;
    somespgemm::convert(Ch, *C);
    if (cfg.check_correctness) {
      auto gt = somespgemm::loadMatrix(cfg.ground_truth_file);
      std::cout << kBanner << "\nComparing CSR matrices" << std::endl;
      oc_compare_result r;
      int rc = oc_csr_compare(&Ch.m, &gt->m, cfg.compare_relative_error,
                              cfg.compare_autofail_relative_error, &r);
      if (rc != OC_OK) {
        std::cerr << oc_last_error() << std::endl;
        return 2;
      }
      std::cout << r.num_errors << " / " << Ch.nnz() << " elements have relative error greater than "
                << cfg.compare_relative_error << " (max " << r.max_rel_err << ")\n"
                << "Comparison completed\n" << kBanner << std::endl;
    }
    if (cfg.output_stats) {
      std::cout << "Writing to " << cfg.stats_output_file << std::endl;
      std::ofstream out(cfg.stats_output_file);
      out << stats << "\n";
    }
  } catch (const std::exception &e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
