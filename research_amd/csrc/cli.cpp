// spgemm -- command-line driver with the reference's argv grammar
// (src/main.cu:11-84): spgemm A [B] [config.json].  With three arguments a
// ".json" suffix selects a config, anything else a second matrix.  Written
// against include/ocean.hpp in the shape of the reference's driver: load
// the matrices, read the config as JSON, SpGEMM(config).run(A, B) with its
// warmup/bench loop, then the optional comparison against
// ground_truth_file and the stats file.  Exit status: 0 on success, 1 on
// a usage, load or dimension error, 2 when the comparison fails (where the
// reference's uncaught exception aborts).
#include <fstream>
#include <iostream>
#include <memory>
#include <string>

#include "../../include/ocean.hpp"

using json = somespgemm::json;

static bool ends_with(const std::string &s, const std::string &suf) {
  return s.size() >= suf.size() && s.compare(s.size() - suf.size(), suf.size(), suf) == 0;
}

int main(int argc, char **argv) {
  if (argc < 2 || argc > 4) {
    std::cerr << "Usage: " << argv[0] << " <matrix_file> [matrix_b_file] [config_file]\n";
    std::cerr << "Matrices are .csr files, and config should end with .json.\n";
    return 1;
  }
  std::string a_file = argv[1], b_file = argv[1], cfg_file;
  if (argc == 3) {
    const std::string arg2 = argv[2];
    if (ends_with(arg2, ".json")) cfg_file = arg2;
    else b_file = arg2;
  } else if (argc == 4) {
    b_file = argv[2];
    cfg_file = argv[3];
  }
  try {
    auto matrix_a = somespgemm::loadMatrix(a_file);
    auto matrix_b = somespgemm::loadMatrix(b_file);
    if (matrix_a->cols != matrix_b->rows) {
      std::cerr << "Incompatible matrix dimensions!" << std::endl;
      return 1;
    }
    if (matrix_a->m.vtype != matrix_b->m.vtype) {
      std::cerr << "A and B have different value types" << std::endl;
      return 1;
    }
    auto config = somespgemm::Config();
    if (!cfg_file.empty()) {
      std::ifstream config_ifs(cfg_file);
      if (!config_ifs.is_open()) {
        std::cerr << "could not open config " << cfg_file << std::endl;
        return 1;
      }
      json j;
      config_ifs >> j;
      from_json(j, config);
    }
    somespgemm::SpGEMM spgemm{config};
    auto d_A = std::make_shared<somespgemm::cuCSR>(*matrix_a);
    auto d_B = std::make_shared<somespgemm::cuCSR>(*matrix_b);
    auto result = spgemm.run(d_A, d_B);
    auto result_host = std::make_shared<somespgemm::CSR>(*result);
    if (config.check_correctness) {
      auto ground_truth = somespgemm::loadMatrix(config.ground_truth_file);
      try {
        somespgemm::compare(*result_host, *ground_truth, config.compare_relative_error,
                            config.compare_autofail_relative_error);
      } catch (const somespgemm::compare_error &e) {
        std::cerr << e.what() << std::endl;
        return 2;
      }
    }
    if (config.output_stats) {
      std::cout << "Writing to " << config.stats_output_file << std::endl;
      json stats = spgemm.get_stats();
      std::ofstream stats_ofs(config.stats_output_file);
      stats_ofs << stats.dump(4);
    }
  } catch (const std::exception &e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
