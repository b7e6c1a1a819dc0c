// ocean.hpp -- C++ surface of the MI355X Ocean SpGEMM over the C ABI
// (ocean_capi.h).  Mirrors the reference's C++ API so that its callers can
// switch: namespace somespgemm, CSR / cuCSR, loadCSR, loadMatrix, convert,
// compare, Config + from_json, SpGEMM(config).run(A, B) and get_stats()
// (include/CSR.h:13-53, include/Utils.h:17-118, kernels/SpGEMM.cuh:209-343).
// Errors from the C ABI are rethrown as std::runtime_error.
#pragma once

#include <cstring>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>

#include "ocean_capi.h"

namespace somespgemm {

inline void check(int rc) {
  if (rc != OC_OK) throw std::runtime_error(oc_last_error());
}

using Config = oc_config;

inline Config default_config() {
  Config c;
  oc_config_default(&c);
  return c;
}

// Parse the JSON text of a config file into c (src/Utils.cpp:8-37).
inline void from_json(const std::string &text, Config &c) {
  check(oc_config_from_json(&c, text.c_str()));
}

// Host matrix (struct CSR) and device matrix (struct cuCSR) own an oc_csr.
struct CSR {
  oc_csr m{};
  CSR() = default;
  CSR(const CSR &) = delete;
  CSR &operator=(const CSR &) = delete;
  ~CSR() { oc_csr_free(&m); }
  size_t rows() const { return m.rows; }
  size_t cols() const { return m.cols; }
  size_t nnz() const { return m.nnz; }
};

struct cuCSR {
  oc_csr m{};
  cuCSR() = default;
  explicit cuCSR(const CSR &src) { check(oc_csr_upload(&src.m, &m, nullptr)); }
  cuCSR(const cuCSR &) = delete;
  cuCSR &operator=(const cuCSR &) = delete;
  ~cuCSR() { oc_csr_free(&m); }
  void reset() { oc_csr_free(&m); }
};

// convert(CSR&, const cuCSR&): device -> host (src/CSR.cpp:243-250)
inline void convert(CSR &dst, const cuCSR &src) {
  oc_csr_free(&dst.m);
  check(oc_csr_download(&src.m, &dst.m, nullptr));
}

inline std::shared_ptr<CSR> loadCSR(const char *file, int vtype = -1) {
  auto r = std::make_shared<CSR>();
  check(oc_csr_load(file, vtype, &r->m));
  return r;
}

// loadMatrix (src/Utils.cpp:39-57): prints progress, throws on failure.
inline std::shared_ptr<CSR> loadMatrix(const std::string &path, int vtype = -1) {
  std::cout << "trying to load csr file \"" << path << "\"\n";
  auto m = loadCSR(path.c_str(), vtype);
  std::cout << "successfully loaded: \"" << path << "\"\n";
  return m;
}

// compare (src/CSR.cpp:252-335)
inline oc_compare_result compare(const CSR &a, const CSR &b, double rel = 1e-6,
                                 double autofail = 1e-4) {
  oc_compare_result r;
  check(oc_csr_compare(&a.m, &b.m, rel, autofail, &r));
  return r;
}

class SpGEMM {
 public:
  explicit SpGEMM(const Config &cfg) : h_(oc_spgemm_create(&cfg)) {
    if (!h_) throw std::runtime_error(oc_last_error());
  }
  ~SpGEMM() { oc_spgemm_destroy(h_); }
  SpGEMM(const SpGEMM &) = delete;
  SpGEMM &operator=(const SpGEMM &) = delete;

  // One product C = A * B on the device.
  std::shared_ptr<cuCSR> run(const cuCSR &A, const cuCSR &B, void *stream = nullptr) {
    auto C = std::make_shared<cuCSR>();
    check(oc_spgemm_run(h_, &A.m, &B.m, &C->m, stream));
    return C;
  }
  std::string get_stats() const {
    size_t need = 0;
    check(oc_spgemm_stats_json(h_, nullptr, 0, &need));
    std::string s(need, '\0');
    check(oc_spgemm_stats_json(h_, &s[0], need, &need));
    s.resize(std::strlen(s.c_str()));
    return s;
  }
  double timer_ms(const char *name) const { return oc_spgemm_timer_ms(h_, name); }
  oc_spgemm *handle() { return h_; }

 private:
  oc_spgemm *h_;
};

}  // namespace somespgemm
