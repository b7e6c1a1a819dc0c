// ocean.hpp -- C++ surface of the MI355X Ocean SpGEMM over the C ABI
// (ocean_capi.h), header-only.  It keeps the shapes of the reference's C++
// API so that a caller written against it compiles unchanged:
//   namespace somespgemm
//   struct CSR / cuCSR with rows, cols, nnz fields, CSR(cuCSR&) and
//     cuCSR(const CSR&)                           (include/CSR.h:16-51)
//   loadCSR, loadMatrix, convert, compare          (src/CSR.cpp, src/Utils.cpp:39-57)
//   Config with its defaults, from_json(json, Config&)
//                                                  (include/Utils.h:17-118)
//   SpGEMM(config).run(shared_ptr<cuCSR>, shared_ptr<cuCSR>) with the
//     warmup/bench loop, json get_stats()          (kernels/SpGEMM.cuh:209-343)
// `json` is a small JSON value (parse from a stream, members, dump(indent))
// standing in for the nlohmann::json the reference's callers use; keys are
// kept sorted, as nlohmann::json keeps them.  Errors from the C ABI are
// rethrown as std::runtime_error.
#pragma once

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <istream>
#include <iterator>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "ocean_capi.h"

namespace somespgemm {

inline void check(int rc) {
  if (rc != OC_OK) throw std::runtime_error(oc_last_error());
}

// ------------------------------------------------------------------ json
class json {
 public:
  enum class kind { null, boolean, integer, unsigned_integer, number, string, array, object };

  json() = default;
  json(std::nullptr_t) {}
  json(bool b) : k_(kind::boolean), b_(b) {}
  json(int v) : k_(kind::integer), i_(v) {}
  json(long v) : k_(kind::integer), i_(v) {}
  json(long long v) : k_(kind::integer), i_(v) {}
  json(unsigned v) : k_(kind::unsigned_integer), u_(v) {}
  json(unsigned long v) : k_(kind::unsigned_integer), u_(v) {}
  json(unsigned long long v) : k_(kind::unsigned_integer), u_(v) {}
  json(double v) : k_(kind::number), d_(v) {}
  json(const char *s) : k_(kind::string), s_(s) {}
  json(std::string s) : k_(kind::string), s_(std::move(s)) {}

  static json parse(const std::string &text) {
    const char *p = text.c_str();
    json j = parse_value(p);
    skip_ws(p);
    if (*p) throw std::runtime_error("json: trailing characters");
    return j;
  }

  kind type() const { return k_; }
  bool is_null() const { return k_ == kind::null; }
  bool is_object() const { return k_ == kind::object; }
  bool is_array() const { return k_ == kind::array; }
  bool is_string() const { return k_ == kind::string; }
  bool is_number() const {
    return k_ == kind::integer || k_ == kind::unsigned_integer || k_ == kind::number;
  }
  size_t size() const {
    return k_ == kind::array ? a_.size() : k_ == kind::object ? o_.size() : (k_ == kind::null ? 0 : 1);
  }

  bool contains(const std::string &key) const { return k_ == kind::object && find(key) != nullptr; }
  // Object member, created (as null) when missing; a null value becomes an object.
  json &operator[](const std::string &key) {
    if (k_ == kind::null) k_ = kind::object;
    if (k_ != kind::object) throw std::runtime_error("json: not an object");
    auto it = std::lower_bound(o_.begin(), o_.end(), key,
                               [](const std::pair<std::string, json> &e, const std::string &k) {
                                 return e.first < k;
                               });
    if (it == o_.end() || it->first != key) it = o_.insert(it, {key, json()});
    return it->second;
  }
  json &operator[](const char *key) { return (*this)[std::string(key)]; }
  const json &at(const std::string &key) const {
    const json *v = k_ == kind::object ? find(key) : nullptr;
    if (!v) throw std::out_of_range("json: no key \"" + key + "\"");
    return *v;
  }
  const json &operator[](const std::string &key) const { return at(key); }
  const json &operator[](const char *key) const { return at(std::string(key)); }
  json &operator[](size_t i) { return a_.at(i); }
  const json &operator[](size_t i) const { return a_.at(i); }
  void push_back(json v) {
    if (k_ == kind::null) k_ = kind::array;
    if (k_ != kind::array) throw std::runtime_error("json: not an array");
    a_.push_back(std::move(v));
  }

  template <typename T>
  T get() const {
    if constexpr (std::is_same<T, bool>::value) {
      if (k_ != kind::boolean) throw std::runtime_error("json: not a boolean");
      return b_;
    } else if constexpr (std::is_same<T, std::string>::value) {
      if (k_ != kind::string) throw std::runtime_error("json: not a string");
      return s_;
    } else {
      static_assert(std::is_arithmetic<T>::value, "json::get: unsupported type");
      switch (k_) {
        case kind::integer: return static_cast<T>(i_);
        case kind::unsigned_integer: return static_cast<T>(u_);
        case kind::number: return static_cast<T>(d_);
        case kind::boolean: return static_cast<T>(b_);
        default: throw std::runtime_error("json: not a number");
      }
    }
  }

  // indent < 0: compact; otherwise one member per line, `indent` spaces per level.
  std::string dump(int indent = -1) const {
    std::string out;
    write(out, indent, 0);
    return out;
  }

 private:
  kind k_ = kind::null;
  bool b_ = false;
  long long i_ = 0;
  unsigned long long u_ = 0;
  double d_ = 0.0;
  std::string s_;
  std::vector<json> a_;
  std::vector<std::pair<std::string, json>> o_;  // sorted by key

  const json *find(const std::string &key) const {
    auto it = std::lower_bound(o_.begin(), o_.end(), key,
                               [](const std::pair<std::string, json> &e, const std::string &k) {
                                 return e.first < k;
                               });
    return it != o_.end() && it->first == key ? &it->second : nullptr;
  }

  static void skip_ws(const char *&p) {
    while (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r') ++p;
  }
  static std::string parse_string(const char *&p) {
    if (*p != '"') throw std::runtime_error("json: expected a string");
    ++p;
    std::string s;
    while (*p && *p != '"') {
      if (*p != '\\') {
        s += *p++;
        continue;
      }
      ++p;
      switch (*p) {
        case 'n': s += '\n'; break;
        case 't': s += '\t'; break;
        case 'r': s += '\r'; break;
        case 'b': s += '\b'; break;
        case 'f': s += '\f'; break;
        case 'u': {
          unsigned v = 0;
          for (int i = 1; i <= 4; ++i) {
            const char c = p[i];
            if (!std::isxdigit((unsigned char)c)) throw std::runtime_error("json: bad \\u escape");
            v = v * 16 + (unsigned)(std::isdigit((unsigned char)c) ? c - '0' : (std::tolower(c) - 'a' + 10));
          }
          p += 4;
          if (v < 0x80) {
            s += (char)v;
          } else if (v < 0x800) {
            s += (char)(0xC0 | (v >> 6));
            s += (char)(0x80 | (v & 0x3F));
          } else {
            s += (char)(0xE0 | (v >> 12));
            s += (char)(0x80 | ((v >> 6) & 0x3F));
            s += (char)(0x80 | (v & 0x3F));
          }
          break;
        }
        case '\0': throw std::runtime_error("json: unterminated string");
        default: s += *p;
      }
      ++p;
    }
    if (*p != '"') throw std::runtime_error("json: unterminated string");
    ++p;
    return s;
  }
  static json parse_value(const char *&p) {
    skip_ws(p);
    if (*p == '{') {
      ++p;
      json j;
      j.k_ = kind::object;
      skip_ws(p);
      if (*p == '}') {
        ++p;
        return j;
      }
      for (;;) {
        skip_ws(p);
        std::string key = parse_string(p);
        skip_ws(p);
        if (*p != ':') throw std::runtime_error("json: expected ':'");
        ++p;
        j[key] = parse_value(p);
        skip_ws(p);
        if (*p == ',') {
          ++p;
          continue;
        }
        if (*p != '}') throw std::runtime_error("json: expected '}'");
        ++p;
        return j;
      }
    }
    if (*p == '[') {
      ++p;
      json j;
      j.k_ = kind::array;
      skip_ws(p);
      if (*p == ']') {
        ++p;
        return j;
      }
      for (;;) {
        j.a_.push_back(parse_value(p));
        skip_ws(p);
        if (*p == ',') {
          ++p;
          continue;
        }
        if (*p != ']') throw std::runtime_error("json: expected ']'");
        ++p;
        return j;
      }
    }
    if (*p == '"') return json(parse_string(p));
    if (!std::strncmp(p, "true", 4)) { p += 4; return json(true); }
    if (!std::strncmp(p, "false", 5)) { p += 5; return json(false); }
    if (!std::strncmp(p, "null", 4)) { p += 4; return json(); }
    const char *start = p;
    bool real = false;
    if (*p == '-') ++p;
    while (std::isdigit((unsigned char)*p) || *p == '.' || *p == 'e' || *p == 'E' || *p == '+' ||
           *p == '-') {
      real |= *p == '.' || *p == 'e' || *p == 'E';
      ++p;
    }
    if (p == start) throw std::runtime_error("json: unexpected character");
    const std::string num(start, p);
    if (!real) {
      if (num[0] == '-') return json(std::strtoll(num.c_str(), nullptr, 10));
      return json(std::strtoull(num.c_str(), nullptr, 10));
    }
    return json(std::strtod(num.c_str(), nullptr));
  }

  static void write_string(std::string &out, const std::string &s) {
    out += '"';
    for (unsigned char c : s) {
      switch (c) {
        case '"': out += "\\\""; break;
        case '\\': out += "\\\\"; break;
        case '\n': out += "\\n"; break;
        case '\t': out += "\\t"; break;
        case '\r': out += "\\r"; break;
        case '\b': out += "\\b"; break;
        case '\f': out += "\\f"; break;
        default:
          if (c < 0x20) {
            char buf[8];
            std::snprintf(buf, sizeof(buf), "\\u%04x", c);
            out += buf;
          } else {
            out += (char)c;
          }
      }
    }
    out += '"';
  }
  static void write_number(std::string &out, double d) {
    if (!std::isfinite(d)) {
      out += "null";  // as nlohmann::json dumps NaN and infinities
      return;
    }
    char buf[32];
    for (int prec = 15; prec <= 17; ++prec) {  // shortest of these that reads back exactly
      std::snprintf(buf, sizeof(buf), "%.*g", prec, d);
      if (std::strtod(buf, nullptr) == d) break;
    }
    out += buf;
    if (!std::strpbrk(buf, ".eE")) out += ".0";
  }
  void write(std::string &out, int indent, int depth) const {
    const std::string nl = indent >= 0 ? "\n" : "";
    auto pad = [&](int d) { return indent > 0 ? std::string((size_t)(indent * d), ' ') : std::string(); };
    switch (k_) {
      case kind::null: out += "null"; break;
      case kind::boolean: out += b_ ? "true" : "false"; break;
      case kind::integer: out += std::to_string(i_); break;
      case kind::unsigned_integer: out += std::to_string(u_); break;
      case kind::number: write_number(out, d_); break;
      case kind::string: write_string(out, s_); break;
      case kind::array:
        if (a_.empty()) {
          out += "[]";
          break;
        }
        out += "[" + nl;
        for (size_t i = 0; i < a_.size(); ++i) {
          out += pad(depth + 1);
          a_[i].write(out, indent, depth + 1);
          if (i + 1 < a_.size()) out += ",";
          out += nl;
        }
        out += pad(depth) + "]";
        break;
      case kind::object:
        if (o_.empty()) {
          out += "{}";
          break;
        }
        out += "{" + nl;
        for (size_t i = 0; i < o_.size(); ++i) {
          out += pad(depth + 1);
          write_string(out, o_[i].first);
          out += indent >= 0 ? ": " : ":";
          o_[i].second.write(out, indent, depth + 1);
          if (i + 1 < o_.size()) out += ",";
          out += nl;
        }
        out += pad(depth) + "}";
        break;
    }
  }
};

inline std::istream &operator>>(std::istream &in, json &j) {
  std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  j = json::parse(text);
  return in;
}
inline std::ostream &operator<<(std::ostream &out, const json &j) { return out << j.dump(); }

// ---------------------------------------------------------------- config
// Config with the reference's default values (include/Utils.h:22-107).
struct Config : oc_config {
  Config() { oc_config_default(this); }
  Config(const oc_config &c) : oc_config(c) {}
};

// from_json (src/Utils.cpp:8-37): the keys of the reference's config files.
inline void from_json(const json &j, Config &c) { check(oc_config_from_json(&c, j.dump().c_str())); }
inline void from_json(const std::string &text, Config &c) {
  check(oc_config_from_json(&c, text.c_str()));
}

// ------------------------------------------------------------------- CSR
struct cuCSR;

// Host matrix (struct CSR, include/CSR.h:16-30).  rows, cols and nnz mirror
// the owned oc_csr `m`.
struct CSR {
  size_t rows = 0, cols = 0, nnz = 0;
  oc_csr m{};
  CSR() = default;
  inline CSR(const cuCSR &src);  // device -> host (include/CSR.h:27)
  CSR(const CSR &) = delete;
  CSR &operator=(const CSR &) = delete;
  ~CSR() { oc_csr_free(&m); }
  void sync() {
    rows = m.rows;
    cols = m.cols;
    nnz = m.nnz;
  }
};

// Device matrix (struct cuCSR, include/CSR.h:34-51).
struct cuCSR {
  size_t rows = 0, cols = 0, nnz = 0;
  oc_csr m{};
  cuCSR() = default;
  cuCSR(const CSR &src) {  // host -> device (include/CSR.h:44-46)
    check(oc_csr_upload(&src.m, &m, nullptr));
    sync();
  }
  cuCSR(const cuCSR &) = delete;
  cuCSR &operator=(const cuCSR &) = delete;
  virtual ~cuCSR() { oc_csr_free(&m); }
  void reset() {
    oc_csr_free(&m);
    sync();
  }
  void sync() {
    rows = m.rows;
    cols = m.cols;
    nnz = m.nnz;
  }
};

// convert (src/CSR.cpp:226-250)
inline void convert(CSR &dst, const cuCSR &src) {
  oc_csr_free(&dst.m);
  check(oc_csr_download(&src.m, &dst.m, nullptr));
  dst.sync();
}
inline void convert(cuCSR &dst, const CSR &src) {
  oc_csr_free(&dst.m);
  check(oc_csr_upload(&src.m, &dst.m, nullptr));
  dst.sync();
}
inline CSR::CSR(const cuCSR &src) { convert(*this, src); }

inline std::shared_ptr<CSR> loadCSR(const char *file, int vtype = -1) {
  auto r = std::make_shared<CSR>();
  check(oc_csr_load(file, vtype, &r->m));
  r->sync();
  return r;
}

// loadMatrix (src/Utils.cpp:39-57): prints progress; a file that cannot be
// loaded is reported and ends the process with status 1, as there.
inline std::shared_ptr<CSR> loadMatrix(const std::string &path, int vtype = -1) {
  std::cout << "trying to load csr file \"" << path << "\"\n";
  std::shared_ptr<CSR> m;
  try {
    m = loadCSR(path.c_str(), vtype);
  } catch (const std::exception &ex) {
    std::cout << "could not load csr file:\n\t" << ex.what() << "\n";
    std::exit(1);
  }
  std::cout << "successfully loaded: \"" << path << "\"\n";
  return m;
}

// ---------------------------------------------------------------- compare
static const char *const kBanner =
    "================================================================";

// A failed comparison (compare below); a std::runtime_error.
struct compare_error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// compare (src/CSR.cpp:252-335): equal dimensions, row offsets and column
// ids, and no value above autofail relative error; prints the count of
// values above rel.  Throws compare_error with the reference's messages.
inline oc_compare_result compare(const CSR &a, const CSR &b, double rel = 1e-6,
                                 double autofail = 1e-4) {
  std::cout << kBanner << "\nComparing CSR matrices" << std::endl;
  oc_compare_result r;
  const int rc = oc_csr_compare(&a.m, &b.m, rel, autofail, &r);
  if (rc == OC_ERR_MISMATCH) {
    static const char *const what[5] = {"Matrices differ", "Matrices have different dimensions",
                                        "Matrices have different row offsets",
                                        "Matrices have different column IDs",
                                        "Matrices have different data"};
    std::cout << oc_last_error() << std::endl;
    throw compare_error(what[r.kind >= 0 && r.kind <= 4 ? r.kind : 0]);
  }
  check(rc);
  std::cout << r.num_errors << " / " << a.m.nnz << " elements have relative error greater than "
            << rel << " (max " << r.max_rel_err << ")\nComparison completed\n" << kBanner
            << std::endl;
  return r;
}

// ----------------------------------------------------------------- SpGEMM
class SpGEMM {
 public:
  explicit SpGEMM(const Config &cfg) : cfg_(cfg), h_(oc_spgemm_create(&cfg)) {
    if (!h_) throw std::runtime_error(oc_last_error());
  }
  ~SpGEMM() { oc_spgemm_destroy(h_); }
  SpGEMM(const SpGEMM &) = delete;
  SpGEMM &operator=(const SpGEMM &) = delete;

  // SpGEMM::run (kernels/SpGEMM.cuh:213-343): warmup_iters + bench_iters
  // products C = A * B; the C of every iteration but the last is freed.
  // Prints each iteration's time and workflow and the average, and the
  // stage report when track_stage_time is set; get_stats() then holds the
  // last run's stats with total_time and gflops averaged over the bench
  // iterations.
  std::shared_ptr<cuCSR> run(std::shared_ptr<cuCSR> A, std::shared_ptr<cuCSR> B) {
    static const char *const kStages[] = {
        "prologue", "analysis_product_calc", "analysis_reduce", "analysis_mem_cpy",
        "est_sample", "est_malloc", "est_hll_construct", "sampling_var_calc_host",
        "est_hll_merge", "symbolic_binning", "symbolic_hash", "symbolic_dense",
        "symbolic_overflow", "numeric_malloc", "numeric_binning", "numeric_hash",
        "numeric_usparse_outlier", "numeric_usparse_main", "numeric_dense", "numeric_window",
        "numeric_overflow", "epilogue_scan", "epilogue_copy"};
    const int total = std::max(1, cfg_.warmup_iters + cfg_.bench_iters);
    const int first_timed = total - std::max(1, cfg_.bench_iters);
    std::shared_ptr<cuCSR> C;
    double time = 0.0;
    std::map<std::string, double> stages;
    for (int i = 0; i < total; ++i) {
      C.reset();
      C = run(*A, *B);
      const double ms = timer_ms("total");
      // matrix types of kernels/SpGEMM.cuh:294-309 from the stats' workflow ids
      static const int kType[4] = {1, 2, 0, 3};
      const json st = get_stats_last();
      const int wf = st.contains("decision") ? st["decision"]["workflow"].get<int>() : 0;
      std::cout << "Iteration " << i << " done. Time: " << ms << " ms\n"
                << "Matrix type: " << kType[wf & 3] << std::endl;
      if (i >= first_timed) {
        time += ms;
        for (const char *k : kStages) {
          const double v = timer_ms(k);
          if (v > 0) stages[k] += v;
        }
      }
      if (cfg_.print_memory) {
        size_t fr = 0, tot = 0;
        if (oc_device_mem_info(&fr, &tot) == OC_OK)
          std::cout << kBanner << "\nGPU memory usage: used = " << (tot - fr) / 1048576.0
                    << " MB, free = " << fr / 1048576.0 << " MB, total = " << tot / 1048576.0
                    << " MB\n" << kBanner << "\n";
      }
    }
    const int iters = total - first_timed;
    const double avg = time / iters;
    std::cout << "Average SpGEMM time over " << cfg_.bench_iters << " iters: " << avg << " ms\n";
    if (cfg_.track_stage_times) {
      std::cout << kBanner << "\nTiming Report (ms):\n";
      for (auto &kv : stages) std::cout << kv.first << ": " << kv.second / iters << "\n";
      std::cout << kBanner << "\n";
    }
    stats_ = get_stats_last();
    stats_["total_time"] = avg;
    const double products = stats_.contains("stats") && stats_["stats"].contains("intermediate_product")
                                ? stats_["stats"]["intermediate_product"].get<double>()
                                : 0.0;
    stats_["gflops"] = products * 2.0 / avg / 1e6;
    have_stats_ = true;
    return C;
  }

  // One product C = A * B on the device (no warmup loop; no counterpart).
  std::shared_ptr<cuCSR> run(const cuCSR &A, const cuCSR &B, void *stream = nullptr) {
    auto C = std::make_shared<cuCSR>();
    check(oc_spgemm_run(h_, &A.m, &B.m, &C->m, stream));
    C->sync();
    have_stats_ = false;
    return C;
  }

  // Stats of the last run(A, B) loop, or of the last single product.
  json get_stats() const { return have_stats_ ? stats_ : get_stats_last(); }
  double timer_ms(const char *name) const { return oc_spgemm_timer_ms(h_, name); }
  oc_spgemm *handle() { return h_; }

 private:
  json get_stats_last() const {
    size_t need = 0;
    check(oc_spgemm_stats_json(h_, nullptr, 0, &need));
    std::string s(need, '\0');
    check(oc_spgemm_stats_json(h_, &s[0], need, &need));
    s.resize(std::strlen(s.c_str()));
    return json::parse(s);
  }

  Config cfg_;
  oc_spgemm *h_;
  json stats_;
  bool have_stats_ = false;
};

}  // namespace somespgemm
