// Config defaults and the JSON config reader.
//
// Defaults follow Config's member initialisers (include/Utils.h:22-107);
// the accepted keys are those of from_json (src/Utils.cpp:8-37), with the
// same `track_stage_time` spelling, plus the learned-binning keys of this
// library.  Unknown keys are ignored, as nlohmann's assign_if_present does.
// The reader handles any JSON text; nested values are skipped.
#include <cctype>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "oc_internal.hpp"

using namespace oc;

extern "C" void oc_config_default(oc_config *c) {
  std::memset(c, 0, sizeof(*c));
  c->use_estimation_workflow = 1;
  c->use_ultrasparse_workflow = 1;
  c->use_fat_workflow = 1;
  c->use_assist_information = 1;
  c->use_hybrid_accumulator = 1;
  c->use_indirect_sort = 1;
  c->use_segmented_workflow = 0;
  c->segmentation_grid_max = 256;
  c->segmentation_hot_threshold_k = 2.0;
  c->segmentation_min_segment_rows = 1024;
  c->segmentation_cc_max_iters = 32;
  c->usparse_avg_product_threshold = 64;
  c->estimation_input_expansion_threshold = 8;
  c->estimation_output_compaction_threshold = 8;
  c->dense_query_bitmap_compaction_threshold = 2;
  c->numeric_hybrid_hash_kernel_row_length_threshold = 70000;
  c->fat_ratio = 100;
  c->hll_sample_rate = 0.03;
  c->hll_sample_min_rows = 648;
  c->hll_sample_max_rows = 10800;
  c->hll_expansion_low = 2.0;
  c->hll_expansion_high = 1.5;
  c->hll_numeric_malloc_expansion_coe = 1.3;
  c->hll_higher_precision_expansion_threshold = 48;
  c->hll_precision_low = 5;
  c->hll_precision_high = 6;
  c->estimate_compaction_target_confidence_interval = 0.99;
  c->estimate_compaction_target_z_value = 2.53;
  c->symbolic_expansion_coe = 1.25;
  c->numeric_expansion_coe = 1.5;
  c->track_stage_times = 0;
  c->check_correctness = 0;
  c->output_stats = 0;
  c->compare_relative_error = 1e-6;
  c->compare_autofail_relative_error = 1e-4;
  c->warmup_iters = 10;
  c->bench_iters = 10;
  c->print_info = 0;
  c->print_memory = 0;
  c->use_learned_binning = 0;
  c->use_tiled_numeric = 3;
  c->tile_log2_cols = 14;
  c->tile_block = 1024;
  c->tile_products_cap = 0;
  c->learned_quantile = 0.99;
  c->band_log2_cols = 0;
  c->band_min_products = 16384;
  c->learned_min_compression = 1.5;
  for (int b = 0; b < 7; ++b) c->bin_concurrency[b] = c->bin_block[b] = 0;
}

namespace {

struct JValue {
  enum Kind { NUM, BOOL, STR, NUL, ARR, OTHER } kind = OTHER;
  double num = 0;
  bool b = false;
  std::string s;
  std::vector<double> arr;  // ARR: a flat array of numbers
};

struct Parser {
  const char *p;
  std::string err;
  explicit Parser(const char *t) : p(t) {}

  void ws() {
    while (*p && std::isspace((unsigned char)*p)) ++p;
  }
  bool str(std::string &out) {
    if (*p != '"') return fail("expected string");
    ++p;
    out.clear();
    while (*p && *p != '"') {
      if (*p == '\\') {
        ++p;
        switch (*p) {
          case 'n': out += '\n'; break;
          case 't': out += '\t'; break;
          case 'r': out += '\r'; break;
          case 'b': out += '\b'; break;
          case 'f': out += '\f'; break;
          case 'u': {  // keep ASCII, replace the rest
            unsigned v = 0;
            for (int i = 0; i < 4; ++i) {
              ++p;
              if (!std::isxdigit((unsigned char)*p)) return fail("bad \\u escape");
              v = v * 16 + (unsigned)(std::isdigit((unsigned char)*p)
                                          ? *p - '0'
                                          : (std::tolower(*p) - 'a' + 10));
            }
            out += v < 128 ? (char)v : '?';
            break;
          }
          case '\0': return fail("unterminated string");
          default: out += *p;
        }
        ++p;
      } else {
        out += *p++;
      }
    }
    if (*p != '"') return fail("unterminated string");
    ++p;
    return true;
  }
  bool fail(const char *m) {
    if (err.empty()) err = m;
    return false;
  }
  bool value(JValue &v) {
    ws();
    if (*p == '"') {
      v.kind = JValue::STR;
      return str(v.s);
    }
    if (*p == '[') {  // a flat array of numbers, else skipped
      const char *start = p;
      ++p;
      ws();
      v.arr.clear();
      bool flat = true;
      while (*p && *p != ']') {
        char *end = nullptr;
        const double x = std::strtod(p, &end);
        if (end == p) {
          flat = false;
          break;
        }
        v.arr.push_back(x);
        p = end;
        ws();
        if (*p == ',') {
          ++p;
          ws();
        }
      }
      if (flat && *p == ']') {
        ++p;
        v.kind = JValue::ARR;
        return true;
      }
      p = start;
      v.kind = JValue::OTHER;
      return skip_nested();
    }
    if (*p == '{') {  // skip nested structure
      v.kind = JValue::OTHER;
      return skip_nested();
    }
    if (!std::strncmp(p, "true", 4)) {
      p += 4;
      v.kind = JValue::BOOL;
      v.b = true;
      return true;
    }
    if (!std::strncmp(p, "false", 5)) {
      p += 5;
      v.kind = JValue::BOOL;
      v.b = false;
      return true;
    }
    if (!std::strncmp(p, "null", 4)) {
      p += 4;
      v.kind = JValue::NUL;
      return true;
    }
    char *end = nullptr;
    v.num = std::strtod(p, &end);
    if (end == p) return fail("bad value");
    p = end;
    v.kind = JValue::NUM;
    return true;
  }
  bool skip_nested() {
    int depth = 0;
    do {
      if (*p == '"') {
        std::string tmp;
        if (!str(tmp)) return false;
        continue;
      }
      if (*p == '{' || *p == '[') ++depth;
      if (*p == '}' || *p == ']') --depth;
      if (!*p) return fail("unterminated structure");
      ++p;
    } while (depth > 0);
    return true;
  }
};

bool as_bool(const JValue &v, int32_t &out) {
  if (v.kind == JValue::BOOL) {
    out = v.b;
    return true;
  }
  if (v.kind == JValue::NUM) {
    out = v.num != 0;
    return true;
  }
  return false;
}
bool as_int(const JValue &v, int32_t &out) {
  if (v.kind == JValue::NUM) {
    out = (int32_t)v.num;
    return true;
  }
  if (v.kind == JValue::BOOL) {
    out = v.b;
    return true;
  }
  return false;
}
bool as_dbl(const JValue &v, double &out) {
  if (v.kind == JValue::NUM) {
    out = v.num;
    return true;
  }
  return false;
}
bool as_ints(const JValue &v, int32_t *out, size_t n) {
  if (v.kind != JValue::ARR || v.arr.size() != n) return false;
  for (size_t i = 0; i < n; ++i) out[i] = (int32_t)v.arr[i];
  return true;
}
bool as_str(const JValue &v, char *out, size_t cap) {
  if (v.kind != JValue::STR) return false;
  std::strncpy(out, v.s.c_str(), cap - 1);
  out[cap - 1] = '\0';
  return true;
}

}  // namespace

extern "C" int oc_config_check(const oc_config *c) {
  OC_REQUIRE(c, "oc_config_check: null argument");
  // Bins are sized from products (symbolic) or exact counts (numeric) times
  // these coefficients; below 1 a row can be given a table smaller than its
  // output.
  OC_REQUIRE(c->symbolic_expansion_coe >= 1.0, "symbolic_expansion_coe %g is below 1",
             c->symbolic_expansion_coe);
  OC_REQUIRE(c->numeric_expansion_coe >= 1.0, "numeric_expansion_coe %g is below 1",
             c->numeric_expansion_coe);
  OC_REQUIRE(c->fat_ratio >= 1, "fat_ratio must be at least 1");
  OC_REQUIRE((c->hll_precision_low == 5 || c->hll_precision_low == 6) &&
                 (c->hll_precision_high == 5 || c->hll_precision_high == 6),
             "HyperLogLog precisions must be 5 or 6");
  OC_REQUIRE(c->hll_sample_min_rows >= 1 && c->hll_sample_max_rows >= c->hll_sample_min_rows,
             "need 1 <= hll_sample_min_rows <= hll_sample_max_rows");
  OC_REQUIRE(c->use_tiled_numeric >= 0 && c->use_tiled_numeric <= 3,
             "use_tiled_numeric must be 0..3");
  OC_REQUIRE(c->tile_block == 256 || c->tile_block == 512 || c->tile_block == 1024,
             "tile_block must be 256, 512 or 1024");
  OC_REQUIRE(c->tile_log2_cols >= 10 && c->tile_log2_cols <= 16, "tile_log2_cols must be 10..16");
  OC_REQUIRE(c->band_log2_cols == 0 || (c->band_log2_cols >= 10 && c->band_log2_cols <= 15),
             "band_log2_cols must be 0 or 10..15");
  for (int b = 0; b < 7; ++b) {
    const int32_t cc = c->bin_concurrency[b], bb = c->bin_block[b];
    OC_REQUIRE(cc >= 0 && cc <= 64, "bin_concurrency[%d] must be 0 (default) or 1..64", b);
    OC_REQUIRE(bb == 0 || bb == 64 || bb == 128 || bb == 256 || bb == 512 || bb == 1024,
               "bin_block[%d] must be 0 (default) or 64, 128, 256, 512, 1024", b);
  }
  return OC_OK;
}

extern "C" int oc_config_from_json(oc_config *c, const char *text) {
  OC_REQUIRE(c && text, "oc_config_from_json: null argument");
  Parser ps(text);
  ps.ws();
  OC_REQUIRE(*ps.p == '{', "config JSON: expected an object");
  ++ps.p;
  ps.ws();
  bool first = true;
  while (*ps.p && *ps.p != '}') {
    if (!first) {
      OC_REQUIRE(*ps.p == ',', "config JSON: expected ','");
      ++ps.p;
      ps.ws();
    }
    first = false;
    std::string key;
    OC_REQUIRE(ps.str(key), "config JSON: %s", ps.err.c_str());
    ps.ws();
    OC_REQUIRE(*ps.p == ':', "config JSON: expected ':' after \"%s\"", key.c_str());
    ++ps.p;
    JValue v;
    OC_REQUIRE(ps.value(v), "config JSON: %s at key \"%s\"", ps.err.c_str(), key.c_str());
    ps.ws();
    bool ok = true;
    const char *k = key.c_str();
#define B(name, field) else if (!std::strcmp(k, name)) ok = as_bool(v, c->field);
#define I(name, field) else if (!std::strcmp(k, name)) ok = as_int(v, c->field);
#define D(name, field) else if (!std::strcmp(k, name)) ok = as_dbl(v, c->field);
#define S(name, field) \
  else if (!std::strcmp(k, name)) ok = as_str(v, c->field, sizeof(c->field));
#define IA(name, field) else if (!std::strcmp(k, name)) ok = as_ints(v, c->field, 7);
    if (false) {
    }
    // src/Utils.cpp:9-36
    B("track_stage_time", track_stage_times)
    B("check_correctness", check_correctness)
    B("output_stats", output_stats)
    D("compare_relative_error", compare_relative_error)
    D("compare_autofail_relative_error", compare_autofail_relative_error)
    S("ground_truth_file", ground_truth_file)
    S("stats_output_file", stats_output_file)
    I("warmup_iters", warmup_iters)
    I("bench_iters", bench_iters)
    B("print_info", print_info)
    B("print_memory", print_memory)
    B("use_estimation_workflow", use_estimation_workflow)
    B("use_ultrasparse_workflow", use_ultrasparse_workflow)
    B("use_assist_information", use_assist_information)
    B("use_hybrid_accumulator", use_hybrid_accumulator)
    B("use_indirect_sort", use_indirect_sort)
    B("use_fat_workflow", use_fat_workflow)
    D("hll_numeric_malloc_expansion_coe", hll_numeric_malloc_expansion_coe)
    I("estimation_input_expansion_threshold", estimation_input_expansion_threshold)
    I("estimation_output_compaction_threshold", estimation_output_compaction_threshold)
    B("use_segmented_workflow", use_segmented_workflow)
    I("segmentation_grid_max", segmentation_grid_max)
    D("segmentation_hot_threshold_k", segmentation_hot_threshold_k)
    I("segmentation_min_segment_rows", segmentation_min_segment_rows)
    I("segmentation_cc_max_iters", segmentation_cc_max_iters)
    // learned row-size predictor (no reference counterpart)
    B("use_learned_binning", use_learned_binning)
    S("learned_model_file", learned_model_file)
    I("use_tiled_numeric", use_tiled_numeric)
    I("tile_log2_cols", tile_log2_cols)
    I("tile_block", tile_block)
    I("tile_products_cap", tile_products_cap)
    D("learned_quantile", learned_quantile)
    I("band_log2_cols", band_log2_cols)
    I("band_min_products", band_min_products)
    D("learned_min_compression", learned_min_compression)
    IA("bin_concurrency", bin_concurrency)
    IA("bin_block", bin_block)
#undef IA
#undef B
#undef I
#undef D
#undef S
    OC_REQUIRE(ok, "config JSON: key \"%s\" has the wrong type", k);
  }
  OC_REQUIRE(*ps.p == '}', "config JSON: unterminated object");
  return oc_config_check(c);
}

extern "C" int oc_config_load(oc_config *c, const char *path) {
  OC_REQUIRE(c && path, "oc_config_load: null argument");
  std::ifstream in(path);
  if (!in.is_open()) {
    set_error("could not open config \"%s\"", path);
    return OC_ERR_IO;
  }
  std::stringstream ss;
  ss << in.rdbuf();
  std::string text = ss.str();
  return oc_config_from_json(c, text.c_str());
}
