// MatrixMarket coordinate files -> host CSR (the role of loadMTX and the
// COO -> CSR conversion of utils/convert.cpp:83-232).
//
// Accepted as the reference accepts them: "%%MatrixMarket matrix coordinate"
// headers with field real / integer / double / pattern (value 1) / complex
// (the real part), symmetry general / symmetric / Hermitian (off-diagonal
// entries mirrored with the same value); comment lines and blank lines are
// skipped; indices are 1-based and checked against the size line.
// Differences, all on inputs the reference mishandles or where it leaves
// the order open: a 0 index is an error (the reference wraps it to 2^32-1);
// more entries than the size line declares is an error (the reference
// writes past its buffer); entries with the same (row, column) keep file
// order (the reference's std::sort leaves it unspecified) and are summed
// when merge_duplicates is set; "hermitian" is accepted in either case;
// offsets are u64 when nnz >= 2^32.
#include <algorithm>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "oc_internal.hpp"

namespace {

struct Entry {
  uint32_t r, c;
  double v;
};

bool starts_with(const char *s, const char *p) { return std::strncmp(s, p, std::strlen(p)) == 0; }

// the whitespace-separated tokens of one line
std::vector<std::string> tokens(const char *b, const char *e) {
  std::vector<std::string> out;
  while (b < e) {
    while (b < e && (*b == ' ' || *b == '\t' || *b == '\r')) ++b;
    const char *s = b;
    while (b < e && !(*b == ' ' || *b == '\t' || *b == '\r')) ++b;
    if (b > s) out.emplace_back(s, b);
  }
  return out;
}

bool blank(const char *b, const char *e) {
  for (; b < e; ++b)
    if (!(*b == ' ' || *b == '\t' || *b == '\r' || *b == '\v' || *b == '\f')) return false;
  return true;
}

// unsigned integer at *p (leading blanks skipped); false when none
bool read_u64(const char *&p, const char *e, uint64_t &out) {
  while (p < e && (*p == ' ' || *p == '\t')) ++p;
  if (p >= e || *p < '0' || *p > '9') return false;
  uint64_t v = 0;
  while (p < e && *p >= '0' && *p <= '9') {
    v = v * 10 + (uint64_t)(*p - '0');
    if (v > (1ull << 40)) return false;
    ++p;
  }
  out = v;
  return true;
}

// number at *p parsed as T's own type (strtof for float, so a value rounds
// once, as the reference's stream extraction into float does)
bool read_value(const char *&p, const char *e, int vtype, double &out) {
  while (p < e && (*p == ' ' || *p == '\t')) ++p;
  if (p >= e) return false;
  char buf[128];
  size_t n = 0;
  while (p + n < e && n < sizeof(buf) - 1 && !(p[n] == ' ' || p[n] == '\t' || p[n] == '\r'))
    ++n;
  std::memcpy(buf, p, n);
  buf[n] = 0;
  char *end = nullptr;
  errno = 0;
  if (vtype == OC_F32) out = (double)std::strtof(buf, &end);
  else out = std::strtod(buf, &end);
  if (end == buf || *end != 0) return false;
  p += n;
  return true;
}

}  // namespace

extern "C" int oc_mtx_load(const char *path, int vtype, int merge_duplicates, oc_csr *m) {
  OC_REQUIRE(path && m, "oc_mtx_load: null argument");
  OC_REQUIRE(vtype == OC_F64 || vtype == OC_F32, "oc_mtx_load: bad vtype %d", vtype);
  std::memset(m, 0, sizeof(*m));
  FILE *f = std::fopen(path, "rb");
  OC_REQUIRE(f != nullptr, "could not open \"%s\"", path);
  std::string text;
  {
    char buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, got);
    std::fclose(f);
  }
  const char *p = text.data(), *end = p + text.size();
  auto next_line = [&](const char *&b, const char *&e) {
    if (p >= end) return false;
    b = p;
    const char *nl = static_cast<const char *>(std::memchr(p, '\n', (size_t)(end - p)));
    e = nl ? nl : end;
    p = nl ? nl + 1 : end;
    return true;
  };
  const char *lb, *le;
  OC_REQUIRE(next_line(lb, le) && starts_with(lb, "%%MatrixMarket matrix coordinate"),
             "Can only read MatrixMarket format that is in coordinate form");
  const std::vector<std::string> head = tokens(lb, le);
  OC_REQUIRE(head.size() >= 5, "MatrixMarket header has too few fields");
  const std::string &field = head[3], &sym = head[4];
  const bool pattern = field == "pattern";
  OC_REQUIRE(pattern || field == "complex" || field == "real" || field == "integer" ||
                 field == "double",
             "MatrixMarket data type does not match matrix format");
  bool mirror = false;
  if (sym == "general") mirror = false;
  else if (sym == "symmetric" || sym == "Hermitian" || sym == "hermitian") mirror = true;
  else
    OC_REQUIRE(false, "Can only read MatrixMarket format that is either symmetric, general or "
                      "hermitian");
  uint64_t rows = 0, cols = 0, declared = 0;
  size_t line = 0;
  bool have_size = false;
  while (next_line(lb, le)) {
    ++line;
    if (lb < le && *lb == '%') continue;
    const char *q = lb;
    OC_REQUIRE(read_u64(q, le, rows) && read_u64(q, le, cols) && read_u64(q, le, declared),
               "Failed to read matrix market header from \"%s\"", path);
    have_size = true;
    break;
  }
  OC_REQUIRE(have_size, "Failed to read matrix market header from \"%s\"", path);
  OC_REQUIRE(rows < (1ull << 32) && cols < (1ull << 32), "matrix dimensions beyond 2^32 - 1");
  std::vector<Entry> ent;
  ent.reserve((size_t)declared * (mirror ? 2 : 1));
  uint64_t listed = 0;
  while (next_line(lb, le)) {
    ++line;
    if (lb < le && *lb == '%') continue;
    if (blank(lb, le)) continue;
    const char *q = lb;
    uint64_t r = 0, c = 0;
    double v = 1.0;
    const bool ok = read_u64(q, le, r) && read_u64(q, le, c) && (pattern || read_value(q, le, vtype, v));
    OC_REQUIRE(ok, "Failed to read data at line %zu from matrix market file \"%s\"", line, path);
    OC_REQUIRE(r >= 1 && r <= rows, "Row index out of bounds at line %zu in matrix market file \"%s\"",
               line, path);
    OC_REQUIRE(c >= 1 && c <= cols,
               "Column index out of bounds at line %zu in matrix market file \"%s\"", line, path);
    OC_REQUIRE(++listed <= declared,
               "more entries than the size line declares (%llu) in matrix market file \"%s\"",
               (unsigned long long)declared, path);
    ent.push_back(Entry{(uint32_t)(r - 1), (uint32_t)(c - 1), v});
    if (mirror && r != c) ent.push_back(Entry{(uint32_t)(c - 1), (uint32_t)(r - 1), v});
  }
  // rows by a counting pass (file order kept), then each row stably by column
  std::vector<uint64_t> ro(rows + 1, 0);
  for (const Entry &e : ent) ++ro[e.r + 1];
  for (uint64_t i = 0; i < rows; ++i) ro[i + 1] += ro[i];
  std::vector<Entry> sorted(ent.size());
  {
    std::vector<uint64_t> pos(ro.begin(), ro.end() - 1);
    for (const Entry &e : ent) sorted[pos[e.r]++] = e;
  }
  ent.clear();
  ent.shrink_to_fit();
  for (uint64_t i = 0; i < rows; ++i)
    std::stable_sort(sorted.begin() + (long)ro[i], sorted.begin() + (long)ro[i + 1],
                     [](const Entry &a, const Entry &b) { return a.c < b.c; });
  if (merge_duplicates) {
    size_t w = 0;
    std::vector<uint64_t> nro(rows + 1, 0);
    for (uint64_t i = 0; i < rows; ++i) {
      for (uint64_t k = ro[i]; k < ro[i + 1]; ++k) {
        if (w > nro[i] && sorted[w - 1].c == sorted[k].c) sorted[w - 1].v += sorted[k].v;
        else sorted[w++] = sorted[k];
      }
      nro[i + 1] = w;
    }
    sorted.resize(w);
    ro.swap(nro);
  }
  const uint64_t nnz = sorted.size();
  OC_TRY(oc_csr_alloc(m, rows, cols, nnz, vtype, nnz >= (1ull << 32) ? 8 : 4, 0));
  for (uint64_t k = 0; k < nnz; ++k) {
    m->col_ids[k] = sorted[k].c;
    if (vtype == OC_F64) static_cast<double *>(m->values)[k] = sorted[k].v;
    else static_cast<float *>(m->values)[k] = (float)sorted[k].v;
  }
  for (uint64_t i = 0; i <= rows; ++i) {
    if (m->off_bytes == 8) static_cast<uint64_t *>(m->row_offsets)[i] = ro[i];
    else static_cast<uint32_t *>(m->row_offsets)[i] = (uint32_t)ro[i];
  }
  return OC_OK;
}
