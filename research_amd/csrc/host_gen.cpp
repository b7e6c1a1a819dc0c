// Seeded synthetic matrices for the benchmark and test configurations.
// Every generator is deterministic in its seed (splitmix64).  Structure and
// values use separate streams derived from the seed, so the pattern does not
// depend on the value type.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "oc_internal.hpp"

using namespace oc;

namespace {

constexpr uint64_t kValueStream = 0xA5A5A5A5DEADBEEFull;

double draw_value(SplitMix64 &g, int dist) {
  if (dist == 0) return 2.0 * g.uniform() - 1.0;  // uniform [-1, 1)
  // N(0,1) by Box-Muller on (0,1] x [0,1)
  double u1 = 1.0 - g.uniform(), u2 = g.uniform();
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

void fill_values(oc_csr *m, uint64_t seed, int dist) {
  SplitMix64 g(seed ^ kValueStream);
  if (m->vtype == OC_F64) {
    double *v = (double *)m->values;
    for (uint64_t i = 0; i < m->nnz; ++i) v[i] = draw_value(g, dist);
  } else {
    float *v = (float *)m->values;
    for (uint64_t i = 0; i < m->nnz; ++i) v[i] = (float)draw_value(g, dist);
  }
}

// Build a host CSR from per-row sorted unique column lists.
int build(oc_csr *m, uint64_t rows, uint64_t cols,
          const std::vector<uint32_t> &rowptr, const std::vector<uint32_t> &colv,
          int vtype, uint64_t seed, int dist) {
  OC_TRY(oc_csr_alloc(m, rows, cols, colv.size(), vtype, 4, 0));
  std::memcpy(m->row_offsets, rowptr.data(), (rows + 1) * 4);
  if (!colv.empty()) std::memcpy(m->col_ids, colv.data(), colv.size() * 4);
  fill_values(m, seed, dist);
  return OC_OK;
}

uint32_t poisson(SplitMix64 &g, double lambda) {
  if (lambda <= 0) return 0;
  if (lambda < 60.0) {  // Knuth
    double l = std::exp(-lambda), p = 1.0;
    uint32_t k = 0;
    do {
      ++k;
      p *= g.uniform();
    } while (p > l);
    return k - 1;
  }
  double x = std::round(lambda + std::sqrt(lambda) * draw_value(g, 1));
  return x < 0 ? 0u : (uint32_t)x;
}

// LSD radix sort of 64-bit keys on their low `bits` bits.
void radix_sort_u64(std::vector<uint64_t> &a, int bits) {
  std::vector<uint64_t> tmp(a.size());
  for (int shift = 0; shift < bits; shift += 11) {
    size_t cnt[2049] = {0};
    for (uint64_t x : a) cnt[((x >> shift) & 2047) + 1]++;
    for (int i = 0; i < 2048; ++i) cnt[i + 1] += cnt[i];
    for (uint64_t x : a) tmp[cnt[(x >> shift) & 2047]++] = x;
    a.swap(tmp);
  }
}

}  // namespace

extern "C" int oc_gen_laplacian2d(uint32_t nx, uint32_t ny, int vtype,
                                  uint64_t seed, oc_csr *m) {
  OC_REQUIRE(nx > 0 && ny > 0 && (uint64_t)nx * ny < (1ull << 31),
             "laplacian: bad grid");
  uint64_t n = (uint64_t)nx * ny;
  std::vector<uint32_t> rowptr(n + 1, 0), colv;
  colv.reserve(5 * n);
  for (uint32_t y = 0; y < ny; ++y)
    for (uint32_t x = 0; x < nx; ++x) {
      uint32_t i = y * nx + x;
      if (y > 0) colv.push_back(i - nx);
      if (x > 0) colv.push_back(i - 1);
      colv.push_back(i);
      if (x + 1 < nx) colv.push_back(i + 1);
      if (y + 1 < ny) colv.push_back(i + nx);
      rowptr[i + 1] = (uint32_t)colv.size();
    }
  return build(m, n, n, rowptr, colv, vtype, seed, 0);
}

extern "C" int oc_gen_rmat(uint32_t log_n, uint64_t edges, double a, double b,
                           double c, int vtype, uint64_t seed, oc_csr *m) {
  OC_REQUIRE(log_n >= 1 && log_n <= 30, "rmat: log_n must be in [1,30]");
  OC_REQUIRE(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0, "rmat: bad probabilities");
  const uint64_t n = 1ull << log_n;
  // quadrant thresholds on 32-bit uniforms
  const uint64_t ta = (uint64_t)(a * 4294967296.0);
  const uint64_t tb = (uint64_t)((a + b) * 4294967296.0);
  const uint64_t tc = (uint64_t)((a + b + c) * 4294967296.0);
  std::vector<uint64_t> keys(edges);
  SplitMix64 g(seed);
  for (uint64_t e = 0; e < edges; ++e) {
    uint64_t r = 0, col = 0;
    uint64_t bits = 0;
    int have = 0;
    for (uint32_t l = 0; l < log_n; ++l) {
      if (have == 0) {
        bits = g.next();
        have = 2;
      }
      uint64_t u = bits & 0xFFFFFFFFull;
      bits >>= 32;
      --have;
      uint64_t rb = 0, cb = 0;
      if (u < ta) {
      } else if (u < tb) {
        cb = 1;
      } else if (u < tc) {
        rb = 1;
      } else {
        rb = 1;
        cb = 1;
      }
      r = (r << 1) | rb;
      col = (col << 1) | cb;
    }
    keys[e] = (r << log_n) | col;
  }
  radix_sort_u64(keys, 2 * (int)log_n);
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());  // merge duplicates
  std::vector<uint32_t> rowptr(n + 1, 0), colv(keys.size());
  const uint64_t mask = n - 1;
  for (size_t i = 0; i < keys.size(); ++i) {
    uint64_t r = keys[i] >> log_n;
    colv[i] = (uint32_t)(keys[i] & mask);
    rowptr[r + 1]++;
  }
  for (uint64_t i = 0; i < n; ++i) rowptr[i + 1] += rowptr[i];
  return build(m, n, n, rowptr, colv, vtype, seed, 0);
}

extern "C" int oc_gen_erdos_renyi(uint32_t n, double avg_degree, int vtype,
                                  uint64_t seed, oc_csr *m) {
  OC_REQUIRE(n > 0 && avg_degree >= 0 && avg_degree <= n, "erdos_renyi: bad args");
  SplitMix64 g(seed);
  std::vector<uint32_t> rowptr(n + 1, 0), colv, row;
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t d = std::min<uint32_t>(poisson(g, avg_degree), n);
    row.clear();
    for (uint32_t k = 0; k < d; ++k) row.push_back((uint32_t)(g.next() % n));
    std::sort(row.begin(), row.end());
    row.erase(std::unique(row.begin(), row.end()), row.end());
    colv.insert(colv.end(), row.begin(), row.end());
    rowptr[i + 1] = (uint32_t)colv.size();
  }
  return build(m, n, n, rowptr, colv, vtype, seed, 0);
}

extern "C" int oc_gen_banded(uint32_t n, uint32_t bandwidth, int vtype,
                             uint64_t seed, oc_csr *m) {
  OC_REQUIRE(n > 0 && bandwidth >= 1, "banded: bad args");
  uint32_t half = (bandwidth - 1) / 2;
  std::vector<uint32_t> rowptr(n + 1, 0), colv;
  colv.reserve((uint64_t)n * (2 * half + 1));
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t lo = i >= half ? i - half : 0;
    uint32_t hi = std::min<uint64_t>((uint64_t)i + half, n - 1);
    for (uint32_t j = lo; j <= hi; ++j) colv.push_back(j);
    rowptr[i + 1] = (uint32_t)colv.size();
  }
  return build(m, n, n, rowptr, colv, vtype, seed, 0);
}

extern "C" int oc_gen_rect_band(uint32_t rows, uint32_t cols, double mean,
                                double band_frac, uint32_t half_band,
                                uint32_t stride, int vtype, int dist,
                                uint64_t seed, oc_csr *m) {
  OC_REQUIRE(rows > 0 && cols > 0 && band_frac >= 0 && band_frac <= 1,
             "rect_band: bad args");
  SplitMix64 g(seed);
  std::vector<uint32_t> rowptr(rows + 1, 0), colv, row;
  for (uint32_t i = 0; i < rows; ++i) {
    uint32_t d = std::min<uint32_t>(poisson(g, mean), cols);
    row.clear();
    for (uint32_t k = 0; k < d; ++k) {
      if (g.uniform() < band_frac) {
        int64_t center = (int64_t)i * stride;
        int64_t off = (int64_t)(g.next() % (2ull * half_band + 1)) - (int64_t)half_band;
        int64_t cc = ((center + off) % (int64_t)cols + (int64_t)cols) % (int64_t)cols;
        row.push_back((uint32_t)cc);
      } else {
        row.push_back((uint32_t)(g.next() % cols));
      }
    }
    std::sort(row.begin(), row.end());
    row.erase(std::unique(row.begin(), row.end()), row.end());
    colv.insert(colv.end(), row.begin(), row.end());
    rowptr[i + 1] = (uint32_t)colv.size();
  }
  return build(m, rows, cols, rowptr, colv, vtype, seed, dist);
}
