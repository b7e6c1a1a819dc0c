// Host-side CSR handling: allocation, host<->device copies, the .csr binary
// file format, compare() and transpose.
//
// .csr layout (src/CSR.cpp:58-116, utils/convert.cpp:256-309): a 9-byte
// magic {'H','i',1,'C','o','m','p','s','d'}, 7 bytes of padding, eight u64
// {typesize, compresseddir, indexsize, fixedoffset, offsetsize, rows, cols,
// nnz} (80 bytes total), then State<T>{T scaling; bool transpose} padded to
// 16 (f64) or 8 (f32) bytes, then values[nnz], col_ids[nnz] (u32) and
// row_offsets[rows+1] (offsetsize bytes each).  The format has to match
// byte for byte, so the field order is the reference's; the padding bytes,
// uninitialised in the reference writer, are written as zero here.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "oc_internal.hpp"

namespace oc {

namespace {
struct OutBlock {
  int dev;
  size_t size;
};
struct OutCache {
  std::mutex mu;
  std::unordered_map<void *, OutBlock> blocks;       // every block handed out or cached
  std::map<int, std::multimap<size_t, void *>> idle;  // per device: cached blocks by size
  std::map<int, size_t> idle_bytes;
};
OutCache &out_cache() {
  static OutCache *c = new OutCache();  // never destroyed: freed at process exit
  return *c;
}
// Gives the idle blocks of device dev back to the driver, smallest first,
// until at most keep bytes stay cached (lock held).
void out_shrink(OutCache &c, int dev, size_t keep) {
  auto &idle = c.idle[dev];
  while (c.idle_bytes[dev] > keep && !idle.empty()) {
    auto it = idle.begin();
    (void)hipFree(it->second);
    c.idle_bytes[dev] -= it->first;
    c.blocks.erase(it->second);
    idle.erase(it);
  }
}
}  // namespace

int out_alloc(void **p, size_t bytes) {
  const size_t g = ((std::max<size_t>(bytes, 1) + (2u << 20) - 1) >> 21) << 21;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    set_error("out_alloc: no current device");
    return OC_ERR_HIP;
  }
  OutCache &c = out_cache();
  std::lock_guard<std::mutex> lk(c.mu);
  auto &idle = c.idle[dev];
  // the smallest idle block that holds the request; requests under 64 MiB
  // take blocks of at most four times their size
  auto it = idle.lower_bound(g);
  if (it != idle.end() && (g >= (64u << 20) || it->first <= 4 * g)) {
    *p = it->second;
    c.idle_bytes[dev] -= it->first;
    idle.erase(it);
    return OC_OK;
  }
  hipError_t e = hipMalloc(p, g);
  if (e == hipErrorOutOfMemory) {  // give the cached blocks back and retry
    (void)hipGetLastError();
    out_shrink(c, dev, 0);
    e = hipMalloc(p, g);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    set_error("device allocation of %zu bytes for a product output failed: %s", g,
              hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? OC_ERR_NOMEM : OC_ERR_HIP;
  }
  c.blocks[*p] = OutBlock{dev, g};
  return OC_OK;
}

void out_free(void *p) {
  if (!p) return;
  (void)hipDeviceSynchronize();  // as hipFree: no work may still use it
  OutCache &c = out_cache();
  std::lock_guard<std::mutex> lk(c.mu);
  auto b = c.blocks.find(p);
  if (b == c.blocks.end()) {  // not from the cache
    (void)hipFree(p);
    return;
  }
  const int dev = b->second.dev;
  c.idle[dev].emplace(b->second.size, p);
  c.idle_bytes[dev] += b->second.size;
}

void out_release() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    return;
  }
  OutCache &c = out_cache();
  std::lock_guard<std::mutex> lk(c.mu);
  out_shrink(c, dev, 0);
}

static thread_local std::string g_err;

void set_error(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}

const char *get_error() { return g_err.c_str(); }

}  // namespace oc

using namespace oc;

extern "C" const char *oc_last_error(void) { return oc::get_error(); }

static int hip_fail(hipError_t e, const char *what) {
  set_error("%s: %s", what, hipGetErrorString(e));
  return e == hipErrorOutOfMemory ? OC_ERR_NOMEM : OC_ERR_HIP;
}

extern "C" int oc_csr_alloc(oc_csr *m, uint64_t rows, uint64_t cols,
                            uint64_t nnz, int vtype, int off_bytes,
                            int on_device) {
  OC_REQUIRE(m != nullptr, "oc_csr_alloc: null matrix");
  OC_REQUIRE(vtype == OC_F64 || vtype == OC_F32, "oc_csr_alloc: bad vtype %d",
             vtype);
  OC_REQUIRE(off_bytes == 4 || off_bytes == 8,
             "oc_csr_alloc: off_bytes must be 4 or 8");
  std::memset(m, 0, sizeof(*m));
  m->rows = rows;
  m->cols = cols;
  m->nnz = nnz;
  m->vtype = vtype;
  m->off_bytes = off_bytes;
  m->on_device = on_device;
  m->owned = 1;
  size_t ob = (size_t)(rows + 1) * off_bytes;
  size_t cb = (size_t)(nnz ? nnz : 1) * 4;
  size_t vb = (size_t)(nnz ? nnz : 1) * vsize(vtype);
  if (on_device) {
    hipError_t e;
    if ((e = hipMalloc(&m->row_offsets, ob)) != hipSuccess ||
        (e = hipMalloc((void **)&m->col_ids, cb)) != hipSuccess ||
        (e = hipMalloc(&m->values, vb)) != hipSuccess) {
      oc_csr_free(m);
      return hip_fail(e, "oc_csr_alloc hipMalloc");
    }
  } else {
    m->row_offsets = std::malloc(ob);
    m->col_ids = (uint32_t *)std::malloc(cb);
    m->values = std::malloc(vb);
    if (!m->row_offsets || !m->col_ids || !m->values) {
      oc_csr_free(m);
      set_error("oc_csr_alloc: host allocation of %zu bytes failed",
                ob + cb + vb);
      return OC_ERR_NOMEM;
    }
  }
  return OC_OK;
}

extern "C" void oc_csr_free(oc_csr *m) {
  if (!m) return;
  if (m->owned == 3) {  // product output from the output cache
    out_free(m->row_offsets);
    out_free(m->col_ids);
    out_free(m->values);
  } else if (m->owned == 2) {  // stream-ordered pool allocation
    if (m->row_offsets) (void)hipFreeAsync(m->row_offsets, nullptr);
    if (m->col_ids) (void)hipFreeAsync(m->col_ids, nullptr);
    if (m->values) (void)hipFreeAsync(m->values, nullptr);
    (void)hipStreamSynchronize(nullptr);
  } else if (m->owned) {
    if (m->on_device) {
      if (m->row_offsets) (void)hipFree(m->row_offsets);
      if (m->col_ids) (void)hipFree(m->col_ids);
      if (m->values) (void)hipFree(m->values);
    } else {
      std::free(m->row_offsets);
      std::free(m->col_ids);
      std::free(m->values);
    }
  }
  std::memset(m, 0, sizeof(*m));
}

static int copy_csr(const oc_csr *src, oc_csr *dst, int to_device,
                    void *stream) {
  OC_TRY(oc_csr_alloc(dst, src->rows, src->cols, src->nnz, src->vtype,
                      src->off_bytes, to_device));
  hipStream_t s = (hipStream_t)stream;
  hipMemcpyKind kind = to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
  size_t ob = (size_t)(src->rows + 1) * src->off_bytes;
  hipError_t e;
  if ((e = hipMemcpyAsync(dst->row_offsets, src->row_offsets, ob, kind, s)) != hipSuccess ||
      (src->nnz &&
       (e = hipMemcpyAsync(dst->col_ids, src->col_ids, src->nnz * 4, kind, s)) != hipSuccess) ||
      (src->nnz && (e = hipMemcpyAsync(dst->values, src->values,
                                       src->nnz * vsize(src->vtype), kind, s)) != hipSuccess) ||
      (e = hipStreamSynchronize(s)) != hipSuccess) {
    oc_csr_free(dst);
    return hip_fail(e, to_device ? "oc_csr_upload" : "oc_csr_download");
  }
  return OC_OK;
}

extern "C" int oc_csr_upload(const oc_csr *host, oc_csr *dev, void *stream) {
  OC_REQUIRE(host && dev && !host->on_device, "oc_csr_upload: need a host matrix");
  return copy_csr(host, dev, 1, stream);
}

extern "C" int oc_csr_download(const oc_csr *dev, oc_csr *host, void *stream) {
  OC_REQUIRE(dev && host && dev->on_device, "oc_csr_download: need a device matrix");
  return copy_csr(dev, host, 0, stream);
}

// ------------------------------------------------------------------ .csr
namespace {
struct CsrHeader {  // 80 bytes, field order fixed by the file format
  char magic[9];
  char pad[7];
  uint64_t typesize, compresseddir, indexsize, fixedoffset, offsetsize;
  uint64_t rows, cols, nnz;
};
static_assert(sizeof(CsrHeader) == 80, "header must be 80 bytes");
const char kMagic[9] = {'H', 'i', 1, 'C', 'o', 'm', 'p', 's', 'd'};
}  // namespace

extern "C" int oc_csr_load(const char *path, int vtype_expected, oc_csr *m) {
  OC_REQUIRE(path && m, "oc_csr_load: null argument");
  FILE *f = std::fopen(path, "rb");
  if (!f) {
    set_error("could not open \"%s\"", path);
    return OC_ERR_IO;
  }
  CsrHeader h;
  int rc = OC_OK;
  if (std::fread(&h, sizeof(h), 1, f) != 1) {
    set_error("Could not read CSR header");
    rc = OC_ERR_IO;
  } else if (std::memcmp(h.magic, kMagic, 9) != 0) {
    set_error("File does not appear to be a CSR Matrix");
    rc = OC_ERR_IO;
  } else if (h.typesize != 4 && h.typesize != 8) {
    set_error("unsupported value typesize %llu", (unsigned long long)h.typesize);
    rc = OC_ERR_IO;
  } else if (vtype_expected >= 0 && h.typesize != vsize(vtype_expected)) {
    set_error("File does not contain a CSR matrix with matching type (%llu != %zu)",
              (unsigned long long)h.typesize, vsize(vtype_expected));
    rc = OC_ERR_IO;
  } else if (h.offsetsize != 4 && h.offsetsize != 8) {
    set_error("unsupported offsetsize %llu", (unsigned long long)h.offsetsize);
    rc = OC_ERR_IO;
  }
  if (rc != OC_OK) {
    std::fclose(f);
    return rc;
  }
  int vtype = h.typesize == 4 ? OC_F32 : OC_F64;
  size_t state_bytes = h.typesize == 8 ? 16 : 8;
  char state[16];
  if (std::fread(state, state_bytes, 1, f) != 1) {
    std::fclose(f);
    set_error("Could not read CompressedMatrix state");
    return OC_ERR_IO;
  }
  rc = oc_csr_alloc(m, h.rows, h.cols, h.nnz, vtype, (int)h.offsetsize, 0);
  if (rc != OC_OK) {
    std::fclose(f);
    return rc;
  }
  bool ok = std::fread(m->values, h.typesize, h.nnz, f) == h.nnz &&
            std::fread(m->col_ids, 4, h.nnz, f) == h.nnz &&
            std::fread(m->row_offsets, h.offsetsize, h.rows + 1, f) == h.rows + 1;
  std::fclose(f);
  if (!ok) {
    oc_csr_free(m);
    set_error("Could not read CSR matrix data");
    return OC_ERR_IO;
  }
  return OC_OK;
}

static uint64_t host_off(const oc_csr *m, uint64_t i) {
  return m->off_bytes == 8 ? ((const uint64_t *)m->row_offsets)[i]
                           : ((const uint32_t *)m->row_offsets)[i];
}

extern "C" int oc_csr_store(const char *path, const oc_csr *m) {
  OC_REQUIRE(path && m && !m->on_device, "oc_csr_store: need a host matrix");
  FILE *f = std::fopen(path, "wb");
  if (!f) {
    set_error("could not open \"%s\"", path);
    return OC_ERR_IO;
  }
  CsrHeader h;
  std::memset(&h, 0, sizeof(h));
  std::memcpy(h.magic, kMagic, 9);
  h.typesize = vsize(m->vtype);
  h.compresseddir = 0;
  h.indexsize = 4;
  h.fixedoffset = 0;
  // u32 offsets whenever they fit, so the reference loader reads the file
  bool wide = m->nnz >= 0xFFFFFFFFull;
  h.offsetsize = wide ? 8 : 4;
  h.rows = m->rows;
  h.cols = m->cols;
  h.nnz = m->nnz;
  unsigned char state[16];
  std::memset(state, 0, sizeof(state));
  if (m->vtype == OC_F64) {
    double one = 1.0;
    std::memcpy(state, &one, 8);
  } else {
    float one = 1.0f;
    std::memcpy(state, &one, 4);
  }
  size_t state_bytes = m->vtype == OC_F64 ? 16 : 8;
  bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1 &&
            std::fwrite(state, state_bytes, 1, f) == 1 &&
            std::fwrite(m->values, vsize(m->vtype), m->nnz, f) == m->nnz &&
            std::fwrite(m->col_ids, 4, m->nnz, f) == m->nnz;
  if (ok) {
    if (wide == (m->off_bytes == 8)) {
      ok = std::fwrite(m->row_offsets, m->off_bytes, m->rows + 1, f) == m->rows + 1;
    } else if (!wide) {
      std::vector<uint32_t> tmp(m->rows + 1);
      for (uint64_t i = 0; i <= m->rows; ++i) tmp[i] = (uint32_t)host_off(m, i);
      ok = std::fwrite(tmp.data(), 4, m->rows + 1, f) == m->rows + 1;
    } else {
      std::vector<uint64_t> tmp(m->rows + 1);
      for (uint64_t i = 0; i <= m->rows; ++i) tmp[i] = host_off(m, i);
      ok = std::fwrite(tmp.data(), 8, m->rows + 1, f) == m->rows + 1;
    }
  }
  if (std::fclose(f) != 0) ok = false;
  if (!ok) {
    set_error("could not write \"%s\"", path);
    return OC_ERR_IO;
  }
  return OC_OK;
}

static double host_val(const oc_csr *m, uint64_t i) {
  return m->vtype == OC_F64 ? ((const double *)m->values)[i]
                            : (double)((const float *)m->values)[i];
}

extern "C" int oc_csr_compare(const oc_csr *a, const oc_csr *b, double rel_err,
                              double autofail, oc_compare_result *res) {
  OC_REQUIRE(a && b && !a->on_device && !b->on_device,
             "oc_csr_compare: need two host matrices");
  oc_compare_result r;
  std::memset(&r, 0, sizeof(r));
  r.first_bad = ~0ull;
  int rc = OC_OK;
  if (a->rows != b->rows || a->cols != b->cols || a->nnz != b->nnz) {
    r.kind = 1;
    set_error("Matrices have different dimensions");
    rc = OC_ERR_MISMATCH;
  }
  // row_offsets[0..rows) as in src/CSR.cpp:264-272 (the last one is not
  // compared there; nnz equality already covers it)
  for (uint64_t i = 0; rc == OC_OK && i < a->rows; ++i) {
    if (host_off(a, i) != host_off(b, i)) {
      r.kind = 2;
      r.first_bad = i;
      set_error("Matrices have different row offsets (row %llu)",
                (unsigned long long)i);
      rc = OC_ERR_MISMATCH;
    }
  }
  for (uint64_t i = 0; rc == OC_OK && i < a->nnz; ++i) {
    if (a->col_ids[i] != b->col_ids[i]) {
      r.kind = 3;
      r.first_bad = i;
      set_error("Matrices have different column IDs (entry %llu)",
                (unsigned long long)i);
      rc = OC_ERR_MISMATCH;
    }
  }
  for (uint64_t i = 0; rc == OC_OK && i < a->nnz; ++i) {
    double x = host_val(a, i), y = host_val(b, i);
    double den = std::max(std::fabs(x), std::fabs(y));
    double e = den > 0 ? std::fabs(x - y) / den : 0.0;
    if (std::isnan(x) != std::isnan(y)) e = INFINITY;
    if (e > r.max_rel_err) r.max_rel_err = e;
    if (e > rel_err) ++r.num_errors;
    if (e > autofail) {
      r.kind = 4;
      r.first_bad = i;
      set_error("Matrices have different data (entry %llu, rel err %g)",
                (unsigned long long)i, e);
      rc = OC_ERR_MISMATCH;
    }
  }
  if (res) *res = r;
  return rc;
}

extern "C" int oc_csr_transpose(const oc_csr *in, oc_csr *out) {
  OC_REQUIRE(in && out && !in->on_device, "oc_csr_transpose: need a host matrix");
  OC_TRY(oc_csr_alloc(out, in->cols, in->rows, in->nnz, in->vtype, 4, 0));
  uint32_t *ro = (uint32_t *)out->row_offsets;
  std::memset(ro, 0, (out->rows + 1) * 4);
  for (uint64_t i = 0; i < in->nnz; ++i) ro[in->col_ids[i] + 1]++;
  for (uint64_t i = 0; i < out->rows; ++i) ro[i + 1] += ro[i];
  std::vector<uint32_t> pos(ro, ro + out->rows);
  size_t vs = vsize(in->vtype);
  for (uint64_t r = 0; r < in->rows; ++r) {
    for (uint64_t j = host_off(in, r); j < host_off(in, r + 1); ++j) {
      uint32_t c = in->col_ids[j];
      uint32_t p = pos[c]++;
      out->col_ids[p] = (uint32_t)r;
      std::memcpy((char *)out->values + p * vs, (const char *)in->values + j * vs, vs);
    }
  }
  return OC_OK;
}
