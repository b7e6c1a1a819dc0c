// Internal helpers shared by the host and device translation units of
// libocean.so.  Not part of the public ABI (include/ocean_capi.h is).
#pragma once

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/ocean_capi.h"

namespace oc {

// Per-thread message of the last failing call (oc_last_error).
void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
const char *get_error();

// Status-returning error propagation for internal helpers.
#define OC_TRY(expr)                                                          \
  do {                                                                        \
    int oc_try_rc_ = (expr);                                                  \
    if (oc_try_rc_ != OC_OK) return oc_try_rc_;                               \
  } while (0)

#define OC_REQUIRE(cond, ...)                                                 \
  do {                                                                        \
    if (!(cond)) {                                                            \
      ::oc::set_error(__VA_ARGS__);                                           \
      return OC_ERR_INVALID;                                                  \
    }                                                                         \
  } while (0)

inline size_t vsize(int vtype) { return vtype == OC_F32 ? 4 : 8; }

// Device buffers of product outputs (C's arrays, oc_csr.owned == 3): freed
// buffers stay in a per-device cache and are reused best fit (2 MiB units;
// requests under 64 MiB take buffers of at most four times their size).
// Device memory handed back to the driver is cleared before it is handed
// out again, at tens of GB/s, so a series of differently sized products
// that freed and allocated its outputs spent up to 1.5 s per product in the
// allocation.  The cached buffers go back to the driver only when a device
// allocation of the library fails (out_release, then one retry), as the
// stream-ordered pool keeps freed memory with its release threshold at
// the maximum.  out_free waits for the device.
int out_alloc(void **p, size_t bytes);
void out_free(void *p);
void out_release();  // the current device's cached buffers back to the driver

// splitmix64: the seeded generator behind every synthetic matrix.
struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  // uniform double in [0, 1) with 53 random bits
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};

}  // namespace oc
