// convert <input.mtx> <output.csr> [float|double] [--merge-duplicates]
//
// The .mtx -> .csr converter of the reference (utils/convert.cpp:312-369):
// the same arguments, messages and exit codes, over oc_mtx_load and
// oc_csr_store.  --merge-duplicates (no counterpart) sums entries with the
// same (row, column).
#include <cstdio>
#include <cstring>
#include <string>

#include "ocean_capi.h"

int main(int argc, char **argv) {
  std::string type = "double";
  int merge = 0;
  const char *in = nullptr, *out = nullptr;
  int npos = 0;
  for (int i = 1; i < argc; ++i) {
    if (std::strcmp(argv[i], "--merge-duplicates") == 0) {
      merge = 1;
      continue;
    }
    if (npos == 0) in = argv[i];
    else if (npos == 1) out = argv[i];
    else if (npos == 2) type = argv[i];
    ++npos;
  }
  if (npos < 2) {
    std::printf("Usage: %s <input.mtx> <output.csr> [float|double] [--merge-duplicates]\n",
                argv[0]);
    std::printf("  Default data type is double if not specified\n");
    return 1;
  }
  int vtype;
  if (type == "float") vtype = OC_F32;
  else if (type == "double") vtype = OC_F64;
  else {
    std::fprintf(stderr, "Invalid data type: %s\n", type.c_str());
    std::fprintf(stderr, "Supported types: float, double\n");
    return 1;
  }
  std::printf("Loading MTX file: %s\n", in);
  std::fflush(stdout);
  oc_csr m;
  if (oc_mtx_load(in, vtype, merge, &m) != OC_OK) {
    std::fprintf(stderr, "Error: %s\n", oc_last_error());
    return 1;
  }
  std::printf("  Matrix dimensions: %llu x %llu\n", (unsigned long long)m.rows,
              (unsigned long long)m.cols);
  std::printf("  Non-zeros: %llu\n", (unsigned long long)m.nnz);
  std::printf("Converting to CSR format...\n");
  std::printf("Writing CSR to file: %s\n", out);
  std::fflush(stdout);
  if (oc_csr_store(out, &m) != OC_OK) {
    std::fprintf(stderr, "Error: %s\n", oc_last_error());
    oc_csr_free(&m);
    return 1;
  }
  oc_csr_free(&m);
  std::printf("Conversion complete!\n");
  return 0;
}
