"""Writes rmat20_structure.json: the oracle's structure of C = A*A for the
R-MAT n=2^20 benchmark matrix (16 nnz/row, (.57,.19,.19), seed 1): total
nnz, SHA-256 of the per-row nnz array (u32 little-endian) and the
position-weighted checksum sum_r h_r*(2r+1) mod 2^64 of the per-row column
set hashes h_r = sum_j mix64(col_j).  Run from the repository root:
    python tests/golden/make_rmat20_structure.py
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as ora  # noqa: E402
import research_amd as ra  # noqa: E402


def position_checksum(h: np.ndarray) -> int:
    w = (2 * np.arange(len(h), dtype=np.uint64) + 1)
    with np.errstate(over="ignore"):
        return int(np.sum(h.astype(np.uint64) * w, dtype=np.uint64))


def main():
    A = ra.CSR.rmat(20, 16 << 20, 0.57, 0.19, 0.19, vtype=0, seed=1)
    t0 = time.time()
    nnz, h = ora.symbolic(A, A)
    out = {
        "matrix": "rmat log_n=20 edges=16<<20 a=.57 b=.19 c=.19 f64 seed=1, C=A*A",
        "a_nnz": A.nnz,
        "c_nnz": int(nnz.astype(np.uint64).sum()),
        "row_nnz_sha256": hashlib.sha256(nnz.astype("<u4").tobytes()).hexdigest(),
        "row_hash_checksum": position_checksum(h),
        "oracle_seconds": round(time.time() - t0, 1),
    }
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "rmat20_structure.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(out)


if __name__ == "__main__":
    main()
