"""Shared checks of GPU results against the CPU oracle (test helpers)."""
import sys

import numpy as np

from oracle import oracle as ora


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def check_product(A, B, C, tol, name="", plain_exact=False):
    """C (host CSR from the GPU) against the FP64 oracle.

    Structure: row offsets and column ids bit-exact.  Values: the
    condition-normalised error |c - c_ref| / sum|a_ik b_kj| <= tol; the max
    plain relative error is returned as well.  plain_exact: values must equal
    the oracle's product-order sums bit for bit (the ESC path).
    """
    rp, cols, vals, cond = ora.numeric(A, B, compensated=0 if plain_exact else 1)
    assert C.rows == A.rows and C.cols == B.cols, name
    assert C.nnz == len(cols), f"{name}: nnz {C.nnz} != oracle {len(cols)}"
    assert np.array_equal(C.row_offsets.astype(np.uint64), rp), f"{name}: row offsets differ"
    assert np.array_equal(C.col_ids, cols), f"{name}: column ids differ"
    got = C.values.astype(np.float64)
    err = np.abs(got - vals)
    if plain_exact:
        assert np.array_equal(got, vals), f"{name}: ESC values not bit-exact"
    norm = np.where(cond > 0, err / np.where(cond > 0, cond, 1.0), err)
    worst = float(norm.max()) if len(norm) else 0.0
    assert worst <= tol, f"{name}: normalised error {worst:.3e} > {tol:.1e}"
    nz = np.abs(vals) > 0
    plain = float((err[nz] / np.abs(vals[nz])).max()) if nz.any() else 0.0
    return worst, plain


def csr_from_entries(gpu, rows, cols, entries, vtype=0):
    """Host CSR from (row, col, value) triples, sorted by row, then column."""
    entries = sorted(entries)
    ro = np.zeros(rows + 1, np.uint32)
    for r, _, _ in entries:
        ro[r + 1] += 1
    ro = np.cumsum(ro, dtype=np.uint64).astype(np.uint32)
    ci = np.array([c for _, c, _ in entries], np.uint32)
    va = np.array([v for _, _, v in entries], np.float32 if vtype else np.float64)
    return gpu.CSR(rows, cols, ro, ci, va)
