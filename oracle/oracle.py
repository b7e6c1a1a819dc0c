"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker and CPU baseline, never by the
product.  Each wrapped function restates the reference algorithm at the
file:line cited in oracle.h.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(_HERE, "liboracle.so")


def _load():
    if not os.path.exists(_PATH):
        raise RuntimeError(f"{_PATH} not built (make -C oracle)")
    L = C.CDLL(_PATH)
    P, U64, U32, I, D = C.c_void_p, C.c_uint64, C.c_uint32, C.c_int, C.c_double
    sig = {
        "ora_murmur3_x86_32": (U32, [P, C.c_size_t, U32]),
        "ora_murmur3_col": (U32, [U32]),
        "ora_hll_lc_table": (None, [I, P]),
        "ora_hll_construct": (None, [I, U64, P, P, P]),
        "ora_hll_estimate": (C.c_int32, [I, P, P]),
        "ora_hll_merge": (None, [I, U64, P, P, P, D, P, P]),
        "ora_sample_row": (U32, [U32, U64]),
        "ora_hll_sample": (None, [I, U64, P, P, P, U32, P, P]),
        "ora_analysis": (None, [U64, P, P, U64, P, P, P, P, P, P, P, P]),
        "ora_bin_symbolic": (None, [U64, P, P, P, P, P, I, I, D, D, P]),
        "ora_bin_numeric": (None, [U64, P, P, P, P, P, I, D, I, I, U32, P]),
        "ora_decide_cfg_default": (None, [P]),
        "ora_decide_analysis": (None, [P, U64, U64, U64, U64, U64, U64, U32, P]),
        "ora_num_samples": (I, [P, U64]),
        "ora_decide_sampling": (None, [P, U64, U64, I, P, P, P]),
        "ora_mix64": (U64, [U64]),
        "ora_symbolic": (None, [U64, U64, P, P, U64, P, P, P, P, I]),
        "ora_numeric": (None, [U64, U64, I, P, P, P, U64, P, P, P, P, P, P, P, I, I]),
        "ora_numeric_rows": (None, [U64, P, I, P, P, P, U64, P, P, P, P, P, P, P, I]),
        "ora_symbolic_rows": (U64, [U64, P, P, P, U64, P, P, P, I]),
        "ora_baseline_spgemm": (U64, [U64, U64, I, P, P, P, U64, P, P, P, P, I]),
    }
    for n, (r, a) in sig.items():
        f = getattr(L, n)
        f.restype, f.argtypes = r, a
    return L


lib = _load()


def _p(a: np.ndarray):
    return a.ctypes.data


class DecideCfg(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "use_estimation_workflow", "use_ultrasparse_workflow", "use_fat_workflow",
        "use_assist_information", "usparse_avg_product_threshold",
        "estimation_input_expansion_threshold", "estimation_output_compaction_threshold",
        "dense_query_bitmap_compaction_threshold", "fat_ratio",
        "hll_higher_precision_expansion_threshold", "hll_precision_low", "hll_precision_high",
        "hll_sample_min_rows", "hll_sample_max_rows")] + [
        (n, C.c_double) for n in ("hll_sample_rate", "hll_expansion_low", "hll_expansion_high",
                                  "estimate_compaction_target_z_value")] + [
        ("fat_dense_limit", C.c_uint32)]


class Decision(C.Structure):
    _fields_ = [
        ("ana1_type", C.c_int), ("spark_kernel_size", C.c_int), ("spark_need_outlier", C.c_int),
        ("spark_outlier_kernel_size", C.c_int), ("hll_log2_precision", C.c_int),
        ("hll_expansion_factor", C.c_double), ("num_sample_rows", C.c_int),
        ("compaction_1", C.c_double), ("compaction_2", C.c_double), ("ana2_type", C.c_int),
        ("dense_query_bitmap", C.c_int), ("compaction_avg", C.c_double),
        ("compaction_var", C.c_double), ("safe_compaction", C.c_double),
        ("use_estimated_symbolic", C.c_int),
    ]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _vals(a):
    a = np.ascontiguousarray(a)
    return a, (1 if a.dtype == np.float32 else 0)


# ------------------------------------------------------------------ murmur
def murmur3(data: bytes, seed: int) -> int:
    buf = (C.c_uint8 * max(len(data), 1)).from_buffer_copy(data or b"\0")
    return lib.ora_murmur3_x86_32(buf, len(data), seed)


def murmur3_col(col: int) -> int:
    return lib.ora_murmur3_col(col)


# --------------------------------------------------------------------- HLL
def hll_lc_table(log_p: int) -> np.ndarray:
    out = np.zeros((1 << log_p) + 1, dtype=np.int32)
    lib.ora_hll_lc_table(log_p, _p(out))
    return out


def hll_construct(log_p: int, B) -> np.ndarray:
    rp, ci = _u32(B.row_offsets), _u32(B.col_ids)
    out = np.zeros(B.rows * (1 << log_p), dtype=np.uint8)
    lib.ora_hll_construct(log_p, B.rows, _p(rp), _p(ci), _p(out))
    return out


def hll_estimate(log_p: int, regs: np.ndarray) -> int:
    regs = np.ascontiguousarray(regs, dtype=np.uint8)
    lc = hll_lc_table(log_p)
    return lib.ora_hll_estimate(log_p, _p(regs), _p(lc))


def hll_merge(log_p: int, A, hll_b: np.ndarray, expansion: float, num_sizes) -> np.ndarray:
    rp, ci = _u32(A.row_offsets), _u32(A.col_ids)
    ns = _u32(num_sizes)
    hb = np.ascontiguousarray(hll_b, dtype=np.uint8)
    out = np.zeros(A.rows, dtype=np.uint32)
    lib.ora_hll_merge(log_p, A.rows, _p(rp), _p(ci), _p(hb), expansion, _p(ns), _p(out))
    return out


def hll_sample(log_p: int, A, hll_b: np.ndarray, nsamples: int):
    rp, ci = _u32(A.row_offsets), _u32(A.col_ids)
    hb = np.ascontiguousarray(hll_b, dtype=np.uint8)
    rows = np.zeros(nsamples, dtype=np.uint32)
    est = np.zeros(nsamples, dtype=np.int32)
    lib.ora_hll_sample(log_p, A.rows, _p(rp), _p(ci), _p(hb), nsamples, _p(rows), _p(est))
    return rows, est


def sample_row(i: int, rows: int) -> int:
    return lib.ora_sample_row(i, rows)


# ---------------------------------------------------------------- analysis
def analysis(A, B) -> dict:
    arp, aci, brp, bci = _u32(A.row_offsets), _u32(A.col_ids), _u32(B.row_offsets), _u32(B.col_ids)
    m = A.rows
    out = {k: np.zeros(m, dtype=np.uint32) for k in ("products", "max_b", "avg_b", "left", "right")}
    mx = np.zeros(1, dtype=np.uint32)
    lib.ora_analysis(m, _p(arp), _p(aci), B.rows, _p(brp), _p(bci), _p(out["products"]),
                     _p(out["max_b"]), _p(out["avg_b"]), _p(out["left"]), _p(out["right"]), _p(mx))
    out["max_a_nnz"] = int(mx[0])
    return out


def bin_symbolic(products, left, right, hash_sizes, dense_sizes, max_dense_bin,
                 use_compaction=0, expand=1.25, avg_compaction=0.0) -> np.ndarray:
    p, l, r = _u32(products), _u32(left), _u32(right)
    hs, ds = _u32(hash_sizes), _u32(dense_sizes)
    out = np.zeros(len(p), dtype=np.uint8)
    lib.ora_bin_symbolic(len(p), _p(p), _p(l), _p(r), _p(hs), _p(ds), max_dense_bin,
                         use_compaction, expand, avg_compaction, _p(out))
    return out


def bin_numeric(inp, left, right, hash_sizes, dense_sizes, expand=1, expand_factor=1.5,
                modify_dense_est=0, use_largest=0, hybrid_max=0):
    x = _u32(inp).copy()
    l, r = _u32(left), _u32(right)
    hs, ds = _u32(hash_sizes), _u32(dense_sizes)
    out = np.zeros(len(x), dtype=np.uint8)
    lib.ora_bin_numeric(len(x), _p(x), _p(l), _p(r), _p(hs), _p(ds), expand, expand_factor,
                        modify_dense_est, use_largest, hybrid_max, _p(out))
    return out, x


def decide_cfg(**kw) -> DecideCfg:
    c = DecideCfg()
    lib.ora_decide_cfg_default(C.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def decide_analysis(cfg, a_rows, a_cols, a_nnz, b_rows, b_cols, total_products, max_products):
    d = Decision()
    lib.ora_decide_analysis(C.byref(cfg), a_rows, a_cols, a_nnz, b_rows, b_cols,
                            total_products, max_products, C.byref(d))
    return d


def num_samples(cfg, a_rows) -> int:
    return lib.ora_num_samples(C.byref(cfg), a_rows)


def decide_sampling(cfg, a_nnz, total_products, sample_products, sample_est, d: Decision):
    sp, se = _u32(sample_products), np.ascontiguousarray(sample_est, dtype=np.int32)
    lib.ora_decide_sampling(C.byref(cfg), a_nnz, total_products, len(sp), _p(sp), _p(se), C.byref(d))
    return d


# -------------------------------------------------------------- Gustavson
def mix64(x: int) -> int:
    return lib.ora_mix64(x)


def symbolic(A, B, r0=0, r1=None, nthreads=0):
    r1 = A.rows if r1 is None else r1
    arp, aci, brp, bci = _u32(A.row_offsets), _u32(A.col_ids), _u32(B.row_offsets), _u32(B.col_ids)
    nnz = np.zeros(r1 - r0, dtype=np.uint32)
    h = np.zeros(r1 - r0, dtype=np.uint64)
    lib.ora_symbolic(r0, r1, _p(arp), _p(aci), B.cols, _p(brp), _p(bci), _p(nnz), _p(h), nthreads)
    return nnz, h


def numeric(A, B, r0=0, r1=None, compensated=1, with_cond=True, nthreads=0):
    """Rows [r0, r1) of A*B: (rowptr u64 relative, cols, vals f64, cond)."""
    r1 = A.rows if r1 is None else r1
    nnz, _ = symbolic(A, B, r0, r1, nthreads)
    rowptr = np.zeros(r1 - r0 + 1, dtype=np.uint64)
    np.cumsum(nnz, out=rowptr[1:])
    n = int(rowptr[-1])
    cols = np.zeros(max(n, 1), dtype=np.uint32)
    vals = np.zeros(max(n, 1), dtype=np.float64)
    cond = np.zeros(max(n, 1), dtype=np.float64) if with_cond else None
    arp, aci, brp, bci = _u32(A.row_offsets), _u32(A.col_ids), _u32(B.row_offsets), _u32(B.col_ids)
    av, vt = _vals(A.values)
    bv, _ = _vals(B.values)
    lib.ora_numeric(r0, r1, vt, _p(arp), _p(aci), _p(av), B.cols, _p(brp), _p(bci), _p(bv),
                    _p(rowptr), _p(cols), _p(vals), _p(cond) if cond is not None else None,
                    compensated, nthreads)
    return rowptr, cols[:n], vals[:n], (cond[:n] if cond is not None else None)


def baseline_spgemm(A, B, r0=0, r1=None, nthreads=0):
    """CPU baseline (not a checker): plain row-wise Gustavson over rows
    [r0, r1) of A*B.  Returns (nnz, sum of the values)."""
    r1 = A.rows if r1 is None else r1
    arp, aci, brp, bci = _u32(A.row_offsets), _u32(A.col_ids), _u32(B.row_offsets), _u32(B.col_ids)
    av, vt = _vals(A.values)
    bv, _ = _vals(B.values)
    cs = C.c_double(0.0)
    n = lib.ora_baseline_spgemm(r0, r1, vt, _p(arp), _p(aci), _p(av), B.cols, _p(brp), _p(bci),
                                _p(bv), C.byref(cs), nthreads)
    return int(n), cs.value


def numeric_rows(A, B, rows, nthreads=0):
    """Listed rows of A*B (compensated sums): (rowptr, cols, vals, cond)."""
    rows = _u32(rows)
    arp, aci, brp, bci = _u32(A.row_offsets), _u32(A.col_ids), _u32(B.row_offsets), _u32(B.col_ids)
    nnz = np.zeros(len(rows), dtype=np.uint32)
    lib.ora_symbolic_rows(len(rows), _p(rows), _p(arp), _p(aci), B.cols, _p(brp), _p(bci),
                          _p(nnz), nthreads)
    rowptr = np.zeros(len(rows) + 1, dtype=np.uint64)
    np.cumsum(nnz, out=rowptr[1:])
    n = int(rowptr[-1])
    cols = np.zeros(max(n, 1), dtype=np.uint32)
    vals = np.zeros(max(n, 1), dtype=np.float64)
    cond = np.zeros(max(n, 1), dtype=np.float64)
    av, vt = _vals(A.values)
    bv, _ = _vals(B.values)
    lib.ora_numeric_rows(len(rows), _p(rows), vt, _p(arp), _p(aci), _p(av), B.cols, _p(brp),
                         _p(bci), _p(bv), _p(rowptr), _p(cols), _p(vals), _p(cond), nthreads)
    return rowptr, cols[:n], vals[:n], cond[:n]
