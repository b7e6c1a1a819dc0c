"""Full-size parity of the benchmark product (R-MAT n=2^20, C = A*A, 9.7e9
output nonzeros) through size-independent properties, plus the estimation
workflow's overflow path forced from the host."""
import ctypes as C
import hashlib
import json
import os

import numpy as np
import pytest

from helpers import check_product, log
from oracle import oracle as ora

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _rows_from_device(gpu, dC, off, rows):
    """(cols, vals) of the listed rows of a device result, row by row."""
    vdt = np.float32 if dC.m.vtype == gpu.OC_F32 else np.float64
    out = []
    for r in rows:
        s, e = int(off[r]), int(off[r + 1])
        c = np.empty(e - s, np.uint32)
        v = np.empty(e - s, vdt)
        if e > s:
            gpu._check(gpu.lib.oc_memcpy_dtoh(c.ctypes.data, dC.m.col_ids + 4 * s, c.nbytes))
            gpu._check(gpu.lib.oc_memcpy_dtoh(v.ctypes.data, dC.m.values + v.itemsize * s,
                                              v.nbytes))
        out.append((c, v))
    return out


@pytest.mark.timeout(300)
@pytest.mark.parametrize("config", ["default", "tuned", "one_gather"])
def test_rmat20_full_size(gpu, config):
    """config "tuned": the autotuner's configuration for this matrix, the
    one the benchmark runs with; "one_gather": the same with the long rows
    through the one-gather rank tiles (use_tiled_numeric 3)."""
    gold = json.load(open(os.path.join(HERE, "golden", "rmat20_structure.json")))
    log(f"case: rmat 2^20 deg 16 f64 seed 1, C=A*A (full benchmark size), {config} config")
    A = gpu.CSR.rmat(20, 16 << 20, 0.57, 0.19, 0.19, vtype=0, seed=1)
    assert A.nnz == gold["a_nnz"]
    dA = A.upload()
    knobs = {}
    if config in ("tuned", "one_gather"):
        with open(os.path.join(HERE, "..", "research_amd", "tuned", "rmat20.json")) as f:
            knobs = json.load(f)
    if config == "one_gather":
        knobs["use_tiled_numeric"] = 3
    sp = gpu.SpGEMM(gpu.Config(**knobs))
    dC = sp.run(dA, dA)
    log(f"rmat20: nnz(C)={dC.nnz}, workflow {sp.get_stats()['decision']['workflow']}")
    # structure: sorted, in range, per-row counts and column sets
    assert dC.validate() == 0
    assert dC.nnz == gold["c_nnz"]
    off = dC.row_offsets().astype(np.uint64)
    row_nnz = np.diff(off).astype("<u4")
    assert hashlib.sha256(row_nnz.tobytes()).hexdigest() == gold["row_nnz_sha256"]
    h = dC.row_set_hash()
    w = 2 * np.arange(A.rows, dtype=np.uint64) + 1
    with np.errstate(over="ignore"):
        assert int(np.sum(h * w, dtype=np.uint64)) == gold["row_hash_checksum"]
    # values: C x == A (A x) for a random x, normalised by |A| (|A| |x|)
    x = np.random.default_rng(7).uniform(-1, 1, A.cols)
    y = dC.spmv(x)
    S = A.to_scipy()
    z = S @ (S @ x)
    absS = abs(S)
    bound = absS @ (absS @ np.abs(x))
    err = np.abs(y - z) / np.where(bound > 0, bound, 1.0)
    log(f"rmat20: spmv identity max normalised error {err.max():.2e}")
    assert err.max() <= 1e-10
    # sampled rows (the heaviest and a random set) against the oracle
    rl = np.diff(A.row_offsets.astype(np.int64))
    ro = A.row_offsets.astype(np.int64)
    cs = np.concatenate(([0], np.cumsum(rl[A.col_ids.astype(np.int64)])))
    prod = cs[ro[1:]] - cs[ro[:-1]]
    heavy = np.argsort(prod)[-3:]
    rnd = np.random.default_rng(11).choice(A.rows, 1500, replace=False)
    rows = np.unique(np.concatenate([heavy, rnd])).astype(np.uint32)
    rp, cols, vals, cond = ora.numeric_rows(A, A, rows)
    got = _rows_from_device(gpu, dC, off, rows)
    worst = 0.0
    for i, r in enumerate(rows):
        s, e = int(rp[i]), int(rp[i + 1])
        c, v = got[i]
        assert np.array_equal(c, cols[s:e]), int(r)
        nerr = np.abs(v - vals[s:e]) / np.where(cond[s:e] > 0, cond[s:e], 1.0)
        worst = max(worst, float(nerr.max()) if e > s else 0.0)
    log(f"rmat20: {len(rows)} sampled rows, max normalised error {worst:.2e}")
    assert worst <= 1e-12
    dC.free()


@pytest.mark.parametrize("kind", ["exact", "double", "partial", "long_rows", "all_low"])
def test_estimation_overflow_forced(gpu, kind):
    """The estimation workflow with host-given per-row estimates: exact
    counts (no row overflows), generous estimates, generous estimates with
    every fifth multi-entry row given an estimate of 1, or three long rows
    (beyond the LDS kernels) given half their count, so those rows overflow
    into the spare region, or every multi-entry row given an estimate of 1,
    whose overflow rows need more room than the spare region sized from the
    estimates holds (they then get a buffer of their own).  The product must
    equal the oracle's in every
    case; the overflow count is the rows whose count exceeds the estimate
    after binning raised dense-bin-0 estimates to that bin's width."""
    if kind == "long_rows":
        log("case: rmat 2^15 deg 32 f64, workflow 1, long rows under-estimated")
        A = gpu.CSR.rmat(15, 32 << 15, seed=5)
    else:
        log(f"case: rmat 2^13 deg 16 f64, workflow 1, estimates {kind}")
        A = gpu.CSR.rmat(13, 16 << 13, seed=3)
    exact, _ = ora.symbolic(A, A)
    est = (2 * exact + 8).astype(np.uint32) if kind != "exact" else exact.astype(np.uint32)
    ref = ora.analysis(A, A)
    t = gpu.tables()
    if kind == "partial":
        cand = np.nonzero(exact > 1)[0]
        est[cand[::5]] = 1
    if kind == "all_low":
        est[exact > 1] = 1
    if kind == "long_rows":
        bins0, _ = ora.bin_numeric(est, ref["left"], ref["right"], t["hash_num"], t["dense_num"],
                                   1, 1.5, modify_dense_est=1)
        longr = np.nonzero(bins0 == 13)[0]
        assert len(longr) >= 3, "no long rows in this matrix"
        est[longr[:3]] = exact[longr[:3]] // 2
    bins, est_mod = ora.bin_numeric(est, ref["left"], ref["right"], t["hash_num"], t["dense_num"],
                                    1, 1.5, modify_dense_est=1)
    want_ovf = int(np.count_nonzero(exact > est_mod))
    dA = A.upload()
    sp = gpu.SpGEMM(gpu.Config(use_ultrasparse_workflow=0))
    sp.set_workflow(1)
    dest = gpu.DeviceBuffer.from_array(est)
    sp.set_estimates(dest)
    dC = sp.run(dA, dA)
    st = sp.get_stats()
    assert st["decision"]["workflow"] == 1
    assert dC.validate() == 0
    Cm = dC.download()
    check_product(A, A, Cm, 1e-12, f"est-{kind}")
    ovf = st["estimation"]["overflow_rows"]
    log(f"est {kind}: overflow rows {ovf} (expected {want_ovf})")
    assert ovf == want_ovf
    if kind == "partial":
        assert want_ovf > 100
    if kind == "all_low":
        assert want_ovf > A.rows // 4 and st["estimation"]["spare_separate"]
    if kind in ("exact", "double"):
        assert not st["estimation"]["spare_separate"]
    if kind == "long_rows":
        assert want_ovf == 3 and st["binning_2"][14] - st["binning_2"][13] >= 3
    sp.set_estimates(None)
