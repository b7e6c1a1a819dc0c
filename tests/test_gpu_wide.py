"""Symbolic rows beyond the LDS bitmap, the estimated-compaction symbolic
path, and the workflow decisions, against the oracle.

Rows whose column span exceeds the widest dense symbolic bitmap
(tables()["dense_sym"][6] columns) and whose products exceed the hash bins take the
windowed symbolic kernel (the role of hashSymbolicGMEMKernel,
kernels/AccumulatorHash.cuh:297-433, launched at kernels/Wrappers.cuh:75-99);
so do the rows whose shrunken hash table overflows under the estimated
compaction (kernels/SpGEMM.cuh:759-923).  The decisions of
kernels/SpGEMM.cuh:455-540 and 587-686 (decideMatrixType,
kernels/Analysis.cuh:10-29) are compared with the oracle's on the GPU's own
sampled rows and estimates."""
import ctypes as C

import numpy as np
import pytest

from helpers import check_product, log
from oracle import oracle as ora

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


def wide_product(gpu, seed=21):
    """A (1500 x 8192) * B (8192 x 2^21), f64.  B: 512 long rows of ~2000
    columns spread over all 2^21 columns, one row of 60,000, and short rows
    of 6.  A: 500 rows of 16 entries on long B rows (32k products, spans of
    ~2M columns), 24 rows of 4,500-6,000 entries on short B rows (more A
    entries than the kernel keeps cursors for in LDS), 2 rows on the 60k
    row, and short rows of 4 entries anywhere."""
    rng = np.random.default_rng(seed)
    ncols, nb = 1 << 21, 8192
    blen = np.full(nb, 6, np.int64)
    long_rows = rng.choice(nb, 512, replace=False)
    blen[long_rows] = rng.integers(1500, 2500, 512)
    hub = int(np.setdiff1d(np.arange(nb), long_rows)[0])
    blen[hub] = 60000
    bcols = [np.sort(rng.choice(ncols, int(L), replace=False)).astype(np.uint32) for L in blen]
    brp = np.zeros(nb + 1, np.uint32)
    np.cumsum(blen, out=brp[1:])
    B = gpu.CSR(nb, ncols, brp, np.concatenate(bcols), rng.uniform(-1, 1, int(brp[-1])))
    short = np.setdiff1d(np.arange(nb), np.append(long_rows, hub))
    arows = []
    for r in range(1500):
        if r < 500:
            c = rng.choice(long_rows, 16, replace=False)
        elif r < 524:
            c = rng.choice(short, int(rng.integers(4500, 6000)), replace=False)
        elif r < 526:
            c = np.array([hub, long_rows[r % 512]])
        else:
            c = rng.choice(nb, 4, replace=False)
        arows.append(np.sort(c).astype(np.uint32))
    arp = np.zeros(1501, np.uint32)
    np.cumsum([len(c) for c in arows], out=arp[1:])
    A = gpu.CSR(1500, nb, arp, np.concatenate(arows), rng.uniform(-1, 1, int(arp[-1])))
    return A, B


def est_sym_matrix(gpu, n=1 << 17, bw=11, nspread=24, per=300, spacing=400, seed=8):
    """Banded n x n (bandwidth 11: 121 products and 21 outputs per row, so a
    sampled compaction near 5.8 with a small variance) with 24 rows replaced
    by 300 entries 400 columns apart (no two referenced B rows overlap, so
    those rows have compaction 1 and outgrow the tables the estimated
    compaction gives them)."""
    A = gpu.CSR.banded(n, bw, seed=seed)
    rng = np.random.default_rng(seed)
    ro = A.row_offsets.astype(np.int64)
    lens = np.diff(ro)
    special = np.sort(rng.choice(np.arange(1000, n - 1000), nspread, replace=False))
    spread = {}
    for r in special:
        start = int(rng.integers(0, n - per * spacing))
        spread[int(r)] = np.arange(start, start + per * spacing, spacing, dtype=np.uint32)
        lens[r] = per
    rp = np.zeros(n + 1, np.uint32)
    np.cumsum(lens, out=rp[1:])
    ci = np.empty(int(rp[-1]), np.uint32)
    va = np.empty(int(rp[-1]), np.float64)
    for r in range(n):
        s, e = int(rp[r]), int(rp[r + 1])
        if r in spread:
            ci[s:e] = spread[r]
            va[s:e] = rng.uniform(-1, 1, e - s)
        else:
            ci[s:e] = A.col_ids[ro[r]:ro[r + 1]]
            va[s:e] = A.values[ro[r]:ro[r + 1]]
    return gpu.CSR(n, n, rp, ci, va)


def _window_rows(gpu, A, B, use_comp=0, comp=0.0):
    ref = ora.analysis(A, B)
    t = gpu.tables(A.vtype)
    bins = ora.bin_symbolic(ref["products"], ref["left"], ref["right"], t["hash_sym"],
                            t["dense_sym"], t["max_dense_sym_bin"], use_comp, 1.25, comp)
    return np.nonzero(bins == 6)[0], ref


def test_symbolic_window_wide_spans(gpu):
    """oc_symbolic on rows spanning ~2M columns: the windowed bitmap
    (several windows per row, cursors in LDS and in global scratch) gives
    the oracle's exact counts bit for bit."""
    A, B = wide_product(gpu)
    rows, ref = _window_rows(gpu, A, B)
    nA = np.diff(A.row_offsets.astype(np.int64))
    span = ref["right"].astype(np.int64) - ref["left"].astype(np.int64) + 1
    assert len(rows) >= 500 and (nA[rows] > 4096).sum() >= 20, "window rows not reached"
    assert span[rows].max() > gpu.tables()["dense_sym"][6]
    log(f"case: wide symbolic, {len(rows)} window rows, max span {span[rows].max()}, "
        f"max A entries {nA[rows].max()}")
    dA, dB = A.upload(), B.upload()
    nz = gpu.DeviceBuffer(4 * A.rows)
    gpu._check(gpu.lib.oc_symbolic(None, C.byref(dA.m), C.byref(dB.m), nz.ptr, None))
    want, _ = ora.symbolic(A, B)
    got = nz.to_array(np.uint32, A.rows)
    assert np.array_equal(got, want), f"{(got != want).sum()} rows differ"


def test_wide_product_precise(gpu):
    """The whole product on the same operands, precise workflow: window
    rows in the symbolic pass, long rows through the rank-tile kernel with
    B wider than the compact tile-start table covers (2^21 columns)."""
    log("case: wide product, precise workflow")
    A, B = wide_product(gpu)
    dA, dB = A.upload(), B.upload()
    sp = gpu.SpGEMM(gpu.Config(track_stage_times=1))
    sp.set_workflow(0)
    dC = sp.run(dA, dB)
    assert dC.validate() == 0
    st = sp.get_stats()
    assert st["decision"]["workflow"] == 0
    assert st["binning_1"][7] - st["binning_1"][6] >= 500
    C_ = dC.download()
    worst, plain = check_product(A, B, C_, TOL64, "wide")
    log(f"wide: nnz {C_.nnz}, window rows {st['binning_1'][7] - st['binning_1'][6]}, "
        f"big rows {st['binning_2'][14] - st['binning_2'][13]}, err {worst:.2e}")


def test_estimated_symbolic_overflow(gpu):
    """Automatic workflow on a matrix whose sampled compaction turns on the
    estimated symbolic tables (S - S/8 slots per bin): the rows that outgrow
    them are recounted by the windowed kernel and the product matches the
    oracle."""
    M = est_sym_matrix(gpu)
    ref = ora.analysis(M, M)
    cfg = ora.decide_cfg()
    tp = int(ref["products"].astype(np.int64).sum())
    d = ora.decide_analysis(cfg, M.rows, M.cols, M.nnz, M.rows, M.cols, tp,
                            int(ref["products"].max()))
    hb = ora.hll_construct(d.hll_log2_precision, M)
    sr, se = ora.hll_sample(d.hll_log2_precision, M, hb, ora.num_samples(cfg, M.rows))
    d = ora.decide_sampling(cfg, M.nnz, tp, ref["products"][sr], se, d)
    assert d.ana2_type == 0 and d.use_estimated_symbolic == 1
    t = gpu.tables()
    bins = ora.bin_symbolic(ref["products"], ref["left"], ref["right"], t["hash_sym"],
                            t["dense_sym"], t["max_dense_sym_bin"], 1, 1.25, d.safe_compaction)
    nnz, _ = ora.symbolic(M, M)
    lim = np.array([s - s // 8 for s in t["hash_sym"]])
    expect_ovf = int(sum(1 for r in np.nonzero(bins <= 5)[0] if nnz[r] > lim[bins[r]]))
    assert expect_ovf > 0
    log(f"case: estimated symbolic, safe compaction {d.safe_compaction:.3f}, "
        f"{expect_ovf} rows overflow")
    dM = M.upload()
    sp = gpu.SpGEMM(gpu.Config(track_stage_times=1))
    dC = sp.run(dM, dM)
    st = sp.get_stats()
    assert st["decision"]["workflow"] == 0 and st["decision"]["estimated_symbolic"] is True
    assert st["decision"]["symbolic_overflow_rows"] == expect_ovf
    C_ = dC.download()
    worst, _ = check_product(M, M, C_, TOL64, "est-symbolic")
    log(f"est-symbolic: nnz {C_.nnz}, overflow rows {expect_ovf}, err {worst:.2e}")


def _decision_mats(gpu):
    return {
        "laplacian": lambda: gpu.CSR.laplacian2d(96, 80, seed=0),
        "rmat13": lambda: gpu.CSR.rmat(13, 16 << 13, seed=1),
        "rmat20": lambda: gpu.CSR.rmat(20, 16 << 20, seed=1),
        "banded": lambda: gpu.CSR.banded(60000, 65, seed=6),
        "rect": lambda: gpu.CSR.rect_band(1 << 14, 1 << 17, 8.0, 0.3, 4096, 8, vtype=1, seed=2),
        "est_sym": lambda: est_sym_matrix(gpu),
    }


@pytest.mark.parametrize("name", ["laplacian", "rmat13", "rmat20", "banded", "rect", "est_sym"])
def test_decisions_match_oracle(gpu, name):
    """The library's decisions (stats `decision` and `sampling` blocks)
    equal ora_decide_analysis / ora_decide_sampling fed with the GPU's own
    analysis and sampled (row, estimate) pairs."""
    log(f"case: decisions {name}")
    A = _decision_mats(gpu)[name]()
    B = A.transpose() if name == "rect" else A
    dA, dB = A.upload(), B.upload()
    lib = gpu.lib
    m = A.rows
    outs = [gpu.DeviceBuffer(4 * m) for _ in range(5)]
    psum, sc = gpu.DeviceBuffer(8 * (m + 1)), gpu.DeviceBuffer(8)
    gpu._check(lib.oc_analysis(C.byref(dA.m), C.byref(dB.m), *[o.ptr for o in outs], psum.ptr,
                               sc.ptr, sc.ptr + 4, None))
    prod = outs[0].to_array(np.uint32, m)
    tp = int(psum.to_array(np.uint64, m + 1)[-1])
    cfg = ora.decide_cfg()
    d = ora.decide_analysis(cfg, A.rows, A.cols, A.nnz, B.rows, B.cols, tp, int(prod.max()))
    sp = gpu.SpGEMM(gpu.Config())
    dC = sp.run(dA, dB)
    st = sp.get_stats()
    dC.free()
    dec = st["decision"]
    if d.ana1_type == 0:
        assert dec["workflow"] == 2
        assert st["ultra_sparse"]["kernel_size"] == d.spark_kernel_size
        if d.spark_need_outlier:
            assert st["ultra_sparse"]["outlier_size"] == d.spark_outlier_kernel_size
        log(f"{name}: ultra-sparse, kernel {d.spark_kernel_size}")
        return
    assert d.ana1_type == 1
    lp = d.hll_log2_precision
    ns = ora.num_samples(cfg, m)
    hb = gpu.DeviceBuffer(B.rows << lp)
    gpu._check(lib.oc_hll_build(lp, C.byref(dB.m), hb.ptr, None))
    rows, se = gpu.DeviceBuffer(4 * ns), gpu.DeviceBuffer(4 * ns)
    gpu._check(lib.oc_hll_sample(lp, C.byref(dA.m), hb.ptr, ns, rows.ptr, se.ptr, None))
    srows = rows.to_array(np.uint32, ns)
    d = ora.decide_sampling(cfg, A.nnz, tp, prod[srows], se.to_array(np.int32, ns), d)
    smp = st["sampling"]
    assert smp["log2_precision"] == lp and smp["num_samples"] == ns

    def same(x, y):
        return (x is None and np.isnan(y)) or x == y

    assert same(smp["compaction_1"], d.compaction_1)
    assert same(smp["compaction_2"], d.compaction_2)
    assert dec["workflow"] == (1 if d.ana2_type == 1 else 0)
    assert dec["dense_bitmap_query"] == bool(d.dense_query_bitmap)
    if d.ana2_type == 0:
        assert same(smp["safe_estimated_compaction"], d.safe_compaction)
        assert dec["estimated_symbolic"] == bool(d.use_estimated_symbolic)
    log(f"{name}: workflow {dec['workflow']}, p {lp}, compaction {d.compaction_1:.3f} / "
        f"{d.compaction_2:.3f}, safe {d.safe_compaction:.3f}, "
        f"estimated symbolic {d.use_estimated_symbolic}")
