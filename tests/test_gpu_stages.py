"""Stage-level GPU parity through the C-ABI stage calls: every stage of the
path against the oracle on the same seeded inputs, bit-exact (all of these
are integer/index work)."""
import ctypes as C

import numpy as np
import pytest

from helpers import log
from oracle import oracle as ora

pytestmark = pytest.mark.gpu


def _buf(gpu, a):
    return gpu.DeviceBuffer.from_array(np.ascontiguousarray(a))


def _out(gpu, nbytes):
    return gpu.DeviceBuffer(max(int(nbytes), 8))


def _chk(gpu, rc):
    gpu._check(rc)


def _hubs(ra):
    """Rows of every length class of the analysis: short rows, rows just
    below and above the long-row cut of each lane-group width, and one hub
    row of 40000 entries; a third of the rows (so of B's rows) are empty."""
    rng = np.random.default_rng(11)
    n = 50000
    lens = rng.integers(0, 24, n)
    lens[rng.random(n) < 0.33] = 0
    for i, L in enumerate([127, 128, 129, 511, 512, 513, 1024, 1025, 2049, 6000, 40000]):
        lens[100 + 97 * i] = L
    ro = np.zeros(n + 1, np.uint32)
    np.cumsum(lens, out=ro[1:])
    ci = np.concatenate([np.sort(rng.choice(n, L, replace=False)) for L in lens]).astype(np.uint32)
    return ra.CSR(rows=n, cols=n, row_offsets=ro, col_ids=ci,
                  values=rng.random(len(ci)))


MATS = {
    "laplacian": lambda ra: ra.CSR.laplacian2d(96, 80, seed=0),
    "hubs": _hubs,
    "rmat13": lambda ra: ra.CSR.rmat(13, 16 << 13, seed=1),
    "er": lambda ra: ra.CSR.erdos_renyi(5000, 30.0, seed=2),
    "rect": lambda ra: ra.CSR.rect_band(3000, 50000, 8.0, 0.3, 4096, 8, vtype=1, seed=3),
}


def _pair(gpu, name):
    A = MATS[name](gpu)
    B = A.transpose() if name == "rect" else A
    return A, B


@pytest.mark.parametrize("name", list(MATS))
def test_analysis(gpu, name):
    log(f"case: analysis {name}")
    A, B = _pair(gpu, name)
    dA, dB = A.upload(), B.upload()
    m = A.rows
    outs = [_out(gpu, 4 * m) for _ in range(5)]
    psum = _out(gpu, 8 * (m + 1))
    sc = _out(gpu, 8)
    lib = gpu.lib
    _chk(gpu, lib.oc_analysis(C.byref(dA.m), C.byref(dB.m), *[o.ptr for o in outs], psum.ptr,
                              sc.ptr, sc.ptr + 4, None))
    ref = ora.analysis(A, B)
    for o, k in zip(outs, ("products", "max_b", "avg_b", "left", "right")):
        assert np.array_equal(o.to_array(np.uint32, m), ref[k]), k
    ps = psum.to_array(np.uint64, m + 1)
    assert ps[0] == 0 and np.array_equal(np.diff(ps), ref["products"].astype(np.uint64))
    s = sc.to_array(np.uint32, 2)
    assert s[0] == ref["max_a_nnz"] and s[1] == ref["products"].max()


@pytest.mark.parametrize("n", [0, 1, 1023, 1024, 65537, 3_000_001])
def test_scan_u32(gpu, n):
    log(f"case: scan n={n}")
    x = np.random.default_rng(n).integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    dx, dy = _buf(gpu, x if n else np.zeros(1, np.uint32)), _out(gpu, 8 * (n + 1))
    _chk(gpu, gpu.lib.oc_scan_u32(dx.ptr, dy.ptr, n, None))
    want = np.zeros(n + 1, np.uint64)
    np.cumsum(x.astype(np.uint64), out=want[1:])
    assert np.array_equal(dy.to_array(np.uint64, n + 1), want)


def test_murmur3_cols(gpu):
    log("case: murmur3 cols")
    cols = np.concatenate([np.arange(4096, dtype=np.uint32),
                           np.random.default_rng(0).integers(0, 1 << 32, 20000, dtype=np.uint64)
                           .astype(np.uint32), np.array([0xFFFFFFFF], np.uint32)])
    dc, dh = _buf(gpu, cols), _out(gpu, 4 * len(cols))
    _chk(gpu, gpu.lib.oc_murmur3_cols(dc.ptr, dh.ptr, len(cols), None))
    want = np.array([ora.murmur3_col(int(c)) for c in cols], np.uint32)
    assert np.array_equal(dh.to_array(np.uint32, len(cols)), want)


@pytest.mark.parametrize("log_p", [5, 6])
@pytest.mark.parametrize("name", ["rmat13", "er", "rect", "hubs"])
def test_hll_build_estimate_sample(gpu, name, log_p):
    """Sketches, per-row estimates and samples bit-exact.  `hubs` has rows
    of 2049, 6000 and 40000 entries: rows above 2048 entries are merged by
    a whole workgroup each (the heavy-row kernel), in both modes."""
    log(f"case: hll {name} log_p={log_p}")
    A, B = _pair(gpu, name)
    if name == "hubs":
        assert np.diff(A.row_offsets.astype(np.int64)).max() > 2048
    dA, dB = A.upload(), B.upload()
    lib = gpu.lib
    P = 1 << log_p
    hb = _out(gpu, B.rows * P)
    _chk(gpu, lib.oc_hll_build(log_p, C.byref(dB.m), hb.ptr, None))
    got_b = hb.to_array(np.uint8, B.rows * P)
    ref_b = ora.hll_construct(log_p, B)
    assert np.array_equal(got_b, ref_b)
    sizes = gpu.tables()["hash_num"]
    for expansion in (1.0, 1.5, 2.0):
        est = _out(gpu, 4 * A.rows)
        _chk(gpu, lib.oc_hll_estimate(log_p, C.byref(dA.m), hb.ptr, expansion, est.ptr, None))
        want = ora.hll_merge(log_p, A, ref_b, expansion, sizes)
        assert np.array_equal(est.to_array(np.uint32, A.rows), want), expansion
    samples = (1, 648, 3000)
    if name == "hubs":
        # enough samples that the LCG reaches rows above 2048 entries
        heavy = np.nonzero(np.diff(A.row_offsets.astype(np.int64)) > 2048)[0]
        ns = 1
        while not np.isin([ora.sample_row(i, A.rows) for i in range(ns)], heavy).any():
            ns *= 2
        samples = samples + (ns,)
    for ns in samples:
        rows, se = _out(gpu, 4 * ns), _out(gpu, 4 * ns)
        _chk(gpu, lib.oc_hll_sample(log_p, C.byref(dA.m), hb.ptr, ns, rows.ptr, se.ptr, None))
        wr, we = ora.hll_sample(log_p, A, ref_b, ns)
        assert np.array_equal(rows.to_array(np.uint32, ns), wr)
        assert np.array_equal(se.to_array(np.int32, ns), we)


def _check_grouping(bins, bin_rows, offs):
    m = len(bins)
    assert offs[0] == 0 and offs[15] == m
    for b in range(15):
        want = np.nonzero(bins == b)[0].astype(np.uint32)
        assert np.array_equal(bin_rows[offs[b]:offs[b + 1]], want), b


@pytest.mark.parametrize("name", list(MATS))
@pytest.mark.parametrize("comp", [(0, 0.0), (1, 3.5)])
def test_bin_symbolic(gpu, name, comp):
    log(f"case: bin_symbolic {name} compaction={comp}")
    A, B = _pair(gpu, name)
    ref = ora.analysis(A, B)
    t = gpu.tables(A.vtype)
    m = A.rows
    dp, dl, dr = _buf(gpu, ref["products"]), _buf(gpu, ref["left"]), _buf(gpu, ref["right"])
    bor, brows, boff = _out(gpu, m), _out(gpu, 4 * m), _out(gpu, 64)
    _chk(gpu, gpu.lib.oc_bin_symbolic(A.vtype, m, dp.ptr, dl.ptr, dr.ptr, comp[0], 1.25,
                                            comp[1], bor.ptr, brows.ptr, boff.ptr, None))
    want = ora.bin_symbolic(ref["products"], ref["left"], ref["right"], t["hash_sym"],
                            t["dense_sym"], t["max_dense_sym_bin"], comp[0], 1.25, comp[1])
    bins = bor.to_array(np.uint8, m)
    assert np.array_equal(bins, want)
    _check_grouping(bins, brows.to_array(np.uint32, m), boff.to_array(np.uint32, 16))


@pytest.mark.parametrize("name", list(MATS))
@pytest.mark.parametrize("mode", ["precise", "est", "hybrid"])
def test_bin_numeric(gpu, name, mode):
    log(f"case: bin_numeric {name} {mode}")
    A, B = _pair(gpu, name)
    ref = ora.analysis(A, B)
    t = gpu.tables(A.vtype)
    m = A.rows
    if mode == "precise":
        inp, kw = ora.symbolic(A, B)[0], dict(expand=1, expand_factor=1.5)
    elif mode == "est":
        hb = ora.hll_construct(5, B)
        inp = ora.hll_merge(5, A, hb, 2.0, t["hash_num"])
        kw = dict(expand=0, expand_factor=0.0, modify_dense_est=1)
    else:
        inp = ora.symbolic(A, B)[0]
        kw = dict(expand=1, expand_factor=1.5, use_largest=1, hybrid_max=max(1, m // 50))
    din = _buf(gpu, inp.astype(np.uint32))
    dl, dr = _buf(gpu, ref["left"]), _buf(gpu, ref["right"])
    bor, brows, boff = _out(gpu, m), _out(gpu, 4 * m), _out(gpu, 64)
    _chk(gpu, gpu.lib.oc_bin_numeric(
        A.vtype, m, din.ptr, dl.ptr, dr.ptr, kw.get("expand", 1), kw.get("expand_factor", 1.5),
        kw.get("modify_dense_est", 0), kw.get("use_largest", 0), kw.get("hybrid_max", 0),
        bor.ptr, brows.ptr, boff.ptr, None))
    want, want_in = ora.bin_numeric(inp, ref["left"], ref["right"], t["hash_num"], t["dense_num"],
                                    **kw)
    bins = bor.to_array(np.uint8, m)
    assert np.array_equal(bins, want)
    assert np.array_equal(din.to_array(np.uint32, m), want_in)
    _check_grouping(bins, brows.to_array(np.uint32, m), boff.to_array(np.uint32, 16))


@pytest.mark.parametrize("name", list(MATS) + ["rmat15_32", "banded"])
def test_symbolic_nnz(gpu, name):
    log(f"case: symbolic {name}")
    if name == "rmat15_32":
        A = B = gpu.CSR.rmat(15, 32 << 15, seed=5)
    elif name == "banded":
        A = B = gpu.CSR.banded(20000, 301, seed=6)
    else:
        A, B = _pair(gpu, name)
    dA, dB = A.upload(), B.upload()
    nz = _out(gpu, 4 * A.rows)
    _chk(gpu, gpu.lib.oc_symbolic(None, C.byref(dA.m), C.byref(dB.m), nz.ptr, None))
    want, _ = ora.symbolic(A, B)
    assert np.array_equal(nz.to_array(np.uint32, A.rows), want)


@pytest.mark.parametrize("wide,vt", [(False, np.float64), (True, np.float64), (False, np.float32)])
def test_sort_rows(gpu, wide, vt):
    """LDS path (rows up to 6144 entries) and global path (longer rows),
    all rows and a listed subset; the sort is stable."""
    log(f"case: sort_rows wide={wide} {np.dtype(vt).name}")
    rng = np.random.default_rng(9)
    lens = [0, 1, 2, 3, 64, 1000, 6143, 6144, 6145, 9000, 30000, 5, 77]
    ncols = 1 << 20
    cols = [rng.integers(0, ncols if i % 3 else 4096, n, dtype=np.int64).astype(np.uint32)
            for i, n in enumerate(lens)]  # some rows with repeated keys
    rp = np.zeros(len(lens) + 1, np.uint64 if wide else np.uint32)
    rp[1:] = np.cumsum(lens)
    ci = np.concatenate(cols)
    va = rng.standard_normal(len(ci)).astype(vt)
    M = gpu.CSR(len(lens), ncols, rp, ci, va)
    for subset in (None, np.array([1, 3, 6, 8, 10], np.uint32)):
        dM = M.upload()
        if subset is None:
            _chk(gpu, gpu.lib.oc_sort_rows(C.byref(dM.m), None, 0, None))
        else:
            ds = _buf(gpu, subset)
            _chk(gpu, gpu.lib.oc_sort_rows(C.byref(dM.m), ds.ptr, len(subset), None))
        got = dM.download()
        for r in range(len(lens)):
            s, e = int(rp[r]), int(rp[r + 1])
            if subset is None or r in subset:
                o = np.argsort(ci[s:e], kind="stable")
            else:
                o = np.arange(e - s)
            assert np.array_equal(got.col_ids[s:e], ci[s:e][o]), r
            assert np.array_equal(got.values[s:e], va[s:e][o]), r
