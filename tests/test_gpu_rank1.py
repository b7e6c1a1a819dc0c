"""GPU parity of the one-gather rank tiles (use_tiled_numeric = 3): the
kernel for rows beyond the LDS hash/dense kernels (the role of
denseNumericIterKernel, kernels/AccumulatorDense.cuh:610-806).

In the precise workflow the dense symbolic pass stores the long rows'
bitmaps and the numeric pass gathers every product once; rows without a
stored bitmap (hash-symbolic rows, the estimation workflow) mark their
tiles first.  Each case checks the product against the FP64 oracle
(structure bit-exact, values within 1e-12 / 1e-5 condition-normalised)."""
import numpy as np
import pytest

from helpers import check_product, log

pytestmark = pytest.mark.gpu

TOL64, TOL32 = 1e-12, 1e-5


def _run(gpu, A, B, cfg, workflow=None):
    dA, dB = A.upload(), B.upload()
    sp = gpu.SpGEMM(cfg)
    if workflow is not None:
        sp.set_workflow(workflow)
    dC = sp.run(dA, dB)
    assert dC.validate() == 0
    C = dC.download()
    dC.free()
    return C, sp.get_stats()


def hubs(gpu, seed=3, vtype=0):
    """2^15 x 2^15 B with 50 long rows (2,000-20,000 entries) among short
    ones, and an A of rows with 3000, 1025, 1024, 2049, 64, 5, 1 and 600
    entries: rows of more than 1024 entries are staged in chunks, and the
    long B rows give tiles with more outputs than LDS slots."""
    rng = np.random.default_rng(seed)
    n = 1 << 15
    blen = rng.integers(0, 40, n)
    blen[:50] = rng.integers(2000, 20000, 50)
    bcols = [np.sort(rng.choice(n, int(L), replace=False)) for L in blen]
    brp = np.zeros(n + 1, np.uint32)
    np.cumsum(blen, out=brp[1:])
    dt = np.float32 if vtype else np.float64
    B = gpu.CSR(n, n, brp, np.concatenate(bcols).astype(np.uint32),
                rng.uniform(-1, 1, int(brp[-1])).astype(dt))
    alen = np.array([3000, 1025, 1024, 2049, 64, 5, 1, 600])
    acols = [np.sort(rng.choice(n, int(L), replace=False)) for L in alen]
    arp = np.zeros(len(alen) + 1, np.uint32)
    np.cumsum(alen, out=arp[1:])
    A = gpu.CSR(len(alen), n, arp, np.concatenate(acols).astype(np.uint32),
                rng.uniform(-1, 1, int(arp[-1])).astype(dt))
    return A, B


def short_segments(gpu, seed=5):
    """B of 2^14 rows over 2^20 columns: rows of 1-3 entries and eight rows
    of 30,000-60,000; A rows of up to 1,500 entries on the short rows (many
    segments per 64 positions) plus two hub rows (several widest tiles)."""
    rng = np.random.default_rng(seed)
    n, nb = 1 << 20, 1 << 14
    blen = rng.integers(1, 4, nb)
    blen[:8] = rng.integers(30000, 60000, 8)
    bcols = [np.sort(rng.choice(n, int(L), replace=False)) for L in blen]
    brp = np.zeros(nb + 1, np.uint32)
    np.cumsum(blen, out=brp[1:])
    B = gpu.CSR(nb, n, brp, np.concatenate(bcols).astype(np.uint32),
                rng.uniform(-1, 1, int(brp[-1])))
    alen = np.array([1000, 1024, 1500, 300, 8, 2])
    acols = []
    for i, L in enumerate(alen):
        c = rng.choice(np.arange(8, nb), int(L), replace=False)
        if i >= 3:
            c[:2] = [0, 1]
        acols.append(np.sort(c))
    arp = np.zeros(len(alen) + 1, np.uint32)
    np.cumsum(alen, out=arp[1:])
    A = gpu.CSR(len(alen), nb, arp, np.concatenate(acols).astype(np.uint32),
                rng.uniform(-1, 1, int(arp[-1])))
    return A, B


@pytest.mark.parametrize("block", [1024, 512])
@pytest.mark.parametrize("workflow", [0, 1])
@pytest.mark.parametrize("log_t", [14, 11])
def test_rank1_rmat15(gpu, workflow, log_t, block):
    """R-MAT 2^15, degree 32: precise (stored bitmaps) and estimation
    (tiles marked first, capacities checked) workflows, 1024-thread
    workgroups (one per CU) and 512-thread ones (two per CU)."""
    log(f"case: rank1 rmat 2^15 deg 32, workflow {workflow}, tiles 2^{log_t}, block {block}")
    A = gpu.CSR.rmat(15, 32 << 15, seed=5)
    cfg = gpu.Config(use_tiled_numeric=3, tile_log2_cols=log_t, tile_block=block,
                     track_stage_times=1)
    C, st = _run(gpu, A, A, cfg, workflow)
    assert st["binning_2"][14] - st["binning_2"][13] > 0, "no rows reached the big-row bin"
    if workflow == 0:
        assert st["decision"]["stored_row_bitmap_bytes"] > 0
    worst, _ = check_product(A, A, C, TOL64, f"rank1 rmat15 w{workflow}")
    log(f"rank1 rmat15 workflow {workflow} 2^{log_t}: big rows "
        f"{st['binning_2'][14] - st['binning_2'][13]}, err {worst:.2e}")


@pytest.mark.parametrize("workflow", [0, 1])
@pytest.mark.parametrize("vtype", [0, 1])
def test_rank1_hubs(gpu, workflow, vtype):
    """Rows of more than 1024 (or 512) A entries (chunked) and tiles with
    more outputs than slots (rank ranges), f64 and f32, both workgroup
    sizes."""
    log(f"case: rank1 hubs, workflow {workflow}, vtype {vtype}")
    A, B = hubs(gpu, vtype=vtype)
    for log_t, block in ((10, 1024), (14, 1024), (12, 512)):
        cfg = gpu.Config(use_tiled_numeric=3, tile_log2_cols=log_t, tile_block=block)
        C, st = _run(gpu, A, B, cfg, workflow)
        worst, _ = check_product(A, B, C, TOL64 if vtype == 0 else TOL32, f"hubs 2^{log_t}")
        log(f"rank1 hubs 2^{log_t}: nnz {C.nnz}, err {worst:.2e}")


@pytest.mark.parametrize("workflow", [0, 1])
def test_rank1_short_segments(gpu, workflow):
    """Many one-to-three-entry segments per 64 positions (the owner window
    reloads) and hub rows spanning 2^20 columns (widest tiles)."""
    log(f"case: rank1 short segments, workflow {workflow}")
    A, B = short_segments(gpu)
    for log_t in (11, 14):
        cfg = gpu.Config(use_tiled_numeric=3, tile_log2_cols=log_t)
        C, st = _run(gpu, A, B, cfg, workflow)
        worst, _ = check_product(A, B, C, TOL64, f"short segments 2^{log_t}")
        log(f"rank1 short segments 2^{log_t}: nnz {C.nnz}, err {worst:.2e}")


def test_rank1_plan_replay(gpu):
    """Value-only replays reuse the stored bitmaps of the planned run."""
    log("case: rank1 plan/replay on rmat 2^15 deg 32")
    A = gpu.CSR.rmat(15, 32 << 15, seed=5)
    dA = A.upload()
    sp = gpu.SpGEMM(gpu.Config(use_tiled_numeric=3))
    dC = sp.plan(dA, dA)
    assert sp.get_stats()["decision"]["stored_row_bitmap_bytes"] > 0
    check_product(A, A, dC.download(), TOL64, "rank1 plan")
    rng = np.random.default_rng(2)
    for it in range(2):
        vals = np.ascontiguousarray(rng.uniform(-1, 1, A.nnz))
        gpu._check(gpu.lib.oc_memcpy_htod(dA.m.values, vals.ctypes.data, vals.nbytes))
        A.values = vals
        sp.replay()
        worst, _ = check_product(A, A, dC.download(), TOL64, f"rank1 replay {it}")
        log(f"rank1 replay {it}: err {worst:.2e}")


def heavy_rows(gpu, log_cols, seed=9):
    """B of 2^15 rows over 2^log_cols columns, 64-255 entries each; A rows
    of 32,000, 20,000, 700 and 3 entries: the first two have about 5M and
    3M products (more than the 2^21 of one piece) and spread over all
    columns."""
    rng = np.random.default_rng(seed)
    nb, n = 1 << 15, 1 << log_cols
    blen = rng.integers(64, 256, nb)
    bcols = [np.sort(rng.choice(n, int(L), replace=False)) for L in blen]
    brp = np.zeros(nb + 1, np.uint32)
    np.cumsum(blen, out=brp[1:])
    B = gpu.CSR(nb, n, brp, np.concatenate(bcols).astype(np.uint32),
                rng.uniform(-1, 1, int(brp[-1])))
    alen = np.array([32000, 20000, 700, 3])
    acols = [np.sort(rng.choice(nb, int(L), replace=False)) for L in alen]
    arp = np.zeros(len(alen) + 1, np.uint32)
    np.cumsum(alen, out=arp[1:])
    A = gpu.CSR(len(alen), nb, arp, np.concatenate(acols).astype(np.uint32),
                rng.uniform(-1, 1, int(arp[-1])))
    return A, B


@pytest.mark.parametrize("log_cols", [20, 21])
def test_rank1_row_pieces(gpu, log_cols):
    """Rows of more than 2^21 products are split into pieces of consecutive
    tiles on different workgroups.  Precise workflow: each piece starts from
    the outputs counted in the stored words before it; over 2^20 columns the
    bitmaps come from the dense symbolic kernel, over 2^21 (wider than its
    LDS bitmap) from the windowed symbolic kernel; then a value-only replay
    of the plan.  Estimation workflow (no stored bitmaps): a count pass
    gives each piece its outputs, and a row over its estimated capacity is
    listed once for the recount."""
    log(f"case: rank1 row pieces, B 2^15 x 2^{log_cols}")
    A, B = heavy_rows(gpu, log_cols)
    C, st = _run(gpu, A, B, gpu.Config(use_tiled_numeric=3), 0)
    assert st["decision"]["stored_row_bitmap_bytes"] > 0
    worst, _ = check_product(A, B, C, TOL64, f"row pieces 2^{log_cols}")
    log(f"row pieces 2^{log_cols}: nnz {C.nnz}, err {worst:.2e}")
    for coe in (None, 0.5):  # 0.5: capacities below the counts, rows overflow
        cfg = gpu.Config(use_tiled_numeric=3)
        if coe:
            cfg = gpu.Config(use_tiled_numeric=3, hll_expansion_low=coe, hll_expansion_high=coe)
        C, st = _run(gpu, A, B, cfg, 1)
        assert st["decision"]["workflow"] == 1
        ovf = (st.get("estimation") or {}).get("overflow_rows")
        if coe:
            assert ovf and ovf > 0, "no row overflowed its halved capacity"
        worst, _ = check_product(A, B, C, TOL64, f"row pieces est 2^{log_cols} coe {coe}")
        log(f"row pieces estimation 2^{log_cols} expansion {coe}: overflow rows {ovf}, "
            f"err {worst:.2e}")
    dA, dB = A.upload(), B.upload()
    sp = gpu.SpGEMM(gpu.Config(use_tiled_numeric=3))
    sp.set_workflow(0)
    dC = sp.plan(dA, dB)
    vals = np.ascontiguousarray(np.random.default_rng(3).uniform(-1, 1, A.nnz))
    gpu._check(gpu.lib.oc_memcpy_htod(dA.m.values, vals.ctypes.data, vals.nbytes))
    A.values = vals
    sp.replay()
    worst, _ = check_product(A, B, dC.download(), TOL64, f"row pieces replay 2^{log_cols}")
    log(f"row pieces replay 2^{log_cols}: err {worst:.2e}")
