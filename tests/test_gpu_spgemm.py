"""End-to-end GPU parity of C = A*B against the FP64 oracle (structure
bit-exact, values within the condition-normalised tolerance)."""
import numpy as np
import pytest

from helpers import check_product, log

pytestmark = pytest.mark.gpu

TOL64, TOL32 = 1e-12, 1e-5


def _run(gpu, A, B, cfg=None, workflow=None):
    dA, dB = A.upload(), B.upload()
    sp = gpu.SpGEMM(cfg or gpu.Config())
    if workflow is not None:
        sp.set_workflow(workflow)
    dC = sp.run(dA, dB)
    assert dC.validate() == 0
    C = dC.download()
    return C, sp.get_stats()


def test_laplacian_ultrasparse_bit_exact(gpu):
    log("case: laplacian 256x256 f64 seed0 (ESC)")
    A = gpu.CSR.laplacian2d(256, 256, seed=0)
    C, st = _run(gpu, A, A)
    assert st["decision"]["workflow"] == 2
    assert C.nnz == 846852 and st["stats"]["intermediate_product"] == 1629192
    worst, plain = check_product(A, A, C, TOL64, "laplacian", plain_exact=True)
    log(f"laplacian: normalised {worst:.2e} plain {plain:.2e}")


@pytest.mark.parametrize("log_n,deg", [(10, 8), (12, 16), (14, 16)])
def test_rmat_default_workflow(gpu, log_n, deg):
    log(f"case: rmat 2^{log_n} deg {deg} f64 (auto workflow)")
    A = gpu.CSR.rmat(log_n, deg << log_n, seed=1)
    C, st = _run(gpu, A, A)
    worst, plain = check_product(A, A, C, TOL64, f"rmat{log_n}")
    log(f"rmat {log_n}: workflow {st['decision']['workflow']} err {worst:.2e} plain {plain:.2e}")


@pytest.mark.parametrize("workflow", [0, 1])
def test_rmat_forced_workflows(gpu, workflow):
    log(f"case: rmat 2^13 deg 16 f64 forced workflow {workflow}")
    A = gpu.CSR.rmat(13, 16 << 13, seed=3)
    cfg = gpu.Config(use_ultrasparse_workflow=0)
    C, st = _run(gpu, A, A, cfg, workflow)
    assert st["decision"]["workflow"] == workflow
    worst, plain = check_product(A, A, C, TOL64, f"rmat13 wf{workflow}")
    log(f"forced {workflow}: err {worst:.2e}, overflow rows {st.get('estimation', {}).get('overflow_rows')}")


def _drop_rows(gpu, M, every):
    """M with every `every`-th row emptied."""
    ro = M.row_offsets.astype(np.int64)
    keep = np.ones(M.nnz, bool)
    lens = np.diff(ro)
    for r in range(0, M.rows, every):
        keep[ro[r]:ro[r + 1]] = False
        lens[r] = 0
    rp = np.zeros(M.rows + 1, np.uint32)
    np.cumsum(lens, out=rp[1:])
    return gpu.CSR(M.rows, M.cols, rp, M.col_ids[keep], M.values[keep])


@pytest.mark.parametrize("rows,bcols,vt,tol", [(40, 700, 0, TOL64), (3, 2900, 1, TOL32),
                                               (300, 64, 0, TOL64)])
def test_fat_workflow(gpu, rows, bcols, vt, tol):
    """Short-and-wide A times tall-and-narrow B (A.cols / fat_ratio > A.rows,
    B.rows / fat_ratio > B.cols, B.cols below the widest dense bin): the
    FAT workflow (reported as 3), checked against the oracle and against
    the same product with use_fat_workflow off.  A has empty rows and
    B empty rows and never-hit columns."""
    inner = 300000
    log(f"case: fat {rows} x {inner} times {inner} x {bcols}, vtype {vt}")
    A = gpu.CSR.rect_band(rows, inner, 1500.0, 0.2, 4096, inner // rows, vtype=vt, seed=5)
    if rows > 3:
        A = _drop_rows(gpu, A, 7)
    Bt = gpu.CSR.rect_band(bcols, inner, 2.5 * inner / bcols, 0.5, 2048, inner // bcols,
                           vtype=vt, seed=6)
    B = Bt.transpose()
    C, st = _run(gpu, A, B, gpu.Config(track_stage_times=1))
    assert st["decision"]["workflow"] == 3
    worst, plain = check_product(A, B, C, tol, "fat")
    C2, st2 = _run(gpu, A, B, gpu.Config(use_fat_workflow=0))
    assert st2["decision"]["workflow"] != 3
    assert np.array_equal(C.row_offsets, C2.row_offsets) and np.array_equal(C.col_ids, C2.col_ids)
    log(f"fat: nnz {C.nnz}, products {st['stats']['intermediate_product']}, err {worst:.2e}, "
        f"timing {st['timing']['numeric']['fat']:.3f} ms")


@pytest.mark.parametrize("log_rows", [14, 18])
def test_rectangular_f32(gpu, log_rows):
    """A*A^T with A 2^k x 2^(k+3), Poisson(8) entries per row, 30% of them
    within +-4096 of row*8, FP32 N(0,1) values, seed 2; k = 18 is the
    benchmark's rectangular configuration at full size (ESC<32> with hash
    outliers, kernels/AccumulatorESC.cuh:129-183)."""
    log(f"case: rectangular 2^{log_rows} x 2^{log_rows + 3} poisson(8) f32, A*A^T")
    A = gpu.CSR.rect_band(1 << log_rows, 1 << (log_rows + 3), 8.0, 0.3, 4096, 8, vtype=1,
                          value_dist=1, seed=2)
    B = A.transpose()
    C, st = _run(gpu, A, B)
    worst, plain = check_product(A, B, C, TOL32, "rect")
    log(f"rect 2^{log_rows}: workflow {st['decision']['workflow']}, nnz {C.nnz}, "
        f"products {st['stats']['intermediate_product']}, err {worst:.2e}, max plain rel {plain:.2e}")


@pytest.mark.parametrize("kind", ["er", "banded"])
def test_small_types_f32(gpu, kind):
    log(f"case: {kind} f32")
    if kind == "er":
        A = gpu.CSR.erdos_renyi(3000, 12.0, vtype=1, seed=4)
    else:
        A = gpu.CSR.banded(5000, 33, vtype=1, seed=4)
    C, st = _run(gpu, A, A)
    check_product(A, A, C, TOL32, kind)


def test_precise_dense_and_hash_bins(gpu):
    log("case: precise on ER 4000 deg 40 (hash + dense bins)")
    A = gpu.CSR.erdos_renyi(4000, 40.0, seed=7)
    C, st = _run(gpu, A, A, gpu.Config(use_ultrasparse_workflow=0), 0)
    check_product(A, A, C, TOL64, "er-precise")
    log(f"bins: {st.get('binning_1')} / {st.get('binning_2')}")


@pytest.mark.parametrize("tiled,block", [(3, 1024), (2, 1024), (2, 512), (2, 256), (1, 1024),
                                         (0, 1024)])
def test_big_rows_tiled_and_windowed(gpu, tiled, block):
    """Rows beyond the LDS hash/dense kernels (numeric bin 13) through the
    one-gather rank tiles, the rank-tile kernel (three workgroup shapes),
    the dense-tile kernel and the cursor-windowed kernel."""
    log(f"case: rmat 2^15 deg 32 f64, big rows, use_tiled_numeric={tiled} block={block}")
    A = gpu.CSR.rmat(15, 32 << 15, seed=5)
    cfg = gpu.Config(use_tiled_numeric=tiled, tile_block=block, band_log2_cols=0,
                     track_stage_times=1)
    C, st = _run(gpu, A, A, cfg)
    assert st["binning_2"][14] - st["binning_2"][13] > 0, "no rows reached the big-row bin"
    worst, plain = check_product(A, A, C, TOL64, f"bigrows tiled={tiled} block={block}")
    log(f"tiled={tiled} block={block}: big rows {st['binning_2'][14] - st['binning_2'][13]} "
        f"err {worst:.2e}")


@pytest.mark.parametrize("log_n,deg,log_w,min_products,vtype",
                         [(15, 32, 14, 16384, 0), (15, 32, 12, 2000, 0), (16, 16, 15, 4000, 0),
                          (15, 32, 13, 1000, 1)])
def test_band_long_rows(gpu, log_n, deg, log_w, min_products, vtype):
    """Long rows of the precise workflow through the band-major kernels
    (count, then numeric), with other rows in the usual bins (a row takes
    the band path when it is wider than one band)."""
    log(f"case: rmat 2^{log_n} deg {deg} vtype {vtype}, bands 2^{log_w}, min products {min_products}")
    A = gpu.CSR.rmat(log_n, deg << log_n, vtype=vtype, seed=5)
    cfg = gpu.Config(band_log2_cols=log_w, band_min_products=min_products, track_stage_times=1)
    C, st = _run(gpu, A, A, cfg, 0)
    skipped = st["binning_2"][15] - st["binning_2"][14]
    assert skipped > 0 and "band_numeric" in st["kernels"], "no rows took the band path"
    worst, plain = check_product(A, A, C, TOL64 if vtype == 0 else TOL32, f"bands 2^{log_w}")
    log(f"bands 2^{log_w}: {skipped} long rows, err {worst:.2e}")


@pytest.mark.parametrize("conc,block", [([16, 16, 8, 4, 2, 1, 1], [128, 128, 256, 512, 1024, 1024, 1024]),
                                        ([32, 32, 32, 16, 8, 4, 2], [64, 64, 64, 128, 256, 256, 512])])
def test_bin_shapes(gpu, conc, block):
    """Non-default per-bin shapes (twice or half the LDS per workgroup,
    other workgroup sizes): the tables, bins and kernel instances follow,
    and the product still equals the oracle's, in both workflows."""
    log(f"case: bin shapes conc {conc} block {block}")
    A = gpu.CSR.rmat(14, 24 << 14, seed=9)
    for wf in (0, 1):
        C, st = _run(gpu, A, A, gpu.Config(bin_concurrency=conc, bin_block=block), wf)
        worst, _ = check_product(A, A, C, TOL64, f"shapes wf {wf}")
        log(f"shapes {block} wf {wf}: bins {st['binning_2'][:14]}, err {worst:.2e}")
