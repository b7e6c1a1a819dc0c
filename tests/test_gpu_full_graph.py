"""The whole path captured into one HIP graph (oc_spgemm_capture_full,
oc_batch_capture_full): analysis, binning, symbolic and numeric with every
bin kernel persistent over its device-side row range, so structure and
values are recomputed on each launch.  The reference reads bin sizes and
totals back between these stages (kernels/SpGEMM.cuh:509-514, 611-615,
809-810, 1578-1580).

Inputs change structure between launches by symmetric permutations
P A P^T, which keep every dimension, entry count and nnz(C) while moving
every entry; each launch is checked against the oracle (structure
bit-exact, values 1e-12 / 1e-5 condition-normalised), and graph launches
alternate with eager work, as the graph rules here require."""
import numpy as np
import pytest
import scipy.sparse as sps

from helpers import check_product, log

pytestmark = pytest.mark.gpu

TOL64, TOL32 = 1e-12, 1e-5


def _permuted(gpu, M, seed):
    """P M P^T for a random permutation P (same shape, nnz and nnz(M*M))."""
    p = np.random.default_rng(seed).permutation(M.rows)
    S = M.to_scipy().tocoo()
    T = sps.csr_matrix((S.data, (p[S.row], p[S.col])), shape=S.shape)
    T.sort_indices()
    return gpu.CSR(M.rows, M.cols, T.indptr.astype(np.uint32), T.indices.astype(np.uint32),
                   T.data.astype(M.values.dtype))


def _overwrite(gpu, dM, M):
    """Replace a device matrix's structure and values in place."""
    for dst, arr in ((dM.m.row_offsets, M.row_offsets), (dM.m.col_ids, M.col_ids),
                     (dM.m.values, M.values)):
        a = np.ascontiguousarray(arr)
        gpu._check(gpu.lib.oc_memcpy_htod(dst, a.ctypes.data, a.nbytes))


@pytest.mark.parametrize("tiled", [2, 3])
def test_full_graph_single_product(gpu, tiled):
    """R-MAT 2^14, degree 24 (hash, dense and long-row bins): the captured
    whole path on the planned input, then on three permutations of it, with
    eager products between graph launches; the eager whole-path replay too."""
    log(f"case: whole-path graph, rmat 2^14 deg 24, use_tiled_numeric {tiled}")
    A = gpu.CSR.rmat(14, 24 << 14, seed=5)
    dA = A.upload()
    sp = gpu.SpGEMM(gpu.Config(use_tiled_numeric=tiled))
    dC = sp.plan(dA, dA)
    nnz0 = dC.nnz
    st = sp.get_stats()
    assert st["binning_2"][14] - st["binning_2"][13] > 0, "no long rows"
    sp.capture_full()
    sp.graph_launch_full()
    assert sp.full_status() == 0
    check_product(A, A, dC.download(), TOL64, "full graph, planned input")
    for it in range(3):
        M = _permuted(gpu, A, 100 + it)
        _overwrite(gpu, dA, M)
        if it == 1:
            sp.full_replay()
        else:
            sp.graph_launch_full()
        assert sp.full_status() == 0
        got = dC.download()
        assert got.nnz == nnz0
        worst, _ = check_product(M, M, got, TOL64, f"full graph, permutation {it}")
        # eager work between launches
        eager = gpu.SpGEMM(gpu.Config()).run(dA, dA).download()
        assert np.array_equal(eager.col_ids, got.col_ids)
        log(f"permutation {it} ({'eager' if it == 1 else 'graph'}): err {worst:.2e}")


def test_full_graph_checks(gpu):
    """Inputs the plan cannot hold set the check bits and leave C unwritten:
    a product with more outputs than the planned capacity (bit 1) and an A
    row longer than the planned longest (bit 2)."""
    log("case: whole-path graph checks")
    n = 4000
    A = gpu.CSR.banded(n, 9, seed=3)  # 9 entries per row but the ends
    dA = A.upload()
    sp = gpu.SpGEMM(gpu.Config(use_ultrasparse_workflow=0))
    dC = sp.plan(dA, dA)
    sp.capture_full()
    # same nnz, entries spread: more outputs per row than the band gives
    rng = np.random.default_rng(4)
    lens = np.diff(A.row_offsets.astype(np.int64))
    cols = np.concatenate([np.sort(rng.choice(n, int(L), replace=False)) for L in lens])
    M = gpu.CSR(n, n, A.row_offsets.copy(), cols.astype(np.uint32), A.values.copy())
    _overwrite(gpu, dA, M)
    sp.graph_launch_full()
    assert sp.full_status() & 1, "capacity overflow not flagged"
    # one row holding 20 entries (the planned longest has 9)
    lens2 = lens.copy()
    lens2[100] += 11
    lens2[200:211] -= 1
    rp = np.zeros(n + 1, np.uint32)
    np.cumsum(lens2, out=rp[1:])
    cols2 = np.concatenate([np.sort(rng.choice(n, int(L), replace=False)) for L in lens2])
    M2 = gpu.CSR(n, n, rp, cols2.astype(np.uint32), A.values.copy())
    _overwrite(gpu, dA, M2)
    sp.graph_launch_full()
    assert sp.full_status() & 2, "long A row not flagged"
    # the planned input again: no bits, the right product
    _overwrite(gpu, dA, A)
    sp.graph_launch_full()
    assert sp.full_status() == 0
    check_product(A, A, dC.download(), TOL64, "full graph after the checks")


def test_full_graph_batch(gpu):
    """40 mixed products (ER, banded, R-MAT; f32) as one batch: the whole
    path of every product in one graph, launched on the planned inputs and
    on two sets of permuted inputs; every product's structure against a
    fresh batched run, and a sample of products against the oracle."""
    log("case: whole-path graph of a batch of 40 products")
    rng = np.random.default_rng(8)
    mats = []
    for i in range(40):
        nn = int(np.exp(rng.uniform(np.log(300), np.log(4000))))
        if i % 3 == 0:
            M = gpu.CSR.erdos_renyi(nn, float(rng.integers(2, 33)), vtype=1, seed=800 + i)
        elif i % 3 == 1:
            M = gpu.CSR.banded(nn, int(rng.integers(3, 66)) | 1, vtype=1, seed=800 + i)
        else:
            lg = max(8, int(np.log2(nn)))
            M = gpu.CSR.rmat(lg, 8 << lg, vtype=1, seed=800 + i)
        mats.append(M)
    dev = [M.upload() for M in mats]
    b = gpu.Batch(gpu.Config())
    b.set_inputs(dev)
    b.plan()
    b.capture_full()
    for it in range(3):
        if it:
            mats = [_permuted(gpu, M, 1000 * it + i) for i, M in enumerate(mats)]
            for d, M in zip(dev, mats):
                _overwrite(gpu, d, M)
        b.launch_full()
        assert b.full_status() == 0
        view, _, _ = b.result()
        h = view.row_set_hash()
        ref = gpu.Batch(gpu.Config())
        ref.set_inputs(dev)
        ref.run()
        rview, _, _ = ref.result()
        assert np.array_equal(h, rview.row_set_hash()), f"launch {it}: structure differs"
        for i in (0, 7, 20, 39):
            check_product(mats[i], mats[i], b.extract(i).download(), TOL32,
                          f"batch launch {it} product {i}")
        log(f"batch whole-path launch {it}: structure of 40 products identical, "
            f"4 products against the oracle")


def test_full_graph_estimated_symbolic(gpu):
    """A product whose sampled compaction turns on the shrunken symbolic
    tables (the banded matrix with spread rows of test_gpu_wide.py): in the
    captured whole path the rows that fill their tables are recounted by the
    windowed kernel over the device-side overflow list, and the product of
    permuted inputs still equals the oracle's."""
    from test_gpu_wide import est_sym_matrix
    log("case: whole-path graph with estimated symbolic tables")
    M = est_sym_matrix(gpu)
    dM = M.upload()
    sp = gpu.SpGEMM(gpu.Config())
    dC = sp.plan(dM, dM)
    st = sp.get_stats()
    assert st["decision"]["estimated_symbolic"] is True
    assert st["decision"]["symbolic_overflow_rows"] > 0
    sp.capture_full()
    for it in range(2):
        X = M if it == 0 else _permuted(gpu, M, 77)
        _overwrite(gpu, dM, X)
        sp.graph_launch_full()
        assert sp.full_status() == 0
        worst, _ = check_product(X, X, dC.download(), TOL64, f"estimated symbolic, launch {it}")
        gpu.SpGEMM(gpu.Config()).run(dM, dM).free()  # eager work between launches
        log(f"estimated symbolic whole path, launch {it}: err {worst:.2e}")
