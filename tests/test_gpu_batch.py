"""GPU parity of value-only replays (plan, eager replay, HIP graph) and of
batched products (block-diagonal concatenation) against the oracle."""
import numpy as np
import pytest

from helpers import check_product, log

pytestmark = pytest.mark.gpu


def _set_values(gpu, dM, M, vals):
    """Overwrite the values of a device matrix in place (same structure)."""
    vals = np.ascontiguousarray(vals, dtype=M.values.dtype)
    gpu._check(gpu.lib.oc_memcpy_htod(dM.m.values, vals.ctypes.data, vals.nbytes))
    M.values = vals


@pytest.mark.parametrize("kind", ["rmat13", "rmat15_long_rows"])
def test_plan_replay_graph(gpu, kind):
    log(f"case: plan/replay/graph on {kind}")
    A = gpu.CSR.rmat(13, 16 << 13, seed=3) if kind == "rmat13" else gpu.CSR.rmat(15, 32 << 15, seed=5)
    dA = dA_keep = A.upload()
    sp = gpu.SpGEMM(gpu.Config())
    dC = sp.plan(dA, dA)
    assert sp.get_stats()["decision"]["workflow"] == 0
    check_product(A, A, dC.download(), 1e-12, f"{kind} plan")
    rng = np.random.default_rng(1)
    sp.capture()
    for it in range(3):
        _set_values(gpu, dA_keep, A, rng.uniform(-1, 1, A.nnz))
        if it == 0:
            sp.replay()
        else:
            sp.graph_launch()
        got = dC.download()
        worst, _ = check_product(A, A, got, 1e-12, f"{kind} replay {it}")
        # the eager product of the same values agrees
        eager = gpu.SpGEMM(gpu.Config()).run(dA, dA).download()
        assert np.array_equal(eager.col_ids, got.col_ids)
        log(f"{kind} replay {it} ({'eager' if it == 0 else 'graph'}): err {worst:.2e}")
    # graph replays with unchanged values keep the result (sums may be
    # reordered by the LDS FP64 atomics, so compare against the oracle)
    sp.graph_launch()
    sp.graph_launch()
    worst, _ = check_product(A, A, dC.download(), 1e-12, f"{kind} repeated graph")
    log(f"{kind} repeated graph launches: err {worst:.2e}")


def _mixed_batch(gpu, count, vtype, seed):
    rng = np.random.default_rng(seed)
    mats = []
    for i in range(count):
        n = int(np.exp(rng.uniform(np.log(200), np.log(3000))))
        t = i % 3
        if t == 0:
            M = gpu.CSR.erdos_renyi(n, float(rng.integers(2, 33)), vtype=vtype, seed=seed * 1000 + i)
        elif t == 1:
            M = gpu.CSR.banded(n, int(rng.integers(3, 66)) | 1, vtype=vtype, seed=seed * 1000 + i)
        else:
            M = gpu.CSR.rmat(max(8, int(np.log2(n))), 8 << max(8, int(np.log2(n))), vtype=vtype,
                             seed=seed * 1000 + i)
        mats.append(M)
    return mats


@pytest.mark.parametrize("vtype,tol", [(0, 1e-12), (1, 1e-5)])
def test_batch_run_matches_per_product_oracle(gpu, vtype, tol):
    log(f"case: batch of 40 mixed products, vtype {vtype}")
    mats = _mixed_batch(gpu, 40, vtype, 7 + vtype)
    dev = [M.upload() for M in mats]
    b = gpu.Batch(gpu.Config())
    b.set_inputs(dev)
    b.run()
    view, rb, cb = b.result()
    assert view.validate() == 0
    worst = 0.0
    for i, M in enumerate(mats):
        Ci = b.extract(i).download()
        assert Ci.rows == M.rows and Ci.cols == M.cols
        w, _ = check_product(M, M, Ci, tol, f"batch product {i}")
        worst = max(worst, w)
    assert rb[-1] == sum(M.rows for M in mats) and cb[-1] == sum(M.cols for M in mats)
    log(f"batch vtype {vtype}: workflow {b.get_stats()['decision']['workflow']}, "
        f"max normalised error {worst:.2e}")


def test_batch_rectangular_pairs(gpu):
    log("case: batch of 12 rectangular A_i * A_i^T, f32")
    As = [gpu.CSR.rect_band(300 + 50 * i, 4000 + 300 * i, 6.0, 0.3, 512, 8, vtype=1, seed=60 + i)
          for i in range(12)]
    Bs = [A.transpose() for A in As]
    b = gpu.Batch(gpu.Config())
    b.set_inputs([A.upload() for A in As], [B.upload() for B in Bs])
    b.run()
    for i in range(12):
        check_product(As[i], Bs[i], b.extract(i).download(), 1e-5, f"rect pair {i}")


def test_batch_graph_refresh(gpu):
    log("case: batch plan + captured graph, values refreshed 3 times")
    mats = _mixed_batch(gpu, 24, 1, 21)
    dev = [M.upload() for M in mats]
    b = gpu.Batch(gpu.Config())
    b.set_inputs(dev)
    b.plan()
    b.capture()
    rng = np.random.default_rng(5)
    for it in range(3):
        for M, dM in zip(mats, dev):
            _set_values(gpu, dM, M, rng.uniform(-1, 1, M.nnz).astype(np.float32))
        b.launch()
        for i in (0, 5, 11, 23):
            check_product(mats[i], mats[i], b.extract(i).download(), 1e-5, f"graph {it} product {i}")
        # the same values through an eager batch run give the same structure
        e = gpu.Batch(gpu.Config())
        e.set_inputs(dev)
        e.run()
        v1, _, _ = e.result()
        v2, _, _ = b.result()
        assert np.array_equal(v1.row_offsets(), v2.row_offsets())
        log(f"graph refresh {it}: ok")


@pytest.mark.parametrize("mode", ["run", "graph", "learned_est"])
def test_batch_long_rows_block_tiles(gpu, mode):
    """A batch whose products have rows beyond the LDS kernels: the
    rank-tile kernel takes them with tiles counted from each product's
    column block (the concatenated B is far wider than one block).  Each
    product equals the oracle's, for a plain run, a graph replay after new
    values, and the estimation workflow with every estimate set to 1 (all
    long rows overflow and are recomputed by the same kernel)."""
    log(f"case: batch with long rows ({mode})")
    mats = [gpu.CSR.rmat(14, 32 << 14, seed=60 + i) for i in range(2)]
    mats += [gpu.CSR.erdos_renyi(3000, 6.0, seed=70), gpu.CSR.rmat(13, 16 << 13, seed=71)]
    dev = [M.upload() for M in mats]
    b = gpu.Batch(gpu.Config(track_stage_times=1))
    b.set_inputs(dev)
    if mode == "graph":
        b.plan()
        b.capture()
        rng = np.random.default_rng(3)
        for d, M in zip(dev, mats):
            _set_values(gpu, d, M, rng.uniform(-1, 1, M.nnz))
        b.launch()
    elif mode == "learned_est":
        rows = sum(M.rows for M in mats)
        est = gpu.DeviceBuffer.from_array(np.ones(rows, np.uint32))
        h = gpu.lib.oc_batch_handle(b.h)
        gpu._check(gpu.lib.oc_spgemm_set_workflow(h, 1))
        gpu._check(gpu.lib.oc_spgemm_set_estimates(h, est.ptr))
        b.run()
        gpu._check(gpu.lib.oc_spgemm_set_estimates(h, None))
    else:
        b.run()
    st = b.get_stats()
    if mode != "graph":
        assert st["binning_2"][14] - st["binning_2"][13] > 0 or mode == "learned_est"
        fam = set(st.get("kernels", {}))
        log(f"{mode}: workflow {st['decision']['workflow']}, long rows "
            f"{st['binning_2'][14] - st['binning_2'][13]}, families {sorted(fam)}")
    worst = 0.0
    for i, M in enumerate(mats):
        Ci = b.extract(i).download()
        w, _ = check_product(M, M, Ci, 1e-12, f"{mode} product {i}")
        worst = max(worst, w)
    log(f"{mode}: max normalised error {worst:.2e}")


@pytest.mark.parametrize("mode", ["run", "graph"])
def test_batch_degenerate_products(gpu, mode):
    """A batch holding a product with no rows, one with rows but no
    entries, a 1x1 product, one whose A entries meet only empty B rows,
    and two ordinary products: each extracted product equals the oracle's,
    for a run and for a captured graph replayed after new values.  A
    batch of one product is checked too."""
    from test_gpu_edges import _cases
    log(f"case: batch of degenerate products ({mode})")
    cs = _cases(gpu)
    pairs = [cs[k] for k in ("no_rows", "no_entries", "one_entry", "only_empty_b_rows", "ragged")]
    # A and B of this pair are separate host objects: new values go to A only
    pairs.insert(2, (gpu.CSR.rmat(10, 8 << 10, seed=81), gpu.CSR.rmat(10, 8 << 10, seed=81)))
    As, Bs = [p[0] for p in pairs], [p[1] for p in pairs]
    dAs, dBs = [A.upload() for A in As], [B.upload() for B in Bs]
    b = gpu.Batch(gpu.Config())
    b.set_inputs(dAs, dBs)
    if mode == "graph":
        b.plan()
        b.capture()
        rng = np.random.default_rng(8)
        for A, dA in zip(As, dAs):
            if A.nnz:
                _set_values(gpu, dA, A, rng.uniform(-1, 1, A.nnz))
        b.launch()
    else:
        b.run()
    view, rb, cb = b.result()
    assert rb[-1] == sum(A.rows for A in As) and cb[-1] == sum(B.cols for B in Bs)
    for i, (A, B) in enumerate(zip(As, Bs)):
        check_product(A, B, b.extract(i).download(), 1e-12, f"degenerate batch {mode} {i}")
    one = gpu.Batch(gpu.Config())
    one.set_inputs([dAs[5]], [dBs[5]])
    one.run()
    check_product(As[5], Bs[5], one.extract(0).download(), 1e-12, "batch of one")
