"""The autotuner on the GPU: every candidate of its space runs, the chosen
configuration is never slower than the default by its own score, and every
accepted candidate kept the default's output structure (checked here again
against the oracle for the chosen configuration)."""
import numpy as np
import pytest

from helpers import check_product, log

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("workload", ["product", "batched"])
def test_autotune_runs_space_and_keeps_structure(gpu, workload):
    from research_amd import autotune
    log(f"case: autotune over its space, {workload}")
    if workload == "product":
        A = gpu.CSR.rmat(12, 16 << 12, seed=11)
        w = autotune._Product(A)
    else:
        mats = [gpu.CSR.erdos_renyi(600 + 50 * i, 6.0, vtype=1, seed=40 + i) for i in range(6)]
        mats += [gpu.CSR.rmat(10, 8 << 10, vtype=1, seed=60 + i) for i in range(2)]
        w = autotune._Batched(mats)
    res = autotune.tune(w, reps=1, passes=1, log=None)
    n_cand = sum(len(v) - 1 for v in autotune.SPACE.values())
    assert res["evaluations"] == 1 + n_cand
    assert res["ms"] <= res["default_ms"]
    # scalar knobs always, per-bin shapes as 7-entry arrays when chosen
    assert {k for k in autotune.SPACE if isinstance(k, str)} <= set(res["config"])
    assert set(res["config"]) <= {k if isinstance(k, str) else k[0] for k in autotune.SPACE}
    for h in res["history"][1:]:
        assert "error" not in h, h
    accepted = [h for h in res["history"][1:] if h["config"] == res["config"]]
    if res["config"] != res["history"][0]["config"]:
        assert accepted and accepted[-1]["same_structure"]
    # the chosen configuration's product against the oracle
    cfg = gpu.Config(**res["config"])
    if workload == "product":
        C = gpu.SpGEMM(cfg).run(w.dA, w.dA).download()
        check_product(A, A, C, 1e-12, "autotuned product")
    else:
        b = gpu.Batch(cfg)
        b.set_inputs(w.dev)
        b.run()
        for i in (0, len(mats) - 1):
            check_product(mats[i], mats[i], b.extract(i).download(), 1e-5, f"autotuned batch {i}")
    log(f"autotune {workload}: {res['default_ms']:.2f} -> {res['ms']:.2f} ms, {res['config']}")
