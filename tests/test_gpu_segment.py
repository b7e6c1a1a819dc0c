"""Spy-plot segmentation on the GPU (stats JSON `segmentation` block)
against the oracle: integers exact, FP64 statistics to 1e-12 relative; C
itself unchanged by the analysis."""
import numpy as np
import pytest

from helpers import log
from oracle import oracle as ora
from oracle import segment as seg

pytestmark = pytest.mark.gpu


def _two_blocks(gpu):
    rng = np.random.default_rng(5)
    n = 4096
    rows, cols = [], []
    for b0 in (0, 2048):
        for r in range(b0, b0 + 1024):
            c = b0 + rng.choice(1024, 200, replace=False)
            rows += [r] * len(c)
            cols += c.tolist()
    for r in range(n):
        c = rng.choice(n, 2, replace=False)
        rows += [r] * 2
        cols += c.tolist()
    key = np.unique(np.array(rows, np.int64) * n + np.array(cols, np.int64))
    r, c = key // n, key % n
    rp = np.zeros(n + 1, np.uint32)
    np.cumsum(np.bincount(r, minlength=n), out=rp[1:])
    return gpu.CSR(n, n, rp, c.astype(np.uint32), rng.uniform(-1, 1, len(c)))


CASES = {
    "rmat14": (lambda g: (g.CSR.rmat(14, 16 << 14, seed=1), None), {}),
    "two_blocks": (lambda g: (_two_blocks(g), None), {"segmentation_grid_max": 64}),
    "banded16": (lambda g: (g.CSR.banded(1 << 16, 33, seed=2), None),
                 {"segmentation_grid_max": 128, "segmentation_min_segment_rows": 256}),
    "rect_grid8": (lambda g: (lambda A: (A, A.transpose()))(
        g.CSR.rect_band(1 << 13, 1 << 15, 8.0, 0.3, 4096, 4, vtype=1, seed=3)),
        {"segmentation_grid_max": 8, "segmentation_hot_threshold_k": 0.5}),
    "er_capped": (lambda g: (g.CSR.erdos_renyi(1 << 13, 8.0, seed=4), None),
                  {"segmentation_cc_max_iters": 2, "segmentation_hot_threshold_k": 1.0,
                   "segmentation_min_segment_rows": 64}),
}


@pytest.mark.parametrize("name", list(CASES))
def test_segmentation_matches_oracle(gpu, name):
    log(f"case: segmentation {name}")
    make, cfg = CASES[name]
    A, B = make(gpu)
    B = A if B is None else B
    dA, dB = A.upload(), B.upload()
    sp = gpu.SpGEMM(gpu.Config(use_segmented_workflow=1, **cfg))
    dC = sp.run(dA, dB)
    stats = sp.get_stats()
    got = stats["segmentation"]
    # the keys utils/plot_results.py reads (:43, :61, :64-67, :111-120)
    assert isinstance(stats["gflops"], float) and isinstance(stats["total_time"], float)
    for s in got["segments"]:
        assert {"class", "col_left", "col_right", "row_start", "row_end"} <= set(s)
    an = ora.analysis(A, B)
    want = seg.segmentation(A, B, an, grid_max=cfg.get("segmentation_grid_max", 256),
                            k=cfg.get("segmentation_hot_threshold_k", 2.0),
                            min_rows=cfg.get("segmentation_min_segment_rows", 1024),
                            exp_threshold=8, cc_max_iters=cfg.get("segmentation_cc_max_iters", 32))
    for k in ("grid_h", "grid_w", "tile_rows", "tile_cols", "num_components", "cc_iterations"):
        assert got[k] == want[k], k
    for k in ("density_mean", "density_stddev"):
        assert got[k] == pytest.approx(want[k], rel=1e-12, abs=1e-12), k
    assert len(got["segments"]) == len(want["segments"])
    for g, w in zip(got["segments"], want["segments"]):
        for k, v in w.items():
            if isinstance(v, float):
                assert g[k] == pytest.approx(v, rel=1e-12, abs=1e-12), k
            else:
                assert g[k] == v, k
    for k in ("time_density_ms", "time_threshold_ms", "time_cc_ms", "time_classify_ms"):
        assert got[k] >= 0.0
    # analysis only: C equals the product without segmentation
    C1 = dC.download()
    sp2 = gpu.SpGEMM(gpu.Config())
    C2 = sp2.run(dA, dB).download()
    assert "segmentation" not in sp2.get_stats()
    assert np.array_equal(C1.row_offsets, C2.row_offsets) and np.array_equal(C1.col_ids, C2.col_ids)
    log(f"segmentation {name}: {len(got['segments'])} segments, "
        f"{got['num_components']} components, {got['cc_iterations']} sweeps, classes "
        f"{sorted({s['class'] for s in got['segments']})}")
