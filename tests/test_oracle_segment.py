"""The segmentation oracle (oracle/segment.py) pinned by independent
computations: tile counts by numpy.histogram2d, components by
scipy.ndimage.label, a hand-built two-block matrix, and segment sums by
brute force."""
import numpy as np
import pytest
import scipy.ndimage

import research_amd as ra
from oracle import oracle as ora
from oracle import segment as seg


def _csr(rows, cols, entries, vtype=np.float64):
    entries = sorted(set(entries))
    rp = np.zeros(rows + 1, np.uint32)
    for r, _ in entries:
        rp[r + 1] += 1
    rp = np.cumsum(rp).astype(np.uint32)
    ci = np.array([c for _, c in entries], np.uint32)
    return ra.CSR(rows, cols, rp, ci, np.ones(len(ci), vtype))


@pytest.mark.parametrize("rows,cols,gmax", [(1000, 700, 64), (37, 5000, 256), (5000, 37, 8)])
def test_density_against_histogram2d(rows, cols, gmax):
    A = ra.CSR.rect_band(rows, cols, 12.0, 0.4, 300, max(1, cols // rows), seed=3)
    gh, gw, tr, tc = seg.grid_shape(rows, cols, gmax)
    d = seg.density(A, gh, gw, tr, tc)
    r = np.repeat(np.arange(rows), np.diff(A.row_offsets.astype(np.int64)))
    c = A.col_ids.astype(np.int64)
    # bin edges of the tiles, the last bin open-ended (indices past the grid clip into it)
    er = np.append(np.arange(gh) * tr, np.inf)
    ec = np.append(np.arange(gw) * tc, np.inf)
    h, _, _ = np.histogram2d(r, c, bins=[er, ec])
    assert np.array_equal(d, h.astype(np.uint64)) and int(d.sum()) == A.nnz


@pytest.mark.parametrize("p,seed", [(0.2, 0), (0.45, 1), (0.6, 2), (0.8, 3)])
def test_components_against_scipy(p, seed):
    rng = np.random.default_rng(seed)
    mask = rng.random((48, 61)) < p
    lab0 = np.where(mask.ravel(), np.arange(mask.size), seg.EMPTY).reshape(mask.shape)
    lab, sweeps = seg.components(lab0, max_sweeps=mask.size)  # to convergence
    ref, n = scipy.ndimage.label(mask)  # default structure: 4-connectivity
    flat = np.arange(mask.size).reshape(mask.shape)
    assert np.all(lab[~mask] == seg.EMPTY)
    for k in range(1, n + 1):
        comp = ref == k
        assert np.all(lab[comp] == flat[comp].min())
    assert len(np.unique(lab[mask])) == n
    assert sweeps >= 0 and (n == 0 or sweeps < mask.size)


def test_two_dense_blocks():
    """Two dense diagonal blocks in a sparse background: 2 components, four
    bands of 1024 rows (block, background, block, background)."""
    rng = np.random.default_rng(5)
    n = 4096
    ent = set()
    for b0 in (0, 2048):
        for r in range(b0, b0 + 1024):
            for c in rng.choice(1024, 200, replace=False):
                ent.add((r, b0 + int(c)))
    for r in range(n):
        for c in rng.choice(n, 2, replace=False):
            ent.add((r, int(c)))
    A = _csr(n, n, list(ent))
    an = ora.analysis(A, A)
    out = seg.segmentation(A, A, an, grid_max=64, k=2.0, min_rows=1024)
    assert (out["grid_h"], out["grid_w"], out["tile_rows"], out["tile_cols"]) == (64, 64, 64, 64)
    assert out["num_components"] == 2
    hot = out["_labels"] != seg.EMPTY
    assert hot[:16, :16].all() and hot[32:48, 32:48].all() and hot.sum() == 512
    got = [(s["row_start"], s["row_end"]) for s in out["segments"]]
    assert got == [(0, 1024), (1024, 2048), (2048, 3072), (3072, 4096)]
    assert set(np.unique(out["_labels"][hot]).tolist()) == {0, 32 * 64 + 32}


def test_segment_sums_brute_force():
    A = ra.CSR.rmat(11, 8 << 11, seed=4)
    an = ora.analysis(A, A)
    segs = [(0, 100), (100, 101), (101, 2048), (700, 700)]
    out = seg.classify(segs, A, an, A.cols, 8)
    ro = A.row_offsets.astype(np.int64)
    for (s0, s1), o in zip(segs, out):
        rows = range(s0, s1)
        assert o["sum_products"] == sum(int(an["products"][r]) for r in rows)
        assert o["sum_a_nnz"] == sum(int(ro[r + 1] - ro[r]) for r in rows)
        assert o["max_a_row_len"] == max([int(ro[r + 1] - ro[r]) for r in rows], default=0)
        assert o["row_count"] == s1 - s0
        lefts = [int(an["left"][r]) for r in rows if an["left"][r] != seg.EMPTY]
        assert o["col_left"] == (min(lefts) if lefts else 0)
        if s1 > s0:
            p = np.array([float(an["products"][r]) for r in rows])
            assert o["mean_row_len"] == pytest.approx(p.mean(), rel=1e-12)
            assert o["var_row_len"] == pytest.approx(p.var(), rel=1e-9, abs=1e-9)
        assert o["class"] in seg.CLASSES


def test_class_rule_cases():
    """One matrix per class, each for a reason stated without the rule's
    code: 1024 rows of a banded matrix (they reach about 1060 of its
    100000 columns, under 5%, with 33 products per A entry) are banded; rows whose products fill most of their span are dense; an
    R-MAT graph's heavy-tailed row products (coefficient of variation above
    1.5, measured here with numpy) are power law; Erdos-Renyi rows are
    sparse uniform."""
    cases = []
    B = ra.CSR.banded(100000, 33, seed=1)
    cases.append((B, "BANDED", (50000, 51024)))
    rng = np.random.default_rng(2)
    n = 256
    D = _csr(n, n, [(r, int(c)) for r in range(n) for c in rng.choice(n, 128, replace=False)])
    cases.append((D, "DENSE", (0, n)))
    R = ra.CSR.rmat(14, 16 << 14, seed=3)
    anR = ora.analysis(R, R)
    p = anR["products"].astype(np.float64)
    assert p.std() / p.mean() >= 1.5
    cases.append((R, "POWER_LAW", (0, R.rows)))
    E = ra.CSR.erdos_renyi(20000, 8.0, seed=4)
    cases.append((E, "SPARSE_UNIFORM", (0, E.rows)))
    for M, want, rows in cases:
        an = ora.analysis(M, M)
        (o,) = seg.classify([rows], M, an, M.cols, 8)
        assert o["class"] == want, (want, o)


@pytest.mark.parametrize("sweeps", [1, 2, 5])
def test_capped_propagation_closed_form(sweeps):
    """A 1 x L line of hot tiles: after s sweeps tile x holds max(0, x - s);
    each pointer-jumping round doubles the reach, so after the 8 rounds
    tile x holds max(0, x - 256 s) (rounds stop early only once all reach
    0)."""
    L = 3000
    lab0 = np.arange(L, dtype=np.int64).reshape(1, L)
    lab, done = seg.components(lab0, max_sweeps=sweeps)
    assert done == sweeps
    x = np.arange(L)
    assert np.array_equal(lab.ravel(), np.maximum(0, x - 256 * sweeps))
