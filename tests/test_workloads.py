"""Host-side helpers of the predictor evaluation, the batched benchmark and
the autotuner (no GPU)."""
import numpy as np

import research_amd as ra
from research_amd import autotune, predictor, workloads


def test_panels_by_products_cover_rows_in_order():
    A = ra.CSR.rmat(12, 16 << 12, seed=1)
    rl = np.diff(A.row_offsets.astype(np.int64))
    ro = A.row_offsets.astype(np.int64)
    cs = np.concatenate(([0], np.cumsum(rl[A.col_ids.astype(np.int64)])))
    prod = cs[ro[1:]] - cs[ro[:-1]]
    for limit in (1.0, 5e5, 2e6, 1e12):
        p = predictor.panels_by_products(A, limit)
        assert p[0][0] == 0 and p[-1][1] == A.rows
        assert all(a[1] == b[0] for a, b in zip(p, p[1:]))
        for r0, r1 in p:  # within the limit unless a single row exceeds it
            assert r1 > r0 and (prod[r0:r1].sum() <= limit or r1 - r0 == 1)


def test_batched_products_deterministic():
    a = workloads.batched_products(ra, 9, seed=4)
    b = workloads.batched_products(ra, 9, seed=4)
    assert [m.rows for m in a] == [m.rows for m in b]
    for x, y in zip(a, b):
        assert np.array_equal(x.col_ids, y.col_ids) and np.array_equal(x.values, y.values)
        assert x.values.dtype == np.float32 and 500 <= x.rows <= 65536
    assert workloads.product_count(a[0]) == int(
        np.diff(a[0].row_offsets.astype(np.int64))[a[0].col_ids.astype(np.int64)].sum())


def test_autotune_space_uses_config_keys():
    c = ra.Config()
    for knob, values in autotune.SPACE.items():
        assert values[0] == getattr(c, knob), knob  # the first value is the default
        for v in values:
            setattr(c, knob, v)
            assert getattr(c, knob) == v
