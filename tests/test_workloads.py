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
    """Every knob is a config key (per-bin knobs index a 7-entry array whose
    0 means the gfx950 default); the first value is the default, and every
    point of the space is a valid configuration."""
    c = ra.Config()
    for knob, values in autotune.SPACE.items():
        if isinstance(knob, str):
            assert values[0] == getattr(c, knob), knob  # the first value is the default
            for v in values:
                setattr(c, knob, v)
                assert getattr(c, knob) == v
        else:
            name, b = knob
            assert values[0] == 0 and getattr(c, name)[b] == 0, knob
            for v in values:
                cur = {k: vv[0] for k, vv in autotune.SPACE.items()}
                cur[knob] = v
                cfg = ra.Config(**autotune._config(cur))
                assert ra.ocean.lib.oc_config_check(ra.ocean.C.byref(cfg._c)) == 0, (knob, v)
                assert (getattr(cfg, name)[b] == v) or v == 0
                # the shape's tables fit its LDS (checked at handle creation)
                h = ra.ocean.lib.oc_spgemm_create(ra.ocean.C.byref(cfg._c))
                assert h, (knob, v, ra.ocean.lib.oc_last_error())
                ra.ocean.lib.oc_spgemm_destroy(h)
