"""Synthetic workloads shared by bench.py and the autotuner."""
from __future__ import annotations

import numpy as np


def batched_products(ra, count: int, seed: int = 4):
    """The batched configuration: `count` independent products C_i = A_i*A_i,
    n log-uniform in [500, 50000], type cycling Erdos-Renyi (degree 2..32),
    banded (bandwidth 3..65) and R-MAT (degree 8, n rounded to 2^k), FP32
    values U[-1,1)."""
    rng = np.random.default_rng(seed)
    mats = []
    for i in range(count):
        n = int(np.exp(rng.uniform(np.log(500), np.log(50000))))
        s = seed * 1_000_000 + i
        if i % 3 == 0:
            M = ra.CSR.erdos_renyi(n, float(rng.integers(2, 33)), vtype=1, seed=s)
        elif i % 3 == 1:
            M = ra.CSR.banded(n, int(rng.integers(1, 33)) * 2 + 1, vtype=1, seed=s)
        else:
            lg = max(9, int(round(np.log2(n))))
            M = ra.CSR.rmat(lg, 8 << lg, vtype=1, seed=s)
        mats.append(M)
    return mats


def product_count(M) -> int:
    rl = np.diff(M.row_offsets.astype(np.int64))
    return int(rl[M.col_ids.astype(np.int64)].sum())
