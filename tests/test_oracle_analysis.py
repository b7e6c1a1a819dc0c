"""Oracle pinning: per-row analysis against an independent brute force."""
import numpy as np
import pytest

from oracle import oracle as ora


def _rand_csr(rng, m, n, max_len, empty_frac=0.2):
    class M:
        pass
    rows = []
    for _ in range(m):
        k = 0 if rng.random() < empty_frac else int(rng.integers(1, max_len + 1))
        rows.append(np.sort(rng.choice(n, size=min(k, n), replace=False)).astype(np.uint32))
    x = M()
    x.rows, x.cols = m, n
    x.row_offsets = np.zeros(m + 1, np.uint32)
    x.row_offsets[1:] = np.cumsum([len(r) for r in rows])
    x.col_ids = np.concatenate(rows) if rows else np.zeros(0, np.uint32)
    return x, rows


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_analysis_brute_force(seed):
    rng = np.random.default_rng(seed)
    B, brows = _rand_csr(rng, 60, 500, 40)
    A, arows = _rand_csr(rng, 45, 60, 12)
    got = ora.analysis(A, B)
    for r, ar in enumerate(arows):
        lens = [len(brows[k]) for k in ar]
        nonempty = [brows[k] for k in ar if len(brows[k])]
        assert got["products"][r] == sum(lens)
        assert got["max_b"][r] == (max(lens) if lens else 0)
        assert got["avg_b"][r] == (sum(lens) // len(ar) if len(ar) else 0)
        # empty rows keep the sentinels left = 0xFFFFFFFF, right = 0
        assert got["left"][r] == (min(int(b[0]) for b in nonempty) if nonempty else 0xFFFFFFFF)
        assert got["right"][r] == (max(int(b[-1]) for b in nonempty) if nonempty else 0)
    assert got["max_a_nnz"] == max(len(a) for a in arows)
