"""Oracle pinning: FP64 Gustavson against closed forms, scipy and brute force."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as ora
import research_amd as ra


def _from_scipy(M, dtype=np.float64):
    M = sp.csr_matrix(M)
    M.sort_indices()
    return ra.CSR(M.shape[0], M.shape[1], M.indptr.astype(np.uint32),
                  M.indices.astype(np.uint32), M.data.astype(dtype))


def test_laplacian_closed_form():
    # 5-point stencil on 256x256: products and nnz(C) in closed form
    n = 256
    A = ra.CSR.laplacian2d(n, n, seed=0)
    assert A.nnz == 5 * n * n - 4 * n
    an = ora.analysis(A, A)
    assert int(an["products"].sum()) == 1629192
    nnz, _ = ora.symbolic(A, A)
    # C = A^2 has the 13-point pattern: 65536 diagonal + 4*255*256 (+-1,+-n)
    # + 4*254*256 (+-2, +-2n) + 4*255^2 (diagonal neighbours)
    assert int(nnz.sum()) == 65536 + 4 * 255 * 256 + 4 * 254 * 256 + 4 * 255 * 255 == 846852


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_against_scipy(seed):
    rng = np.random.default_rng(seed)
    A = sp.random(300, 200, density=0.05, random_state=rng, format="csr")
    B = sp.random(200, 250, density=0.05, random_state=rng, format="csr")
    A.data = rng.uniform(-1, 1, A.nnz)
    B.data = rng.uniform(-1, 1, B.nnz)
    C = (A @ B).tocsr()
    C.sort_indices()
    rp, cols, vals, cond = ora.numeric(_from_scipy(A), _from_scipy(B))
    # random continuous values: no exact cancellation, so scipy (which drops
    # exact zeros) has the same pattern
    assert np.array_equal(rp, C.indptr.astype(np.uint64))
    assert np.array_equal(cols, C.indices.astype(np.uint32))
    assert np.max(np.abs(vals - C.data) / cond) < 1e-14


def test_identity_and_permutation():
    rng = np.random.default_rng(3)
    A = sp.random(120, 120, density=0.08, random_state=rng, format="csr")
    I = sp.identity(120, format="csr")
    rp, cols, vals, _ = ora.numeric(_from_scipy(A), _from_scipy(I))
    A.sort_indices()
    assert np.array_equal(cols, A.indices.astype(np.uint32)) and np.array_equal(vals, A.data)
    P = sp.csr_matrix((np.ones(120), (np.arange(120), rng.permutation(120))), shape=(120, 120))
    rp, cols, vals, _ = ora.numeric(_from_scipy(P), _from_scipy(A))
    ref = (P @ A).tocsr()
    ref.sort_indices()
    assert np.array_equal(cols, ref.indices.astype(np.uint32)) and np.array_equal(vals, ref.data)


def test_cancellation_keeps_structure():
    # A row 0 = [1, 1]; B row 0 has +1 at column 3, B row 1 has -1 at column 3:
    # the product is an explicit zero at (0, 3), which stays in the pattern.
    A = _from_scipy(sp.csr_matrix(np.array([[1.0, 1.0]])))
    B = _from_scipy(sp.csr_matrix(np.array([[0, 0, 0, 1.0, 2.0], [0, 0, 0, -1.0, 0]])))
    rp, cols, vals, cond = ora.numeric(A, B)
    assert list(cols) == [3, 4] and vals[0] == 0.0 and vals[1] == 2.0 and cond[0] == 2.0


def test_brute_force_tiny():
    rng = np.random.default_rng(4)
    for _ in range(20):
        m, k, n = rng.integers(1, 12, size=3)
        Ad = rng.uniform(-1, 1, (m, k)) * (rng.random((m, k)) < 0.4)
        Bd = rng.uniform(-1, 1, (k, n)) * (rng.random((k, n)) < 0.4)
        rp, cols, vals, _ = ora.numeric(_from_scipy(Ad), _from_scipy(Bd))
        struct = (np.abs(Ad) > 0).astype(int) @ (np.abs(Bd) > 0).astype(int) > 0
        dense = Ad @ Bd
        r = 0
        for i in range(m):
            row = np.nonzero(struct[i])[0]
            assert list(cols[rp[i]:rp[i + 1]]) == list(row)
            assert np.allclose(vals[rp[i]:rp[i + 1]], dense[i, row], atol=1e-14)


def test_plain_order_matches_sequential_sum():
    # plain mode sums products in A-entry-then-B-entry order starting at 0.0
    A = _from_scipy(sp.csr_matrix(np.array([[0.1, 0.2, 0.1]])))
    B = _from_scipy(sp.csr_matrix(np.array([[1e16], [1.0], [-1e16]])))
    rp, cols, vals, _ = ora.numeric(A, B, compensated=0)
    assert vals[0] == ((0.0 + 0.1 * 1e16) + 0.2 * 1.0) + 0.1 * -1e16 == 0.25
    rp, cols, vals_c, _ = ora.numeric(A, B, compensated=1)
    assert vals_c[0] == 0.2  # the compensated sum recovers the small term


def test_float32_inputs_promoted_exactly():
    rng = np.random.default_rng(5)
    A = sp.random(50, 60, density=0.2, random_state=rng, format="csr", dtype=np.float32)
    B = sp.random(60, 40, density=0.2, random_state=rng, format="csr", dtype=np.float32)
    rp, cols, vals, _ = ora.numeric(_from_scipy(A, np.float32), _from_scipy(B, np.float32))
    ref = (A.astype(np.float64) @ B.astype(np.float64)).tocsr()
    ref.sort_indices()
    assert np.allclose(vals, ref.data, rtol=1e-14)


def test_symbolic_set_hash_and_rows():
    A = ra.CSR.rmat(10, 8 << 10, seed=2)
    nnz, h = ora.symbolic(A, A)
    rp, cols, _, _ = ora.numeric(A, A)
    for r in range(0, A.rows, 37):
        seg = cols[rp[r]:rp[r + 1]]
        assert len(seg) == nnz[r]
        assert int(h[r]) == sum(ora.mix64(int(c)) for c in seg) % (1 << 64)
    rows = np.array([5, 0, 900, 5], np.uint32)
    rrp, rcols, rvals, _ = ora.numeric_rows(A, A, rows)
    for i, r in enumerate(rows):
        assert np.array_equal(rcols[rrp[i]:rrp[i + 1]], cols[rp[r]:rp[r + 1]])


def test_baseline_spgemm_matches_oracle():
    """The CPU baseline's nnz and value sum equal the oracle's plain
    Gustavson on the same rows (dense rows take the range scan, sparse rows
    the sort)."""
    import research_amd as ra
    for A, r0, r1 in ((ra.CSR.rmat(12, 16 << 12, seed=1), 0, None),
                      (ra.CSR.banded(3000, 65, seed=2), 100, 2900),
                      (ra.CSR.erdos_renyi(5000, 6.0, seed=3), 7, 4000)):
        n, cs = ora.baseline_spgemm(A, A, r0, r1, nthreads=4)
        rp, cols, vals, _ = ora.numeric(A, A, r0, r1, compensated=0, with_cond=False)
        assert n == len(cols)
        assert abs(cs - vals.sum()) <= 1e-9 * max(1.0, np.abs(vals).sum())
