"""Oracle pinning: the learned row-size predictor's CPU restatement."""
import math

import numpy as np
import pytest

from oracle import mlp
from oracle import oracle as ora


def test_bf16_round_nearest_even():
    cases = {
        1.0: 1.0,
        1.0 + 2.0 ** -8: 1.0,                       # tie -> even mantissa (0)
        1.0 + 3 * 2.0 ** -8: 1.0 + 2.0 ** -6,       # tie -> even mantissa (2)
        1.0 + 2.0 ** -8 + 2.0 ** -20: 1.0 + 2.0 ** -7,  # above the tie
        -2.5: -2.5,
        3.0e38: 1.765625 * 2.0 ** 127,              # 1.7632 * 2^127: 225.7/128 -> 226/128
        0.0: 0.0,
    }
    for x, want in cases.items():
        got = float(mlp.bf16(np.float32(x)).reshape(-1)[0])
        assert got == pytest.approx(want, rel=1e-7, abs=0), x
    # every bf16 value is a fixed point
    u = np.arange(0, 1 << 16, 97, dtype=np.uint32) << 16
    v = u.view(np.float32)
    v = v[np.isfinite(v)]
    assert np.array_equal(mlp.bf16(v), v.astype(np.float64))


def _rand_params(seed):
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0, math.sqrt(2 / 16), (16, 64))
    W2 = rng.normal(0, math.sqrt(2 / 64), (64, 64))
    W3 = rng.normal(0, math.sqrt(1 / 64), (64, 8))
    return mlp.pack(W1, rng.normal(0, .1, 64), W2, rng.normal(0, .1, 64), W3, rng.normal(0, .1, 8))


def test_exact_gradients_finite_differences():
    rng = np.random.default_rng(3)
    p = _rand_params(3)
    X = rng.normal(0, 1, (7, 16))
    y = rng.integers(0, 8, 7)
    loss, g = mlp.loss_and_grads(p, X, y, rnd=mlp.ident)
    ends = {"W1": "b1", "b1": "W2", "W2": "b2", "b2": "W3", "W3": "b3"}
    idx = np.concatenate([rng.choice(np.arange(mlp.OFF[k], mlp.OFF[ends[k]] if k in ends else mlp.NPARAM),
                                     6, replace=False) for k in ("W1", "b1", "W2", "b2", "W3", "b3")])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mlp.loss_and_grads(pp, X, y, rnd=mlp.ident)[0] -
              mlp.loss_and_grads(pm, X, y, rnd=mlp.ident)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), i


def test_emulated_close_to_exact():
    rng = np.random.default_rng(4)
    p = _rand_params(4)
    X = rng.uniform(0, 1.5, (256, 16))
    y = rng.integers(0, 8, 256)
    l0, g0 = mlp.loss_and_grads(p, X, y, rnd=mlp.ident)
    l1, g1 = mlp.loss_and_grads(p, X, y, rnd=mlp.bf16)
    assert abs(l0 - l1) < 0.02 * abs(l0)
    assert np.linalg.norm(g0 - g1) < 0.05 * np.linalg.norm(g0)


def test_adam_closed_form():
    p = np.array([1.0, -2.0])
    g = np.array([0.5, -0.25])
    p1, m1, m2 = mlp.adam(p, np.zeros(2), np.zeros(2), g, 1, 0.1)
    # first step moves every parameter by lr * sign(g) (up to eps)
    assert np.allclose(p1, p - 0.1 * np.sign(g), atol=1e-7)
    assert np.allclose(m1, 0.1 * g) and np.allclose(m2, 0.001 * g * g)


class _M:
    pass


def _csr(rows, ncols):
    m = _M()
    m.rows, m.cols = len(rows), ncols
    m.row_offsets = np.zeros(len(rows) + 1, np.uint32)
    m.row_offsets[1:] = np.cumsum([len(r) for r in rows])
    m.col_ids = np.array([c for r in rows for c in r], np.uint32)
    return m


def test_features_brute_force():
    B = _csr([[0, 5, 9], [], [2], [1, 2, 3, 4, 5, 6, 7]], 10)
    A = _csr([[0, 1, 3], [], [2, 2], [1]], 4)
    f = mlp.features(A, B)
    # row 0: B rows 0 (3), 1 (empty), 3 (7): P = 10, nA = 3, mb = 7, avg = 10 // 3
    P, nA, mb, ab, lo, hi = 10, 3, 7, 3, 0, 9
    span = hi - lo + 1
    lg = lambda v: math.log2(1 + v) / 16  # noqa: E731
    ub, uni = min(P, span), span * (1 - math.exp(-P / span))
    want = [lg(P), lg(nA), lg(mb), lg(ab), lg(span), lg(ub), lg(uni), lg(P) - lg(mb),
            P / (nA * mb), min(P / span, 4) / 4, mb / span, lg(ub) - lg(mb),
            lg(uni) - lg(mb), lg(ub) - lg(uni), 0.0, 1.0]
    assert np.allclose(f[0], want, rtol=1e-12, atol=1e-15)
    # empty A row: span 0 (sentinels), everything but the bias is 0
    assert np.allclose(f[1, :15], 0.0) and f[1, 15] == 1.0
    # row 2 references B row 2 twice: P = 2, span 1, min(P, span) = 1
    assert f[2, 4] == pytest.approx(lg(1)) and f[2, 5] == pytest.approx(lg(1))


def test_labels_and_capacities():
    thr = [200, 488, 1064, 2216, 4776, 9640, 65536, 0xFFFFFFFF]
    nnz = np.array([0, 200, 201, 9640, 9641, 65536, 65537, 10 ** 9])
    assert list(mlp.labels(nnz, thr)) == [0, 0, 1, 5, 6, 6, 7, 7]
    # certain logits: class 1 for row 0, class 7 for row 1
    lg = np.full((3, 8), -50.0)
    lg[0, 1] = lg[1, 7] = lg[2, 0] = 50.0
    est, cls = mlp.capacities(lg, thr, products=[1000, 90000, 50], max_b=[10, 10, 70],
                              left=[0, 0, 0], right=[99999, 99999, 99], q=0.99)
    assert list(cls) == [1, 7, 0]
    # row 0: threshold 488; row 1: min(products, span) = 90000; row 2: threshold
    # 200 raised to max_b 70 is 200, clipped to min(50, 100) = 50
    assert list(est) == [488, 90000, 50]
