"""GPU parity of the learned row-size predictor against its CPU oracle
(oracle/mlp.py), and of SpGEMM run on the predicted capacities."""
import numpy as np
import pytest

from helpers import check_product, log
from oracle import mlp as omlp
from oracle import oracle as ora

pytestmark = pytest.mark.gpu


def _rand_params(seed):
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0, np.sqrt(2 / 16), (16, 64))
    W2 = rng.normal(0, np.sqrt(2 / 64), (64, 64))
    W3 = rng.normal(0, np.sqrt(1 / 64), (64, 8))
    p = omlp.pack(W1, rng.normal(0, .1, 64), W2, rng.normal(0, .1, 64), W3, rng.normal(0, .1, 8))
    return p.astype(np.float32)


@pytest.mark.parametrize("kind", ["rmat", "er", "rect"])
def test_features_and_labels(gpu, kind):
    log(f"case: mlp features/labels {kind}")
    if kind == "rmat":
        A = B = gpu.CSR.rmat(12, 16 << 12, seed=1)
    elif kind == "er":
        A = B = gpu.CSR.erdos_renyi(3000, 12.0, seed=2)
    else:
        A = gpu.CSR.rect_band(2000, 30000, 8.0, 0.3, 4096, 8, vtype=1, seed=3)
        B = A.transpose()
    dA, dB = A.upload(), B.upload()
    feat = gpu.Mlp.features(dA, dB).to_array(np.float32, A.rows * 16).reshape(A.rows, 16)
    want = omlp.features(A, B)
    err = np.abs(feat - want)
    log(f"features {kind}: max abs error {err.max():.2e}")
    assert np.all(err <= 2e-6 + 2e-6 * np.abs(want))
    nnz, _ = ora.symbolic(A, B)
    dn = gpu.DeviceBuffer.from_array(nnz.astype(np.uint32))
    lab = gpu.Mlp.labels(dn, A.rows).to_array(np.uint8, A.rows)
    assert np.array_equal(lab, omlp.labels(nnz, gpu.Mlp.thresholds()))


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.uniform(0, 1.5, (n, 16)).astype(np.float32)
    y = rng.integers(0, 8, n).astype(np.uint8)
    return X, y


@pytest.mark.parametrize("n", [1, 64, 1000, 65536, 200001])
def test_forward_matches_oracle(gpu, n):
    log(f"case: mlp forward n={n}")
    p = _rand_params(5)
    m = gpu.Mlp(0)
    m.set_params(p)
    X, _ = _data(n, 6)
    z = m.forward(gpu.DeviceBuffer.from_array(X), n)
    want, _ = omlp.forward(p.astype(np.float64), X.astype(np.float64), omlp.bf16)
    d = np.abs(z - want)
    log(f"forward n={n}: max |dz| {d.max():.2e}, median {np.median(d):.2e}")
    assert np.all(d <= 0.02 + 0.01 * np.abs(want))
    assert np.median(d) <= 1e-3


@pytest.mark.parametrize("n", [100, 4096, 65536])
def test_backward_matches_oracle(gpu, n):
    log(f"case: mlp backward n={n}")
    p = _rand_params(7)
    m = gpu.Mlp(0)
    m.set_params(p)
    X, y = _data(n, 8)
    g, loss = m.backward(gpu.DeviceBuffer.from_array(X), gpu.DeviceBuffer.from_array(y), n)
    wl, wg = omlp.loss_and_grads(p.astype(np.float64), X.astype(np.float64), y.astype(np.int64),
                                 omlp.bf16)
    assert abs(loss - wl) <= 1e-4 * abs(wl)
    for k, e in (("W1", "b1"), ("b1", "W2"), ("W2", "b2"), ("b2", "W3"), ("W3", "b3"),
                 ("b3", None)):
        a, b = omlp.OFF[k], omlp.OFF[e] if e else omlp.NPARAM
        rel = np.linalg.norm(g[a:b] - wg[a:b]) / max(np.linalg.norm(wg[a:b]), 1e-30)
        log(f"backward n={n} {k}: relative error {rel:.2e}")
        assert rel <= 2e-2, k
    # the exact (unrounded) gradient agrees to the BF16 level
    _, eg = omlp.loss_and_grads(p.astype(np.float64), X.astype(np.float64), y.astype(np.int64),
                                omlp.ident)
    assert np.linalg.norm(g - eg) <= 0.05 * np.linalg.norm(eg)


def test_adam_step_matches_oracle(gpu):
    log("case: mlp adam step")
    p = _rand_params(9)
    m = gpu.Mlp(0)
    m.set_params(p)
    X, y = _data(3000, 10)
    g, _ = m.backward(gpu.DeviceBuffer.from_array(X), gpu.DeviceBuffer.from_array(y), 3000)
    m.step(1e-3)
    got = m.params()
    want, _, _ = omlp.adam(p.astype(np.float64), 0.0, 0.0, g.astype(np.float64), 1, 1e-3)
    assert np.allclose(got, want, rtol=1e-5, atol=1e-7)


def _training_set(gpu):
    mats = [gpu.CSR.rmat(12, 8 << 12, seed=s) for s in (21, 22)] + \
           [gpu.CSR.rmat(13, 24 << 13, seed=23)] + \
           [gpu.CSR.erdos_renyi(6000, d, seed=24 + i) for i, d in enumerate((6.0, 20.0, 45.0))] + \
           [gpu.CSR.banded(8000, bw, seed=30 + bw) for bw in (9, 65, 301)]
    feats, labs = [], []
    for A in mats:
        dA = A.upload()
        feats.append(gpu.Mlp.features(dA, dA).to_array(np.float32, A.rows * 16).reshape(-1, 16))
        nnz, _ = ora.symbolic(A, A)
        labs.append(omlp.labels(nnz, gpu.Mlp.thresholds()))
    X = np.concatenate(feats)
    y = np.concatenate(labs)
    perm = np.random.default_rng(0).permutation(len(y))
    return X[perm], y[perm]


def test_training_and_learned_spgemm(gpu):
    log("case: mlp training on 9 matrices, then learned SpGEMM on held-out rmat 2^13")
    X, y = _training_set(gpu)
    n = len(y)
    m = gpu.Mlp(1)
    dX, dy = gpu.DeviceBuffer.from_array(X), gpu.DeviceBuffer.from_array(y)
    losses = m.train(dX, dy, n, 4096, 12, 3e-3)
    logits = m.forward(dX, n)
    acc = float((logits.argmax(1) == y).mean())
    majority = float(np.bincount(y, minlength=8).max() / n)
    log(f"training: loss {losses[0]:.3f} -> {losses[-1]:.3f}, accuracy {acc:.3f} "
        f"(majority class {majority:.3f})")
    assert losses[-1] < 0.6 * losses[0]
    assert acc > majority + 0.05
    # held-out product on the predicted capacities: output equals the oracle's
    A = gpu.CSR.rmat(13, 16 << 13, seed=41)
    dA = A.upload()
    est = m.estimates(dA, dA, 0.99).to_array(np.uint32, A.rows)
    exact, _ = ora.symbolic(A, A)
    # learned_min_compression 0: the estimation workflow on any capacities
    sp = gpu.SpGEMM(gpu.Config(use_ultrasparse_workflow=0, learned_min_compression=0.0))
    sp.set_model(m, 0.99)
    dC = sp.run(dA, dA)
    st = sp.get_stats()
    assert st["decision"]["learned_estimates"] and st["decision"]["workflow"] == 1
    assert st["decision"]["learned_total_capacity"] == int(est.astype(np.uint64).sum())
    assert dC.validate() == 0
    check_product(A, A, dC.download(), 1e-12, "learned rmat13")
    under = int(np.count_nonzero(exact > est))
    log(f"learned: capacities below the exact count for {under} rows, "
        f"overflow rows {st['estimation']['overflow_rows']}, "
        f"sum capacity / sum nnz {est.sum() / max(exact.sum(), 1):.2f}")
    assert st["estimation"]["overflow_rows"] <= under + 0
    # default rule: estimation only when products >= 1.5 x the capacities
    sp2 = gpu.SpGEMM(gpu.Config(use_ultrasparse_workflow=0))
    sp2.set_model(m, 0.99)
    dC2 = sp2.run(dA, dA)
    st2 = sp2.get_stats()
    ratio = st2["stats"]["intermediate_product"] / st2["decision"]["learned_total_capacity"]
    assert abs(ratio - st2["decision"]["learned_products_per_capacity"]) < 1e-9 * ratio
    assert (st2["decision"]["workflow"] == 1) == (ratio >= 1.5)
    assert st2["decision"]["learned_estimates"] == (ratio >= 1.5)
    log(f"learned rule: products / capacity {ratio:.3f} -> workflow {st2['decision']['workflow']}")
    if st2["decision"]["workflow"] == 0:
        # exact counts picked by the predictor: no HyperLogLog sampling
        assert st2["sampling"]["num_samples"] == 0
    check_product(A, A, dC2.download(), 1e-12, "learned rmat13, default rule")


def test_learned_gain_on_batch(gpu):
    """What the trained network gains: one warmed Batch handle over 256
    mixed products (the batched workload's generator), timed with the
    HyperLogLog-sampled path and then with the trained model's capacities;
    both give the same structure, and the learned run is faster."""
    import os
    import time
    from research_amd import predictor
    from research_amd.workloads import batched_products
    import research_amd as ra
    log("case: learned gain on a batch of 256 mixed products")
    mats = batched_products(ra, 256)
    dev = [M.upload() for M in mats]
    b = gpu.Batch(gpu.Config())
    b.set_inputs(dev)

    def timed(reps=5):
        b.run()
        gpu._check(gpu.lib.oc_device_synchronize())
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            b.run()
            gpu._check(gpu.lib.oc_device_synchronize())
            ts.append(time.perf_counter() - t0)
        view, _, _ = b.result()
        return float(np.median(ts)), view.row_set_hash(), b.get_stats()

    t_hll, h_hll, st_hll = timed()
    m = gpu.Mlp(0)
    assert os.path.exists(predictor.DEFAULT_MODEL)
    m.load(predictor.DEFAULT_MODEL)
    gpu._check(gpu.lib.oc_spgemm_set_model(gpu.lib.oc_batch_handle(b.h), m.h, 0.99))
    t_lrn, h_lrn, st_lrn = timed()
    assert np.array_equal(h_hll, h_lrn), "learned capacities changed the structure"
    gain = t_hll / t_lrn
    log(f"learned gain: HyperLogLog {t_hll * 1e3:.2f} ms (workflow "
        f"{st_hll['decision']['workflow']}), learned {t_lrn * 1e3:.2f} ms (workflow "
        f"{st_lrn['decision']['workflow']}), factor {gain:.3f}")
    assert gain > 1.0
