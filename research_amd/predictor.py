"""Training and evaluation of the learned row-size predictor.

The predictor (k_mlp.hip, oc_mlp_* in include/ocean_capi.h) is trained on
the GPU by its own forward/backward/Adam kernels, on features and exact
per-row output counts of C = A*A produced by the library itself: features
by oc_mlp_features, counts by the plain symbolic pass (oc_symbolic).

    python -m research_amd.predictor train [--out MODEL] [--matrices 200]
    python -m research_amd.predictor eval  [--model MODEL] [--log-n 21 --degree 32]

The training set is 200 seeded matrices: R-MAT (a, b, c = .57, .19, .19),
Erdos-Renyi and banded, n = 2^14 .. 2^18, average degree 4 .. 64.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

from . import ocean

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_MODEL = os.path.join(HERE, "models", "rowsize_mlp.bin")


def training_matrix(i: int, seed: int = 2024):
    """The i-th training matrix (type, n, degree drawn from a seeded stream)."""
    rng = np.random.default_rng([seed, i])
    kind = i % 3
    log_n = int(rng.integers(14, 19))
    n = 1 << log_n
    deg = float(np.exp(rng.uniform(np.log(4), np.log(64))))
    s = seed * 1000 + i
    if kind == 0:
        return "rmat", ocean.CSR.rmat(log_n, int(deg * n), 0.57, 0.19, 0.19, vtype=1, seed=s)
    if kind == 1:
        return "er", ocean.CSR.erdos_renyi(n, deg, vtype=1, seed=s)
    return "banded", ocean.CSR.banded(n, max(3, int(deg) | 1), vtype=1, seed=s)


def exact_counts(dA, dB) -> ocean.DeviceBuffer:
    nz = ocean.DeviceBuffer(4 * max(dA.rows, 1))
    ocean._check(ocean.lib.oc_symbolic(None, C.byref(dA.m), C.byref(dB.m), nz.ptr, None))
    return nz


def build_dataset(count: int, rows_per_matrix: int, seed: int = 2024, log=print):
    """Features and labels of a sample of rows of each training matrix
    (host arrays; rows drawn uniformly, the heaviest rows always kept)."""
    feats, labels = [], []
    for i in range(count):
        kind, A = training_matrix(i, seed)
        dA = A.upload()
        f = ocean.Mlp.features(dA, dA).to_array(np.float32, A.rows * ocean.MLP_FEATURES)
        f = f.reshape(A.rows, ocean.MLP_FEATURES)
        nz = exact_counts(dA, dA)
        lab = ocean.Mlp.labels(nz, A.rows).to_array(np.uint8, A.rows)
        rng = np.random.default_rng([seed, i, 1])
        take = rng.choice(A.rows, size=min(rows_per_matrix, A.rows), replace=False)
        heavy = np.argsort(f[:, 0])[-256:]
        idx = np.unique(np.concatenate([take, heavy]))
        feats.append(f[idx])
        labels.append(lab[idx])
        if log and (i % 20 == 0 or i == count - 1):
            log(f"[predictor] matrix {i}: {kind} n={A.rows} nnz={A.nnz}", flush=True)
        dA.free()
    X = np.concatenate(feats)
    y = np.concatenate(labels)
    perm = np.random.default_rng(seed).permutation(len(y))
    return X[perm], y[perm]


def train(out: str, count: int = 200, rows_per_matrix: int = 16384, epochs: int = 20,
          batch: int = 8192, lr: float = 2e-3, seed: int = 2024, log=print) -> dict:
    t0 = time.time()
    X, y = build_dataset(count, rows_per_matrix, seed, log)
    n = len(y)
    hold = n // 10  # the last tenth is held out
    m = ocean.Mlp(seed)
    dX = ocean.DeviceBuffer.from_array(np.ascontiguousarray(X))
    dy = ocean.DeviceBuffer.from_array(np.ascontiguousarray(y))
    losses = m.train(dX, dy, n - hold, batch, epochs, lr)
    logits = m.forward(dX, n)
    pred = logits.argmax(1)
    acc_train = float((pred[:n - hold] == y[:n - hold]).mean())
    acc_hold = float((pred[n - hold:] == y[n - hold:]).mean())
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    m.save(out)
    info = {"rows": int(n), "held_out": int(hold), "epochs": epochs, "batch": batch, "lr": lr,
            "loss_first_epoch": float(losses[0]), "loss_last_epoch": float(losses[-1]),
            "accuracy_train": acc_train, "accuracy_held_out": acc_hold,
            "class_histogram": np.bincount(y, minlength=ocean.MLP_CLASSES).tolist(),
            "seconds": round(time.time() - t0, 1), "model": os.path.relpath(out)}
    if log:
        log("[predictor] " + json.dumps(info), flush=True)
    return info


def _panel(dM: "ocean.DeviceCSR", r0: int, r1: int, ro: np.ndarray) -> "ocean.DeviceCSR":
    """Rows [r0, r1) of a device CSR as a borrowed view (same entry arrays)."""
    v = ocean.DeviceCSR()
    v.m.rows, v.m.cols = r1 - r0, dM.m.cols
    v.m.nnz = int(ro[r1] - ro[r0])
    v.m.vtype, v.m.off_bytes = dM.m.vtype, 4
    v.m.row_offsets = dM.m.row_offsets + 4 * r0
    v.m.col_ids, v.m.values = dM.m.col_ids, dM.m.values
    v.m.on_device, v.m.owned = 1, 0
    return v


def panels_by_products(A, limit: float):
    """Row panels [r0, r1) whose products sum to at most `limit` (a panel
    holds at least one row)."""
    ro = A.row_offsets.astype(np.int64)
    rl = np.diff(ro)
    cs = np.concatenate(([0], np.cumsum(rl[A.col_ids.astype(np.int64)])))
    prod = cs[ro[1:]] - cs[ro[:-1]]
    cum = np.concatenate(([0], np.cumsum(prod)))
    out, r0 = [], 0
    while r0 < A.rows:
        r1 = int(np.searchsorted(cum, cum[r0] + limit, side="right")) - 1
        r1 = min(A.rows, max(r0 + 1, r1))
        out.append((r0, r1))
        r0 = r1
    return out


def evaluate(model: str, log_n: int = 21, degree: int = 32, panel_products: float = 4e9,
             quantile: float = 0.99, log=print) -> dict:
    """Held-out R-MAT product C = A*A in row panels (C of the whole product
    exceeds one GPU's memory), learned capacities against the plain
    HyperLogLog-sampled path: time per panel, overflow rows and capacity
    totals, and the same output (per-row column-set hashes and row counts)
    from both."""
    A = ocean.CSR.rmat(log_n, degree << log_n, 0.57, 0.19, 0.19, vtype=0, seed=7)
    dA = A.upload()
    ro = A.row_offsets.astype(np.int64)
    m = ocean.Mlp(0)
    m.load(model)
    plain = ocean.SpGEMM(ocean.Config(track_stage_times=1))
    learned = ocean.SpGEMM(ocean.Config(track_stage_times=1))
    learned.set_model(m, quantile)
    t_plain = t_learned = 0.0
    # output allocation per path: device memory freed and allocated again is
    # cleared by the driver first, which can take hundreds of ms for a
    # panel's output whatever the path; reported apart from the totals
    a_plain = a_learned = 0.0
    ovf = est_total = nnz_total = products = 0
    same = True
    panels = panels_by_products(A, panel_products)
    for i, (r0, r1) in enumerate(panels):
        P = _panel(dA, r0, r1, ro)
        if log:
            log(f"[predictor] panel {i}: rows {r0}..{r1}, plain path", flush=True)
        C1 = plain.run(P, dA)
        t_plain += plain.timer_ms("total")
        a_plain += max(0.0, plain.timer_ms("numeric_malloc"))
        h1, o1 = C1.row_set_hash(), C1.row_offsets()
        st1 = plain.get_stats()
        C1.free()
        if log:
            log(f"[predictor] panel {i}: learned path", flush=True)
        C2 = learned.run(P, dA)
        t_learned += learned.timer_ms("total")
        a_learned += max(0.0, learned.timer_ms("numeric_malloc"))
        st2 = learned.get_stats()
        same = same and np.array_equal(h1, C2.row_set_hash()) and np.array_equal(o1, C2.row_offsets())
        ovf += st2.get("estimation", {}).get("overflow_rows", 0)
        est_total += st2.get("estimation", {}).get("total_est", 0)
        nnz_total += int(st2["C"]["nnz"])
        products += int(st1["stats"]["intermediate_product"])
        C2.free()
        if log:
            log(f"[predictor] rows {r0}..{r1}: plain {plain.timer_ms('total'):.1f} ms "
                f"(workflow {st1['decision']['workflow']}), learned "
                f"{learned.timer_ms('total'):.1f} ms", flush=True)
    res = {"matrix": f"R-MAT n=2^{log_n}, degree {degree}, f64, seed 7, C = A*A in "
                     f"{len(panels)} row panels of <= {panel_products:.0e} products",
           "products": products, "nnz_c": nnz_total, "plain_ms": round(t_plain, 2),
           "learned_ms": round(t_learned, 2), "speedup": round(t_plain / t_learned, 3),
           "plain_output_alloc_ms": round(a_plain, 2),
           "learned_output_alloc_ms": round(a_learned, 2),
           "speedup_without_output_alloc":
               round((t_plain - a_plain) / max(t_learned - a_learned, 1e-9), 3),
           "plain_gflops": round(2 * products / t_plain / 1e6, 2),
           "learned_gflops": round(2 * products / t_learned / 1e6, 2),
           "overflow_rows": int(ovf), "capacity_over_nnz": round(est_total / max(nnz_total, 1), 3),
           "identical_structure": bool(same), "quantile": quantile}
    if log:
        log("[predictor] " + json.dumps(res), flush=True)
    return res


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m research_amd.predictor")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--out", default=DEFAULT_MODEL)
    t.add_argument("--matrices", type=int, default=200)
    t.add_argument("--rows-per-matrix", type=int, default=16384)
    t.add_argument("--epochs", type=int, default=20)
    e = sub.add_parser("eval")
    e.add_argument("--model", default=DEFAULT_MODEL)
    e.add_argument("--log-n", type=int, default=21)
    e.add_argument("--degree", type=int, default=32)
    e.add_argument("--panel-products", type=float, default=4e9)
    e.add_argument("--quantile", type=float, default=0.99)
    a = ap.parse_args(argv)
    if a.cmd == "train":
        info = train(a.out, a.matrices, a.rows_per_matrix, a.epochs)
    else:
        info = evaluate(a.model, a.log_n, a.degree, a.panel_products, a.quantile)
    print(json.dumps(info), flush=True)


if __name__ == "__main__":
    sys.exit(main())
