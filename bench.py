#!/usr/bin/env python3
"""SpGEMM benchmark on one MI355X: C = A*A with A the R-MAT n=2^20 matrix
(16 nnz/row on average, (a,b,c,d) = (.57,.19,.19,.05), duplicates merged,
FP64 values uniform[-1,1) from seed 1, int32 indices).

A "step" is one full C = A*A through libocean.so (analysis, HLL sampling,
binning, symbolic, numeric, output allocation; the output is freed between
steps).  A and B are resident in HBM before the timed region.  The product
runs with the configuration the autotuner picked for this matrix
(research_amd/tuned/rmat20.json: a symbolic binning threshold, the same output;
--config default gives the reference's defaults).  Prints one JSON line:
  value     GFLOP/s = 2 * intermediate products / step time
  roofline  dominant kernel family, its algorithmic bytes per launch over its
            average launch time (HIP events recorded on the launch stream by
            the library during the timed steps) against 8 TB/s
  cpu_baseline  FP64 row-wise Gustavson (the oracle's, OpenMP) on a bounded
            sample of the same rows, on this box's host cores
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SpGEMM GFLOP/s + % HBM roofline; learned+batched vs plain HLL path speedup"
PEAK_HBM_GBPS = 8000.0


def workload(name: str, ra):
    if name == "rmat20":
        A = ra.CSR.rmat(20, 16 << 20, 0.57, 0.19, 0.19, vtype=0, seed=1)
        desc = "R-MAT n=2^20, 16 nnz/row avg, (a,b,c,d)=(.57,.19,.19,.05), dedup, FP64 U[-1,1) seed 1, C=A*A"
        return A, A, desc, "f64"
    if name == "laplacian":
        A = ra.CSR.laplacian2d(256, 256, vtype=0, seed=0)
        return A, A, "2D 5-point Laplacian 256x256, FP64 U[-1,1) seed 0, C=A*A", "f64"
    if name == "rect":
        A = ra.CSR.rect_band(1 << 18, 1 << 21, 8.0, 0.3, 4096, 8, vtype=1, value_dist=1, seed=2)
        return A, A.transpose(), "2^18 x 2^21 Poisson(8)/row, 30% in row*8+-4096, FP32 N(0,1) seed 2, C=A*A^T", "f32"
    raise SystemExit(f"unknown workload {name}")


def run_batched(ra, ocean, count: int, steps: int, model_path: str | None):
    """Plain path (one oc_spgemm_run per product, HyperLogLog-binned) against
    one block-diagonal batched run (HyperLogLog and learned capacities) and
    the captured graph that recomputes every product's values."""
    t0 = time.perf_counter()
    from research_amd.workloads import batched_products, product_count
    mats = batched_products(ra, count)
    products = sum(product_count(M) for M in mats)
    dev = [M.upload() for M in mats]
    gen_s = time.perf_counter() - t0
    ocean.lib.oc_device_synchronize()
    # plain: one product at a time, on a handle warmed by an untimed pass
    # over a few products (its first run allocates the handle's memory)
    sp = ra.SpGEMM(ra.Config())
    for d in dev[:8]:
        sp.run(d, d).free()
    ocean.lib.oc_device_synchronize()
    lat = []
    t0 = time.perf_counter()
    for d in dev:
        t1 = time.perf_counter()
        C = sp.run(d, d)
        C.free()
        lat.append((time.perf_counter() - t1) * 1e3)
    plain_s = time.perf_counter() - t0
    # batched run (whole path once over the block-diagonal operands), with
    # the autotuner's configuration for batches when it is there
    tuned = os.path.join(ROOT, "research_amd", "tuned", "batched1024.json")
    knobs = {}
    if os.path.exists(tuned):
        with open(tuned) as f:
            knobs = json.load(f)
    b = ra.Batch(ra.Config(**knobs))
    b.set_inputs(dev)
    b.run()
    ocean.lib.oc_device_synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.run()
    ocean.lib.oc_device_synchronize()
    batched_s = (time.perf_counter() - t0) / steps
    ref_view, _, _ = b.result()
    ref_hash = ref_view.row_set_hash()
    out = {"products_count": count, "products": products, "generation_s": round(gen_s, 1),
           "batched_config": knobs or None,
           "plain_gflops": round(2 * products / plain_s / 1e9, 3),
           "plain_latency_ms_p50": round(float(np.percentile(lat, 50)), 3),
           "plain_latency_ms_p99": round(float(np.percentile(lat, 99)), 3),
           "plain_sum_of_latencies_ms": round(float(np.sum(lat)), 3),
           "batched_gflops": round(2 * products / batched_s / 1e9, 3),
           "batched_latency_ms": round(batched_s * 1e3, 3),
           "batched_workflow": b.get_stats()["decision"]["workflow"]}
    # learned capacities on the batched run
    if model_path and os.path.exists(model_path):
        m = ra.Mlp(0)
        m.load(model_path)
        ocean._check(ocean.lib.oc_spgemm_set_model(ocean.lib.oc_batch_handle(b.h), m.h, 0.99))
        b.run()
        ocean.lib.oc_device_synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            b.run()
        ocean.lib.oc_device_synchronize()
        learned_s = (time.perf_counter() - t0) / steps
        view, _, _ = b.result()
        st = b.get_stats()
        out.update({"learned_batched_gflops": round(2 * products / learned_s / 1e9, 3),
                    "learned_batched_latency_ms": round(learned_s * 1e3, 3),
                    "learned_workflow": st["decision"]["workflow"],
                    "learned_products_per_capacity":
                        st["decision"].get("learned_products_per_capacity"),
                    "learned_overflow_rows": (st.get("estimation") or {}).get("overflow_rows"),
                    "learned_same_structure": bool(np.array_equal(view.row_set_hash(), ref_hash)),
                    "learned_batched_vs_plain": round(plain_s / learned_s, 2),
                    "learned_batched_vs_batched": round(batched_s / learned_s, 3)})
        ocean._check(ocean.lib.oc_spgemm_set_model(ocean.lib.oc_batch_handle(b.h), None, 0.99))
    # value-only graph replays of the planned batch
    b.plan()
    b.capture()
    b.launch()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.launch()
    graph_s = (time.perf_counter() - t0) / steps
    view, _, _ = b.result()
    out.update({"graph_gflops": round(2 * products / graph_s / 1e9, 3),
                "graph_latency_ms": round(graph_s * 1e3, 3),
                "graph_same_structure": bool(np.array_equal(view.row_set_hash(), ref_hash)),
                "batched_vs_plain": round(plain_s / batched_s, 2),
                "graph_vs_plain": round(plain_s / graph_s, 2)})
    # the whole path of every product in one graph (analysis to numeric,
    # structure recomputed on each launch): launch latencies over repeated
    # launches; every product of a launch completes with it
    b.capture_full()
    b.launch_full()
    full_lat = []
    for _ in range(max(steps, 5)):
        t1 = time.perf_counter()
        b.launch_full()
        full_lat.append((time.perf_counter() - t1) * 1e3)
    view, _, _ = b.result()
    full_s = float(np.mean(full_lat)) / 1e3
    out.update({"full_graph_gflops": round(2 * products / full_s / 1e9, 3),
                "full_graph_products_per_s": round(count / full_s, 1),
                "full_graph_latency_ms_p50": round(float(np.percentile(full_lat, 50)), 3),
                "full_graph_latency_ms_p99": round(float(np.percentile(full_lat, 99)), 3),
                "full_graph_status": b.full_status(),
                "full_graph_same_structure": bool(np.array_equal(view.row_set_hash(), ref_hash)),
                "full_graph_vs_plain": round(plain_s / full_s, 2)})
    return out


def compulsory(A, B, st, dtype: str, ms: float) -> dict:
    """The compulsory traffic of one product: A and B read once, C written
    once (u32 offsets and columns of the operands, u64 offsets of C), and
    the share of 8 TB/s it is at the measured step time."""
    s = 8 if dtype == "f64" else 4
    ops = sum((M.rows + 1) * 4 + M.nnz * (4 + s) for M in (A, B))
    c = (st["C"]["rows"] + 1) * 8 + st["C"]["nnz"] * (4 + s)
    b = ops + c
    return {"bytes": int(b), "achieved_gbps": round(b / (ms * 1e-3) / 1e9, 2),
            "frac": round(b / (ms * 1e-3) / 8e12, 4)}


def measured_traffic(workload: str, family: str):
    """HBM bytes per launch of the dominant kernel from the rocprofv3
    FETCH_SIZE / WRITE_SIZE runs summarised under profiles/ (a line
    `traffic_bytes_per_launch <workload> <family> <bytes>`), or None."""
    d = os.path.join(ROOT, "profiles")
    if not os.path.isdir(d):
        return None
    for name in sorted(os.listdir(d)):
        if not name.endswith(".md"):
            continue
        for line in open(os.path.join(d, name)):
            f = line.strip().strip("`").split()
            if len(f) == 4 and f[0] == "traffic_bytes_per_launch" and f[1] == workload and f[2] == family:
                return int(f[3])
    return None


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(A, B, budget_s: float, nthreads: int):
    """Plain row-wise Gustavson (dense FP64 accumulator + touched list, rows
    in column order, OpenMP over rows) on row panels spread over the matrix;
    a product too small to fill the budget is run whole, repeatedly."""
    from oracle import oracle as ora
    m = A.rows
    npan = 64
    pan = max(1, m // (npan * 8))  # 1/8 of the rows in 64 panels
    starts = [int(i * (m // npan)) for i in range(npan)]
    rl = np.diff(B.row_offsets.astype(np.int64))
    ro = A.row_offsets.astype(np.int64)
    cs = np.concatenate(([0], np.cumsum(rl[A.col_ids.astype(np.int64)])))
    prod_row = cs[ro[1:]] - cs[ro[:-1]]
    ora.baseline_spgemm(A, B, 0, min(m, pan), nthreads=nthreads)  # warm the threads
    done_prod, done_rows, t_used = 0, 0, 0.0
    for s in starts:
        e = min(m, s + pan)
        t0 = time.perf_counter()
        ora.baseline_spgemm(A, B, s, e, nthreads=nthreads)
        t_used += time.perf_counter() - t0
        done_prod += int(prod_row[s:e].sum())
        done_rows += e - s
        if t_used > budget_s:
            break
    how = f"{done_rows} rows in {pan}-row panels spread over the matrix"
    if t_used < 0.05 * budget_s:
        # small product: whole passes until a tenth of the budget is used
        done_prod, done_rows, t_used, passes = 0, 0, 0.0, 0
        while t_used < 0.1 * budget_s:
            t0 = time.perf_counter()
            ora.baseline_spgemm(A, B, 0, m, nthreads=nthreads)
            t_used += time.perf_counter() - t0
            done_prod += int(prod_row.sum())
            done_rows += m
            passes += 1
        how = f"{passes} whole passes over the {m} rows"
    gflops = 2.0 * done_prod / t_used / 1e9
    return {
        "value": round(gflops, 3),
        "unit": "GFLOP/s",
        "cores": nthreads,
        "kind": "port",
        "cpu": cpu_model(),
        "sample": f"{how} ({done_prod} products, {t_used:.1f} s); plain row-wise Gustavson, "
                  f"dense FP64 accumulator + touched list, rows in column order (range scan "
                  f"or sort), per-thread output buffers, OpenMP over rows on {nthreads} threads",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="rmat20",
                    choices=["rmat20", "laplacian", "rect", "batched"])
    ap.add_argument("--batch-count", type=int, default=None,
                    help="products in the batched comparison (4096 for --workload batched, "
                         "512 beside the default workload, 0 to skip)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--config", default=None,
                    help="JSON of configuration keys for the product (default for rmat20: the "
                         "autotuner's research_amd/tuned/rmat20.json; 'default': the "
                         "reference's defaults)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run three 1 GiB device-to-device copies (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the library has no multi-GPU path "
              "(row panels of A would be the shards, one per GPU).", file=sys.stderr)
        sys.exit(2)

    import research_amd as ra
    from research_amd import ocean

    if args.calibrate_copy:
        # a plain copy of known size: FETCH_SIZE / WRITE_SIZE of this run give
        # the factors that turn those counters into bytes on this device
        n = 1 << 30
        src, dst = ra.DeviceBuffer(n), ra.DeviceBuffer(n)
        src.zero()
        for _ in range(3):
            ocean._check(ocean.lib.oc_memcpy_dtod(dst.ptr, src.ptr, n))
        ocean.lib.oc_device_synchronize()
        print(json.dumps({"calibration": "device-to-device copy", "bytes_read": n,
                          "bytes_written": n, "copies": 3}), flush=True)
        return

    model_path = os.path.join(ROOT, "research_amd", "models", "rowsize_mlp.bin")
    if args.workload == "batched":
        n = args.batch_count or 4096
        res = run_batched(ra, ocean, n, args.steps, model_path)
        # the configuration's own form: every product's whole path in one
        # captured graph launch (structure and values recomputed each time)
        full_ms = 2 * res["products"] / (res["full_graph_gflops"] * 1e9) * 1e3
        out = {"metric": METRIC, "value": res["full_graph_gflops"], "unit": "GFLOP/s",
               "n_gpus": 1, "steps": max(args.steps, 5), "warmup": 1,
               "ms_per_step": round(full_ms, 3),
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
               "data": "synthetic",
               "products_per_s": res["full_graph_products_per_s"],
               "latency_ms_p50": res["full_graph_latency_ms_p50"],
               "latency_ms_p99": res["full_graph_latency_ms_p99"],
               "config": {"workload": f"{n} independent products C_i = A_i*A_i (ER / banded / "
                                      f"R-MAT, n log-uniform 500..50000, FP32, seed 4), whole "
                                      f"path of all products in one captured graph launch",
                          "parallelism": "single GPU"},
               "batched": res}
        print(json.dumps(out), flush=True)
        return

    t0 = time.perf_counter()
    A, B, desc, dtype = workload(args.workload, ra)
    print(f"[bench] workload built in {time.perf_counter() - t0:.1f} s: {desc}", file=sys.stderr, flush=True)
    dA = A.upload()
    dB = dA if B is A else B.upload()
    cfg_path = args.config
    if cfg_path is None and args.workload == "rmat20":
        cfg_path = os.path.join(ROOT, "research_amd", "tuned", "rmat20.json")
    knobs = {}
    if cfg_path and cfg_path != "default":
        with open(cfg_path) as f:
            knobs = json.load(f)
    sp = ra.SpGEMM(ra.Config(track_stage_times=1, **knobs))

    for i in range(args.warmup):
        C = sp.run(dA, dB)
        print(f"[bench] warmup {i}: {sp.timer_ms('total'):.1f} ms, nnz(C)={C.nnz}", file=sys.stderr, flush=True)
        C.free()
    ocean.lib.oc_device_synchronize()

    fam_ms, fam_bytes, fam_info = {}, {}, {}
    step_ms = []
    t_start = time.perf_counter()
    for i in range(args.steps):
        C = sp.run(dA, dB)
        step_ms.append(sp.timer_ms("total"))
        st = sp.get_stats()
        for k, v in st.get("kernels", {}).items():
            fam_ms[k] = fam_ms.get(k, 0.0) + v["ms"]
            fam_bytes[k] = v["bytes_alg"]
            fam_info[k] = v
        nnz_c = C.nnz
        print(f"[bench] step {i}: {step_ms[-1]:.1f} ms; stages " +
              ", ".join(f"{k}={v:.1f}" for k, v in st.get("timing_flat", {}).items()),
              file=sys.stderr, flush=True)
        if i < args.steps - 1:
            C.free()
    ocean.lib.oc_device_synchronize()
    t_total = time.perf_counter() - t_start
    bad = C.validate()
    C.free()
    ms_per_step = t_total * 1e3 / args.steps
    products = st["stats"]["intermediate_product"]
    value = 2.0 * products / (ms_per_step * 1e-3) / 1e9

    dom = max(fam_ms, key=lambda k: fam_ms[k]) if fam_ms else None
    roof = None
    if dom:
        avg_ms = fam_ms[dom] / args.steps
        ach = fam_bytes[dom] / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "kernel": dom, "achieved": round(ach, 2), "peak": PEAK_HBM_GBPS,
                "unit": "GB/s", "frac": round(ach / PEAK_HBM_GBPS, 4),
                "traffic": measured_traffic(args.workload, dom),
                "avg_launch_ms": round(avg_ms, 3), "bytes_alg_per_launch": fam_bytes[dom]}
    whole = st.get("roofline", {})
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": {"workload": desc, "products": products, "nnz_c": st["C"]["nnz"],
                   "workflow": st["decision"]["workflow"], "parallelism": "single GPU",
                   "tuned_config": knobs or None},
        "roofline": roof,
        "whole_product_roofline": {
            "bytes_alg": whole.get("bytes_alg"),
            "achieved_gbps": round(whole.get("bytes_alg", 0) / (ms_per_step * 1e-3) / 1e9, 2),
            "frac": round(whole.get("bytes_alg", 0) / (ms_per_step * 1e-3) / 8e12, 4)},
        "compulsory_model": compulsory(A, B, st, dtype, ms_per_step),
        "kernel_ms_per_step": {k: round(v / args.steps, 3) for k, v in fam_ms.items()},
        "library_ms_per_step": round(float(np.mean(step_ms)), 3),
        "library_ms_per_step_median": round(float(np.median(step_ms)), 3),
        "structure_invalid_rows": bad,
    }
    if not args.no_cpu_baseline:
        nth = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
        out["cpu_baseline"] = cpu_baseline(A, B, args.cpu_seconds, nth)
    nb = 512 if args.batch_count is None else args.batch_count
    if nb > 0:
        dA.free()
        if dB is not dA:
            dB.free()
        out["batched"] = run_batched(ra, ocean, nb, max(1, min(args.steps, 3)), model_path)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
