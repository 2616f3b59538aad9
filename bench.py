"""Benchmark of the MI355X dispersion / depth-kernel path.

Workload (one step = one full solve): 256 x 256 columns, each 64 layers of
1 km over a half-space (65 entries), depth trend 2.6 + 0.035 z km/s with a
+-8 % Vs checkerboard of 40 km cells, Brocher Vp/rho; 60 periods 1-80 s;
Rayleigh modes 0-3 and Love modes 0-1; phase + group velocity + dc/dVs and
dU/dVs kernels (the Rg/Lg form: three root chains per wave, spherical Earth).
Synthetic, seeded inputs resident in HBM before the timed region.

Prints one JSON line: throughput in column-periods/s (one column-period = c,
U, dc/dVs, dU/dVs for every wave type and mode at one period of one column),
the roofline of the dominant kernel (algorithmic FP64 flops from the frozen
model jdsurfg_amd/flops_model.json over its event-timed duration), and the
CPU baseline (the reference numerics compiled from source, OpenMP over
columns on the host cores, on a bounded column sample).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="grid256")
    ap.add_argument("--cpu-sample", type=int, default=384,
                    help="columns in the CPU baseline sample (0 skips it)")
    ap.add_argument("--no-cpu", action="store_true")
    return ap.parse_args()


def flop_model(w, plan, nevals):
    """Algorithmic flops of the chain launch and the two eigen launches."""
    import numpy as np
    with open(os.path.join(ROOT, "jdsurfg_amd", "flops_model.json")) as f:
        fm = json.load(f)
    N = np.diff(w.layer_off).astype(np.float64)
    K = np.diff(w.per_off).astype(np.float64)
    neig = 3 if w.group else 1
    itemsR = (w.nmode[:, 0] * K * N).sum()
    itemsL = (w.nmode[:, 1] * K * N).sum()
    eigR = itemsR * neig * fm["f_eig"]["rayleigh"]
    eigL = itemsL * neig * fm["f_eig"]["love"]
    return fm, eigR, eigL


def cpu_baseline(w, ncols, nthreads):
    """Time the reference numerics (oracle/_ref, compiled from the reference
    sources) or, if that build is absent, the plain-C restatement, on a
    strided sample of columns; returns column-periods/s and the kind."""
    import numpy as np
    from oracle.pyoracle import Oracle, available
    kind = "reference" if available("ref") else "port"
    o = Oracle("ref" if kind == "reference" else "rest")
    step = max(1, w.ncol // ncols)
    cols = np.arange(0, w.ncol, step)[:ncols]
    sub = w.subset(cols)
    t0 = time.perf_counter()
    o.batch_solve(sub.layer_off, sub.thk, sub.vp, sub.vs, sub.rho, sub.per_off, sub.periods,
                  sub.nmode.ravel(), sub.group, nthreads=nthreads, kernels=False)
    dt = time.perf_counter() - t0
    return sub.column_periods() / dt, kind, len(cols), dt


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(columns are independent: a multi-GPU run would block-shard "
                          "them with no exchange)"}))
        sys.exit(2)
    import numpy as np
    import torch
    from jdsurfg_amd import swd
    from jdsurfg_amd import workloads as W

    os.environ["JDSG_TIMING"] = "1"
    w = W.BUILDERS[args.workload]()
    plan = swd.Plan(w.layer_off, w.per_off, w.nmode, w.group,
                    "grid" if w.grid is not None else "csr")
    inp = swd.upload(w)
    out = plan.alloc_outputs(kernels=("dcdb", "dudb"))
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        plan.execute(*inp, out)
    torch.cuda.synchronize()
    stage_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.execute(*inp, out)   # each execute synchronises on its last event
        stage_ms.append(plan.timings())
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / args.steps
    st = out["status"].cpu().numpy()
    nevals = plan.eval_count()
    cps = w.column_periods()
    value = cps / (ms_step * 1e-3)
    sm = np.mean(np.array(stage_ms), axis=0)  # prep, chains, eigR, eigL (ms)
    fm, eigR, eigL = flop_model(w, plan, nevals)
    N = float(np.mean(np.diff(w.layer_off)))
    # chain flops: evaluations are counted by the kernel; split R/L by the
    # per-layer cost is not observable, so the Rayleigh figure bounds it.
    chain_flops = nevals * N * fm["f_disp"]["rayleigh"]
    stages = {"chains": (sm[1], chain_flops), "eigen_rayleigh": (sm[2], eigR),
              "eigen_love": (sm[3], eigL)}
    dom = max(stages, key=lambda k: stages[k][0])
    dms, dfl = float(stages[dom][0]), float(stages[dom][1])
    achieved = dfl / (dms * 1e-3) / 1e12
    peak = fm["peak_fp64_tflops"]
    line = {
        "metric": "column-periods/s (c, U, dc/dVs, dU/dVs), plain FP64 method",
        "value": float(value),
        "unit": "column-periods/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": float(ms_step),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{w.name}: {w.ncol} columns x {int(N)} layer entries, "
                               f"{int(np.diff(w.per_off)[0])} periods 1-80 s, "
                               "Rayleigh modes 0-3 + Love modes 0-1, phase+group+kernels",
                   "columns": w.ncol, "column_periods": cps,
                   "parallelism": "single GPU"},
        "stage_ms": {"prep": float(sm[0]), "chains": float(sm[1]),
                     "eigen_rayleigh": float(sm[2]), "eigen_love": float(sm[3])},
        "period_equation_evals": int(nevals),
        "status_counts": {str(i): int(c) for i, c in enumerate(np.bincount(st)) if c},
        "roofline": {"bound": "fp64-valu", "kernel": dom, "achieved": achieved, "peak": peak,
                     "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None},
    }
    if not args.no_cpu and args.cpu_sample > 0:
        nthr = min(16, os.cpu_count() or 1)
        v, kind, ncols, dt = cpu_baseline(w, args.cpu_sample, nthr)
        line["cpu_baseline"] = {"value": float(v), "unit": "column-periods/s", "cores": nthr,
                                "kind": kind,
                                "sample": f"{ncols} strided columns of the same workload "
                                          f"({ncols * int(np.diff(w.per_off)[0])} column-periods, "
                                          f"{dt:.1f} s, OpenMP over columns)"}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
