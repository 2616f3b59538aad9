"""Batched uneven solve across inversion iterations.

One plan holds columns of any layer count, period set and mode count
(CSR offsets); each stage is one persistent launch whose lanes fetch work.
`run_iterations` picks the launch widths with the library's autotuner,
captures one iteration -- an in-place 1 % Vs model update (Vp and rho by
Brocher's relations) followed by the whole solve -- into a hipGraph, and
replays it, timing each replay on the stream.  The reference re-runs
SurfTime::get_1d_kernel once per inversion iteration (surftomo.cpp:27-138);
the graph turns that into one launch per iteration.
"""
import json
import os
import time

import numpy as np

from . import swd
from . import workloads as W


def run_iterations(w, iters=10, amp=0.01, seed=1, kernels=("dcdb", "dudb"), tune=True,
                   log=None):
    """Autotune, capture and replay `iters` iterations of workload w.
    Returns a report dict (times in ms)."""
    import torch
    plan = swd.Plan(w.layer_off, w.per_off, w.nmode, w.group, "csr")
    inp = swd.upload(w)
    out = plan.alloc_outputs(kernels)
    st = torch.cuda.Stream()
    rep = {"workload": w.name, "columns": w.ncol, "column_periods": w.column_periods(),
           "layers": int(w.layer_off[-1]), "iterations": iters, "perturbation": amp}
    with torch.cuda.stream(st):
        t0 = time.perf_counter()
        plan.execute(*inp, out, stream=st)
        st.synchronize()
        rep["first_solve_ms"] = (time.perf_counter() - t0) * 1e3
        if tune:
            t0 = time.perf_counter()
            cw, ew = plan.autotune(*inp, out, stream=st)
            st.synchronize()
            rep["autotune_ms"] = (time.perf_counter() - t0) * 1e3
            rep["chain_waves"], rep["eigen_waves"] = cw, ew
        # eager reference timing at the tuned widths
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record(st)
        plan.execute(*inp, out, stream=st)
        ev[1].record(st)
        st.synchronize()
        rep["eager_ms"] = ev[0].elapsed_time(ev[1])
        g = plan.capture(*inp, out, st, perturb_amp=amp, seed=seed)
        times = []
        vs0 = inp[2].clone()
        for i in range(iters):
            ev[0].record(st)
            g.launch(st)
            ev[1].record(st)
            st.synchronize()
            times.append(ev[0].elapsed_time(ev[1]))
            if log:
                log(f"iteration {i} {times[-1]:.1f} ms")
        drift = (inp[2] / torch.where(vs0 > 0, vs0, torch.ones_like(vs0)) - 1.0).abs().max()
    rep["graph_ms"] = times
    rep["graph_ms_mean"] = float(np.mean(times))
    rep["column_periods_per_s"] = w.column_periods() / (rep["graph_ms_mean"] * 1e-3)
    rep["max_vs_drift"] = float(drift.item())
    rep["status_counts"] = {str(i): int(c) for i, c in
                            enumerate(np.bincount(out["status"].cpu().numpy())) if c}
    return rep


def run_surrogate_iterations(w, iters=10, amp=0.01, seed=1, kernels=("dcdb", "dudb"), mlp=None,
                             log=None):
    """The same iterations with the learned bracket: the plan in exact-root
    mode with the network bound to it, each iteration (1 % Vs update, network,
    bracket/polish, fallback chains, eigen stage) captured once and replayed.
    Returns a report dict (times in ms) with the fallback rate."""
    import torch
    from . import surrogate as S
    wpath = S.weights_for(w.nmode) if mlp is None else "given"
    mlp = mlp or S.MLP.load(wpath)
    plan = swd.Plan(w.layer_off, w.per_off, w.nmode, w.group, "csr")
    plan.set_tolerance(1e-10)
    S.bind_network(plan, mlp)
    inp = swd.upload(w)
    out = plan.alloc_outputs(kernels)
    st = torch.cuda.Stream()
    rep = {"workload": w.name, "columns": w.ncol, "column_periods": w.column_periods(),
           "iterations": iters, "perturbation": amp, "network_width": mlp.hid,
           "weights": os.path.relpath(wpath, os.path.dirname(os.path.dirname(
               os.path.abspath(__file__)))) if wpath != "given" else wpath}
    with torch.cuda.stream(st):
        S.surrogate_solve_net(plan, inp, out, stream=st.cuda_stream)
        st.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record(st)
        S.surrogate_solve_net(plan, inp, out, stream=st.cuda_stream)
        ev[1].record(st)
        st.synchronize()
        rep["eager_ms"] = ev[0].elapsed_time(ev[1])
        g = S.capture_surrogate(plan, inp, out, st, perturb_amp=amp, seed=seed)
        times = []
        for i in range(iters):
            ev[0].record(st)
            g.launch(st)
            ev[1].record(st)
            st.synchronize()
            times.append(ev[0].elapsed_time(ev[1]))
            if log:
                log(f"surrogate iteration {i} {times[-1]:.1f} ms")
    stats = S.surrogate_stats(plan)
    nvar = 3 if w.group else 1
    rep["graph_ms"] = times
    rep["graph_ms_mean"] = float(np.mean(times))
    rep["column_periods_per_s"] = w.column_periods() / (rep["graph_ms_mean"] * 1e-3)
    rep["fallback_rate"] = stats["roots_left_to_chain"] / (nvar * plan.nent)
    rep["status_counts"] = {str(i): int(c) for i, c in
                            enumerate(np.bincount(out["status"].cpu().numpy())) if c}
    return rep


def main(argv=None):
    """python -m jdsurfg_amd.batch [--ncol N] [--iters K]: the uneven batch
    (50,000 columns, 16-128 layers, 8-64 periods, 1-3 modes per wave type)
    autotuned, graph-captured and replayed with a 1 % Vs update per
    iteration; prints one JSON report."""
    import argparse
    ap = argparse.ArgumentParser(prog="python -m jdsurfg_amd.batch")
    ap.add_argument("--ncol", type=int, default=50000)
    ap.add_argument("--iters", type=int, default=10)
    ap.add_argument("--amp", type=float, default=0.01)
    a = ap.parse_args(argv)
    w = W.uneven(ncol=a.ncol)
    print(json.dumps(run_iterations(w, a.iters, a.amp, log=lambda m: print(m, flush=True))),
          flush=True)


if __name__ == "__main__":
    main()
