"""Benchmark of the MI355X dispersion / depth-kernel path.

Workload (one step = one full solve): 256 x 256 columns, each 64 layers of
1 km over a half-space (65 entries), depth trend 2.6 + 0.035 z km/s with a
+-8 % Vs checkerboard of 40 km cells, Brocher Vp/rho; 60 periods 1-80 s;
Rayleigh modes 0-3 and Love modes 0-1; phase + group velocity + dc/dVs and
dU/dVs kernels (the Rg/Lg form: three root chains per wave, spherical Earth).
Synthetic, seeded inputs resident in HBM before the timed region.

Two legs on the same inputs:
  plain      the reference method on the GPU (dc-grid scan + Neville chains,
             float-rounded roots), then the eigen/kernel stage;
  surrogate  the learned-bracket solve: BF16 MFMA network prediction of every
             root, +-0.5 % bracket, FP64 polish to 1e-10, the plain chain for
             roots the checks reject, then the same eigen/kernel stage.
`value` is the surrogate leg (the quoted metric is surrogate-bracket against
plain); the plain leg, the ratio and the accuracy of the surrogate against
the plain method in exact-root mode ride along.

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
    ap.add_argument("--workload", default="grid256",
                    help="grid256 (the default line), grid128, uneven (the 50,000-column "
                         "batch, replayed graph), lsmr (LSMR on Frechet rows), smooth "
                         "(gradient smoothing) or fmst (traveltimes and rays of 3040 station "
                         "pairs)")
    ap.add_argument("--cpu-sample", type=int, default=384,
                    help="columns in the CPU baseline sample (0 skips it)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--surftime-steps", type=int, default=2,
                    help="timed calls of the drop-in SurfTime leg (0 skips it)")
    ap.add_argument("--weights", default=None,
                    help="surrogate network (.npz); default: the 4 x 512 network for grid128 "
                         "(its 128 x 128 sediment columns), the 4 x 256 one otherwise")
    ap.add_argument("--fmst-steps", type=int, default=2,
                    help="timed calls of the FMST leg (0 skips it)")
    return ap.parse_args()


def flop_model():
    """The frozen algorithmic flop model (op-counted on the oracle
    restatement, oracle/flopcount/freeze.py)."""
    with open(os.path.join(ROOT, "jdsurfg_amd", "flops_model.json")) as f:
        return json.load(f)


def eigen_pass_flops(w, status, fm):
    """Algorithmic flops of the Rayleigh eigen pass launches of one solve:
    every Rayleigh entry with a T root takes the T pass (status 0, 2, 8) and,
    with group kernels, those with T(1+-0.005) roots (status 0) two more;
    N layer entries x f_eig per pass."""
    import numpy as np
    N = np.diff(w.layer_off).astype(np.int64)
    K = np.diff(w.per_off).astype(np.int64)
    nR = w.nmode[:, 0].astype(np.int64)
    ne = (w.nmode.sum(axis=1) * K).astype(np.int64)
    col = np.repeat(np.arange(w.ncol), ne)
    j = np.arange(len(status)) - np.repeat(np.concatenate([[0], np.cumsum(ne)[:-1]]), ne)
    ray = j < (nR * K)[col]
    t_pass = ray & np.isin(status, (0, 2, 8))
    pm_pass = ray & (status == 0) if w.group else np.zeros_like(ray)
    lay = N[col].astype(np.float64)
    passes = float((lay * t_pass).sum() + 2.0 * (lay * pm_pass).sum())
    return passes * fm["f_eig"]["rayleigh"], int(t_pass.sum()), int(pm_pass.sum())


def kernel_counters(name):
    """Per-dispatch HBM bytes and hardware FP64 flops of a kernel from the
    committed rocprofv3 counter runs (jdsurfg_amd/traffic_model.json)."""
    p = os.path.join(ROOT, "jdsurfg_amd", "traffic_model.json")
    if not os.path.exists(p):
        return None, None
    with open(p) as f:
        d = json.load(f)
    return d["kernels"].get(name), d.get("source")


def cpu_baseline(w, ncols, nthreads):
    """Time the reference numerics (oracle/_ref, compiled from the reference
    sources) and the project's restatement (oracle/liboracle.so) on the same
    strided sample of columns, OpenMP over columns on nthreads host threads;
    returns both rates in column-periods/s."""
    import numpy as np
    from oracle.pyoracle import Oracle, available
    step = max(1, w.ncol // ncols)
    cols = np.arange(0, w.ncol, step)[:ncols]
    sub = w.subset(cols)
    out = {}
    for kind in (("ref", "rest") if available("ref") else ("rest",)):
        o = Oracle(kind)
        t0 = time.perf_counter()
        o.batch_solve(sub.layer_off, sub.thk, sub.vp, sub.vs, sub.rho, sub.per_off, sub.periods,
                      sub.nmode.ravel(), sub.group, nthreads=nthreads, kernels=False)
        out[kind] = (sub.column_periods() / (time.perf_counter() - t0),
                     time.perf_counter() - t0)
    return out, len(cols), sub.column_periods()


def lsmr_main(args, nx=256, ny=256, nz=17, kmax=120, nrows=20000, it_gpu=40, it_cpu=3):
    """LSMR (lsmrModule_csr.f90) on the Frechet rows of 20000 straight rays
    over the 256 x 256 x 17 grid (the rows jdsg_frechet_rows builds from random
    depth kernels; 254*254*16 model cells), b random, damp 0, run for a fixed
    iteration count (atol = btol = conlim = 0, so istop = 7).  Prints one JSON
    line: GPU iterations/s of the cell-major solve (jdsg_lsmr_csr_layers, 16
    layers) from the difference of a 1- and an it_gpu-iteration solve (the
    one-time set-up reported apart; the solve in the caller's layer-major
    numbering rides along), the algorithmic bytes of an
    iteration over its time, and the compiled reference LSMR on the host
    threads over it_cpu iterations, with the relative difference of the two
    solutions after it_cpu iterations."""
    import numpy as np
    import torch
    from jdsurfg_amd import frechet as F
    from jdsurfg_amd import lsmr as G
    rng = np.random.default_rng(0)
    g = torch.Generator(device="cuda").manual_seed(0)
    kernel = torch.randn(nx * ny * nz * kmax, dtype=torch.float64, device="cuda", generator=g)
    ray_off, ray_cell, ray_w = F.straight_rays(nx, ny, nrows, 0)
    row_idx = rng.integers(0, kmax, nrows).astype(np.int32)
    ro, col, val = F.frechet_rows(kernel, nx, ny, nz, kmax, row_idx, ray_off, ray_cell, ray_w)
    del kernel
    ptr = ro.to(torch.int32)
    m, n, nnz = nrows, (nx - 2) * (ny - 2) * (nz - 1), int(col.numel())
    b = torch.randn(m, device="cuda", generator=g)
    print(f"lsmr: {m} rows, {n} columns, {nnz} nonzeros", file=sys.stderr, flush=True)

    def solve(itn, layers):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x, info = G.lsmr(m, n, val, col, ptr, b, 0.0, 0.0, 0.0, 0.0, itn, 10, layers=layers)
        torch.cuda.synchronize()
        return x, info, time.perf_counter() - t0

    def timing(layers):
        solve(1, layers)                       # warm the library and the allocator
        t1 = min(solve(1, layers)[2] for _ in range(3))
        xs, infos, tk = zip(*[solve(it_gpu, layers) for _ in range(max(1, args.steps // 2))])
        assert all(i["itn"] == it_gpu and i["istop"] == 7 for i in infos)
        per_it = (min(tk) - t1) / (it_gpu - 1)
        return per_it, t1 - per_it, xs[0]

    # the caller's (layer-major) numbering, and the cell-major solve
    per_it_plain, setup_plain, x_plain = timing(1)
    per_it, setup, x_cell = timing(nz - 1)
    # an iteration reads each of A and A^T once (value 4 B + index 4 B per
    # nonzero, 4 B per row pointer), gathers x and writes y once (4 B each),
    # and streams n-vectors for the norms, the reorthogonalisation against 10
    # vectors (2 reads of v and q, 1 write of v) and the h/hbar/x update
    alg = 2 * (8 * nnz) + 4 * (m + 1) + 4 * (n + 1) + 8 * (m + n) * 2 \
        + 4 * n * (10 * 3 + 7 + 2) + 4 * m * 2
    rep = {"metric": "LSMR iterations/s on Frechet rows (float, local reorthogonalisation 10)",
           "value": 1.0 / per_it, "unit": "iterations/s", "n_gpus": 1, "steps": max(1, args.steps // 2),
           "warmup": 1, "ms_per_step": per_it * 1e3, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"lsmr_frechet_{m}x{n}", "rows": m, "columns": n,
                      "nnz": nnz, "iterations": it_gpu, "layers": nz - 1},
           "setup_ms": setup * 1e3,
           "layer_major": {"ms_per_step": per_it_plain * 1e3, "setup_ms": setup_plain * 1e3},
           "x_rel_diff_cell_vs_layer_major": float((x_cell - x_plain).abs().max() /
                                                   x_plain.abs().max()),
           "roofline": {"bound": "hbm", "achieved": alg / per_it / 1e9, "peak": 8000.0,
                        "unit": "GB/s", "frac": alg / per_it / 8e12, "traffic": None,
                        "note": "whole iteration (both products, norms, host syncs)"}}
    if not args.no_cpu:
        from oracle import pyoracle
        if pyoracle.available("ref") and os.path.exists(pyoracle.LIB_LSMR_REF):
            kind, lib = "reference", "ref"
        else:
            kind, lib = "port", "rest"
        threads = 16 if lib == "ref" else 1
        hv, hi, hp, hb = (t.cpu().numpy() for t in (val, col, ptr, b))
        t0 = time.perf_counter()
        xr, ir = pyoracle.lsmr(m, n, hv, hi, hp, hb, 0.0, 0.0, 0.0, 0.0, it_cpu, 10, kind=lib,
                               threads=threads)
        dt = time.perf_counter() - t0
        xg = solve(it_cpu, nz - 1)[0].cpu().numpy()
        rep["cpu_baseline"] = {"value": it_cpu / dt, "unit": "iterations/s", "cores": threads,
                               "kind": kind, "sample": f"{it_cpu} iterations of the same solve"}
        rep["x_rel_diff_vs_cpu"] = float(np.abs(xg - xr).max() / np.abs(xr).max())
    print(json.dumps(rep), flush=True)


def smooth_main(args, nx=256, ny=256):
    """DSurfTomo::smoothing (surftomo.cpp:147-178) of a random gradient over a
    256 x 256-node model with 17 irregular depths 0-100 km (finest step 1 km,
    so 80 regular depths): the spherical form (sigma_h 20 km, sigma_v 5 km,
    0.1 degree nodes) and the Cartesian form (sigma_h 3, sigma_v 2 nodes).
    Prints one JSON line: node-steps/s of the spherical driver (each explicit
    diffusion step updates every node of the 254 x 254 x 80 regular grid), the
    Cartesian driver beside it, and the oracle restatement on the host (one
    thread) over a few steps of the same grid."""
    import numpy as np
    import torch
    from jdsurfg_amd import smooth as S
    dep = np.array([0, 1, 2, 4, 6, 8, 10, 15, 20, 25, 30, 40, 50, 60, 70, 80, 100], np.float32)
    lat = (40.0 - 0.1 * np.arange(nx)).astype(np.float32)
    lon = (100.0 + 0.1 * np.arange(ny)).astype(np.float32)
    g0 = torch.randn((dep.size - 1, ny - 2, nx - 2), generator=torch.Generator().manual_seed(0))
    g0 = g0.cuda()
    nz1 = int((dep[-2] - dep[0]) / (1.0 * 0.99))
    nodes = (nx - 2) * (ny - 2) * nz1

    def run(in_km, sh, sv):
        g = g0.clone()
        S.smooth_gradient(g, dep, lat, lon, in_km, sh, sv)      # warm-up
        torch.cuda.synchronize()
        ts = []
        for _ in range(max(1, args.steps // 2)):
            g = g0.clone()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            S.smooth_gradient(g, dep, lat, lon, in_km, sh, sv)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return min(ts)

    from oracle import pyoracle as P
    dz = float(np.float32((dep[-2] - dep[0]) / np.float32(nz1 - 1)))
    geo = P.sph_geometry(nx - 2, ny - 2, nz1, 0.1, 0.1, dz, float(lat[1]), float(lon[1]), 0.0,
                         20.0, 5.0)
    nstep_sph = geo["nstep"]
    nstep_cart = int(float(np.float32(9.0)) / (2.0 * float(np.float32(1.0 / 12))))
    t_sph = run(1, 20.0, 5.0)
    t_cart = run(0, 3.0, 2.0)
    rep = {"metric": "gradient smoothing node-steps/s (spherical PDE smoother, 254x254x80)",
           "value": nodes * nstep_sph / t_sph, "unit": "node-steps/s", "n_gpus": 1,
           "steps": max(1, args.steps // 2), "warmup": 1, "ms_per_step": t_sph * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32/f64",
           "data": "synthetic",
           "config": {"workload": "smooth_gradient_256x256x17", "regular_depths": nz1,
                      "sph_steps": nstep_sph, "cart_steps": nstep_cart},
           "cartesian": {"ms": t_cart * 1e3, "node_steps_per_s": nodes * nstep_cart / t_cart}}
    if not args.no_cpu:
        g = np.random.default_rng(0).normal(size=(nz1, ny - 2, nx - 2)).astype(np.float32)
        # sigma_v = sigma_h giving 4 steps of the same grid
        h2 = float(geo["Ch"]) / float(np.float32(1.0 / 7)) * (20.0 ** 2 / 20.0 ** 2)
        sig = float(np.sqrt(4.5 * 2 * (1.0 / 7) * h2))
        ns = P.sph_geometry(nx - 2, ny - 2, nz1, 0.1, 0.1, dz, float(lat[1]), float(lon[1]), 0.0,
                            sig, sig)["nstep"]
        t0 = time.perf_counter()
        P.smooth_sph_pde(g, 0.1, 0.1, dz, float(lat[1]), float(lon[1]), 0.0, sig, sig)
        dt = time.perf_counter() - t0
        rep["cpu_baseline"] = {"value": nodes * ns / dt, "unit": "node-steps/s", "cores": 1,
                               "kind": "port", "sample": f"{ns} steps of the spherical smoother "
                               "on the same grid (the reference needs Eigen, absent)"}
    print(json.dumps(rep), flush=True)


def timed(fn, steps, warmup):
    """Run fn warmup times, then time `steps` calls between device syncs."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    per = []
    for _ in range(steps):
        per.append(fn())
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps, per


def uneven_main(args):
    """The batched uneven workload (50,000 columns, 16-128 layers of
    0.5-4 km, 8-64 periods in 1-80 s, 1-3 modes per wave type, group kernels):
    one plan, autotuned launch widths, one inversion iteration (1 % Vs update
    with Brocher Vp/rho, then the whole solve) captured in a hipGraph and
    replayed `steps` times.  Prints one JSON line: column-periods/s of a
    replay (mean over the replays)."""
    from jdsurfg_amd import batch as B
    from jdsurfg_amd import workloads as W
    w = W.uneven()
    rep = B.run_iterations(w, iters=max(2, args.steps * 2), amp=0.01,
                           log=lambda m: print(m, file=sys.stderr, flush=True))
    line = {"metric": "column·periods/s per replayed inversion iteration (uneven batch)",
            "value": rep["column_periods_per_s"], "unit": "column·periods/s", "n_gpus": 1,
            "steps": len(rep["graph_ms"]), "warmup": 1, "ms_per_step": rep["graph_ms_mean"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": f"uneven: {w.ncol} columns, 16-128 layers, 8-64 periods, "
                                   "1-3 modes per wave, group kernels, hipGraph replay",
                       "columns": w.ncol, "column_periods": rep["column_periods"]},
            "detail": {k: rep[k] for k in ("eager_ms", "autotune_ms", "chain_waves",
                                            "eigen_waves", "graph_ms", "status_counts",
                                            "max_vs_drift") if k in rep}}
    # the same iterations with the learned bracket (exact-root mode), captured
    # with the network in the library
    sr = B.run_surrogate_iterations(w, iters=max(2, args.steps * 2), amp=0.01,
                                    log=lambda m: print(m, file=sys.stderr, flush=True))
    line["surrogate"] = {"value": sr["column_periods_per_s"], "unit": "column·periods/s",
                         "ms_per_step": sr["graph_ms_mean"], "graph_ms": sr["graph_ms"],
                         "eager_ms": sr["eager_ms"], "fallback_rate": sr["fallback_rate"],
                         "network_width": sr["network_width"], "weights": sr["weights"],
                         "speedup_vs_plain": rep["graph_ms_mean"] / sr["graph_ms_mean"]}
    print(json.dumps(line), flush=True)


def surftime_leg(steps):
    """The drop-in SurfTime form (get_1d_kernel, SurfaceWave.cpp:458-502) on a
    grid like grid256: 256 x 256 columns of 17 nodes 0-64 km (4 km apart,
    refined to 65 layer entries), the grid256 Vs (trend with the +-8 % checkerboard
    at the node depths), Rg tables of 4 modes and Lg tables of 2 modes at the
    60 periods 1-80 s; Vp and rho by Brocher inside, the a/b/rho kernels,
    the chain rule and the grid epilogue; Vs and outputs resident in HBM."""
    import numpy as np
    import torch
    from jdsurfg_amd import swd
    nx = ny = 256
    dep = (4.0 * np.arange(17)).astype(np.float32)
    x = (np.arange(nx) + 0.5) * 10.0
    sgn = (-1.0) ** ((np.floor(x / 40.0)[:, None, None] + np.floor(x / 40.0)[None, :, None] +
                      np.floor(dep / 40.0)[None, None, :]) % 2)
    vs = ((2.6 + 0.035 * dep)[None, None, :] * (1.0 + 0.08 * sgn)).astype(np.float32)
    per = np.geomspace(1.0, 80.0, 60)
    tabs = {"Rg": np.tile(per[:, None], (1, 4)), "Lg": np.tile(per[:, None], (1, 2))}
    S = swd.SurfTime(nx, ny, 17, dep, tabs)
    dvs = torch.from_numpy(np.ascontiguousarray(vs.ravel(order="F"))).cuda()
    out = S.alloc(kernel=True)

    def run():
        S.kernel(dvs, out)
    ms, _ = timed(run, steps, 1)
    cps = nx * ny * len(per)
    st = out["status"].cpu().numpy()
    return {"value": float(cps / (ms * 1e-3)), "unit": "column·periods/s",
            "ms_per_step": float(ms), "steps": steps,
            "config": "SurfTime get_1d_kernel, 256x256 columns x 17 nodes (65 layer entries), "
                      "Rg modes 0-3 + Lg modes 0-1, 60 periods 1-80 s, reference tolerance",
            "status_counts": {str(i): int(c) for i, c in enumerate(np.bincount(st)) if c}}


def fmst_leg(steps, cpu_pairs, nthreads):
    """FMST at the reference example's scale (jdsurfg_amd.workloads.
    station_pairs: 33 x 36 x 41 model, 3040 station pairs, 41,033 receivers,
    Rayleigh phase and group at 19 periods): the phase/group maps and depth
    kernels of the model (get_1d_kernel), then every pair's fast-marching
    traveltimes and every receiver's ray and weights (the FMST loop of
    SurfTime::frechet_matrix, SurfaceWave.cpp:598-672), inputs resident in
    HBM.  Timed: the FMST pass alone (jdsg_fmst_run mode 1 + jdsg_fmst_rays)
    and the whole frechet_matrix call (kernels, FMST, Fréchet rows).  CPU: the
    reference FMST compiled from source in the reference's OpenMP pair loop
    (oracle/_ref, all host threads) on the first `cpu_pairs` pairs."""
    import numpy as np
    import torch
    from jdsurfg_amd import fmst, swd
    from jdsurfg_amd import workloads as W
    d = W.station_pairs()
    nx, ny, lat0, lon0, dlat, dlon = d["grid"]
    S = swd.SurfTime(nx, ny, len(d["dep"]), d["dep"], {w: d["tables"][w][0] for w in ("Rc", "Rg")})
    vs = torch.from_numpy(np.ascontiguousarray(d["vs"].ravel(order="F"))).cuda()
    F = fmst.from_data(d, nx, ny, lat0, lon0, dlat, dlon)
    out = S.alloc(kernel=True)
    S.kernel(vs, out)
    torch.cuda.synchronize()
    res = {}

    def run_fmst():
        res["o"] = F.run(out["vc"], out["vout"], S.kmax, mode=1)
        res["r"] = F.rays()
    ms, _ = timed(run_fmst, steps, 1)
    nz = int(res["r"][1].numel())

    def run_frechet():
        res["f"] = fmst.frechet_matrix(S, F, vs)
    ms_f, _ = timed(run_frechet, steps, 1)
    st = res["o"]["rstatus"].cpu().numpy()
    line = {"value": float(F.npair / (ms * 1e-3)), "unit": "station pairs/s",
            "ms_per_step": float(ms), "steps": steps,
            "receivers_per_s": float(F.nrcv / (ms * 1e-3)),
            "frechet_matrix_ms": float(ms_f),
            "frechet_rows_nonzeros": int(res["f"][3].numel()),
            "ray_weights": nz,
            "config": f"FMST, {nx}x{ny} nodes (propagation grid {(nx - 3) * 8 + 1}x{(ny - 3) * 8 + 1}, "
                      f"source boxes refined 8x), {F.npair} station pairs, {F.nrcv} receivers, "
                      "38 period maps; times, rays and weights for every receiver",
            "status_counts": {str(i): int(c) for i, c in enumerate(np.bincount(st)) if c}}
    if cpu_pairs > 0:
        from oracle import pyoracle as P
        n = min(cpu_pairs, F.npair)
        k = F._keep
        vc = out["vc"].cpu().numpy().reshape(S.kmax, nx * ny)
        vo = out["vout"].cpu().numpy().reshape(S.kmax, nx * ny)
        sub = (k[0][:n], k[1][:n], k[2][:n], k[3][:n + 1], k[4][:k[3][n]], k[5][:k[3][n]])
        kind = "reference" if P.FmstRef.available() else "port"
        fn = P.fmst_pairs_ref if kind == "reference" else P.fmst_pairs
        t0 = time.perf_counter()
        r = fn(d["grid"], vc, vo, *sub, want_fdm=False, nthreads=nthreads)
        dt = time.perf_counter() - t0
        gt = res["o"]["ttime"].cpu().numpy()[:k[3][n]]
        exp = r[0].copy()
        for p in range(n):      # Rg pairs report the ray time (SurfaceWave.cpp:616-618)
            if k[6][p]:
                exp[k[3][p]:k[3][p + 1]] = r[1][k[3][p]:k[3][p + 1]]
        line["cpu_baseline"] = {"value": float(n / dt), "unit": "station pairs/s", "cores": nthreads,
                                "kind": kind,
                                "sample": f"the first {n} pairs ({k[3][n]} receivers), {dt:.1f} s, "
                                          f"OpenMP over pairs on {nthreads} host threads",
                                "gpu_times_identical": bool(np.array_equal(gt, exp))}
    return line


def fmst_main(args):
    nthr = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    leg = fmst_leg(args.steps, 0 if args.no_cpu else 256, nthr)
    line = {"metric": "station pairs/s (FMST traveltimes + ray weights, example scale)",
            "value": leg["value"], "unit": "station pairs/s", "n_gpus": 1, "steps": args.steps,
            "warmup": 1, "ms_per_step": leg["ms_per_step"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": leg["config"]}, "detail": leg}
    if "cpu_baseline" in leg:
        line["cpu_baseline"] = leg.pop("cpu_baseline")
    print(json.dumps(line), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(columns are independent: a multi-GPU run would block-shard "
                          "them with no exchange)"}))
        sys.exit(2)
    if args.workload == "lsmr":
        return lsmr_main(args)
    if args.workload == "smooth":
        return smooth_main(args)
    if args.workload == "uneven":
        return uneven_main(args)
    if args.workload == "fmst":
        return fmst_main(args)
    import numpy as np
    import torch
    from jdsurfg_amd import swd
    from jdsurfg_amd import workloads as W

    os.environ["JDSG_TIMING"] = "1"
    w = W.BUILDERS[args.workload]()
    layout = "grid" if w.grid is not None else "csr"
    kern = ("dcdb", "dudb")
    inp = swd.upload(w)
    cps = w.column_periods()
    N = float(np.mean(np.diff(w.layer_off)))
    fm = flop_model()
    peak = fm["peak_fp64_tflops"]

    # ---- plain leg (the reference method)
    plain = swd.Plan(w.layer_off, w.per_off, w.nmode, w.group, layout)
    pout = plain.alloc_outputs(kernels=kern)

    def run_plain():
        plain.execute(*inp, pout)   # synchronises on its last event (timing on)
        return list(plain.timings()) + [float(plain.pass_timings()[0][0])]
    ms_plain, st_plain = timed(run_plain, args.steps, args.warmup)
    sm_plain = np.mean(np.array(st_plain), axis=0)  # prep, chains, eigR, eigL, R passes
    nevals = plain.eval_count()
    status = pout["status"].cpu().numpy()
    line = {
        # the headline metric: the surrogate-bracketed solve's throughput,
        # with the plain solve and the ratio beside it
        "metric": "column·periods/s (c, U, dc/dVs) at ≤1e-6 rel err: surrogate-bracket vs plain",
        "value": None, "unit": "column·periods/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{w.name}: {w.ncol} columns x {int(N)} layer entries, "
                               f"{int(np.diff(w.per_off)[0])} periods {w.periods.min():g}-"
                               f"{w.periods.max():g} s, Rayleigh modes 0-"
                               f"{int(np.asarray(w.nmode).reshape(w.ncol, -1)[:, 0].max()) - 1} + "
                               f"Love modes 0-"
                               f"{int(np.asarray(w.nmode).reshape(w.ncol, -1)[:, 1].max()) - 1}, "
                               "phase+group+kernels",
                   "columns": w.ncol, "column_periods": cps, "parallelism": "single GPU"},
        "plain": {"value": float(cps / (ms_plain * 1e-3)), "ms_per_step": float(ms_plain),
                  "stage_ms": {"prep": float(sm_plain[0]), "chains": float(sm_plain[1]),
                               "eigen_rayleigh": float(sm_plain[2]),
                               "eigen_love": float(sm_plain[3])},
                  "period_equation_evals": int(nevals),
                  "status_counts": {str(i): int(c) for i, c in enumerate(np.bincount(status))
                                    if c}},
    }
    # roofline of the plain leg's chain launches (the plain method's dominant
    # kernel): E evaluations x N x f_disp, Rayleigh and Love apart would need
    # per-wave counts, so the Rayleigh figure bounds it from above
    chain_flops = nevals * N * fm["f_disp"]["rayleigh"]
    line["plain"]["chain_roofline"] = {
        "achieved": float(chain_flops / (sm_plain[1] * 1e-3) / 1e12), "peak": peak,
        "unit": "TFLOP/s", "note": "upper bound: every evaluation counted at the Rayleigh f_disp"}

    dom_name = "jdsg_eigen_pass<0>"
    dom_status, dom_ms, dom_n = status, float(sm_plain[4]), int(plain.pass_timings()[1][0])
    if not args.no_surrogate:
        from jdsurfg_amd import surrogate as S
        wpath = args.weights or (os.path.join(os.path.dirname(S.DEFAULT_WEIGHTS),
                                              "surrogate_mlp_512.npz")
                                 if args.workload == "grid128" else S.weights_for(w.nmode))
        mlp = S.MLP.load(wpath)
        sp = swd.Plan(w.layer_off, w.per_off, w.nmode, w.group, layout)
        sp.set_tolerance(1e-10)
        sout = sp.alloc_outputs(kernels=kern)
        # the whole iteration in the library (network, bracket/polish,
        # fallback, eigen stage) with no host work between its launches
        S.bind_network(sp, mlp)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]

        def run_sur():
            ev[0].record()
            S.surrogate_solve_net(sp, inp, sout)
            ev[1].record()
            ev[1].synchronize()
            t = [float(x) for x in sp.timings()]
            # the network runs first, before the plan's own stage events
            net = ev[0].elapsed_time(ev[1]) - sum(t)
            return [net] + t + [float(sp.pass_timings()[0][0])]
        ms_sur, st_sur = timed(run_sur, args.steps, args.warmup)
        sm = np.mean(np.array(st_sur), axis=0)  # predict, prep+polish, chains, eigR, eigL, passes
        stats = S.surrogate_stats(sp)
        nvar = 3 if w.group else 1
        # accuracy against the plain method in exact-root mode, same inputs
        ex = swd.Plan(w.layer_off, w.per_off, w.nmode, w.group, layout)
        ex.set_tolerance(1e-10)
        eout = ex.alloc_outputs(kernels=kern)
        ex.execute(*inp, eout)
        torch.cuda.synchronize()
        ok = (eout["status"] == 0) & (sout["status"] == 0)
        relc = ((sout["c"] - eout["c"]).abs() / eout["c"].abs())[ok]
        relu = ((sout["u"] - eout["u"]).abs() / eout["u"].abs())[ok]
        okp = (pout["status"] == 0) & (sout["status"] == 0)
        relcp = ((sout["c"] - pout["c"]).abs() / pout["c"].abs())[okp]
        relup = ((sout["u"] - pout["u"]).abs() / pout["u"].abs())[okp]
        # MLP inference as computed: the first layer once per column (its
        # per-entry part is four vector terms), the hidden layers per entry on
        # MFMA, and the one output row of the head
        H = mlp.hid
        mlp_flops = 2.0 * (sp.ncol * mlp.dims[0][1] * H +
                           sp.nent * (sum(o * i for o, i in mlp.dims[1:4]) + H))
        line.update({
            "value": float(cps / (ms_sur * 1e-3)), "ms_per_step": float(ms_sur),
            "speedup_vs_plain": float(ms_plain / ms_sur),
            "stage_ms": {"mlp_predict": float(sm[0]), "prep_polish": float(sm[1]),
                         "fallback_chains": float(sm[2]), "eigen_rayleigh": float(sm[3]),
                         "eigen_love": float(sm[4])},
            "surrogate": {"roots": nvar * sp.nent,
                          "roots_left_to_chain": stats["roots_left_to_chain"],
                          "fallback_rate": stats["roots_left_to_chain"] / (nvar * sp.nent),
                          "polish_evals": stats["polish_evals"],
                          "chain_evals": stats["chain_evals"],
                          "status_identical_to_plain_exact":
                              bool(torch.equal(eout["status"], sout["status"])),
                          "c_max_rel_vs_plain_exact": float(relc.max().item()),
                          "u_max_rel_vs_plain_exact": float(relu.max().item()),
                          "c_max_rel_vs_plain": float(relcp.max().item()),
                          "u_max_rel_vs_plain": float(relup.max().item()),
                          "u_frac_above_1e-6_vs_plain": float((relup > 1e-6).double().mean()
                                                              .item()),
                          "mlp_tflops": float(mlp_flops / (sm[0] * 1e-3) / 1e12),
                          "mlp_mfma_frac_bf16": float(mlp_flops / (sm[0] * 1e-3) / 2.5e15),
                          "weights": os.path.relpath(wpath, ROOT)},
        })
        dom_status = sout["status"].cpu().numpy()
        dom_ms, dom_n = float(sm[5]), int(sp.pass_timings()[1][0])
        # the same iteration as device work only, captured into a graph and
        # replayed on the same model (a copy of the inputs); the tests replay
        # it with a 1 % Vs update per iteration against eager solves
        gp = swd.Plan(w.layer_off, w.per_off, w.nmode, w.group, layout)
        gp.set_tolerance(1e-10)
        S.bind_network(gp, mlp)
        ginp = tuple(t.clone() for t in inp)
        gout = gp.alloc_outputs(kernels=kern)
        gst = torch.cuda.Stream()
        with torch.cuda.stream(gst):
            S.surrogate_solve_net(gp, ginp, gout, stream=gst.cuda_stream)
            gst.synchronize()
            graph = S.capture_surrogate(gp, ginp, gout, gst, perturb_amp=0.0)
            gev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
            graph.launch(gst)
            gst.synchronize()
            gev[0].record(gst)
            for _ in range(args.steps):
                graph.launch(gst)
            gev[1].record(gst)
            gst.synchronize()
        g_ms = gev[0].elapsed_time(gev[1]) / args.steps
        line["surrogate"]["graph_replay"] = {
            "ms_per_iteration": float(g_ms), "value": float(cps / (g_ms * 1e-3)),
            "unit": "column·periods/s", "replays": args.steps,
            "iteration": "network + bracket/polish + fallback chains + eigen stage, one graph "
                         "launch",
            "identical_to_eager": bool(all(torch.equal(gout[n], sout[n])
                                           for n in ("c", "u", "status") + kern))}
        del gp, ginp, gout, graph
    else:
        line.update({"value": line["plain"]["value"], "ms_per_step": float(ms_plain),
                     "metric": "column-periods/s (c, U, dc/dVs, dU/dVs), plain FP64 method"})
    # roofline of the dominant kernel: the Rayleigh eigen pass launches'
    # algorithmic flops (valid passes only) over their own event-timed time
    flops, n_t, n_pm = eigen_pass_flops(w, dom_status, fm)
    achieved = flops / (dom_ms * 1e-3) / 1e12
    kc, ksrc = kernel_counters(dom_name)
    line["roofline"] = {
        "bound": "fp64-valu", "kernel": dom_name, "achieved": achieved, "peak": peak,
        "unit": "TFLOP/s", "frac": achieved / peak,
        "traffic": (kc["hbm_read_bytes_per_dispatch"] + kc["hbm_write_bytes_per_dispatch"])
        if kc else None,
        "launches": dom_n, "avg_launch_ms": dom_ms / max(dom_n, 1),
        "algorithmic_flops_per_launch": flops / max(dom_n, 1),
        "valid_passes": {"T": n_t, "T_pm": 2 * n_pm},
        "f_eig_per_layer_pass": fm["f_eig"]["rayleigh"],
        "source": "flops: jdsurfg_amd/flops_model.json (op-counted restatement); time: HIP "
                  "events around each pass launch on its stream; traffic: " + (ksrc or "none")}
    if kc and kc.get("hw_fp64_flops_per_dispatch"):
        hw = kc["hw_fp64_flops_per_dispatch"] * dom_n / (dom_ms * 1e-3) / 1e12
        line["roofline"]["hw_fp64_tflops"] = hw
        line["roofline"]["hw_frac"] = hw / peak
    if args.surftime_steps > 0:
        line["surftime"] = surftime_leg(args.surftime_steps)
    if args.fmst_steps > 0:
        nthr = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        line["fmst"] = fmst_leg(args.fmst_steps, 0 if args.no_cpu else 96, nthr)
    if not args.no_cpu and args.cpu_sample > 0:
        nthr = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        res, ncols, ncp = cpu_baseline(w, args.cpu_sample, nthr)
        kind = "reference" if "ref" in res else "port"
        v, dt = res["ref" if kind == "reference" else "rest"]
        line["cpu_baseline"] = {
            "value": float(v), "unit": "column·periods/s", "cores": nthr, "kind": kind,
            "sample": f"{ncols} strided columns of the same workload ({ncp} column-periods, "
                      f"{dt:.1f} s, OpenMP over columns on all {nthr} host threads this process "
                      "may use)"}
        if kind == "reference":
            line["cpu_baseline"]["restatement_value"] = float(res["rest"][0])
            line["cpu_baseline"]["restatement_over_reference"] = float(res["rest"][0] / v)
            line["cpu_baseline"]["note"] = (
                "reference timed only: its fixed-form surfdisp96.f leaves the threadprivate "
                "lines of del1st/dhalf as comments, so OpenMP columns race on them")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
