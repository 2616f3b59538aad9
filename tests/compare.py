"""Shared comparison helpers for parity tests (GPU path vs the CPU oracle)."""
import numpy as np

KERNELS = ("dcda", "dcdb", "dcdr", "dcdh", "duda", "dudb", "dudr", "dudh")


def oracle_solve(oracle, w, nthreads=4):
    return oracle.batch_solve(w.layer_off, w.thk, w.vp, w.vs, w.rho, w.per_off, w.periods,
                              w.nmode.ravel(), w.group, nthreads=nthreads)


def gpu_solve(w, layout=None):
    """Run the HIP path; returns numpy dict in csr order (like the oracle)."""
    import torch
    from jdsurfg_amd import swd
    plan, out = swd.solve(w, layout=layout)
    res = {k: v.cpu().numpy() for k, v in out.items()}
    res["cflat"] = plan.flat_roots().ravel()
    if plan.layout == 0:
        N = int(w.layer_off[1] - w.layer_off[0])
        for k in KERNELS:
            res[k] = swd.grid_to_csr(plan, res[k], N)
    res["ent_off"] = plan.ent_off
    res["kern_off"] = plan.kern_off
    return res


def same_roots(a, b, w):
    """Entries whose flat roots (T and, with group kernels, T(1+-.005)) are
    bit-identical in both solutions."""
    ne = len(b["c"])
    nv = 3 if w.group else 1
    same = np.ones(ne, bool)
    for v in range(nv):
        same &= a["cflat"][v * ne:(v + 1) * ne] == b["cflat"][v * ne:(v + 1) * ne]
    return same


def kernel_errors(a, b, w, names=KERNELS, mask=None):
    """max over entries of max|a-b| / max|b| per entry (entries with status 0)."""
    ok = (a["status"] == 0) & (b["status"] == 0)
    if mask is not None:
        ok &= mask
    N = np.diff(w.layer_off)
    out = {}
    for n in names:
        worst = 0.0
        x, y = a[n], b[n]
        for c in range(w.ncol):
            e0, e1 = b["ent_off"][c], b["ent_off"][c + 1]
            nn = N[c]
            for e in range(e0, e1):
                if not ok[e]:
                    continue
                q = b["kern_off"][c] + (e - e0) * nn
                ys = y[q:q + nn]
                m = np.abs(ys).max()
                if m > 0:
                    worst = max(worst, float(np.abs(x[q:q + nn] - ys).max() / m))
        out[n] = worst
    return out


def scalar_errors(a, b, mask=None):
    ok = (a["status"] == 0) & (b["status"] == 0)
    if mask is not None:
        ok &= mask
    res = {}
    for n in ("c", "u"):
        if ok.any():
            r = np.abs(a[n][ok] - b[n][ok]) / np.abs(b[n][ok])
            res[n] = float(r.max())
            res[n + "_exact"] = float(np.mean(a[n][ok] == b[n][ok]))
        else:
            res[n] = 0.0
            res[n + "_exact"] = 1.0
    return res


def at_gpu_roots(oracle, w, a, nthreads=4):
    """The oracle's eigen stage evaluated at the GPU's flat roots: what the
    reference's sregn96_/sregnpu_/slegn96_/slegnpu_ return for those roots."""
    r = np.zeros(3 * len(a["c"]))
    r[:len(a["cflat"])] = a["cflat"]
    return oracle.batch_solve(w.layer_off, w.thk, w.vp, w.vs, w.rho, w.per_off, w.periods,
                              w.nmode.ravel(), w.group, nthreads=nthreads, roots=r)


def differing_root_report(oracle, w, a, b, nthreads=4):
    """Entries (status 0) whose flat roots differ between the GPU (a) and the
    oracle (b): U and kernel errors of the GPU against the oracle evaluated
    at the GPU's roots (the bar), and against the oracle's own roots (the
    spread the reference's 1e-6 stop band allows)."""
    same = same_roots(a, b, w)
    diff = (~same) & (a["status"] == 0) & (b["status"] == 0)
    rep = {"n": int(diff.sum())}
    if not diff.any():
        return rep, diff
    g = at_gpu_roots(oracle, w, a, nthreads)
    assert np.array_equal(g["status"][diff], a["status"][diff])
    rep["at_gpu_roots"] = scalar_errors(a, g, mask=diff)
    rep["at_gpu_roots"].update(kernel_errors(a, g, w, mask=diff))
    rep["vs_own_roots"] = scalar_errors(a, b, mask=diff)
    rep["vs_own_roots"].update(kernel_errors(a, b, w, mask=diff))
    return rep, diff
