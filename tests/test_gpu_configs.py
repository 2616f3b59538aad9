"""The benchmark configurations on the GPU against the oracle, on strided
column subsets sized so the CPU oracle finishes in seconds.

* grid256 (256x256 columns x 64 layers of 1 km with the +-8 % checkerboard, 60
  periods 1-80 s, Rayleigh modes 0-3 + Love 0-1, group kernels): the plain
  path at the reference tolerance, the plain path in exact-root mode, and the
  shipped network's surrogate solve, each against the oracle.
* grid128 (128x128 columns x 48 layers of 1.5 km with sediments, 40 periods
  1-60 s, Rayleigh 0-2 + Love 0-1, group kernels): the same three, the
  surrogate with the 4 x 512 network configured for it, reporting its
  fallback rate.

Bars: status identical; c and U within 1e-6 and kernels within 1e-4 (of
the entry's largest layer value) of the oracle at the reference tolerance,
U and kernels checked at the GPU's roots where a root differs in its last
bits; exact-root mode and the surrogate against the exact-root oracle
(tolerance 1e-12 for both): c within 1e-8, U within 1e-7, dc/dm and dU/dm
within 1e-4.
"""
import numpy as np
import pytest

from jdsurfg_amd import workloads as W
from tests.compare import (KERNELS, differing_root_report, gpu_solve, kernel_errors, oracle_solve,
                           same_roots, scalar_errors)

pytestmark = pytest.mark.gpu
# exact-root stop rule: at 1e-10 the oracle's own dU/dm still moves by
# 2e-4 on the grid128 sediment columns when the tolerance is tightened (the
# difference quotient over T(1+-0.005) amplifies the root's last digits near
# mode osculations); at 1e-12 it moves by at most 5e-5
EXACT_TOL = 1e-12


def _c5():
    return W.grid256().subset(np.arange(0, 65536, 1021)[:64])


def _c4():
    return W.grid128().subset(np.arange(0, 16384, 257)[:64])


CASES = {"C5_grid256": _c5, "C4_grid128": _c4}
NETS = {"C5_grid256": None, "C4_grid128": "surrogate_mlp_512.npz"}


def _gpu_csr(plan, out, w):
    from jdsurfg_amd import swd
    g = {k: v.cpu().numpy() for k, v in out.items()}
    if plan.layout == 0:
        N = int(w.layer_off[1] - w.layer_off[0])
        for k in KERNELS:
            if k in g:
                g[k] = swd.grid_to_csr(plan, g[k], N)
    g["ent_off"] = plan.ent_off
    g["kern_off"] = plan.kern_off
    g["cflat"] = plan.flat_roots().ravel()
    return g


@pytest.mark.parametrize("name", list(CASES))
def test_plain_matches_oracle(name, oracle_rest):
    w = CASES[name]()
    print(f"case {name} plain: ncol={w.ncol} entries={int(w.nmode.sum(1) @ np.diff(w.per_off))}",
          flush=True)
    b = oracle_solve(oracle_rest, w, nthreads=16)
    a = gpu_solve(w)
    assert np.array_equal(a["status"], b["status"])
    same = same_roots(a, b, w)
    ok = b["status"] == 0
    frac = float(same[ok].mean())
    s_all = scalar_errors(a, b)
    s = scalar_errors(a, b, mask=same)
    k = kernel_errors(a, b, w, mask=same)
    rep, _ = differing_root_report(oracle_rest, w, a, b, nthreads=16)
    print(name, f"identical roots {frac:.5f}", s_all, s, k, rep, flush=True)
    assert s_all["c"] <= 1e-6 and frac >= 0.999 and s["u"] <= 1e-6
    assert max(k.values()) <= 1e-4
    if rep["n"]:
        assert rep["at_gpu_roots"]["u"] <= 1e-6
        assert max(v for n, v in rep["at_gpu_roots"].items() if n in KERNELS) <= 1e-4


def _exact_oracle(oracle_rest, w):
    oracle_rest.set_tolerance(EXACT_TOL)
    try:
        return oracle_solve(oracle_rest, w, nthreads=16)
    finally:
        oracle_rest.set_tolerance(1e-6)


def _check_exact(name, what, g, ref, w):
    assert np.array_equal(g["status"], ref["status"]), what
    s = scalar_errors(g, ref)
    k = kernel_errors(g, ref, w)
    print(name, what, s, {n: f"{v:.2e}" for n, v in k.items()}, flush=True)
    assert s["c"] <= 1e-8 and s["u"] <= 1e-7
    for n in ("dcda", "dcdb", "dcdr", "dcdh"):
        assert k[n] <= 1e-4, (n, k[n])
    return k


@pytest.mark.parametrize("name", list(CASES))
def test_exact_and_surrogate_match_exact_oracle(name, oracle_rest):
    import os

    import torch
    from jdsurfg_amd import swd, surrogate as S
    w = CASES[name]()
    print(f"case {name} exact + surrogate: ncol={w.ncol}", flush=True)
    ref = _exact_oracle(oracle_rest, w)
    plan = swd.Plan(w.layer_off, w.per_off, w.nmode, w.group, "grid")
    plan.set_tolerance(EXACT_TOL)
    inp = swd.upload(w)
    out = plan.alloc_outputs()
    plan.execute(*inp, out)
    torch.cuda.synchronize()
    ke = _check_exact(name, "exact-root plain", _gpu_csr(plan, out, w), ref, w)
    net = NETS[name]
    path = S.DEFAULT_WEIGHTS if net is None else os.path.join(os.path.dirname(S.DEFAULT_WEIGHTS),
                                                              net)
    mlp = S.MLP.load(path)
    sout = plan.alloc_outputs()
    S.surrogate_solve(plan, inp, sout, mlp=mlp)
    torch.cuda.synchronize()
    st = S.surrogate_stats(plan)
    nroot = (3 if w.group else 1) * plan.nent
    print(name, f"network width {mlp.hid}: fallback {st['roots_left_to_chain']} of {nroot} "
          f"({st['roots_left_to_chain'] / nroot:.4f})", st, flush=True)
    ks = _check_exact(name, "surrogate", _gpu_csr(plan, sout, w), ref, w)
    # dU/dm: the difference quotient over T(1+-0.005) (sregn96.f90:1831-1834)
    # divides root differences by 0.01 of the period
    for n in ("duda", "dudb", "dudr", "dudh"):
        assert ke[n] <= 1e-4 and ks[n] <= 1e-4, (n, ke[n], ks[n])
