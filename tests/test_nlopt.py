"""The optimiser oracle (oracle/nlopt_oracle.py, NonlinOPT of
include/nonlinear_opt.hpp:15-390) pinned by what the mathematics fixes:
the L-BFGS two-loop recursion is the BFGS inverse-Hessian product, the
Polak-Ribiere directions with exact line searches on a quadratic are
mutually conjugate and finish in n steps, and full updates on a log-space
least-squares problem descend to its solution with both Wolfe conditions
met by every accepted step.  Plus the C-ABI symbols of the device path."""
import numpy as np
import pytest

from oracle.nlopt_oracle import NonlinOPTOracle, fdot
from tests.nlopt_problem import LogQuad


@pytest.mark.parametrize("it,start", [(1, 0), (3, 0), (5, 0), (7, 0), (7, 4)])
def test_two_loop_is_the_bfgs_inverse_hessian_product(it, start):
    rng = np.random.default_rng(it + 10 * start)
    n = 12
    B = rng.standard_normal((n, n))
    B = B @ B.T + n * np.eye(n)                       # SPD: y^T s > 0 for every pair
    o = NonlinOPTOracle(0.03, start, it)
    x = {k: rng.standard_normal(n).astype(np.float32) for k in range(start, it + 1)}
    g0 = rng.standard_normal(n)
    for k in range(start, it + 1):
        o.logmd_h[k] = x[k]
        o.grad_h[k] = (B @ x[k].astype(np.float64) + g0).astype(np.float32)
    got = o.lbfgs_direction(it)
    lo = max(it - 5, start)
    S = [x[k + 1].astype(np.float64) - x[k] for k in range(lo, it)]
    Y = [o.grad_h[k + 1].astype(np.float64) - o.grad_h[k] for k in range(lo, it)]
    H = (S[-1] @ Y[-1]) / (Y[-1] @ Y[-1]) * np.eye(n)
    for s, y in zip(S, Y):                            # oldest pair first
        r = 1.0 / (y @ s)
        V = np.eye(n) - r * np.outer(y, s)
        H = V.T @ H @ V + r * np.outer(s, s)
    want = -H @ o.grad_h[it].astype(np.float64)
    assert np.allclose(got, want, rtol=2e-4, atol=2e-4 * np.abs(want).max())


def test_polak_ribiere_with_exact_line_search_is_linear_cg():
    rng = np.random.default_rng(5)
    n = 6
    Q = rng.standard_normal((n, n))
    H = Q @ Q.T + np.eye(n)
    b = rng.standard_normal(n)
    o = NonlinOPTOracle(1.0, 0, 0)
    u = np.zeros(n)
    dirs = []
    for k in range(n):
        g = (H @ u - b).astype(np.float32)
        o.grad_h[k] = g
        d = o.cg_direction(k, g).astype(np.float64)
        dirs.append(d)
        u = u - (g @ d) / (d @ H @ d) * d
    for k in range(1, n):
        for j in range(k):
            c = dirs[k] @ H @ dirs[j] / (np.linalg.norm(dirs[k]) * np.linalg.norm(H @ dirs[j]))
            assert abs(c) < 1e-4
    assert np.linalg.norm(H @ u - b) < 1e-4 * np.linalg.norm(b)


@pytest.mark.parametrize("method,iters", [("LBFGS", 25), ("CG", 30)])
def test_updates_descend_to_the_least_squares_solution(method, iters):
    p = LogQuad(n=60, m=80)
    o = NonlinOPTOracle(0.2, 0, 0)
    x = p.x0.copy()
    f_prev = None
    for _ in range(iters):
        x_old = x.copy()
        x, dsyn = o.update(p, x, method)
        f0, f1, q, q1, step = o.accepted
        # the accepted model is the one whose misfit and gradient were tested
        ds, g = p.grad_np(x)
        assert f1 == p.misfit_np(ds) and fdot(g, g) >= 0
        m2 = np.float32(0.1 if method == "CG" else 0.9)
        assert f1 <= f0 + np.float32(1e-4) * step * q and q1 >= m2 * q
        assert np.all(np.abs(np.log(x) - np.log(x_old)) <= 0.2 * 1.0001)
        assert f_prev is None or f0 <= f_prev
        f_prev = f1
    u_ls = np.linalg.lstsq(p.A, p.d, rcond=None)[0]
    err = np.abs(np.log(x.astype(np.float64)) - u_ls).max()
    assert err < (2e-3 if method == "LBFGS" else 2e-2), err
    assert o.ITER == iters


def test_library_exports_the_optimiser():
    from jdsurfg_amd import nlopt
    L = nlopt._L()
    assert L.jdsg_nlopt_work_floats(300, 400) >= 20 * 300 + 2 * 400
    for name in ("jdsg_nlopt_create", "jdsg_nlopt_update", "jdsg_nlopt_state",
                 "jdsg_nlopt_destroy", "jdsg_nlopt_last_error"):
        assert hasattr(L, name)


def test_capped_step_that_never_meets_the_curvature_condition_stops():
    # the step grows tenfold after a curvature failure and is capped back
    # at max_step / dmax: the reference's loop would not end
    p = LogQuad()
    o = NonlinOPTOracle(0.03, 0, 0)
    with pytest.raises(RuntimeError, match="repeats"):
        o.update(p, p.x0.copy(), "CG")
    # the first trial is at the cap; its tenfold is capped back to it: stop
    assert o.trials == 1 and o.STEP_FAC * np.float32(10) > 0


def test_create_refuses_a_resumed_history():
    """The history ring starts empty, so a create with iter_current past
    iter_start is refused before any device work (an equal pair starts
    fresh at that iteration)."""
    import ctypes as C
    from jdsurfg_amd import nlopt
    L = nlopt._L()
    h = C.c_void_p()
    rc = L.jdsg_nlopt_create(300, 400, 0.3, 0, 3, None, None, C.byref(h))
    assert rc != 0 and b"iter_current" in L.jdsg_nlopt_last_error()
