"""The device-resident optimiser (jdsg_nlopt_*, NonlinOPT of
include/nonlinear_opt.hpp:15-390) against the oracle on the same
log-space least-squares problem: after every update the models agree to
float rounding, and the iteration counters, the line-search trial counts
and the restarts (ITER_START) are the same.  Device expf/logf and numpy's
differ by an ulp here and there; the cases run only as many updates as the
oracle itself stays within these tolerances when its starting model moves
by one ulp (CG with a 100 % cap on this problem amplifies such a change to
1e-3 in the trial step by its seventh update, and splits after that).
Tolerances: x 2e-5 relative, dsyn 1e-4, trial step 1e-4."""
import numpy as np
import pytest

from oracle.nlopt_oracle import NonlinOPTOracle
from tests.nlopt_problem import LogQuad, TorchLogQuad

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method,max_step,iters", [("CG", 0.3, 12), ("CG", 1.0, 6),
                                                    ("LBFGS", 0.03, 10), ("LBFGS", 1.0, 12)])
def test_device_optimiser_matches_the_oracle(method, max_step, iters):
    import torch
    from jdsurfg_amd.nlopt import NonlinOPT
    print(f"nlopt {method} max_step={max_step} iters={iters}", flush=True)
    p = LogQuad()
    o = NonlinOPTOracle(max_step, 0, 0)
    dev = NonlinOPT(p.n, p.m, max_step, 0, 0)
    tp = TorchLogQuad(p)
    xo = p.x0.copy()
    xd = torch.from_numpy(p.x0).cuda()
    dd = torch.zeros(p.m, dtype=torch.float32, device="cuda")
    for k in range(iters):
        xo, dso = o.update(p, xo, method)
        st = dev.update(tp, xd, dd, method)
        torch.cuda.synchronize()
        assert (st["iter"], st["iter_start"], st["trials"]) == (o.ITER, o.ITER_START, o.trials), k
        assert np.allclose(xd.cpu().numpy(), xo, rtol=2e-5, atol=0), k
        assert np.allclose(dd.cpu().numpy(), dso, rtol=1e-4, atol=1e-4), k
        assert abs(st["step_fac"] - o.STEP_FAC) <= 1e-4 * abs(o.STEP_FAC), k


def test_capped_step_that_fails_the_curvature_condition_stops_both():
    """CG with a 3 % cap on this problem: the capped step never meets the
    curvature condition and the reference's line search would not end; the
    oracle and the device path both stop at the first repeat of the capped
    trial step, after the one trial at the cap."""
    import torch
    from jdsurfg_amd.nlopt import NonlinOPT
    print("nlopt CG max_step=0.03 stuck line search", flush=True)
    p = LogQuad()
    o = NonlinOPTOracle(0.03, 0, 0)
    with pytest.raises(RuntimeError, match="Wolfe"):
        o.update(p, p.x0.copy(), "CG")
    assert o.trials == 1
    dev = NonlinOPT(p.n, p.m, 0.03, 0, 0)
    xd = torch.from_numpy(p.x0).cuda()
    dd = torch.zeros(p.m, dtype=torch.float32, device="cuda")
    with pytest.raises(RuntimeError, match="Wolfe"):
        dev.update(TorchLogQuad(p), xd, dd, "CG")
    assert dev.state()["trials"] == 1
    assert torch.equal(xd.cpu(), torch.from_numpy(p.x0))


def test_restart_at_a_later_iteration_matches_the_oracle():
    """A run that starts at iteration 2 (iter_start = iter_current = 2, an
    empty history) updates as the oracle started there."""
    import torch
    from jdsurfg_amd.nlopt import NonlinOPT
    print("nlopt LBFGS started at iteration 2", flush=True)
    p = LogQuad()
    o = NonlinOPTOracle(0.3, 2, 2)
    dev = NonlinOPT(p.n, p.m, 0.3, 2, 2)
    tp = TorchLogQuad(p)
    xo = p.x0.copy()
    xd = torch.from_numpy(p.x0).cuda()
    dd = torch.zeros(p.m, dtype=torch.float32, device="cuda")
    for k in range(4):
        xo, dso = o.update(p, xo, "LBFGS")
        st = dev.update(tp, xd, dd, "LBFGS")
        torch.cuda.synchronize()
        assert (st["iter"], st["iter_start"], st["trials"]) == (o.ITER, o.ITER_START, o.trials), k
        assert np.allclose(xd.cpu().numpy(), xo, rtol=2e-5, atol=0), k
