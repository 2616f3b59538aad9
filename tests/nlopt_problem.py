"""A log-space least-squares problem for the optimiser tests: u = ln x,
dsyn = A u, misfit = 0.5 |dsyn - d|^2, grad = A^T (dsyn - d) (the gradient
with respect to ln x, the variable NonlinOPT steps in); smoothing scales the
direction by a fixed positive weight.  Every value is computed in float64
on the host and rounded to float32, so the optimisers under test see the
same numbers for the same x."""
import numpy as np


class LogQuad:
    def __init__(self, n=300, m=400, seed=3):
        rng = np.random.default_rng(seed)
        self.A = (np.eye(m, n) * 1.0 + rng.standard_normal((m, n)) * (0.4 / np.sqrt(n)))
        self.u_true = np.log(3.5 + 0.3 * rng.random(n))
        self.d = self.A @ self.u_true
        self.w = (0.75 + 0.5 * rng.random(n)).astype(np.float32)
        self.x0 = np.full(n, 3.3, np.float32)
        self.n, self.m = n, m

    def grad_np(self, x):
        u = np.log(np.asarray(x, np.float64))
        dsyn = self.A @ u
        return dsyn.astype(np.float32), (self.A.T @ (dsyn - self.d)).astype(np.float32)

    def misfit_np(self, dsyn):
        r = np.asarray(dsyn, np.float64) - self.d
        return np.float32(0.5 * np.dot(r, r))

    # the oracle's model interface
    def compute_grad(self, x):
        return self.grad_np(x)

    def compute_misfit(self, dsyn):
        return self.misfit_np(dsyn)

    def smoothing(self, direc):
        return (direc * self.w).astype(np.float32)


class TorchLogQuad:
    """The same problem behind the device optimiser's callbacks."""

    def __init__(self, prob):
        import torch
        self.p = prob
        self.w = torch.from_numpy(prob.w).cuda()

    def compute_grad(self, x, dsyn, grad):
        import torch
        ds, g = self.p.grad_np(x.cpu().numpy())
        dsyn.copy_(torch.from_numpy(ds))
        grad.copy_(torch.from_numpy(g))

    def compute_misfit(self, dsyn):
        return float(self.p.misfit_np(dsyn.cpu().numpy()))

    def smoothing(self, direc):
        direc.mul_(self.w)
