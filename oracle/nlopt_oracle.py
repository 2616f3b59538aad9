"""Oracle -- test infrastructure only, never imported by the product path.

numpy restatement of NonlinOPT<T> (include/nonlinear_opt.hpp:15-390): the
CG (Polak-Ribiere) and L-BFGS directions in log-model space, the angle
check, the trial step, the multiplicative update and the Wolfe line search,
with the reference's float arithmetic (numpy float32 scalars and arrays)
and its file history kept in dicts keyed by iteration.  Inner products are
float64 sums of float32 differences rounded to float32, as the device path
forms them (the reference sums in float in Eigen's order).

Pinned by closed forms in tests/test_nlopt.py: the two-loop recursion
equals the textbook BFGS inverse-Hessian product for 1..5 pairs, the CG
direction equals -g + beta d with the Polak-Ribiere beta, and the whole
update converges to the least-squares solution of a log-space quadratic
with every accepted step meeting both Wolfe conditions."""
import math

import numpy as np

F = np.float32


def fdot(a, b):
    return F(np.dot(np.asarray(a, np.float64), np.asarray(b, np.float64)))


class NonlinOPTOracle:
    MSTORE = 5

    def __init__(self, max_step, iter_start, iter_current):
        self.MAX_STEP = F(max_step)
        self.ITER_START = iter_start
        self.ITER = iter_current
        self.ITER_LS = 0
        self.FLAG = "INIT"
        self.STEP_FAC = F(-1.0)
        self.STEP_L = F(0.0)
        self.STEP_R = F(0.0)
        self.grad_h, self.logmd_h, self.direc_h, self.syn_h = {}, {}, {}, {}
        self.trials = 0

    # nonlinear_opt.hpp:65-90
    def cg_direction(self, it, grad):
        if it == self.ITER_START:
            direc = -grad
        else:
            gp, dp = self.grad_h[it - 1], self.direc_h[it - 1]
            beta = fdot(grad, grad - gp) / fdot(gp, gp)
            direc = -grad + beta * dp
        self.direc_h[it] = direc.copy()
        return direc

    # nonlinear_opt.hpp:92-188
    def lbfgs_direction(self, it):
        if it == self.ITER_START:
            return -self.grad_h[it]
        iter_store = max(it - self.MSTORE, self.ITER_START)
        a, p = {}, {}
        q = self.grad_h[it].copy()
        for s in range(it - 1, iter_store - 1, -1):
            gd = self.grad_h[s + 1] - self.grad_h[s]
            xd = self.logmd_h[s + 1] - self.logmd_h[s]
            p[s] = F(1.0 / float(fdot(gd, xd)))
            a[s] = p[s] * fdot(xd, q)
            q = q - a[s] * gd
        gd = self.grad_h[it] - self.grad_h[it - 1]
        xd = self.logmd_h[it] - self.logmd_h[it - 1]
        r = (fdot(gd, xd) / fdot(gd, gd)) * q
        for s in range(iter_store, it):
            gd = self.grad_h[s + 1] - self.grad_h[s]
            xd = self.logmd_h[s + 1] - self.logmd_h[s]
            b = p[s] * fdot(gd, r)
            r = r + xd * (a[s] - b)
        return -r

    # nonlinear_opt.hpp:201-268
    def next_model(self, model, grad, x, chi, method):
        it = self.ITER
        direc = self.cg_direction(it, grad) if method == "CG" else self.lbfgs_direction(it)
        if it != self.ITER_START:
            l1, l2 = np.sqrt(fdot(direc, direc)), np.sqrt(fdot(grad, grad))
            ca = F(-float(np.dot(np.asarray(direc, np.float64), np.asarray(grad, np.float64)))
                   / (float(l1) * float(l2)))
            with np.errstate(invalid="ignore"):
                theta = F(float(np.arccos(ca)) * 180.0 / math.pi)
            if theta > 92:
                direc = -grad
                self.ITER_START = it
        direc = model.smoothing(direc)
        dmax = F(np.max(np.abs(direc)))
        step_fac = self.STEP_FAC
        if it == self.ITER_START and self.ITER_LS == 0:
            step_fac = F(-1.0)
        if it == self.ITER_START + 1 and self.ITER_LS == 0:
            if method == "CG":
                step_fac = F(-2.0 / float(fdot(grad, direc)) * float(chi))
            else:
                step_fac = F(1.0)
        if step_fac < 0 or step_fac * dmax > self.MAX_STEP:
            step_fac = self.MAX_STEP / dmax
        self.STEP_FAC = F(step_fac)
        x1 = (x * np.exp(self.STEP_FAC * direc)).astype(F)
        if self.FLAG != "LS":
            self.FLAG = "LS"
            self.ITER_LS = 0
        return x1, direc

    # nonlinear_opt.hpp:271-389; returns (x, dsyn)
    def update(self, model, x, method="CG"):
        x = np.asarray(x, F)
        self.trials = 0
        if self.FLAG == "INIT":
            dsyn, grad = model.compute_grad(x)
            self.grad_h[self.ITER] = grad
            self.logmd_h[self.ITER] = np.log(x)
        else:
            grad = self.grad_h[self.ITER]
            dsyn = self.syn_h[self.ITER]
        fcost = F(model.compute_misfit(dsyn))
        self.fcost = fcost
        x1, direc = self.next_model(model, grad, x, fcost, method)
        m1, m2 = F(1.0e-4), F(0.1 if method == "CG" else 0.9)
        while self.FLAG == "LS":
            self.trials += 1
            if self.trials > 100:
                raise RuntimeError("no step met the Wolfe conditions")
            dsyn1, grad1 = model.compute_grad(x1)
            if self.ITER_LS == 0:
                self.STEP_L = self.STEP_R = F(0.0)
            fcost1 = F(model.compute_misfit(dsyn1))
            self.fcost1 = fcost1
            q, q1 = fdot(grad, direc), fdot(grad1, direc)
            cond1 = fcost1 <= fcost + m1 * self.STEP_FAC * q
            cond2 = q1 >= m2 * q
            if cond1 and cond2:
                self.FLAG = "GRAD"
                self.ITER += 1
                self.ITER_LS = 0
                self.grad_h[self.ITER] = grad1
                self.syn_h[self.ITER] = dsyn1
                x = x1
                self.logmd_h[self.ITER] = np.log(x)
                self.accepted = (fcost, fcost1, q, q1, self.STEP_FAC)
                break
            elif not cond1:
                self.STEP_R = self.STEP_FAC
                self.STEP_FAC = F(0.5 * float(F(self.STEP_R + self.STEP_L)))
                self.ITER_LS += 1
            else:
                self.STEP_L = self.STEP_FAC
                tried, grow = self.STEP_FAC, self.STEP_R == 0
                if not grow:
                    self.STEP_FAC = F(0.5 * float(F(self.STEP_L + self.STEP_R)))
                else:
                    self.STEP_FAC = F(self.STEP_FAC * F(10.0))
                self.ITER_LS += 1
                x1, direc = self.next_model(model, grad, x, fcost, method)
                # the tenfold step capped back to the one just tried repeats
                # forever in the reference (nonlinear_opt.hpp:360-380)
                if grow and self.STEP_FAC == tried:
                    raise RuntimeError("no step met the Wolfe conditions (the capped trial "
                                       "step repeats)")
                continue
            x1, direc = self.next_model(model, grad, x, fcost, method)
        return x, dsyn
