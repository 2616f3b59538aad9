"""Learned root-bracket surrogate for the phase-velocity roots.

An MLP (192 inputs -> 4 x 512 GELU -> 1) predicts the flat phase velocity of
every (column, wave, mode, period) from a 48-point Vs/Vp/rho profile, ln T,
the wave type and the mode.  Weights and activations are bf16, products run
on the MFMA units with fp32 accumulation; an fp32 master copy and Adam
moments are kept for training.  All arithmetic runs in this library's HIP
kernels (swd_mlp.hip, swd_surrogate.hip); PyTorch only holds the buffers.

Training data come from the plain GPU solver (the reference method) on
seeded random layered columns.  At solve time each predicted root is
bracketed at +-half_width, checked for a sign change of the period equation,
polished in FP64 to the exact root, and the roots that fail go back to the
plain chain (see Plan.execute_surrogate).
"""
import ctypes as C
import os
import time

import numpy as np

from . import swd
from . import workloads as W

FEAT, HID, OUTP = 192, 512, 64
DIMS = [(HID, FEAT), (HID, HID), (HID, HID), (HID, HID), (OUTP, HID)]
C_SHIFT, C_SCALE = 3.5, 1.0


def _L():
    L = swd.lib()
    vp = C.c_void_p
    L.jdsg_gemm_bf16.argtypes = [C.c_int] * 5 + [vp, C.c_int, vp, C.c_int, vp, C.c_int, C.c_int,
                                                 vp, vp, C.c_int, vp]
    L.jdsg_gelu_bwd.argtypes = [C.c_longlong, vp, vp, vp, vp, vp]
    L.jdsg_colsum.argtypes = [C.c_int, C.c_int, vp, vp, vp]
    L.jdsg_mse.argtypes = [C.c_int, C.c_int, vp, vp, vp, vp, vp, vp]
    L.jdsg_adam.argtypes = [C.c_longlong, vp, vp, vp, vp, C.c_float, C.c_float, C.c_float,
                            C.c_float, C.c_int, C.c_float, vp, vp]
    L.jdsg_f32_to_bf16.argtypes = [C.c_longlong, vp, vp, vp]
    L.jdsg_plan_features.argtypes = [vp] * 6 + [C.c_longlong, C.c_longlong, vp, vp, vp, C.c_int,
                                                vp]
    L.jdsg_plan_roots.argtypes = [vp] * 9
    L.jdsg_denorm.argtypes = [C.c_longlong, vp, C.c_int, C.c_double, C.c_double, vp,
                              C.c_longlong, vp]
    L.jdsg_plan_enable_surrogate.argtypes = [vp]
    L.jdsg_plan_execute_surrogate.argtypes = [vp] * 7 + [C.c_double] + [vp] * 12
    L.jdsg_plan_surrogate_stats.argtypes = [vp, C.POINTER(C.c_longlong),
                                            C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
    return L


def _p(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _chk(rc, what):
    if rc:
        raise RuntimeError(f"{what}: {swd.last_error()}")


class MLP:
    """The surrogate network; weights live on the GPU."""

    def __init__(self, seed=0):
        import torch
        g = torch.Generator().manual_seed(seed)
        self.W, self.b = [], []
        for li, (o, i) in enumerate(DIMS):
            w = torch.randn(o, i, generator=g) * (2.0 / i) ** 0.5
            if li == len(DIMS) - 1:
                w = w * 0.1
                w[1:] = 0.0   # padded output rows stay zero
            self.W.append(w.cuda().contiguous())
            self.b.append(torch.zeros(o, device="cuda"))
        self.Wbf = [w.to(torch.bfloat16) for w in self.W]
        self.mW = [torch.zeros_like(w) for w in self.W]
        self.vW = [torch.zeros_like(w) for w in self.W]
        self.mb = [torch.zeros_like(b) for b in self.b]
        self.vb = [torch.zeros_like(b) for b in self.b]
        self.gW = [torch.zeros_like(w) for w in self.W]
        self.gb = [torch.zeros_like(b) for b in self.b]
        self.step = 0
        self._buf = {}

    def _bufs(self, M, train):
        import torch
        key = (M, train)
        if key in self._buf:
            return self._buf[key]
        bf = torch.bfloat16
        d = {"Y": [None] + [torch.empty(M, HID, dtype=bf, device="cuda") for _ in range(4)],
             "out": torch.zeros(M, OUTP, device="cuda")}
        if train:
            d["Z"] = [torch.empty(M, HID, device="cuda") for _ in range(4)]
            d["dOut"] = torch.zeros(M, OUTP, device="cuda")
            d["dOutbf"] = torch.zeros(M, OUTP, dtype=bf, device="cuda")
            d["dYf"] = torch.empty(M, HID, device="cuda")
            d["dZbf"] = torch.empty(M, HID, dtype=bf, device="cuda")
            d["dZf"] = torch.empty(M, HID, device="cuda")
            d["loss"] = torch.zeros((M + 255) // 256, device="cuda")
        self._buf[key] = d
        return d

    def forward(self, X, train=False, stream=None):
        """X: (M, 192) bf16 CUDA tensor, M a multiple of 64.  Returns the
        (M, 64) fp32 output (column 0 is the prediction)."""
        L = _L()
        M = X.shape[0]
        d = self._bufs(M, train)
        Y = d["Y"]
        Y[0] = X
        K = FEAT
        for li in range(4):
            if train:
                _chk(L.jdsg_gemm_bf16(0, 1, M, HID, K, _p(Y[li]), K, _p(self.Wbf[li]), K,
                                      _p(d["Z"][li]), HID, 1, _p(self.b[li]), _p(Y[li + 1]), HID,
                                      stream), "gemm")
            else:
                _chk(L.jdsg_gemm_bf16(0, 1, M, HID, K, _p(Y[li]), K, _p(self.Wbf[li]), K, None,
                                      HID, 4, _p(self.b[li]), _p(Y[li + 1]), HID, stream), "gemm")
            K = HID
        _chk(L.jdsg_gemm_bf16(0, 1, M, OUTP, HID, _p(Y[4]), HID, _p(self.Wbf[4]), HID,
                              _p(d["out"]), OUTP, 2, _p(self.b[4]), None, 0, stream), "gemm")
        return d["out"]

    def loss_backward(self, target, weight=None, stream=None):
        """MSE on column 0 against target (M,) fp32; fills the gradients and
        returns the loss tensor (partial sums)."""
        L = _L()
        M = target.shape[0]
        d = self._bufs(M, True)
        _chk(L.jdsg_mse(M, OUTP, _p(d["out"]), _p(target), _p(weight), _p(d["dOut"]),
                        _p(d["loss"]), stream), "mse")
        _chk(L.jdsg_f32_to_bf16(M * OUTP, _p(d["dOut"]), _p(d["dOutbf"]), stream), "cvt")
        Y = d["Y"]
        # output layer
        _chk(L.jdsg_gemm_bf16(1, 0, OUTP, HID, M, _p(d["dOutbf"]), OUTP, _p(Y[4]), HID,
                              _p(self.gW[4]), HID, 0, None, None, 0, stream), "gemm dW")
        _chk(L.jdsg_colsum(M, OUTP, _p(d["dOut"]), _p(self.gb[4]), stream), "colsum")
        _chk(L.jdsg_gemm_bf16(0, 0, M, HID, OUTP, _p(d["dOutbf"]), OUTP, _p(self.Wbf[4]), HID,
                              _p(d["dYf"]), HID, 0, None, None, 0, stream), "gemm dX")
        for li in range(3, -1, -1):
            K = FEAT if li == 0 else HID
            _chk(L.jdsg_gelu_bwd(M * HID, _p(d["dYf"]), _p(d["Z"][li]), _p(d["dZbf"]),
                                 _p(d["dZf"]), stream), "gelu bwd")
            _chk(L.jdsg_gemm_bf16(1, 0, HID, K, M, _p(d["dZbf"]), HID, _p(Y[li]), K,
                                  _p(self.gW[li]), K, 0, None, None, 0, stream), "gemm dW")
            _chk(L.jdsg_colsum(M, HID, _p(d["dZf"]), _p(self.gb[li]), stream), "colsum")
            if li > 0:
                _chk(L.jdsg_gemm_bf16(0, 0, M, HID, HID, _p(d["dZbf"]), HID, _p(self.Wbf[li]),
                                      HID, _p(d["dYf"]), HID, 0, None, None, 0, stream),
                     "gemm dX")
        return d["loss"]

    def adam(self, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0, stream=None):
        L = _L()
        self.step += 1
        for li in range(len(DIMS)):
            _chk(L.jdsg_adam(self.W[li].numel(), _p(self.W[li]), _p(self.mW[li]),
                             _p(self.vW[li]), _p(self.gW[li]), lr, b1, b2, eps, self.step, wd,
                             _p(self.Wbf[li]), stream), "adam")
            _chk(L.jdsg_adam(self.b[li].numel(), _p(self.b[li]), _p(self.mb[li]),
                             _p(self.vb[li]), _p(self.gb[li]), lr, b1, b2, eps, self.step, 0.0,
                             None, stream), "adam")

    def save(self, path):
        np.savez(path, **{f"W{i}": w.cpu().numpy() for i, w in enumerate(self.W)},
                 **{f"b{i}": b.cpu().numpy() for i, b in enumerate(self.b)})

    @classmethod
    def load(cls, path):
        import torch
        m = cls()
        z = np.load(path)
        for i in range(len(DIMS)):
            m.W[i].copy_(torch.from_numpy(z[f"W{i}"]))
            m.b[i].copy_(torch.from_numpy(z[f"b{i}"]))
            m.Wbf[i] = m.W[i].to(torch.bfloat16)
        return m


def training_columns(ncol, seed):
    """Seeded random layered columns spanning the benchmark families:
    40-70 layers of 0.8-1.6 km, the depth trend with a smooth random field,
    a +-8 % checkerboard-like step or none, optional soft sediments; 40
    log-uniform periods in 1-80 s; Rayleigh modes 0-3, Love modes 0-1."""
    rng = np.random.default_rng(seed)
    nl = rng.integers(40, 71, ncol)
    dz = rng.uniform(0.8, 1.6, ncol)
    layer_off = np.concatenate([[0], np.cumsum(nl + 1)]).astype(np.int64)
    tot = int(layer_off[-1])
    thk = np.empty(tot, np.float32)
    vs = np.empty(tot, np.float32)
    K = 40
    periods = np.sort(np.exp(rng.uniform(0.0, np.log(80.0), (ncol, K))), axis=1)
    amp = rng.uniform(0.0, 0.12, ncol)
    wl = rng.uniform(10.0, 60.0, ncol)
    ph = rng.uniform(0, 2 * np.pi, ncol)
    step = rng.uniform(-0.08, 0.08, ncol) * (rng.random(ncol) < 0.5)
    zstep = rng.uniform(5, 60, ncol)
    sed = rng.random(ncol) < 0.3
    hsed = rng.uniform(0.0, 4.0, ncol)
    vsed = rng.uniform(0.8, 2.0, ncol)
    for c in range(ncol):
        n = nl[c]
        z = np.concatenate([(np.arange(n) + 0.5) * dz[c], [n * dz[c]]])
        v = (2.6 + 0.035 * z) * (1 + amp[c] * np.sin(2 * np.pi * z / wl[c] + ph[c]))
        v = v * (1 + step[c] * np.where(z < zstep[c], 1.0, -1.0))
        if sed[c]:
            v = np.where(z < hsed[c], vsed[c], v)
        v = np.clip(v, 0.6, 4.8)
        s = slice(layer_off[c], layer_off[c + 1])
        thk[s] = dz[c]
        thk[layer_off[c + 1] - 1] = 0.0
        vs[s] = v
    vp, rho = W.brocher(vs)
    return W.Workload("train", layer_off, thk, vp, vs, rho,
                      np.arange(ncol + 1) * K, periods.ravel(),
                      np.tile(np.array([4, 2], np.int32), (ncol, 1)), group=False)


def make_dataset(w):
    """Solve the roots of w with the plain GPU solver; returns the plan, its
    inputs and the flat roots (0 where none) as CUDA tensors."""
    import torch
    plan = swd.Plan(w.layer_off, w.per_off, w.nmode, False, "csr")
    inp = swd.upload(w)
    roots = plain_roots(plan, inp)
    torch.cuda.synchronize()
    return plan, inp, roots


class Features:
    """Feature builder bound to a plan and its inputs."""

    def __init__(self, plan, inp):
        import torch
        self.plan, self.inp = plan, inp
        self.colf = torch.empty(plan.ncol * 144, device="cuda")
        self.first = True

    def rows(self, rows=None, e0=0, n=None, X=None, stream=None):
        import torch
        L = _L()
        n = len(rows) if rows is not None else n
        if X is None:
            X = torch.empty(n, FEAT, dtype=torch.bfloat16, device="cuda")
        _chk(L.jdsg_plan_features(self.plan.h, *[_p(t) for t in self.inp], e0, n, _p(rows),
                                  _p(X), _p(self.colf), int(self.first), stream), "features")
        self.first = False
        return X


def train(ncol=200000, steps=3000, batch=8192, lr=1e-3, seed=7, log=print, valid_cols=4000):
    """Generate data with the plain GPU solver and train the MLP on the GPU.
    Returns (mlp, report)."""
    import torch
    t0 = time.time()
    w = training_columns(ncol + valid_cols, seed)
    plan, inp, roots = make_dataset(w)
    t_data = time.time() - t0
    ok = torch.nonzero(roots > 0).squeeze(1)
    # the last valid_cols columns are held out
    e_split = int(plan.ent_off[ncol])
    tr = ok[ok < e_split]
    va = ok[ok >= e_split]
    feats = Features(plan, inp)
    mlp = MLP(seed)
    g = torch.Generator(device="cuda").manual_seed(seed)
    t1 = time.time()
    losses = []
    for it in range(steps):
        idx = tr[torch.randint(0, tr.numel(), (batch,), device="cuda", generator=g)]
        X = feats.rows(rows=idx.contiguous())
        tgt = ((roots[idx] - C_SHIFT) / C_SCALE).float().contiguous()
        mlp.forward(X, train=True)
        loss = mlp.loss_backward(tgt)
        frac = it / max(1, steps - 1)
        lr_t = lr * (0.02 + 0.98 * 0.5 * (1 + np.cos(np.pi * frac)))
        mlp.adam(lr_t)
        if it % 250 == 0 or it == steps - 1:
            lv = float(loss.sum().item()) / batch
            losses.append((it, lv))
            log(f"train step {it} loss {lv:.3e}")
    torch.cuda.synchronize()
    t_train = time.time() - t1
    rep = evaluate(mlp, feats, roots, va)
    rep.update({"train_columns": ncol, "train_entries": int(tr.numel()), "steps": steps,
                "batch": batch, "t_data_s": t_data, "t_train_s": t_train, "losses": losses})
    return mlp, rep


def predict(mlp, feats, rows=None, e0=0, n=None, chunk=1 << 20):
    """Predicted flat c for entries (rows, or e0..e0+n) as a float64 tensor."""
    import torch
    L = _L()
    if rows is not None:
        n = rows.numel()
    out = torch.empty(n, dtype=torch.float64, device="cuda")
    Xb = None
    for s in range(0, n, chunk):
        m = min(chunk, n - s)
        mp = (m + 63) // 64 * 64
        if Xb is None or Xb.shape[0] != mp:
            Xb = torch.zeros(mp, FEAT, dtype=torch.bfloat16, device="cuda")
        if rows is not None:
            r = rows[s:s + m]
            if mp > m:
                r = torch.cat([r, r[-1:].expand(mp - m)])
            feats.rows(rows=r.contiguous(), X=Xb)
        else:
            feats.rows(e0=e0 + s, n=m, X=Xb[:m])
        o = mlp.forward(Xb, train=False)
        _chk(L.jdsg_denorm(m, _p(o), OUTP, C_SCALE, C_SHIFT, _p(out), s, None), "denorm")
    return out


def evaluate(mlp, feats, roots, rows):
    """Held-out accuracy: relative error of the predicted roots and the share
    inside +-0.5 % (the bracket)."""
    pred = predict(mlp, feats, rows=rows)
    c = roots[rows]
    rel = ((pred - c).abs() / c).cpu().numpy()
    return {"valid_entries": int(rows.numel()), "rel_err_median": float(np.median(rel)),
            "rel_err_p90": float(np.percentile(rel, 90)),
            "rel_err_p99": float(np.percentile(rel, 99)),
            "within_0.5pct": float(np.mean(rel < 0.005)),
            "within_0.25pct": float(np.mean(rel < 0.0025))}


def plain_roots(plan, inp, stream=None):
    """Flat roots (nvar*nent float64, 0 where none) of the plain GPU solver at
    the plan's tolerance."""
    import torch
    L = _L()
    nvar = 3 if plan.group else 1
    roots = torch.zeros(nvar * plan.nent, dtype=torch.float64, device="cuda")
    _chk(L.jdsg_plan_roots(plan.h, *[_p(t) for t in inp], _p(roots), None, stream), "roots")
    return roots


def surrogate_solve(plan, inp, out, mlp=None, pred=None, half_width=0.005, stream=None):
    """The learned-bracket solve on a plan in exact-root mode (tol < 1e-6):
    bracket + FP64 polish of every root around its predicted value (the MLP's,
    or `pred`, a float64 CUDA tensor of nent flat roots), plain chains for the
    roots left open, then the eigen/kernel stage.  Returns `out`."""
    import torch
    L = _L()
    _chk(L.jdsg_plan_enable_surrogate(plan.h), "enable surrogate")
    if pred is None:
        if not hasattr(plan, "_feats"):
            plan._feats = Features(plan, inp)
        plan._feats.first = True
        pred = predict(mlp, plan._feats, e0=0, n=plan.nent)
    ks = [_p(out.get(n)) for n in swd.KERNEL_NAMES]
    st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    _chk(L.jdsg_plan_execute_surrogate(plan.h, *[_p(t) for t in inp], _p(pred), half_width,
                                       _p(out["c"]), _p(out["u"]), _p(out["status"]), *ks,
                                       C.c_void_p(st)), "execute_surrogate")
    return out


def surrogate_stats(plan):
    L = _L()
    f, pe, ce = C.c_longlong(), C.c_longlong(), C.c_longlong()
    _chk(L.jdsg_plan_surrogate_stats(plan.h, C.byref(f), C.byref(pe), C.byref(ce)), "stats")
    return {"roots_left_to_chain": f.value, "polish_evals": pe.value, "chain_evals": ce.value}


DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights",
                               "surrogate_mlp.npz")
