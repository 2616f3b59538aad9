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

FEAT, HID, OUTP = 192, 512, 64   # inputs, default hidden width, padded output
C_SHIFT, C_SCALE = float(np.log(3.5)), 0.3   # target: (ln c - C_SHIFT) / C_SCALE


def _L():
    L = swd.lib()
    vp = C.c_void_p
    L.jdsg_gemm_bf16.argtypes = [C.c_int] * 5 + [vp, C.c_int, vp, C.c_int, vp, C.c_int, C.c_int,
                                                 vp, vp, C.c_int, vp]
    L.jdsg_gelu_bwd.argtypes = [C.c_longlong, vp, vp, vp, vp, vp]
    L.jdsg_colsum.argtypes = [C.c_int, C.c_int, vp, vp, vp]
    L.jdsg_mse.argtypes = [C.c_int, C.c_int, vp, vp, vp, C.c_float, vp, vp, vp]
    L.jdsg_adam.argtypes = [C.c_longlong, vp, vp, vp, vp, C.c_float, C.c_float, C.c_float,
                            C.c_float, C.c_int, C.c_float, vp, vp]
    L.jdsg_f32_to_bf16.argtypes = [C.c_longlong, vp, vp, vp]
    L.jdsg_plan_features.argtypes = [vp] * 6 + [C.c_longlong, C.c_longlong, vp, vp, vp, C.c_int,
                                                vp]
    L.jdsg_plan_roots.argtypes = [vp] * 9
    L.jdsg_denorm.argtypes = [C.c_longlong, vp, C.c_int, C.c_double, C.c_double, vp,
                              C.c_longlong, vp]
    L.jdsg_mlp_head.argtypes = [C.c_longlong, C.c_int, vp, vp, C.c_float, C.c_double, C.c_double,
                                vp, C.c_longlong, vp]
    L.jdsg_mlp_column_rows.argtypes = [C.c_int, vp, vp, vp]
    L.jdsg_plan_mlp_expand.argtypes = [vp, vp, C.c_longlong, C.c_longlong, vp, C.c_int, vp,
                                       C.c_int, vp, vp]
    L.jdsg_plan_mlp_fused.argtypes = [vp, vp, C.c_longlong, C.c_longlong, vp, C.c_int] + \
        [vp] * 7 + [vp, C.c_float, C.c_double, C.c_double, vp, C.c_longlong, vp, C.c_longlong,
                    vp]
    L.jdsg_mlp_fused_work.argtypes = [C.c_longlong]
    L.jdsg_mlp_fused_work.restype = C.c_longlong
    L.jdsg_plan_enable_surrogate.argtypes = [vp]
    L.jdsg_plan_scan_jobs.argtypes = [vp, C.POINTER(C.c_longlong)]
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
    """The surrogate network; weights live on the GPU.  `hid` is the width of
    the four hidden layers (a multiple of 128; the stored weights carry it)."""

    def __init__(self, seed=0, hid=HID):
        import torch
        if hid % 128:
            raise ValueError("hidden width must be a multiple of 128")
        self.hid = hid
        self.dims = [(hid, FEAT), (hid, hid), (hid, hid), (hid, hid), (OUTP, hid)]
        g = torch.Generator().manual_seed(seed)
        self.W, self.b = [], []
        for li, (o, i) in enumerate(self.dims):
            w = torch.randn(o, i, generator=g) * (2.0 / i) ** 0.5
            if li == len(self.dims) - 1:
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
        self._consts = None   # inference constants (ext, output bias), see inference_consts

    def _bufs(self, M, train):
        import torch
        key = (M, train)
        if key in self._buf:
            return self._buf[key]
        bf = torch.bfloat16
        H = self.hid
        d = {"Y": [None] + [torch.empty(M, H, dtype=bf, device="cuda") for _ in range(4)],
             "out": torch.zeros(M, OUTP, device="cuda")}
        if train:
            d["Z"] = [torch.empty(M, self.hid, device="cuda") for _ in range(4)]
            d["dOut"] = torch.zeros(M, OUTP, device="cuda")
            d["dOutbf"] = torch.zeros(M, OUTP, dtype=bf, device="cuda")
            d["dYf"] = torch.empty(M, self.hid, device="cuda")
            d["dZbf"] = torch.empty(M, self.hid, dtype=bf, device="cuda")
            d["dZf"] = torch.empty(M, self.hid, device="cuda")
            d["loss"] = torch.zeros((M + 255) // 256, device="cuda")
        self._buf[key] = d
        return d

    def forward(self, X, train=False, stream=None, hidden_only=False, y1=None):
        """X: (M, 192) bf16 CUDA tensor, M a multiple of 64.  Returns the
        (M, 64) fp32 output (column 0 is the prediction), or with hidden_only
        the last hidden layer (M, hid) bf16 for the output head.  With y1
        (inference only) X is None and the first hidden layer (M, hid) bf16
        is given."""
        L = _L()
        H = self.hid
        M = X.shape[0] if y1 is None else y1.shape[0]
        d = self._bufs(M, train)
        Y = d["Y"]
        Y[0] = X
        K = FEAT
        first = 0
        if y1 is not None:
            assert not train
            Y[1], K, first = y1, H, 1
        for li in range(first, 4):
            if train:
                _chk(L.jdsg_gemm_bf16(0, 1, M, H, K, _p(Y[li]), K, _p(self.Wbf[li]), K,
                                      _p(d["Z"][li]), H, 1, _p(self.b[li]), _p(Y[li + 1]), H,
                                      stream), "gemm")
            else:
                _chk(L.jdsg_gemm_bf16(0, 1, M, H, K, _p(Y[li]), K, _p(self.Wbf[li]), K, None,
                                      H, 4, _p(self.b[li]), _p(Y[li + 1]), H, stream), "gemm")
            K = H
        if hidden_only:
            return Y[4]
        _chk(L.jdsg_gemm_bf16(0, 1, M, OUTP, H, _p(Y[4]), H, _p(self.Wbf[4]), H,
                              _p(d["out"]), OUTP, 2, _p(self.b[4]), None, 0, stream), "gemm")
        return d["out"]

    def loss_backward(self, target, weight=None, delta=0.0, stream=None):
        """Squared-error (or, delta > 0, Huber) loss on column 0 against
        target (M,) fp32; fills the gradients and returns the loss tensor
        (partial sums)."""
        L = _L()
        M, H = target.shape[0], self.hid
        d = self._bufs(M, True)
        _chk(L.jdsg_mse(M, OUTP, _p(d["out"]), _p(target), _p(weight), float(delta),
                        _p(d["dOut"]), _p(d["loss"]), stream), "loss")
        _chk(L.jdsg_f32_to_bf16(M * OUTP, _p(d["dOut"]), _p(d["dOutbf"]), stream), "cvt")
        Y = d["Y"]
        # output layer
        _chk(L.jdsg_gemm_bf16(1, 0, OUTP, H, M, _p(d["dOutbf"]), OUTP, _p(Y[4]), H,
                              _p(self.gW[4]), H, 0, None, None, 0, stream), "gemm dW")
        _chk(L.jdsg_colsum(M, OUTP, _p(d["dOut"]), _p(self.gb[4]), stream), "colsum")
        _chk(L.jdsg_gemm_bf16(0, 0, M, H, OUTP, _p(d["dOutbf"]), OUTP, _p(self.Wbf[4]), H,
                              _p(d["dYf"]), H, 0, None, None, 0, stream), "gemm dX")
        for li in range(3, -1, -1):
            K = FEAT if li == 0 else H
            _chk(L.jdsg_gelu_bwd(M * H, _p(d["dYf"]), _p(d["Z"][li]), _p(d["dZbf"]),
                                 _p(d["dZf"]), stream), "gelu bwd")
            _chk(L.jdsg_gemm_bf16(1, 0, H, K, M, _p(d["dZbf"]), H, _p(Y[li]), K,
                                  _p(self.gW[li]), K, 0, None, None, 0, stream), "gemm dW")
            _chk(L.jdsg_colsum(M, H, _p(d["dZf"]), _p(self.gb[li]), stream), "colsum")
            if li > 0:
                _chk(L.jdsg_gemm_bf16(0, 0, M, H, H, _p(d["dZbf"]), H, _p(self.Wbf[li]),
                                      H, _p(d["dYf"]), H, 0, None, None, 0, stream),
                     "gemm dX")
        return d["loss"]

    def adam(self, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0, stream=None):
        L = _L()
        self.step += 1
        self._consts = None
        for li in range(len(self.dims)):
            _chk(L.jdsg_adam(self.W[li].numel(), _p(self.W[li]), _p(self.mW[li]),
                             _p(self.vW[li]), _p(self.gW[li]), lr, b1, b2, eps, self.step, wd,
                             _p(self.Wbf[li]), stream), "adam")
            _chk(L.jdsg_adam(self.b[li].numel(), _p(self.b[li]), _p(self.mb[li]),
                             _p(self.vb[li]), _p(self.gb[li]), lr, b1, b2, eps, self.step, 0.0,
                             None, stream), "adam")

    def save(self, path):
        """Weights as bf16 bit patterns (what inference uses) plus fp32 biases."""
        import torch
        np.savez(path, **{f"W{i}": w.view(torch.int16).cpu().numpy()
                          for i, w in enumerate(self.Wbf)},
                 **{f"b{i}": b.cpu().numpy() for i, b in enumerate(self.b)})

    @classmethod
    def load(cls, path):
        import torch
        z = np.load(path)
        m = cls(hid=int(z["W0"].shape[0]))
        for i in range(len(m.dims)):
            wb = torch.from_numpy(z[f"W{i}"]).view(torch.bfloat16).cuda()
            m.Wbf[i] = wb.contiguous()
            m.W[i].copy_(wb.float())
            m.b[i].copy_(torch.from_numpy(z[f"b{i}"]))
        m._consts = None
        return m

    def inference_consts(self):
        """The weight columns of the per-entry inputs with the first bias
        (8 x hid fp32) and the output bias as a host float, formed once per
        set of weights (the first call reads the bias back; later calls do
        no device work)."""
        import torch
        if self._consts is None:
            W1 = self.Wbf[0].float()
            ext = torch.cat([W1[:, 144:151].t(), self.b[0][None, :]]).contiguous()
            self._consts = (ext, float(self.b[4][0].item()))
        return self._consts


def training_columns(ncol, seed, K=40, love_modes=3):
    """Seeded random layered columns spanning the workload families, built
    vectorised over all columns.  Each column draws
      - 16-128 finite layers: uniform thickness 0.8-2 km (grid models) or
        per-layer thickness 0.5-4 km (uneven stacks);
      - Vs = 2.6 + 0.035 z (capped at 4.6 for half of the uneven stacks) plus
        one of: a correlated random profile (AR(1) along depth, sd up to
        0.2 km/s, correlation 5-40 km), four sinusoids (20-120 km, rms up to
        0.15 km/s), or nothing; a +-(0-10) % step at 20-60 km (checkerboards)
        with probability 1/2; per-layer noise (sd 0.075) for uneven stacks;
        and a soft sediment top (0-6 km, 0.7-2.0 km/s) with probability 1/2;
        Vs kept in 0.6-5.6 km/s;
      - K log-uniform periods in 1-80 s; Rayleigh modes 0-3, Love modes 0 to
        love_modes - 1 (the batched workloads ask for up to three of each).
    Vp and rho follow Brocher's relations as in the workloads."""
    rng = np.random.default_rng(seed)
    uneven = rng.random(ncol) < 0.3
    nl = rng.integers(16, 129, ncol)
    nl = np.where(uneven, nl, rng.integers(30, 81, ncol))
    n1 = nl + 1
    layer_off = np.concatenate([[0], np.cumsum(n1)]).astype(np.int64)
    tot = int(layer_off[-1])
    col = np.repeat(np.arange(ncol), n1)
    li = np.arange(tot) - layer_off[col]
    last = li == nl[col]
    dz = rng.uniform(0.8, 2.0, ncol)
    h = np.where(uneven[col], rng.uniform(0.5, 4.0, tot), dz[col])
    h[last] = 0.0
    ctop = np.cumsum(h) - h                      # running top depth (global)
    ztop = ctop - ctop[layer_off[:-1]][col]      # per column
    z = np.where(last, ztop, ztop + 0.5 * h)     # mid-points; half-space at the base
    trend = 2.6 + 0.035 * z
    cap = uneven & (rng.random(ncol) < 0.5)
    trend = np.where(cap[col], np.minimum(trend, 4.6), trend)
    kind = rng.integers(0, 3, ncol)              # 0 AR(1), 1 sinusoids, 2 none
    # AR(1) along depth: x_i = r x_{i-1} + sqrt(1 - r^2) e_i, r = exp(-dz/l)
    sd = rng.uniform(0.0, 0.2, ncol)
    ell = rng.uniform(5.0, 40.0, ncol)
    eps = rng.normal(0.0, 1.0, tot)
    ar = np.empty(tot)
    first = layer_off[:-1]
    ar[first] = eps[first]
    dzl = np.diff(np.concatenate([z, [0.0]]))    # spacing to the next mid-point
    for i in range(1, int(n1.max())):
        idx = first[n1 > i] + i
        r = np.exp(-np.abs(dzl[idx - 1]) / ell[col[idx]])
        ar[idx] = r * ar[idx - 1] + np.sqrt(1.0 - r * r) * eps[idx]
    rms = rng.uniform(0.0, 0.15, ncol)
    fld = np.zeros(tot)
    for _ in range(4):
        lam = rng.uniform(20.0, 120.0, ncol)
        ph = rng.uniform(0.0, 2 * np.pi, ncol)
        fld += np.sin(2 * np.pi * z / lam[col] + ph[col])
    pert = np.where(kind[col] == 0, sd[col] * ar,
                    np.where(kind[col] == 1, rms[col] * np.sqrt(0.5) * fld, 0.0))
    v = trend + pert
    stp = rng.random(ncol) < 0.5
    sgn = rng.choice([-1.0, 1.0], ncol) * rng.uniform(0.0, 0.10, ncol)
    zst = rng.uniform(20.0, 60.0, ncol)
    v = v * (1.0 + np.where(stp[col], sgn[col] * np.where(z < zst[col], 1.0, -1.0), 0.0))
    v = v + np.where(uneven[col], rng.normal(0.0, 0.075, tot), 0.0)
    sed = rng.random(ncol) < 0.5
    hsed = rng.uniform(0.0, 6.0, ncol)
    vsed = rng.uniform(0.7, 2.0, ncol)
    v = np.where(sed[col] & (z < hsed[col]), vsed[col], v)
    vs = np.clip(v, 0.6, 5.6).astype(np.float32)
    thk = h.astype(np.float32)
    periods = np.sort(np.exp(rng.uniform(0.0, np.log(80.0), (ncol, K))), axis=1)
    vp, rho = W.brocher(vs)
    return W.Workload("train", layer_off, thk, vp, vs, rho,
                      np.arange(ncol + 1) * K, periods.ravel(),
                      np.tile(np.array([4, love_modes], np.int32), (ncol, 1)), group=False)


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
        self.colf = torch.empty(plan.ncol * 148, device="cuda")
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


def train(ncol=200000, steps=3000, batch=8192, lr=1e-3, seed=7, log=print, valid_cols=4000,
          huber=0.02, hid=HID, love_modes=3):
    """Generate data with the plain GPU solver and train the MLP on the GPU.
    Returns (mlp, report)."""
    import torch
    t0 = time.time()
    w = training_columns(ncol + valid_cols, seed, love_modes=love_modes)
    plan, inp, roots = make_dataset(w)
    t_data = time.time() - t0
    ok = torch.nonzero(roots > 0).squeeze(1)
    # the last valid_cols columns are held out
    e_split = int(plan.ent_off[ncol])
    tr = ok[ok < e_split]
    va = ok[ok >= e_split]
    feats = Features(plan, inp)
    mlp = MLP(seed, hid=hid)
    g = torch.Generator(device="cuda").manual_seed(seed)
    t1 = time.time()
    losses = []
    for it in range(steps):
        idx = tr[torch.randint(0, tr.numel(), (batch,), device="cuda", generator=g)]
        X = feats.rows(rows=idx.contiguous())
        tgt = ((torch.log(roots[idx]) - C_SHIFT) / C_SCALE).float().contiguous()
        mlp.forward(X, train=True)
        loss = mlp.loss_backward(tgt, delta=huber)
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


def first_layer_parts(mlp, feats, stream=None):
    """Per-column pre-activations of the first hidden layer (ncol x hid fp32:
    one GEMM over the column feature rows) and the weight columns of the
    per-entry inputs (8 x hid fp32: ln T, two wave and four mode one-hots,
    then the bias) -- what jdsg_plan_mlp_expand combines per entry."""
    import torch
    L = _L()
    if feats.first:   # column features only (no entry rows)
        feats.rows(e0=0, n=0, X=torch.empty(0, FEAT, dtype=torch.bfloat16, device="cuda"),
                   stream=stream)
    ncol, H = feats.plan.ncol, mlp.hid
    ncp = (ncol + 127) // 128 * 128
    key = (ncp, H)
    if getattr(feats, "_fl_key", None) != key:
        # rows past ncol only feed rows of acol nothing reads
        feats._xc = torch.empty(ncp, FEAT, dtype=torch.bfloat16, device="cuda")
        feats._acol = torch.empty(ncp, H, device="cuda")
        feats._fl_key = key
    Xc, acol = feats._xc, feats._acol
    _chk(L.jdsg_mlp_column_rows(ncol, _p(feats.colf), _p(Xc), stream), "column rows")
    _chk(L.jdsg_gemm_bf16(0, 1, ncp, H, FEAT, _p(Xc), FEAT, _p(mlp.Wbf[0]), FEAT, _p(acol), H,
                          0, None, None, 0, stream), "gemm")
    ext, _ = mlp.inference_consts()
    return acol, ext


def predict(mlp, feats, rows=None, e0=0, n=None, chunk=1 << 20, split=True, fused=None):
    """Predicted flat c for entries (rows, or e0..e0+n) as a float64 tensor.
    For a range of entries (split) the first layer is the per-column GEMM
    plus the per-entry terms, and with `fused` (default: for width 256) the
    whole network then runs in one kernel with the same result; for given
    rows it is the full GEMM of their input rows."""
    import torch
    L = _L()
    if rows is not None:
        n = rows.numel()
        split = False
    if fused is None:
        fused = mlp.hid == 256
    out = torch.empty(n, dtype=torch.float64, device="cuda")
    _, b4 = mlp.inference_consts()   # bias of the output used (the rest is padding)
    if split and fused:
        if mlp.hid != 256:
            raise ValueError("the fused network kernel is for hidden width 256")
        acol, ext = first_layer_parts(mlp, feats)
        nw = int(L.jdsg_mlp_fused_work(n))
        if getattr(feats, "_work_n", None) != nw:
            feats._work = torch.empty(max(nw, 1), dtype=torch.int32, device="cuda")
            feats._work_n = nw
        work = feats._work
        _chk(L.jdsg_plan_mlp_fused(feats.plan.h, _p(feats.inp[4]), e0, n, _p(acol), mlp.hid,
                                   _p(ext), *[_p(mlp.Wbf[i]) for i in (1, 2, 3)],
                                   *[_p(mlp.b[i]) for i in (1, 2, 3)], _p(mlp.Wbf[4]), b4,
                                   C_SCALE, C_SHIFT, _p(out), 0, _p(work), nw, None),
             "fused network")
        return out
    if split:
        acol, ext = first_layer_parts(mlp, feats)
        for s in range(0, n, chunk):
            m = min(chunk, n - s)
            mp = (m + 63) // 64 * 64
            d = mlp._bufs(mp, False)
            if "Y1in" not in d:
                d["Y1in"] = torch.zeros(mp, mlp.hid, dtype=torch.bfloat16, device="cuda")
            Y1 = d["Y1in"]
            _chk(L.jdsg_plan_mlp_expand(feats.plan.h, _p(feats.inp[4]), e0 + s, m, _p(acol),
                                        mlp.hid, _p(ext), mlp.hid, _p(Y1), None), "expand")
            Yh = mlp.forward(None, train=False, hidden_only=True, y1=Y1)
            _chk(L.jdsg_mlp_head(m, mlp.hid, _p(Yh), _p(mlp.Wbf[4]), b4, C_SCALE, C_SHIFT,
                                 _p(out), s, None), "head")
        return out
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
        Yh = mlp.forward(Xb, train=False, hidden_only=True)
        _chk(L.jdsg_mlp_head(m, mlp.hid, _p(Yh), _p(mlp.Wbf[4]), b4, C_SCALE, C_SHIFT, _p(out),
                             s, None), "head")
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


def bind_network(plan, mlp):
    """Bind the network to the plan for the device-only surrogate iteration
    (jdsg_plan_set_network); the plan keeps a reference to the weights."""
    L = _L()
    W = (C.c_void_p * 5)(*[w.data_ptr() for w in mlp.Wbf])
    B = (C.c_void_p * 5)(*[b.data_ptr() for b in mlp.b])
    L.jdsg_plan_set_network.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_double,
                                        C.c_double]
    _chk(L.jdsg_plan_set_network(plan.h, mlp.hid, W, B, C_SCALE, C_SHIFT), "set network")
    plan._net = mlp


def surrogate_solve_net(plan, inp, out, half_width=0.005, stream=None):
    """The surrogate iteration of a plan with a bound network, all device
    work (no host synchronisation): network, bracket + polish, fallback,
    eigen stage."""
    import torch
    L = _L()
    L.jdsg_plan_execute_surrogate_net.argtypes = [C.c_void_p] * 6 + [C.c_double] + \
        [C.c_void_p] * 12
    ks = [_p(out.get(n)) for n in swd.KERNEL_NAMES]
    st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    _chk(L.jdsg_plan_execute_surrogate_net(plan.h, *[_p(t) for t in inp], half_width,
                                           _p(out["c"]), _p(out["u"]), _p(out["status"]), *ks,
                                           C.c_void_p(st)), "execute_surrogate_net")
    return out


def capture_surrogate(plan, inp, out, stream, half_width=0.005, perturb_amp=0.0, seed=0):
    """Record one surrogate iteration (optional in-place model update, then
    surrogate_solve_net) on `stream` into a graph; returns a swd.Graph."""
    L = _L()
    L.jdsg_plan_capture_surrogate.argtypes = [C.c_void_p] * 6 + [C.c_double] + \
        [C.c_void_p] * 11 + [C.c_float, C.c_uint, C.c_void_p, C.POINTER(C.c_void_p)]
    g = C.c_void_p()
    ks = [_p(out.get(n)) for n in swd.KERNEL_NAMES]
    _chk(L.jdsg_plan_capture_surrogate(plan.h, *[_p(t) for t in inp], half_width,
                                       _p(out["c"]), _p(out["u"]), _p(out["status"]), *ks,
                                       float(perturb_amp), int(seed),
                                       C.c_void_p(stream.cuda_stream), C.byref(g)),
         "capture_surrogate")
    return swd.Graph(g, plan)


def network_pred(plan):
    L = _L()
    h = np.zeros(plan.nent, np.float64)
    L.jdsg_plan_network_pred.argtypes = [C.c_void_p, C.c_void_p]
    _chk(L.jdsg_plan_network_pred(plan.h, h.ctypes.data), "network pred")
    return h


def surrogate_stats(plan):
    L = _L()
    f, pe, ce = C.c_longlong(), C.c_longlong(), C.c_longlong()
    _chk(L.jdsg_plan_surrogate_stats(plan.h, C.byref(f), C.byref(pe), C.byref(ce)), "stats")
    sj = C.c_longlong()
    _chk(L.jdsg_plan_scan_jobs(plan.h, C.byref(sj)), "scan jobs")
    return {"roots_left_to_chain": f.value, "polish_evals": pe.value, "chain_evals": ce.value,
            "scan_jobs": sj.value}


DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights",
                               "surrogate_mlp.npz")
# the same width and training, with Love modes 0-2 in the training set
L3_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights",
                          "surrogate_mlp_l3.npz")


def weights_for(nmode):
    """The shipped width-256 network for a batch's modes: the one trained on
    Rayleigh modes 0-3 and Love modes 0-1, which is the more accurate on
    those, unless a column asks for a third Love mode, which only the network
    trained on Love modes 0-2 as well predicts (the other one never saw that
    mode's input and misses most of its roots)."""
    nm = np.asarray(nmode).reshape(-1, 2)
    return L3_WEIGHTS if nm.size and int(nm[:, 1].max()) > 2 else DEFAULT_WEIGHTS


def _timed(fn, reps=1):
    import torch
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        r = fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps, r


def compare_solves(w, mlp, half_width=0.005, kernels=("dcdb", "dudb")):
    """Plain solve (reference tolerance), plain exact-root solve and the
    surrogate solve of one workload: times (ms, after one warm-up each), the
    surrogate's fallback statistics, its identity with the plain exact-root
    roots, and its relative differences to the plain (reference-mode)
    outputs."""
    import torch
    os.environ["JDSG_TIMING"] = "1"
    layout = "grid" if w.grid is not None else "csr"
    plan = swd.Plan(w.layer_off, w.per_off, w.nmode, w.group, layout)
    inp = swd.upload(w)
    ref = plan.alloc_outputs(kernels)
    out = plan.alloc_outputs(kernels)
    rep = {"workload": w.name, "columns": w.ncol, "column_periods": w.column_periods()}
    plan.execute(*inp, ref)
    rep["plain_ms"], _ = _timed(lambda: plan.execute(*inp, ref))
    rep["plain_stage_ms"] = [float(x) for x in plan.timings()]
    plan.set_tolerance(1e-10)
    plan.execute(*inp, out)
    rep["plain_exact_ms"], _ = _timed(lambda: plan.execute(*inp, out))
    exact_roots = torch.from_numpy(plan.flat_roots()[0]).cuda()
    exact_status = out["status"].clone()
    surrogate_solve(plan, inp, out, mlp=mlp, half_width=half_width)
    plan._feats.first = True
    rep["predict_ms"], pred = _timed(lambda: predict(mlp, plan._feats, e0=0, n=plan.nent))
    rep["surrogate_ms"], _ = _timed(lambda: surrogate_solve(plan, inp, out, mlp=mlp,
                                                            half_width=half_width))
    rep["surrogate_stage_ms"] = [float(x) for x in plan.timings()]
    st = surrogate_stats(plan)
    nvar = 3 if w.group else 1
    rep.update(st)
    rep["roots"] = nvar * plan.nent
    has = exact_roots > 0
    rel = ((pred - exact_roots).abs() / exact_roots)[has].cpu().numpy()
    rep["pred_rel_err_median"] = float(np.median(rel))
    rep["pred_rel_err_p99"] = float(np.percentile(rel, 99))
    rep["pred_within_bracket"] = float(np.mean(rel < half_width))
    sroots = torch.from_numpy(plan.flat_roots()[0]).cuda()
    okm = (exact_status == 0) & (out["status"] == 0)
    same = ((sroots - exact_roots).abs() <= 1e-9 * exact_roots.abs()) | ~has
    rep["status_identical"] = bool(torch.equal(exact_status, out["status"]))
    rep["entries_other_root"] = int((~same & okm).sum().item())
    ok = (ref["status"] == 0) & (out["status"] == 0)
    for n in ("c", "u"):
        d = ((out[n] - ref[n]).abs() / ref[n].abs())[ok]
        rep[f"max_rel_{n}_vs_plain"] = float(d.max().item()) if d.numel() else 0.0
    rep["speedup_vs_plain"] = rep["plain_ms"] / rep["surrogate_ms"]
    return rep


def main(argv=None):
    """python -m jdsurfg_amd.surrogate train|eval ...: train the network on
    the GPU and write its weights, or compare the surrogate solve with the
    plain one on a benchmark workload."""
    import argparse
    import json
    import torch
    ap = argparse.ArgumentParser(prog="python -m jdsurfg_amd.surrogate")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--ncol", type=int, default=2_000_000)
    t.add_argument("--steps", type=int, default=30000)
    t.add_argument("--batch", type=int, default=8192)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--seed", type=int, default=7)
    t.add_argument("--hid", type=int, default=HID, help="hidden width (multiple of 128)")
    t.add_argument("--love-modes", type=int, default=3, help="Love modes in the training set")
    t.add_argument("--out", default=DEFAULT_WEIGHTS)
    e = sub.add_parser("eval")
    e.add_argument("--weights", default=None, help="default: the network the workload's "
                   "modes call for (weights_for)")
    e.add_argument("--workload", default="grid128")
    a = ap.parse_args(argv)
    if not torch.cuda.is_available():
        raise SystemExit("the surrogate runs on the GPU")
    if a.cmd == "train":
        mlp, rep = train(ncol=a.ncol, steps=a.steps, batch=a.batch, lr=a.lr, seed=a.seed,
                         log=lambda m: print(m, flush=True), hid=a.hid, love_modes=a.love_modes)
        os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
        mlp.save(a.out)
        rep["weights"] = a.out
        print(json.dumps(rep), flush=True)
    else:
        w = W.BUILDERS[a.workload]()
        mlp = MLP.load(a.weights or weights_for(w.nmode))
        print(json.dumps(compare_solves(w, mlp)), flush=True)


if __name__ == "__main__":
    main()
