"""The nonlinear model update of an inversion iteration on the GPU:
NonlinOPT<T> (include/nonlinear_opt.hpp:15-390) through libjdsg_swd.so
(jdsg_nlopt_*).  The model object keeps the reference's three calls, on
torch tensors in device memory:

    model.compute_grad(x, dsyn, grad)   # fill dsyn (m) and grad (n) at x
    model.compute_misfit(dsyn) -> float
    model.smoothing(direc)              # in place; optional

They run on the optimiser's stream, as torch's current stream."""
import ctypes as C

from jdsurfg_amd import swd

GRAD_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p)
MISFIT_FN = C.CFUNCTYPE(C.c_float, C.c_void_p, C.c_void_p, C.c_void_p)
SMOOTH_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_void_p)


def _L():
    L = swd.lib()
    vp = C.c_void_p
    L.jdsg_nlopt_work_floats.argtypes = [C.c_int, C.c_int]
    L.jdsg_nlopt_work_floats.restype = C.c_longlong
    L.jdsg_nlopt_create.argtypes = [C.c_int, C.c_int, C.c_float, C.c_int, C.c_int, vp, vp,
                                    C.POINTER(vp)]
    L.jdsg_nlopt_update.argtypes = [vp, vp, vp, C.c_int, GRAD_FN, MISFIT_FN, SMOOTH_FN, vp]
    L.jdsg_nlopt_state.argtypes = [vp] * 7
    L.jdsg_nlopt_destroy.argtypes = [vp]
    L.jdsg_nlopt_last_error.restype = C.c_char_p
    return L


class NonlinOPT:
    """NonlinOPT(outdir, max_step, iter_start, iter_current) for n model
    values and m data; the history lives in one device workspace."""

    def __init__(self, n, m, max_step, iter_start=0, iter_current=0, device="cuda"):
        import torch
        self.L = L = _L()
        self.n, self.m = n, m
        self.work = torch.empty(int(L.jdsg_nlopt_work_floats(n, m)), dtype=torch.float32,
                                device=device)
        self.stream = torch.cuda.current_stream(self.work.device)
        self.h = C.c_void_p()
        if L.jdsg_nlopt_create(n, m, max_step, iter_start, iter_current,
                               C.c_void_p(self.work.data_ptr()),
                               C.c_void_p(self.stream.cuda_stream), C.byref(self.h)):
            raise RuntimeError(L.jdsg_nlopt_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            self.L.jdsg_nlopt_destroy(self.h)
            self.h = None

    def _view(self, ptr, length, known):
        for t in known:
            if t.data_ptr() == ptr:
                return t
        off, rem = divmod(ptr - self.work.data_ptr(), 4)
        if rem or off < 0 or off + length > self.work.numel():
            raise RuntimeError("nlopt: a callback buffer lies outside the workspace")
        return self.work[off:off + length]

    def update(self, model, x, dsyn, method="CG"):
        """NonlinOPT::update: x (n float32, device) becomes the accepted
        model; dsyn (m float32, device) the synthetics of the model on entry."""
        import torch
        if x.numel() != self.n or dsyn.numel() != self.m or x.dtype != torch.float32 \
                or dsyn.dtype != torch.float32 or not x.is_contiguous() or not dsyn.is_contiguous():
            raise ValueError("x and dsyn must be contiguous float32 of sizes n and m")
        known, err = (x, dsyn), []

        def grad_cb(user, px, pd, pg, st):
            try:
                with torch.cuda.stream(self.stream):
                    model.compute_grad(self._view(px, self.n, known),
                                       self._view(pd, self.m, known),
                                       self._view(pg, self.n, known))
                return 0
            except Exception as e:  # reported after the call returns
                err.append(e)
                return 1

        def misfit_cb(user, pd, st):
            try:
                with torch.cuda.stream(self.stream):
                    return float(model.compute_misfit(self._view(pd, self.m, known)))
            except Exception as e:
                err.append(e)
                return float("nan")

        def smooth_cb(user, pv, st):
            try:
                with torch.cuda.stream(self.stream):
                    if hasattr(model, "smoothing"):
                        model.smoothing(self._view(pv, self.n, known))
                return 0
            except Exception as e:
                err.append(e)
                return 1

        cbs = (GRAD_FN(grad_cb), MISFIT_FN(misfit_cb), SMOOTH_FN(smooth_cb))
        rc = self.L.jdsg_nlopt_update(self.h, C.c_void_p(x.data_ptr()), C.c_void_p(dsyn.data_ptr()),
                                      0 if method == "CG" else 1, *cbs, None)
        if err:
            raise err[0]
        if rc:
            raise RuntimeError(self.L.jdsg_nlopt_last_error().decode())
        return self.state()

    def state(self):
        it, its, tr = C.c_int(), C.c_int(), C.c_int()
        sf, f0, f1 = C.c_float(), C.c_float(), C.c_float()
        self.L.jdsg_nlopt_state(self.h, C.byref(it), C.byref(its), C.byref(tr), C.byref(sf),
                                C.byref(f0), C.byref(f1))
        return {"iter": it.value, "iter_start": its.value, "trials": tr.value,
                "step_fac": sf.value, "fcost": f0.value, "fcost1": f1.value}
