"""Host-side interface to libjdsg_swd.so (the HIP dispersion/kernel library).

PyTorch provides device memory and the stream; every computation runs in the
library's HIP kernels.  If the library is missing this module raises -- there
is no CPU fallback.

Main objects
------------
Plan     -- a layer-stack batch (the per-column call sequence of the reference's
            kernel1D, SurfaceWave.cpp:266-412, for many columns at once);
solve()  -- convenience: plan + execute for a Workload (jdsurfg_amd.workloads).
Per-column functions mirroring src/SWD/swd.hpp (surfdisp96, sregn96, ...)
are exported by the library itself under the reference's symbol names.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libjdsg_swd.so")

_lib = None


def lib():
    """Load libjdsg_swd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() "
                           "(python jdsurfg_amd/build.py)")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.jdsg_last_error.restype = C.c_char_p
    L.jdsg_plan_create.argtypes = [C.c_int, vp, vp, vp, C.c_int, C.c_int, C.c_int,
                                   C.POINTER(vp)]
    L.jdsg_plan_sizes.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                                  C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.jdsg_plan_offsets.argtypes = [vp, vp, vp]
    L.jdsg_plan_execute.argtypes = [vp] * 17 + [vp]
    L.jdsg_plan_set_waves.argtypes = [vp, C.c_int, C.c_int]
    L.jdsg_plan_destroy.argtypes = [vp]
    L.jdsg_plan_timings.argtypes = [vp, vp]
    L.jdsg_plan_eval_count.argtypes = [vp, C.POINTER(C.c_longlong)]
    L.jdsg_plan_set_tolerance.argtypes = [vp, C.c_double]
    L.jdsg_plan_set_fallback.argtypes = [vp, C.c_int]
    L.jdsg_surfdisp.argtypes = [vp, vp, vp, vp, C.c_int, vp, vp, C.c_int, C.c_int, C.c_int,
                                C.c_int, C.c_int]
    _lib = L
    return L


def last_error():
    return lib().jdsg_last_error().decode()


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


KERNEL_NAMES = ("dcda", "dcdb", "dcdr", "dcdh", "duda", "dudb", "dudr", "dudh")


class Plan:
    """A planned layer-stack batch.

    layer_off, per_off: CSR offsets (ncol+1); nmode: (ncol, 2) Rayleigh/Love
    mode counts; group: U kernels (three root chains per wave) or phase
    kernels only; layout: "grid" (column-fastest, uniform batches) or "csr".
    """

    def __init__(self, layer_off, per_off, nmode, group=True, layout="grid", iflsph=1):
        L = lib()
        self.layer_off = np.ascontiguousarray(layer_off, np.int32)
        self.per_off = np.ascontiguousarray(per_off, np.int32)
        self.nmode = np.ascontiguousarray(np.asarray(nmode, np.int32).reshape(-1))
        self.ncol = len(self.layer_off) - 1
        self.group = bool(group)
        self.layout = {"grid": 0, "csr": 1}[layout]
        h = C.c_void_p()
        rc = L.jdsg_plan_create(self.ncol, self.layer_off.ctypes.data, self.per_off.ctypes.data,
                                self.nmode.ctypes.data, int(group), self.layout, int(iflsph),
                                C.byref(h))
        if rc:
            raise RuntimeError("jdsg_plan_create: " + last_error())
        self.h = h
        ne, nk = C.c_longlong(), C.c_longlong()
        nmax, kmax = C.c_int(), C.c_int()
        L.jdsg_plan_sizes(h, C.byref(ne), C.byref(nk), C.byref(nmax), C.byref(kmax))
        self.nent, self.nkern, self.nmax, self.kmax = ne.value, nk.value, nmax.value, kmax.value
        self.ent_off = np.zeros(self.ncol + 1, np.int64)
        self.kern_off = np.zeros(self.ncol + 1, np.int64)
        L.jdsg_plan_offsets(h, self.ent_off.ctypes.data, self.kern_off.ctypes.data)

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.jdsg_plan_destroy(self.h)
            self.h = None

    def set_waves(self, chain_waves=0, eigen_waves=0):
        if lib().jdsg_plan_set_waves(self.h, int(chain_waves), int(eigen_waves)):
            raise RuntimeError(last_error())

    def alloc_outputs(self, kernels=KERNEL_NAMES, device="cuda"):
        import torch
        out = {"c": torch.empty(self.nent, dtype=torch.float64, device=device),
               "u": torch.empty(self.nent, dtype=torch.float64, device=device),
               "status": torch.empty(self.nent, dtype=torch.int32, device=device)}
        for n in kernels:
            out[n] = torch.empty(self.nkern, dtype=torch.float64, device=device)
        return out

    def execute(self, thk, vp, vs, rho, periods, out, stream=None):
        """Launch the solve.  Inputs/outputs are CUDA tensors (float32 model,
        float64 periods); out comes from alloc_outputs().  Asynchronous."""
        import torch
        for t, dt in ((thk, torch.float32), (vp, torch.float32), (vs, torch.float32),
                      (rho, torch.float32), (periods, torch.float64)):
            if not t.is_cuda or t.dtype != dt or not t.is_contiguous():
                raise ValueError("inputs must be contiguous CUDA tensors of the documented dtype")
        if stream is None:
            stream = torch.cuda.current_stream()
        ks = [_ptr(out.get(n)) for n in KERNEL_NAMES]
        rc = lib().jdsg_plan_execute(self.h, _ptr(thk), _ptr(vp), _ptr(vs), _ptr(rho),
                                     _ptr(periods), _ptr(out["c"]), _ptr(out["u"]),
                                     _ptr(out["status"]), *ks, C.c_void_p(stream.cuda_stream))
        if rc:
            raise RuntimeError("jdsg_plan_execute: " + last_error())
        return out

    def autotune(self, thk, vp, vs, rho, periods, out, stream=None):
        """Pick the chain and eigen launch widths by timing candidates on
        these inputs; returns (chain_waves, eigen_waves).  Leaves a solve in
        out."""
        import torch
        L = lib()
        L.jdsg_plan_autotune.argtypes = [C.c_void_p] * 18 + [C.POINTER(C.c_int)] * 2
        if stream is None:
            stream = torch.cuda.current_stream()
        cw, ew = C.c_int(), C.c_int()
        ks = [_ptr(out.get(n)) for n in KERNEL_NAMES]
        rc = L.jdsg_plan_autotune(self.h, _ptr(thk), _ptr(vp), _ptr(vs), _ptr(rho), _ptr(periods),
                                  _ptr(out["c"]), _ptr(out["u"]), _ptr(out["status"]), *ks,
                                  C.c_void_p(stream.cuda_stream), C.byref(cw), C.byref(ew))
        if rc:
            raise RuntimeError("jdsg_plan_autotune: " + last_error())
        return cw.value, ew.value

    def capture(self, thk, vp, vs, rho, periods, out, stream, perturb_amp=0.0, seed=0):
        """Record one iteration (optional in-place model update, then the
        solve) on `stream` (a torch.cuda.Stream, not the default stream) into
        a graph; returns a Graph whose launch() replays it."""
        L = lib()
        L.jdsg_plan_capture.argtypes = [C.c_void_p] * 17 + [C.c_float, C.c_uint, C.c_void_p,
                                                            C.POINTER(C.c_void_p)]
        g = C.c_void_p()
        ks = [_ptr(out.get(n)) for n in KERNEL_NAMES]
        rc = L.jdsg_plan_capture(self.h, _ptr(thk), _ptr(vp), _ptr(vs), _ptr(rho), _ptr(periods),
                                 _ptr(out["c"]), _ptr(out["u"]), _ptr(out["status"]), *ks,
                                 float(perturb_amp), int(seed), C.c_void_p(stream.cuda_stream),
                                 C.byref(g))
        if rc:
            raise RuntimeError("jdsg_plan_capture: " + last_error())
        return Graph(g, self)

    def set_tolerance(self, tol):
        if lib().jdsg_plan_set_tolerance(self.h, float(tol)):
            raise RuntimeError(last_error())

    def set_fallback(self, form):
        """Chains of the roots the surrogate's brackets left open: 0 one lane
        per chain, 1 the same with suspended scans, 2 as 1 with the resumed
        chains one per wave (the default)."""
        if lib().jdsg_plan_set_fallback(self.h, int(form)):
            raise RuntimeError(last_error())

    def eval_count(self):
        v = C.c_longlong()
        if lib().jdsg_plan_eval_count(self.h, C.byref(v)):
            raise RuntimeError(last_error())
        return v.value

    def flat_roots(self):
        """Flat roots of the last execute: (nvar, nent) array (T, T+, T-)."""
        nvar = 3 if self.group else 1
        h = np.zeros(nvar * self.nent, np.float64)
        L = lib()
        L.jdsg_plan_flat_roots.argtypes = [C.c_void_p, C.c_void_p]
        if L.jdsg_plan_flat_roots(self.h, h.ctypes.data):
            raise RuntimeError(last_error())
        return h.reshape(nvar, self.nent)

    def pass_timings(self):
        """The eigen pass kernels' own milliseconds and launch counts of the
        last execute, per wave type (Rayleigh, Love), when JDSG_TIMING=1 was
        set before the plan was created."""
        ms = np.zeros(2, np.float32)
        n = np.zeros(2, np.int32)
        L = lib()
        L.jdsg_plan_pass_timings.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        if L.jdsg_plan_pass_timings(self.h, ms.ctypes.data, n.ctypes.data):
            return None
        return ms, n

    def timings(self):
        """Per-stage milliseconds of the last execute (prep, chains, eigen R,
        eigen L) when JDSG_TIMING=1 was set before the plan was created."""
        ms = np.zeros(4, np.float32)
        if lib().jdsg_plan_timings(self.h, ms.ctypes.data):
            return None
        return ms


class Graph:
    """A captured iteration (see Plan.capture)."""

    def __init__(self, h, plan):
        self.h, self.plan = h, plan

    def launch(self, stream):
        L = lib()
        L.jdsg_graph_launch.argtypes = [C.c_void_p, C.c_void_p]
        if L.jdsg_graph_launch(self.h, C.c_void_p(stream.cuda_stream)):
            raise RuntimeError("jdsg_graph_launch: " + last_error())

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.jdsg_graph_destroy.argtypes = [C.c_void_p]
            _lib.jdsg_graph_destroy(self.h)
            self.h = None


def upload(w, device="cuda"):
    import torch
    f = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(device=device, dtype=dt)
    return (f(w.thk, torch.float32), f(w.vp, torch.float32), f(w.vs, torch.float32),
            f(w.rho, torch.float32), f(w.periods, torch.float64))


def solve(w, layout=None, kernels=KERNEL_NAMES, iflsph=1):
    """Plan + execute a Workload; returns (plan, dict of CUDA tensors)."""
    import torch
    uniform = w.grid is not None
    if layout is None:
        layout = "grid" if uniform else "csr"
    p = Plan(w.layer_off, w.per_off, w.nmode, w.group, layout, iflsph)
    inp = upload(w)
    out = p.alloc_outputs(kernels)
    p.execute(*inp, out)
    torch.cuda.synchronize()
    return p, out


def grid_to_csr(plan, arr, N):
    """Reorder a grid-layout kernel array (uniform batch) into csr order."""
    ncol = plan.ncol
    a = np.asarray(arr).reshape(-1, N, ncol)     # (entry_local, layer, col)
    return np.ascontiguousarray(a.transpose(2, 0, 1)).reshape(-1)


class SurfTime:
    """The SurfTime grid forms on the GPU for a fixed depth grid and period
    tables (SurfaceWave.hpp:102-115): kernel() is get_1d_kernel, map() is
    get_2d_map, both on a device-resident (nx, ny, nz) Vs model.

    tables: dict with keys 'Rc','Rg','Lc','Lg' -> (nt, nmode) period arrays
    padded with -1.  A column whose top node has Vs = 0 has a water top."""

    def __init__(self, nx, ny, nz, dep, tables):
        L = lib()
        self.nx, self.ny, self.nz = nx, ny, nz
        args, self._keep = [], []
        for w in ("Rc", "Rg", "Lc", "Lg"):
            a = tables.get(w)
            if a is None or np.size(a) == 0:
                args += [None, 0, 0]
            else:
                a = np.asfortranarray(np.asarray(a, np.float64))
                self._keep.append(a)
                args += [C.c_void_p(a.ctypes.data), a.shape[0], a.shape[1]]
        h = C.c_void_p()
        depa = np.ascontiguousarray(dep, np.float32)
        L.jdsg_surftime_create.argtypes = [C.c_int] * 3 + [C.c_void_p] + \
            [C.c_void_p, C.c_int, C.c_int] * 4 + [C.POINTER(C.c_void_p)]
        if L.jdsg_surftime_create(nx, ny, nz, depa.ctypes.data, *args, C.byref(h)):
            raise RuntimeError("jdsg_surftime_create: " + last_error())
        self.h = h
        km = C.c_int()
        L.jdsg_surftime_kmax.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.jdsg_surftime_kmax(h, C.byref(km))
        self.kmax = km.value

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.jdsg_surftime_destroy.argtypes = [C.c_void_p]
            _lib.jdsg_surftime_destroy(self.h)
            self.h = None

    def alloc(self, kernel=True):
        import torch
        nxy = self.nx * self.ny
        o = {"vc": torch.zeros(nxy * self.kmax, dtype=torch.float64, device="cuda"),
             "vout": torch.zeros(nxy * self.kmax, dtype=torch.float64, device="cuda"),
             "status": torch.zeros(nxy * self.kmax, dtype=torch.int32, device="cuda")}
        if kernel:
            o["kernel"] = torch.zeros(nxy * self.nz * self.kmax, dtype=torch.float64,
                                      device="cuda")
        return o

    def kernel(self, vs, out, stream=None):
        """get_1d_kernel on vs (CUDA float32, (nx, ny, nz) column-major
        flattened); fills out from alloc()."""
        import torch
        L = lib()
        L.jdsg_surftime_kernel.argtypes = [C.c_void_p] * 7
        st = (stream or torch.cuda.current_stream()).cuda_stream
        if L.jdsg_surftime_kernel(self.h, _ptr(vs), _ptr(out["vc"]), _ptr(out["vout"]),
                                  _ptr(out.get("kernel")), _ptr(out["status"]), C.c_void_p(st)):
            raise RuntimeError("jdsg_surftime_kernel: " + last_error())
        return out

    def map(self, vs, out, stream=None):
        """get_2d_map on vs; fills vc, vout, status of out."""
        import torch
        L = lib()
        L.jdsg_surftime_map.argtypes = [C.c_void_p] * 6
        st = (stream or torch.cuda.current_stream()).cuda_stream
        if L.jdsg_surftime_map(self.h, _ptr(vs), _ptr(out["vc"]), _ptr(out["vout"]),
                               _ptr(out["status"]), C.c_void_p(st)):
            raise RuntimeError("jdsg_surftime_map: " + last_error())
        return out

    def to_host(self, out):
        nxy = self.nx * self.ny
        f = lambda t, shp: t.cpu().numpy().reshape(shp, order="F")
        r = [f(out["vc"], (nxy, self.kmax)), f(out["vout"], (nxy, self.kmax))]
        if "kernel" in out:
            r.append(f(out["kernel"], (nxy, self.nz, self.kmax)))
        r.append(f(out["status"], (nxy, self.kmax)))
        return tuple(r)


def _upload_vs(vs):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(vs, np.float32).ravel(order="F"))).cuda()


def get_1d_kernel(vs, dep, tables):
    """SurfTime::get_1d_kernel on the GPU (host arrays in, host arrays out).

    vs: (nx, ny, nz) float array; dep: (nz,) node depths; tables: dict with
    keys 'Rc','Rg','Lc','Lg' -> (nt, nmode) period arrays padded with -1.
    Returns vc, vout (nx*ny, kmax), kernel (nx*ny, nz, kmax) in the reference's
    column-major order, and status (nx*ny, kmax)."""
    import torch
    nx, ny, nz = vs.shape
    S = SurfTime(nx, ny, nz, dep, tables)
    out = S.kernel(_upload_vs(vs), S.alloc(kernel=True))
    torch.cuda.synchronize()
    return S.to_host(out)


def get_2d_map(vs, dep, tables):
    """SurfTime::get_2d_map on the GPU: vc, vout (nx*ny, kmax) and status."""
    import torch
    nx, ny, nz = vs.shape
    S = SurfTime(nx, ny, nz, dep, tables)
    out = S.map(_upload_vs(vs), S.alloc(kernel=False))
    torch.cuda.synchronize()
    return S.to_host(out)
