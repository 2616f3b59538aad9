"""ctypes bindings for the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, as the checker.  Two builds are bound:

* ``rest`` -- oracle/liboracle.so, the project's plain-C restatement of
  src/SWD/surfdisp96.f, sregn96.f90, slegn96.f90, surfdisp.cpp and the
  SurfTime glue of SurfaceWave.cpp;
* ``ref``  -- oracle/_ref/libswdref.so, the same glue bound to the reference
  Fortran/C++ compiled from its own sources (present only where it was built).
"""
import ctypes as C
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_REST = os.path.join(HERE, "liboracle.so")
LIB_REF = os.path.join(HERE, "_ref", "libswdref.so")
LIB_LSMR_REF = os.path.join(HERE, "_ref", "liblsmrref.so")

_f = np.float32
_d = np.float64
_fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")

WAVES = {"Rc": 0, "Rg": 1, "Lc": 2, "Lg": 3}


class _Tables(C.Structure):
    _fields_ = [("t", C.c_void_p * 4), ("ntmax", C.c_int * 4), ("nmode", C.c_int * 4)]


class _Batch(C.Structure):
    _fields_ = [("ncol", C.c_int), ("layer_off", C.c_void_p), ("thk", C.c_void_p),
                ("vp", C.c_void_p), ("vs", C.c_void_p), ("rho", C.c_void_p),
                ("per_off", C.c_void_p), ("periods", C.c_void_p),
                ("nmode", C.c_void_p), ("group", C.c_int)]


class _BatchOut(C.Structure):
    _fields_ = [("ent_off", C.c_void_p), ("kern_off", C.c_void_p), ("c", C.c_void_p),
                ("u", C.c_void_p)] + [(n, C.c_void_p) for n in
                                       ("dcda", "dcdb", "dcdr", "dcdh",
                                        "duda", "dudb", "dudr", "dudh")] + \
               [("status", C.c_void_p), ("cflat", C.c_void_p)]


def available(kind="rest"):
    return os.path.exists(LIB_REST if kind == "rest" else LIB_REF)


class Oracle:
    """One loaded oracle build; ``kind`` is 'rest' or 'ref'."""

    def __init__(self, kind="rest"):
        self.kind = kind
        path = LIB_REST if kind == "rest" else LIB_REF
        if not os.path.exists(path):
            raise FileNotFoundError(f"oracle library {path} is not built")
        self.lib = L = C.CDLL(path)
        L.or_surfdisp96.argtypes = [_fp, _fp, _fp, _fp, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_int, C.c_int, _dp, _dp, C.POINTER(C.c_int),
                                    C.POINTER(C.c_longlong)]
        L.or_sregn96.argtypes = [_fp] * 4 + [C.c_int, _dp, _dp, _dp] + [_dp] * 8 + [C.c_int]
        L.or_sregnpu.argtypes = [_fp] * 4 + [C.c_int, _dp, _dp, _dp] + [_dp] * 4 + \
                                [_dp] * 4 + [_dp] * 8 + [C.c_int]
        L.or_slegn96.argtypes = [_fp] * 3 + [C.c_int, _dp, _dp, _dp] + [_dp] * 5 + [C.c_int]
        L.or_slegnpu.argtypes = [_fp] * 3 + [C.c_int, _dp, _dp, _dp, _dp, _dp] + \
                                [_dp] * 4 + [_dp] * 6 + [C.c_int]
        L.or_surfdisp.argtypes = [_fp] * 4 + [C.c_int, _dp, _dp, C.c_int, C.c_int, C.c_int,
                                              C.c_int, C.c_int]
        L.or_surfdisp.restype = C.c_int
        L.or_empirical_relation.argtypes = [C.c_float, C.POINTER(C.c_float),
                                            C.POINTER(C.c_float)]
        L.or_empirical_deriv.argtypes = [C.c_float, C.c_float, C.POINTER(C.c_float),
                                         C.POINTER(C.c_float)]
        L.or_grid2layer.argtypes = [_fp] * 4 + [C.c_int, C.c_int] + [_fp] * 4
        L.or_grid2layer.restype = C.c_int
        L.or_param_convert.argtypes = [_fp, C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.or_get_1d_kernel.argtypes = [C.c_void_p, _fp, C.c_int, C.c_int, C.c_int, _fp,
                                       C.POINTER(_Tables), _dp, _dp, _dp, C.c_int]
        L.or_get_1d_kernel.restype = C.c_int
        L.or_get_2d_map.argtypes = [C.c_void_p, _fp, C.c_int, C.c_int, C.c_int, _fp,
                                    C.POINTER(_Tables), _dp, _dp, C.c_int]
        L.or_get_2d_map.restype = C.c_int
        L.or_batch_offsets.argtypes = [C.POINTER(_Batch), C.c_void_p, C.c_void_p]
        L.or_batch_solve.argtypes = [C.c_void_p, C.POINTER(_Batch), C.POINTER(_BatchOut),
                                     C.c_int]
        L.or_batch_eigen.argtypes = [C.c_void_p, C.POINTER(_Batch), C.POINTER(_BatchOut),
                                     C.c_void_p, C.c_int]
        L.or_restatement_backend.restype = C.c_void_p
        L.or_set_tolerance.argtypes = [C.c_double]
        _lp = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
        L.or_frechet_rows.argtypes = [C.c_int] * 4 + [_dp, C.c_int, _ip, _fp, _lp, _ip, _fp,
                                                      C.c_longlong]
        L.or_frechet_rows.restype = C.c_longlong
        L.or_frechet_grad.argtypes = [C.c_int] * 4 + [_dp, C.c_int, _ip, _fp, _fp, _fp, _fp]
        if kind == "ref":
            L.ref_backend.restype = C.c_void_p
            self.backend = L.ref_backend()
            L.ref_surfdisp96.argtypes = [_fp, _fp, _fp, _fp, C.c_int, C.c_int, C.c_int,
                                         C.c_int, C.c_int, C.c_int, _dp, _dp,
                                         C.POINTER(C.c_int)]
        else:
            self.backend = L.or_restatement_backend()

    def set_tolerance(self, tol=1e-6):
        """Neville stop rule; below 1e-6 the exact-root mode (restatement only)."""
        self.lib.or_set_tolerance(float(tol))

    # ---------------- per-column ABI ----------------
    def surfdisp96(self, thk, vp, vs, rho, t, iflsph=1, iwave=2, mode=1, igr=0):
        thk, vp, vs, rho = (np.ascontiguousarray(a, _f) for a in (thk, vp, vs, rho))
        t = np.ascontiguousarray(t, _d)
        cg = np.zeros(len(t), _d)
        ierr = C.c_int(0)
        nev = C.c_longlong(0)
        if self.kind == "ref":
            self.lib.ref_surfdisp96(thk, vp, vs, rho, len(thk), iflsph, iwave, mode, igr,
                                    len(t), t, cg, C.byref(ierr))
        else:
            self.lib.or_surfdisp96(thk, vp, vs, rho, len(thk), iflsph, iwave, mode, igr,
                                   len(t), t, cg, C.byref(ierr), C.byref(nev))
        return cg, ierr.value, nev.value

    def surfdisp(self, thk, vp, vs, rho, t, wave="Rc", mode=0, sphere=True, keep_flat=True):
        thk, vp, vs, rho = (np.ascontiguousarray(a, _f) for a in (thk, vp, vs, rho))
        t = np.ascontiguousarray(t, _d)
        cg = np.zeros(len(t), _d)
        if self.kind == "ref":
            raise NotImplementedError("the reference wrapper exit(1)s; use surfdisp96")
        rc = self.lib.or_surfdisp(thk, vp, vs, rho, len(thk), t, cg, len(t), WAVES[wave],
                                  mode, int(sphere), int(keep_flat))
        return cg, rc

    def _be(self):
        return C.cast(self.backend, C.c_void_p)

    def sregn96(self, thk, vp, vs, rho, t, cp, iflsph=1):
        n = len(thk)
        thk, vp, vs, rho = (np.ascontiguousarray(a, _f) for a in (thk, vp, vs, rho))
        out = [np.zeros(n, _d) for _ in range(8)]
        tt = np.array([t], _d)
        c = np.array([cp], _d)
        u = np.zeros(1, _d)
        fn = self.lib.or_sregn96 if self.kind == "rest" else None
        if fn is None:
            return self._ref_call("sregn96", thk, vp, vs, rho, tt, c, u, out, iflsph)
        fn(thk, vp, vs, rho, n, tt, c, u, *out, iflsph)
        return c[0], u[0], dict(zip(("ur", "uz", "tr", "tz", "dcda", "dcdb", "dcdh", "dcdr"), out))

    def _ref_call(self, *a):
        raise NotImplementedError("per-column reference calls go through the glue")

    # ---------------- SurfTime glue ----------------
    @staticmethod
    def make_tables(tabs):
        """tabs: dict wave -> 2-D array (ntmax, nmode) padded with -1."""
        tb = _Tables()
        keep = []
        for w, i in WAVES.items():
            a = tabs.get(w)
            if a is None or np.size(a) == 0:
                tb.t[i] = None
                tb.ntmax[i] = 0
                tb.nmode[i] = 0
                continue
            a = np.asfortranarray(np.asarray(a, _d))
            keep.append(a)
            tb.t[i] = a.ctypes.data
            tb.ntmax[i] = a.shape[0]
            tb.nmode[i] = a.shape[1]
        return tb, keep

    @staticmethod
    def tables_kmax(tabs):
        k = 0
        for w in WAVES:
            a = tabs.get(w)
            if a is None:
                continue
            a = np.asarray(a)
            for m in range(a.shape[1]):
                col = a[:, m]
                n = 0
                while n < len(col) and col[n] != -1:
                    n += 1
                k += n
        return k

    def get_1d_kernel(self, vs, dep, tabs, nthreads=1):
        """vs: (nx,ny,nz) array; returns vc (nxy,kmax), vout, kernel (nxy,nz,kmax),
        all Fortran-ordered like the reference's Eigen arrays, and the error count."""
        nx, ny, nz = vs.shape
        vsf = np.ascontiguousarray(np.asarray(vs, _f).ravel(order="F"))
        tb, keep = self.make_tables(tabs)
        kmax = self.tables_kmax(tabs)
        vc = np.zeros(nx * ny * kmax, _d)
        vout = np.zeros(nx * ny * kmax, _d)
        ker = np.zeros(nx * ny * nz * kmax, _d)
        nerr = self.lib.or_get_1d_kernel(self._be(), vsf, nx, ny, nz,
                                         np.ascontiguousarray(dep, _f), C.byref(tb),
                                         vc, vout, ker, nthreads)
        del keep
        return (vc.reshape((nx * ny, kmax), order="F"), vout.reshape((nx * ny, kmax), order="F"),
                ker.reshape((nx * ny, nz, kmax), order="F"), nerr)

    def get_2d_map(self, vs, dep, tabs, nthreads=1):
        nx, ny, nz = vs.shape
        vsf = np.ascontiguousarray(np.asarray(vs, _f).ravel(order="F"))
        tb, keep = self.make_tables(tabs)
        kmax = self.tables_kmax(tabs)
        vc = np.zeros(nx * ny * kmax, _d)
        vout = np.zeros(nx * ny * kmax, _d)
        nerr = self.lib.or_get_2d_map(self._be(), vsf, nx, ny, nz,
                                      np.ascontiguousarray(dep, _f), C.byref(tb), vc, vout,
                                      nthreads)
        del keep
        return vc.reshape((nx * ny, kmax), order="F"), vout.reshape((nx * ny, kmax), order="F"), nerr

    # ---------------- layer-stack batch ----------------
    def batch_solve(self, layer_off, thk, vp, vs, rho, per_off, periods, nmode, group,
                    nthreads=1, kernels=True, roots=None):
        """The layer-stack batch (kernel1D's call sequence per column).  With
        roots ([3*nent] flat roots T, T(1+.005), T(1-.005)) the chains are
        skipped and the eigen stage is evaluated at those roots."""
        layer_off = np.ascontiguousarray(layer_off, np.int32)
        per_off = np.ascontiguousarray(per_off, np.int32)
        nmode = np.ascontiguousarray(nmode, np.int32)
        arrs = [np.ascontiguousarray(a, _f) for a in (thk, vp, vs, rho)]
        periods = np.ascontiguousarray(periods, _d)
        ncol = len(layer_off) - 1
        b = _Batch(ncol, layer_off.ctypes.data, arrs[0].ctypes.data, arrs[1].ctypes.data,
                   arrs[2].ctypes.data, arrs[3].ctypes.data, per_off.ctypes.data,
                   periods.ctypes.data, nmode.ctypes.data, int(group))
        ent_off = np.zeros(ncol + 1, np.int64)
        kern_off = np.zeros(ncol + 1, np.int64)
        self.lib.or_batch_offsets(C.byref(b), ent_off.ctypes.data, kern_off.ctypes.data)
        ne, nk = int(ent_off[-1]), int(kern_off[-1])
        res = {"c": np.zeros(ne, _d), "u": np.zeros(ne, _d), "status": np.zeros(ne, np.int32),
               "ent_off": ent_off, "kern_off": kern_off, "cflat": np.zeros(3 * ne, _d)}
        names = ("dcda", "dcdb", "dcdr", "dcdh", "duda", "dudb", "dudr", "dudh")
        for n in names:
            res[n] = np.zeros(nk if kernels else 0, _d)
        o = _BatchOut(ent_off.ctypes.data, kern_off.ctypes.data, res["c"].ctypes.data,
                      res["u"].ctypes.data,
                      *[(res[n].ctypes.data if kernels else None) for n in names],
                      res["status"].ctypes.data, res["cflat"].ctypes.data)
        if roots is None:
            self.lib.or_batch_solve(self._be(), C.byref(b), C.byref(o), nthreads)
        else:
            roots = np.ascontiguousarray(roots, _d)
            if roots.size != 3 * ne:
                raise ValueError("roots must hold 3*nent flat roots")
            res["cflat"][:] = roots
            self.lib.or_batch_eigen(self._be(), C.byref(b), C.byref(o), roots.ctypes.data,
                                    nthreads)
        return res


    def frechet_rows(self, kernel, row_idx, fdm):
        """SurfaceWave.cpp:601-641: CSR rows (row_off, col, val) from the
        kernels (nx*ny, nz, kmax) and dense ray sensitivities fdm (nrows, nx, ny)
        (fdm[r, ix, iy]), both in the reference's column-major sense."""
        nrows, nx, ny = fdm.shape
        kernel = np.asarray(kernel, _d)
        nz, kmax = kernel.shape[1], kernel.shape[2]
        kf = np.ascontiguousarray(kernel.ravel(order="F"))
        ff = np.ascontiguousarray(np.stack([f.ravel(order="F") for f in fdm]).astype(_f))
        ri = np.ascontiguousarray(row_idx, np.int32)
        ro = np.zeros(nrows + 1, np.int64)
        nnz = self.lib.or_frechet_rows(nx, ny, nz, kmax, kf, nrows, ri, ff, ro,
                                       np.zeros(1, np.int32), np.zeros(1, _f), 0)
        col = np.zeros(max(nnz, 1), np.int32)
        val = np.zeros(max(nnz, 1), _f)
        self.lib.or_frechet_rows(nx, ny, nz, kmax, kf, nrows, ri, ff, ro, col, val, nnz)
        return ro, col[:nnz], val[:nnz]

    def frechet_grad(self, kernel, row_idx, fdm, resid, vs):
        """SurfaceWave.cpp:738-773 with one thread: the gradient over the
        interior cells ((nx-2)*(ny-2)*(nz-1), c = iz*(nx-2)*(ny-2)+...)."""
        nrows, nx, ny = fdm.shape
        kernel = np.asarray(kernel, _d)
        nz, kmax = kernel.shape[1], kernel.shape[2]
        kf = np.ascontiguousarray(kernel.ravel(order="F"))
        ff = np.ascontiguousarray(np.stack([f.ravel(order="F") for f in fdm]).astype(_f))
        g = np.zeros((nx - 2) * (ny - 2) * (nz - 1), _f)
        self.lib.or_frechet_grad(nx, ny, nz, kmax, kf, nrows,
                                 np.ascontiguousarray(row_idx, np.int32), ff,
                                 np.ascontiguousarray(resid, _f),
                                 np.ascontiguousarray(np.asarray(vs, _f).ravel(order="F")), g)
        return g


def lsmr(m, n, val, ind, ptr, b, damp=0.0, atol=1e-5, btol=1e-5, conlim=1e6, itnlim=None,
         local_size=10, kind="rest", threads=1):
    """LSMR on a CSR matrix (lsmrModule_csr.f90): the restatement ('rest'),
    its double twin ('f64', x returned in float64) or the compiled reference
    ('ref', shared/lsmrModule_csr.f90 with shared/parallel.cpp, `threads`
    OpenMP threads).  Returns x and a dict of
    istop, itn, normA, condA, normr, normAr, normx."""
    if itnlim is None:
        itnlim = 2 * n
    val = np.ascontiguousarray(val, _f)
    ind = np.ascontiguousarray(ind, np.int32)
    ptr = np.ascontiguousarray(ptr, np.int32)
    b = np.ascontiguousarray(b, _f)
    x = np.zeros(n, _f)
    ist, itn = C.c_int(), C.c_int()
    outs = [C.c_float() for _ in range(5)]
    ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    if kind == "f64":  # the double twin of the restatement (a yardstick)
        d = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        L = C.CDLL(LIB_REST)
        fn = L.or_lsmr_csr_f64
        fn.argtypes = [C.c_int, C.c_int, d, ip, ip, d] + [C.c_double] * 4 + \
            [C.c_int, C.c_int, d] + [C.POINTER(C.c_int)] * 2 + [C.POINTER(C.c_double)] * 5
        x = np.zeros(n, np.float64)
        outs = [C.c_double() for _ in range(5)]
        fn(m, n, val.astype(np.float64), ind, ptr, b.astype(np.float64), damp, atol, btol, conlim,
           itnlim, local_size, x, C.byref(ist), C.byref(itn), *[C.byref(o) for o in outs])
    elif kind == "rest":
        L = C.CDLL(LIB_REST)
        fn = L.or_lsmr_csr
        fn.argtypes = [C.c_int, C.c_int, _fp, ip, ip, _fp] + [C.c_float] * 4 + \
            [C.c_int, C.c_int, _fp] + [C.POINTER(C.c_int)] * 2 + [C.POINTER(C.c_float)] * 5
        fn(m, n, val, ind, ptr, b, damp, atol, btol, conlim, itnlim, local_size, x,
           C.byref(ist), C.byref(itn), *[C.byref(o) for o in outs])
    else:
        L = C.CDLL(LIB_LSMR_REF)
        fn = L.LSMR_csr
        fn.argtypes = [C.c_int, C.c_int, _fp, ip, ip, _fp] + [C.c_float] * 4 + \
            [C.c_int, C.c_int, _fp] + [C.POINTER(C.c_int)] * 2 + \
            [C.POINTER(C.c_float)] * 5 + [C.c_int, C.c_int]
        # LSMR_csr keeps h, hbar, u, v, w and localV(n, localVecs) as
        # automatic arrays (lsmrModule_csr.f90:789), which flang places on the
        # stack: call it on a thread whose stack holds them
        lv = min(local_size, m, n)
        need = 4 * (5 * n + m + n * lv) + (64 << 20)
        run = lambda: fn(m, n, val, ind, ptr, b, damp, atol, btol, conlim, itnlim, local_size, x,
                         C.byref(ist), C.byref(itn), *[C.byref(o) for o in outs], 0, threads)
        old = threading.stack_size()
        threading.stack_size(-(-need // 4096) * 4096)
        try:
            t = threading.Thread(target=run)
            t.start()
        finally:
            threading.stack_size(old)
        t.join()
    names = ("normA", "condA", "normr", "normAr", "normx")
    info = {"istop": ist.value, "itn": itn.value}
    info.update({k: o.value for k, o in zip(names, outs)})
    return x, info


def empirical_relation(vs):
    """Brocher Vp, rho for an array of Vs (float32), empirical.cpp:9-18."""
    o = Oracle("rest")
    vs = np.asarray(vs, _f)
    vp = np.zeros_like(vs)
    rho = np.zeros_like(vs)
    a, b = C.c_float(), C.c_float()
    for i, v in enumerate(vs.ravel()):
        o.lib.or_empirical_relation(float(v), C.byref(a), C.byref(b))
        vp.ravel()[i] = a.value
        rho.ravel()[i] = b.value
    return vp, rho


# ---- model smoothing and regularisation rows (smooth_oracle.c)
def _smooth_lib():
    L = C.CDLL(LIB_REST)
    fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    L.or_fdcoefs.argtypes = [C.c_int, C.c_int, fp]
    L.or_interp_irregular_z.argtypes = [fp, fp] + [C.c_int] * 4 + [fp, C.c_int]
    L.or_smooth_cart_pde.argtypes = [fp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float]
    L.or_sph_geometry.argtypes = [C.c_int] * 3 + [C.c_float] * 8 + [fp] * 5
    L.or_sph_geometry.restype = C.c_int
    L.or_smooth_sph_pde.argtypes = [fp, C.c_int, C.c_int, C.c_int] + [C.c_float] * 8
    L.or_add_regularization.argtypes = [C.c_int, ip, ip, fp, C.c_float, C.c_int, C.c_int, C.c_int]
    L.or_add_regularization.restype = C.c_longlong
    L.or_smooth_gradient.argtypes = [fp, C.c_int, C.c_int, C.c_int, fp, fp, fp, C.c_int,
                                     C.c_float, C.c_float]
    L.or_smooth_gradient.restype = C.c_int
    L.or_smooth_sph_conv.argtypes = [fp, C.c_int, C.c_int, C.c_int] + [C.c_float] * 8
    return L


def fdcoefs(acc, deriv):
    """Central finite-difference coefficients (smoothing.cpp:84-110)."""
    c = np.zeros(acc + 1, np.float32)
    _smooth_lib().or_fdcoefs(acc, deriv, c)
    return c


def interp_irregular_z(grad, nz1, dep, fwd, gradr=None):
    """grad: (nz, ny, nx) float32 (x fastest).  fwd: returns gradr (nz1, ny, nx);
    else fills and returns grad from gradr (smoothing.cpp:122-170)."""
    nz, ny, nx = grad.shape
    grad = np.ascontiguousarray(grad, np.float32).copy()
    gradr = np.zeros((nz1, ny, nx), np.float32) if gradr is None else \
        np.ascontiguousarray(gradr, np.float32).copy()
    _smooth_lib().or_interp_irregular_z(grad, gradr, nx, ny, nz, nz1,
                                        np.ascontiguousarray(dep, np.float32), int(fwd))
    return gradr if fwd else grad


def smooth_cart_pde(grad, sigma_h, sigma_v):
    """grad: (nz, ny, nx) float32; returns the smoothed copy (smoothing.cpp:177-249)."""
    g = np.ascontiguousarray(grad, np.float32).copy()
    nz, ny, nx = g.shape
    _smooth_lib().or_smooth_cart_pde(g, nx, ny, nz, sigma_h, sigma_v)
    return g


def sph_geometry(nx, ny, nz, dx, dy, dz, lat0, lon0, z0, sigma_h, sigma_v):
    """Padded-grid radius, colatitude, sin, cos, (Ch, Cv, dr, dth, dphi) and
    the step count of smooth_sph_pde."""
    radi = np.zeros(nz + 6, np.float32)
    th, s, c = (np.zeros(nx + 6, np.float32) for _ in range(3))
    sc = np.zeros(5, np.float32)
    nstep = _smooth_lib().or_sph_geometry(nx, ny, nz, dx, dy, dz, lat0, lon0, z0, sigma_h,
                                          sigma_v, radi, th, s, c, sc)
    return {"radi": radi, "theta": th, "sin": s, "cos": c, "Ch": sc[0], "Cv": sc[1],
            "dr": sc[2], "dth": sc[3], "dphi": sc[4], "nstep": nstep}


def smooth_sph_pde(grad, dx, dy, dz, lat0, lon0, z0, sigma_h, sigma_v):
    """grad: (nz, ny, nx) float32 over (r, lon, colat); returns the smoothed
    copy (smoothing.cpp:256-362)."""
    g = np.ascontiguousarray(grad, np.float32).copy()
    nz, ny, nx = g.shape
    _smooth_lib().or_smooth_sph_pde(g, nx, ny, nz, dx, dy, dz, lat0, lon0, z0, sigma_h, sigma_v)
    return g


def add_regularization(ptr, col, val, m, weight, nx, ny, nz):
    """Append the nx*ny*nz regularisation rows after m rows of a CSR whose
    arrays have room (csr_matrix.cpp:268-328); returns (ptr, col, val, nnz)."""
    ptr = np.ascontiguousarray(ptr, np.int32).copy()
    col = np.ascontiguousarray(col, np.int32).copy()
    val = np.ascontiguousarray(val, np.float32).copy()
    nnz = _smooth_lib().or_add_regularization(m, ptr, col, val, weight, nx, ny, nz)
    return ptr, col, val, int(nnz)


def smooth_gradient(grad, nx, ny, nz, dep, lat, lon, smooth_in_km, sigma_h, sigma_v):
    """DSurfTomo::smoothing (surftomo.cpp:147-178) on grad of shape
    (nz-1, ny-2, nx-2) (x fastest); returns the smoothed copy."""
    g = np.ascontiguousarray(grad, np.float32).copy()
    f = lambda a: np.ascontiguousarray(a, np.float32)
    if _smooth_lib().or_smooth_gradient(g, nx, ny, nz, f(dep), f(lat), f(lon), int(smooth_in_km),
                                        sigma_h, sigma_v):
        raise ValueError("depths do not give a regular grid")
    return g


def smooth_sph_conv(grad, dx, dy, dz, lat0, lon0, z0, sigma_h, sigma_v):
    """The direct Gaussian convolution smooth_sph (smoothing.cpp:21-63), in
    double; grad (nz, ny, nx) float32; returns the smoothed copy."""
    g = np.ascontiguousarray(grad, np.float32).copy()
    nz, ny, nx = g.shape
    _smooth_lib().or_smooth_sph_conv(g, nx, ny, nz, dx, dy, dz, lat0, lon0, z0, sigma_h, sigma_v)
    return g


# ---------------------------------------------------------------- FMST
LIB_FMST_REF = os.path.join(HERE, "_ref", "libfmstref.so")


def _fmst_rest():
    L = C.CDLL(LIB_REST)
    fp, dp, ip = _fp, _dp, _ip
    f = C.c_float
    L.fo_pair.argtypes = [C.c_int, C.c_int, f, f, f, f, dp, dp, f, f, C.c_int, fp, fp, fp, fp,
                          C.c_void_p]
    L.fo_pair.restype = C.c_int
    L.fo_pairs.argtypes = [C.c_int, C.c_int, f, f, f, f, C.c_int, dp, dp, C.c_int, ip, fp, fp,
                           ip, fp, fp, fp, fp, C.c_void_p, ip, C.c_int]
    L.fo_pairs.restype = C.c_int
    return L


def fmst_pairs(grid, vc, vout, idx, srcx, srcz, roff, rcx, rcz, want_fdm=True, nthreads=1):
    """The restated FMST over station pairs, as SurfTime::frechet_matrix
    drives it (SurfaceWave.cpp:612-624), OpenMP over pairs.

    grid = (nx, ny, lat0, lon0, dlat, dlon); vc, vout: (kmax, nx*ny) maps
    (column n = ix + iy*nx); idx[p] the pair's period index; srcx/srcz the
    source colatitude/longitude in rad; roff[p]:roff[p+1] its receivers.
    Returns ttime (srtimes), tsyn (ray-integrated), fdm (nrcv, nx*ny) or
    None, status per pair."""
    nx, ny, lat0, lon0, dlat, dlon = grid
    L = _fmst_rest()
    vc = np.ascontiguousarray(vc, _d)
    vout = np.ascontiguousarray(vout, _d)
    npair = len(idx)
    nrcv = int(roff[-1])
    tt = np.zeros(nrcv, _f)
    ts = np.zeros(nrcv, _f)
    fdm = np.zeros((nrcv, nx * ny), _f) if want_fdm else None
    st = np.zeros(npair, np.int32)
    L.fo_pairs(nx, ny, lat0, lon0, dlat, dlon, vc.shape[0], vc, vout, npair,
               np.ascontiguousarray(idx, np.int32), np.ascontiguousarray(srcx, _f),
               np.ascontiguousarray(srcz, _f), np.ascontiguousarray(roff, np.int32),
               np.ascontiguousarray(rcx, _f), np.ascontiguousarray(rcz, _f), tt, ts,
               None if fdm is None else fdm.ctypes.data, st, nthreads)
    return tt, ts, fdm, st


class FmstRef:
    """The reference fmst_init / fmst_run / fmst_reset / fmst_raypath /
    fmst_finalize (SWD/fmst.f90) compiled from source, called in
    SurfTime::frechet_matrix's order.  Its module state is per thread, so
    one instance is used from one thread."""

    def __init__(self):
        self.lib = L = C.CDLL(LIB_FMST_REF)
        f = C.c_float
        L.fmst_init.argtypes = [C.c_int, C.c_int, f, f, f, f]
        L.fmst_run.argtypes = [_dp, f, f, _fp, _fp, C.c_int, _fp]
        L.fmst_reset.argtypes = [_dp]
        L.fmst_raypath.argtypes = [f, f, f, f, _fp]
        L.fmst_raypath.restype = f
        L.fmst_finalize.argtypes = []

    @staticmethod
    def available():
        return os.path.exists(LIB_FMST_REF)

    def pair(self, grid, vc_col, vout_col, srcx, srcz, rcx, rcz):
        nx, ny, lat0, lon0, dlat, dlon = grid
        L = self.lib
        rcx = np.ascontiguousarray(rcx, _f)
        rcz = np.ascontiguousarray(rcz, _f)
        nr = len(rcx)
        tt = np.zeros(nr, _f)
        ts = np.zeros(nr, _f)
        fdm = np.zeros((nr, nx * ny), _f)
        L.fmst_init(nx, ny, lat0, lon0, dlat, dlon)
        L.fmst_run(np.ascontiguousarray(vc_col, _d), srcx, srcz, rcx, rcz, nr, tt)
        L.fmst_reset(np.ascontiguousarray(vout_col, _d))
        for i in range(nr):
            row = np.zeros(nx * ny, _f)
            ts[i] = L.fmst_raypath(srcx, srcz, rcx[i], rcz[i], row)
            fdm[i] = row
        L.fmst_finalize()
        return tt, ts, fdm


def fmst_pairs_ref(grid, vc, vout, idx, srcx, srcz, roff, rcx, rcz, want_fdm=True, nthreads=1):
    """fmst_pairs with the reference FMST compiled from source, driven by the
    OpenMP pair loop of SurfTime::frechet_matrix (oracle/fmst_ref_driver.c);
    same arguments and results (statuses are not reported: the reference
    stops on a station outside the model)."""
    nx, ny, lat0, lon0, dlat, dlon = grid
    L = C.CDLL(LIB_FMST_REF)
    f = C.c_float
    L.fr_pairs.argtypes = [C.c_int, C.c_int, f, f, f, f, _dp, _dp, C.c_int, _ip, _fp, _fp, _ip,
                           _fp, _fp, _fp, _fp, C.c_void_p, C.c_void_p, C.c_int]
    vc = np.ascontiguousarray(vc, _d)
    vout = np.ascontiguousarray(vout, _d)
    nrcv = int(roff[-1])
    tt = np.zeros(nrcv, _f)
    ts = np.zeros(nrcv, _f)
    fdm = np.zeros((nrcv, nx * ny), _f) if want_fdm else None
    scratch = np.zeros(max(nthreads, 1) * nx * ny, _f)
    L.fr_pairs(nx, ny, lat0, lon0, dlat, dlon, vc, vout, len(idx),
               np.ascontiguousarray(idx, np.int32), np.ascontiguousarray(srcx, _f),
               np.ascontiguousarray(srcz, _f), np.ascontiguousarray(roff, np.int32),
               np.ascontiguousarray(rcx, _f), np.ascontiguousarray(rcz, _f), tt, ts,
               None if fdm is None else fdm.ctypes.data, scratch.ctypes.data, nthreads)
    return tt, ts, fdm
