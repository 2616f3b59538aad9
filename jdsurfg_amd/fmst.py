"""FMST on the GPU: first-arrival traveltimes of the station pairs and the
ray Fréchet weights, the per-pair fmst_init / fmst_run / fmst_reset /
fmst_raypath sequence of SurfTime::travel_time, frechet_matrix and
compute_grad (src/SWD/SurfaceWave.cpp:504-676, 695-781; src/SWD/fmst.f90,
src/SWD/fmm.f90), run for every pair at once by libjdsg_swd.so.

A station pair is a source (colatitude, longitude in rad) with its receivers
and the period index of its dispersion map (get_period_index order).  The
maps are the column-major (nx*ny, kmax) vc / vout that SurfTime.map() or
SurfTime.kernel() leave on the GPU.  run() gives each receiver's time: the
bilinear traveltime of the fast-marching field (srtimes), or, for pairs whose
`ray_time` flag is set (Rg, Lg), the group time integrated along the ray;
mode 1 also keeps every receiver's ray weights, which rays() returns in the
sparse form jdsurfg_amd.frechet.frechet_rows takes.

Statuses: 0 ok; 1 source or receiver outside the model (the reference
stops); 2 the fast-marching band outgrew the reference's heap; 3 the ray
did not reach the source.
"""
import ctypes as C

import numpy as np

from . import swd

_f32 = np.float32


def _L():
    L = swd.lib()
    vp = C.c_void_p
    f = C.c_float
    L.jdsg_fmst_last_error.restype = C.c_char_p
    L.jdsg_fmst_create.argtypes = [C.c_int, C.c_int, f, f, f, f, C.c_int] + [vp] * 7 + \
        [C.POINTER(vp)]
    L.jdsg_fmst_sizes.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.jdsg_fmst_run.argtypes = [vp, vp, vp, C.c_int, C.c_int, vp, vp, vp, vp, vp]
    L.jdsg_fmst_rays.argtypes = [vp, vp, vp, vp, C.c_longlong, C.POINTER(C.c_longlong), vp]
    L.jdsg_fmst_destroy.argtypes = [vp]
    L.jdsg_surftime_travel_time.argtypes = [vp] * 6
    L.jdsg_surftime_frechet.argtypes = [vp] * 8 + [C.c_longlong, C.POINTER(C.c_longlong), vp]
    L.jdsg_surftime_compute_grad.argtypes = [vp] * 8
    return L


def last_error():
    return _L().jdsg_fmst_last_error().decode()


def _p(t):
    return None if t is None else C.c_void_p(t.data_ptr())


class Fmst:
    """Station pairs on one grid (fmst_init's nx, ny, lat0, lon0, dlat,
    dlon: the model's upper-left node and spacing in degrees)."""

    def __init__(self, nx, ny, lat0, lon0, dlat, dlon, idx, src_colat, src_lon, rcv_off,
                 rcv_colat, rcv_lon, ray_time=None):
        L = _L()
        self.nx, self.ny = nx, ny
        a = lambda x, t: np.ascontiguousarray(x, t)
        self._keep = [a(idx, np.int32), a(src_colat, _f32), a(src_lon, _f32),
                      a(rcv_off, np.int32), a(rcv_colat, _f32), a(rcv_lon, _f32),
                      a(np.zeros(len(idx)) if ray_time is None else ray_time, np.int32)]
        self.npair = len(idx)
        self.nrcv = int(self._keep[3][-1]) if self.npair else 0
        h = C.c_void_p()
        if L.jdsg_fmst_create(nx, ny, lat0, lon0, dlat, dlon, self.npair,
                              *[C.c_void_p(k.ctypes.data) for k in self._keep], C.byref(h)):
            raise RuntimeError("jdsg_fmst_create: " + last_error())
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) is not None and swd._lib is not None:
            swd._lib.jdsg_fmst_destroy(self.h)
            self.h = None

    def run(self, vc, vout, kmax, mode=0, stream=None):
        """vc, vout: CUDA float64 maps, plane idx at [idx*nx*ny, (idx+1)*nx*ny).
        Returns ttime, tsyn (float32 per receiver), rstatus, pstatus."""
        import torch
        L = _L()
        dev = vc.device
        o = {"ttime": torch.zeros(max(self.nrcv, 1), dtype=torch.float32, device=dev),
             "tsyn": torch.zeros(max(self.nrcv, 1), dtype=torch.float32, device=dev),
             "rstatus": torch.zeros(max(self.nrcv, 1), dtype=torch.int32, device=dev),
             "pstatus": torch.zeros(max(self.npair, 1), dtype=torch.int32, device=dev)}
        st = (stream or torch.cuda.current_stream()).cuda_stream
        if L.jdsg_fmst_run(self.h, _p(vc), _p(vout), kmax, mode, _p(o["ttime"]), _p(o["tsyn"]),
                           _p(o["rstatus"]), _p(o["pstatus"]), C.c_void_p(st)):
            raise RuntimeError("jdsg_fmst_run: " + last_error())
        return {k: v[:self.nrcv] if k != "pstatus" else v[:self.npair] for k, v in o.items()}

    def rays(self, stream=None):
        """The sparse ray weights of the last mode-1 run: ray_off (int64,
        nrcv + 1), ray_cell (int32, n = ix + iy*nx increasing), ray_w."""
        import torch
        L = _L()
        st = (stream or torch.cuda.current_stream()).cuda_stream
        nnz = C.c_longlong()
        off = torch.zeros(self.nrcv + 1, dtype=torch.int64, device="cuda")
        if L.jdsg_fmst_rays(self.h, _p(off), None, None, 0, C.byref(nnz), C.c_void_p(st)):
            if nnz.value <= 0:
                raise RuntimeError("jdsg_fmst_rays: " + last_error())
        n = max(int(nnz.value), 1)
        cell = torch.zeros(n, dtype=torch.int32, device="cuda")
        w = torch.zeros(n, dtype=torch.float32, device="cuda")
        if L.jdsg_fmst_rays(self.h, _p(off), _p(cell), _p(w), n, C.byref(nnz), C.c_void_p(st)):
            raise RuntimeError("jdsg_fmst_rays: " + last_error())
        return off, cell[:nnz.value], w[:nnz.value]


def _stream(stream):
    import torch
    return C.c_void_p((stream or torch.cuda.current_stream()).cuda_stream)


def travel_time(S, F, vs, stream=None):
    """SurfTime::travel_time on the GPU: S a swd.SurfTime, F an Fmst on the
    same grid, vs the CUDA float32 model ((nx, ny, nz) column-major flat).
    Returns data (float32 per receiver) and rstatus."""
    import torch
    L = _L()
    data = torch.zeros(max(F.nrcv, 1), dtype=torch.float32, device="cuda")
    rst = torch.zeros(max(F.nrcv, 1), dtype=torch.int32, device="cuda")
    if L.jdsg_surftime_travel_time(S.h, F.h, _p(vs), _p(data), _p(rst), _stream(stream)):
        raise RuntimeError("jdsg_surftime_travel_time: " + last_error())
    return data[:F.nrcv], rst[:F.nrcv]


def frechet_matrix(S, F, vs, cap=None, stream=None):
    """SurfTime::frechet_matrix on the GPU: data, rstatus and the rows in CSR
    (row_off int64, col int32, val float32), rows in data order."""
    import torch
    L = _L()
    data = torch.zeros(max(F.nrcv, 1), dtype=torch.float32, device="cuda")
    rst = torch.zeros(max(F.nrcv, 1), dtype=torch.int32, device="cuda")
    off = torch.zeros(F.nrcv + 1, dtype=torch.int64, device="cuda")
    nnz = C.c_longlong()
    if cap is None:
        cap = F.nrcv * (S.nx - 2) * (S.ny - 2) * (S.nz - 1) // 8 + 1
    col = torch.zeros(cap, dtype=torch.int32, device="cuda")
    val = torch.zeros(cap, dtype=torch.float32, device="cuda")
    if L.jdsg_surftime_frechet(S.h, F.h, _p(vs), _p(data), _p(rst), _p(off), _p(col), _p(val), cap,
                               C.byref(nnz), _stream(stream)):
        if nnz.value > cap:
            return frechet_matrix(S, F, vs, cap=nnz.value, stream=stream)
        raise RuntimeError("jdsg_surftime_frechet: " + last_error())
    n = nnz.value
    return data[:F.nrcv], rst[:F.nrcv], off, col[:n], val[:n]


def compute_grad(S, F, vs, obst, stream=None):
    """SurfTime::compute_grad on the GPU: data, grad ((nx-2)(ny-2)(nz-1)
    float32), rstatus; obst the observed times (CUDA float32 per receiver)."""
    import torch
    L = _L()
    data = torch.zeros(max(F.nrcv, 1), dtype=torch.float32, device="cuda")
    rst = torch.zeros(max(F.nrcv, 1), dtype=torch.int32, device="cuda")
    grad = torch.zeros((S.nx - 2) * (S.ny - 2) * (S.nz - 1), dtype=torch.float32, device="cuda")
    if L.jdsg_surftime_compute_grad(S.h, F.h, _p(vs), _p(obst), _p(data), _p(grad), _p(rst),
                                    _stream(stream)):
        raise RuntimeError("jdsg_surftime_compute_grad: " + last_error())
    return data[:F.nrcv], grad, rst[:F.nrcv]


def period_index(kmax_per_mode, wave, mode, period):
    """SurfTime::get_period_index (SurfaceWave.cpp:14-48): the map column of
    a (wave, mode, period) triple.  kmax_per_mode: four arrays (Rc, Rg, Lc,
    Lg) of the period count of each mode; wave 0 Rc, 1 Rg, 2 Lc, 3 Lg."""
    base = 0
    for w in range(wave):
        base += int(np.sum(kmax_per_mode[w]))
    return base + int(np.sum(kmax_per_mode[wave][:mode])) + int(period)


def from_data(d, nx, ny, lat0, lon0, dlat, dlon):
    """An Fmst for every station pair of a data file read by
    jdsurfg_amd.data.read_swd_data, on the model grid (fmst_init's lat0,
    lon0 = the model's upper-left node, dlat, dlon its spacing); Rg and Lg
    pairs report the ray time, as SurfTime does."""
    kpm = [np.asarray(d["tables"][w][1]) for w in ("Rc", "Rg", "Lc", "Lg")]
    idx = np.array([period_index(kpm, w, m, t) for w, m, t in
                    zip(d["wave"], d["mode"], d["period"])], np.int32)
    ray_time = np.isin(d["wave"], (1, 3)).astype(np.int32)
    return Fmst(nx, ny, lat0, lon0, dlat, dlon, idx, d["src_colat"], d["src_lon"], d["pair_off"],
                d["rcv_colat"], d["rcv_lon"], ray_time)
