"""The surface-wave data file (surfdataSC.dat) through libjdsg_swd.so:
SurfTime::read_swd_data (src/SWD/io.cpp:29-238).  Host code; returns numpy
arrays ready to upload."""
import ctypes as C

import numpy as np

from . import swd

WAVES = ("Rc", "Rg", "Lc", "Lg")


def _L():
    L = swd.lib()
    vp = C.c_void_p
    L.jdsg_read_swd_data.argtypes = [C.c_char_p, C.POINTER(vp)]
    L.jdsg_swd_data_sizes.argtypes = [vp, vp, vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
    L.jdsg_swd_data_table.argtypes = [vp, C.c_int, vp, vp]
    L.jdsg_swd_data_arrays.argtypes = [vp] * 11
    L.jdsg_swd_data_free.argtypes = [vp]
    L.jdsg_io_last_error.restype = C.c_char_p
    L.jdsg_read_model_header.argtypes = [C.c_char_p] + [vp] * 7
    L.jdsg_read_model.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, vp, vp]
    L.jdsg_write_model.argtypes = [C.c_char_p] + [C.c_int] * 3 + [C.c_float] * 4 + [vp, vp]
    return L


def read_swd_data(path):
    """Returns a dict: tables {Rc, Rg, Lc, Lg: (periods nt x nmode, -1 padded,
    column-major as (nt, nmode)), kmax}, pair_off, wave, mode, period,
    src_colat, src_lon (per pair), rcv_colat, rcv_lon, dist, obst (per
    measurement)."""
    L = _L()
    h = C.c_void_p()
    if L.jdsg_read_swd_data(str(path).encode(), C.byref(h)):
        raise RuntimeError(L.jdsg_io_last_error().decode())
    try:
        nmode, nt = np.zeros(4, np.int32), np.zeros(4, np.int32)
        npair, nobs = C.c_longlong(), C.c_longlong()
        ptr = lambda a: a.ctypes.data_as(C.c_void_p)
        L.jdsg_swd_data_sizes(h, ptr(nmode), ptr(nt), C.byref(npair), C.byref(nobs))
        out = {"tables": {}}
        for w, name in enumerate(WAVES):
            t = np.zeros(nt[w] * nmode[w], np.float32)
            k = np.zeros(nmode[w], np.int32)
            L.jdsg_swd_data_table(h, w, ptr(t), ptr(k))
            out["tables"][name] = (t.reshape(nmode[w], nt[w]).T.copy(), k)
        npr, no = npair.value, nobs.value
        arrs = {"pair_off": np.zeros(npr + 1, np.int32), "wave": np.zeros(npr, np.int32),
                "mode": np.zeros(npr, np.int32), "period": np.zeros(npr, np.int32),
                "src_colat": np.zeros(npr, np.float32), "src_lon": np.zeros(npr, np.float32),
                "rcv_colat": np.zeros(no, np.float32), "rcv_lon": np.zeros(no, np.float32),
                "dist": np.zeros(no, np.float32), "obst": np.zeros(no, np.float32)}
        L.jdsg_swd_data_arrays(h, *[ptr(arrs[k]) for k in ("pair_off", "wave", "mode", "period",
                                                          "src_colat", "src_lon", "rcv_colat",
                                                          "rcv_lon", "dist", "obst")])
        out.update(arrs)
        return out
    finally:
        L.jdsg_swd_data_free(h)


def read_model(path):
    """DSurfTomo::read_model (src/surftomo/io.cpp:27-114): returns a dict with
    nx, ny, nz, goxd, gozd, dvxd, dvzd, dep (float32[nz]) and vs, a float32
    array of shape (nx, ny, nz) in Fortran order (i fastest), as the batch
    calls take it; lat = goxd - i*dvxd, lon = gozd + j*dvzd (float32)."""
    L = _L()
    n = np.zeros(3, np.int32)
    g = np.zeros(4, np.float32)
    ptr = lambda a, i=0: C.c_void_p(a.ctypes.data + i * a.itemsize)
    if L.jdsg_read_model_header(str(path).encode(), ptr(n, 0), ptr(n, 1), ptr(n, 2),
                                ptr(g, 0), ptr(g, 1), ptr(g, 2), ptr(g, 3)):
        raise RuntimeError(L.jdsg_io_last_error().decode())
    nx, ny, nz = (int(v) for v in n)
    dep = np.zeros(nz, np.float32)
    vs = np.zeros((nx, ny, nz), np.float32, order="F")
    if L.jdsg_read_model(str(path).encode(), nx, ny, nz, ptr(dep), ptr(vs)):
        raise RuntimeError(L.jdsg_io_last_error().decode())
    i, j = np.arange(nx, dtype=np.float32), np.arange(ny, dtype=np.float32)
    return {"nx": nx, "ny": ny, "nz": nz, "goxd": g[0], "gozd": g[1], "dvxd": g[2],
            "dvzd": g[3], "dep": dep, "vs": vs,
            "lat": (g[0] - i * g[2]).astype(np.float32), "lon": (g[1] + j * g[3]).astype(np.float32)}


def write_model(path, vs, dep, goxd, gozd, dvxd, dvzd):
    """The per-iteration snapshot results/mod_iter*.dat of
    src/surftomo/main.cpp:53-59: "lon lat dep vs" per node."""
    L = _L()
    vs = np.asfortranarray(vs, np.float32)
    dep = np.ascontiguousarray(dep, np.float32)
    nx, ny, nz = vs.shape
    if dep.shape != (nz,):
        raise ValueError("dep must hold nz depths")
    if L.jdsg_write_model(str(path).encode(), nx, ny, nz, goxd, gozd, dvxd, dvzd,
                          dep.ctypes.data_as(C.c_void_p), vs.ctypes.data_as(C.c_void_p)):
        raise RuntimeError(L.jdsg_io_last_error().decode())
