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
