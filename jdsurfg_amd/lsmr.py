"""LSMR on the GPU for the least-squares step of an inversion iteration: the
solve csr_matrix::lsmr_solver runs after the Frechet rows are assembled
(src/shared/csr_matrix.cpp:114-128, src/shared/lsmrModule_csr.f90)."""
import ctypes as C

from jdsurfg_amd import swd


def _L():
    L = swd.lib()
    vp = C.c_void_p
    L.jdsg_lsmr_csr.argtypes = [C.c_int, C.c_int, vp, vp, vp, vp] + [C.c_float] * 4 + \
        [C.c_int, C.c_int, vp] + [C.POINTER(C.c_int)] * 2 + [C.POINTER(C.c_float)] * 5 + [vp]
    L.jdsg_lsmr_csr_layers.argtypes = L.jdsg_lsmr_csr.argtypes[:-1] + [C.c_int, vp]
    L.jdsg_lsmr_last_error.restype = C.c_char_p
    return L


def lsmr(m, n, val, indices, indptr, b, damp=0.0, atol=1e-5, btol=1e-5, conlim=1e6,
         itnlim=None, local_size=10, x=None, stream=None, layers=1):
    """Solve min ||A x - b||^2 + damp^2 ||x||^2 for a CSR matrix in device
    memory (torch tensors: val float32, indices/indptr int32, b float32).
    Returns x (float32 tensor) and a dict of istop, itn and the norm
    estimates, as LSMR_csr returns them.  layers > 1: the columns are
    numbered layer-major (c = layer * (n / layers) + cell, as Frechet rows
    are) and the solve runs cell-major (jdsg_lsmr_csr_layers)."""
    import torch
    if itnlim is None:
        itnlim = 2 * n
    if x is None:
        x = torch.empty(n, dtype=torch.float32, device=b.device)
    ist, itn = C.c_int(), C.c_int()
    outs = [C.c_float() for _ in range(5)]
    p = lambda t: C.c_void_p(t.data_ptr())
    L = _L()
    args = [m, n, p(val), p(indices), p(indptr), p(b), damp, atol, btol, conlim, int(itnlim),
            int(local_size), p(x), C.byref(ist), C.byref(itn), *[C.byref(o) for o in outs]]
    rc = L.jdsg_lsmr_csr_layers(*args, int(layers), stream) if layers != 1 else \
        L.jdsg_lsmr_csr(*args, stream)
    if rc:
        raise RuntimeError(L.jdsg_lsmr_last_error().decode())
    info = {"istop": ist.value, "itn": itn.value}
    info.update({k: o.value for k, o in zip(("normA", "condA", "normr", "normAr", "normx"), outs)})
    return x, info
