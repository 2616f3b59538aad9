"""Model smoothing and regularisation rows on the GPU (libjdsg_swd.so):
smooth_cart_pde / smooth_sph_pde / interp_irregular_z of
src/shared/smoothing.cpp, DSurfTomo::smoothing (src/surftomo/surftomo.cpp:
147-178) and add_regularization (src/shared/csr_matrix.cpp:268-328).
Arrays are CUDA tensors; 3-D fields are (nz, ny, nx) tensors (x fastest, the
reference's column-major (nx, ny, nz))."""
import ctypes as C

import numpy as np

from . import swd


def _L():
    L = swd.lib()
    vp, fl, it = C.c_void_p, C.c_float, C.c_int
    L.jdsg_smooth_cart_pde.argtypes = [vp, it, it, it, fl, fl, vp]
    L.jdsg_smooth_sph_pde.argtypes = [vp, it, it, it] + [fl] * 8 + [vp]
    L.jdsg_interp_irregular_z.argtypes = [vp, vp, it, it, it, it, vp, it, vp]
    L.jdsg_smooth_gradient.argtypes = [vp, it, it, it, vp, vp, vp, it, fl, fl, vp]
    L.jdsg_add_regularization.argtypes = [it, vp, vp, vp, C.c_longlong, fl, it, it, it,
                                          C.POINTER(C.c_longlong), vp]
    L.jdsg_smooth_last_error.restype = C.c_char_p
    return L


def _p(t):
    return C.c_void_p(t.data_ptr())


def _st(stream):
    import torch
    return C.c_void_p(stream if stream is not None else torch.cuda.current_stream().cuda_stream)


def _check(L, rc, what):
    if rc:
        raise RuntimeError(f"{what}: {L.jdsg_smooth_last_error().decode()}")


def _host(a):
    return np.ascontiguousarray(np.asarray(a, np.float32))


def smooth_cart_pde(grad, sigma_h, sigma_v, stream=None):
    """Smooth a contiguous float32 CUDA tensor (nz, ny, nx) in place."""
    L = _L()
    nz, ny, nx = grad.shape
    _check(L, L.jdsg_smooth_cart_pde(_p(grad), nx, ny, nz, sigma_h, sigma_v, _st(stream)),
           "smooth_cart_pde")
    return grad


def smooth_sph_pde(grad, dx, dy, dz, lat0, lon0, z0, sigma_h, sigma_v, stream=None):
    """Spherical form of smooth_cart_pde (nodes dx, dy degrees, dz km apart)."""
    L = _L()
    nz, ny, nx = grad.shape
    _check(L, L.jdsg_smooth_sph_pde(_p(grad), nx, ny, nz, dx, dy, dz, lat0, lon0, z0, sigma_h,
                                    sigma_v, _st(stream)), "smooth_sph_pde")
    return grad


def interp_irregular_z(grad, gradr, dep, fwd, stream=None):
    """grad (nz, ny, nx) <-> gradr (nz1, ny, nx) on the host depths dep[nz]."""
    L = _L()
    nz, ny, nx = grad.shape
    d = _host(dep)
    _check(L, L.jdsg_interp_irregular_z(_p(grad), _p(gradr), nx, ny, nz, gradr.shape[0],
                                        d.ctypes.data_as(C.c_void_p), int(fwd), _st(stream)),
           "interp_irregular_z")


def smooth_gradient(grad, dep, lat, lon, smooth_in_km, sigma_h, sigma_v, stream=None):
    """DSurfTomo::smoothing on grad (nz-1, ny-2, nx-2) of a model with host axes
    lat[nx], lon[ny], dep[nz]; in place."""
    L = _L()
    d, la, lo = _host(dep), _host(lat), _host(lon)
    _check(L, L.jdsg_smooth_gradient(_p(grad), la.size, lo.size, d.size,
                                     d.ctypes.data_as(C.c_void_p), la.ctypes.data_as(C.c_void_p),
                                     lo.ctypes.data_as(C.c_void_p), int(smooth_in_km), sigma_h,
                                     sigma_v, _st(stream)), "smooth_gradient")
    return grad


def add_regularization(ptr, col, val, m, weight, nx, ny, nz, stream=None):
    """Append the Laplacian rows after the m rows of a CSR held in int32 /
    float32 CUDA tensors with room (ptr: m + nx*ny*nz + 1); returns nnz."""
    L = _L()
    nnz = C.c_longlong(0)
    _check(L, L.jdsg_add_regularization(m, _p(ptr), _p(col), _p(val), min(col.numel(),
                                                                          val.numel()),
                                        weight, nx, ny, nz, C.byref(nnz), _st(stream)),
           "add_regularization")
    return nnz.value
