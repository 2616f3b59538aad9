"""The GPU column preparation (Earth flattening of the root-search model and
the starting velocity, surfdisp96.f:148-221,499-571) is bit-identical to the
oracle's."""
import ctypes as C

import numpy as np
import pytest

from jdsurfg_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_chain_model_bitwise(oracle_rest):
    from jdsurfg_amd import swd
    w = W.grid128(nx=8, ny=8)
    plan, out = swd.solve(w)
    L = swd.lib()
    N = int(w.layer_off[1])
    npad = C.c_int()
    R = np.zeros((plan.nmax * 64 * ((w.ncol + 63) // 64), 4), np.float32)
    Lv = np.zeros_like(R)
    cc = np.zeros(w.ncol, np.float32)
    bx = np.zeros(w.ncol, np.float32)
    L.jdsg_plan_chain_model.argtypes = [C.c_void_p] * 5 + [C.POINTER(C.c_int)]
    assert L.jdsg_plan_chain_model(plan.h, R.ctypes.data, Lv.ctypes.data, cc.ctypes.data,
                                   bx.ctypes.data, C.byref(npad)) == 0
    O = oracle_rest.lib
    O.or_chain_model.argtypes = [C.c_void_p] * 4 + [C.c_int, C.c_int] + [C.c_void_p] * 7
    print("case grid128 8x8 chain model", flush=True)
    for col in range(w.ncol):
        s = slice(col * N, (col + 1) * N)
        arrs = [np.ascontiguousarray(a[s]) for a in (w.thk, w.vp, w.vs, w.rho)]
        d, a, b, rR, rL = (np.zeros(N, np.float32) for _ in range(5))
        c1 = np.zeros(1, np.float32)
        b1 = np.zeros(1, np.float32)
        O.or_chain_model(*[x.ctypes.data for x in arrs], N, 1, d.ctypes.data, a.ctypes.data,
                         b.ctypes.data, rR.ctypes.data, rL.ctypes.data, c1.ctypes.data,
                         b1.ctypes.data)
        g = R.reshape(-1, npad.value, 4)[:N, col]
        gl = Lv.reshape(-1, npad.value, 4)[:N, col]
        assert np.array_equal(g[:, 0], d) and np.array_equal(g[:, 1], a)
        assert np.array_equal(g[:, 2], b)
        assert np.array_equal(g[:, 3], rR), (col, np.nonzero(g[:, 3] != rR))
        assert np.array_equal(gl[:, 3], rL)
        assert cc[col] == c1[0] and bx[col] == b1[0]
