"""The allocation-free batch solve (jdsg_plan_execute_static) against the
planned solve (jdsg_plan_create + jdsg_plan_execute) through the C ABI:
identical outputs on two calls in a row, the second with a smaller batch
(the static scratch keeps the first call's contents), in both kernel
layouts; a batch beyond the scratch is refused with an error and leaves the
library usable."""
import ctypes as C

import numpy as np
import pytest

from jdsurfg_amd import workloads as W

pytestmark = pytest.mark.gpu
NAMES = ("c", "u", "status", "dcda", "dcdb", "dcdr", "dcdh", "duda", "dudb", "dudr", "dudh")


def _lib():
    from jdsurfg_amd import swd
    L = swd.lib()
    vp = C.c_void_p
    L.jdsg_plan_execute_static.argtypes = [C.c_int, vp, vp, vp, C.c_int, C.c_int, C.c_int] + \
        [vp] * 16 + [vp]
    L.jdsg_static_scratch_bytes.restype = C.c_longlong
    return L


def _solve_both(w, layout):
    import torch
    from jdsurfg_amd import swd
    L = _lib()
    plan = swd.Plan(w.layer_off, w.per_off, w.nmode, w.group, layout)
    inp = swd.upload(w)
    ref = plan.alloc_outputs()
    plan.execute(*inp, ref)
    out = plan.alloc_outputs()
    for t in out.values():
        t.fill_(7)   # stale values the solve must overwrite
    lo = np.ascontiguousarray(w.layer_off, np.int32)
    po = np.ascontiguousarray(w.per_off, np.int32)
    nm = np.ascontiguousarray(w.nmode.ravel(), np.int32)
    p = lambda t: C.c_void_p(t.data_ptr())
    rc = L.jdsg_plan_execute_static(w.ncol, lo.ctypes.data, po.ctypes.data, nm.ctypes.data,
                                    int(w.group), {"grid": 0, "csr": 1}[layout], 1,
                                    *[p(t) for t in inp], *[p(out[n]) for n in NAMES],
                                    C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0, swd.last_error()
    torch.cuda.synchronize()
    return ref, out


@pytest.mark.parametrize("layout", ["csr", "grid"])
def test_static_matches_planned_solve_twice(layout):
    import torch
    big = W.random_columns(12, 30, 31, nmode=(3, 2), group=True)
    small = W.random_columns(5, 20, 32, nmode=(2, 1), group=True, sediments=True)
    if layout == "csr":
        big = W.uneven(ncol=10, seed=33)
    for name, w in (("first", big), ("second, smaller", small)):
        print(f"case static solve {layout} {name}: ncol={w.ncol}", flush=True)
        ref, out = _solve_both(w, layout)
        for n in NAMES:
            assert torch.equal(ref[n], out[n]), (layout, name, n)


def test_static_rejects_a_batch_beyond_its_scratch():
    import torch
    from jdsurfg_amd import swd
    L = _lib()
    cap = L.jdsg_static_scratch_bytes()
    ncol, N, K = 20000, 129, 8
    lo = (np.arange(ncol + 1) * N).astype(np.int32)
    po = (np.arange(ncol + 1) * K).astype(np.int32)
    nm = np.tile(np.array([1, 0], np.int32), ncol)
    print(f"case static solve beyond its {cap >> 20} MiB scratch", flush=True)
    d = torch.zeros(8, device="cuda")
    rc = L.jdsg_plan_execute_static(ncol, lo.ctypes.data, po.ctypes.data, nm.ctypes.data, 0, 1,
                                    1, *[C.c_void_p(d.data_ptr())] * 16, None)
    assert rc != 0 and "too large" in swd.last_error()
    # the library still solves afterwards
    ref, out = _solve_both(W.random_columns(3, 12, 34, nmode=(1, 1), group=False), "grid")
    for n in ("c", "u", "status", "dcdb"):
        assert torch.equal(ref[n], out[n])
