"""GPU LSMR (jdsg_lsmr_csr) against the oracle restatement of
src/shared/lsmrModule_csr.f90, which gives the compiled reference's bits:
the same stopping code, the iteration count within two, and the solution to
the accuracy the stop rules (atol = btol = 1e-5) leave in float; on random
sparse systems (over- and underdetermined, damped, without local
reorthogonalisation) and on a system of Frechet rows built on the GPU."""
import numpy as np
import pytest

from oracle import pyoracle
from tests.test_lsmr import sparse_problem

pytestmark = pytest.mark.gpu


def _gpu(m, n, val, ind, ptr, b, **kw):
    import torch
    from jdsurfg_amd import lsmr as G
    t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dt)
    x, info = G.lsmr(m, n, t(val, torch.float32), t(ind, torch.int32), t(ptr, torch.int32),
                     t(b, torch.float32), **kw)
    torch.cuda.synchronize()
    return x.cpu().numpy(), info


def _compare(m, n, val, ind, ptr, b, **kw):
    x, info = _gpu(m, n, val, ind, ptr, b, **kw)
    xr, ir = pyoracle.lsmr(m, n, val, ind, ptr, b, **{k: v for k, v in kw.items()
                                                    if k != "local_size"},
                           local_size=kw.get("local_size", 10))
    err = np.abs(x - xr).max() / np.abs(xr).max()
    print(f"gpu {info}\noracle {ir}\nmax |dx| / max |x| {err:.2e}", flush=True)
    assert info["istop"] == ir["istop"]
    assert abs(info["itn"] - ir["itn"]) <= 2
    assert err <= 1e-3
    assert abs(info["normr"] - ir["normr"]) <= 1e-3 * max(ir["normr"], 1e-12) + 1e-6


@pytest.mark.parametrize("m,n,density,damp,local",
                         [(300, 120, 0.05, 0.0, 10), (300, 120, 0.05, 0.5, 10),
                          (80, 200, 0.1, 0.0, 10), (2000, 700, 0.01, 0.1, 10),
                          (2000, 700, 0.01, 0.0, 0),
                          # rows of ~1200 entries (64 lanes per row) / columns of ~1200
                          (64, 4000, 0.3, 0.0, 10), (4000, 64, 0.3, 0.1, 10)])
def test_gpu_lsmr_matches_oracle(m, n, density, damp, local):
    print(f"case lsmr {m}x{n} density {density} damp {damp} local {local}", flush=True)
    A, ptr, ind, val, b = sparse_problem(m, n, density, m + n)
    _compare(m, n, val, ind, ptr, b, damp=damp, local_size=local)


def test_gpu_lsmr_on_frechet_rows():
    """Rows of straight rays through a 16x14x6 grid (the Frechet assembly of
    the GPU path), b from a known model plus noise, damped."""
    from jdsurfg_amd import frechet as F
    import torch
    nx, ny, nz, kmax, nrows = 16, 14, 6, 4, 400
    print("case lsmr on frechet rows", flush=True)
    rng = np.random.default_rng(5)
    K = rng.normal(size=(nx * ny, nz, kmax)) * 0.1
    off, cell, w = F.straight_rays(nx, ny, nrows, seed=6)
    idx = rng.integers(0, kmax, nrows).astype(np.int32)
    Kd = torch.from_numpy(np.ascontiguousarray(K.ravel(order="F"))).cuda()
    ro, col, val = F.frechet_rows(Kd, nx, ny, nz, kmax, idx, off, cell, w)
    ptr = ro.cpu().numpy().astype(np.int32)
    ind = col.cpu().numpy().astype(np.int32)
    v = val.cpu().numpy().astype(np.float32)
    n = (nx - 2) * (ny - 2) * (nz - 1)
    xt = rng.normal(size=n).astype(np.float32)
    A = np.zeros((nrows, n), np.float64)
    for r in range(nrows):
        np.add.at(A[r], ind[ptr[r]:ptr[r + 1]], v[ptr[r]:ptr[r + 1]])
    b = (A @ xt + 0.01 * rng.normal(size=nrows)).astype(np.float32)
    _compare(nrows, n, v, ind, ptr, b, damp=0.05, local_size=10)


def test_gpu_lsmr_rejects_bad_matrix():
    A, ptr, ind, val, b = sparse_problem(50, 20, 0.1, 3)
    bad = ind.copy()
    bad[5] = 20                       # column index out of range
    with pytest.raises(RuntimeError, match="column index"):
        _gpu(50, 20, val, bad, ptr, b)
    x, info = _gpu(50, 20, val, ind, ptr, b)   # a good call still works
    assert info["istop"] in (1, 2, 3)
