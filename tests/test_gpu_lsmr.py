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
                                                    if k not in ("local_size", "layers")},
                           local_size=kw.get("local_size", 10))
    err = np.abs(x - xr).max() / np.abs(xr).max()
    print(f"gpu {info}\noracle {ir}\nmax |dx| / max |x| {err:.2e}", flush=True)
    assert info["istop"] == ir["istop"]
    assert abs(info["itn"] - ir["itn"]) <= 2
    assert err <= 1e-3
    # the residual estimate of a consistent system ends at rounding noise,
    # which is measured against ||b|| (the scale of the btol test)
    assert abs(info["normr"] - ir["normr"]) <= 1e-3 * ir["normr"] + 1e-5 * np.linalg.norm(b)


@pytest.mark.parametrize("m,n,density,damp,local",
                         [(300, 120, 0.05, 0.0, 10), (300, 120, 0.05, 0.5, 10),
                          (80, 200, 0.1, 0.0, 10), (2000, 700, 0.01, 0.1, 10),
                          (2000, 700, 0.01, 0.0, 0),
                          # rows of ~1200 entries (64 lanes per row) / columns of ~1200
                          (64, 4000, 0.3, 0.0, 10), (4000, 64, 0.3, 0.1, 10),
                          # rows of ~6000 entries (a block of 256 lanes per row)
                          (32, 20000, 0.3, 0.0, 10)])
def test_gpu_lsmr_matches_oracle(m, n, density, damp, local):
    print(f"case lsmr {m}x{n} density {density} damp {damp} local {local}", flush=True)
    A, ptr, ind, val, b = sparse_problem(m, n, density, m + n)
    _compare(m, n, val, ind, ptr, b, damp=damp, local_size=local)


@pytest.mark.parametrize("layers", [1, 5])
def test_gpu_lsmr_on_frechet_rows(layers):
    """Rows of straight rays through a 16x14x6 grid (the Frechet assembly of
    the GPU path), b from a known model plus noise, damped."""
    from jdsurfg_amd import frechet as F
    import torch
    nx, ny, nz, kmax, nrows = 16, 14, 6, 4, 400
    print(f"case lsmr on frechet rows, layers {layers}", flush=True)
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
    # layers = nz - 1: the solve runs cell-major, x comes back layer-major
    _compare(nrows, n, v, ind, ptr, b, damp=0.05, local_size=10, layers=layers)


def test_gpu_lsmr_rejects_bad_matrix():
    A, ptr, ind, val, b = sparse_problem(50, 20, 0.1, 3)
    bad = ind.copy()
    bad[5] = 20                       # column index out of range
    with pytest.raises(RuntimeError, match="column index"):
        _gpu(50, 20, val, bad, ptr, b)
    x, info = _gpu(50, 20, val, ind, ptr, b)   # a good call still works
    assert info["istop"] in (1, 2, 3)


def test_gpu_lsmr_unaligned_arrays_take_the_scalar_products():
    """Value and index arrays that are not 16-byte aligned (views one element
    into their buffers) are read one entry per lane; the solve matches the
    oracle as the aligned one does."""
    import torch
    from jdsurfg_amd import lsmr as G
    m, n = 2000, 700
    print("case lsmr unaligned 2000x700", flush=True)
    A, ptr, ind, val, b = sparse_problem(m, n, 0.01, 9)
    vbuf = torch.zeros(val.size + 1, dtype=torch.float32, device="cuda")
    ibuf = torch.zeros(ind.size + 1, dtype=torch.int32, device="cuda")
    vbuf[1:] = torch.from_numpy(val).cuda()
    ibuf[1:] = torch.from_numpy(ind).cuda()
    vv, iv = vbuf[1:], ibuf[1:]
    assert vv.data_ptr() % 16 and iv.data_ptr() % 16
    x, info = G.lsmr(m, n, vv, iv, torch.from_numpy(ptr).cuda(), torch.from_numpy(b).cuda(),
                     damp=0.1)
    xr, ir = pyoracle.lsmr(m, n, val, ind, ptr, b, damp=0.1)
    x = x.cpu().numpy()
    assert info["istop"] == ir["istop"] and abs(info["itn"] - ir["itn"]) <= 2
    assert np.abs(x - xr).max() <= 1e-3 * np.abs(xr).max()


def test_gpu_lsmr_at_frechet_width_tracks_double():
    """At a million model cells (254 x 254 x 16, the rows of 1500 straight
    rays with random depth kernels), three iterations: the GPU iterate stays
    within 1e-4 of the double-precision iterate, while the float restatement
    (the reference's own bits) drifts further, its million-term float sums
    being the larger error."""
    import torch
    from jdsurfg_amd import frechet as F
    nx = ny = 256
    nz, nrows = 17, 1500
    print("case lsmr frechet width 1500 x 1032256", flush=True)
    off, cell, w = F.straight_rays(nx, ny, nrows, 0)
    rng = np.random.default_rng(0)
    nin = (nx - 2) * (ny - 2)
    ptr, inds, vals = [0], [], []
    for r in range(nrows):
        c, ww = cell[off[r]:off[r + 1]], w[off[r]:off[r + 1]]
        ix, iy = c % nx, c // nx
        ok = (ix >= 1) & (ix <= nx - 2) & (iy >= 1) & (iy <= ny - 2)
        ix, iy, ww = ix[ok], iy[ok], ww[ok]
        for iz in range(nz - 1):
            inds.append(iz * nin + (iy - 1) * (nx - 2) + (ix - 1))
            vals.append(ww * rng.normal(size=ww.size))
        ptr.append(ptr[-1] + ww.size * (nz - 1))
    ind = np.concatenate(inds).astype(np.int32)
    val = np.concatenate(vals).astype(np.float32)
    ptr = np.array(ptr, np.int32)
    m, n = nrows, nin * (nz - 1)
    b = rng.normal(size=m).astype(np.float32)
    kw = dict(damp=0.0, atol=0.0, btol=0.0, conlim=0.0, itnlim=3, local_size=10)
    x64, i64 = pyoracle.lsmr(m, n, val, ind, ptr, b, kind="f64", **kw)
    xf, _ = pyoracle.lsmr(m, n, val, ind, ptr, b, kind="rest", **kw)
    scale = np.abs(x64).max()
    e_ref = np.abs(xf - x64).max() / scale
    for layers in (1, nz - 1):
        x, info = _gpu(m, n, val, ind, ptr, b, layers=layers, **kw)
        e_gpu = np.abs(x - x64).max() / scale
        print(f"vs double: gpu (layers {layers}) {e_gpu:.2e}, float restatement {e_ref:.2e}",
              flush=True)
        assert info["itn"] == 3 and info["istop"] == 7
        assert e_gpu <= 1e-4
        assert e_gpu <= e_ref


def test_gpu_lsmr_layers_must_divide_columns():
    A, ptr, ind, val, b = sparse_problem(50, 21, 0.1, 4)
    with pytest.raises(RuntimeError, match="layers"):
        _gpu(50, 21, val, ind, ptr, b, layers=4)
