"""GPU Fréchet rows and misfit gradient (SurfaceWave.cpp:601-641, 738-773)
against the oracle: rows bit for bit (the value is one double product rounded
to float, so any correct evaluation gives the same bits), the gradient to
float summation order, and the error paths (capacity, invalid rays)."""
import numpy as np
import pytest

from oracle import pyoracle

pytestmark = pytest.mark.gpu


def _case(nx, ny, nz, kmax, nrows, seed):
    from jdsurfg_amd import frechet as F
    rng = np.random.default_rng(seed)
    K = rng.normal(size=(nx * ny, nz, kmax))
    K[rng.random(K.shape) < 0.05] = 0.0     # some exact zeros drop out of the rows
    off, cell, w = F.straight_rays(nx, ny, nrows, seed)
    idx = rng.integers(0, kmax, nrows).astype(np.int32)
    return K, off, cell, w, idx


def _upload_kernel(K):
    import torch
    return torch.from_numpy(np.ascontiguousarray(K.ravel(order="F"))).cuda()


@pytest.mark.parametrize("shape", [(12, 9, 5, 7, 60), (33, 20, 9, 3, 200)])
def test_rows_match_oracle_bitwise(shape):
    from jdsurfg_amd import frechet as F
    nx, ny, nz, kmax, nrows = shape
    print(f"case frechet rows grid {nx}x{ny}x{nz} kmax {kmax} rows {nrows}", flush=True)
    K, off, cell, w, idx = _case(nx, ny, nz, kmax, nrows, seed=nx)
    ro, col, val = F.frechet_rows(_upload_kernel(K), nx, ny, nz, kmax, idx, off, cell, w)
    o = pyoracle.Oracle("rest")
    rro, rcol, rval = o.frechet_rows(K, idx, F.dense_from_rays(nx, ny, off, cell, w))
    assert np.array_equal(ro.cpu().numpy(), rro)
    assert np.array_equal(col.cpu().numpy(), rcol)
    assert np.array_equal(val.cpu().numpy(), rval)


def test_gradient_matches_oracle():
    import torch
    from jdsurfg_amd import frechet as F
    nx, ny, nz, kmax, nrows = 24, 18, 7, 5, 300
    print("case frechet gradient", flush=True)
    K, off, cell, w, idx = _case(nx, ny, nz, kmax, nrows, seed=9)
    rng = np.random.default_rng(10)
    resid = rng.normal(size=nrows).astype(np.float32)
    vs = rng.uniform(2.5, 4.5, (nx, ny, nz)).astype(np.float32)
    g = F.frechet_grad(_upload_kernel(K), nx, ny, nz, kmax, idx, off, cell, w, resid, vs)
    torch.cuda.synchronize()
    o = pyoracle.Oracle("rest")
    ref = o.frechet_grad(K, idx, F.dense_from_rays(nx, ny, off, cell, w), resid, vs)
    g = g.cpu().numpy()
    # float sums of up to nrows terms in another order
    assert np.abs(g - ref).max() <= 2e-5 * np.abs(ref).max()


def test_rows_from_surftime_kernels():
    """The kernels of the SurfTime entry point, as they stay on the GPU, feed
    the rows: same bits as the oracle rows over the same kernel values."""
    from jdsurfg_amd import frechet as F
    from jdsurfg_amd import swd
    nx, ny, nz = 7, 6, 6
    print("case frechet rows over SurfTime kernels 7x6x6", flush=True)
    dep = np.linspace(0.0, 40.0, nz).astype(np.float32)
    rng = np.random.default_rng(11)
    vs = (2.8 + 0.03 * dep[None, None, :] + 0.1 * rng.normal(size=(nx, ny, nz))).astype(np.float32)
    t = np.geomspace(5.0, 40.0, 6).reshape(-1, 1)
    vc, vout, kernel, status = swd.get_1d_kernel(vs, dep, {"Rc": t, "Rg": t})
    assert (status == 0).all()
    kmax = kernel.shape[2]
    off, cell, w = F.straight_rays(nx, ny, 40, seed=12)
    idx = rng.integers(0, kmax, 40).astype(np.int32)
    ro, col, val = F.frechet_rows(_upload_kernel(kernel), nx, ny, nz, kmax, idx, off, cell, w)
    o = pyoracle.Oracle("rest")
    rro, rcol, rval = o.frechet_rows(kernel, idx, F.dense_from_rays(nx, ny, off, cell, w))
    assert np.array_equal(ro.cpu().numpy(), rro)
    assert np.array_equal(col.cpu().numpy(), rcol)
    assert np.array_equal(val.cpu().numpy(), rval)


def test_capacity_and_invalid_rays_are_errors():
    from jdsurfg_amd import frechet as F
    nx, ny, nz, kmax = 10, 8, 4, 3
    print("case frechet error paths", flush=True)
    K, off, cell, w, idx = _case(nx, ny, nz, kmax, 10, seed=13)
    Kd = _upload_kernel(K)
    with pytest.raises(RuntimeError, match="capacity"):
        F.frechet_rows(Kd, nx, ny, nz, kmax, idx, off, cell, w, cap=5)
    bad = cell.copy()
    bad[off[3]], bad[off[3] + 1] = bad[off[3] + 1], bad[off[3]]    # not increasing
    with pytest.raises(RuntimeError, match="increasing"):
        F.frechet_rows(Kd, nx, ny, nz, kmax, idx, off, bad, w)
    bidx = idx.copy()
    bidx[2] = kmax                                                 # period index out of range
    with pytest.raises(RuntimeError, match="period index"):
        F.frechet_rows(Kd, nx, ny, nz, kmax, bidx, off, cell, w)
    ro, col, val = F.frechet_rows(Kd, nx, ny, nz, kmax, idx, off, cell, w)   # still fine
    assert int(ro[-1].item()) == col.numel() > 0
