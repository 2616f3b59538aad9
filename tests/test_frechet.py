"""CPU checks of the Fréchet-row oracle (SurfaceWave.cpp:601-641, 738-773),
the sparse ray form and the frechet.bin writer.

The oracle is pinned by what the definition fixes: a ray on one interior
cell gives that cell's kernel column down the depths (so a swapped ix/iy or
a wrong depth stride fails), boundary cells contribute nothing, rows of rays
with disjoint cells add, and the gradient is the residual-weighted sum of the
rows times vs.
"""
import os

import numpy as np
import pytest

from oracle import pyoracle

pytestmark = pytest.mark.skipif(not pyoracle.available("rest"), reason="oracle not built")


def _kernel(nx, ny, nz, kmax, seed=0):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(nx * ny, nz, kmax))


def _dense(rows):
    return np.stack(rows).astype(np.float32)


def test_single_cell_ray_is_the_kernel_column():
    o = pyoracle.Oracle("rest")
    nx, ny, nz, kmax = 7, 5, 6, 4
    K = _kernel(nx, ny, nz, kmax)
    ix, iy, idx, w = 4, 2, 3, np.float32(-0.37)
    f = np.zeros((nx, ny), np.float32)
    f[ix, iy] = w
    ro, col, val = o.frechet_rows(K, [idx], f[None])
    n = ix + iy * nx
    nin = (nx - 2) * (ny - 2)
    want_c = np.arange(nz - 1) * nin + (iy - 1) * (nx - 2) + (ix - 1)
    want_v = (np.float64(w) * K[n, :nz - 1, idx]).astype(np.float32)
    assert ro.tolist() == [0, nz - 1]
    assert np.array_equal(col, want_c)
    assert np.array_equal(val, want_v)


def test_boundary_cells_and_zero_kernels_drop_out():
    o = pyoracle.Oracle("rest")
    nx, ny, nz, kmax = 6, 6, 4, 2
    K = _kernel(nx, ny, nz, kmax, 1)
    K[2 + 3 * nx, 1, 0] = 0.0           # one zero kernel value
    f = np.zeros((nx, ny), np.float32)
    f[0, 3] = f[5, 2] = f[3, 0] = f[2, 5] = 1.0   # boundary: no entries
    f[2, 3] = 2.0
    ro, col, val = o.frechet_rows(K, [0], f[None])
    assert ro[1] == nz - 2              # nz-1 depths minus the zero one
    assert np.all(val != 0)


def test_rows_of_disjoint_rays_add():
    o = pyoracle.Oracle("rest")
    nx, ny, nz, kmax = 9, 8, 5, 3
    K = _kernel(nx, ny, nz, kmax, 2)
    rng = np.random.default_rng(3)
    a = np.zeros((nx, ny), np.float32)
    b = np.zeros((nx, ny), np.float32)
    a[1:4, 1:7] = rng.normal(size=(3, 6))
    b[5:8, 1:7] = rng.normal(size=(3, 6))
    _, ca, va = o.frechet_rows(K, [1], a[None])
    _, cb, vb = o.frechet_rows(K, [1], b[None])
    _, cs, vs = o.frechet_rows(K, [1], (a + b)[None])
    order = np.argsort(np.concatenate([ca, cb]), kind="stable")
    assert np.array_equal(cs, np.concatenate([ca, cb])[order])
    assert np.array_equal(vs, np.concatenate([va, vb])[order])


def test_gradient_is_the_weighted_sum_of_rows():
    o = pyoracle.Oracle("rest")
    nx, ny, nz, kmax = 8, 7, 5, 3
    K = _kernel(nx, ny, nz, kmax, 4)
    rng = np.random.default_rng(5)
    f = rng.normal(size=(6, nx, ny)).astype(np.float32)
    f[f < 0.3] = 0.0
    idx = rng.integers(0, kmax, 6)
    resid = rng.normal(size=6).astype(np.float32)
    vs = rng.uniform(2, 4, (nx, ny, nz)).astype(np.float32)
    ro, col, val = o.frechet_rows(K, idx, f)
    g = o.frechet_grad(K, idx, f, resid, vs)
    nin = (nx - 2) * (ny - 2)
    want = np.zeros(nin * (nz - 1))
    for r in range(6):
        sl = slice(ro[r], ro[r + 1])
        np.add.at(want, col[sl], val[sl].astype(np.float64) * resid[r])
    c = np.arange(nin * (nz - 1))
    iz, rem = c // nin, c % nin
    iy, ix = rem // (nx - 2) + 1, rem % (nx - 2) + 1
    want *= vs[ix, iy, iz]
    assert np.allclose(g, want, rtol=1e-5, atol=1e-5 * np.abs(want).max())


def test_sparse_ray_form_round_trips():
    from jdsurfg_amd import frechet as F
    nx, ny = 11, 9
    off, cell, w = F.straight_rays(nx, ny, 20, seed=6)
    assert np.all(np.diff(off) > 0)
    for r in range(20):
        assert np.all(np.diff(cell[off[r]:off[r + 1]]) > 0)
    d = F.dense_from_rays(nx, ny, off, cell, w)
    off2, cell2, w2 = F.rays_from_dense(d)
    assert np.array_equal(off, off2) and np.array_equal(cell, cell2) and np.array_equal(w, w2)


def test_frechet_bin_round_trip(tmp_path):
    from jdsurfg_amd import frechet as F
    rng = np.random.default_rng(7)
    row_off = np.array([0, 3, 3, 7], np.int64)
    col = np.array([1, 5, 9, 0, 2, 4, 8], np.int32)
    val = rng.normal(size=7).astype(np.float32)
    path = os.path.join(tmp_path, "frechet.bin")
    F.write_frechet(path, 40, 123, [10, 11, 12], row_off, col, val)
    m, n, rows = F.read_frechet(path)
    assert (m, n) == (40, 123) and [r[0] for r in rows] == [10, 11, 12]
    for r, (_, c, v) in enumerate(rows):
        assert np.array_equal(c, col[row_off[r]:row_off[r + 1]])
        assert np.array_equal(v, val[row_off[r]:row_off[r + 1]])
    assert os.path.getsize(path) == 8 + 3 * 8 + 7 * 8
