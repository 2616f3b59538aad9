"""GPU smoothing and regularisation rows against the oracle restatement of
src/shared/smoothing.cpp, src/surftomo/surftomo.cpp:147-178 and
src/shared/csr_matrix.cpp:268-328: the same operations in the same order
(float stencil products, double sums, contraction off), so the fields agree
to float rounding and the rows bit for bit."""
import numpy as np
import pytest

from oracle import pyoracle

pytestmark = pytest.mark.gpu


def _cuda(a, dt=None):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _close(got, want, what):
    err = np.abs(got.astype(np.float64) - want).max() / max(np.abs(want).max(), 1e-30)
    same = float((got == want).mean())
    print(f"{what}: max rel err {err:.2e}, bit-identical {same:.4f}", flush=True)
    assert err <= 1e-6


@pytest.mark.parametrize("sh,sv", [(2.0, 1.0), (1.5, 3.0)])
def test_gpu_cart_smoothing_matches_oracle(sh, sv):
    from jdsurfg_amd import smooth as S
    print(f"case smooth cart sigma {sh} {sv}", flush=True)
    g = np.random.default_rng(3).normal(size=(12, 25, 30)).astype(np.float32)
    d = _cuda(g)
    S.smooth_cart_pde(d, sh, sv)
    _close(d.cpu().numpy(), pyoracle.smooth_cart_pde(g, sh, sv), "cart")


def test_gpu_sph_smoothing_matches_oracle():
    from jdsurfg_amd import smooth as S
    print("case smooth sph", flush=True)
    g = np.random.default_rng(4).normal(size=(10, 18, 20)).astype(np.float32)
    args = (0.5, 0.5, 5.0, 40.0, 100.0, 0.0, 30.0, 10.0)
    d = _cuda(g)
    S.smooth_sph_pde(d, *args)
    _close(d.cpu().numpy(), pyoracle.smooth_sph_pde(g, *args), "sph")


def test_gpu_depth_interpolation_matches_oracle():
    import torch
    from jdsurfg_amd import smooth as S
    print("case interp z", flush=True)
    dep = np.array([0, 1, 2.5, 4, 7, 12, 20, 33], np.float32)
    g = np.random.default_rng(5).normal(size=(dep.size, 7, 9)).astype(np.float32)
    nz1 = 34
    d, dr = _cuda(g), torch.zeros((nz1, 7, 9), dtype=torch.float32, device="cuda")
    S.interp_irregular_z(d, dr, dep, True)
    want_r = pyoracle.interp_irregular_z(g, nz1, dep, True)
    assert np.array_equal(dr.cpu().numpy(), want_r)
    back = torch.zeros_like(d)
    S.interp_irregular_z(back, dr, dep, False)
    want_b = pyoracle.interp_irregular_z(np.zeros_like(g), nz1, dep, False, gradr=want_r)
    assert np.array_equal(back.cpu().numpy(), want_b)


@pytest.mark.parametrize("in_km", [0, 1])
def test_gpu_smooth_gradient_driver_matches_oracle(in_km):
    from jdsurfg_amd import smooth as S
    print(f"case smooth gradient in_km {in_km}", flush=True)
    nx, ny = 22, 20
    dep = np.array([0, 2, 4, 7, 10, 15, 20, 30, 40, 60], np.float32)
    lat = (35.0 - 0.25 * np.arange(nx)).astype(np.float32)
    lon = (100.0 + 0.25 * np.arange(ny)).astype(np.float32)
    g = np.random.default_rng(6).normal(size=(dep.size - 1, ny - 2, nx - 2)).astype(np.float32)
    sh, sv = (25.0, 8.0) if in_km else (2.0, 1.5)
    d = _cuda(g)
    S.smooth_gradient(d, dep, lat, lon, in_km, sh, sv)
    want = pyoracle.smooth_gradient(g, nx, ny, dep.size, dep, lat, lon, in_km, sh, sv)
    _close(d.cpu().numpy(), want, f"driver in_km={in_km}")


def test_gpu_regularization_rows_are_the_oracle_rows():
    import torch
    from jdsurfg_amd import smooth as S
    print("case regularization rows", flush=True)
    nx, ny, nz, m, w = 9, 7, 5, 3, 0.35
    n = nx * ny * nz
    ptr = np.zeros(m + n + 1, np.int32)
    ptr[1:m + 1] = [4, 9, 11]
    cap = 11 + 7 * n
    col = np.zeros(cap, np.int32)
    val = np.zeros(cap, np.float32)
    col[:11] = np.arange(11)
    val[:11] = 1.0
    wp, wc, wv, wn = pyoracle.add_regularization(ptr, col, val, m, w, nx, ny, nz)
    dp, dc, dv = _cuda(ptr), _cuda(col), _cuda(val)
    nnz = S.add_regularization(dp, dc, dv, m, w, nx, ny, nz)
    assert nnz == wn
    assert np.array_equal(dp.cpu().numpy(), wp)
    assert np.array_equal(dc.cpu().numpy()[:nnz], wc[:nnz])
    assert np.array_equal(dv.cpu().numpy()[:nnz], wv[:nnz])
    small = torch.zeros(nnz - 1, dtype=torch.int32, device="cuda")
    with pytest.raises(RuntimeError, match="sparse ratio"):
        S.add_regularization(dp, small, dv, m, w, nx, ny, nz)


def test_gpu_sph_smoother_on_the_reference_test_case():
    """The reference's own smoothing case (test/smooth): the GPU smoother gives
    the oracle's bits and stays within 1 % of the direct convolution."""
    from jdsurfg_amd import smooth as S
    from tests.test_smooth import REF_CASE, box_field
    print("case smooth reference test/smooth", flush=True)
    g = box_field()
    d = _cuda(g)
    S.smooth_sph_pde(d, *REF_CASE)
    got = d.cpu().numpy()
    assert np.array_equal(got, pyoracle.smooth_sph_pde(g, *REF_CASE))
    assert np.abs(got - pyoracle.smooth_sph_conv(g, *REF_CASE)).max() <= 0.02
