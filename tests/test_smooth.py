"""CPU checks of the smoothing oracle (src/shared/smoothing.cpp) and the
regularisation rows (src/shared/csr_matrix.cpp:268-328), pinned by what the
mathematics fixes: the central-difference tables, a constant field kept,
mass conserved and the variance of a delta grown by exactly 2 C per step and
axis (the Gaussian width the PDE smoother is built for), one step of the
spherical smoother on fields whose Laplacian terms are known in closed form,
linear fields through the depth interpolation, and the row sums of the
Laplacian rows.  The reference itself needs Eigen (absent) and is not run."""
from fractions import Fraction

import numpy as np
import pytest

from oracle import pyoracle

pytestmark = pytest.mark.skipif(not pyoracle.available("rest"), reason="oracle not built")

TABLES = {(8, 2): ["-1/560", "8/315", "-1/5", "8/5", "-205/72", "8/5", "-1/5", "8/315", "-1/560"],
          (6, 2): ["1/90", "-3/20", "3/2", "-49/18", "3/2", "-3/20", "1/90"],
          (6, 1): ["-1/60", "3/20", "-3/4", "0", "3/4", "-3/20", "1/60"]}


@pytest.mark.parametrize("acc,deriv", list(TABLES))
def test_fd_coefficients_are_the_central_tables(acc, deriv):
    want = np.array([float(Fraction(t)) for t in TABLES[(acc, deriv)]], np.float32)
    assert np.array_equal(pyoracle.fdcoefs(acc, deriv), want)


def _cart_steps(sh, sv):
    eta = np.float32(1.0 / 12)
    sh2, sv2 = np.float32(sh) * np.float32(sh), np.float32(sv) * np.float32(sv)
    mx = max(sh2, sv2)
    nstep = int(float(mx) / (2.0 * float(eta)))
    return nstep, float(np.float32(sh2 / mx) * eta), float(np.float32(sv2 / mx) * eta)


def test_cart_smoother_keeps_constants_and_mass_and_grows_variance_by_2C_per_step():
    g = np.full((12, 9, 10), 3.25, np.float32)
    assert np.abs(pyoracle.smooth_cart_pde(g, 2.0, 1.0) - 3.25).max() <= 2e-6
    n = 41
    d = np.zeros((n, n, n), np.float32)
    d[20, 20, 20] = 1.0
    s = pyoracle.smooth_cart_pde(d, 2.0, 1.0).astype(np.float64)
    nstep, Ch, Cv = _cart_steps(2.0, 1.0)
    assert nstep == 23  # 4 / (2 * float(1/12)) truncates below 24
    assert abs(s.sum() - 1.0) <= 1e-5
    ax = np.arange(n) - 20.0
    var = [(s.sum(axis=tuple(a for a in range(3) if a != k)) * ax ** 2).sum() for k in range(3)]
    # axes of the (z, y, x) array: z gets Cv, y and x get Ch
    assert abs(var[0] - 2 * Cv * nstep) <= 1e-4 * 2 * Cv * nstep
    assert abs(var[1] - 2 * Ch * nstep) <= 1e-4 * 2 * Ch * nstep
    assert abs(var[2] - 2 * Ch * nstep) <= 1e-4 * 2 * Ch * nstep


# angular spacing comparable with the radial one, so that every term of the
# spherical step moves a grid-scale field well above float resolution
# (dx != dy, so that a dtheta put for a dphi shows)
SPH = dict(dx=0.02, dy=0.03, dz=2.0, lat0=40.0, lon0=100.0, z0=0.0)


def _one_step_sph(field, nx=24, ny=22, nz=20):
    """Run smooth_sph_pde for exactly one step on field(r, theta, phi) at the
    nodes' float coordinates; returns (before, after, geometry, coords)."""
    args = (nx, ny, nz, SPH["dx"], SPH["dy"], SPH["dz"], SPH["lat0"], SPH["lon0"], SPH["z0"])
    geo = pyoracle.sph_geometry(*args, 1.0, 1.0)
    h2 = geo["Ch"] / np.float32(1.0 / 7)                  # Ch = eta h2 when sigma_h = sigma_v
    sig = float(np.sqrt(1.5 * 2 * (1.0 / 7) * h2))        # max / (2 eta h2) = 1.5 -> 1 step
    geo = pyoracle.sph_geometry(*args, sig, sig)
    assert geo["nstep"] == 1
    r = geo["radi"][3:-3].astype(np.float64)
    th = geo["theta"][3:-3].astype(np.float64)
    ph0 = np.float32(np.deg2rad(np.float32(SPH["lon0"])))
    ph = ph0 + np.arange(ny) * np.float64(geo["dphi"])
    R, P, T = np.meshgrid(r, ph, th, indexing="ij")
    f = field(R, T, P).astype(np.float32)
    out = pyoracle.smooth_sph_pde(f, SPH["dx"], SPH["dy"], SPH["dz"], SPH["lat0"], SPH["lon0"],
                                  SPH["z0"], sig, sig)
    return f, out, geo, (R, T, P)


def _inner(a):
    return a[4:-4, 4:-4, 4:-4]


def _symbols(a):
    """S2(a) = sum_l c2_l cos(l a), S1(a) = sum_l c1_l sin(l a) of the
    sixth-order tables: the discrete second and first derivatives of
    cos(k x) are cos(k x) S2(k h) / h^2 and -sin(k x) S1(k h) / h."""
    l = np.arange(-3, 4)
    c2 = np.array([float(Fraction(t)) for t in TABLES[(6, 2)]])
    c1 = np.array([float(Fraction(t)) for t in TABLES[(6, 1)]])
    return (c2 * np.cos(l * a)).sum(), (c1 * np.sin(l * a)).sum()


@pytest.mark.parametrize("axis", ["theta", "phi", "r"])
def test_sph_smoother_one_step_matches_the_discrete_operator_on_cosines(axis):
    """One step adds Ch (f_tt / r^2 + cot(t) f_t / r^2) for f(theta),
    Ch f_pp / (r sin t)^2 for f(phi) and Cv (1/r) (r f)_rr for f(r), each a
    closed form of the stencil symbols on a cosine."""
    k = 1.1
    geo0 = pyoracle.sph_geometry(24, 22, 20, SPH["dx"], SPH["dy"], SPH["dz"], SPH["lat0"],
                                 SPH["lon0"], SPH["z0"], 1.0, 1.0)
    dth, dphi, dr = float(geo0["dth"]), float(geo0["dphi"]), float(geo0["dr"])
    fields = {"theta": lambda R, T, P: np.cos(k * T / dth),
              "phi": lambda R, T, P: np.cos(k * P / dphi),
              "r": lambda R, T, P: np.cos(k * R / dr) / R * 6371.0}
    f, out, geo, (R, T, P) = _one_step_sph(fields[axis])
    Ch, Cv = float(geo["Ch"]), float(geo["Cv"])
    S2, S1 = _symbols(k)
    if axis == "theta":
        want = Ch / R ** 2 * (np.cos(k * T / dth) * S2 / dth ** 2
                              - np.cos(T) / np.sin(T) * np.sin(k * T / dth) * S1 / dth)
    elif axis == "phi":
        want = Ch * np.cos(k * P / dphi) * S2 / dphi ** 2 / (R * np.sin(T)) ** 2
    else:
        want = Cv * np.cos(k * R / dr) * S2 * 6371.0 / R / dr ** 2
    got = _inner(out.astype(np.float64) - f.astype(np.float64))
    w = _inner(want)
    assert np.abs(w).max() > 1e-3                       # the step moves the field visibly
    assert np.abs(got - w).max() <= 2e-3 * np.abs(w).max(), np.abs(got - w).max()


def test_sph_smoother_cot_term_and_radial_weights():
    """Fields whose other terms vanish exactly, kept small so that one step
    shows above float resolution: f = theta - theta_c (second differences
    zero) gains exactly Ch cot(theta) / r^2; f = 6371/r - 1 ((r f) linear in
    r) does not change."""
    dth = float(pyoracle.sph_geometry(24, 22, 20, SPH["dx"], SPH["dy"], SPH["dz"], SPH["lat0"],
                                      SPH["lon0"], SPH["z0"], 1.0, 1.0)["dth"])

    def lin_theta(R, T, P):  # exactly linear in the node index (float colatitudes are not)
        i = np.rint((T - T.min()) / dth)
        return (i - 12.0) * dth
    f, out, geo, (R, T, P) = _one_step_sph(lin_theta)
    want = _inner(float(geo["Ch"]) * np.cos(T) / np.sin(T) / R ** 2)
    got = _inner(out.astype(np.float64) - f.astype(np.float64))
    assert want.min() > 5e-9
    assert np.abs(got - want).max() <= 0.1 * np.abs(want).max(), (np.abs(got - want).max(),
                                                                  want.max())
    f, out, geo, _ = _one_step_sph(lambda R, T, P: 6371.0 / R - 1.0)
    got = _inner(out.astype(np.float64) - f.astype(np.float64))
    assert np.abs(got).max() <= 3e-9, np.abs(got).max()


def test_sph_smoother_keeps_constants():
    f = np.full((10, 9, 8), -1.5, np.float32)
    out = pyoracle.smooth_sph_pde(f, SPH["dx"], SPH["dy"], SPH["dz"], SPH["lat0"], SPH["lon0"],
                                  SPH["z0"], 20.0, 5.0)
    assert np.abs(out + 1.5).max() <= 1e-5


def test_depth_interpolation_round_trips_linear_fields():
    dep = np.array([0, 1, 2.5, 4, 7, 12, 20], np.float32)
    nz, ny, nx = dep.size, 3, 4
    a = np.random.default_rng(1).normal(size=(ny, nx)).astype(np.float32)
    g = (2.0 + a[None] * dep[:, None, None]).astype(np.float32)
    nz1 = int((dep[-1] - dep[0]) / (1.0 * 0.99))
    gr = pyoracle.interp_irregular_z(g, nz1, dep, True)
    zr = dep[0] + np.arange(nz1, dtype=np.float32) * np.float32((dep[-1] - dep[0]) / (nz1 - 1))
    assert np.abs(gr - (2.0 + a[None] * zr[:, None, None])).max() <= 1e-4
    back = pyoracle.interp_irregular_z(np.zeros_like(g), nz1, dep, False, gradr=gr)
    assert np.abs(back - g).max() <= 1e-4


def test_regularization_rows():
    nx, ny, nz, w, m = 5, 4, 3, 0.7, 2
    n = nx * ny * nz
    interior = (nx - 2) * (ny - 2) * (nz - 2)
    ptr = np.zeros(m + n + 1, np.int32)
    ptr[1:m + 1] = [3, 5]
    cap = 5 + 7 * n
    col = np.zeros(cap, np.int32)
    val = np.zeros(cap, np.float32)
    ptr, col, val, nnz = pyoracle.add_regularization(ptr, col, val, m, w, nx, ny, nz)
    assert nnz == 5 + (n - interior) + 7 * interior and ptr[-1] == nnz
    ones = np.ones(n)
    for r in range(n):
        lo, hi = ptr[m + r], ptr[m + r + 1]
        s = float(np.dot(val[lo:hi].astype(np.float64), ones[col[lo:hi]]))
        k, rem = divmod(r, nx * ny)
        j, i = divmod(rem, nx)
        edge = i in (0, nx - 1) or j in (0, ny - 1) or k in (0, nz - 1)
        assert (hi - lo) == (1 if edge else 7)
        assert abs(s - (2 * w if edge else 0.0)) <= 1e-6
        if not edge:
            assert list(col[lo:hi]) == [r, r - 1, r + 1, r - nx, r + nx, r - nx * ny, r + nx * ny]


def box_field(nx=33, ny=36, nz=25):
    """The field of the reference's own smoothing test (test/smooth/main.cpp):
    -1 with a box of +1 in the middle third of every axis."""
    g = -np.ones((nz, ny, nx), np.float32)
    ix, iy, iz = np.arange(nx), np.arange(ny), np.arange(nz)
    inx = (ix > nx // 3) & (ix < nx // 3 * 2)
    iny = (iy > ny // 3) & (iy < ny // 3 * 2)
    inz = (iz > nz // 3) & (iz < nz // 3 * 2)
    g[np.ix_(inz, iny, inx)] = 1.0
    return g


# test/smooth/smooth.in: 33 x 36 x 25 nodes, 0.3 deg / 4 km, from 35.3 N 99.7 E,
# sigma_h 120 km, sigma_v 6 km
REF_CASE = (0.3, 0.3, 4.0, 35.3, 99.7, 0.0, 120.0, 6.0)


def test_pde_smoother_matches_the_direct_gaussian_convolution_on_the_reference_case():
    """The reference's test smooths its box field both ways (smooth_sph, the
    direct convolution, and smooth_sph_pde) and compares them by eye; here
    they agree to 1 % of the field's range everywhere (3149 explicit steps)."""
    g = box_field()
    pde = pyoracle.smooth_sph_pde(g, *REF_CASE)
    conv = pyoracle.smooth_sph_conv(g, *REF_CASE)
    assert np.abs(pde - conv).max() <= 0.02
    assert abs(pde[12, 18, 16] - conv[12, 18, 16]) <= 0.005
    assert 0.2 < conv[12, 18, 16] < 0.5          # the box is visibly spread


def test_float_lu_coefficients_drift_is_small():
    """The reference solves for its finite-difference coefficients with a
    float LU (smoothing.cpp:84-110, Eigen partialPivLu on A(i,j) =
    (float)pow(j-n, i) / i!); the smoothers here use the exact rationals
    rounded to float.  Eigen is absent, so LAPACK's float LU (scipy) on the
    same float matrix stands in: its coefficients are within 3e-6 of the
    rationals, and the smoothed field over a full run of the spherical
    smoother (1,399 steps at sigma_h 20 km, sigma_v 5 km on 0.1 deg x 1 km
    nodes) and of the Cartesian one moves by under 1e-5 of its largest value."""
    import ctypes as C
    from math import factorial
    import scipy.linalg as sl

    def lu_coefs(acc, deriv):
        n = acc // 2
        A = np.array([[np.float32(float(j - n) ** i) / np.float32(factorial(i))
                       for j in range(acc + 1)] for i in range(acc + 1)], np.float32)
        b = np.zeros(acc + 1, np.float32)
        b[deriv] = 1.0
        x = sl.lu_solve(sl.lu_factor(A), b).astype(np.float32)
        if deriv % 2:
            x[n] = 0.0
        return x

    keys = ((8, 2), (6, 2), (6, 1))
    lu = [lu_coefs(*k) for k in keys]
    for k, c in zip(keys, lu):
        assert np.abs(c - pyoracle.fdcoefs(*k)).max() < 3e-6
    L = pyoracle._smooth_lib()
    L.or_fdcoefs_override.argtypes = [C.c_void_p] * 3
    rng = np.random.default_rng(0)
    g = rng.standard_normal((24, 48, 48)).astype(np.float32)
    args = (0.1, 0.1, 1.0, 35.0, 100.0, 0.0, 20.0, 5.0)
    a = pyoracle.smooth_sph_pde(g, *args)
    a2 = pyoracle.smooth_cart_pde(g, 3.0, 2.0)
    try:
        L.or_fdcoefs_override(*[c.ctypes.data for c in lu])
        b = pyoracle.smooth_sph_pde(g, *args)
        b2 = pyoracle.smooth_cart_pde(g, 3.0, 2.0)
    finally:
        L.or_fdcoefs_override(None, None, None)
    d_sph = np.abs(a - b).max() / np.abs(a).max()
    d_cart = np.abs(a2 - b2).max() / np.abs(a2).max()
    print("float-LU drift: spherical", d_sph, "cartesian", d_cart)
    assert d_sph < 1e-5 and d_cart < 1e-5
