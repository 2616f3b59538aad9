"""The CPU oracle pinned against the reference's outputs and the mathematics.

1. Golden vectors: outputs of the reference numerics compiled from their own
   sources (tests/golden/make_golden.py) -- phase velocities identical,
   U and kernels to rounding.
2. Closed forms: a uniform half-space (Rayleigh root of the Rayleigh
   function, no dispersion), a layer over a half-space (Love dispersion
   relation), both on a flat Earth.
3. Invariants: U = c / (1 + (T/c) dc/dT) against finite differences of the
   root chain; the flat-Earth scaling identities
   sum(a dc/da + b dc/db + h dc/dh) = c and sum(rho dc/drho) = 0; the kernels
   against +-0.5 % finite differences of the root chain (test/swd/main.cpp's
   check, given a tolerance here).
"""
import os

import numpy as np
import pytest

from jdsurfg_amd import workloads as W

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KERN = ("dcda", "dcdb", "dcdr", "dcdh", "duda", "dudb", "dudr", "dudh")


def _load(name):
    return dict(np.load(os.path.join(GOLD, name)))


@pytest.mark.parametrize("name", ["batch_grid20.npz", "batch_higher_modes.npz",
                                  "batch_sediment_lvz.npz", "batch_water_top.npz",
                                  "batch_thin16_thick128.npz", "batch_grid128_sub.npz"])
def test_batch_matches_reference_fixture(name, oracle_rest):
    g = _load(name)
    r = oracle_rest.batch_solve(g["layer_off"], g["thk"], g["vp"], g["vs"], g["rho"],
                                g["per_off"], g["periods"], g["nmode"].ravel(),
                                bool(g["group"]), nthreads=4)
    assert np.array_equal(r["status"], g["status"])
    ok = g["status"] == 0
    assert ok.sum() > 0
    assert np.array_equal(r["c"][ok], g["c"][ok]), "phase velocity must be identical"
    rel_u = np.abs(r["u"][ok] - g["u"][ok]) / np.abs(g["u"][ok])
    assert rel_u.max() <= 1e-9
    N = np.diff(g["layer_off"])
    q = 0
    for c in g["kern_cols"]:
        n = int(g["kern_off"][c + 1] - g["kern_off"][c])
        e0, e1 = g["ent_off"][c], g["ent_off"][c + 1]
        for nm in KERN:
            a = r[nm][g["kern_off"][c]:g["kern_off"][c + 1]].reshape(-1, N[c])
            b = g[nm][q:q + n].reshape(-1, N[c])
            okc = ok[e0:e1]
            if not okc.any():
                continue
            scale = np.abs(b[okc]).max(axis=1)
            scale[scale == 0] = 1.0
            err = (np.abs(a[okc] - b[okc]).max(axis=1) / scale).max()
            assert err <= 1e-8, (nm, c, err)
        q += n


def test_surftime_glue_matches_reference(oracle_rest):
    g = _load("surftime_model_txt.npz")
    t = g["periods"]
    tabs = {w: t.reshape(-1, 1) for w in ("Rc", "Rg", "Lc", "Lg")}
    vc, vout, ker, nerr = oracle_rest.get_1d_kernel(g["vs"].reshape(1, 1, -1), g["z"], tabs)
    assert nerr == 0
    assert np.array_equal(vc, g["vc"])
    np.testing.assert_allclose(vout, g["vout"], rtol=1e-12)
    np.testing.assert_allclose(ker, g["kernel"], rtol=0, atol=1e-8 * np.abs(g["kernel"]).max())
    mvc, mvout, nerr2 = oracle_rest.get_2d_map(g["vs"].reshape(1, 1, -1), g["z"], tabs)
    np.testing.assert_allclose(mvc, g["map_vc"], rtol=1e-15)
    np.testing.assert_allclose(mvout, g["map_vout"], rtol=1e-12)
    tabs3 = {"Rc": g["tR"], "Rg": g["tR"], "Lc": g["tL"], "Lg": g["tL"]}
    vc3, vout3, ker3, _ = oracle_rest.get_1d_kernel(g["vs3"], g["z"], tabs3)
    np.testing.assert_allclose(vc3, g["g_vc"], rtol=1e-15)
    np.testing.assert_allclose(vout3, g["g_vout"], rtol=1e-12)
    np.testing.assert_allclose(ker3, g["g_kernel"], rtol=0,
                               atol=1e-8 * np.abs(g["g_kernel"]).max())


def test_surfdisp96_calls_match_reference(oracle_rest):
    g = _load("surfdisp96_calls.npz")
    for key, cg in zip(g["keys"], g["cg"]):
        iflsph, iwave, mode, igr, ierr = (int(v) for v in key)
        mine, myerr, _ = oracle_rest.surfdisp96(g["thk"], g["vp"], g["vs"], g["rho"],
                                                g["periods"], iflsph, iwave, mode, igr)
        assert myerr == ierr
        assert np.array_equal(mine, cg), key
    one, e1, _ = oracle_rest.surfdisp96(g["thk"], g["vp"], g["vs"], g["rho"],
                                        g["periods"][5:6], 1, 2, 3, 0)
    assert np.array_equal(one, g["single"]) and e1 == int(g["single_ierr"])


# ------------------------------------------------------------ closed forms
def _rayleigh_root(a, b):
    """Rayleigh speed of a Poisson-type half-space by bisection on
    (2-k^2)^2 - 4 sqrt(1-k^2 g^2) sqrt(1-k^2), k = c/b, g = b/a."""
    g = b / a
    f = lambda k: (2 - k * k) ** 2 - 4 * np.sqrt(1 - k * k * g * g) * np.sqrt(1 - k * k)
    lo, hi = 0.5, 0.9999999
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(lo) * f(mid) <= 0:
            hi = mid
        else:
            lo = mid
    return 0.5 * (lo + hi) * b


def test_uniform_halfspace_rayleigh_is_nondispersive(oracle_rest):
    a, b, rho = 6.0, 3.464, 2.7
    n = 12
    thk = np.full(n, 2.0, np.float32)
    thk[-1] = 0
    vp = np.full(n, a, np.float32)
    vs = np.full(n, b, np.float32)
    rh = np.full(n, rho, np.float32)
    t = np.geomspace(2, 40, 8)
    cg, ierr, _ = oracle_rest.surfdisp96(thk, vp, vs, rh, t, 0, 2, 1, 0)
    cr = _rayleigh_root(float(np.float32(a)), float(np.float32(b)))
    assert ierr == 0
    assert np.abs(cg / cr - 1).max() < 2e-6
    # a planted mistake (dropped gamma in the Rayleigh function) is caught
    wrong = _rayleigh_root(float(np.float32(b)) * 1.0001, float(np.float32(b)))
    assert np.abs(cg / wrong - 1).max() > 1e-3


def _love_root(h, b1, r1, b2, r2, T, mode=0):
    """Love layer over half-space: tan(w h s1) = mu2 s2 / (mu1 s1),
    s1 = sqrt(1/b1^2 - 1/c^2), s2 = sqrt(1/c^2 - 1/b2^2)."""
    w = 2 * np.pi / T
    mu1, mu2 = r1 * b1 * b1, r2 * b2 * b2

    def f(c):
        s1 = np.sqrt(1 / b1 ** 2 - 1 / c ** 2)
        s2 = np.sqrt(1 / c ** 2 - 1 / b2 ** 2)
        return np.sin(w * h * s1) * mu1 * s1 - mu2 * s2 * np.cos(w * h * s1)

    cs = np.linspace(b1 * 1.0000001, b2 * 0.9999999, 20001)
    fv = f(cs)
    idx = np.nonzero(np.sign(fv[:-1]) != np.sign(fv[1:]))[0]
    roots = []
    for i in idx:
        lo, hi = cs[i], cs[i + 1]
        for _ in range(100):
            mid = 0.5 * (lo + hi)
            if np.sign(f(lo)) == np.sign(f(mid)):
                lo = mid
            else:
                hi = mid
        roots.append(0.5 * (lo + hi))
    # mode n has w h s1 in [n pi, (n+1/2) pi): the fundamental is the slowest
    return sorted(roots)[mode]


def test_love_layer_over_halfspace(oracle_rest):
    h, b1, r1, b2, r2 = 30.0, 3.2, 2.6, 4.5, 3.3
    thk = np.array([h, 0], np.float32)
    vs = np.array([b1, b2], np.float32)
    vp = (vs * 1.73).astype(np.float32)
    rho = np.array([r1, r2], np.float32)
    t = np.array([10.0, 20.0, 40.0])
    for mode in (1, 2):
        cg, ierr, _ = oracle_rest.surfdisp96(thk, vp, vs, rho, t[:1], 0, 1, mode, 0)
        ref = _love_root(h, float(vs[0]), float(rho[0]), float(vs[1]), float(rho[1]), t[0],
                         mode - 1)
        assert ierr == 0
        assert abs(cg[0] / ref - 1) < 2e-6, (mode, cg, ref)
    cg, ierr, _ = oracle_rest.surfdisp96(thk, vp, vs, rho, t, 0, 1, 1, 0)
    ref = [_love_root(h, float(vs[0]), float(rho[0]), float(vs[1]), float(rho[1]), T) for T in t]
    assert ierr == 0
    assert np.abs(cg / np.array(ref) - 1).max() < 2e-6


# ------------------------------------------------------------ invariants
def _col(seed=4, nlay=20):
    w = W.random_columns(1, nlay, seed, nmode=(1, 1))
    return w.thk.copy(), w.vp.copy(), w.vs.copy(), w.rho.copy()


@pytest.mark.parametrize("iwave", [2, 1])
def test_group_velocity_against_fd_of_phase(oracle_rest, iwave):
    thk, vp, vs, rho = _col()
    T = 12.0
    dT = 0.05
    ts = np.array([T - dT, T, T + dT])
    c, _, _ = oracle_rest.surfdisp96(thk, vp, vs, rho, ts, 0, iwave, 1, 0)
    dcdT = (c[2] - c[0]) / (2 * dT)
    u_fd = c[1] / (1 + (T / c[1]) * dcdT)
    n = len(thk)
    o = [np.zeros(n) for _ in range(8)]
    cc = np.array([c[1]])
    uu = np.zeros(1)
    tt = np.array([T])
    L = oracle_rest.lib
    if iwave == 2:
        L.or_sregn96(thk, vp, vs, rho, n, tt, cc, uu, *o, 0)
    else:
        L.or_slegn96(thk, vs, rho, n, tt, cc, uu, *o[:5], 0)
    assert abs(uu[0] / u_fd - 1) < 2e-3


def test_flat_scaling_identities(oracle_rest):
    thk, vp, vs, rho = _col(seed=6)
    n = len(thk)
    T = 15.0
    c, _, _ = oracle_rest.surfdisp96(thk, vp, vs, rho, np.array([T]), 0, 2, 1, 0)
    ur, uz, tr, tz, da, db, dh, dr = (np.zeros(n) for _ in range(8))
    cc = c.copy()
    uu = np.zeros(1)
    oracle_rest.lib.or_sregn96(thk, vp, vs, rho, n, np.array([T]), cc, uu, ur, uz, tr, tz,
                               da, db, dh, dr, 0)
    a, b, h, r = (x.astype(np.float64) for x in (vp, vs, thk, rho))
    s_vel = (a * da + b * db + h * dh).sum()
    assert abs(s_vel / c[0] - 1) < 1e-5
    s_rho = (r * dr).sum()
    assert abs(s_rho) < 1e-5 * np.abs(r * dr).sum()
    # sign flip of one term (a plausible transcription slip) breaks it
    assert abs((a * da - b * db + h * dh).sum() / c[0] - 1) > 1e-2


def test_kernel_against_finite_difference(oracle_rest):
    thk, vp, vs, rho = _col(seed=8, nlay=12)
    n = len(thk)
    T = 10.0
    c0, _, _ = oracle_rest.surfdisp96(thk, vp, vs, rho, np.array([T]), 0, 2, 1, 0)
    ur, uz, tr, tz, da, db, dh, dr = (np.zeros(n) for _ in range(8))
    cc = c0.copy()
    oracle_rest.lib.or_sregn96(thk, vp, vs, rho, n, np.array([T]), cc, np.zeros(1), ur, uz,
                               tr, tz, da, db, dh, dr, 0)
    fd = np.zeros(n)
    for i in range(n):
        v0 = vs[i]
        vs[i] = np.float32(v0 * (1 - 0.005))
        cm, _, _ = oracle_rest.surfdisp96(thk, vp, vs, rho, np.array([T]), 0, 2, 1, 0)
        vs[i] = np.float32(v0 * (1 + 0.005))
        cp_, _, _ = oracle_rest.surfdisp96(thk, vp, vs, rho, np.array([T]), 0, 2, 1, 0)
        fd[i] = (cp_[0] - cm[0]) / (float(vs[i]) - float(np.float32(v0 * (1 - 0.005))))
        vs[i] = v0
    assert np.abs(fd - db).max() < 0.02 * np.abs(db).max()


def test_brocher_relation(oracle_rest):
    vs = np.linspace(1.0, 4.6, 50).astype(np.float32)
    vp, rho = W.brocher(vs)
    import ctypes as C
    a, b = C.c_float(), C.c_float()
    for i, v in enumerate(vs):
        oracle_rest.lib.or_empirical_relation(float(v), C.byref(a), C.byref(b))
        assert a.value == vp[i] and b.value == rho[i]
    # crustal anchor of the Brocher (2005) fits: Vs 3.5 km/s -> Vp near 6 km/s
    # and rho near 2.7 g/cc
    p, r = W.brocher(np.array([3.5]))
    assert abs(p[0] - 6.0) < 0.1 and abs(r[0] - 2.7) < 0.1


def test_param_convert_constant_kernel(oracle_rest):
    """A unit kernel on every refined layer maps to 2 at the top node, 4 at
    interior nodes and 2 + 7/8 + 4 at the bottom node (the half-space value is
    added once per sublayer, SurfaceWave.cpp:82)."""
    nz = 6
    rmax = nz + 3 * (nz - 1)
    vs = np.ones(nz, np.float32)
    out = np.zeros(nz)
    oracle_rest.lib.or_param_convert(vs, nz, 3, rmax, np.ones(rmax), out)
    expect = np.full(nz, 4.0)
    expect[0] = 2.0
    expect[-1] = 2.0 + 7.0 / 8.0 + 4.0
    np.testing.assert_allclose(out, expect, rtol=1e-15)


@pytest.mark.parametrize("group", [True, False])
def test_eigen_at_given_roots_reproduces_the_solve(oracle_rest, group):
    """The eigen stage at given roots (used to check GPU kernels where a root
    differs in its last bits) returns the solve's own outputs when given the
    solve's own roots."""
    from tests.compare import KERNELS, at_gpu_roots
    w = W.random_columns(5, 20, 4, nmode=(2, 2), group=group)
    b = oracle_rest.batch_solve(w.layer_off, w.thk, w.vp, w.vs, w.rho, w.per_off, w.periods,
                                w.nmode.ravel(), w.group, nthreads=2)
    a = dict(b)
    if not group:
        a["cflat"] = b["cflat"][:len(b["c"])]
    g = at_gpu_roots(oracle_rest, w, a, nthreads=2)
    ok = b["status"] != 4
    assert np.array_equal(g["status"][ok], b["status"][ok])
    for n in ("c", "u") + KERNELS:
        assert np.array_equal(g[n], b[n]), n
