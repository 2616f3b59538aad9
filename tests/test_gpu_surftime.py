"""GPU parity of the SurfTime grid entry point (get_1d_kernel) against the
oracle glue (SurfaceWave.cpp:458-502) and the reference fixture."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _check(a, b, ok):
    vc, vout, ker, st = a
    rvc, rvout, rker = b
    assert np.abs(vc[ok] / rvc[ok] - 1).max() <= 1e-6
    assert np.abs(vout[ok] / rvout[ok] - 1).max() <= 1e-6
    nxy, nz, kmax = ker.shape
    for n in range(nxy):
        for i in range(kmax):
            if not ok[n, i]:
                continue
            ref = rker[n, :, i]
            m = np.abs(ref).max()
            assert np.abs(ker[n, :, i] - ref).max() <= 1e-4 * m, (n, i)


def test_model_txt_all_wave_types():
    from jdsurfg_amd import swd
    g = dict(np.load(os.path.join(GOLD, "surftime_model_txt.npz")))
    t = g["periods"]
    tabs = {w: t.reshape(-1, 1) for w in ("Rc", "Rg", "Lc", "Lg")}
    print("case model.txt Rc/Rg/Lc/Lg", flush=True)
    a = swd.get_1d_kernel(g["vs"].reshape(1, 1, -1), g["z"], tabs)
    ok = a[3] == 0
    assert ok.all()
    _check(a, (g["vc"], g["vout"], g["kernel"]), ok)


def test_grid_with_two_mode_tables(oracle_rest):
    from jdsurfg_amd import swd
    g = dict(np.load(os.path.join(GOLD, "surftime_model_txt.npz")))
    tabs = {"Rc": g["tR"], "Rg": g["tR"], "Lc": g["tL"], "Lg": g["tL"]}
    print("case 3x2 grid, two-mode tables", flush=True)
    a = swd.get_1d_kernel(g["vs3"], g["z"], tabs)
    b = oracle_rest.get_1d_kernel(g["vs3"], g["z"], tabs)
    ok = a[3] == 0
    assert ok.sum() > 0.9 * ok.size
    _check(a, b[:3], ok)
    _check(a, (g["g_vc"], g["g_vout"], g["g_kernel"]), ok)


def test_map_matches_reference_fixture():
    """get_2d_map (disp1D: one chain per wave and mode, _flat2sphere, the
    analytic group velocity of groupvel_r/_l) against the reference's map of
    test/swd/model.txt and of the 3x2 grid with two-mode tables."""
    from jdsurfg_amd import swd
    g = dict(np.load(os.path.join(GOLD, "surftime_model_txt.npz")))
    t = g["periods"]
    tabs = {w: t.reshape(-1, 1) for w in ("Rc", "Rg", "Lc", "Lg")}
    print("case get_2d_map model.txt", flush=True)
    vc, vout, st = swd.get_2d_map(g["vs"].reshape(1, 1, -1), g["z"], tabs)
    assert (st == 0).all()
    assert np.abs(vc / g["map_vc"] - 1).max() <= 1e-6
    assert np.abs(vout / g["map_vout"] - 1).max() <= 1e-6
    print("map model.txt: vc identical", float((vc == g["map_vc"]).mean()), flush=True)
    tabs3 = {"Rc": g["tR"], "Rg": g["tR"], "Lc": g["tL"], "Lg": g["tL"]}
    print("case get_2d_map 3x2 grid", flush=True)
    vc, vout, st = swd.get_2d_map(g["vs3"], g["z"], tabs3)
    ok = st == 0
    assert ok.sum() > 0.9 * ok.size
    assert np.array_equal(g["g_map_vc"][~ok], np.zeros((~ok).sum()))
    assert np.abs(vc[ok] / g["g_map_vc"][ok] - 1).max() <= 1e-6
    assert np.abs(vout[ok] / g["g_map_vout"][ok] - 1).max() <= 1e-6


def _coastal_grid():
    """A 4x3 grid on 21 nodes (0-40 km): land columns with gradients and a
    low-velocity zone, and water columns (vs = 0 at the top node) over slow
    sediments, mixed in one grid."""
    rng = np.random.default_rng(17)
    dep = (2.0 * np.arange(21)).astype(np.float32)
    base = 2.6 + 0.04 * dep
    vs = np.zeros((4, 3, 21), np.float32)
    for i in range(4):
        for j in range(3):
            v = base + rng.normal(0, 0.05, 21)
            if (i + j) % 3 == 0:
                v[0] = 0.0
                v[1:3] = rng.uniform(1.0, 1.8, 2)
            vs[i, j] = v
    return vs, dep


def test_mixed_water_and_land_grid(oracle_rest):
    """A grid mixing water-top and land columns: get_1d_kernel and
    get_2d_map against the oracle's glue, Rayleigh modes 0-1 (phase and
    group).  Love waves are compared on the land columns: below water the
    reference's slegn96 reads an unset value (uu(1) of a fluid top layer)."""
    from jdsurfg_amd import swd
    vs, dep = _coastal_grid()
    t = np.array([4.0, 6.0, 9.0, 14.0, 22.0, 35.0])
    two = np.full((6, 2), -1.0)
    two[:, 0] = t
    two[:4, 1] = t[:4]
    tabs = {"Rc": two, "Rg": two, "Lc": t.reshape(-1, 1), "Lg": t.reshape(-1, 1)}
    nxy = 12
    water = (vs[:, :, 0] == 0).ravel(order="F")
    print("case coastal grid get_1d_kernel", flush=True)
    a = swd.get_1d_kernel(vs, dep, tabs)
    b = oracle_rest.get_1d_kernel(vs, dep, tabs)
    nR = 6 + 4
    love = np.zeros(a[0].shape, bool)
    love[:, 2 * nR:] = True
    ok = (a[3] == 0) & ~(love & water[:, None])
    assert ok.sum() > 0.8 * ok.size
    assert water.sum() >= 3 and (~water).sum() >= 3
    # the water column's top node carries no kernel (SurfaceWave.cpp:495)
    kb = b[2].copy()
    kb[water, 0, :] = 0.0
    _check(a, (b[0], b[1], kb), ok)
    print("case coastal grid get_2d_map", flush=True)
    m = swd.get_2d_map(vs, dep, tabs)
    mo = oracle_rest.get_2d_map(vs, dep, tabs)
    okm = (m[2] == 0) & ~(love & water[:, None])
    assert np.abs(m[0][okm] / mo[0][okm] - 1).max() <= 1e-6
    assert np.abs(m[1][okm] / mo[1][okm] - 1).max() <= 1e-6
    # the map's phase velocities are the kernel form's (same chains)
    both = okm & ok
    assert np.abs(m[0][both] / a[0][both] - 1).max() <= 1e-12
    print("coastal: Love in water columns (kernel form) GPU", a[1][water][:, 2 * nR:][:2],
          "oracle", b[1][water][:, 2 * nR:][:2], flush=True)


def test_ragged_period_table_through_surftime(tmp_path):
    """A ragged two-mode header read by the data reader gives SurfTime the
    period index count of the header (3 + 1), with get_nt_row stopping at
    the -1 padding."""
    from jdsurfg_amd import data as D, swd
    p = tmp_path / "ragged.dat"
    p.write_text("2 3 1\n10 20 30\n40\n0\n0\n0\n# 31 108 0 1 2 0\n33.7 105.7 3.1\n")
    t, k = D.read_swd_data(p)["tables"]["Rc"]
    print("case ragged table through SurfTime", flush=True)
    dep = (2.0 * np.arange(11)).astype(np.float32)
    S = swd.SurfTime(2, 1, 11, dep, {"Rc": t})
    assert S.kmax == sum(k) == 4
    vs = np.tile((2.8 + 0.04 * dep)[None, None, :], (2, 1, 1)).astype(np.float32)
    vc, vout, st = swd.get_2d_map(vs, dep, {"Rc": t})
    assert vc.shape == (2, 4) and (vc[:, :3] > 0).all()
