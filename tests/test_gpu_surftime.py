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
