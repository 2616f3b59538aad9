"""FMST traveltimes and ray weights on the GPU (libjdsg_swd.so, jdsg_fmst_*)
against the reference compiled from source (the fixtures of
tests/golden/make_golden_fmst.py) and against the oracle restatement
(oracle/fmst_oracle.c, bit-identical to the reference on its own tests).

Bars: statuses equal; fast-marching traveltimes bit-identical (the march
runs the reference's single-precision operations in its order, with
correctly rounded division and square root, and the grid sines come from
the host's sinf as the reference's do); ray times and weights bit-identical
for at least 99 % of the receivers and otherwise within 1e-4 relative (time)
and 1e-3 of the receiver's largest weight: a ray takes the sine of its own
points on the device, correctly rounded, where the host's sinf may be off by
an ulp in rare cases, which moves a ray point by that much.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fmst_cases.npz")
CASES = ("deriv", "example", "edge")


def _case(name):
    z = np.load(GOLD)
    return {k.split("__", 1)[1]: z[k] for k in z.files if k.startswith(name + "__")}


def _gpu(c, mode=1, ray_time=None):
    import torch
    from jdsurfg_amd import fmst
    nx, ny = int(c["nx"]), int(c["ny"])
    lat0, lon0, dlat, dlon = (float(v) for v in c["grid"])
    F = fmst.Fmst(nx, ny, lat0, lon0, dlat, dlon, c["idx"], c["srcx"], c["srcz"], c["roff"],
                  c["rcx"], c["rcz"], ray_time)
    kmax = c["vc"].shape[0]
    vc = torch.from_numpy(np.ascontiguousarray(c["vc"], np.float64).ravel()).cuda()
    vout = torch.from_numpy(np.ascontiguousarray(c["vout"], np.float64).ravel()).cuda()
    o = F.run(vc, vout, kmax, mode=mode)
    r = {k: v.cpu().numpy() for k, v in o.items()}
    if mode == 1:
        off, cell, w = F.rays()
        r["ray_off"], r["ray_cell"], r["ray_w"] = off.cpu().numpy(), cell.cpu().numpy(), w.cpu().numpy()
    torch.cuda.synchronize()
    return r


def _dense(off, cell, w, nxy):
    out = np.zeros((len(off) - 1, nxy), np.float32)
    for q in range(len(off) - 1):
        out[q, cell[off[q]:off[q + 1]]] = w[off[q]:off[q + 1]]
    return out


def _compare(name, g, ref, nxy):
    assert (g["rstatus"] == 0).all() and (g["pstatus"] == 0).all(), (g["rstatus"], g["pstatus"])
    rel = lambda a, b: np.abs(a.astype(np.float64) - b) / np.abs(b)
    same_t = float((g["ttime"] == ref["ttime"]).mean())
    same_s = float((g["tsyn"] == ref["tsyn"]).mean())
    fa = _dense(g["ray_off"], g["ray_cell"], g["ray_w"], nxy)
    fb = _dense(ref["ray_off"], ref["ray_cell"], ref["ray_w"], nxy)
    scale = np.maximum(np.abs(fb).max(axis=1), 1e-30)
    werr = (np.abs(fa - fb).max(axis=1) / scale)
    same_w = float((werr == 0).mean())
    print(f"{name}: ttime identical {same_t:.3f} max rel {rel(g['ttime'], ref['ttime']).max():.2e}; "
          f"tsyn identical {same_s:.3f} max rel {rel(g['tsyn'], ref['tsyn']).max():.2e}; "
          f"weights identical rows {same_w:.3f} max rel {werr.max():.2e}", flush=True)
    assert np.array_equal(g["ttime"], ref["ttime"])
    assert same_s >= 0.99 and same_w >= 0.99
    assert rel(g["tsyn"], ref["tsyn"]).max() <= 1e-4
    assert werr.max() <= 1e-3


@pytest.mark.parametrize("name", CASES)
def test_fixture_cases_match_reference(name):
    c = _case(name)
    print(f"case fmst {name}: pairs {len(c['idx'])} receivers {len(c['ttime'])}", flush=True)
    g = _gpu(c, mode=1)
    _compare(name, g, c, int(c["nx"]) * int(c["ny"]))


def test_travel_time_mode_takes_ray_times_for_group_pairs():
    """mode 0 (SurfTime::travel_time, SurfaceWave.cpp:524-536): srtimes for
    phase pairs, the ray time for group pairs, and no weights kept."""
    c = _case("example")
    rt = (np.arange(len(c["idx"])) % 2).astype(np.int32)
    print("case fmst travel_time mode on example", flush=True)
    g = _gpu(c, mode=0, ray_time=rt)
    ref = c["ttime"].copy()
    for p in range(len(c["idx"])):
        if rt[p]:
            ref[c["roff"][p]:c["roff"][p + 1]] = c["tsyn"][c["roff"][p]:c["roff"][p + 1]]
    rel = np.abs(g["ttime"].astype(np.float64) - ref) / ref
    print("travel_time mode max rel", rel.max(), flush=True)
    assert rel.max() <= 1e-4
    assert (g["rstatus"] == 0).all()


def test_random_pairs_match_oracle():
    """Random smooth maps and stations on three grids, the oracle
    restatement as the yardstick (OpenMP over pairs)."""
    from oracle import pyoracle as P
    rng = np.random.default_rng(7)
    for (nx, ny, lat0, lon0, dlat, dlon, npair, nr) in [(33, 36, 35.3, 99.7, 0.3, 0.3, 40, 6),
                                                       (24, 30, 10.0, 20.0, 0.5, 0.25, 24, 5),
                                                       (19, 19, -5.0, 140.0, 0.4, 0.4, 16, 4)]:
        ix = np.arange(nx)[:, None] / nx
        iy = np.arange(ny)[None, :] / ny
        maps = []
        for k in range(3):
            f = np.sin(2 * np.pi * (1 + k) * ix + k) * np.cos(2 * np.pi * (2 - 0.5 * k) * iy)
            maps.append((3.0 + 0.3 * k + 0.2 * f + 0.03 * rng.standard_normal((nx, ny))).T.ravel())
        vc = np.stack(maps)
        vout = vc * 0.9
        lat = lambda n: rng.uniform(lat0 - dlat * (nx - 2) + 0.01, lat0 - dlat - 0.01, n)
        lon = lambda n: rng.uniform(lon0 + dlon + 0.01, lon0 + dlon * (ny - 2) - 0.01, n)
        d2r = np.float32(np.pi / 180)
        srcx = ((90 - lat(npair)) * d2r).astype(np.float32)
        srcz = (lon(npair) * d2r).astype(np.float32)
        rcx = ((90 - lat(npair * nr)) * d2r).astype(np.float32)
        rcz = (lon(npair * nr) * d2r).astype(np.float32)
        roff = np.arange(npair + 1, dtype=np.int32) * nr
        idx = rng.integers(0, 3, npair).astype(np.int32)
        tt, ts, fdm, st = P.fmst_pairs((nx, ny, lat0, lon0, dlat, dlon), vc, vout, idx, srcx, srcz,
                                       roff, rcx, rcz, nthreads=16)
        assert (st == 0).all()
        off, cell, w = (np.zeros(npair * nr + 1, np.int64), [], [])
        for q in range(npair * nr):
            nzq = np.nonzero(fdm[q])[0]
            off[q + 1] = off[q] + len(nzq)
            cell.append(nzq)
            w.append(fdm[q][nzq])
        ref = {"ttime": tt, "tsyn": ts, "ray_off": off, "ray_cell": np.concatenate(cell),
               "ray_w": np.concatenate(w)}
        c = {"nx": nx, "ny": ny, "grid": np.array([lat0, lon0, dlat, dlon], np.float32), "vc": vc,
             "vout": vout, "idx": idx, "srcx": srcx, "srcz": srcz, "roff": roff, "rcx": rcx,
             "rcz": rcz}
        print(f"case fmst random grid {nx}x{ny}: pairs {npair} receivers {npair * nr}", flush=True)
        g = _gpu(c, mode=1)
        _compare(f"random {nx}x{ny}", g, ref, nx * ny)


def test_outside_source_and_receiver_report_status():
    """A source or receiver outside the propagation grid: the reference
    prints and stops (fmst.f90:165-171, fmm.f90:1010-1016, fmst.f90:488-495);
    here the pair or receiver gets status 1 and the rest are solved."""
    c = _case("example")
    srcx = c["srcx"].copy()
    rcx = c["rcx"].copy()
    srcx[1] = np.float32((90 - 50.0) * np.pi / 180)      # north of the grid
    rcx[0] = np.float32((90 - 45.0) * np.pi / 180)
    c2 = dict(c, srcx=srcx, rcx=rcx)
    print("case fmst outside source/receiver", flush=True)
    g = _gpu(c2, mode=1)
    assert g["pstatus"][1] == 1
    assert (g["rstatus"][c["roff"][1]:c["roff"][2]] == 1).all()
    assert g["rstatus"][0] == 1
    ok = np.ones(len(rcx), bool)
    ok[c["roff"][1]:c["roff"][2]] = False
    ok[0] = False
    assert (g["rstatus"][ok] == 0).all()
    assert np.array_equal(g["ttime"][ok], c["ttime"][ok])


def test_global_memory_heap_matches(monkeypatch):
    """Pairs whose fast-marching band outgrows the heap kept in LDS are
    marched again with the heap in global memory; forcing a 64-entry LDS
    heap sends every pair there, and the results stay the reference's."""
    monkeypatch.setenv("JDSG_FMST_LDS_HEAP", "64")
    for name in CASES:
        c = _case(name)
        print(f"case fmst {name} with the global-memory heap", flush=True)
        g = _gpu(c, mode=1)
        _compare(name + " (global heap)", g, c, int(c["nx"]) * int(c["ny"]))
