"""The FMST oracle (oracle/fmst_oracle.c, the restatement of src/SWD/fmst.f90
and fmm.f90) pinned on the CPU:

* against the reference compiled from source: the committed fixtures
  (tests/golden/make_golden_fmst.py) and, where oracle/_ref/libfmstref.so is
  built, random maps and stations on three grids -- bit for bit;
* by what the mathematics fixes, independently of the reference:
  - doubling both maps halves every traveltime and quarters every weight
    exactly (every term scales with the slowness; powers of two are exact);
  - in a uniform medium the first-arrival time and the ray time equal the
    great-circle distance over the speed to the method's accuracy;
  - reciprocity: swapping source and receiver keeps the traveltime;
  - the weights are the derivatives of the ray time with respect to the
    output map's B-spline control values (the ray follows the phase map, so
    it stays put when the output map moves): central differences agree;
  - in a uniform output map sum(fdm) * v = -(ray length traced) / v, i.e.
    the ray time less its final segment into the source, which the weights
    leave out as the reference does.
A planted error of sign, orientation (x/y swapped) or slowness power fails
one of these.
"""
import os

import numpy as np
import pytest

from oracle import pyoracle as P

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fmst_cases.npz")
D2R = np.float32(np.pi / 180)
GRID = (33, 36, 35.3, 99.7, 0.3, 0.3)


def _case(name):
    z = np.load(GOLD)
    return {k.split("__", 1)[1]: z[k] for k in z.files if k.startswith(name + "__")}


def _dense(off, cell, w, nxy):
    out = np.zeros((len(off) - 1, nxy), np.float32)
    for q in range(len(off) - 1):
        out[q, cell[off[q]:off[q + 1]]] = w[off[q]:off[q + 1]]
    return out


@pytest.mark.parametrize("name", ["deriv", "example", "edge"])
def test_restatement_matches_reference_fixtures(name):
    c = _case(name)
    nx, ny = int(c["nx"]), int(c["ny"])
    grid = (nx, ny) + tuple(float(v) for v in c["grid"])
    tt, ts, fdm, st = P.fmst_pairs(grid, c["vc"], c["vout"], c["idx"], c["srcx"], c["srcz"],
                                   c["roff"], c["rcx"], c["rcz"], nthreads=4)
    assert (st == 0).all()
    assert np.array_equal(tt, c["ttime"]) and np.array_equal(ts, c["tsyn"])
    assert np.array_equal(fdm, _dense(c["ray_off"], c["ray_cell"], c["ray_w"], nx * ny))


@pytest.mark.skipif(not P.FmstRef.available(), reason="reference FMST not built here")
def test_restatement_matches_reference_build():
    R = P.FmstRef()
    rng = np.random.default_rng(11)
    for (nx, ny, lat0, lon0, dlat, dlon) in [(33, 36, 35.3, 99.7, 0.3, 0.3),
                                             (19, 21, 40.0, 100.0, 0.5, 0.5),
                                             (20, 45, 30.0, 90.0, 0.2, 0.4)]:
        grid = (nx, ny, lat0, lon0, dlat, dlon)
        ix = np.arange(nx)[:, None]
        iy = np.arange(ny)[None, :]
        vc = (3.0 + 0.3 * np.sin(ix / 3.0) * np.cos(iy / 4.0) +
              0.1 * rng.standard_normal((nx, ny))).T.ravel()[None].copy()
        vout = vc * 0.85 + 0.05 * rng.standard_normal(vc.shape)
        n = 6 * 5
        lats = lat0 - dlat * rng.uniform(1.2, nx - 2.2, n)
        lons = lon0 + dlon * rng.uniform(1.2, ny - 2.2, n)
        cx = ((90 - lats) * D2R).astype(np.float32)
        cz = (lons * D2R).astype(np.float32)
        srcx, srcz, rcx, rcz = cx[:6], cz[:6], cx[6:], cz[6:]
        roff = np.arange(7, dtype=np.int32) * 4
        tt, ts, fdm, st = P.fmst_pairs(grid, vc, vout, np.zeros(6, np.int32), srcx, srcz, roff,
                                       rcx, rcz, nthreads=4)
        for p in range(6):
            s = slice(roff[p], roff[p + 1])
            a, b, f = R.pair(grid, vc[0], vout[0], srcx[p], srcz[p], rcx[s], rcz[s])
            assert np.array_equal(tt[s], a) and np.array_equal(ts[s], b)
            assert np.array_equal(fdm[s], f)


def _run(vc, vout, sx, sz, rx, rz):
    return P.fmst_pairs(GRID, np.atleast_2d(vc), np.atleast_2d(vout), [0], [sx], [sz],
                        [0, len(rx)], np.asarray(rx, np.float32), np.asarray(rz, np.float32))


def _stations():
    sx, sz = np.float32((90 - 30) * D2R), np.float32(102 * D2R)
    rx = ((90 - np.array([30, 27, 33.5, 26.5, 34.5])) * D2R).astype(np.float32)
    rz = (np.array([109, 104, 101, 108, 106.0]) * D2R).astype(np.float32)
    return sx, sz, rx, rz


def _smooth(a, b, c, d):
    nx, ny = GRID[:2]
    ix = np.arange(nx)[:, None] / nx
    iy = np.arange(ny)[None, :] / ny
    return (a + b * np.sin(2 * np.pi * c * ix) * np.cos(2 * np.pi * d * iy)).T.ravel()


def _gc(c1, l1, c2, l2):
    a = (np.sin((c2 - c1) / 2) ** 2 +
         np.sin((l2 - l1) / 2) ** 2 * np.cos(np.pi / 2 - c1) * np.cos(np.pi / 2 - c2))
    return 6371.0 * 2 * np.arctan2(np.sqrt(a), np.sqrt(1 - a))


def test_power_of_two_scaling_is_exact():
    sx, sz, rx, rz = _stations()
    vc, vout = _smooth(3.2, 0.25, 1.5, 2.0), _smooth(2.9, 0.2, 1.0, 1.3)
    t1, s1, f1, _ = _run(vc, vout, sx, sz, rx, rz)
    t2, s2, f2, _ = _run(2 * vc, 2 * vout, sx, sz, rx, rz)
    assert np.array_equal(t2, t1 / 2) and np.array_equal(s2, s1 / 2)
    assert np.array_equal(f2, f1 / 4)


def test_uniform_medium_times_are_distances():
    sx, sz, rx, rz = _stations()
    v = np.full(GRID[0] * GRID[1], 3.0)
    tt, ts, _, st = _run(v, v, sx, sz, rx, rz)
    assert (st == 0).all()
    t_exact = _gc(np.float64(sx), np.float64(sz), rx.astype(np.float64), rz.astype(np.float64)) / 3.0
    assert np.abs(tt / t_exact - 1).max() < 2e-3
    assert np.abs(ts / t_exact - 1).max() < 2e-3


def test_reciprocity():
    sx, sz, rx, rz = _stations()
    v = _smooth(3.2, 0.25, 1.5, 2.0)
    ta = _run(v, v, sx, sz, rx, rz)[0]
    tb = np.array([_run(v, v, rx[i], rz[i], [sx], [sz])[0][0] for i in range(len(rx))])
    assert np.abs(ta / tb - 1).max() < 2e-3


def test_weights_are_ray_time_derivatives():
    sx, sz, rx, rz = _stations()
    vc, g = _smooth(3.2, 0.25, 1.5, 2.0), _smooth(2.9, 0.2, 1.0, 1.3)
    _, _, f0, _ = _run(vc, g, sx, sz, rx[:1], rz[:1])
    f0 = f0[0]
    big = np.nonzero(np.abs(f0) > 0.05 * np.abs(f0).max())[0]
    h = 1e-2
    fd = []
    for k in big:
        gp, gm = g.copy(), g.copy()
        gp[k] += h
        gm[k] -= h
        tp = _run(vc, gp, sx, sz, rx[:1], rz[:1])[1][0]
        tm = _run(vc, gm, sx, sz, rx[:1], rz[:1])[1][0]
        fd.append((np.float64(tp) - tm) / (2 * h))
    err = np.abs(np.array(fd) - f0[big]) / np.abs(f0).max()
    # the weights leave out the final segment into the source, which the ray
    # time holds: the nodes there differ by a few per cent
    assert np.median(err) < 5e-3 and err.max() < 0.1
    # the weights read in the other orientation (y fastest) are far off
    nx, ny = GRID[:2]
    ft = f0.reshape(ny, nx).T.ravel()
    assert np.abs(ft[big] - np.array(fd)).max() > 0.1 * np.abs(f0).max()


def test_weight_sum_is_traced_ray_time():
    sx, sz, rx, rz = _stations()
    vc = _smooth(3.2, 0.25, 1.5, 2.0)
    v = np.full(GRID[0] * GRID[1], 3.0)
    _, ts, f, _ = _run(vc, v, sx, sz, rx, rz)
    r = -(f.sum(axis=1).astype(np.float64) * 3.0) / ts
    assert ((r > 0.98) & (r < 1.0)).all()
