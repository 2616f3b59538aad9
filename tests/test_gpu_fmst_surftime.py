"""SurfTime::travel_time, frechet_matrix and compute_grad as one library call
each (jdsg_surftime_travel_time / _frechet / _compute_grad), on a 33 x 36 x 9
model over the reference example's grid (35.3 N, 99.7 E, 0.3 deg) with the
station pairs of the example data file that lie inside it.

The composition is checked against the oracle run on the same maps and
kernels the GPU made (the maps themselves are checked against the oracle in
the SurfTime tests): the FMST restatement (bit-identical to the reference)
for the times and rays, the restated row and gradient loops of
SurfaceWave.cpp for the rows and the gradient.  Bars: data and rows bit for
bit; the gradient within 1e-5 of its largest entry (the GPU adds the rows'
contributions with atomics, in no fixed order).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NX, NY, LAT0, LON0, DLAT, DLON = 33, 36, 35.3, 99.7, 0.3, 0.3
DEP = np.array([0, 2, 5, 10, 15, 20, 30, 40, 60], np.float32)


def _data_inside():
    from jdsurfg_amd import data as D
    d = D.read_swd_data(os.path.join(GOLD, "surfdataSC_head.dat"))
    d2r = np.pi / 180
    lo_c, hi_c = (90 - (LAT0 - DLAT)) * d2r + 1e-4, (90 - (LAT0 - DLAT * (NX - 2))) * d2r - 1e-4
    lo_l, hi_l = (LON0 + DLON) * d2r + 1e-4, (LON0 + DLON * (NY - 2)) * d2r - 1e-4
    inside = lambda c, lo: (c > lo_c) & (c < hi_c) & (lo > lo_l) & (lo < hi_l)
    keep_p, rc, rl, obs, off = [], [], [], [], [0]
    po = d["pair_off"]
    for p in range(len(po) - 1):
        s = slice(po[p], po[p + 1])
        k = inside(d["rcv_colat"][s], d["rcv_lon"][s])
        if not inside(d["src_colat"][p], d["src_lon"][p]) or not k.any():
            continue
        keep_p.append(p)
        rc += list(d["rcv_colat"][s][k])
        rl += list(d["rcv_lon"][s][k])
        obs += list(d["obst"][s][k])
        off.append(off[-1] + int(k.sum()))
    kp = np.array(keep_p)
    out = dict(d)
    for n in ("wave", "mode", "period", "src_colat", "src_lon"):
        out[n] = d[n][kp]
    out["pair_off"] = np.array(off, np.int32)
    out["rcv_colat"] = np.array(rc, np.float32)
    out["rcv_lon"] = np.array(rl, np.float32)
    out["obst"] = np.array(obs, np.float32)
    return out


def _model():
    ix = np.arange(NX)[:, None, None] / NX
    iy = np.arange(NY)[None, :, None] / NY
    z = DEP[None, None, :]
    vs = 2.9 + 0.035 * z + 0.15 * np.sin(2 * np.pi * 1.3 * ix) * np.cos(2 * np.pi * 0.8 * iy)
    return vs.astype(np.float32)


def _setup():
    import torch
    from jdsurfg_amd import fmst, swd
    d = _data_inside()
    tables = {w: d["tables"][w][0] for w in ("Rc", "Rg", "Lc", "Lg") if d["tables"][w][0].size}
    S = swd.SurfTime(NX, NY, len(DEP), DEP, tables)
    F = fmst.from_data(d, NX, NY, LAT0, LON0, DLAT, DLON)
    vs = _model()
    vs_d = torch.from_numpy(np.ascontiguousarray(vs.ravel(order="F"))).cuda()
    return d, S, F, vs, vs_d


def _oracle_fmst(S, F, vc, vout, d):
    from oracle import pyoracle as P
    c = F._keep
    tt, ts, fdm, st = P.fmst_pairs((NX, NY, LAT0, LON0, DLAT, DLON), vc.T, vout.T, c[0], c[1],
                                   c[2], c[3], c[4], c[5], nthreads=16)
    assert (st == 0).all()
    exp = tt.copy()
    for p in range(F.npair):
        if d["wave"][p] in (1, 3):
            exp[c[3][p]:c[3][p + 1]] = ts[c[3][p]:c[3][p + 1]]
    return exp, fdm.reshape(-1, NY, NX).transpose(0, 2, 1)


def test_travel_time():
    import torch
    from jdsurfg_amd import fmst
    d, S, F, vs, vs_d = _setup()
    print(f"case surftime travel_time: pairs {F.npair} receivers {F.nrcv} kmax {S.kmax}",
          flush=True)
    data, rst = fmst.travel_time(S, F, vs_d)
    m = S.map(vs_d, S.alloc(kernel=False))
    torch.cuda.synchronize()
    vc, vout, _ = S.to_host(m)
    exp, _ = _oracle_fmst(S, F, vc, vout, d)
    assert (rst.cpu().numpy() == 0).all()
    assert np.array_equal(data.cpu().numpy(), exp)


def test_frechet_matrix_and_gradient(oracle_rest):
    import torch
    from jdsurfg_amd import fmst
    d, S, F, vs, vs_d = _setup()
    print(f"case surftime frechet_matrix/compute_grad: pairs {F.npair} receivers {F.nrcv}",
          flush=True)
    data, rst, off, col, val = fmst.frechet_matrix(S, F, vs_d)
    k = S.kernel(vs_d, S.alloc(kernel=True))
    torch.cuda.synchronize()
    vc, vout, ker, _ = S.to_host(k)
    exp, fdm = _oracle_fmst(S, F, vc, vout, d)
    assert (rst.cpu().numpy() == 0).all()
    assert np.array_equal(data.cpu().numpy(), exp)
    row_idx = F._keep[0][np.repeat(np.arange(F.npair), np.diff(F._keep[3]))]
    ro, co, va = oracle_rest.frechet_rows(ker, row_idx, fdm)
    print("rows", len(ro) - 1, "nonzeros", len(co), flush=True)
    assert np.array_equal(off.cpu().numpy(), ro)
    assert np.array_equal(col.cpu().numpy(), co) and np.array_equal(val.cpu().numpy(), va)
    obst = torch.from_numpy(d["obst"]).cuda()
    data2, grad, rst2 = fmst.compute_grad(S, F, vs_d, obst)
    assert np.array_equal(data2.cpu().numpy(), exp)
    g_ref = oracle_rest.frechet_grad(ker, row_idx, fdm, exp - d["obst"], vs)
    g = grad.cpu().numpy()
    err = np.abs(g - g_ref).max() / np.abs(g_ref).max()
    print("gradient max rel", err, flush=True)
    assert err <= 1e-5
