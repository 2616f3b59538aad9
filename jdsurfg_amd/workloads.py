"""Seeded synthetic layered-Earth workloads for the dispersion/kernel path.

Stack convention: a column of N entries holds N-1 finite layers plus the
half-space (thickness 0).  Layer properties are taken at the layer mid-depth;
the half-space takes the value at the bottom depth.  Vp and rho follow from
Vs through Brocher (2005), mirrored from src/SWD/empirical.cpp:9-18 in float
arithmetic (Vp, rho rounded to float like the reference).

Named workloads (sizes as the benchmark configurations state them):
  grid20    20x20 columns x 30 x 2 km layers, Vs = 2.6+0.035 z plus a seeded
            Gaussian random field (sigma 0.15 km/s, 30 km correlation),
            20 log-spaced periods 2-40 s, Rayleigh mode 0.
  grid128   128x128 columns x 48 x 1.5 km, same field plus a random
            sediment top (0-4 km thick, Vs 0.8-2.0 km/s), 40 periods 1-60 s,
            Rayleigh modes 0-2, Love modes 0-1.
  uneven    50,000 columns, 16-128 layers of 0.5-4 km (Vs trend capped at
            4.6 km/s, the range of the Brocher fit), per-column sorted period
            sets of 8-64 periods in 1-80 s, 1-3 modes.
  grid256   256x256 columns x 64 x 1 km, depth trend with a +-8 % Vs
            checkerboard of 40 km cells, 60 periods 1-80 s, Rayleigh modes
            0-3, Love modes 0-1.
"""
import numpy as np


def brocher(vs):
    """Vs (km/s) -> (Vp, rho) as float32, empirical.cpp:9-18 evaluated in
    double and rounded to float exactly as the reference does."""
    vs = np.asarray(vs, np.float32).astype(np.float64)
    vp = (0.9409 + 2.0947 * vs - 0.8206 * vs ** 2 + 0.2683 * vs ** 3
          - 0.0251 * vs ** 4).astype(np.float32)
    v = vp.astype(np.float64)
    rho = (1.6612 * v - 0.4721 * v ** 2 + 0.0671 * v ** 3 - 0.0043 * v ** 4
           + 0.000106 * v ** 5).astype(np.float32)
    return vp, rho


def brocher_deriv(vp, vs):
    """(drho/dVp, dVp/dVs) as float32, empirical.cpp:26-34."""
    vp = np.asarray(vp, np.float32).astype(np.float64)
    vs = np.asarray(vs, np.float32).astype(np.float64)
    drda = (1.6612 - 0.4721 * 2 * vp + 0.0671 * 3 * vp ** 2 - 0.0043 * 4 * vp ** 3
            + 0.000106 * 5 * vp ** 4).astype(np.float32)
    dadb = (2.0947 - 0.8206 * 2 * vs + 0.2683 * 3 * vs ** 2
            - 0.0251 * 4 * vs ** 3).astype(np.float32)
    return drda, dadb


def gaussian_field(shape, spacing, corr_km, sigma, rng):
    """Stationary Gaussian random field by FFT filtering of white noise."""
    noise = rng.standard_normal(shape)
    f = np.fft.fftn(noise)
    k2 = np.zeros(shape)
    for ax, (n, h) in enumerate(zip(shape, spacing)):
        k = np.fft.fftfreq(n, d=h) * 2 * np.pi
        sh = [1] * len(shape)
        sh[ax] = n
        k2 = k2 + (k.reshape(sh)) ** 2
    filt = np.exp(-0.25 * k2 * corr_km ** 2)
    out = np.real(np.fft.ifftn(f * filt))
    out -= out.mean()
    s = out.std()
    if s > 0:
        out *= sigma / s
    return out


class Workload:
    """A batch of layer stacks plus period sets and mode counts.

    layer_off/per_off are CSR offsets (ncol+1); thk, vp, vs, rho are float32
    flat arrays; periods float64; nmode is (ncol, 2) int32 = [Rayleigh, Love]
    mode counts; group says whether U kernels (three root chains) are wanted.
    """

    def __init__(self, name, layer_off, thk, vp, vs, rho, per_off, periods, nmode,
                 group, grid=None):
        self.name = name
        self.layer_off = np.asarray(layer_off, np.int32)
        self.thk = np.asarray(thk, np.float32)
        self.vp = np.asarray(vp, np.float32)
        self.vs = np.asarray(vs, np.float32)
        self.rho = np.asarray(rho, np.float32)
        self.per_off = np.asarray(per_off, np.int32)
        self.periods = np.asarray(periods, np.float64)
        self.nmode = np.asarray(nmode, np.int32).reshape(-1, 2)
        self.group = bool(group)
        self.grid = grid

    @property
    def ncol(self):
        return len(self.layer_off) - 1

    def column_periods(self):
        """Number of (column, period) units in the batch."""
        return int(self.per_off[-1])

    def subset(self, cols):
        """A new workload holding the listed columns in the given order."""
        cols = np.asarray(cols)
        lo, po = self.layer_off, self.per_off
        lsz = lo[cols + 1] - lo[cols]
        psz = po[cols + 1] - po[cols]
        lidx = np.concatenate([np.arange(lo[c], lo[c + 1]) for c in cols])
        pidx = np.concatenate([np.arange(po[c], po[c + 1]) for c in cols])
        return Workload(self.name + "-sub", np.concatenate([[0], np.cumsum(lsz)]),
                        self.thk[lidx], self.vp[lidx], self.vs[lidx], self.rho[lidx],
                        np.concatenate([[0], np.cumsum(psz)]), self.periods[pidx],
                        self.nmode[cols], self.group)


def _uniform_grid(name, nx, ny, nlay, dz, vs_cols, periods, nmode, group):
    """Columns with nlay finite layers of thickness dz plus the half-space."""
    ncol = nx * ny
    n = nlay + 1
    thk = np.full((ncol, n), dz, np.float32)
    thk[:, -1] = 0.0
    vs = np.asarray(vs_cols, np.float32).reshape(ncol, n)
    vp, rho = brocher(vs)
    periods = np.asarray(periods, np.float64)
    K = len(periods)
    return Workload(name, np.arange(ncol + 1) * n, thk.ravel(), vp.ravel(), vs.ravel(),
                    rho.ravel(), np.arange(ncol + 1) * K, np.tile(periods, ncol),
                    np.tile(np.asarray(nmode, np.int32), (ncol, 1)), group,
                    grid=(nx, ny, n))


def _trend(z):
    return 2.6 + 0.035 * z


def grid20(seed=20, nx=20, ny=20, group=True):
    nlay, dz = 30, 2.0
    zmid = np.concatenate([(np.arange(nlay) + 0.5) * dz, [nlay * dz]])
    rng = np.random.default_rng(seed)
    fld = gaussian_field((nx, ny, nlay + 1), (10.0, 10.0, dz), 30.0, 0.15, rng)
    vs = _trend(zmid)[None, None, :] + fld
    periods = np.geomspace(2.0, 40.0, 20)
    return _uniform_grid("grid20", nx, ny, nlay, dz, vs.reshape(-1, nlay + 1), periods,
                         (1, 0), group)


def grid128(seed=128, nx=128, ny=128, group=True):
    nlay, dz = 48, 1.5
    zmid = np.concatenate([(np.arange(nlay) + 0.5) * dz, [nlay * dz]])
    rng = np.random.default_rng(seed)
    fld = gaussian_field((nx, ny, nlay + 1), (10.0, 10.0, dz), 30.0, 0.15, rng)
    vs = _trend(zmid)[None, None, :] + fld
    hsed = rng.uniform(0.0, 4.0, (nx, ny))
    vsed = rng.uniform(0.8, 2.0, (nx, ny))
    sed = zmid[None, None, :] < hsed[:, :, None]
    vs = np.where(sed, vsed[:, :, None], vs)
    periods = np.geomspace(1.0, 60.0, 40)
    return _uniform_grid("grid128", nx, ny, nlay, dz, vs.reshape(-1, nlay + 1), periods,
                         (3, 2), group)


def grid256(seed=256, nx=256, ny=256, group=True):
    nlay, dz = 64, 1.0
    zmid = np.concatenate([(np.arange(nlay) + 0.5) * dz, [nlay * dz]])
    x = (np.arange(nx) + 0.5) * 10.0
    y = (np.arange(ny) + 0.5) * 10.0
    cell = 40.0
    sx = np.floor(x / cell).astype(int)
    sy = np.floor(y / cell).astype(int)
    sz = np.floor(zmid / cell).astype(int)
    sign = (-1.0) ** ((sx[:, None, None] + sy[None, :, None] + sz[None, None, :]) % 2)
    rng = np.random.default_rng(seed)
    phase = rng.integers(0, 2)
    sign = sign * (1 if phase == 0 else -1)
    vs = _trend(zmid)[None, None, :] * (1.0 + 0.08 * sign)
    periods = np.geomspace(1.0, 80.0, 60)
    return _uniform_grid("grid256", nx, ny, nlay, dz, vs.reshape(-1, nlay + 1), periods,
                         (4, 2), group)


def uneven(seed=50000, ncol=50000, group=True):
    rng = np.random.default_rng(seed)
    nl = rng.integers(16, 129, ncol)  # finite layers
    kk = rng.integers(8, 65, ncol)
    nm = rng.integers(1, 4, (ncol, 2))
    layer_sizes = nl + 1
    layer_off = np.concatenate([[0], np.cumsum(layer_sizes)])
    per_off = np.concatenate([[0], np.cumsum(kk)])
    thk = np.empty(layer_off[-1], np.float32)
    vs = np.empty(layer_off[-1], np.float32)
    periods = np.empty(per_off[-1], np.float64)
    for c in range(ncol):
        n = nl[c]
        h = rng.uniform(0.5, 4.0, n)
        ztop = np.concatenate([[0.0], np.cumsum(h)])
        zmid = np.concatenate([ztop[:-1] + 0.5 * h, [ztop[-1]]])
        v = np.clip(np.minimum(_trend(zmid), 4.6) + rng.normal(0.0, 0.075, n + 1), 1.0, 4.8)
        s = slice(layer_off[c], layer_off[c + 1])
        thk[s] = np.concatenate([h, [0.0]])
        vs[s] = v
        periods[per_off[c]:per_off[c + 1]] = np.sort(
            np.exp(rng.uniform(np.log(1.0), np.log(80.0), kk[c])))
    vp, rho = brocher(vs)
    return Workload("uneven", layer_off, thk, vp, vs, rho, per_off, periods, nm, group)


def random_columns(ncol, nlay, seed, nmode=(1, 1), group=True, periods=None,
                   sediments=False, water=False):
    """Small seeded batches for parity tests: trend + noise, optional soft
    sediments (reversed dispersion) or a water top layer (Vs = 0)."""
    rng = np.random.default_rng(seed)
    dz = 60.0 / nlay
    zmid = np.concatenate([(np.arange(nlay) + 0.5) * dz, [nlay * dz]])
    vs = _trend(zmid)[None, :] + rng.normal(0, 0.1, (ncol, nlay + 1))
    if sediments:
        vs[:, :2] = rng.uniform(0.8, 2.0, (ncol, 1))
    if periods is None:
        periods = np.geomspace(2.0, 40.0, 12)
    w = _uniform_grid("random", ncol, 1, nlay, dz, vs, periods, nmode, group)
    if water:
        vs2 = w.vs.reshape(ncol, -1)
        vp2 = w.vp.reshape(ncol, -1)
        rho2 = w.rho.reshape(ncol, -1)
        vs2[:, 0] = 0.0
        vp2[:, 0] = 1.5
        rho2[:, 0] = 1.0
    return w


BUILDERS = {"grid20": grid20, "grid128": grid128, "grid256": grid256, "uneven": uneven}


def station_pairs(seed=3046, nx=33, ny=36, lat0=35.3, lon0=99.7, dlat=0.3, dlon=0.3, nz=41,
                  nsrc=80, nsta=240, nrcv=(13, 14)):
    """The travel-time workload at the reference example's scale
    (example/MOD: 33 x 36 nodes from 35.3 N, 99.7 E at 0.3 deg, 41 depth
    nodes; example/surfdataSC.dat: 3046 station pairs, 41,068 times, Rayleigh
    phase and group at 19 periods 4-40 s).  A seeded synthetic model (depth
    trend plus smooth lateral variation), 240 stations inside the grid,
    `nsrc` sources each paired with 13-14 receivers at every one of the 38
    period indices (Rc periods 0-18, then Rg).  Returns a dict in the form
    jdsurfg_amd.data.read_swd_data gives, plus grid, dep and vs."""
    rng = np.random.default_rng(seed)
    dep = np.linspace(0.0, 80.0, nz).astype(np.float32)
    ix = np.arange(nx)[:, None, None] / nx
    iy = np.arange(ny)[None, :, None] / ny
    z = dep[None, None, :]
    vs = (2.9 + 0.02 * z + 0.12 * np.sin(2 * np.pi * 1.2 * ix + 0.3) * np.cos(2 * np.pi * 0.9 * iy) *
          np.exp(-z / 60.0)).astype(np.float32)
    per = np.arange(4.0, 41.0, 2.0)
    tables = {"Rc": (per[:, None].copy(), np.array([len(per)], np.int32)),
              "Rg": (per[:, None].copy(), np.array([len(per)], np.int32)),
              "Lc": (np.zeros((0, 0)), np.zeros(0, np.int32)),
              "Lg": (np.zeros((0, 0)), np.zeros(0, np.int32))}
    d2r = np.pi / 180.0
    lat = rng.uniform(lat0 - dlat * (nx - 2) + 0.05, lat0 - dlat - 0.05, nsta)
    lon = rng.uniform(lon0 + dlon + 0.05, lon0 + dlon * (ny - 2) - 0.05, nsta)
    colat = ((90.0 - lat) * d2r).astype(np.float32)
    lonr = (lon * d2r).astype(np.float32)
    wave, mode, period, sc, sl, rc, rl, off = [], [], [], [], [], [], [], [0]
    for s in range(nsrc):
        for k in range(2 * len(per)):
            n = int(rng.integers(nrcv[0], nrcv[1] + 1))
            others = rng.choice(np.delete(np.arange(nsta), s), n, replace=False)
            wave.append(0 if k < len(per) else 1)
            mode.append(0)
            period.append(k % len(per))
            sc.append(colat[s])
            sl.append(lonr[s])
            rc += list(colat[others])
            rl += list(lonr[others])
            off.append(off[-1] + n)
    return {"grid": (nx, ny, lat0, lon0, dlat, dlon), "dep": dep, "vs": vs, "tables": tables,
            "pair_off": np.array(off, np.int32), "wave": np.array(wave, np.int32),
            "mode": np.array(mode, np.int32), "period": np.array(period, np.int32),
            "src_colat": np.array(sc, np.float32), "src_lon": np.array(sl, np.float32),
            "rcv_colat": np.array(rc, np.float32), "rcv_lon": np.array(rl, np.float32)}
