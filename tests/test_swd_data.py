"""The surfdataSC.dat reader (jdsg_read_swd_data, SurfTime::read_swd_data of
src/SWD/io.cpp:29-238) on the head of the reference's example data
(tests/golden/surfdataSC_head.dat: its period header and first 25 source
stations), against an independent tokenisation of the same file: tables,
modes, period indices, wave codes and positions equal (colatitude as the
reference forms it, (90.0 - lat) in double times the float DEG2RAD);
distances equal to the reference's own delsph (shared/spherical.cpp built
from source into oracle/_ref) bit for bit, and within float rounding of the
haversine in double; times = distance / velocity.  A ragged two-mode header
keeps the -1 padding past each mode's periods."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HEAD = os.path.join(HERE, "golden", "surfdataSC_head.dat")
FULL = "/root/reference/example/surfdataSC.dat"

def tokenise(path):
    lines = [ln for ln in open(path).read().splitlines() if ln.strip()]
    pos, tables = 0, []
    for _ in range(4):
        head = lines[pos].split("#")[0].split()
        pos += 1
        nmode = int(float(head[0]))
        kmax = [int(float(x)) for x in head[1:1 + nmode]]
        nt = max(kmax) if nmode else 0
        t = -np.ones((nt, nmode), np.float32)
        for im in range(nmode):
            vals = [float(x) for x in lines[pos].split()]
            pos += 1
            t[:kmax[im], im] = vals[:kmax[im]]
        tables.append((t, kmax))
    pairs, recv = [], []
    while pos < len(lines):
        f = lines[pos][1:].split()
        pos += 1
        lat, lon, mode, per, wt, vt = float(f[0]), float(f[1]), *map(int, f[2:6])
        code = {(2, 0): 0, (2, 1): 1, (1, 0): 2}.get((wt, vt), 3)
        rs = []
        while pos < len(lines) and not lines[pos].startswith("#"):
            rs.append([float(x) for x in lines[pos].split()[:3]])
            pos += 1
        pairs.append((lat, lon, mode, per - 1, code, len(rs)))
        recv.extend(rs)
    return tables, pairs, np.array(recv, np.float64).reshape(-1, 3)


def colat_lon(lat, lon):
    d2r = np.float32(np.pi / 180.0)
    lat = np.asarray(lat, np.float32).astype(np.float64)
    return ((90.0 - lat) * np.float64(d2r)).astype(np.float32), \
        (np.asarray(lon, np.float32) * d2r).astype(np.float32)


def ref_delsph():
    import ctypes as C
    p = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libsphref.so")
    if not os.path.exists(p):
        return None
    L = C.CDLL(p)
    L.delsph.argtypes = [C.c_float] * 4
    L.delsph.restype = C.c_float
    return L.delsph


def test_reader_matches_an_independent_tokenisation():
    from jdsurfg_amd import data as D
    got = D.read_swd_data(HEAD)
    tables, pairs, recv = tokenise(HEAD)
    for w, name in enumerate(D.WAVES):
        t, k = got["tables"][name]
        assert np.array_equal(t, tables[w][0]) and list(k) == tables[w][1]
    assert len(got["wave"]) == len(pairs) == 25
    assert list(got["wave"]) == [p[4] for p in pairs]
    assert list(got["mode"]) == [p[2] for p in pairs]
    assert list(got["period"]) == [p[3] for p in pairs]
    assert list(np.diff(got["pair_off"])) == [p[5] for p in pairs]
    sc, sl = colat_lon(np.array([p[0] for p in pairs]), np.array([p[1] for p in pairs]))
    assert np.array_equal(got["src_colat"], sc) and np.array_equal(got["src_lon"], sl)
    rc, rl = colat_lon(recv[:, 0], recv[:, 1])
    assert np.array_equal(got["rcv_colat"], rc) and np.array_equal(got["rcv_lon"], rl)
    # haversine in double on the same float positions
    src_idx = np.repeat(np.arange(len(pairs)), np.diff(got["pair_off"]))
    c1, l1 = sc[src_idx].astype(np.float64), sl[src_idx].astype(np.float64)
    c2, l2 = rc.astype(np.float64), rl.astype(np.float64)
    a = np.sin((c2 - c1) / 2) ** 2 + np.sin((l2 - l1) / 2) ** 2 * np.sin(c1) * np.sin(c2)
    d = 6371.0 * 2 * np.arctan2(np.sqrt(a), np.sqrt(1 - a))
    assert np.abs(got["dist"] - d).max() <= 1e-5 * d.max() + 1e-3
    f = ref_delsph()
    if f is not None:
        ref = np.array([f(float(sc[s]), float(sl[s]), float(a_), float(b_))
                        for s, a_, b_ in zip(src_idx, rc, rl)], np.float32)
        assert np.array_equal(got["dist"], ref)
    assert np.array_equal(got["obst"], (got["dist"] / recv[:, 2].astype(np.float32)).astype(np.float32))


@pytest.mark.skipif(not os.path.exists(FULL), reason="reference example data not present")
def test_reader_counts_the_full_example():
    from jdsurfg_amd import data as D
    got = D.read_swd_data(FULL)
    lines = [ln for ln in open(FULL).read().splitlines() if ln.strip()]
    n_src = sum(1 for ln in lines if ln.startswith("#"))
    assert len(got["wave"]) == n_src == 3046
    assert got["pair_off"][-1] == len(lines) - n_src - 6   # 6 header lines (2 period rows)
    assert got["tables"]["Rc"][1].tolist() == [19] and got["tables"]["Lc"][1].tolist() == []
    assert (got["obst"] > 0).all() and (got["period"] >= 0).all()


def test_ragged_mode_rows_keep_their_padding(tmp_path):
    """Mode 0 with three periods and mode 1 with one: the table is 3 x 2
    with -1 below mode 1's period (the reference's strtof would store 0
    there, io.cpp:69-75)."""
    from jdsurfg_amd import data as D
    p = tmp_path / "ragged.dat"
    p.write_text("2 3 1\n10 20 30\n40\n0\n0\n0\n# 31 108 0 1 2 0\n33.7 105.7 3.1\n")
    got = D.read_swd_data(p)
    t, k = got["tables"]["Rc"]
    assert list(k) == [3, 1]
    assert np.array_equal(t, np.array([[10, 40], [20, -1], [30, -1]], np.float32))


def test_reader_rejects_a_malformed_file(tmp_path):
    from jdsurfg_amd import data as D
    p = tmp_path / "bad.dat"
    p.write_text("1 2\n4 6\n0\n0\n0\n# 31 108 0 1 2 0\n33.7 105.7\n")
    with pytest.raises(RuntimeError, match="lat lon velocity"):
        D.read_swd_data(p)
    with pytest.raises(RuntimeError, match="cannot open"):
        D.read_swd_data(tmp_path / "missing.dat")
