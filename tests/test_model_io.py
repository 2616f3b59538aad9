"""The model file (MOD) reader and the per-iteration snapshot writer
(jdsg_read_model*, jdsg_write_model: DSurfTomo::read_model of
src/surftomo/io.cpp:27-114 and the mod_iter*.dat writer of
src/surftomo/main.cpp:53-59).  Checked against an independent tokenisation
of the same files: the reference's example MOD and MOD.true when that
directory is present, and small files written here with the example's
header comments and ragged line breaks."""
import os

import numpy as np
import pytest

EXAMPLE = "/root/reference/example"


def tokenise(path):
    lines = open(path).read().splitlines()
    nx, ny, nz = (int(v) for v in lines[0].split()[:3])
    g = [np.float32(v) for v in lines[1].split()[:2] + lines[2].split()[:2]]
    dep = np.array(lines[3].split()[:nz], np.float32)
    vals = np.array(" ".join(lines[4:]).split()[:nx * ny * nz], np.float32)
    # the file runs i fastest, then j, then k
    vs = vals.reshape(nz, ny, nx).transpose(2, 1, 0)
    return (nx, ny, nz), g, dep, vs


def write_mod(path, vs, dep, g, per_line=7):
    nx, ny, nz = vs.shape
    flat = vs.transpose(2, 1, 0).ravel()
    with open(path, "w") as f:
        f.write(f"{nx} {ny} {nz} \t\t\t\t#nlat nlon nz\n")
        f.write(f"{g[0]} {g[1]} \t\t\t\t# lat0 lon0\n")
        f.write(f"{g[2]} {g[3]}  \t\t\t\t# dlat dlon\n")
        f.write(" ".join(f"{d:g}" for d in dep) + " \n")
        for s in range(0, flat.size, per_line):
            f.write(" ".join(repr(float(v)) for v in flat[s:s + per_line]) + " \n")


def check(path):
    from jdsurfg_amd import data as D
    m = D.read_model(path)
    dims, g, dep, vs = tokenise(path)
    assert (m["nx"], m["ny"], m["nz"]) == dims
    assert [m["goxd"], m["gozd"], m["dvxd"], m["dvzd"]] == g
    assert np.array_equal(m["dep"], dep)
    assert m["vs"].shape == vs.shape and np.array_equal(m["vs"], vs)
    assert m["vs"].flags.f_contiguous
    assert np.array_equal(m["lat"], np.float32(g[0]) - np.arange(dims[0], dtype=np.float32) * g[2])
    return m


@pytest.mark.parametrize("name", ["MOD", "MOD.true"])
def test_reference_example_models(name):
    path = os.path.join(EXAMPLE, name)
    if not os.path.exists(path):
        pytest.skip("the reference's example directory is not present")
    m = check(path)
    assert (m["nx"], m["ny"], m["nz"]) == (33, 36, 41)
    assert m["dep"][0] == 0 and m["dep"][-1] == 80


def test_small_model_with_ragged_lines(tmp_path):
    rng = np.random.default_rng(7)
    vs = rng.uniform(2.0, 4.6, (5, 4, 3)).astype(np.float32)
    dep = np.array([0, 2.5, 7], np.float32)
    p = tmp_path / "MOD"
    write_mod(p, vs, dep, (35.3, 99.7, 0.3, 0.25))
    m = check(p)
    assert np.array_equal(m["vs"], vs)
    # a node's position in the array is its (lat, lon, depth) index
    assert m["vs"][4, 1, 2] == vs[4, 1, 2]


def test_snapshot_writer_round_trip(tmp_path):
    from jdsurfg_amd import data as D
    rng = np.random.default_rng(11)
    nx, ny, nz = 4, 3, 2
    vs = rng.uniform(2.0, 4.6, (nx, ny, nz)).astype(np.float32)
    dep = np.array([0, 3.5], np.float32)
    g = [np.float32(v) for v in (35.3, 99.7, 0.3, 0.25)]
    out = tmp_path / "mod_iter0.dat"
    D.write_model(out, vs, dep, *g)
    rows = open(out).read().splitlines()
    assert len(rows) == nx * ny * nz
    r = 0
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                lat = np.float32(g[0] - np.float32(i) * g[2])
                lon = np.float32(g[1] + np.float32(j) * g[3])
                want = "%g %g %g %g" % (lon, lat, dep[k], vs[i, j, k])
                assert rows[r] == want
                r += 1


def test_errors(tmp_path):
    from jdsurfg_amd import data as D
    with pytest.raises(RuntimeError, match="cannot open"):
        D.read_model(tmp_path / "absent")
    vs = np.ones((3, 2, 2), np.float32)
    p = tmp_path / "MOD"
    write_mod(p, vs, np.array([0, 1], np.float32), (1, 2, 3, 4))
    text = open(p).read().split("\n")
    open(p, "w").write("\n".join(text[:-3]))  # drop the last values
    with pytest.raises(RuntimeError, match="expected 12 velocities"):
        D.read_model(p)
    open(p, "w").write("3 2\n1 2\n3 4\n0 1\n")
    with pytest.raises(RuntimeError, match="bad header"):
        D.read_model(p)
