"""Per-layer flop counts of the frozen model from the op-counting build of
the oracle restatement (TEST INFRASTRUCTURE ONLY).

Runs oracle/_count/flopcount (make -C oracle flopcount) on strided columns of
the 256x256x64 benchmark grid at its 60 periods, Rayleigh modes 0-3 and Love
modes 0-1, and prints the counts:
  f_disp(wave)       flops of one period-equation evaluation per layer entry
  f_eig_phase(wave)  sregn96 / slegn96 per layer entry
  f_eig(wave)        sregnpu / slegnpu per layer entry and period pass
(+, -, *, /, sqrt, exp, log, sin, cos 1 each; pow 3; complex operations as
their real operations -- oracle/flopcount/flopd.hpp).  With --write the
counts replace those in jdsurfg_amd/flops_model.json.
The op-counting build is made here: the restatement's C sources are
rewritten as C++ into oracle/_count/ (double -> FlopD, double complex ->
FlopC, CMPLX -> FlopC, and intijr marked so the part its six calls per layer
share -- E, E^-1, the amplitude projections, the layer functions -- is
tallied apart) and compiled with g++ against flopd.hpp.  f_eig counts that
shared part once per layer (the non-redundant work); f_eig_as_executed
counts it six times, as sregn96.f90:1198-1317 executes it.
Usage: python oracle/flopcount/freeze.py [--columns 8] [--write]
"""
import argparse
import json
import os
import re
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def _to_cpp(src):
    s = src.replace("#include <complex.h>", "")
    s = s.replace("typedef double complex cplx;", "typedef FlopC cplx;")
    s = re.sub(r"\bdouble\b", "FlopD", s)
    s = s.replace("CMPLX(", "CMPLX_(")
    return s


def build():
    """Generate and compile the op-counting build; returns the executable."""
    ora = os.path.join(ROOT, "oracle")
    out = os.path.join(ora, "_count")
    os.makedirs(out, exist_ok=True)
    for f in ("disp", "ray", "love"):
        s = _to_cpp(open(os.path.join(ora, f"swd_oracle_{f}.c")).read())
        if f == "ray":
            # tally intijr's shared part: from its start to the combination c
            a = s.index("static FlopD intijr(")
            b = s.index("\n}\n", a)
            body = s[a:b]
            first = "  evalg0(x, m, NULL, wvno, om, om2, wvno2);"
            assert body.count(first) == 1 and body.count("      c = ") == 4
            body = body.replace(first, "  unsigned long long fs_ = g_flops;\n" + first)
            body = body.replace("      c = ", "      g_shared += g_flops - fs_; c = ")
            s = "extern unsigned long long g_shared;\n" + s[:a] + body + s[b:]
        open(os.path.join(out, f"swd_oracle_{f}.cpp"), "w").write(s)
    h = re.sub(r"\bdouble\b", "FlopD", open(os.path.join(ora, "swd_oracle.h")).read())
    open(os.path.join(out, "swd_oracle.h"), "w").write(h)
    exe = os.path.join(out, "flopcount")
    srcs = [os.path.join(HERE, f) for f in ("count_main.cpp", "count_disp.cpp", "count_ray.cpp",
                                              "count_love.cpp")]
    subprocess.run(["g++", "-O1", "-std=c++17", "-w", "-include",
                    os.path.join(HERE, "flopd.hpp"), "-I" + out, "-I" + HERE, "-o", exe] + srcs,
                   check=True)
    return exe


def count(columns=8):
    from jdsurfg_amd import workloads as W
    exe = build()
    w = W.grid256()
    cols = np.linspace(0, w.ncol - 1, columns).astype(int)
    lines = [str(len(cols))]
    for c in cols:
        s = slice(w.layer_off[c], w.layer_off[c + 1])
        lines.append(str(s.stop - s.start))
        for i in range(s.start, s.stop):
            lines.append(" ".join("%.9g" % float(a[i]) for a in (w.thk, w.vp, w.vs, w.rho)))
    per = w.periods[w.per_off[0]:w.per_off[1]]
    lines.append(str(len(per)))
    lines.append(" ".join(repr(float(t)) for t in per))
    r = subprocess.run([exe], input="\n".join(lines) + "\n", capture_output=True, text=True,
                       check=True)
    return json.loads(r.stdout)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--columns", type=int, default=8)
    ap.add_argument("--write", action="store_true")
    a = ap.parse_args()
    c = count(a.columns)
    print(json.dumps(c))
    if a.write:
        p = os.path.join(ROOT, "jdsurfg_amd", "flops_model.json")
        fm = json.load(open(p))
        for k in ("f_disp", "f_eig_phase", "f_eig", "f_eig_as_executed"):
            fm[k] = {w: round(v, 1) for w, v in c[k].items()}
        fm["source"] = ("op-counted: oracle/flopcount/freeze.py --columns %d (the oracle "
                        "restatement compiled with the counting scalar of "
                        "oracle/flopcount/flopd.hpp) on the 256x256x64 benchmark grid" % a.columns)
        with open(p, "w") as f:
            json.dump(fm, f, indent=1)
            f.write("\n")


if __name__ == "__main__":
    main()
