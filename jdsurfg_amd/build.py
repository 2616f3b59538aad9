"""Build the HIP library in-tree: hipcc for gfx950, one shared object."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib", "libjdsg_swd.so")
SOURCES = ["swd_chain.hip", "swd_eigen.hip", "swd_surftime.hip", "swd_mlp.hip", "swd_surrogate.hip",
           "swd_batch.hip", "swd_frechet.hip", "swd_lsmr.hip", "swd_smooth.hip", "swd_nlopt.hip", "swd_fmst.hip", "swd_io.cpp", "swd_api.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-munsafe-fp-atomics"]


def build(verbose=True):
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    objs = []
    try:
        for src in SOURCES:
            s = os.path.join(CSRC, src)
            o = os.path.join(HERE, "lib", src + ".o")
            lang = ["-x", "hip"]
            cmd = [HIPCC] + FLAGS + lang + ["-c", s, "-o", o]
            if verbose:
                print(" ".join(cmd), flush=True)
            objs.append(o)
            subprocess.run(cmd, check=True)
        cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB + ".tmp"] + objs
        subprocess.run(cmd, check=True)
        os.replace(LIB + ".tmp", LIB)
    finally:
        for o in objs:
            if os.path.exists(o):
                os.remove(o)
    return LIB


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
