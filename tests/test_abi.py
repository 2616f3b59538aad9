"""The C-ABI library loads and exports every function include/jdsg_swd.h
declares (no GPU needed: nothing is called)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "jdsg_swd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:void|int|char|long long)\s*\*?\s*(\w+)\s*\(", src,
                       flags=re.M)
    return sorted(set(names))


def test_header_declares_reference_abi():
    names = declared_functions()
    for n in ("surfdisp96_", "sregn96_", "slegn96_", "sregnpu_", "slegnpu_", "jdsg_surfdisp",
              "jdsg_plan_create", "jdsg_plan_execute"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from jdsurfg_amd import swd
    L = ctypes.CDLL(swd.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def declared_arity():
    """Parameter count of every function the header declares."""
    src = open(os.path.join(ROOT, "include", "jdsg_swd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    out = {}
    for m in re.finditer(r"^\s*(?:const\s+)?(?:void|int|char|long long)\s*\*?\s*(\w+)\s*\(([^)]*)\)\s*;",
                         src, flags=re.M):
        args = m.group(2).strip()
        out[m.group(1)] = 0 if args in ("", "void") else args.count(",") + 1
    return out


def test_python_bindings_match_the_header():
    """Every argtypes list the Python modules set has the header's parameter
    count (a missing or extra argument would shift every later one)."""
    from jdsurfg_amd import data, frechet, surrogate, swd
    L = swd.lib()
    surrogate._L()
    frechet._L()
    data._L()
    arity = declared_arity()
    assert len(arity) >= len(declared_functions()) - 2
    bad = []
    for name, n in arity.items():
        at = getattr(L, name).argtypes
        if at is not None and len(at) != n:
            bad.append((name, len(at), n))
    assert not bad, bad


def test_product_does_not_reach_the_oracle():
    """Nothing under jdsurfg_amd/ imports or links the checker."""
    pkg = os.path.join(ROOT, "jdsurfg_amd")
    for dp, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".hpp")):
                txt = open(os.path.join(dp, f)).read()
                assert "oracle" not in txt.lower().replace("the oracle", ""), f
