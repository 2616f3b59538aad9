"""The frozen flop model (jdsurfg_amd/flops_model.json) is the op count of
the oracle restatement: recount it with the counting build and check the
counting scalar on expressions whose counts are known."""
import importlib.util
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _freeze():
    spec = importlib.util.spec_from_file_location(
        "freeze", os.path.join(ROOT, "oracle", "flopcount", "freeze.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_counting_scalar_counts_known_expressions(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text('''
#include <cstdio>
unsigned long long g_flops = 0;
int main() {
  FlopD a = 1.5, b = 2.0, c = 3.0;
  unsigned long long n[6];
  FlopD r = a * b + c; n[0] = g_flops;                 // 2
  r = sqrt(a) / exp(b) - log(c); n[1] = g_flops;       // +5
  FlopC z(FlopD(1.0), FlopD(2.0)), y(FlopD(0.5), FlopD(-1.0));
  FlopC p = z * y; n[2] = g_flops;                     // +6
  p = z / y; n[3] = g_flops;                           // +9
  p = cexp(z); n[4] = g_flops;                         // +5
  bool t = a < b && fabs(c) > 1.0; n[5] = g_flops;     // +0
  printf("%llu %llu %llu %llu %llu %llu %d\\n", n[0], n[1], n[2], n[3], n[4], n[5], (int)t);
  return 0;
}''')
    exe = tmp_path / "t"
    subprocess.run(["g++", "-std=c++17", "-include",
                    os.path.join(ROOT, "oracle", "flopcount", "flopd.hpp"), str(src), "-o",
                    str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    assert [int(x) for x in out[:6]] == [2, 7, 13, 22, 27, 27] and out[6] == "1"


def test_frozen_model_is_the_op_count():
    fm = json.load(open(os.path.join(ROOT, "jdsurfg_amd", "flops_model.json")))
    c = _freeze().count(columns=4)
    for k in ("f_disp", "f_eig", "f_eig_phase", "f_eig_as_executed"):
        for w in ("rayleigh", "love"):
            assert c[k][w] == pytest.approx(fm[k][w], rel=0.01), (k, w)
    # the survey's estimates of the reference routines (183, 25; ~3000
    # non-redundant of ~3950 spent) bracket the counts
    assert 150 < c["f_disp"]["rayleigh"] < 220
    assert 2500 < c["f_eig"]["rayleigh"] < 3500 < c["f_eig_as_executed"]["rayleigh"]
