"""The device limiters (DMM, DM4, mp5 in koral_amd/csrc/dev_grid.hpp)
compiled for the host are bit-identical to the sign-arithmetic forms of
finite.c:55-63 and finite.c:228-289 on special values (signed zeros, inf,
NaN) and a million random stencils."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_limiters_bit_identical(tmp_path):
    exe = tmp_path / "check_limiters"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", os.path.join(ROOT, "koral_amd", "csrc"),
                    "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "host", "check_limiters.cpp"),
                    "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
