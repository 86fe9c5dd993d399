"""The C-ABI library loads without a GPU and exports every function that
include/koral_hip.h declares (no compute calls: this runs on the CPU)."""
import ctypes
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "koral_hip.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", " ", src, flags=re.S)
    return sorted(set(re.findall(r"\b(koral_hip_\w+)\s*\(", src)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for must in ("koral_hip_op_explicit", "koral_hip_calc_u2p", "koral_hip_op_implicit",
                 "koral_hip_step_rk2imex", "koral_hip_implicit_batch"):
        assert must in names
    assert len(names) > 40


def test_library_exports_every_declared_function():
    from koral_amd import _lib
    L = _lib.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIBPATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (koral_hip_\w+)$", out, flags=re.M))
    assert set(declared_functions()) <= exported
    # the struct the header declares has the size the Python mirror uses
    from koral_amd import params as KP
    assert L.koral_hip_params_size() == ctypes.sizeof(KP.KoralParams)
