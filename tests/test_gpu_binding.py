"""KORAL's own operator signatures (op_explicit, op_implicit, calc_u2p,
do_correct; koral_amd/binding/koral_ops.c) compiled with gcc against
include/koral_hip.h and driven by a C program that plays ko.c: RK2IMEX steps
with ko.c's stage algebra on the host arrays, every operator in parity mode
through the C-ABI.  The state after each run equals the CPU oracle's steps bit
for bit, flags and time-step densities included."""
import os
import subprocess
import sys

import numpy as np
import pytest

import host_grid as H
from koral_amd import params as KP, problems as PR, grid as GR

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "host", "binding_driver")


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _setup(kind):
    if kind == "tube":
        P = KP.params_tube()
        spec, p = PR.tube_setup(P, nx=64)
    else:
        P = KP.params_torus_mhd()
        if kind == "torus_mhd_polar":
            P.polar_nc = 2
        spec, p = PR.torus_setup(P, n=(16, 16, 8))
    return P, spec, GR.to_aos(p)


def _rel_max(a, b):
    """max over variables of |a - b| / (|a| + |b| + 1e-4 max|b|) (the grid
    tests' measure: a quantity that cancels to zero carries the rounding of
    the terms that cancel)"""
    worst = 0.
    for iv in range(a.shape[-1]):
        x, y = a[..., iv], b[..., iv]
        r = np.abs(x - y) / (np.abs(x) + np.abs(y) + 1e-4 * np.abs(y).max() + 1e-300)
        worst = max(worst, float(r.max()))
    return worst


@pytest.mark.parametrize("kind,nsteps", [("tube", 4), ("torus_mhd", 2), ("torus_mhd_polar", 2)])
def test_koral_signature_binding_steps(oracle, tmp_path, kind, nsteps):
    assert os.path.exists(DRIVER), "binding driver not built (__graft_entry__.build())"
    P, spec, p = _setup(kind)
    _say("KORAL-signature binding, %s, %d RK2IMEX steps" % (kind, nsteps))
    S = oracle.GridState(P, oracle.Grid(spec), p)
    S.init_u()
    S.set_bc()
    dt = 1. / S.tstepden()
    t = 0.
    tms = []
    for _ in range(nsteps):
        tm, _c = S.step(t, dt)
        tms.append(tm)
        t += dt
        dt = 1. / tm
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    H.write_input(inp, P, spec, p)
    r = subprocess.run([DRIVER, str(inp), str(out), str(nsteps)], capture_output=True, text=True, timeout=300)
    _say(r.stderr[-2000:])
    assert r.returncode == 0
    pd, ud, fd, recs = H.read_output(out, spec, P.nv, nsteps)
    rp, ru = _rel_max(pd, S.p), _rel_max(ud, S.u)
    _say("  max rel p %.3e u %.3e; tstepdenmax %s vs oracle %s" % (rp, ru, [x[0] for x in recs], tms))
    # the device rounds as the oracle does: identical bit for bit
    assert np.array_equal(pd, S.p) and np.array_equal(ud, S.u)
    assert np.array_equal(fd, S.flags)
    assert [x[0] for x in recs] == tms
