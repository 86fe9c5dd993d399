"""GPU parity of the grid operators (ghost fill, explicit operator, u2p sweep
with fixups, implicit sweep, RK2IMEX step) against the CPU oracle on small
grids of the three step workloads, through the C-ABI."""
import sys

import numpy as np
import pytest

from koral_amd import params as KP, problems as PR, grid as GR

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return np.abs(a - b) / (np.abs(a) + np.abs(b) + 1e-300)


def _setup(kind):
    if kind == "tube":
        P = KP.params_tube()
        spec, p = PR.tube_setup(P, nx=64)
    elif kind == "torus_mhd":
        P = KP.params_torus_mhd()
        spec, p = PR.torus_setup(P, n=(16, 16, 8))
    else:
        P = KP.params_torus_rad()
        spec, p = PR.torus_setup(P, n=(12, 12, 8))
    return P, spec, GR.to_aos(p)


def _oracle_state(oracle, P, spec, p):
    G = oracle.Grid(spec)
    S = oracle.GridState(P, G, p)
    S.init_u()
    S.set_bc()
    return G, S


@pytest.mark.parametrize("kind", ["tube", "torus_mhd", "torus_rad"])
def test_grid_init_and_bc(oracle, kind):
    from koral_amd.gridrun import DeviceGrid
    P, spec, p = _setup(kind)
    _say("grid init/bc parity", kind, spec.nx, spec.ny, spec.nz)
    G, S = _oracle_state(oracle, P, spec, p)
    dg = DeviceGrid(P, spec)
    dg.upload(p)
    dg.p2u()
    dg.set_bc()
    pd, ud, fd = dg.download()
    assert _rel(pd, S.p).max() < 1e-14
    assert _rel(ud, S.u).max() < 1e-13
    td_o, td_d = S.tstepden(), dg.tstepden()
    assert abs(td_o - td_d) / td_o < 1e-12


@pytest.mark.parametrize("kind", ["tube", "torus_mhd", "torus_rad"])
def test_grid_explicit_parity(oracle, kind):
    from koral_amd.gridrun import DeviceGrid
    P, spec, p = _setup(kind)
    _say("op_explicit parity", kind)
    G, S = _oracle_state(oracle, P, spec, p)
    dt = 1. / S.tstepden()
    dg = DeviceGrid(P, spec)
    dg.upload(S.p, S.u)
    tm_o = S.explicit(0., dt)
    tm_d = dg.explicit(0., dt)
    pd, ud, fd = dg.download()
    r = _rel(ud, S.u)
    _say("  max rel u %.3e" % r.max())
    assert r.max() < 1e-10
    assert abs(tm_o - tm_d) / tm_o < 1e-12


@pytest.mark.parametrize("kind,nsteps", [("tube", 5), ("torus_mhd", 2), ("torus_rad", 2)])
def test_grid_step_parity(oracle, kind, nsteps):
    from koral_amd.gridrun import DeviceGrid
    P, spec, p = _setup(kind)
    _say("RK2IMEX step parity", kind, nsteps)
    G, S = _oracle_state(oracle, P, spec, p)
    dt = 1. / S.tstepden()
    dg = DeviceGrid(P, spec)
    dg.upload(p)
    dg.init()
    assert abs(dg.dt - dt) / dt < 1e-12
    dg.dt = dt
    t = 0.
    inner = (Ellipsis,)
    for k in range(nsteps):
        tm_o, c_o = S.step(t, dt)
        dg.step()
        pd, ud, fd = dg.download()
        rp = _rel(pd, S.p)
        ru = _rel(ud, S.u)
        _say("  step %d: max rel p %.3e u %.3e, counters oracle %s device %s" %
             (k, rp.max(), ru.max(), list(c_o), list(dg.counters)))
        assert rp.max() < 1e-10 * (k + 1)
        assert ru.max() < 1e-10 * (k + 1)
        assert np.array_equal(fd, S.flags)
        assert np.array_equal(np.asarray(c_o), dg.counters)
        assert abs(1. / tm_o - dg.dt) * tm_o < 1e-11
        t += dt
        dt = 1. / tm_o
        dg.dt = dt
