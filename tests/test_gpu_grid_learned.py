"""Learned Newton starts in the grid's implicit operator (no reference
counterpart): each interior cell solved from the starts the networks give,
against the oracle's solve from the same starts, bit for bit; with every
cell gated off the step is the plain step."""
import sys

import numpy as np
import pytest

from koral_amd import params as KP, problems as PR, grid as GR

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _same(a, b):
    return bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


def _setup(kind):
    if kind == "tube":
        P = KP.params_tube()
        spec, p = PR.tube_setup(P, nx=64)
    else:
        P = KP.params_torus_rad()
        spec, p = PR.torus_setup(P, n=(12, 12, 8))
    return P, spec, GR.to_aos(p)


def _old_energy(S, spec, P):
    q = S.p.reshape(spec.sz, spec.sy, spec.sx, P.nv)[spec.ngz:spec.ngz + spec.nz, spec.ngy:spec.ngy + spec.ny,
                                                   spec.ngx:spec.ngx + spec.nx]
    return q[..., 1].reshape(-1)


def _nets(seed):
    """A gas network whose output is constant (zero last-layer weights): the
    start of every cell is 1.01 times its old gas energy with its old
    velocities, a Newton path other than the plain one that converges for
    most cells; random radiation and gate networks."""
    from koral_amd import learn as LN
    th = LN.init_params(seed)
    th[-(LN.NOUT * LN.NH + LN.NOUT):] = 0.
    th[-LN.NOUT] = np.log(1.01)
    return (LN.Net(LN.GAS, theta=th), LN.Net(LN.RAD, seed=seed + 1), LN.Net(LN.GATE, seed=seed + 2))


@pytest.mark.parametrize("kind", ["tube", "torus_rad"])
@pytest.mark.parametrize("gated", [False, True])
def test_grid_learned_implicit_matches_oracle(oracle, kind, gated):
    from koral_amd.gridrun import DeviceGrid
    P, spec, p = _setup(kind)
    _say("learned op_implicit parity", kind, "gated" if gated else "")
    S = oracle.GridState(P, oracle.Grid(spec), p)
    S.init_u()
    S.set_bc()
    dt = 1. / S.tstepden()
    gas, rad, gate = _nets(5)
    dg = DeviceGrid(P, spec)
    dg.upload(S.p, S.u)
    dg.set_learned(gas, rad, gate if gated else None, thr=0., maxiter=8)
    dg.implicit(0., dt)
    gxg, gxr = dg.download_guess()
    nan = np.isnan(gxg[0])
    _say("  cells %d, gated off %d" % (gxg.shape[1], int(nan.sum())))
    if not gated:
        assert not nan.any()
    if not gated:
        assert np.allclose(gxg[0], 1.01 * _old_energy(S, spec, P), rtol=1e-2)
    plain = oracle.GridState(P, oracle.Grid(spec), S.p, S.u, S.flags)
    S.implicit(0., dt, gxg, gxr, maxiter=8)
    pd, ud, fd = dg.download()
    assert _same(pd, S.p) and _same(ud, S.u) and np.array_equal(fd, S.flags)
    if not gated and kind == "torus_rad":
        # the learned tries changed the Newton path of some cells (the
        # tube's cells end on the plain bits from these starts)
        plain.implicit(0., dt)
        assert not _same(plain.p, S.p)


def test_grid_learned_step_all_gated_is_plain(oracle):
    """Every cell scored below the threshold starts the plain chain, so the
    RK2IMEX steps equal the plain ones bit for bit."""
    from koral_amd.gridrun import DeviceGrid
    P, spec, p = _setup("tube")
    S = oracle.GridState(P, oracle.Grid(spec), p)
    S.init_u()
    S.set_bc()
    dt = 1. / S.tstepden()
    gas, rad, gate = _nets(7)
    dg = DeviceGrid(P, spec)
    dg.upload(p)
    dg.init()
    dg.dt = dt
    dg.set_learned(gas, rad, gate, thr=float("inf"), maxiter=8)
    t = 0.
    for _ in range(3):
        tm, c = S.step(t, dt)
        dg.step()
        t += dt
        dt = 1. / tm
    gxg, _ = dg.download_guess()
    assert np.isnan(gxg[0]).all()
    pd, ud, fd = dg.download()
    assert _same(pd, S.p) and _same(ud, S.u) and np.array_equal(fd, S.flags)
