"""GPU parity of the grid operators (ghost fill, explicit operator, u2p sweep
with fixups, implicit sweep, RK2IMEX step) against the CPU oracle on small
grids of the three step workloads, through the C-ABI."""
import os
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


# The device rounds every operation as the oracle does, so the grid
# operators' results are compared bit for bit (NaN equal to NaN); _check
# returns 0 when they are identical and prints the first differing points
# (variable, cell, both values) and returns inf otherwise.
def _check(what, dev, ora, P, tol=None):
    same = (dev == ora) | (np.isnan(dev) & np.isnan(ora))
    if same.all():
        return 0.
    bad = np.argwhere(~same)
    for q in bad[:5]:
        q = tuple(q)
        _say("  %s differs at %s: device %.17e oracle %.17e" % (what, q, dev[q], ora[q]))
    _say("  %s: %d of %d values differ" % (what, len(bad), same.size))
    return float("inf")


# The radiative torus faulted on the device in op_explicit (illegal address
# in the flux/update kernels); its explicit and step cases stay off until the
# cause is found.
RAD_TORUS_OFF = pytest.mark.skip(reason="device fault in op_explicit on the radiative torus under investigation")
KINDS_EXPLICIT = ["tube", "torus_mhd", "torus_mhd_bfloor", pytest.param("torus_rad", marks=RAD_TORUS_OFF)]


def _setup(kind):
    if kind == "tube":
        P = KP.params_tube()
        spec, p = PR.tube_setup(P, nx=64)
    elif kind == "torus_mhd_2d":
        # axisymmetric grid (one x3 cell): faces and the x1-x2 edges only
        P = KP.params_torus_mhd()
        spec, p = PR.torus_setup(P, n=(16, 12, 1))
    elif kind in ("torus_mhd", "torus_mhd_polar", "torus_mhd_bfloor"):
        P = KP.params_torus_mhd()
        if kind == "torus_mhd_polar":
            P.polar_nc = 2
        if kind == "torus_mhd_bfloor":
            # b^2/rho limit low enough that about 1 % of the x1 and x2 face
            # states take the ZAMO mass-injection floor
            P.b2rhoratiomax = 1e-2
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
    assert _check("p", pd, S.p, P) == 0.
    assert _check("u", ud, S.u, P) == 0.
    td_o, td_d = S.tstepden(), dg.tstepden()
    assert td_o == td_d


@pytest.mark.parametrize("kind", KINDS_EXPLICIT)
def test_grid_wavespeeds_fluxes(oracle, kind):
    """Wave speeds and face fluxes (after flux-CT) of one op_explicit."""
    from koral_amd.gridrun import DeviceGrid
    P, spec, p = _setup(kind)
    _say("wavespeeds/fluxes parity", kind)
    G, S = _oracle_state(oracle, P, spec, p)
    dt = 1. / S.tstepden()
    dg = DeviceGrid(P, spec)
    dg.upload(S.p, S.u)
    tm_o, ws_o, fl_o = S.explicit_parts(0., dt)
    dg.explicit(0., dt)
    ws_d, fl_d = dg.work()
    # wave speeds are written on the cells the sweep uses; compare there,
    # the radiation slots only with radiation (the device stores them then)
    used = np.any(ws_o != 0., axis=-1) | np.any(ws_d != 0., axis=-1)
    slots = list(range(18)) if P.rad else [0, 1, 2, 3, 4, 5, 12, 13, 14]
    rw = _check("ws", np.where(used[..., None], ws_d, 0.)[..., slots], ws_o[..., slots], P)
    worst = 0.
    for d in range(3):
        if fl_o[d] is None:
            continue
        worst = max(worst, _check("flux%d" % d, fl_d[d], fl_o[d], P))
    assert rw == 0.
    assert worst == 0.


@pytest.mark.parametrize("kind", KINDS_EXPLICIT)
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
    assert _check("p", pd, S.p, P) == 0.
    assert _check("u", ud, S.u, P) == 0.
    assert tm_o == tm_d


@pytest.mark.parametrize("kind,nsteps", [("tube", 5), ("torus_mhd", 2), ("torus_mhd_polar", 2), ("torus_mhd_bfloor", 2),
                                         ("torus_mhd_2d", 2), pytest.param("torus_rad", 2, marks=RAD_TORUS_OFF)])
def test_grid_step_parity(oracle, kind, nsteps):
    from koral_amd.gridrun import DeviceGrid
    P, spec, p = _setup(kind)
    _say("RK2IMEX step parity", kind, nsteps)
    G, S = _oracle_state(oracle, P, spec, p)
    dt = 1. / S.tstepden()
    dg = DeviceGrid(P, spec)
    dg.upload(p)
    dg.init()
    assert dg.dt == dt
    t = 0.
    for k in range(nsteps):
        tm_o, c_o = S.step(t, dt)
        dg.step()
        pd, ud, fd = dg.download()
        rp = _check("p", pd, S.p, P)
        ru = _check("u", ud, S.u, P)
        _say("  step %d: p %s, u %s, counters oracle %s device %s" %
             (k, "identical" if rp == 0. else "differ", "identical" if ru == 0. else "differ", list(c_o),
              list(dg.counters)))
        nfl = np.argwhere(fd != S.flags)
        for q in nfl[:5]:
            _say("  flag mismatch at", tuple(q), "device", fd[tuple(q)], "oracle", S.flags[tuple(q)])
        assert rp == 0. and ru == 0.
        assert len(nfl) == 0
        assert np.array_equal(np.asarray(c_o), dg.counters)
        assert 1. / tm_o == dg.dt
        t += dt
        dt = 1. / tm_o
        dg.dt = dt


@pytest.mark.parametrize("kind", ["tube", "torus_mhd_polar", "torus_mhd_bfloor"])
def test_grid_run_matches_steps(kind):
    """Steps with the time kept on the device (koral_hip_run_rk2imex) give
    the same state, time, step and counters as steps driven from the host."""
    from koral_amd.gridrun import DeviceGrid
    P, spec, p = _setup(kind)
    _say("device-clock run vs host steps", kind)
    a, b = DeviceGrid(P, spec), DeviceGrid(P, spec)
    for dg in (a, b):
        dg.upload(p)
        dg.init()
    tot = np.zeros(6, np.int64)
    for _ in range(3):
        a.step()
        tot += a.counters
    b.run(3)
    pa, ua, fa = a.download()
    pb, ub, fb = b.download()
    _say("  t %.17g vs %.17g, dt %.17g vs %.17g" % (a.t, b.t, a.dt, b.dt))
    assert a.t == b.t and a.dt == b.dt
    assert np.array_equal(pa, pb) and np.array_equal(ua, ub) and np.array_equal(fa, fb)
    assert np.array_equal(tot, np.asarray(b.counters, np.int64))


@pytest.mark.parametrize("kind", ["tube", "torus_mhd_polar", "torus_mhd_bfloor"])
def test_grid_graph_replay(kind):
    """The step captured as a HIP graph and replayed: three runs of two
    steps each, with eager work and synchronisations between the replays,
    each compared bit-for-bit with the same steps launched eagerly."""
    import torch
    from koral_amd.gridrun import DeviceGrid
    P, spec, p = _setup(kind)
    _say("graph replay vs eager steps", kind)
    st = torch.cuda.Stream()
    a = DeviceGrid(P, spec)
    b = DeviceGrid(P, spec, stream=st.cuda_stream)
    for dg in (a, b):
        dg.upload(p)
        dg.init()
    b.use_graph(True)
    for k in range(3):
        a.run(2)
        b.run(2)
        torch.cuda.synchronize()
        pa, ua, fa = a.download()
        pb, ub, fb = b.download()
        b.tstepden()  # eager work on the graph's stream between replays
        _say("  replay %d: t %.17g vs %.17g" % (k, a.t, b.t))
        assert a.t == b.t and a.dt == b.dt
        assert np.array_equal(pa, pb) and np.array_equal(ua, ub) and np.array_equal(fa, fb)
        assert np.array_equal(np.asarray(a.counters), np.asarray(b.counters))
    b.use_graph(False)


@pytest.mark.parametrize("kind", ["torus_mhd", "torus_rad"])
def test_grid_polar_axis(oracle, kind):
    """do_correct with the polar-axis correction (NCCORRECTPOLAR = 2)."""
    from koral_amd.gridrun import DeviceGrid
    P, spec, p = _setup(kind)
    P.polar_nc = 2
    _say("polar-axis correction parity", kind)
    G, S = _oracle_state(oracle, P, spec, p)
    S.p[..., 3] += 0.01   # non-zero theta velocities, so the scaling matters
    if P.rad:
        S.p[..., P.ee + 2] -= 0.02
    dg = DeviceGrid(P, spec)
    dg.upload(S.p, S.u)
    S.do_correct()
    dg.do_correct()
    pd, ud, fd = dg.download()
    assert np.array_equal(pd, S.p)
    assert np.array_equal(ud, S.u)


def test_field_radiation_explicit_refused():
    """The field + radiation explicit operator is refused (no launch) while
    its device fault is unresolved."""
    from koral_amd.gridrun import DeviceGrid
    from koral_amd._lib import KoralError
    P, spec, p = _setup("torus_rad")
    dg = DeviceGrid(P, spec)
    dg.upload(p)
    with pytest.raises(KoralError, match="disabled"):
        dg.explicit(0., 1e-3)
    with pytest.raises(KoralError, match="disabled"):
        dg.dt = 1e-3
        dg.step()


def _read_restart_files(folder, n, nv, shape):
    """Independent reader of KORAL's restart format (fileop.c:911-982)."""
    import re
    head = open(os.path.join(folder, "res%04d.head" % n)).read()
    m = re.match(r"## +(-?\d+) +(-?\d+) +(\S+) +(-?\d+) +(\d+) +(\d+) +(\d+)\n$", head)
    assert m, head
    nx, ny, nz = shape
    raw = open(os.path.join(folder, "res%04d.dat" % n), "rb").read()
    assert len(raw) == nx * ny * nz * (3 * 4 + nv * 8)
    idx = np.frombuffer(raw, np.int32, 3 * nx * ny * nz).reshape(-1, 3)
    prim = np.frombuffer(raw, np.float64, offset=12 * nx * ny * nz).reshape(-1, nv)
    return m, head, idx, prim


def test_grid_restart_files(oracle, tmp_path):
    """Restart files written from the device follow KORAL's format, read
    back bit-for-bit, and a file written independently (cells in shuffled
    order) is read into the right cells with u = p2u(p)."""
    from koral_amd.gridrun import DeviceGrid
    P, spec, p = _setup("torus_mhd")
    _say("restart files torus_mhd")
    a = DeviceGrid(P, spec)
    a.upload(p)
    a.init()
    a.step()
    a.write_restart(tmp_path, 3, nout2=1, problem=7)
    nx, ny, nz = spec.nx, spec.ny, spec.nz
    m, head, idx, prim = _read_restart_files(tmp_path, 3, P.nv, (nx, ny, nz))
    assert [int(m.group(k)) for k in (1, 2, 4, 5, 6, 7)] == [3, 1, 7, nx, ny, nz]
    assert head == "## %5d %5d %10.6e %5d %5d %5d %5d\n" % (3, 1, a.t, 7, nx, ny, nz)
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    assert np.array_equal(idx, np.stack([ix.ravel(), iy.ravel(), iz.ravel()], 1))
    pa, ua, _ = a.download()
    inner = pa[spec.interior()]                       # [z, y, x, nv]
    assert np.array_equal(prim, inner.transpose(2, 1, 0, 3).reshape(-1, P.nv))
    assert os.readlink(os.path.join(tmp_path, "reslast.dat")) == "res0003.dat"
    # read back (reslast) into a fresh grid
    b = DeviceGrid(P, spec)
    b.upload(p)
    found, nxt = b.read_restart(tmp_path)
    assert found and nxt == 4 and b.t == float("%10.6e" % a.t)
    pb, ub, _ = b.download()
    assert np.array_equal(pb[spec.interior()], inner)
    gc, Gc = spec.cell_tables()
    sl = spec.interior()
    u_ref = oracle.p2u(P, inner.reshape(-1, P.nv), np.ascontiguousarray(gc[sl]).reshape(-1, 20),
                       np.ascontiguousarray(Gc[sl]).reshape(-1, 20))
    assert np.array_equal(ub[sl].reshape(-1, P.nv), u_ref)
    # a file in KORAL's format written here, cells in shuffled order
    rng = np.random.default_rng(5)
    order = rng.permutation(nx * ny * nz)
    pnew = inner.transpose(2, 1, 0, 3).reshape(-1, P.nv) * (1. + 1e-3 * rng.random((nx * ny * nz, 1)))
    d = tmp_path / "other"
    d.mkdir()
    (d / "res0000.head").write_text("## %5d %5d %10.6e %5d %5d %5d %5d\n" % (0, 0, 2.5, 7, nx, ny, nz))
    body = np.ascontiguousarray(np.stack([ix.ravel(), iy.ravel(), iz.ravel()], 1)[order], np.int32).tobytes()
    body += np.ascontiguousarray(pnew[order]).tobytes()
    (d / "res0000.dat").write_bytes(body)
    found, nxt = b.read_restart(d, 0)
    assert found and nxt == 1 and b.t == 2.5
    pb, _, _ = b.download()
    assert np.array_equal(pb[sl].transpose(2, 1, 0, 3).reshape(-1, P.nv), pnew)
    # no file: the reference starts from scratch
    assert b.read_restart(tmp_path / "none", 9) == (False, 0)


def _steps(dg, n):
    tot = np.zeros(6, np.int64)
    for _ in range(n):
        dg.step()
        tot += dg.counters
    pd, ud, fd = dg.download()
    return pd, ud, fd, tot, dg.dt


@pytest.mark.parametrize("kind", ["tube", "torus_mhd_bfloor"])
def test_grid_launch_choices_bit_identical(kind):
    """Every launch choice of the autotuner's space computes the same state:
    reconstruction inside the flux kernel, one lane instead of one wave per
    deferred face or cell, other block sizes and waves in flight; two steps
    of each, bit for bit the state, flags, counters and next dt of the
    default choices."""
    from koral_amd.gridrun import DeviceGrid
    P, spec, p = _setup(kind)
    ref = DeviceGrid(P, spec)
    ref.upload(p)
    ref.init()
    want = _steps(ref, 2)

    def variant(fused, wave, block, floor_nw, u2p_nw):
        dg = DeviceGrid(P, spec)
        t = dg.get_tune()
        for d in range(3):
            t.fused[d] = fused
        t.floor_wave = wave
        t.update_pre = fused
        t.comb_nv = wave
        t.u2p_budget = 3 if fused else 0
        t.u2p_group = 32 if block == 128 else 64
        for k in range(32):
            t.block[k] = block
        t.floor_nw, t.u2p_nw = floor_nw, u2p_nw
        dg.set_tune(t)
        dg.upload(p)
        dg.init()
        return _steps(dg, 2)

    for args in [(1, 1, 256, 1024, 4096), (1, 0, 64, 256, 1024), (0, 0, 128, 4096, 16384), (1, 1, 128, 64, 64)]:
        _say("launch choices fused=%d wave=%d block=%d floor_nw=%d u2p_nw=%d" % args)
        got = variant(*args)
        for a, b in zip(got[:3], want[:3]):
            assert np.array_equal(a, b, equal_nan=True)
        assert np.array_equal(got[3], want[3]) and got[4] == want[4]


def test_grid_autotune_keeps_state_and_results():
    """The autotuner searches on a snapshot of the state: the state after it
    is the state before it, and steps with the chosen launch choices equal
    steps with the defaults bit for bit."""
    from koral_amd.gridrun import DeviceGrid
    P, spec, p = _setup("torus_mhd_bfloor")
    a, b = DeviceGrid(P, spec), DeviceGrid(P, spec)
    for dg in (a, b):
        dg.upload(p)
        dg.init()
    before = a.download()
    rows = a.autotune(steps=2)
    after = a.download()
    for x, y in zip(before, after):
        assert np.array_equal(x, y, equal_nan=True)
    chosen = [r for r in rows if r[0].startswith("chose")]
    _say("autotune: %d trials, %d coordinates; start %.4f ms, end %.4f ms" %
         (len(rows) - len(chosen) - 1, len(chosen), rows[0][2], chosen[-1][2]))
    assert rows[0][0] == "start" and len(chosen) >= 5
    t = a.get_tune()
    _say("  chose fused %s floor_wave %d floor_nw %d u2p_nw %d" % (list(t.fused), t.floor_wave, t.floor_nw, t.u2p_nw))
    ga, gb = _steps(a, 2), _steps(b, 2)
    for x, y in zip(ga[:3], gb[:3]):
        assert np.array_equal(x, y, equal_nan=True)
    assert np.array_equal(ga[3], gb[3]) and ga[4] == gb[4]
