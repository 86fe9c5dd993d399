"""Grid-level oracle checks that the mathematics fixes: a uniform flow stays
uniform, the conservative update telescopes on a periodic box, flux-CT keeps
the corner-centred divergence of B (magn.c:593-630) at round-off, MP5
reproduces quartics and PPM linear data at the faces (finite.c:90-289)."""
import numpy as np
import pytest

from koral_amd import params as KP, problems as PR, grid as GR
from koral_amd import metric as M


def _mink_spec(P, n, periodic=True, lims=((0., 1.), (0., 1.), (0., 1.))):
    ng = GR.ng_for(P.int_order)
    bc = [GR.BC_PERIODIC if periodic else GR.BC_COPY] * 6
    return GR.GridSpec(M.Coords(M.MINK), n, lims, ng, bc)


def _state(P, spec, rho, u, v, B=None, E=None, F=None):
    sh = spec.shape
    p = np.zeros((P.nv,) + sh)
    p[0] = np.broadcast_to(rho, sh)
    p[1] = np.broadcast_to(u, sh)
    for k in range(3):
        p[2 + k] = np.broadcast_to(v[k], sh)
    p[5] = PR._entropy(P, p[0], p[1])
    if P.mag:
        for k in range(3):
            p[6 + k] = np.broadcast_to(B[k], sh)
    if P.rad:
        p[P.ee] = np.broadcast_to(E, sh)
        for k in range(3):
            p[P.ee + 1 + k] = np.broadcast_to(F[k], sh)
    return GR.to_aos(p)


def _params(mag, rad, order):
    P = KP.base_params(mag=mag, rad=rad, gamma=5. / 3., mass_msun=10.)
    P.int_order = order
    if rad:
        P.kappa_model = 0
        P.kappa_a_coef = 1e-3
        P.kappa_es_coef = 0.
    return P


def _interior(spec, a):
    return a[spec.interior()]


@pytest.mark.parametrize("mag,rad", [(1, 0), (0, 1), (1, 1)])
def test_uniform_flow_stays_uniform(oracle, mag, rad):
    P = _params(mag, rad, 4)
    spec = _mink_spec(P, (8, 8, 8))
    p = _state(P, spec, 1.0, 0.1, (0.1, -0.05, 0.2), B=(0.1, 0.2, -0.1), E=0.05, F=(0.1, -0.05, 0.2))
    S = oracle.GridState(P, oracle.Grid(spec), p)
    S.init_u()
    S.set_bc()
    u0 = S.u.copy()
    dt = 0.5 / S.tstepden()
    S.explicit(0., dt)
    rel = np.abs(S.u - u0) / (np.abs(u0) + 1e-300)
    assert _interior(spec, rel).max() < 1e-13


def test_periodic_update_telescopes(oracle):
    """sum over a periodic Minkowski box of the conserved gas variables is
    kept by op_explicit (no metric sources in flat space).  The field is left
    out: the EMFs on the domain edges use fluxes through corner ghosts, which
    KORAL fills by its copy/average recipe (finite.c:3335-3700), not as
    periodic images."""
    P = _params(1, 0, 4)
    spec = _mink_spec(P, (8, 8, 8))
    X, Y, Z = spec.centres()
    k = 2 * np.pi
    rho = 1. + 0.2 * np.sin(k * X) * np.cos(k * Y)
    u = 0.1 * (1. + 0.3 * np.cos(k * Z))
    v = (0.1 * np.sin(k * Y), 0.05 * np.cos(k * Z), -0.1 * np.sin(k * X))
    B = (0.05 + 0.02 * np.sin(k * Z), 0.03 * np.cos(k * X), 0.04 + 0.01 * np.sin(k * Y))
    S = oracle.GridState(P, oracle.Grid(spec), _state(P, spec, rho, u, v, B=B))
    S.init_u()
    S.set_bc()
    u0 = _interior(spec, S.u)[..., :6].copy()
    S.explicit(0., 0.5 / S.tstepden())
    u1 = _interior(spec, S.u)[..., :6]
    before, after = u0.sum(axis=(0, 1, 2)), u1.sum(axis=(0, 1, 2))
    scale = np.abs(u1).sum(axis=(0, 1, 2))
    assert np.all(np.abs(after - before) <= 1e-13 * scale)
    assert np.abs(u1 - u0).max() > 1e-6 * np.abs(u0).max()  # cells change, sums do not


def _corner_divb(spec, u):
    """magn.c:593-630 (NZ > 1) on the conserved field u[..., 6:9] (GDETIN=1
    stores sqrt(-g) B^i), at corners whose eight cells are interior."""
    a = u[spec.interior()]
    b1, b2, b3 = a[..., 6], a[..., 7], a[..., 8]
    dx = 1. / spec.nx

    def s8(f, sx, sy, sz):
        # f(ix,iy,iz) +/- neighbours: signs per offset in x, y, z
        out = 0.
        for ox in (0, 1):
            for oy in (0, 1):
                for oz in (0, 1):
                    sign = (sx[ox]) * (sy[oy]) * (sz[oz])
                    out = out + sign * f[1 - oz:f.shape[0] - oz, 1 - oy:f.shape[1] - oy, 1 - ox:f.shape[2] - ox]
        return out

    d1 = s8(b1, (-1, 1), (1, 1), (1, 1)) / (4 * dx)
    d2 = s8(b2, (1, 1), (-1, 1), (1, 1)) / (4 * dx)
    d3 = s8(b3, (1, 1), (1, 1), (-1, 1)) / (4 * dx)
    return d1 + d2 + d3


def test_flux_ct_keeps_corner_divergence(oracle):
    P = _params(1, 0, 4)
    spec = _mink_spec(P, (10, 10, 10))
    X, Y, Z = spec.centres()
    k = 2 * np.pi
    rng = np.random.default_rng(5)
    rho = 1. + 0.2 * rng.uniform(-1, 1, X.shape)
    u = 0.1 * (1. + 0.3 * rng.uniform(-1, 1, X.shape))
    v = tuple(0.1 * rng.uniform(-1, 1, X.shape) for _ in range(3))
    B = (0.05 + 0.02 * np.sin(k * Z), 0.03 * np.cos(k * X), 0.04 + 0.01 * np.sin(k * Y))
    S = oracle.GridState(P, oracle.Grid(spec), _state(P, spec, rho, u, v, B=B))
    S.init_u()
    S.set_bc()
    d0 = _corner_divb(spec, S.u)
    S.explicit(0., 0.5 / S.tstepden())
    d1 = _corner_divb(spec, S.u)
    bscale = np.abs(S.u[..., 6:9]).max() * spec.nx
    assert np.abs(d1 - d0).max() < 1e-13 * bscale
    # the field itself did change
    assert np.abs(_interior(spec, S.u)[..., 6:9] - 0).max() > 0


def _face_density(oracle, P, spec, rho, vx):
    p = _state(P, spec, rho, 1e-3, (vx, 0., 0.), B=(0., 0., 0.), E=1e-8, F=(vx, 0., 0.))
    S = oracle.GridState(P, oracle.Grid(spec), p)
    S.init_u()
    S.set_bc()
    _, ws, fl = S.explicit_parts(0., 1e-4)
    ux = vx  # VELR in flat space: u^x = v~^x
    return fl[0][0, 0, :, 0] / ux  # rho at x faces (gdet = 1)


def test_mp5_reproduces_quartic_at_faces(oracle):
    P = _params(0, 0, 4)
    n = 32
    spec = _mink_spec(P, (n, 1, 1), periodic=False, lims=((0., 1.), (0., 1.), (0., 1.)))
    X, _, _ = spec.centres()
    poly = lambda x: 1.0 + 0.5 * x + 0.3 * x ** 2 + 0.2 * x ** 3 + 0.1 * x ** 4  # increasing on [0,1]
    rho_face = _face_density(oracle, P, spec, poly(X), 0.01)
    # faces 3..n-3 have their six-cell stencil inside the domain
    xb = spec.xb[0][spec.ng + 3: spec.ng + n - 2]
    rel = np.abs(rho_face[spec.ngx + 3: spec.ngx + n - 2] - poly(xb)) / poly(xb)
    assert rel.max() < 1e-13


def test_ppm_reproduces_linear_at_faces(oracle):
    P = _params(0, 0, 2)
    n = 32
    spec = _mink_spec(P, (n, 1, 1), periodic=False)
    X, _, _ = spec.centres()
    lin = lambda x: 1.0 + 0.7 * x
    rho_face = _face_density(oracle, P, spec, lin(X), 0.01)
    xb = spec.xb[0][spec.ng + 3: spec.ng + n - 2]
    rel = np.abs(rho_face[spec.ngx + 3: spec.ngx + n - 2] - lin(xb)) / lin(xb)
    assert rel.max() < 1e-13


@pytest.mark.parametrize("rad", [0, 1])
def test_polar_axis_correction(oracle, rad):
    """do_correct with CORRECT_POLARAXIS (finite.c:4924-5060) on a torus in
    MKS1 coordinates, whose x2 = theta faces are uniform: the nc cells next to
    each polar face take rho, u, S, v^r, v^phi (and E, F^r, F^phi) of cell nc
    (NY-1-nc), their theta components scaled by (ic + 1/2) / (nc + 1/2); the
    field is left alone; u = p2u(p) there; every other cell is untouched."""
    P = KP.params_torus_rad() if rad else KP.params_torus_mhd()
    P.polar_nc = 2
    spec, p = PR.torus_setup(P, n=(8, 12, 4))
    S = oracle.GridState(P, oracle.Grid(spec), GR.to_aos(p))
    S.init_u()
    S.set_bc()
    # make the theta components non-trivial so the scaling is visible
    S.p[..., 3] += 0.01
    if rad:
        S.p[..., P.ee + 2] -= 0.02
    p0, u0 = S.p.copy(), S.u.copy()
    S.do_correct()
    nc, ny, ngy = P.polar_nc, spec.ny, spec.ngy
    zs, xs = spec.interior()[0], spec.interior()[2]
    gas = [0, 1, 5, 2, 4] + ([P.ee, P.ee + 1, P.ee + 3] if rad else [])
    theta = [3] + ([P.ee + 2] if rad else [])
    touched = np.zeros(S.p.shape[:-1], bool)
    for side in (0, 1):
        src = ngy + (nc if side == 0 else ny - 1 - nc)
        for ic in range(nc):
            iy = ngy + (ic if side == 0 else ny - 1 - ic)
            touched[zs, iy, xs] = True
            a, b = S.p[zs, iy, xs], p0[zs, src, xs]
            assert np.array_equal(a[..., gas], b[..., gas])
            fac = (ic + 0.5) / (nc + 0.5)
            np.testing.assert_allclose(a[..., theta], fac * b[..., theta], rtol=1e-12, atol=0)
            assert np.array_equal(a[..., 6:9], p0[zs, iy, xs][..., 6:9])  # field kept
            assert np.array_equal(S.u[zs, iy, xs][..., 6:9], u0[zs, iy, xs][..., 6:9])
    # u = p2u(p) on the corrected cells (independent per-point p2u)
    gc, Gc = spec.cell_tables()
    uu = oracle.p2u(P, S.p[touched], np.ascontiguousarray(gc[touched]), np.ascontiguousarray(Gc[touched]))
    keep = [iv for iv in range(P.nv) if not 6 <= iv <= 8]
    assert np.array_equal(S.u[touched][:, keep], uu[:, keep])
    assert np.array_equal(S.p[~touched], p0[~touched]) and np.array_equal(S.u[~touched], u0[~touched])


@pytest.mark.parametrize("kind", ["torus_mhd", "torus_rad"])
def test_threaded_step_is_bit_identical(oracle, kind):
    """The oracle's per-cell loops (wave speeds, sweeps, update, u2p, implicit,
    stage algebra) on several threads give the serial result bit for bit:
    two RK2IMEX steps, state, flags, counters and the next time step."""
    if kind == "torus_mhd":
        P = KP.params_torus_mhd()
        spec, p = PR.torus_setup(P, n=(16, 16, 8))
    else:
        P = KP.params_torus_rad()
        spec, p = PR.torus_setup(P, n=(12, 12, 8))
    runs = []
    try:
        for threads in (1, 4):
            oracle.grid_set_threads(threads)
            S = oracle.GridState(P, oracle.Grid(spec), GR.to_aos(p))
            S.init_u()
            S.set_bc()
            dt = 1. / S.tstepden()
            tm1, c1 = S.step(0., dt)
            tm2, c2 = S.step(dt, 1. / tm1)
            runs.append((S.p.copy(), S.u.copy(), S.flags.copy(), (tm1, tm2), (list(c1), list(c2))))
    finally:
        oracle.grid_set_threads(1)
    a, b = runs
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)
    assert a[3] == b[3] and a[4] == b[4]


def test_implicit_with_gated_starts_is_plain(oracle):
    """The grid implicit operator with learned starts: cells whose gas start
    is NaN (left out by the gate) take the unchanged chain, so all-NaN starts
    give the plain operator bit for bit; finite starts are used (the state
    differs from the plain one somewhere)."""
    P = KP.params_torus_rad()
    spec, p = PR.torus_setup(P, n=(8, 8, 4))
    p = GR.to_aos(p)
    n = spec.nx * spec.ny * spec.nz
    out = []
    for gx in (None, np.full((4, n), np.nan), "old"):
        S = oracle.GridState(P, oracle.Grid(spec), p)
        S.init_u()
        S.set_bc()
        dt = 1. / S.tstepden()
        if isinstance(gx, str):
            # starts at 1.01 times the old gas energy, the old velocities: a
            # different Newton path from the plain one
            q = S.p.reshape(spec.sz, spec.sy, spec.sx, P.nv)[spec.ngz:spec.ngz + spec.nz, spec.ngy:spec.ngy + spec.ny,
                                                           spec.ngx:spec.ngx + spec.nx]
            gx = np.ascontiguousarray(np.moveaxis(q[..., 1:5], -1, 0).reshape(4, n))
            gx[0] *= 1.01
        S.implicit(0., dt, gx, None, maxiter=16)
        out.append((S.p.copy(), S.u.copy(), S.flags.copy()))
    same = lambda a, b: np.array_equal(a, b, equal_nan=True)
    assert all(same(x, y) for x, y in zip(out[0], out[1]))
    assert not same(out[0][0], out[2][0])
