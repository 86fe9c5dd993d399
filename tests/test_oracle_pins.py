"""Checks of the oracle against what the mathematics fixes, for the parts of
the path that no reference vector covers:

- wave speeds (physics.c:10-241, rad.c:2451-2560) in flat space against the
  relativistic velocity addition of the fluid-frame speed
  w^2 = cs^2 + va^2 - cs^2 va^2 (gas) and w^2 = min(1/3, (4/(3 tau))^2)
  (radiation), with the co-going speeds zeroed, and the time-step density;
- cell_fixup (finite.c:4468-4640) against a brute force over the cells in
  the reference's loop order;
- set_bc (finite.c:2752-3700) on analytic fields: face ghosts of the copy,
  periodic and reflect kinds, the edge copies and diagonal averages of the
  corner recipe, the corner blocks;
- the RK2IMEX stage algebra (ko.c:444-570) on a linear stiff source, whose
  amplification factor has a closed form;
- the HLL and LF face fluxes (finite.c:1297-1415) at a contact between two
  flat-space states, from closed-form fluxes, conserved states and wave
  speeds.
"""
import numpy as np
import pytest

from koral_amd import params as KP, problems as PR, grid as GR
from koral_amd import metric as M

GAMMA_RK = 1. - 1. / np.sqrt(2.)


def _mink(P, n, bc=None, lims=((0., 1.), (0., 1.), (0., 1.))):
    ng = GR.ng_for(P.int_order)
    bc = [GR.BC_PERIODIC] * 6 if bc is None else bc
    return GR.GridSpec(M.Coords(M.MINK), n, lims, ng, bc)


def _params(mag, rad, order=1, kappa=0., fluxmethod=1):
    P = KP.base_params(mag=mag, rad=rad, gamma=5. / 3., mass_msun=10.)
    P.int_order = order
    P.fluxmethod = fluxmethod
    if rad:
        P.kappa_model = 0
        P.kappa_a_coef = kappa
        P.kappa_es_coef = 0.
    return P


def _boost_speeds(v, w2, d):
    """Lab-frame left/right speeds along d of a wave of fluid-frame speed
    sqrt(w2) in a fluid of 3-velocity v (flat space)."""
    v2 = (v ** 2).sum(0)
    vd = v[d]
    root = np.sqrt(w2) * np.sqrt((1. - v2) * (1. - v2 * w2 - vd ** 2 * (1. - w2)))
    den = 1. - v2 * w2
    return (vd * (1. - w2) - root) / den, (vd * (1. - w2) + root) / den


def _threevel(ut):
    """VELR u~^i (flat space, no shift) -> 3-velocity v^i and Lorentz factor"""
    gam = np.sqrt(1. + (ut ** 2).sum(0))
    return ut / gam, gam


# ---------------------------------------------------------------------------
# wave speeds
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("mag,rad", [(0, 0), (1, 0), (1, 1)])
def test_flat_space_wave_speeds(oracle, mag, rad):
    P = _params(mag, rad, order=1, kappa=40.)
    n = (6, 5, 4)
    spec = _mink(P, n)
    rng = np.random.default_rng(11 + 2 * mag + rad)
    sh = spec.shape
    p = np.zeros((P.nv,) + sh)
    p[0] = 10 ** rng.uniform(-2, 1, sh)
    p[1] = p[0] * 10 ** rng.uniform(-3, 1, sh)
    dirs = rng.normal(size=(3,) + sh)
    dirs /= np.sqrt((dirs ** 2).sum(0))
    gam = rng.uniform(1., 5., sh)
    p[2:5] = dirs * np.sqrt(gam ** 2 - 1.)
    p[5] = PR._entropy(P, p[0], p[1])
    if mag:
        p[6:9] = rng.normal(size=(3,) + sh) * np.sqrt(p[0])
    if rad:
        ee = P.ee
        p[ee] = p[1] * 10 ** rng.uniform(-2, 2, sh)
        rd = rng.normal(size=(3,) + sh)
        rd /= np.sqrt((rd ** 2).sum(0))
        p[ee + 1:ee + 4] = rd * np.sqrt(rng.uniform(1., 4., sh) ** 2 - 1.)
    S = oracle.GridState(P, oracle.Grid(spec), GR.to_aos(p))
    S.init_u()
    tm, ws, _ = S.explicit_parts(0., 1e-6)
    ii = spec.interior()
    ws = np.moveaxis(ws[ii], -1, 0)  # [18, z, y, x]
    q = p[(slice(None),) + ii]
    v, gm = _threevel(q[2:5])
    pre = (P.gamma - 1.) * q[1]
    cs2 = P.gamma * pre / (q[0] + q[1] + pre)
    va2 = 0.
    if mag:
        B = q[6:9]
        bsq = (B ** 2).sum(0) / gm ** 2 + ((B * v).sum(0)) ** 2
        va2 = bsq / (bsq + q[0] + P.gamma * q[1])
    w2 = cs2 + va2 - cs2 * va2
    dx = 1. / np.array(n)
    tsd = 0.
    amax = []
    for d in range(3):
        lo, hi = _boost_speeds(v, w2, d)
        al, ar = np.minimum(lo, 0.), np.maximum(hi, 0.)
        np.testing.assert_allclose(ws[2 * d], al, rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(ws[2 * d + 1], ar, rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(ws[12 + d], np.maximum(-al, ar), rtol=1e-12, atol=1e-15)
        a = np.maximum(-al, ar)
        if rad:
            vr, _ = _threevel(q[P.ee + 1:P.ee + 4])
            tau = P.kappa_a_coef * q[0] * dx[d]
            rv2 = np.minimum(1. / 3., (4. / 3. / tau) ** 2)
            lo, hi = _boost_speeds(vr, rv2, d)
            rl, rr = np.minimum(lo, 0.), np.maximum(hi, 0.)
            np.testing.assert_allclose(ws[6 + 2 * d], rl, rtol=1e-12, atol=1e-15)
            np.testing.assert_allclose(ws[7 + 2 * d], rr, rtol=1e-12, atol=1e-15)
            np.testing.assert_allclose(ws[15 + d], np.maximum(-rl, rr), rtol=1e-12, atol=1e-15)
            assert (rv2 < 1. / 3.).mean() > 0.2 and (rv2 == 1. / 3.).mean() > 0.2  # both regimes
            a = np.maximum(a, np.maximum(-rl, rr))
        amax.append(a)
        tsd = tsd + a / dx[d]
    np.testing.assert_allclose(tm, (tsd / P.tsteplim).max(), rtol=1e-12)
    assert (ws[:6] == 0.).mean() < 0.5  # both signs occur


# ---------------------------------------------------------------------------
# cell_fixup
# ---------------------------------------------------------------------------
def _brute_fixup(P, p, flags, ftype, n):
    """The rule of cell_fixup stated on its own: cells flagged ftype are
    visited in the order (ix outer, iz inner); a visited cell's flag is
    cleared; its good neighbours are the in-domain face neighbours whose flag
    is clear at that moment; with at least 1, 2 or 3 of them in 1-D, 2-D,
    3-D the fixed variables become the mean over them of the old primitives.
    Returns the new primitives of the flagged cells (iv < NVMHD written;
    dict cell -> vector) and the cleared flags."""
    nx, ny, nz = n
    fl = flags.copy()
    out = {}
    need = 3 if (ny > 1 and nz > 1) else (2 if (ny > 1 or nz > 1) else 1)
    ee = 6 + 3 * P.mag
    fixed = [iv for iv in range(P.nv) if iv != 0 and (
        (ftype == 2 and iv < 6) or (ftype == 3 and P.rad and ee <= iv <= ee + 3) or
        (ftype == 1 and (iv < 6 or (P.rad and ee <= iv <= ee + 3))))]
    for ix in range(nx):
        for iy in range(ny):
            for iz in range(nz):
                if fl[iz, iy, ix] == 0:
                    continue
                fl[iz, iy, ix] = 0
                good = []
                for dx_, dy_, dz_ in ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)):
                    jx, jy, jz = ix + dx_, iy + dy_, iz + dz_
                    if 0 <= jx < nx and 0 <= jy < ny and 0 <= jz < nz and fl[jz, jy, jx] == 0:
                        good.append(p[jz, jy, jx])
                if len(good) < need:
                    continue
                new = p[iz, iy, ix].copy()
                mean = np.zeros(P.nv)
                for g in good:
                    mean += g
                new[fixed] = mean[fixed] / len(good)
                out[(iz, iy, ix)] = new
    return out, fl


@pytest.mark.parametrize("dims,ftype,rad", [((7, 6, 5), 2, 0), ((7, 6, 5), 1, 1), ((9, 8, 1), 2, 1),
                                             ((17, 1, 1), 1, 1), ((7, 6, 5), 3, 1)])
def test_cell_fixup_brute_force(oracle, dims, ftype, rad):
    P = _params(1, rad)
    spec = _mink(P, dims, bc=[GR.BC_COPY] * 6)
    rng = np.random.default_rng(sum(dims) + 10 * ftype + rad)
    sh = spec.shape
    p = np.zeros((P.nv,) + sh)
    p[0] = rng.uniform(0.5, 2., sh)
    p[1] = rng.uniform(0.05, 0.2, sh)
    p[2:5] = rng.uniform(-0.2, 0.2, (3,) + sh)
    p[5] = PR._entropy(P, p[0], p[1])
    p[6:9] = rng.uniform(-0.1, 0.1, (3,) + sh)
    if rad:
        p[P.ee] = rng.uniform(0.01, 0.1, sh)
        p[P.ee + 1:P.ee + 4] = rng.uniform(-0.2, 0.2, (3,) + sh)
    p = GR.to_aos(p)
    S = oracle.GridState(P, oracle.Grid(spec), p)
    S.init_u()
    ii = spec.interior()
    fl = np.zeros(sh, np.int32)
    fl[ii] = (rng.uniform(size=fl[ii].shape) < 0.3).astype(np.int32)
    S.flags[..., ftype] = fl
    p0, u0 = S.p.copy(), S.u.copy()
    S.cell_fixup(ftype)
    want, _ = _brute_fixup(P, p0[ii], fl[ii], ftype, dims)
    assert not S.flags[..., ftype].any()  # every flag cleared
    nvmhd = 6 + 3 * P.mag
    gc, Gc = spec.cell_tables()
    interior = S.p[ii]
    changed = 0
    for iz in range(interior.shape[0]):
        for iy in range(interior.shape[1]):
            for ix in range(interior.shape[2]):
                got = interior[iz, iy, ix]
                old = p0[ii][iz, iy, ix]
                if (iz, iy, ix) in want:
                    exp = want[(iz, iy, ix)]
                    np.testing.assert_allclose(got[:nvmhd], exp[:nvmhd], rtol=1e-15, atol=0)
                    assert np.array_equal(got[nvmhd:], old[nvmhd:])  # only iv < NVMHD written back
                    changed += int(not np.array_equal(got, old))
                else:
                    assert np.array_equal(got, old)
    # u = p2u(p) on the fixed cells (MHD part)
    cells = np.array(list(want.keys()), dtype=int).reshape(-1, 3)
    if len(cells):
        zz, yy, xx = cells.T + np.array([spec.ngz, spec.ngy, spec.ngx])[:, None]
        uu = oracle.p2u(P, S.p[zz, yy, xx], np.ascontiguousarray(gc[zz, yy, xx]), np.ascontiguousarray(Gc[zz, yy, xx]))
        assert np.array_equal(S.u[zz, yy, xx][:, :nvmhd], uu[:, :nvmhd])
    if ftype != 3:
        assert changed > 0
    ghost = np.ones(sh, bool)
    ghost[ii] = False
    assert np.array_equal(S.p[ghost], p0[ghost]) and np.array_equal(S.u[ghost], u0[ghost])


# ---------------------------------------------------------------------------
# set_bc
# ---------------------------------------------------------------------------
def _analytic(P, spec, kinds):
    """Primitive field over the ghost-inclusive box whose face ghosts follow
    from each boundary kind in closed form: along a periodic dimension the
    field has period n; along a reflecting one the scalars are even and the
    normal components odd about both boundary faces (cos / sin of
    pi m (i + 1/2) / n); along a copy dimension any smooth profile (the ghosts
    take the edge cell's value).  Returns (p_at, field) where p_at(ix, iy, iz)
    is the closed-form value of a face ghost or interior cell."""
    n = (spec.nx, spec.ny, spec.nz)
    nv = P.nv
    rng = np.random.default_rng(3)
    amp = rng.uniform(0.01, 0.05, (nv, 3))
    ph = rng.uniform(0, 2 * np.pi, (nv, 3))
    base = np.array([1.0, 0.1, 0.05, -0.03, 0.02, 0.0, 0.05, 0.04, -0.03] + ([0.05, 0.02, -0.01, 0.03] if P.rad else []))
    vec = {2: 0, 3: 1, 4: 2, 6: 0, 7: 1, 8: 2}
    if P.rad:
        vec.update({P.ee + 1: 0, P.ee + 2: 1, P.ee + 3: 2})

    def one(iv, d, i):
        k = kinds[d]
        if k == GR.BC_PERIODIC:
            return amp[iv, d] * np.cos(2 * np.pi * i / n[d] + ph[iv, d])
        if k == GR.BC_REFLECT:
            m = 1 + iv % 3
            return amp[iv, d] * np.cos(np.pi * m * (i + 0.5) / n[d])
        ic = min(max(i, 0), n[d] - 1)
        return amp[iv, d] * np.cos(0.7 * ic + ph[iv, d])

    def p_at(ix, iy, iz):
        out = np.empty(nv)
        for iv in range(nv):
            if iv == 5:
                continue
            f = base[iv]
            for d, i in enumerate((ix, iy, iz)):
                if n[d] > 1:
                    f = f + one(iv, d, i)
            # the normal component across a reflecting dimension is odd there:
            # the whole value times sin(pi m (i + 1/2) / n)
            d = vec.get(iv)
            if d is not None and kinds[d] == GR.BC_REFLECT:
                i = (ix, iy, iz)[d]
                f = f * np.sin(np.pi * (1 + iv % 3) * (i + 0.5) / n[d])
            out[iv] = f
        out[5] = PR._entropy(P, out[0], out[1])
        return out
    return p_at


@pytest.mark.parametrize("kinds", [(0, 0, 0), (1, 1, 1), (0, 2, 1)])
def test_set_bc_analytic_fields(oracle, kinds):
    P = _params(1, 1, order=4)
    n = (6, 5, 4)
    bc = [k for k in kinds for _ in (0, 1)]
    spec = _mink(P, n, bc=bc)
    NG = spec.ng
    p_at = _analytic(P, spec, kinds)
    sh = spec.shape
    p = np.full(sh + (P.nv,), np.nan)
    for iz in range(n[2]):
        for iy in range(n[1]):
            for ix in range(n[0]):
                p[iz + NG, iy + NG, ix + NG] = p_at(ix, iy, iz)
    S = oracle.GridState(P, oracle.Grid(spec), p)
    S.u[...] = np.nan
    S.set_bc()

    def got(ix, iy, iz):
        return S.p[iz + NG, iy + NG, ix + NG]

    def face(ix, iy, iz):
        return p_at(ix, iy, iz)
    rng_all = [range(-NG, n[d] + NG) for d in range(3)]
    outside = lambda i, d: i < 0 or i >= n[d]  # noqa: E731
    seen = 0
    for iz in rng_all[2]:
        for iy in rng_all[1]:
            for ix in rng_all[0]:
                q = (ix, iy, iz)
                out = [d for d in range(3) if outside(q[d], d)]
                if len(out) == 1:
                    np.testing.assert_allclose(got(*q), face(*q), rtol=1e-14, atol=1e-16)
                elif len(out) == 3:
                    c = tuple(0 if q[d] < 0 else n[d] - 1 for d in range(3))
                    assert np.array_equal(got(*q), got(*c))
                elif len(out) == 2:
                    a, b = out
                    ahi, bhi = q[a] >= n[a], q[b] >= n[b]
                    ea, ia, ea2 = (n[a], n[a] - 1, n[a] + 1) if ahi else (-1, 0, -2)
                    eb, ib, eb2 = (n[b], n[b] - 1, n[b] + 1) if bhi else (-1, 0, -2)

                    def at(va, vb):
                        r = list(q)
                        r[a], r[b] = va, vb
                        return face(*r)
                    if q[b] == eb and q[a] != ea:
                        exp = at(q[a], ib)
                    elif q[a] == ea and q[b] != eb:
                        exp = at(ia, q[b])
                    elif q[a] == ea and q[b] == eb:
                        exp = .5 * (at(ea, ib) + at(ia, eb))
                    elif q[a] == ea2 and q[b] == eb2:
                        exp = .5 * (at(ea2, ib) + at(ia, eb2))
                    else:
                        assert np.all(np.isnan(got(*q)))  # the recipe leaves these cells
                        continue
                    np.testing.assert_allclose(got(*q), exp, rtol=1e-14, atol=1e-16)
                    seen += 1
    assert seen > 100
    # u = p2u(p) wherever a ghost was set
    m = ~np.isnan(S.p[..., 0])
    gc, Gc = spec.cell_tables()
    uu = oracle.p2u(P, S.p[m], np.ascontiguousarray(gc[m]), np.ascontiguousarray(Gc[m]))
    ghost = m.copy()
    ghost[spec.interior()] = False
    assert np.array_equal(S.u[ghost], uu[ghost[m]])


# ---------------------------------------------------------------------------
# RK2IMEX stage algebra on a linear stiff source
# ---------------------------------------------------------------------------
def _rk2imex_factor(z):
    g = GAMMA_RK
    u1 = 1. / (1. - g * z)
    mid = 1. + (1. - 2. * g) * z * u1
    u2 = mid / (1. - g * z)
    return 1. + 0.5 * z * (u1 + u2)


@pytest.mark.parametrize("dtfac", [0.3, 2.0, 12.0])
def test_rk2imex_amplification_on_linear_source(oracle, dtfac):
    """A uniform radiating gas at rest in a periodic flat box, radiation
    slightly out of equilibrium: the explicit operator does nothing (uniform
    fluxes, no metric sources) and the implicit source is linear in the
    deviation x = R^t_t - R^t_t(eq) to second order.  The backward-Euler
    solve over dt gives the rate z = lambda dt (x1 = x0 / (1 - z)); one
    RK2IMEX step must multiply x by the scheme's amplification factor
    R(z) = 1 + z/2 (U1 + U2), U1 = 1/(1 - g z), U2 = (1 + (1 - 2g) z U1)/(1 - g z),
    g = 1 - 1/sqrt(2)."""
    P = _params(0, 1, order=1, kappa=1e3)
    spec = _mink(P, (4, 4, 4))
    ee = P.ee
    # gas at 1e7 K, radiation near its LTE energy 4 sigma T^4 (calc_LTE_EfromT)
    rho, Tg = 1e-3, 1e7
    ugas = Tg * P.k_boltz * rho / ((P.gamma - 1.) * P.mu_gas * P.m_proton)
    Elte = 4. * P.sigma_rad * Tg ** 4

    def state(dev):
        sh = spec.shape
        p = np.zeros((P.nv,) + sh)
        p[0], p[1] = rho, ugas
        p[5] = PR._entropy(P, p[0], p[1])
        p[ee] = Elte * (1. + dev)
        return GR.to_aos(p)

    def run(dev, what, dt):
        S = oracle.GridState(P, oracle.Grid(spec), state(dev))
        S.init_u()
        S.set_bc()
        if what == "implicit":
            c = S.implicit(0., dt)
            assert c[2] == 0  # every solve succeeded
        elif what == "step":
            _, c = S.step(0., dt)
            assert c[2] == 0
        return S.u[spec.interior()][..., ee]

    # the rate from a short backward-Euler solve, the equilibrium from a very long one
    dev = 1e-3
    ueq = run(dev, "implicit", 1e30).mean()
    x0 = run(dev, None, 0.).mean() - ueq
    assert abs(x0) > 1e-4 * abs(ueq)
    h = 1e-4
    x1 = run(dev, "implicit", h).mean() - ueq
    lam = (1. - x0 / x1) / h
    assert lam < 0.
    dt = dtfac / -lam
    xs = run(dev, "step", dt).mean() - ueq
    want = _rk2imex_factor(lam * dt)
    got = xs / x0
    assert abs(got - want) < 1e-6 * max(abs(want), 0.05), (got, want)
    # a dropped term would be visible: the factor without the (1 - 2g) coupling
    wrong = 1. + 0.5 * lam * dt * (1. / (1. - GAMMA_RK * lam * dt)) * 2.
    assert abs(wrong - want) > 1e-3


# ---------------------------------------------------------------------------
# HLL / LF at a contact of two flat-space states
# ---------------------------------------------------------------------------
def _closed_form_hydro(P, q):
    """Conserved state U and x-flux F of the gas (+ M1 radiation) in flat
    space from primitives q [nv] (VELR velocities, gdet = 1)."""
    rho, uu = q[0], q[1]
    ut = q[2:5]
    gam = np.sqrt(1. + ut @ ut)
    ucon = np.array([gam, *ut])
    ucov = np.array([-gam, *ut])
    pre = (P.gamma - 1.) * uu
    w = rho + uu + pre

    def T(mu, nu):
        return w * ucon[mu] * ucov[nu] + pre * (1. if mu == nu else 0.)
    U = np.zeros(P.nv)
    F = np.zeros(P.nv)
    U[0], F[0] = rho * ucon[0], rho * ucon[1]
    U[1], F[1] = T(0, 0) + rho * ucon[0], T(1, 0) + rho * ucon[1]
    for i in range(3):
        U[2 + i], F[2 + i] = T(0, 1 + i), T(1, 1 + i)
    U[5], F[5] = q[5] * ucon[0], q[5] * ucon[1]
    if P.rad:
        ee = P.ee
        E = q[ee]
        rt = q[ee + 1:ee + 4]
        rg = np.sqrt(1. + rt @ rt)
        rcon = np.array([rg, *rt])
        rcov = np.array([-rg, *rt])

        def R(mu, nu):
            return 4. / 3. * E * rcon[mu] * rcov[nu] + 1. / 3. * E * (1. if mu == nu else 0.)
        for k in range(4):
            U[ee + k], F[ee + k] = R(0, k), R(1, k)
    return U, F


@pytest.mark.parametrize("fluxmethod", [0, 1])
@pytest.mark.parametrize("rad", [0, 1])
def test_face_flux_at_contact(oracle, fluxmethod, rad):
    P = _params(0, rad, order=1, kappa=0., fluxmethod=fluxmethod)
    n = 16
    spec = _mink(P, (n, 1, 1), bc=[GR.BC_COPY] * 6)
    sh = spec.shape
    NG = spec.ng
    p = np.zeros((P.nv,) + sh)
    left = np.arange(sh[2]) < NG + n // 2
    for iv, (a, b) in enumerate([(1.0, 0.3), (0.2, 0.05), (0.4, -0.6), (0.1, 0.3), (-0.2, 0.0)]):
        p[iv] = np.where(left, a, b)
    p[5] = PR._entropy(P, p[0], p[1])
    if rad:
        for k, (a, b) in enumerate([(0.05, 0.5), (0.3, -0.2), (0.0, 0.1), (0.1, -0.1)]):
            p[P.ee + k] = np.where(left, a, b)
    S = oracle.GridState(P, oracle.Grid(spec), GR.to_aos(p))
    S.init_u()
    S.set_bc()
    _, ws, fl = S.explicit_parts(0., 1e-5)
    qL, qR = GR.to_aos(p)[0, 0, NG + n // 2 - 1], GR.to_aos(p)[0, 0, NG + n // 2]
    UL, FL = _closed_form_hydro(P, qL)
    UR, FR = _closed_form_hydro(P, qR)
    # wave speeds of the two cells, closed form (tau = 0: radiation w^2 = 1/3)
    pre = lambda q: (P.gamma - 1.) * q[1]  # noqa: E731
    sp = {}
    for side, q in (("L", qL), ("R", qR)):
        v, _ = _threevel(q[2:5])
        cs2 = P.gamma * pre(q) / (q[0] + q[1] + pre(q))
        lo, hi = _boost_speeds(v, cs2, 0)
        g = (min(lo, 0.), max(hi, 0.))
        r = g
        if rad:
            vr, _ = _threevel(q[P.ee + 1:P.ee + 4])
            lo, hi = _boost_speeds(vr, 1. / 3., 0)
            r = (min(lo, 0.), max(hi, 0.))
        sp[side] = (g, r)
    face = fl[0][0, 0, NG + n // 2]
    nvmhd = 6
    for iv in range(P.nv):
        k = 1 if (rad and iv >= nvmhd) else 0
        al = min(sp["L"][k][0], sp["R"][k][0])
        ar = max(sp["L"][k][1], sp["R"][k][1])
        ag = max(-sp["L"][k][0], sp["L"][k][1], -sp["R"][k][0], sp["R"][k][1])
        if fluxmethod == 0:
            want = 0.5 * (FR[iv] + FL[iv] - ag * (UR[iv] - UL[iv]))
        else:
            want = (-al * FR[iv] + ar * FL[iv] + al * ar * (UR[iv] - UL[iv])) / (ar - al)
        scale = abs(FL[iv]) + abs(FR[iv]) + abs(ag * (UR[iv] - UL[iv])) + 1e-300
        assert abs(face[iv] - want) < 1e-13 * scale, (iv, face[iv], want)
    # a face between equal states carries the physical flux
    far = fl[0][0, 0, NG + 2]
    np.testing.assert_allclose(far, FL, rtol=1e-13, atol=1e-15)
