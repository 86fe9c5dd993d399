"""Seeded synthetic inputs for the benchmark workloads (host side).

Nothing here is on the hot path: these build primitive states, metric tables
and time steps that the device kernels then consume.
"""
import numpy as np

from . import metric as M
from . import params as KP
from .params import u_from_Trho, T_from_urho


def _ortho_relvel(g20, gamma, dirs):
    """VELR u~^i with Lorentz factor gamma relative to the normal observer in
    direction `dirs` (unit 3-vectors of an orthonormal spatial triad):
    u~ = sqrt(gamma^2-1) L^{-T} n with g_ij = L L^T, so g_ij u~^i u~^j =
    gamma^2 - 1."""
    n = g20.shape[0]
    gij = g20.reshape(n, 4, 5)[:, 1:4, 1:4]
    L = np.linalg.cholesky(gij)
    y = np.linalg.solve(np.swapaxes(L, 1, 2), dirs[..., None])[..., 0]
    return np.sqrt(gamma * gamma - 1.)[:, None] * y


def _random_dirs(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1)[:, None]


def implicit_batch_states(P, n, seed=0, ks_fraction=0.5, spin=0.9, gamma_max=10.0):
    """Independent cells for the batched implicit solve.

    rho log-uniform [1e-6, 1e2]; T_gas log-uniform [1e4, 1e9] K; E_rad/u_gas
    log-uniform [1e-4, 1e4]; gas and radiation-frame velocities in random
    directions with Lorentz factors uniform in [1, gamma_max]; dt chosen so
    that dt*kappa_abs is log-uniform in [1e-3, 1e6]; metric Minkowski or a
    Kerr-Schild point (a=spin) at r uniform in [3, 50], cos(theta) uniform in
    [-0.95, 0.95].  Returns (pp[n,nv], g20[n,20], G20[n,20], dt[n]).
    """
    rng = np.random.default_rng(seed)
    nv, ee = P.nv, P.ee
    lu = lambda a, b: np.exp(rng.uniform(np.log(a), np.log(b), n))
    rho = lu(1e-6, 1e2)
    T = lu(1e4, 1e9)
    u = u_from_Trho(P, T, rho)
    erat = lu(1e-4, 1e4)
    # metric
    ks = rng.uniform(size=n) < ks_fraction
    r = rng.uniform(3., 50., n)
    th = np.arccos(rng.uniform(-0.95, 0.95, n))
    g20 = np.zeros((n, 20))
    G20 = np.zeros((n, 20))
    gm, Gm = M.point_tables(M.Coords(M.MINK), np.zeros(n), np.zeros(n))
    gk, Gk = M.point_tables(M.Coords(M.KS, spin=spin), r, th)
    g20[:] = np.where(ks[:, None], gk, gm)
    G20[:] = np.where(ks[:, None], Gk, Gm)
    pp = np.zeros((n, nv))
    pp[:, 0] = rho
    pp[:, 1] = u
    gg = rng.uniform(1., gamma_max, n)
    pp[:, 2:5] = _ortho_relvel(g20, gg, _random_dirs(rng, n))
    if P.mag:
        pp[:, 6:9] = 0.
    pp[:, ee] = erat * u
    gr = rng.uniform(1., gamma_max, n)
    pp[:, ee + 1:ee + 4] = _ortho_relvel(g20, gr, _random_dirs(rng, n))
    # entropy consistent with (rho, u) (calc_Sfromu)
    n_ = 1.0 / (P.gamma - 1.)
    pp[:, 5] = rho * np.log((P.gamma - 1.) * u) * n_ - rho * np.log(rho) * (n_ + 1.0)
    # time step from the absorption opacity at the initial state
    kap = kappa_abs_np(P, rho, u)
    tau = lu(1e-3, 1e6)
    dt = tau / kap
    return pp, g20, G20, dt


def kappa_abs_np(P, rho, u):
    if P.kappa_model == 0:
        return P.kappa_a_coef * rho
    T = T_from_urho(P, u, rho)
    rhocgs = rho / P.ggg * P.ccc * P.ccc / P.masscm / P.masscm
    k = P.kappa_a_coef * rhocgs / T / T / T / np.sqrt(T)
    return k / P.ggg * P.ccc * P.ccc / P.masscm * rho


def random_mhd_states(P, n, seed=0, spin=0.9, gamma_max=5.0, beta_range=(0.3, 1e3), rad=True):
    """Generic magnetised (and radiating) states at random Kerr-Schild and
    Minkowski points, for inversion round trips."""
    rng = np.random.default_rng(seed)
    nv, ee = P.nv, P.ee
    lu = lambda a, b: np.exp(rng.uniform(np.log(a), np.log(b), n))
    ks = rng.uniform(size=n) < 0.7
    r = rng.uniform(2.0, 50., n)
    th = np.arccos(rng.uniform(-0.98, 0.98, n))
    gm, Gm = M.point_tables(M.Coords(M.MINK), np.zeros(n), np.zeros(n))
    gk, Gk = M.point_tables(M.Coords(M.KS, spin=spin), r, th)
    g20 = np.where(ks[:, None], gk, gm)
    G20 = np.where(ks[:, None], Gk, Gm)
    pp = np.zeros((n, nv))
    rho = lu(1e-4, 1e1)
    u = rho * lu(1e-5, 1e1)
    pp[:, 0], pp[:, 1] = rho, u
    pp[:, 2:5] = _ortho_relvel(g20, rng.uniform(1., gamma_max, n), _random_dirs(rng, n))
    n_ = 1.0 / (P.gamma - 1.)
    pp[:, 5] = rho * np.log((P.gamma - 1.) * u) * n_ - rho * np.log(rho) * (n_ + 1.0)
    if P.mag:
        beta = lu(*beta_range)
        pmag = (P.gamma - 1.) * u / beta
        # B^i from an orthonormal direction with |B| ~ sqrt(2 pmag) (approximate b^2)
        B = _ortho_relvel(g20, np.sqrt(1. + 2. * pmag), _random_dirs(rng, n))
        pp[:, 6:9] = B
    if P.rad:
        pp[:, ee] = u * lu(1e-3, 1e3)
        pp[:, ee + 1:ee + 4] = _ortho_relvel(g20, rng.uniform(1., gamma_max, n), _random_dirs(rng, n))
    return pp, g20, G20


def u2p_guess_states(P, n, seed=0):
    """Points for the u2p W inversion from a stale guess: true magnetised
    states (random_mhd_states), the conserved state is taken from them by
    the caller, and the guess is the true state perturbed as a timestep
    would leave it: rho and u scaled by exp(s z) and each velocity shifted
    by s z (|u~| + 0.1), with s log-uniform in [1e-4, 0.3] per point.
    Returns (pp_true, pp_guess, g20, G20)."""
    rng = np.random.default_rng(seed + 7919)
    pp, g20, G20 = random_mhd_states(P, n, seed=seed, rad=False)
    s = np.exp(rng.uniform(np.log(1e-4), np.log(0.3), n))
    guess = pp.copy()
    guess[:, 0] *= np.exp(s * rng.standard_normal(n))
    guess[:, 1] *= np.exp(s * rng.standard_normal(n))
    vmag = np.sqrt((pp[:, 2:5] ** 2).sum(1))
    guess[:, 2:5] += (s * (vmag + 0.1))[:, None] * rng.standard_normal((n, 3))
    n_ = 1.0 / (P.gamma - 1.)
    guess[:, 5] = guess[:, 0] * np.log((P.gamma - 1.) * guess[:, 1]) * n_ - guess[:, 0] * np.log(guess[:, 0]) * (n_ + 1.0)
    return pp, guess, g20, G20


# ---------------------------------------------------------------------------
# grid problems
# ---------------------------------------------------------------------------
def _relr_to_4vel(v, g, G):
    """VELR -> (u^mu, u_mu) for arrays v[..., 3], g/G[..., 4, 4] (relele.c)."""
    gs = g[..., 1:4, 1:4]
    qsq = np.einsum("...i,...ij,...j->...", v, gs, v)
    gamma2 = 1. + qsq
    alpha2 = -1. / G[..., 0, 0]
    ucon = np.zeros(v.shape[:-1] + (4,))
    ucon[..., 0] = np.sqrt(gamma2 / alpha2)
    ucon[..., 1:] = v - (np.sqrt(alpha2 * gamma2))[..., None] * G[..., 0, 1:4]
    ucov = np.einsum("...ij,...j->...i", g, ucon)
    return ucon, ucov


def _bsq(B, ucon, ucov, g):
    bt = np.einsum("...i,...i->...", B, ucov[..., 1:])
    bcon = np.zeros(ucon.shape)
    bcon[..., 0] = bt
    bcon[..., 1:] = (B + bt[..., None] * ucon[..., 1:]) / ucon[..., :1]
    bcov = np.einsum("...ij,...j->...i", g, bcon)
    return (bcon * bcov).sum(-1)


def _entropy(P, rho, u):
    n_ = 1.0 / (P.gamma - 1.)
    return rho * np.log(((P.gamma - 1.) * u) ** n_ / rho ** (n_ + 1.))


def tube_setup(P, nx=512, seed=0, xlim=(-15., 15.)):
    """1-D flat radiation shock tube.  Left (x<0): rho=1, u=1e-3, E=1e-2,
    v^x=0.1; right: rho=2.4, u=6e-3, E=1e-1, v^x=0.04 (3-velocities); the
    radiation frame moves with the gas; each state gets a seeded uniform
    1e-3 relative perturbation per cell.  COPY boundaries."""
    from .grid import GridSpec, ng_for, BC_COPY
    ng = ng_for(P.int_order)
    spec = GridSpec(M.Coords(M.MINK), (nx, 1, 1), (xlim, (-1., 1.), (-1., 1.)), ng, [BC_COPY] * 6)
    X, _, _ = spec.centres()
    rng = np.random.default_rng(seed)
    left = X < 0.5 * (xlim[0] + xlim[1])
    sh = X.shape
    pert = lambda: 1. + 1e-3 * rng.uniform(-1., 1., sh)
    rho = np.where(left, 1.0, 2.4) * pert()
    u = np.where(left, 1.e-3, 6.e-3) * pert()
    E = np.where(left, 1.e-2, 1.e-1) * pert()
    v = np.where(left, 0.1, 0.04) * pert()
    p = np.zeros((P.nv,) + sh)
    p[0], p[1] = rho, u
    ut = v / np.sqrt(1. - v * v)  # VEL3 -> VELR in flat space
    p[2] = ut
    p[5] = _entropy(P, rho, u)
    p[P.ee] = E
    p[P.ee + 1] = ut
    return spec, p


def torus_setup(P, n=(128, 128, 64), spin=0.9, rin=6.0, rmax=12.0, r_grid=(1.2, 60.), th0=0.02 * np.pi,
                beta=100., seed=0, pert=0.02, rad_frac=0.1, rho_max=None):
    """Magnetised constant-angular-momentum torus around a Kerr black hole in
    MKS1 coordinates (r = exp x1, theta = x2, phi = x3).

    l = l_K(rmax) (Keplerian at the pressure maximum); enthalpy from
    W = ln(-u_t) with the surface through (rin, pi/2); polytrope p = K rho^Gamma
    normalised to rho = 1 at the density maximum.  Atmosphere rho = 1e-4 r^-1.5,
    u = 1e-6 r^-2.5, normal-observer velocity (relele.c:598-625, atmtype 0).
    Density and energy of torus and atmosphere are then multiplied by rho_max
    (code units); ideal MHD is invariant under that scaling (B follows from
    beta), radiation coupling is not.  Default: 1 without radiation; with
    radiation the density maximum is 3e-2 g/cm^3, the scale of KORAL's
    radiative torus set-up (PROBLEMS/KATOTORUS/define.h:219, KT_RHO0).  Seeded uniform +-pert relative noise on u.
    A_phi = max(rho/rho_max - 0.2, 0) at cell corners, B^i = curl A / sqrt(-g),
    scaled so that max(p_gas)/max(p_mag) = beta.  With radiation: E_rf =
    rad_frac * u, comoving with the gas.  Boundaries: copy in r, reflection at
    the polar axis, periodic in phi.
    """
    from .grid import GridSpec, ng_for, BC_COPY, BC_PERIODIC, BC_REFLECT
    if rho_max is None:
        rho_max = KP.rho_cgs2gu(3e-2, P.masscm) if P.rad else 1.0
    a = spin
    nx, ny, nz = n
    ng = ng_for(P.int_order)
    co = M.Coords(M.MKS1, spin=a)
    lims = ((np.log(r_grid[0]), np.log(r_grid[1])), (th0, np.pi - th0), (0., 2. * np.pi))
    spec = GridSpec(co, (nx, ny, nz), lims, ng,
                    [BC_COPY, BC_COPY, BC_REFLECT, BC_REFLECT, BC_PERIODIC, BC_PERIODIC])
    X1, X2, _ = spec.centres()
    r, th = np.exp(X1), X2

    def bl(r, th):
        s2 = np.sin(th) ** 2
        Sig = r * r + a * a * np.cos(th) ** 2
        Del = r * r - 2 * r + a * a
        gtt = -(1. - 2. * r / Sig)
        gtp = -2. * a * r * s2 / Sig
        gpp = s2 * (r * r + a * a + 2. * a * a * r * s2 / Sig)
        return gtt, gtp, gpp

    def ut_of(r, th, ell):
        gtt, gtp, gpp = bl(r, th)
        den = gpp + 2. * ell * gtp + ell * ell * gtt
        num = gtp * gtp - gtt * gpp
        with np.errstate(invalid="ignore", divide="ignore"):
            return -np.sqrt(num / den), den

    # Keplerian angular momentum in the equatorial plane (prograde)
    Om = 1. / (rmax ** 1.5 + a)
    gtt, gtp, gpp = bl(rmax, np.pi / 2)
    ell = -(gtp + Om * gpp) / (gtt + Om * gtp)
    ut_in, _ = ut_of(rin, np.pi / 2, ell)
    W_in = np.log(-ut_in)

    def torus_h(r, th):
        ut, den = ut_of(r, th, ell)
        with np.errstate(invalid="ignore"):
            W = np.log(-ut)
        h = np.exp(W_in - W)
        ok = (den > 0) & np.isfinite(h) & (h > 1.) & (r > rin * 0.999)
        return np.where(ok, h, 1.), ok

    G_ = P.gamma
    h, inside = torus_h(r, th)
    hmax = h.max()
    # rho = ((h-1)(G-1)/(K G))^(1/(G-1)); choose K so rho_max = 1
    K = (hmax - 1.) * (G_ - 1.) / G_
    rho_t = np.where(inside, ((h - 1.) * (G_ - 1.) / (K * G_)) ** (1. / (G_ - 1.)), 0.)
    u_t = K * rho_t ** G_ / (G_ - 1.)
    rho_a = 1e-4 * r ** -1.5
    u_a = 1e-6 * r ** -2.5
    use_t = inside & (rho_t > rho_a)
    rng = np.random.default_rng(seed)
    rho = np.where(use_t, rho_t, rho_a) * rho_max
    u = np.where(use_t, u_t, u_a) * (1. + pert * rng.uniform(-1., 1., r.shape)) * rho_max

    gc, Gc = spec.cell_tables()
    g = gc.reshape(gc.shape[:-1] + (4, 5))[..., :4]
    G = Gc.reshape(Gc.shape[:-1] + (4, 5))[..., :4]
    # torus 4-velocity: u^r = u^th = 0, u_t from W, u_phi = -l u_t (BL = KS
    # components when u^r = 0); relative velocity u~^i = u^i - u^t g^{ti}/g^{tt}
    ut_cov, _ = ut_of(r, th, ell)
    uph_cov = -ell * ut_cov
    gtt_, gtp_, gpp_ = bl(r, th)
    det = gtt_ * gpp_ - gtp_ * gtp_
    Ut = (gpp_ * ut_cov - gtp_ * uph_cov) / det
    Uph = (-gtp_ * ut_cov + gtt_ * uph_cov) / det
    vrel = np.zeros(r.shape + (3,))
    vrel[..., 2] = np.where(use_t, Uph - Ut * G[..., 0, 3] / G[..., 0, 0], 0.)
    vrel[..., 0] = np.where(use_t, -Ut * G[..., 0, 1] / G[..., 0, 0], 0.)
    p = np.zeros((P.nv,) + r.shape)
    p[0], p[1] = rho, u
    p[2:5] = np.moveaxis(vrel, -1, 0)
    p[5] = _entropy(P, rho, u)

    if P.mag:
        # A_phi at the cell corners (x1 faces, x2 faces), curl at centres
        xb1 = spec.g.xb[0][spec.ng - spec.ngx: spec.ng + nx + spec.ngx + 1]
        xb2 = spec.g.xb[1][spec.ng - spec.ngy: spec.ng + ny + spec.ngy + 1]
        C1, C2 = np.meshgrid(xb1, xb2)  # [y, x]
        hc, insc = torus_h(np.exp(C1), C2)
        rhoc = np.where(insc, ((hc - 1.) * (G_ - 1.) / (K * G_)) ** (1. / (G_ - 1.)), 0.)
        A = np.maximum(rhoc - 0.2, 0.)
        dx1 = np.diff(xb1)[None, :]
        dx2 = np.diff(xb2)[:, None]
        Ay = 0.5 * (A[:, 1:] + A[:, :-1])   # averaged along x1: [y+1, x]
        Ax = 0.5 * (A[1:, :] + A[:-1, :])   # averaged along x2: [y, x+1]
        gd = spec.tables["gc"][19]           # sqrt(-g) at centres [y, x]
        B1 = (Ay[1:, :] - Ay[:-1, :]) / dx2 / gd
        B2 = -(Ax[:, 1:] - Ax[:, :-1]) / dx1 / gd
        Bv = np.zeros(r.shape + (3,))
        Bv[..., 0] = B1[None]
        Bv[..., 1] = B2[None]
        ucon, ucov = _relr_to_4vel(vrel, g, G)
        bsq = _bsq(Bv, ucon, ucov, g)
        pg = (G_ - 1.) * u
        inner = spec.interior()
        fac = np.sqrt(pg[inner].max() / (0.5 * bsq[inner].max()) / beta)
        p[6:9] = np.moveaxis(Bv * fac, -1, 0)
    if P.rad:
        p[P.ee] = rad_frac * u
        p[P.ee + 1:P.ee + 4] = p[2:5]
    return spec, p
