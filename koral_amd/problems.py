"""Seeded synthetic inputs for the benchmark workloads (host side).

Nothing here is on the hot path: these build primitive states, metric tables
and time steps that the device kernels then consume.
"""
import numpy as np

from . import metric as M
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
