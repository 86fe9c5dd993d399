"""Pins the CPU oracle's per-point physics against closed forms, invariants
and limits.  The reference cannot be compiled here (GSL headers absent), and
it ships no golden vectors for this path, so these checks are what fixes the
oracle: each would fail on a dropped term, a wrong sign or a transposed
operand in the restated formula."""
import numpy as np
import pytest

from koral_amd import params as KP, problems as PR, metric as M


def mink(n):
    return M.point_tables(M.Coords(M.MINK), np.zeros(n), np.zeros(n))


def ks(r, th, a=0.9):
    return M.point_tables(M.Coords(M.KS, spin=a), np.asarray(r, float), np.asarray(th, float))


def test_params_layout_matches(oracle):
    """oracle_params and koral_params (ctypes) share one layout."""
    import ctypes
    P = KP.KoralParams
    L = oracle.lib()
    assert L.oracle_params_size() == ctypes.sizeof(P)
    names = ["gamma", "kappa_model", "rhofloor", "u2pconv", "radimpmaxiter", "radimpconvrelentrerr",
             "int_order", "doradimpfixups"]
    for k, nm in enumerate(names):
        assert L.oracle_params_offset(k) == getattr(P, nm).offset, nm


def test_metric_tables_invariants():
    """g G = 1, sqrt(-g) = sqrt(-det g), gttpert = g_tt + 1 for KS and MKS1."""
    for c, x1, x2 in [(M.Coords(M.KS, 0.9), np.array([2.5, 7., 40.]), np.array([0.3, 1.5, 2.9])),
                      (M.Coords(M.MKS1, 0.9), np.log(np.array([2.5, 7., 40.])), np.array([0.3, 1.5, 2.9]))]:
        g20, G20 = M.point_tables(c, x1, x2)
        g = g20.reshape(-1, 4, 5)[:, :, :4]
        G = G20.reshape(-1, 4, 5)[:, :, :4]
        for k in range(len(x1)):
            assert np.allclose(g[k] @ G[k], np.eye(4), atol=1e-13)
            assert np.isclose(np.sqrt(-np.linalg.det(g[k])), g20[k, 19], rtol=1e-13)
            assert np.isclose(g[k, 0, 0] + 1., G20[k, 19], rtol=1e-12)


def test_christoffel_vs_finite_difference():
    """Complex-step Christoffels agree with a 4th-order central difference of
    the metric (independent route) in MKS1."""
    c = M.Coords(M.MKS1, 0.9)
    x1, x2 = np.log(6.3), 1.1
    kr = M.christoffel(c, np.array([x1]), np.array([x2]))[0]
    h = 1e-3
    def g(a, b):
        return np.real(c.g(np.array([a]), np.array([b])))[0]
    dg = np.zeros((4, 4, 4))
    dg[1] = (-g(x1 + 2 * h, x2) + 8 * g(x1 + h, x2) - 8 * g(x1 - h, x2) + g(x1 - 2 * h, x2)) / (12 * h)
    dg[2] = (-g(x1, x2 + 2 * h) + 8 * g(x1, x2 + h) - 8 * g(x1, x2 - h) + g(x1, x2 - 2 * h)) / (12 * h)
    Gi = np.linalg.inv(g(x1, x2))
    ref = np.zeros((4, 4, 4))
    for i in range(4):
        for j in range(4):
            for k in range(4):
                ref[i, j, k] = 0.5 * sum(Gi[i, l] * (dg[j, l, k] + dg[k, l, j] - dg[l, j, k]) for l in range(4))
    assert np.allclose(kr, ref, atol=1e-9 * np.abs(ref).max())


def test_p2u_minkowski_closed_form(oracle):
    """Ideal MHD in flat space: U = (rho gamma, -(w+b^2) gamma^2 + p_tot + rho gamma + b^t b^t ...).
    Checked here without field: T^t_t + rho u^t = -(w gamma^2 - p) + rho gamma,
    T^t_i = w gamma^2 v_i, S u^t; radiation: R^t_t = -(4/3 E g^2 - E/3)."""
    P = KP.params_implicit_batch()
    n = 50
    rng = np.random.default_rng(0)
    g20, G20 = mink(n)
    pp = np.zeros((n, P.nv))
    pp[:, 0] = rng.uniform(0.1, 2., n)
    pp[:, 1] = rng.uniform(0.01, 1., n)
    pp[:, 2:5] = rng.uniform(-2, 2, (n, 3))
    pp[:, 5] = rng.uniform(-1, 1, n)
    pp[:, P.ee] = rng.uniform(0.1, 2, n)
    pp[:, P.ee + 1:P.ee + 4] = rng.uniform(-2, 2, (n, 3))
    uu = oracle.p2u(P, pp, g20, G20)
    gam = np.sqrt(1. + (pp[:, 2:5] ** 2).sum(1))
    v = pp[:, 2:5] / gam[:, None]
    rho, u = pp[:, 0], pp[:, 1]
    p = (P.gamma - 1.) * u
    w = rho + u + p
    assert np.allclose(uu[:, 0], rho * gam, rtol=1e-14)
    assert np.allclose(uu[:, 1], -(w * gam ** 2 - p) + rho * gam, rtol=1e-11, atol=1e-13)
    assert np.allclose(uu[:, 2:5], (w * gam ** 2)[:, None] * v, rtol=1e-13)
    assert np.allclose(uu[:, 5], pp[:, 5] * gam, rtol=1e-14)
    E = pp[:, P.ee]
    gr = np.sqrt(1. + (pp[:, P.ee + 1:P.ee + 4] ** 2).sum(1))
    vr = pp[:, P.ee + 1:P.ee + 4] / gr[:, None]
    assert np.allclose(uu[:, P.ee], -(4. / 3. * E * gr ** 2 - E / 3.), rtol=1e-13)
    assert np.allclose(uu[:, P.ee + 1:P.ee + 4], (4. / 3. * E * gr ** 2)[:, None] * vr, rtol=1e-13)


def test_p2u_field_closed_form(oracle):
    """Static magnetised gas in flat space: b^mu = (0, B), U^t_t = -(u + B^2/2)."""
    P = KP.params_torus_mhd()
    n = 10
    g20, G20 = mink(n)
    rng = np.random.default_rng(1)
    pp = np.zeros((n, P.nv))
    pp[:, 0] = 1.
    pp[:, 1] = rng.uniform(0.1, 1, n)
    pp[:, 6:9] = rng.uniform(-1, 1, (n, 3))
    uu = oracle.p2u(P, pp, g20, G20)
    b2 = (pp[:, 6:9] ** 2).sum(1)
    assert np.allclose(uu[:, 1], -(pp[:, 1] + 0.5 * b2), rtol=1e-13)
    assert np.allclose(uu[:, 6:9], pp[:, 6:9], rtol=0, atol=0)
    assert np.allclose(uu[:, 2:5], 0., atol=1e-15)


def _mild_states(P, n, seed):
    pp, g20, G20 = PR.random_mhd_states(P, n, seed=seed, gamma_max=2.5, beta_range=(1., 100.))
    rng = np.random.default_rng(seed + 100)
    pp[:, 1] = pp[:, 0] * rng.uniform(0.05, 2., n)
    pp[:, 5] = oracle_S(P, pp[:, 0], pp[:, 1])
    return pp, g20, G20


def oracle_S(P, rho, u):
    n = 1.0 / (P.gamma - 1.)
    return rho * np.log(((P.gamma - 1.) * u) ** n / rho ** (n + 1.))


@pytest.mark.parametrize("mag,rad", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_u2p_roundtrip(oracle, mag, rad):
    """u2p(p2u(p)) = p for well-conditioned states in KS and flat metrics,
    starting from a perturbed guess (hot solver, no flags raised)."""
    P = KP.base_params(mag, rad, 5. / 3.)
    pp, g20, G20 = _mild_states(P, 400, seed=10 + 2 * mag + rad)
    uu = oracle.p2u(P, pp, g20, G20)
    guess = pp.copy()
    guess[:, 1] *= 1.5
    guess[:, 2:5] *= 0.8
    if rad:
        guess[:, P.ee] *= 3.
    uo, po, ret, fl = oracle.u2p(P, uu, guess, g20, G20)
    assert (ret == 0).all()
    assert (fl == 0).all()
    assert np.allclose(po[:, :5], pp[:, :5], rtol=2e-8, atol=0)
    assert np.allclose(po[:, 5], pp[:, 5], rtol=1e-7, atol=1e-10)
    if rad:
        assert np.allclose(po[:, P.ee:P.ee + 4], pp[:, P.ee:P.ee + 4], rtol=1e-9, atol=1e-12)
    # conserved follow the primitives exactly
    assert np.array_equal(uo, oracle.p2u(P, po, g20, G20))


def test_u2p_entropy_solver_roundtrip(oracle):
    """The entropy equation alone recovers (rho, u, v) from (D, S u^t, T^t_i)."""
    P = KP.base_params(1, 0, 13. / 9.)
    pp, g20, G20 = _mild_states(P, 300, seed=3)
    uu = oracle.p2u(P, pp, g20, G20)
    po, ret = oracle.u2p_solver(P, uu, pp * [1.] + 0., g20, G20, etype=1)
    assert (ret == 0).all()
    assert np.allclose(po[:, :5], pp[:, :5], rtol=1e-7)


def test_u2p_negative_density_requests_fixup(oracle):
    """u2p.c:228-238: negative D sets rho to RHOFLOOR and requests a fixup."""
    P = KP.base_params(1, 0, 5. / 3.)
    pp, g20, G20 = _mild_states(P, 4, seed=4)
    uu = oracle.p2u(P, pp, g20, G20)
    uu[:, 0] *= -1.
    _, po, ret, fl = oracle.u2p(P, uu, pp, g20, G20)
    assert (fl[:, 2] == 1).all()  # HDFIXUPFLAG


def test_u2p_rad_roundtrip_and_cap(oracle):
    P = KP.base_params(0, 1, 5. / 3.)
    pp, g20, G20 = _mild_states(P, 300, seed=5)
    uu = oracle.p2u(P, pp, g20, G20)
    po, ret, corr = oracle.u2p_rad(P, uu, np.zeros_like(pp), g20, G20)
    assert (ret == 0).all() and (corr == 0).all()
    ee = P.ee
    assert np.allclose(po[:, ee:ee + 4], pp[:, ee:ee + 4], rtol=1e-9, atol=1e-12)
    # radiation boosted beyond GAMMAMAXRAD: inversion caps gamma and flags it
    pp2 = pp.copy()
    g = g20.reshape(-1, 4, 5)[:, 1:4, 1:4]
    urf = pp2[:, ee + 1:ee + 4]
    q = np.einsum("ni,nij,nj->n", urf, g, urf)
    scale = np.sqrt((300. ** 2 - 1.) / q)
    pp2[:, ee + 1:ee + 4] = urf * scale[:, None]
    uu2 = oracle.p2u(P, pp2, g20, G20)
    po2, ret2, corr2 = oracle.u2p_rad(P, uu2, np.zeros_like(pp2), g20, G20)
    assert (corr2 == 1).all()
    q2 = np.einsum("ni,nij,nj->n", po2[:, ee + 1:ee + 4], g, po2[:, ee + 1:ee + 4])
    gam2 = np.sqrt(1. + q2)
    ok = ret2 == 0
    assert ok.mean() > 0.9
    assert np.allclose(gam2[ok], P.gammamaxrad, rtol=1e-6)


def test_inverse44(oracle):
    rng = np.random.default_rng(7)
    for _ in range(50):
        a = rng.normal(size=(4, 4))
        r, ia = oracle.inverse44(a)
        assert r == 0
        assert np.allclose(ia, np.linalg.inv(a), rtol=1e-10, atol=1e-12)
    r, _ = oracle.inverse44(np.zeros((4, 4)))
    assert r == -1


def test_Gi_static_closed_form(oracle):
    """Flat, static gas and radiation: G^t = kappa (E - 4 pi B), G^i = 0,
    and the fluid-frame G^t equals the same."""
    P = KP.params_implicit_batch()
    n = 20
    g20, G20 = mink(n)
    rng = np.random.default_rng(8)
    pp = np.zeros((n, P.nv))
    pp[:, 0] = 10 ** rng.uniform(-4, 1, n)
    pp[:, 1] = KP.u_from_Trho(P, 10 ** rng.uniform(5, 8, n), pp[:, 0])
    pp[:, P.ee] = pp[:, 1] * 10 ** rng.uniform(-2, 2, n)
    G = oracle.Gi(P, pp, g20, G20, 1)
    Gff = oracle.Gi(P, pp, g20, G20, 0)
    T = KP.T_from_urho(P, pp[:, 1], pp[:, 0])
    B = P.sigma_rad * T ** 4 / 3.141592654
    kap = PR.kappa_abs_np(P, pp[:, 0], pp[:, 1])
    expect = kap * (pp[:, P.ee] - 4. * 3.141592654 * B)
    assert np.allclose(G[:, 0], expect, rtol=1e-12)
    assert np.allclose(G[:, 1:], 0., atol=0)
    assert np.allclose(Gff[:, 0], expect, rtol=1e-12)


def test_Gi_vanishes_in_lte_comoving(oracle):
    """G^mu = 0 when the radiation is blackbody at T_gas and comoving with the gas."""
    P = KP.params_implicit_batch()
    n = 200
    pp, g20, G20, _ = PR.implicit_batch_states(P, n, seed=9)
    T = KP.T_from_urho(P, pp[:, 1], pp[:, 0])
    pp[:, P.ee] = 4. * P.sigma_rad * T ** 4
    pp[:, P.ee + 1:P.ee + 4] = pp[:, 2:5]
    G = oracle.Gi(P, pp, g20, G20, 1)
    kap = PR.kappa_abs_np(P, pp[:, 0], pp[:, 1]) + P.kappa_es_coef * pp[:, 0]
    scale = kap * pp[:, P.ee] * (1 + (pp[:, 2:5] ** 2).sum(1))
    assert (np.abs(G).max(axis=1) / scale < 1e-12).all()


def test_ff_Rtt_comoving(oracle):
    """Radiation comoving with the gas: the fluid-frame energy is E_rf."""
    P = KP.params_implicit_batch()
    pp, g20, G20, _ = PR.implicit_batch_states(P, 100, seed=11)
    pp[:, P.ee + 1:P.ee + 4] = pp[:, 2:5]
    Rtt = oracle.ff_Rtt(P, pp, g20, G20)
    assert np.allclose(-Rtt, pp[:, P.ee], rtol=1e-10)
