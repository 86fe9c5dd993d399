"""The oracle's W inversion from a given W start (the learned W guess of u2p,
no reference counterpart), pinned by fixed points of the Newton iteration,
and the numpy W-network features and targets by closed forms."""
import sys
import os

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import mlp  # noqa: E402
from koral_amd import params as KP, problems as PR  # noqa: E402


def _states(oracle, n=4000, seed=3):
    P = KP.params_torus_mhd()
    pt, pg, g20, G20 = PR.u2p_guess_states(P, n, seed=seed)
    uu = oracle.p2u(P, pt, g20, G20)
    return P, pt, pg, g20, G20, uu


def test_plain_W_start_reproduces_plain(oracle):
    """Starting from the W that u2p computes from the guess itself gives the
    plain result bit for bit, with the same iteration counts."""
    P, pt, pg, g20, G20, uu = _states(oracle)
    pa, ra, ia, _ = oracle.u2p_solver_from(P, uu, pg, g20, G20, None)
    pb, rb, ib, gb = oracle.u2p_solver_from(P, uu, pg, g20, G20, mlp.plain_W(P, pg, g20, G20))
    assert (ra == 0).mean() > 0.99
    assert np.array_equal(ra, rb) and np.array_equal(ia, ib)
    assert np.array_equal(pa, pb)
    assert gb[ra == 0].all()


def test_exact_W_start_converges_at_once(oracle):
    """From the W of the converged state the first residual is already below
    the tolerance: one iteration, the same state to the solver tolerance."""
    P, pt, pg, g20, G20, uu = _states(oracle)
    pa, ra, ia, _ = oracle.u2p_solver_from(P, uu, pg, g20, G20, None)
    ok = ra == 0
    assert ia[ok].mean() > 3.
    W = np.where(ok, mlp.plain_W(P, pa, g20, G20), 0.)
    pc, rc, ic, gc = oracle.u2p_solver_from(P, uu, pg, g20, G20, W)
    assert (ic[ok] == 1).all() and gc[ok].all()
    rel = np.abs(pc[ok, :5] - pa[ok, :5]) / (np.abs(pa[ok, :5]) + 1e-300)
    assert rel.max() < 1e-8
    # and the solution is the true state the conserved vector came from
    relt = np.abs(pa[ok, :5] - pt[ok, :5]) / (np.abs(pt[ok, :5]) + 1e-12)
    assert np.median(relt) < 1e-9


def test_non_positive_start_is_plain(oracle):
    """W0 <= 0 means no learned start: the plain solver, nothing guessed."""
    P, pt, pg, g20, G20, uu = _states(oracle, n=1000)
    pa, ra, ia, _ = oracle.u2p_solver_from(P, uu, pg, g20, G20, None)
    pd, rd, idd, gd = oracle.u2p_solver_from(P, uu, pg, g20, G20, np.full(len(pg), -1.))
    assert np.array_equal(pa, pd) and np.array_equal(ia, idd) and gd.sum() == 0


def test_failed_start_falls_back_to_plain(oracle):
    """A start far from the root that fails is retried from the plain start:
    such points get the plain result bit for bit, and iteration counts add."""
    P, pt, pg, g20, G20, uu = _states(oracle, n=2000)
    pa, ra, ia, _ = oracle.u2p_solver_from(P, uu, pg, g20, G20, None)
    W = mlp.plain_W(P, pg, g20, G20) * 1e30
    pe, re, ie, ge = oracle.u2p_solver_from(P, uu, pg, g20, G20, W)
    back = ge == 0
    assert back.sum() > 0
    assert np.array_equal(pe[back], pa[back]) and np.array_equal(re[back], ra[back])
    assert (ie[back] >= ia[back]).all()


def test_w_features_and_targets_closed_forms(oracle):
    """Features of a guess equal to the true state: D/D0 = 1, no energy or
    momentum change; targets of identical states are zero; log10(W0/rho)
    and the magnetisation follow from the state."""
    P, pt, pg, g20, G20, uu = _states(oracle, n=500)
    U0 = oracle.p2u(P, pt, g20, G20)
    f = mlp.u2p_raw_features(P, uu, pt, g20, G20, U0)
    assert np.abs(f[:, 5]).max() < 1e-12 and np.abs(f[:, 6:10]).max() < 1e-12
    assert np.allclose(f[:, 0], np.log10(pt[:, 1] / pt[:, 0]))
    assert np.allclose(f[:, 10], np.log10(mlp.plain_W(P, pt, g20, G20) / pt[:, 0]))
    assert (f[:, 1] > -30).all()
    assert np.abs(mlp.u2p_targets(P, pt, pt, g20, G20)).max() == 0.
    t = mlp.u2p_targets(P, pg, pt, g20, G20)
    assert np.allclose(np.exp(t[:, 0]) * mlp.plain_W(P, pg, g20, G20), mlp.plain_W(P, pt, g20, G20))
