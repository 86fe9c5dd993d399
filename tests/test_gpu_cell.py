"""GPU parity of the batched per-point kernels against the CPU oracle, through
the C-ABI (libkoral_hip.so).  The device rounds every operation as the
oracle does (IEEE FP64 without contraction, the same portable exp/log/pow),
so results, flags, return codes and iteration counts are identical bit for
bit; a mismatch is reported with the cells where it happens."""
import sys

import numpy as np
import pytest

from koral_amd import params as KP, problems as PR

pytestmark = pytest.mark.gpu

LAYOUTS = [(0, 0), (1, 0), (0, 1), (1, 1)]


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from koral_amd import cell
    return cell


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return np.abs(a - b) / (np.abs(a) + np.abs(b) + 1e-300)


def _ident(what, dev, ref):
    """dev and ref equal bit for bit (NaN where both are NaN); prints the
    first differing points otherwise"""
    d, r = np.asarray(dev), np.asarray(ref)
    same = (d == r) | (np.isnan(d) & np.isnan(r)) if d.dtype.kind == "f" else (d == r)
    if not same.all():
        bad = np.argwhere(~same)
        _say("  %s: %d of %d values differ, first at %s: device %r oracle %r" %
             (what, len(bad), same.size, tuple(bad[0]), d[tuple(bad[0])], r[tuple(bad[0])]))
    return bool(same.all())


@pytest.mark.parametrize("mag,rad", LAYOUTS)
def test_p2u_parity(dev, oracle, mag, rad):
    P = KP.base_params(mag, rad, 5. / 3.)
    pp, g20, G20 = PR.random_mhd_states(P, 4096, seed=40 + mag + 2 * rad)
    _say("p2u parity mag=%d rad=%d n=%d" % (mag, rad, len(pp)))
    ref = oracle.p2u(P, pp, g20, G20)
    out = dev.to_host(dev.p2u(P, dev.to_dev(pp), dev.to_dev(g20), dev.to_dev(G20)))
    assert _ident("p2u", out, ref)


@pytest.mark.parametrize("mag,rad", LAYOUTS)
def test_u2p_parity(dev, oracle, mag, rad):
    P = KP.base_params(mag, rad, 13. / 9.)
    pp, g20, G20 = PR.random_mhd_states(P, 8192, seed=50 + mag + 2 * rad)
    uu = oracle.p2u(P, pp, g20, G20)
    rng = np.random.default_rng(1)
    guess = pp * rng.uniform(0.7, 1.3, pp.shape)
    # a few pathological cells: negative density, huge field, cold
    uu[:8, 0] *= -1.
    if mag:
        guess[8:16, 6:9] *= 1e3
        uu[8:16, 6:9] *= 1e3
    _say("u2p parity mag=%d rad=%d n=%d" % (mag, rad, len(pp)))
    uo_r, po_r, ret_r, fl_r = oracle.u2p(P, uu, guess, g20, G20)
    U, Pp = dev.to_dev(uu), dev.to_dev(guess)
    ret, fl = dev.u2p(P, U, Pp, dev.to_dev(g20), dev.to_dev(G20))
    ret, fl = ret.cpu().numpy(), fl.cpu().numpy().T
    po, uo = dev.to_host(Pp), dev.to_host(U)
    assert np.array_equal(ret, ret_r)
    assert np.array_equal(fl, fl_r)
    assert _ident("u2p p", po, po_r) and _ident("u2p u", uo, uo_r)


@pytest.mark.parametrize("etype", [0, 1])
def test_u2p_solver_parity(dev, oracle, etype):
    P = KP.base_params(1, 1, 5. / 3.)
    pp, g20, G20 = PR.random_mhd_states(P, 8192, seed=60 + etype)
    uu = oracle.p2u(P, pp, g20, G20)
    guess = pp * np.random.default_rng(2).uniform(0.5, 1.5, pp.shape)
    _say("u2p_solver parity etype=%d" % etype)
    po_r, ret_r = oracle.u2p_solver(P, uu, guess, g20, G20, etype)
    Pp = dev.to_dev(guess)
    ret = dev.u2p_solver(P, dev.to_dev(uu), Pp, dev.to_dev(g20), dev.to_dev(G20), etype).cpu().numpy()
    assert np.array_equal(ret, ret_r)
    ok = ret == 0
    assert _ident("u2p_solver p", dev.to_host(Pp)[ok], po_r[ok])


def test_u2p_rad_parity(dev, oracle):
    P = KP.base_params(1, 1, 5. / 3.)
    pp, g20, G20 = PR.random_mhd_states(P, 8192, seed=70, gamma_max=400.)
    uu = oracle.p2u(P, pp, g20, G20)
    _say("u2p_rad parity")
    po_r, ret_r, corr_r = oracle.u2p_rad(P, uu, pp, g20, G20)
    Pp = dev.to_dev(pp)
    ret, corr = dev.u2p_rad(P, dev.to_dev(uu), Pp, dev.to_dev(g20), dev.to_dev(G20))
    assert np.array_equal(ret.cpu().numpy(), ret_r)
    assert np.array_equal(corr.cpu().numpy(), corr_r)
    assert corr_r.sum() > 100  # the gamma cap path is exercised
    assert _ident("u2p_rad p", dev.to_host(Pp), po_r)


def test_Gi_parity(dev, oracle):
    P = KP.params_implicit_batch()
    pp, g20, G20, _ = PR.implicit_batch_states(P, 8192, seed=80)
    _say("Gi parity")
    for lab in (1, 0):
        ref = oracle.Gi(P, pp, g20, G20, lab)
        out = dev.to_host(dev.Gi(P, dev.to_dev(pp), dev.to_dev(g20), dev.to_dev(G20), lab))
        assert _ident("G^mu lab=%d" % lab, out, ref)


def test_implicit_parity(dev, oracle):
    """Full fallback chain on 2^14 seeded cells of the batched-solve workload."""
    import torch
    P = KP.params_implicit_batch()
    n = 1 << 14
    pp, g20, G20, dt = PR.implicit_batch_states(P, n, seed=90)
    uu = oracle.p2u(P, pp, g20, G20)
    _say("implicit parity n=%d" % n)
    uo_r, po_r, fl_r, info_r = oracle.implicit(P, uu, pp, g20, G20, dt, nthreads=8)
    U, Pp = dev.to_dev(uu), dev.to_dev(pp)
    fl, info = dev.implicit(P, U, Pp, dev.to_dev(g20), dev.to_dev(G20),
                            torch.from_numpy(dt).cuda())
    torch.cuda.synchronize()
    fl, info = fl.cpu().numpy().T, info.cpu().numpy().T
    po, uo = dev.to_host(Pp), dev.to_host(U)
    same = np.all(info == info_r, axis=1)
    _say("implicit: info identical %.5f, success %.4f" % (same.mean(), (info[:, 1] >= 0).mean()))
    for i in np.where(~same)[0][:10]:
        _say("  cell %d: device info %s oracle %s" % (i, info[i], info_r[i]))
    # attempt, iterations and the chosen (primitives, equation, frame) of
    # every cell, its flags, primitives and conserved state: bit for bit
    assert same.all()
    assert _ident("implicit flags", fl, fl_r)
    assert _ident("implicit p", po, po_r) and _ident("implicit u", uo, uo_r)
