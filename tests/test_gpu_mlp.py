"""GPU checks of the learned-guess network kernels against the numpy
reference (oracle/mlp.py): features, FP32 forward, gradient, Adam step, the
BF16 matrix-core inference (which pins the operand lane layout), and the
implicit solve with the trained guess against the plain solve; the W
network of the u2p inversion and the inversion from its starts."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch.device("cuda:0")


def _bf16(x):
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7fff + ((u >> 16) & 1)) & 0xffff0000
    return u.astype(np.uint32).view(np.float32)


def _rand_net(seed):
    import mlp as M
    rng = np.random.default_rng(seed)
    th = M.init_params(seed)
    th += rng.normal(scale=0.05, size=th.size).astype(np.float32)  # non-zero biases too
    return th.astype(np.float32)


def test_features_match_reference(dev, oracle):
    import torch
    import mlp as M
    from koral_amd import params as KP, problems as PR, cell
    from koral_amd.learn import Net
    P = KP.params_implicit_batch()
    pp, g20, G20, dt = PR.implicit_batch_states(P, 4096, seed=7)
    net = Net(seed=0)
    F = net.features(P, cell.to_dev(pp), torch.from_numpy(dt).to(dev), standardise=False).cpu().numpy().T
    ref = M.raw_features(P, pp, dt, PR.kappa_abs_np(P, pp[:, 0], pp[:, 1]), P.kappa_es_coef * pp[:, 0])
    err = np.abs(F - ref) / (np.abs(ref) + 1e-3)
    _say("features max rel err %.3e" % err.max())
    assert err.max() < 1e-5


def test_forward_and_gradient_match_reference(dev):
    import torch
    import mlp as M
    from koral_amd.learn import Net
    rng = np.random.default_rng(2)
    th = _rand_net(1)
    net = Net(theta=th)
    B = 777
    X = rng.normal(size=(B, M.NIN)).astype(np.float32)
    T = rng.normal(size=(B, M.NOUT)).astype(np.float32)
    Xd = torch.from_numpy(X.T.copy()).to(dev)
    Td = torch.from_numpy(T.T.copy()).to(dev)
    Y = net.forward(Xd).cpu().numpy().T
    Yr, _ = M.forward(th.astype(np.float64), X.astype(np.float64))
    ey = np.abs(Y - Yr).max() / np.abs(Yr).max()
    loss, g = net.grad(Xd, Td)
    lr_, gr = M.loss_and_grad(th.astype(np.float64), X.astype(np.float64), T.astype(np.float64))
    g = g.cpu().numpy()
    eg = np.abs(g - gr) / (np.abs(gr) + 1e-3 * np.abs(gr).max())
    _say("forward max rel %.3e, loss %.6f vs %.6f, grad max rel %.3e" % (ey, loss, lr_, eg.max()))
    assert ey < 1e-5
    assert abs(loss - lr_) < 1e-5 * lr_
    assert eg.max() < 1e-3


def test_adam_step_matches_reference(dev):
    import torch
    import mlp as M
    from koral_amd.learn import Net
    rng = np.random.default_rng(3)
    th = _rand_net(4)
    net = Net(theta=th)
    X = rng.normal(size=(512, M.NIN)).astype(np.float32)
    T = rng.normal(size=(512, M.NOUT)).astype(np.float32)
    Xd, Td = torch.from_numpy(X.T.copy()).to(dev), torch.from_numpy(T.T.copy()).to(dev)
    ref, m, v = th.astype(np.float64), np.zeros(th.size), np.zeros(th.size)
    for k in range(1, 4):
        net.step(Xd, Td, lr=1e-3)
        _, gr = M.loss_and_grad(ref, X.astype(np.float64), T.astype(np.float64))
        ref, m, v = M.adam(ref, gr, m, v, k, lr=1e-3)
    got, _ = net.get()
    err = np.abs(got - ref).max()
    _say("3 Adam steps: max |theta - ref| %.3e (lr 1e-3)" % err)
    assert err < 2e-5


def test_bf16_inference_matches_reference(dev):
    """Outputs of the matrix-core inference against a numpy pass with the same
    BF16 rounding of weights and layer inputs (FP32 sums)."""
    import torch
    import mlp as M
    from koral_amd.learn import Net
    rng = np.random.default_rng(5)
    th = _rand_net(6)
    net = Net(theta=th)
    n = 1000  # not a multiple of the 64-cell tile
    X = rng.normal(size=(n, M.NIN)).astype(np.float32)
    Y = net.infer(torch.from_numpy(X.T.copy()).to(dev)).cpu().numpy().T
    W1, b1, W2, b2, W3, b3, W4, b4 = M.split(th)
    h = _bf16(X)
    for W, b in ((W1, b1), (W2, b2), (W3, b3)):
        h = _bf16(M.gelu((h.astype(np.float64) @ _bf16(W).T.astype(np.float64) + b)).astype(np.float32))
    Yr = h.astype(np.float64) @ _bf16(W4).T.astype(np.float64) + b4
    Yf, _ = M.forward(th.astype(np.float64), X.astype(np.float64))
    err = np.abs(Y - Yr).max() / np.abs(Yr).max()
    err32 = np.abs(Y - Yf).max() / np.abs(Yf).max()
    _say("bf16 inference: max rel vs bf16 reference %.3e, vs fp64 network %.3e" % (err, err32))
    assert err < 2e-2
    assert err32 < 5e-2


def test_targets_and_guess_match_reference(dev, oracle):
    """Target kernel of both kinds against the numpy reference, and the guess
    kernel as its inverse (outputs set to the standardised targets)."""
    import torch
    import mlp as M
    from koral_amd import params as KP, problems as PR, cell
    from koral_amd import learn as LN
    P = KP.params_implicit_batch()
    pp, g20, G20, dt = PR.implicit_batch_states(P, 4096, seed=8)
    uu = oracle.p2u(P, pp, g20, G20)
    _, p1, _, info = oracle.implicit(P, uu, pp, g20, G20, dt, nthreads=8)
    ok = info[:, 1] >= 0
    pp, p1 = pp[ok], p1[ok]
    P0, P1 = cell.to_dev(pp), cell.to_dev(p1)
    for kind, base in ((LN.GAS, 1), (LN.RAD, P.ee)):
        T = LN.raw_targets(P0, P1, kind).cpu().numpy().T
        ref = M.targets(pp, p1, base)
        err = np.abs(T - ref) / (np.abs(ref) + 1e-3)
        net = LN.Net(kind=kind)
        norm = np.concatenate([np.zeros(11), np.ones(11), [0.1, -0.2, 0.3, 0.], [2., 3., 4., 5.]]).astype(np.float32)
        net.set(LN.init_params(0), norm)
        Y = ((torch.from_numpy(ref.T.copy()).float().to(dev) - torch.from_numpy(norm[22:26, None]).to(dev))
             / torch.from_numpy(norm[26:30, None]).to(dev)).contiguous()
        gx = net.guess(P0, Y).cpu().numpy().T
        # the targets are float32 (|t| up to ~30): the round trip keeps ~1e-6 of
        # the change and of the value
        x0, x1 = pp[:, base:base + 4], p1[:, base:base + 4]
        gerr = np.abs(gx - x1) / (np.abs(x1) + np.abs(x1 - x0))
        _say("kind %d: targets max rel %.2e, guess from targets max rel %.2e" % (kind, err.max(), gerr.max()))
        assert err.max() < 1e-5
        assert gerr.max() < 1e-4


def _trained_pair(dev, P, n_train, seed, schedule):
    import torch
    from koral_amd import problems as PR, cell
    from koral_amd import learn as LN
    pp, g20, G20, dt = PR.implicit_batch_states(P, n_train, seed=seed)
    Pp, g, G, dtd = cell.to_dev(pp), cell.to_dev(g20), cell.to_dev(G20), torch.from_numpy(dt).to(dev)
    U = cell.p2u(P, Pp, g, G)
    P0 = Pp.clone()
    _, info = cell.implicit(P, U, Pp, g, G, dtd)
    F, Tg, Tr = LN.training_pairs(P, P0, Pp, dtd, info)
    gas, hg = LN.train(P, F, Tg, LN.GAS, schedule=schedule)
    rad, hr = LN.train(P, F, Tr, LN.RAD, schedule=schedule)
    return gas, rad, hg, hr, F.shape[1]


def test_learned_guess_solve(dev, oracle):
    """Train both networks on pairs from the plain solver, then solve a fresh
    batch from the learned guess: every cell the plain chain solves is solved,
    cells the learned attempt leaves to the chain get exactly the plain
    result, cells solved by both from the gas primitives agree to the solver
    tolerance, and the mean number of Newton iterations drops."""
    import torch
    from koral_amd import params as KP, problems as PR, cell
    from koral_amd import learn as LN
    P = KP.params_implicit_batch()
    gas, rad, hg, hr, npairs = _trained_pair(dev, P, 1 << 18, 21, [(1e-3, 12), (3e-4, 6), (1e-4, 3)])
    _say("training pairs %d, loss gas %s rad %s" % (npairs, ["%.3f" % h for h in hg], ["%.3f" % h for h in hr]))
    assert hg[-1] < hg[0] and hr[-1] < hr[0]
    pp, g20, G20, dt = PR.implicit_batch_states(P, 1 << 15, seed=22)
    Pp, g, G, dtd = cell.to_dev(pp), cell.to_dev(g20), cell.to_dev(G20), torch.from_numpy(dt).to(dev)
    U0 = cell.p2u(P, Pp, g, G)
    Ua, Pa = U0.clone(), Pp.clone()
    _, ia = cell.implicit(P, Ua, Pa, g, G, dtd)
    ia = ia.cpu().numpy()
    pa = cell.to_host(Pa)
    ok_a = ia[1] >= 0
    res = {}
    for name, r, budget in (("gas", None, 0), ("gas+rad", rad, 0), ("gas+rad", rad, 16), ("gas+rad", rad, 10)):
        Ub, Pb = U0.clone(), Pp.clone()
        _, ib = LN.implicit_learned(P, gas, r, Ub, Pb, g, G, dtd, maxiter=budget)
        ib = ib.cpu().numpy()
        pb = cell.to_host(Pb)
        ok_b = ib[1] >= 0
        learned = ib[1] == 6
        _say("%s budget %d: plain solved %d, learned solved %d (%d by the learned attempt, %d on radiation "
             "primitives); mean iterations plain %.3f learned %.3f (x%.2f)" %
             (name, budget, ok_a.sum(), ok_b.sum(), learned.sum(), (learned & (ib[2] == 1)).sum(),
              ia[5].mean(), ib[5].mean(), ia[5].mean() / ib[5].mean()))
        assert np.all(ok_b[ok_a]), "the learned attempt may only add solved cells"
        same = (ib[1] == ia[1]) & ok_a
        assert np.array_equal(pa[same], pb[same])
        both = learned & (ia[1] == 0) & (ia[2] == 2) & (ib[2] == 2)
        du = np.abs(pa[:, 1] - pb[:, 1]) / np.abs(pa[:, 1])
        _say("  same system from both starts (%d cells): |du|/u median %.2e p99 %.2e" %
             (both.sum(), np.median(du[both]), np.quantile(du[both], 0.99)))
        res[(name, budget)] = ib[5].mean()
    assert res[("gas+rad", 0)] < ia[5].mean()


def test_learned_solve_matches_oracle(dev, oracle):
    """The device solve with the learned first attempt against the oracle's
    solve from the same starts (downloaded from the network kernels): same
    attempt indices and iteration counts, primitives to the parity bar of the
    plain implicit solve."""
    import torch
    from koral_amd import params as KP, problems as PR, cell
    from koral_amd import learn as LN
    P = KP.params_implicit_batch()
    gas, rad, _, _, _ = _trained_pair(dev, P, 1 << 16, 41, [(1e-3, 3)])
    n = 3000
    pp, g20, G20, dt = PR.implicit_batch_states(P, n, seed=42)
    uu = oracle.p2u(P, pp, g20, G20)
    Pd, Ud, gd, Gd = cell.to_dev(pp), cell.to_dev(uu), cell.to_dev(g20), cell.to_dev(G20)
    dtd = torch.from_numpy(dt).to(dev)
    gxg = gas.start(P, Pd, dtd).cpu().numpy().T.copy()
    gxr = rad.start(P, Pd, dtd).cpu().numpy().T.copy()
    for r, budget in ((None, 0), (rad, 0), (rad, 12)):
        U1, P1 = Ud.clone(), Pd.clone()
        fl, info = LN.implicit_learned(P, gas, r, U1, P1, gd, Gd, dtd, maxiter=budget)
        info = info.cpu().numpy().T
        po = cell.to_host(P1)
        uo_r, po_r, fl_r, info_r = oracle.implicit(P, uu, pp, g20, G20, dt, gx=gxg, maxiter=budget,
                                                   gxr=gxr if r is not None else None, nthreads=8)
        same = np.all(info == info_r, axis=1)
        ok = same & (info[:, 1] >= 0)
        rel = np.abs(po[ok] - po_r[ok]) / (np.abs(po[ok]) + np.abs(po_r[ok]) + 1e-300)
        _say("%s budget %d: %d cells, learned attempt %d (%d radiation primitives), info identical %.4f, "
             "max rel p %.2e" % ("gas" if r is None else "gas+rad", budget, n, (info[:, 1] == 6).sum(),
                                 ((info[:, 1] == 6) & (info[:, 2] == 1)).sum(), same.mean(), rel.max()))
        assert (info[:, 1] == 6).sum() > 0
        # the same starts give the same iterates: identical bit for bit
        assert same.all()
        assert np.array_equal(po, po_r, equal_nan=True)


def test_gated_learned_solve(dev, oracle):
    """A gate network (kind 3), trained on whether the learned tries solve a
    training cell: the cells it scores below zero skip the learned tries and
    give the plain solve's result bit for bit (same attempt, iterations and
    primitives); the device matches the oracle's solve from the same starts
    with those cells' gas start set to NaN (no learned try)."""
    import torch
    from koral_amd import params as KP, problems as PR, cell
    from koral_amd import learn as LN
    P = KP.params_implicit_batch()
    gas, rad, _, _, _ = _trained_pair(dev, P, 1 << 16, 41, [(1e-3, 3)])
    # gate training cells: the learned tries' outcome on a fresh batch
    pp, g20, G20, dt = PR.implicit_batch_states(P, 1 << 16, seed=43)
    P0, g, G, dtd = cell.to_dev(pp), cell.to_dev(g20), cell.to_dev(G20), torch.from_numpy(dt).to(dev)
    U, Pp = cell.p2u(P, P0, g, G), P0.clone()
    _, info_l = LN.implicit_learned(P, gas, rad, U, Pp, g, G, dtd, maxiter=16)
    Fg, Tgt = LN.gate_pairs(P, P0, dtd, info_l)
    gate, hist = LN.train(P, Fg, Tgt, LN.GATE, schedule=[(1e-3, 4)])
    _say("gate trained on %d cells (%.3f positive), loss %s" % (Fg.shape[1], float((Tgt[0] > 0).float().mean()),
                                                                 ["%.3f" % h for h in hist]))
    # fresh evaluation batch
    n = 3000
    pp, g20, G20, dt = PR.implicit_batch_states(P, n, seed=44)
    uu = oracle.p2u(P, pp, g20, G20)
    Pd, Ud, gd, Gd = cell.to_dev(pp), cell.to_dev(uu), cell.to_dev(g20), cell.to_dev(G20)
    dtd = torch.from_numpy(dt).to(dev)
    _, nm = gate.get()
    Y = gate.infer(gate.features(P, Pd, dtd)).cpu().numpy()
    score = Y[0] * nm[2 * LN.NIN + LN.NOUT] + nm[2 * LN.NIN]
    skipped = ~(score >= 0.)
    _say("gate skips %d of %d cells" % (skipped.sum(), n))
    assert 0 < skipped.sum() < n
    U1, P1 = Ud.clone(), Pd.clone()
    fl_p, info_p = cell.implicit(P, U1, P1, gd, Gd, dtd)
    U2, P2 = Ud.clone(), Pd.clone()
    fl_g, info_g = LN.implicit_learned_gated(P, gas, rad, gate, U2, P2, gd, Gd, dtd, maxiter=16, thr=0.0)
    info_p, info_g = info_p.cpu().numpy().T, info_g.cpu().numpy().T
    p_p, p_g = cell.to_host(P1), cell.to_host(P2)
    assert np.array_equal(info_g[skipped], info_p[skipped])
    assert np.array_equal(p_g[skipped], p_p[skipped], equal_nan=True)
    assert np.array_equal(fl_g.cpu().numpy().T[skipped], fl_p.cpu().numpy().T[skipped])
    # the oracle from the same starts, NaN gas start where the gate skips
    gxg = gas.start(P, Pd, dtd).cpu().numpy().T.copy()
    gxr = rad.start(P, Pd, dtd).cpu().numpy().T.copy()
    gxg[skipped, 0] = np.nan
    uo_r, po_r, fl_r, info_r = oracle.implicit(P, uu, pp, g20, G20, dt, gx=gxg, maxiter=16, gxr=gxr, nthreads=8)
    same = np.all(info_g == info_r, axis=1)
    ok = same & (info_g[:, 1] >= 0)
    rel = np.abs(p_g[ok] - po_r[ok]) / (np.abs(p_g[ok]) + np.abs(po_r[ok]) + 1e-300)
    _say("gated solve vs oracle: info identical %.4f, max rel p %.2e; learned attempt %d cells" %
         (same.mean(), rel.max(), (info_g[:, 1] == 6).sum()))
    assert same.all()
    assert np.array_equal(p_g, po_r, equal_nan=True)


# ---- the W network of the u2p inversion ----
def _u2p_batch(dev, oracle, n, seed):
    from koral_amd import params as KP, problems as PR, cell
    P = KP.params_torus_mhd()
    pt, pg, g20, G20 = PR.u2p_guess_states(P, n, seed=seed)
    uu = oracle.p2u(P, pt, g20, G20)
    return P, pt, pg, g20, G20, uu


def test_u2p_features_targets_match_reference(dev, oracle):
    """W-network feature and target kernels against the numpy restatement."""
    import mlp as M
    from koral_amd import cell
    from koral_amd import learn as LN
    P, pt, pg, g20, G20, uu = _u2p_batch(dev, oracle, 4096, 31)
    U, Pg, Pt, g, G = (cell.to_dev(a) for a in (uu, pg, pt, g20, G20))
    F = LN.u2p_raw_features(P, U, Pg, g, G).cpu().numpy().T
    ref = M.u2p_raw_features(P, uu, pg, g20, G20, oracle.p2u(P, pg, g20, G20))
    ferr = np.abs(F - ref) / (np.abs(ref) + 1e-3)
    T = LN.u2p_raw_targets(P, Pg, Pt, g, G).cpu().numpy().T
    tref = M.u2p_targets(P, pg, pt, g20, G20)
    terr = np.abs(T - tref) / (np.abs(tref) + 1e-3)
    _say("W features max rel %.2e, targets max rel %.2e" % (ferr.max(), terr.max()))
    assert ferr.max() < 1e-5 and terr.max() < 1e-5


def test_u2p_solver_from_matches_oracle(dev, oracle):
    """The device W inversion from given W starts (good, poor and failing)
    against the oracle: return codes, iteration counts and the learned-start
    flag identical, primitives to the parity bar."""
    import torch
    import mlp as M
    from koral_amd import cell
    from koral_amd import learn as LN
    P, pt, pg, g20, G20, uu = _u2p_batch(dev, oracle, 8192, 32)
    rng = np.random.default_rng(5)
    W = M.plain_W(P, pt, g20, G20) * np.exp(rng.normal(scale=0.01, size=len(pt)))
    W[:1000] = M.plain_W(P, pg[:1000], g20[:1000], G20[:1000]) * 1e30  # starts that fail
    W[1000:1500] = -1.                                                  # no start
    U, g, G = cell.to_dev(uu), cell.to_dev(g20), cell.to_dev(G20)
    for W0 in (None, W):
        Pd = cell.to_dev(pg)
        r, it, gd = LN.u2p_solver_from(P, U, Pd, g, G, None if W0 is None else torch.from_numpy(W0).to(dev))
        po = cell.to_host(Pd)
        pr, rr, ir, gr = oracle.u2p_solver_from(P, uu, pg, g20, G20, W0)
        r, it, gd = r.cpu().numpy(), it.cpu().numpy(), gd.cpu().numpy()
        ok = rr == 0
        rel = np.abs(po[ok] - pr[ok]) / (np.abs(pr[ok]) + 1e-300)
        _say("W start %s: solved %d/%d, learned start used %d, mean iters %.3f, max rel p %.2e" %
             ("none" if W0 is None else "given", ok.sum(), len(ok), gd.sum(), it.mean(), rel.max()))
        assert np.array_equal(r, rr) and np.array_equal(it, ir) and np.array_equal(gd, gr)
        assert rel.max() < 1e-10


def test_learned_u2p_guess(dev, oracle):
    """Train the W network on pairs from the plain solver and invert a fresh
    batch from its W starts: every point the plain solver inverts is
    inverted, to the solver tolerance, in no more Newton iterations on
    average; points whose learned start fails get the plain result.  (The
    gain is small: the plain start from the previous primitives is already
    within a few per cent of W and the Newton iteration converges
    quadratically, while the network's start is good to about 1 %.)"""
    import torch
    from koral_amd import cell
    from koral_amd import learn as LN
    P, pt, pg, g20, G20, uu = _u2p_batch(dev, oracle, 1 << 18, 33)
    U, Pg, g, G = (cell.to_dev(a) for a in (uu, pg, g20, G20))
    F, T = LN.u2p_training_pairs(P, U, Pg, g, G)
    net, hist = LN.train(P, F, T, LN.WU2P, schedule=[(1e-3, 8), (3e-4, 4), (1e-4, 2)])
    _say("W network: %d pairs, loss %s" % (F.shape[1], ["%.4f" % h for h in hist]))
    assert hist[-1] < 0.5 * hist[0]
    P2, pt2, pg2, g202, G202, uu2 = _u2p_batch(dev, oracle, 1 << 16, 34)
    U2, g2, G2 = cell.to_dev(uu2), cell.to_dev(g202), cell.to_dev(G202)
    Pa = cell.to_dev(pg2)
    ra, ia, _ = LN.u2p_solver_from(P, U2, Pa, g2, G2, None)
    Pb = cell.to_dev(pg2)
    rb, ib, gb = LN.u2p_solver_learned(P, net, U2, Pb, g2, G2)
    ra, ia, rb, ib, gb = (x.cpu().numpy() for x in (ra, ia, rb, ib, gb))
    pa, pb = cell.to_host(Pa), cell.to_host(Pb)
    ok = ra == 0
    # rho and u~^i relative to themselves; u relative to rho + u (u comes from
    # W / gamma^2 - rho, which cancels for cold states)
    scale = np.abs(pa[ok, :5]) + 1e-300
    scale[:, 1] = pa[ok, 0] + pa[ok, 1]
    rel = np.abs(pb[ok, :5] - pa[ok, :5]) / scale
    # the inversion stops on a residual below 1e-10 or on |dW/W| < 1e-10 with
    # a residual below 0.1, so converged answers may differ beyond 1e-10;
    # the learned answer must be as close to the true state as the plain one
    err_a = (np.abs(pa[ok, :5] - pt2[ok, :5]) / scale).max(1)
    err_b = (np.abs(pb[ok, :5] - pt2[ok, :5]) / scale).max(1)
    _say("plain: solved %.4f, mean iters %.3f; learned: solved %.4f (learned start %.4f), mean iters %.3f (x%.3f); "
         "learned vs plain max rel %.2e median %.2e; error vs true state plain max %.2e learned max %.2e" %
         (ok.mean(), ia.mean(), (rb == 0).mean(), gb.mean(), ib.mean(), ia.mean() / ib.mean(), rel.max(),
          np.median(rel), err_a.max(), err_b.max()))
    assert (rb[ok] == 0).all()
    assert np.median(rel) < 1e-10
    assert err_b.max() <= 2. * err_a.max() + 1e-12
    back = gb == 0
    assert np.array_equal(pb[back], pa[back])
    assert gb.mean() > 0.9
    assert ib.mean() <= ia.mean()
