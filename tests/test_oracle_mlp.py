"""Pins the numpy reference of the learned-guess network (oracle/mlp.py):
gradients against central finite differences in float64, the GELU
derivative, Adam's first step, and the feature/target definitions."""
import numpy as np

import mlp as M


def test_gelu_grad_matches_finite_difference():
    z = np.linspace(-6., 6., 241)
    h = 1e-6
    fd = (M.gelu(z + h) - M.gelu(z - h)) / (2 * h)
    assert np.abs(fd - M.gelu_grad(z)).max() < 1e-8
    # known values: gelu(0) = 0, gelu(x) -> x for large x, -> 0 for large -x
    assert M.gelu(np.array([0.]))[0] == 0.
    assert abs(M.gelu(np.array([8.]))[0] - 8.) < 1e-12
    assert abs(M.gelu(np.array([-8.]))[0]) < 1e-12


def test_loss_gradient_matches_finite_difference():
    rng = np.random.default_rng(1)
    theta = M.init_params(3).astype(np.float64) * 0.5
    theta[M.NPARAM - 4:] = rng.normal(size=4) * 0.1  # non-zero output bias
    X = rng.normal(size=(7, M.NIN))
    T = rng.normal(size=(7, M.NOUT))
    loss, g = M.loss_and_grad(theta, X, T)
    idx = rng.choice(M.NPARAM, 60, replace=False)
    idx = np.concatenate([idx, [0, M.NIN * M.NH, M.NPARAM - 1, M.NPARAM - 5]])
    h = 1e-6
    for k in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        fd = (M.loss_and_grad(tp, X, T)[0] - M.loss_and_grad(tm, X, T)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1e-3, abs(fd)), (k, fd, g[k])


def test_adam_first_step_is_signed_lr():
    g = np.array([0.5, -2., 1e-3, 0.])
    th, m, v = M.adam(np.zeros(4), g, np.zeros(4), np.zeros(4), 1, lr=1e-3)
    # m_hat = g, v_hat = g^2: the first update is lr * g / (|g| + eps)
    np.testing.assert_allclose(th, -1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-12)


def test_features_and_targets():
    class P:
        ee = 9
    pp = np.zeros((2, 13))
    pp[:, 0] = [1e-2, 10.]
    pp[:, 1] = [1e-4, 1.]
    pp[:, 2:5] = [[0.1, 0.2, 0.3], [1., 2., 3.]]
    pp[:, 9] = [1e-3, 100.]
    pp[:, 10:13] = [[-.1, 0., .1], [4., 5., 6.]]
    f = M.raw_features(P, pp, np.array([2., 3.]), np.array([5., 7.]), np.array([0.5, 1. / 3.]))
    np.testing.assert_allclose(f[:, 0], [-2., 1.])
    np.testing.assert_allclose(f[:, 1], [-2., -1.])
    np.testing.assert_allclose(f[:, 2], [1., 2.])
    np.testing.assert_allclose(f[:, 9], [1., np.log10(21.)])
    np.testing.assert_allclose(f[:, 10], [0., 0.], atol=1e-15)
    pp1 = pp.copy()
    pp1[:, 1] *= np.e
    pp1[:, 2:5] += 1e-6
    pp1[:, 10:13] -= 2e-6
    pp1[:, 9] /= np.e ** 2
    s1, s2 = np.arcsinh(1.), np.arcsinh(2.)
    np.testing.assert_allclose(M.targets(pp, pp1), [[1., s1, s1, s1]] * 2, rtol=1e-6)
    np.testing.assert_allclose(M.targets(pp, pp1, base=9), [[-2., -s2, -s2, -s2]] * 2, rtol=1e-6)
    # the start follows back from the targets
    for base in (1, 9):
        np.testing.assert_allclose(M.guess(pp, M.targets(pp, pp1, base), base), pp1[:, base:base + 4],
                                   rtol=1e-9, atol=1e-15)
