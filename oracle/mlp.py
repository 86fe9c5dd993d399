"""Reference (numpy) of the learned initial-guess network: features, targets,
forward and backward passes, Adam.  Test infrastructure: it checks the HIP
training and inference kernels (koral_amd/csrc/k_mlp.hip) and is pinned by
finite differences in tests/test_oracle_mlp.py.

Network: 11 inputs -> 128 -> 128 -> 128 (GELU, tanh form) -> 4 outputs.
Parameters are one flat float32 vector in the order W1[128][11], b1[128],
W2[128][128], b2[128], W3[128][128], b3[128], W4[4][128], b4[4]
(row-major, W[out][in]).

Features of a cell (standardised with the training set's mean and std):
  log10 rho, log10(u/rho), log10(E/u), the three VELR components of the gas
  and of the radiation frame, log10(dt kappa_abs), log10(dt kappa_es)
Targets (standardised with the training set's mean and std) between the
converged state of the energy-equation lab-frame solve (x*) and the initial
state (x0) of the network's four primitives (gas: u, u~^i; radiation: E,
F~^i): log(x*_0 / x0_0) and asinh((x*_i - x0_i) / 1e-6).
"""
import numpy as np

NIN, NH, NOUT = 11, 128, 4
SHAPES = [(NH, NIN), (NH,), (NH, NH), (NH,), (NH, NH), (NH,), (NOUT, NH), (NOUT,)]
NPARAM = sum(int(np.prod(s)) for s in SHAPES)
C_GELU = 0.7978845608028654  # sqrt(2/pi)


def split(theta):
    out, o = [], 0
    for s in SHAPES:
        k = int(np.prod(s))
        out.append(theta[o:o + k].reshape(s))
        o += k
    return out


def init_params(seed=0):
    """He-style uniform initialisation, zero biases (float32)."""
    rng = np.random.default_rng(seed)
    parts = []
    for s in SHAPES:
        if len(s) == 2:
            lim = np.sqrt(6.0 / s[1])
            parts.append(rng.uniform(-lim, lim, s).astype(np.float32).ravel())
        else:
            parts.append(np.zeros(s, np.float32))
    return np.concatenate(parts)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(C_GELU * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    t = np.tanh(C_GELU * (z + 0.044715 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * C_GELU * (1.0 + 3.0 * 0.044715 * z * z)


def forward(theta, X):
    """X [B, 11] -> (Y [B, 4], cache)."""
    W1, b1, W2, b2, W3, b3, W4, b4 = split(theta)
    Z1 = X @ W1.T + b1
    A1 = gelu(Z1)
    Z2 = A1 @ W2.T + b2
    A2 = gelu(Z2)
    Z3 = A2 @ W3.T + b3
    A3 = gelu(Z3)
    Y = A3 @ W4.T + b4
    return Y, (X, Z1, A1, Z2, A2, Z3, A3)


def loss_and_grad(theta, X, T):
    """Mean squared error over batch and outputs, and its gradient."""
    Y, (X, Z1, A1, Z2, A2, Z3, A3) = forward(theta, X)
    W1, b1, W2, b2, W3, b3, W4, b4 = split(theta)
    B = X.shape[0]
    R = Y - T
    loss = (R * R).sum() / (B * NOUT)
    dY = 2.0 * R / (B * NOUT)
    gW4 = dY.T @ A3
    gb4 = dY.sum(0)
    dZ3 = (dY @ W4) * gelu_grad(Z3)
    gW3 = dZ3.T @ A2
    gb3 = dZ3.sum(0)
    dZ2 = (dZ3 @ W3) * gelu_grad(Z2)
    gW2 = dZ2.T @ A1
    gb2 = dZ2.sum(0)
    dZ1 = (dZ2 @ W2) * gelu_grad(Z1)
    gW1 = dZ1.T @ X
    gb1 = dZ1.sum(0)
    g = np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3, gW4, gb4)])
    return loss, g


def adam(theta, g, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (step counts from 1); returns new (theta, m, v)."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** step)
    vh = v / (1.0 - b2 ** step)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v


def raw_features(P, pp, dt, kappa_abs, kappa_es):
    """Unstandardised features [n, 11] of primitive states pp [n, nv]."""
    ee = P.ee
    rho, u, E = pp[:, 0], pp[:, 1], pp[:, ee]
    f = np.empty((pp.shape[0], NIN))
    f[:, 0] = np.log10(rho)
    f[:, 1] = np.log10(u / rho)
    f[:, 2] = np.log10(E / u)
    f[:, 3:6] = pp[:, 2:5]
    f[:, 6:9] = pp[:, ee + 1:ee + 4]
    f[:, 9] = np.log10(dt * kappa_abs)
    f[:, 10] = np.log10(dt * kappa_es)
    return f


VSCALE = 1e-6


def targets(pp0, pp1, base=1):
    """Raw targets [n, 4] between initial and converged primitives of the
    network's kind: base 1 (gas: u, u~^i) or EE (radiation: E, F~^i):
    log(x1_0 / x0_0) and asinh((x1_k - x0_k) / VSCALE)."""
    t = np.empty((pp0.shape[0], NOUT))
    t[:, 0] = np.log(pp1[:, base] / pp0[:, base])
    t[:, 1:4] = np.arcsinh((pp1[:, base + 1:base + 4] - pp0[:, base + 1:base + 4]) / VSCALE)
    return t


def guess(pp0, t, base=1):
    """Newton start [n, 4] from raw targets t (inverse of targets)."""
    g = np.empty((pp0.shape[0], NOUT))
    g[:, 0] = pp0[:, base] * np.exp(t[:, 0])
    g[:, 1:4] = pp0[:, base + 1:base + 4] + VSCALE * np.sinh(t[:, 1:4])
    return g


# ---- the W network of the u2p inversion ----
def _metric(g20, G20):
    """g_{mu nu}, g^{mu nu} [n, 4, 4] from the 20-entry rows (gg[i][j] at
    5 i + j, struct geometry order)."""
    n = g20.shape[0]
    g = g20.reshape(n, 4, 5)[:, :, :4]
    G = G20.reshape(n, 4, 5)[:, :, :4]
    return g, G


def _ucon(v, g, G):
    """VELR u~^i -> u^mu (relele.c conv_vels, VELR -> VEL4)."""
    qsq = np.einsum("ni,nj,nij->n", v, v, g[:, 1:, 1:])
    alpha2 = -1. / G[:, 0, 0]
    alpgam = np.sqrt(alpha2 * (1. + qsq))
    u = np.empty((v.shape[0], 4))
    u[:, 0] = np.sqrt((1. + qsq) / alpha2)
    u[:, 1:] = v - alpgam[:, None] * G[:, 0, 1:]
    return u, qsq


def plain_W(P, pp, g20, G20):
    """The W start of u2p_solver_W (u2p.c:2094-2112): (rho + Gamma u) gamma^2."""
    g, G = _metric(g20, G20)
    _, qsq = _ucon(pp[:, 2:5], g, G)
    return (pp[:, 0] + P.gamma * pp[:, 1]) * (1. + qsq)


def u2p_raw_features(P, uu, pp, g20, G20, U0):
    """Unstandardised W-network features [n, 11] of conserved states uu and
    guesses pp [n, nv]; U0 = p2u(pp) (the oracle's p2u)."""
    g, G = _metric(g20, G20)
    rho, u = pp[:, 0], pp[:, 1]
    ucon, _ = _ucon(pp[:, 2:5], g, G)
    ucov = np.einsum("nij,nj->ni", g, ucon)
    bsq = np.zeros(pp.shape[0])
    if P.mag:
        B = pp[:, 6:9]
        b0 = np.einsum("ni,ni->n", B, ucov[:, 1:])
        bcon = np.empty_like(ucon)
        bcon[:, 0] = b0
        bcon[:, 1:] = (B + b0[:, None] * ucon[:, 1:]) / ucon[:, 0:1]
        bsq = np.einsum("ni,nij,nj->n", bcon, g, bcon)
    f = np.empty((pp.shape[0], NIN))
    f[:, 0] = np.log10(u / rho)
    f[:, 1] = np.log10(bsq / rho + 1e-30)
    for i in range(3):
        f[:, 2 + i] = pp[:, 2 + i] * np.sqrt(g[:, 1 + i, 1 + i])
    f[:, 5] = np.log(uu[:, 0] / U0[:, 0])
    f[:, 6] = np.arcsinh((uu[:, 1] - U0[:, 1]) / np.abs(U0[:, 1]))
    s2 = np.abs(U0[:, 0]) + np.abs(U0[:, 1])
    for i in range(3):
        f[:, 7 + i] = np.arcsinh((uu[:, 2 + i] - U0[:, 2 + i]) / s2)
    f[:, 10] = np.log10(plain_W(P, pp, g20, G20) / rho)
    return f


def u2p_targets(P, pp0, pp1, g20, G20):
    """Raw W targets [n, 4]: log(W(pp1) / W(pp0)) and three zeros."""
    t = np.zeros((pp0.shape[0], NOUT))
    t[:, 0] = np.log(plain_W(P, pp1, g20, G20) / plain_W(P, pp0, g20, G20))
    return t
