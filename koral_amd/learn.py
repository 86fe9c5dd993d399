"""Learned initial guess of the implicit radiation solve: host driver of the
network kernels in the C-ABI (include/koral_hip.h, koral_hip_mlp_*).

    gas, rad = Net(GAS), Net(RAD)                    # parameters on the device
    F, Tg, Tr = training_pairs(P, pp0, pp1, dt, info)  # from the plain solver
    norm = normalisation(F, Tg); gas.set(init_params(0), norm)
    gas.fit(*standardise(F, Tg, norm), epochs=4)       # HIP forward/backward/Adam
    ... the same for rad with Tr ...
    flags, info = implicit_learned(P, gas, rad, uu, pp, g20, G20, dt)

Training pairs come from the plain FP64 solver: the cells it solves with the
energy equation in the lab frame (attempt 0 or 2 of the fallback chain, the
same equation on either primitive set) give (features of the initial state,
targets from the converged state) for both networks.  Arrays are SoA torch
CUDA tensors as in koral_amd.cell; torch only holds memory and shuffles
batches.
"""
import ctypes

import numpy as np
import torch

from ._lib import lib, check
from .cell import _ptr, _stream, _chk

NIN, NH, NOUT = 11, 128, 4
BATCH = 4096
GAS, RAD, WU2P, GATE = 0, 1, 2, 3

_f32p = ctypes.POINTER(ctypes.c_float)


def _sig(L):
    if getattr(L, "_mlp_sigs", False):
        return L
    vp, i, l, f = ctypes.c_void_p, ctypes.c_int, ctypes.c_long, ctypes.c_float
    for name, args in {
        "koral_hip_mlp_create": [i, ctypes.POINTER(vp)],
        "koral_hip_mlp_kind": [vp],
        "koral_hip_mlp_destroy": [vp],
        "koral_hip_mlp_set": [vp, vp, vp],
        "koral_hip_mlp_get": [vp, vp, vp],
        "koral_hip_mlp_features": [vp, vp, l, vp, vp, vp, vp],
        "koral_hip_mlp_targets": [i, l, i, vp, vp, vp, vp],
        "koral_hip_mlp_forward": [vp, vp, i, vp, vp],
        "koral_hip_mlp_grad": [vp, vp, vp, i, vp, vp, vp],
        "koral_hip_mlp_train_step": [vp, vp, vp, i, f, vp, vp],
        "koral_hip_mlp_infer": [vp, vp, l, vp, vp],
        "koral_hip_mlp_guess": [vp, l, i, vp, vp, vp, vp],
        "koral_hip_implicit_batch_learned": [vp, vp, vp, l, vp, vp, vp, vp, vp, ctypes.c_double, i, vp, vp,
                                             vp],
        "koral_hip_implicit_batch_learned_gated": [vp, vp, vp, vp, f, l, vp, vp, vp, vp, vp, ctypes.c_double, i,
                                                   vp, vp, vp],
        "koral_hip_mlp_gate": [vp, l, vp, f, vp, vp],
        "koral_hip_implicit_batch_from": [vp, l, vp, vp, vp, vp, vp, ctypes.c_double, vp, vp, i, vp, vp, vp],
        "koral_hip_mlp_u2p_features": [vp, vp, l, vp, vp, vp, vp, vp, vp],
        "koral_hip_mlp_u2p_targets": [vp, l, vp, vp, vp, vp, vp, vp],
        "koral_hip_mlp_u2p_guess": [vp, vp, l, vp, vp, vp, vp, vp, vp],
        "koral_hip_u2p_solver_batch_from": [vp, l, vp, vp, vp, vp, i, vp, vp, vp, vp, vp],
        "koral_hip_u2p_solver_batch_learned": [vp, vp, l, vp, vp, vp, vp, i, vp, vp, vp, vp],
    }.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = i
    L.koral_hip_mlp_nparam.restype = i
    L._mlp_sigs = True
    return L


def nparam():
    return _sig(lib()).koral_hip_mlp_nparam()


def init_params(seed=0):
    """He-style uniform initial weights, zero biases (float32)."""
    rng = np.random.default_rng(seed)
    parts = []
    for o, i in ((NH, NIN), (NH, NH), (NH, NH), (NOUT, NH)):
        lim = np.sqrt(6.0 / i)
        parts.append(rng.uniform(-lim, lim, (o, i)).astype(np.float32).ravel())
        parts.append(np.zeros(o, np.float32))
    return np.concatenate(parts)


class Net:
    """One network on the device; kind GAS corrects (u, u~^i), RAD (E, F~^i),
    WU2P predicts the W start of the u2p inversion, GATE scores whether the
    learned tries will solve a cell."""

    def __init__(self, kind=GAS, seed=0, theta=None, norm=None):
        self.L = _sig(lib())
        h = ctypes.c_void_p()
        check(self.L.koral_hip_mlp_create(int(kind), ctypes.byref(h)), "mlp_create")
        self.h = h
        self.kind = kind
        theta = init_params(seed) if theta is None else np.asarray(theta, np.float32)
        if norm is None:
            norm = np.concatenate([np.zeros(NIN), np.ones(NIN), np.zeros(NOUT), np.ones(NOUT)])
        self.set(theta, norm)

    def close(self):
        if self.h:
            self.L.koral_hip_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set(self, theta, norm):
        th = np.ascontiguousarray(theta, np.float32)
        nm = np.ascontiguousarray(norm, np.float32)
        assert th.size == nparam() and nm.size == 2 * (NIN + NOUT)
        check(self.L.koral_hip_mlp_set(self.h, th.ctypes.data_as(ctypes.c_void_p),
                                       nm.ctypes.data_as(ctypes.c_void_p)), "mlp_set")

    def get(self):
        th = np.zeros(nparam(), np.float32)
        nm = np.zeros(2 * (NIN + NOUT), np.float32)
        check(self.L.koral_hip_mlp_get(self.h, th.ctypes.data_as(ctypes.c_void_p),
                                       nm.ctypes.data_as(ctypes.c_void_p)), "mlp_get")
        return th, nm

    # ---- FP32 training kernels on one batch [11][B], [4][B] ----
    def forward(self, X):
        B = X.shape[1]
        Y = torch.empty((NOUT, B), dtype=torch.float32, device=X.device)
        check(self.L.koral_hip_mlp_forward(self.h, _ptr(_chk(X, torch.float32)), B, _ptr(Y), _stream()),
              "mlp_forward")
        return Y

    def grad(self, X, T):
        B = X.shape[1]
        g = torch.empty(nparam(), dtype=torch.float32, device=X.device)
        loss = ctypes.c_float(0.)
        check(self.L.koral_hip_mlp_grad(self.h, _ptr(_chk(X, torch.float32)), _ptr(_chk(T, torch.float32)), B,
                                        _ptr(g), ctypes.byref(loss), _stream()), "mlp_grad")
        return loss.value, g

    def step(self, X, T, lr=1e-3, want_loss=False):
        B = X.shape[1]
        loss = ctypes.c_float(0.)
        check(self.L.koral_hip_mlp_train_step(self.h, _ptr(_chk(X, torch.float32)), _ptr(_chk(T, torch.float32)),
                                              B, float(lr), ctypes.byref(loss) if want_loss else None,
                                              _stream()), "mlp_train_step")
        return loss.value if want_loss else None

    def fit(self, X, T, epochs=4, lr=1e-3, batch=BATCH, seed=0):
        """Adam over shuffled mini-batches of the standardised pairs
        X [11][N], T [4][N] (device).  Returns the mean loss of each epoch
        (from every 64th batch)."""
        N = X.shape[1]
        gen = torch.Generator(device=X.device)
        gen.manual_seed(seed)
        hist = []
        for ep in range(epochs):
            perm = torch.randperm(N, device=X.device, generator=gen)
            losses = []
            for k, b0 in enumerate(range(0, N - batch + 1, batch)):
                idx = perm[b0:b0 + batch]
                xb = X[:, idx].contiguous()
                tb = T[:, idx].contiguous()
                l = self.step(xb, tb, lr, want_loss=(k % 64 == 0))
                if l is not None:
                    losses.append(l)
            hist.append(float(np.mean(losses)) if losses else float("nan"))
        return hist

    # ---- inference ----
    def features(self, P, pp, dt, standardise=True):
        n = pp.shape[1]
        F = torch.empty((NIN, n), dtype=torch.float32, device=pp.device)
        check(self.L.koral_hip_mlp_features(ctypes.byref(P), self.h if standardise else None, n, _ptr(_chk(pp)),
                                            _ptr(_chk(dt)), _ptr(F), _stream()), "mlp_features")
        return F

    def infer(self, F):
        n = F.shape[1]
        Y = torch.empty((NOUT, n), dtype=torch.float32, device=F.device)
        check(self.L.koral_hip_mlp_infer(self.h, _ptr(_chk(F, torch.float32)), n, _ptr(Y), _stream()),
              "mlp_infer")
        return Y

    def guess(self, pp, Y):
        n = pp.shape[1]
        gx = torch.empty((4, n), dtype=torch.float64, device=pp.device)
        check(self.L.koral_hip_mlp_guess(self.h, n, pp.shape[0], _ptr(_chk(pp)), _ptr(_chk(Y, torch.float32)),
                                         _ptr(gx), _stream()), "mlp_guess")
        return gx

    def start(self, P, pp, dt):
        """Newton start [4][n] predicted for the cells of pp."""
        return self.guess(pp, self.infer(self.features(P, pp, dt)))


    # ---- the W network ----
    def u2p_features(self, P, uu, pp, g20, G20, standardise=True):
        n = pp.shape[1]
        F = torch.empty((NIN, n), dtype=torch.float32, device=pp.device)
        check(self.L.koral_hip_mlp_u2p_features(ctypes.byref(P), self.h if standardise else None, n,
                                                _ptr(_chk(uu)), _ptr(_chk(pp)), _ptr(_chk(g20)), _ptr(_chk(G20)),
                                                _ptr(F), _stream()), "mlp_u2p_features")
        return F

    def u2p_start(self, P, uu, pp, g20, G20):
        """W start [n] predicted for the points (0 where the plain start is kept)."""
        n = pp.shape[1]
        Y = self.infer(self.u2p_features(P, uu, pp, g20, G20))
        W0 = torch.empty(n, dtype=torch.float64, device=pp.device)
        check(self.L.koral_hip_mlp_u2p_guess(ctypes.byref(P), self.h, n, _ptr(_chk(pp)), _ptr(_chk(g20)),
                                             _ptr(_chk(G20)), _ptr(_chk(Y, torch.float32)), _ptr(W0), _stream()),
              "mlp_u2p_guess")
        return W0


def u2p_raw_features(P, uu, pp, g20, G20):
    L = _sig(lib())
    n = pp.shape[1]
    F = torch.empty((NIN, n), dtype=torch.float32, device=pp.device)
    check(L.koral_hip_mlp_u2p_features(ctypes.byref(P), None, n, _ptr(_chk(uu)), _ptr(_chk(pp)), _ptr(_chk(g20)),
                                       _ptr(_chk(G20)), _ptr(F), _stream()), "mlp_u2p_features")
    return F


def u2p_raw_targets(P, pp0, pp1, g20, G20):
    L = _sig(lib())
    n = pp0.shape[1]
    T = torch.empty((NOUT, n), dtype=torch.float32, device=pp0.device)
    check(L.koral_hip_mlp_u2p_targets(ctypes.byref(P), n, _ptr(_chk(pp0)), _ptr(_chk(pp1)), _ptr(_chk(g20)),
                                      _ptr(_chk(G20)), _ptr(T), _stream()), "mlp_u2p_targets")
    return T


def u2p_solver_from(P, uu, pp, g20, G20, W0=None, etype=0):
    """koral_hip_u2p_solver_batch_from in place on pp: (ret, iters, guessed)."""
    L = _sig(lib())
    n = pp.shape[1]
    ret = torch.zeros(n, dtype=torch.int32, device=pp.device)
    iters = torch.zeros(n, dtype=torch.int32, device=pp.device)
    guessed = torch.zeros(n, dtype=torch.int32, device=pp.device)
    check(L.koral_hip_u2p_solver_batch_from(ctypes.byref(P), n, _ptr(_chk(uu)), _ptr(_chk(pp)), _ptr(_chk(g20)),
                                            _ptr(_chk(G20)), int(etype),
                                            _ptr(_chk(W0)) if W0 is not None else None, _ptr(ret), _ptr(iters),
                                            _ptr(guessed), _stream()), "u2p_solver_from")
    return ret, iters, guessed


def u2p_solver_learned(P, net, uu, pp, g20, G20, etype=0):
    """koral_hip_u2p_solver_batch_learned in place on pp: (ret, iters, guessed)."""
    n = pp.shape[1]
    ret = torch.zeros(n, dtype=torch.int32, device=pp.device)
    iters = torch.zeros(n, dtype=torch.int32, device=pp.device)
    guessed = torch.zeros(n, dtype=torch.int32, device=pp.device)
    check(net.L.koral_hip_u2p_solver_batch_learned(ctypes.byref(P), net.h, n, _ptr(_chk(uu)), _ptr(_chk(pp)),
                                                   _ptr(_chk(g20)), _ptr(_chk(G20)), int(etype), _ptr(ret),
                                                   _ptr(iters), _ptr(guessed), _stream()), "u2p_solver_learned")
    return ret, iters, guessed


def u2p_training_pairs(P, uu, pp0, g20, G20):
    """Raw features [11][m] and W targets [4][m] from the plain solver: the
    points it inverts from the guess pp0 (hot inversion) give (features of
    (uu, pp0), log(W* / W0))."""
    pp1 = pp0.clone()
    ret, _, _ = u2p_solver_from(P, uu, pp1, g20, G20, None)
    F = u2p_raw_features(P, uu, pp0, g20, G20)
    T = u2p_raw_targets(P, pp0, pp1, g20, G20)
    keep = (ret == 0) & torch.isfinite(F).all(0) & torch.isfinite(T).all(0)
    return F[:, keep].contiguous(), T[:, keep].contiguous()


def raw_targets(pp0, pp1, kind=GAS):
    L = _sig(lib())
    n = pp0.shape[1]
    T = torch.empty((NOUT, n), dtype=torch.float32, device=pp0.device)
    check(L.koral_hip_mlp_targets(int(kind), n, pp0.shape[0], _ptr(_chk(pp0)), _ptr(_chk(pp1)), _ptr(T),
                                  _stream()), "mlp_targets")
    return T


def training_pairs(P, pp0, pp1, dt, info):
    """Raw features [11][m] and gas and radiation targets [4][m] of the cells
    the plain solver solved with the energy equation in the lab frame
    (attempts 0 and 2)."""
    sel = (info[1] == 0) | (info[1] == 2)
    L = _sig(lib())
    n = pp0.shape[1]
    F = torch.empty((NIN, n), dtype=torch.float32, device=pp0.device)
    check(L.koral_hip_mlp_features(ctypes.byref(P), None, n, _ptr(_chk(pp0)), _ptr(_chk(dt)), _ptr(F),
                                   _stream()), "mlp_features")
    Tg = raw_targets(pp0, pp1, GAS)
    Tr = raw_targets(pp0, pp1, RAD)
    keep = sel & torch.isfinite(F).all(0) & torch.isfinite(Tg).all(0) & torch.isfinite(Tr).all(0)
    return F[:, keep].contiguous(), Tg[:, keep].contiguous(), Tr[:, keep].contiguous()


def normalisation(F, T):
    """[feature mean, std, target mean, std] (float32) of raw pairs."""
    fm, fs = F.double().mean(1), F.double().std(1)
    tm, ts = T.double().mean(1), T.double().std(1)
    fs = torch.where(fs > 0, fs, torch.ones_like(fs))
    ts = torch.where(ts > 0, ts, torch.ones_like(ts))
    return torch.cat([fm, fs, tm, ts]).float().cpu().numpy()


def standardise(F, T, norm):
    nm = torch.from_numpy(norm).to(F.device)
    Fs = ((F - nm[:NIN, None]) / nm[NIN:2 * NIN, None]).contiguous()
    Ts = ((T - nm[2 * NIN:2 * NIN + NOUT, None]) / nm[2 * NIN + NOUT:, None]).contiguous()
    return Fs, Ts


def implicit_learned(P, gas, rad, uu, pp, g20, G20, dt, global_time=0.0, maxiter=0):
    """koral_hip_implicit_batch_learned in place on uu/pp; (flags, info).
    rad may be None (gas primitives only); maxiter > 0 caps the learned
    attempt's iterations before the plain chain."""
    n = pp.shape[1]
    flags = torch.zeros((7, n), dtype=torch.int32, device=pp.device)
    info = torch.zeros((6, n), dtype=torch.int32, device=pp.device)
    check(gas.L.koral_hip_implicit_batch_learned(ctypes.byref(P), gas.h, rad.h if rad is not None else None, n,
                                                 _ptr(_chk(uu)), _ptr(_chk(pp)), _ptr(_chk(g20)),
                                                 _ptr(_chk(G20)), _ptr(_chk(dt)), float(global_time),
                                                 int(maxiter), _ptr(flags), _ptr(info), _stream()),
          "implicit_learned")
    return flags, info


def implicit_learned_gated(P, gas, rad, gate, uu, pp, g20, G20, dt, global_time=0.0, maxiter=0, thr=0.0):
    """koral_hip_implicit_batch_learned_gated: cells the gate scores below
    thr run the plain chain at once; (flags, info)."""
    n = pp.shape[1]
    flags = torch.zeros((7, n), dtype=torch.int32, device=pp.device)
    info = torch.zeros((6, n), dtype=torch.int32, device=pp.device)
    check(gas.L.koral_hip_implicit_batch_learned_gated(ctypes.byref(P), gas.h, rad.h if rad is not None else None,
                                                       gate.h, float(thr), n, _ptr(_chk(uu)), _ptr(_chk(pp)),
                                                       _ptr(_chk(g20)), _ptr(_chk(G20)), _ptr(_chk(dt)),
                                                       float(global_time), int(maxiter), _ptr(flags), _ptr(info),
                                                       _stream()), "implicit_learned_gated")
    return flags, info


def gate_pairs(P, pp0, dt, info_learned):
    """Raw features [11][m] of every cell and gate targets [4][m]: +1 in the
    first output where the learned tries solved the cell (info[1] = 6), -1
    where they did not (the other outputs 0)."""
    L = _sig(lib())
    n = pp0.shape[1]
    F = torch.empty((NIN, n), dtype=torch.float32, device=pp0.device)
    check(L.koral_hip_mlp_features(ctypes.byref(P), None, n, _ptr(_chk(pp0)), _ptr(_chk(dt)), _ptr(F),
                                   _stream()), "mlp_features")
    T = torch.zeros((NOUT, n), dtype=torch.float32, device=pp0.device)
    T[0] = torch.where(info_learned[1] == 6, 1.0, -1.0)
    keep = torch.isfinite(F).all(0)
    return F[:, keep].contiguous(), T[:, keep].contiguous()


def implicit_from(P, gxg, gxr, uu, pp, g20, G20, dt, global_time=0.0, maxiter=0):
    """koral_hip_implicit_batch_from: the learned first attempt from the
    given starts gxg, gxr [4][n] (gxr may be None)."""
    L = _sig(lib())
    n = pp.shape[1]
    flags = torch.zeros((7, n), dtype=torch.int32, device=pp.device)
    info = torch.zeros((6, n), dtype=torch.int32, device=pp.device)
    check(L.koral_hip_implicit_batch_from(ctypes.byref(P), n, _ptr(_chk(uu)), _ptr(_chk(pp)), _ptr(_chk(g20)),
                                          _ptr(_chk(G20)), _ptr(_chk(dt)), float(global_time),
                                          _ptr(_chk(gxg)), _ptr(_chk(gxr)) if gxr is not None else None,
                                          int(maxiter), _ptr(flags), _ptr(info), _stream()), "implicit_from")
    return flags, info


def train(P, F, T, kind, epochs=4, lr=1e-3, seed=0, schedule=None):
    """A network of the kind fitted to raw pairs F, T; returns (net, loss
    history).  schedule: [(lr, epochs), ...] overrides (lr, epochs)."""
    norm = normalisation(F, T)
    Fs, Ts = standardise(F, T, norm)
    net = Net(kind=kind, seed=seed)
    net.set(init_params(seed), norm)
    hist = []
    for r, e in (schedule or [(lr, epochs)]):
        hist += net.fit(Fs, Ts, epochs=e, lr=r, seed=seed + len(hist))
    return net, hist
