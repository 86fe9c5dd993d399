"""Batched per-point operators on device tensors (host mirror of the
reference's per-cell functions, each launched over a whole batch).

Arrays are torch.float64 CUDA tensors in SoA layout: state [nv, n], metric
[20, n].  Every call runs on torch's current stream through the C-ABI in
include/koral_hip.h.
"""
import ctypes

import torch

from ._lib import lib, check


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _chk(t, dtype=torch.float64):
    if not (t.is_cuda and t.is_contiguous() and t.dtype == dtype):
        raise ValueError("expected a contiguous %s CUDA tensor" % dtype)
    return t


def p2u(P, pp, g20, G20):
    """p2u (p2u.c:40-49) of every point; returns uu [nv, n]."""
    n = pp.shape[1]
    _chk(pp), _chk(g20), _chk(G20)
    uu = torch.empty_like(pp)
    check(lib().koral_hip_p2u_batch(ctypes.byref(P), n, _ptr(pp), _ptr(uu), _ptr(g20), _ptr(G20), _stream()),
          "p2u")
    return uu


def u2p(P, uu, pp, g20, G20):
    """calc_primitives on independent points (in place on uu, pp).
    Returns (ret [n] int32, flags [7, n] int32)."""
    n = pp.shape[1]
    _chk(uu), _chk(pp), _chk(g20), _chk(G20)
    ret = torch.zeros(n, dtype=torch.int32, device=pp.device)
    flags = torch.zeros((7, n), dtype=torch.int32, device=pp.device)
    check(lib().koral_hip_u2p_batch(ctypes.byref(P), n, _ptr(uu), _ptr(pp), _ptr(g20), _ptr(G20), _ptr(ret),
                                    _ptr(flags), _stream()), "u2p")
    return ret, flags


def u2p_solver(P, uu, pp, g20, G20, etype=0):
    n = pp.shape[1]
    ret = torch.zeros(n, dtype=torch.int32, device=pp.device)
    check(lib().koral_hip_u2p_solver_batch(ctypes.byref(P), n, _ptr(_chk(uu)), _ptr(_chk(pp)), _ptr(_chk(g20)),
                                           _ptr(_chk(G20)), etype, _ptr(ret), _stream()), "u2p_solver")
    return ret


def u2p_rad(P, uu, pp, g20, G20):
    n = pp.shape[1]
    ret = torch.zeros(n, dtype=torch.int32, device=pp.device)
    corr = torch.zeros(n, dtype=torch.int32, device=pp.device)
    check(lib().koral_hip_u2p_rad_batch(ctypes.byref(P), n, _ptr(_chk(uu)), _ptr(_chk(pp)), _ptr(_chk(g20)),
                                        _ptr(_chk(G20)), _ptr(ret), _ptr(corr), _stream()), "u2p_rad")
    return ret, corr


def Gi(P, pp, g20, G20, labframe=1):
    n = pp.shape[1]
    out = torch.empty((4, n), dtype=torch.float64, device=pp.device)
    check(lib().koral_hip_Gi_batch(ctypes.byref(P), n, _ptr(_chk(pp)), _ptr(_chk(g20)), _ptr(_chk(G20)),
                                   labframe, _ptr(out), _stream()), "Gi")
    return out


def implicit(P, uu, pp, g20, G20, dt, global_time=0.0):
    """solve_implicit_lab on independent cells, in place on uu/pp.
    Returns (flags [7, n], info [6, n]) int32 tensors."""
    n = pp.shape[1]
    flags = torch.zeros((7, n), dtype=torch.int32, device=pp.device)
    info = torch.zeros((6, n), dtype=torch.int32, device=pp.device)
    check(lib().koral_hip_implicit_batch(ctypes.byref(P), n, _ptr(_chk(uu)), _ptr(_chk(pp)), _ptr(_chk(g20)),
                                         _ptr(_chk(G20)), _ptr(_chk(dt)), float(global_time), _ptr(flags),
                                         _ptr(info), _stream()), "implicit")
    return flags, info


def to_dev(a):
    """AoS numpy [n, k] -> SoA CUDA tensor [k, n]."""
    import numpy as np
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float64).T)).cuda()


def to_host(t):
    """SoA CUDA tensor [k, n] -> AoS numpy [n, k]."""
    return t.cpu().numpy().T.copy()
