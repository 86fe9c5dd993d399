"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU baseline.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        vp = ctypes.c_void_p
        L.oracle_params_size.restype = ctypes.c_int
        L.oracle_params_offset.restype = ctypes.c_int
        L.oracle_params_offset.argtypes = [ctypes.c_int]
        L.oracle_p2u_batch.argtypes = [vp, ctypes.c_long, _dp, _dp, _dp, _dp]
        L.oracle_u2p_batch.argtypes = [vp, ctypes.c_long, _dp, _dp, _dp, _dp, _ip, _ip]
        L.oracle_u2p_solver_batch.argtypes = [vp, ctypes.c_long, _dp, _dp, _dp, _dp, ctypes.c_int, _ip]
        L.oracle_u2p_rad_batch.argtypes = [vp, ctypes.c_long, _dp, _dp, _dp, _dp, _ip, _ip]
        L.oracle_floors_batch.argtypes = [vp, ctypes.c_long, _dp, _dp, _dp, _ip]
        L.oracle_Gi_batch.argtypes = [vp, ctypes.c_long, _dp, _dp, _dp, ctypes.c_int, _dp]
        L.oracle_ff_Rtt_batch.argtypes = [vp, ctypes.c_long, _dp, _dp, _dp, _dp]
        L.oracle_implicit_batch.argtypes = [vp, ctypes.c_long, _dp, _dp, _dp, _dp, _dp, ctypes.c_double,
                                            _ip, _ip, ctypes.c_int]
        L.oracle_inverse44.argtypes = [_dp, _dp]
        L.oracle_inverse44.restype = ctypes.c_int
        _lib = L
    return _lib


def _d(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _i(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def _P(P):
    return ctypes.byref(P)


def p2u(P, pp, g20, G20):
    pp = np.ascontiguousarray(pp, dtype=np.float64)
    n = pp.shape[0]
    uu = np.zeros_like(pp)
    lib().oracle_p2u_batch(_P(P), n, _d(pp), _d(uu), _d(np.ascontiguousarray(g20)), _d(np.ascontiguousarray(G20)))
    return uu


def u2p(P, uu, pp_guess, g20, G20):
    """calc_primitives on independent points: returns (uu_out, pp_out, ret, flags)."""
    uu = np.array(uu, dtype=np.float64, order="C")
    pp = np.array(pp_guess, dtype=np.float64, order="C")
    n = pp.shape[0]
    ret = np.zeros(n, np.int32)
    flags = np.zeros((n, 7), np.int32)
    lib().oracle_u2p_batch(_P(P), n, _d(uu), _d(pp), _d(np.ascontiguousarray(g20)),
                           _d(np.ascontiguousarray(G20)), _i(ret), _i(flags))
    return uu, pp, ret, flags


def u2p_solver(P, uu, pp_guess, g20, G20, etype=0):
    uu = np.ascontiguousarray(uu, dtype=np.float64)
    pp = np.array(pp_guess, dtype=np.float64, order="C")
    n = pp.shape[0]
    ret = np.zeros(n, np.int32)
    lib().oracle_u2p_solver_batch(_P(P), n, _d(uu), _d(pp), _d(np.ascontiguousarray(g20)),
                                  _d(np.ascontiguousarray(G20)), etype, _i(ret))
    return pp, ret


def u2p_rad(P, uu, pp_guess, g20, G20):
    uu = np.ascontiguousarray(uu, dtype=np.float64)
    pp = np.array(pp_guess, dtype=np.float64, order="C")
    n = pp.shape[0]
    ret = np.zeros(n, np.int32)
    corr = np.zeros(n, np.int32)
    lib().oracle_u2p_rad_batch(_P(P), n, _d(uu), _d(pp), _d(np.ascontiguousarray(g20)),
                               _d(np.ascontiguousarray(G20)), _i(ret), _i(corr))
    return pp, ret, corr


def floors(P, pp, g20, G20):
    pp = np.array(pp, dtype=np.float64, order="C")
    n = pp.shape[0]
    ret = np.zeros(n, np.int32)
    lib().oracle_floors_batch(_P(P), n, _d(pp), _d(np.ascontiguousarray(g20)), _d(np.ascontiguousarray(G20)),
                              _i(ret))
    return pp, ret


def Gi(P, pp, g20, G20, labframe=1):
    pp = np.ascontiguousarray(pp, dtype=np.float64)
    n = pp.shape[0]
    out = np.zeros((n, 4))
    lib().oracle_Gi_batch(_P(P), n, _d(pp), _d(np.ascontiguousarray(g20)), _d(np.ascontiguousarray(G20)),
                          labframe, _d(out))
    return out


def ff_Rtt(P, pp, g20, G20):
    pp = np.ascontiguousarray(pp, dtype=np.float64)
    n = pp.shape[0]
    out = np.zeros(n)
    lib().oracle_ff_Rtt_batch(_P(P), n, _d(pp), _d(np.ascontiguousarray(g20)), _d(np.ascontiguousarray(G20)),
                              _d(out))
    return out


def implicit(P, uu, pp, g20, G20, dt, global_time=0.0, nthreads=1):
    """implicit_lab_rad_source_term on independent cells.
    Returns (uu, pp, flags[n,7], info[n,6]) with info columns
    iters, attempt, whichprim, whicheq, whichframe, total_iters."""
    uu = np.array(uu, dtype=np.float64, order="C")
    pp = np.array(pp, dtype=np.float64, order="C")
    n = pp.shape[0]
    flags = np.zeros((n, 7), np.int32)
    info = np.zeros((n, 6), np.int32)
    dt = np.ascontiguousarray(dt, dtype=np.float64)
    lib().oracle_implicit_batch(_P(P), n, _d(uu), _d(pp), _d(np.ascontiguousarray(g20)),
                                _d(np.ascontiguousarray(G20)), _d(dt), global_time, _i(flags), _i(info),
                                nthreads)
    return uu, pp, flags, info


def inverse44(a):
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(16)
    ia = np.zeros(16)
    r = lib().oracle_inverse44(_d(a), _d(ia))
    return r, ia.reshape(4, 4)
