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
        L.oracle_u2p_solver_batch_from.argtypes = [vp, ctypes.c_long, _dp, _dp, _dp, _dp, ctypes.c_int, vp,
                                                   _ip, _ip, _ip]
        L.oracle_u2p_rad_batch.argtypes = [vp, ctypes.c_long, _dp, _dp, _dp, _dp, _ip, _ip]
        L.oracle_floors_batch.argtypes = [vp, ctypes.c_long, _dp, _dp, _dp, _ip]
        L.oracle_Gi_batch.argtypes = [vp, ctypes.c_long, _dp, _dp, _dp, ctypes.c_int, _dp]
        L.oracle_ff_Rtt_batch.argtypes = [vp, ctypes.c_long, _dp, _dp, _dp, _dp]
        L.oracle_implicit_batch.argtypes = [vp, ctypes.c_long, _dp, _dp, _dp, _dp, _dp, ctypes.c_double,
                                            _ip, _ip, ctypes.c_int]
        L.oracle_implicit_batch_guess.argtypes = [vp, ctypes.c_long, _dp, _dp, _dp, _dp, _dp, ctypes.c_double,
                                                  _ip, _ip, ctypes.c_int, _dp, _dp, ctypes.c_int]
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


def u2p_solver_from(P, uu, pp_guess, g20, G20, W0=None, etype=0):
    """The W inversion with a learned W start per point (None: the plain
    start for all): (pp, ret, iters, guessed)."""
    uu = np.ascontiguousarray(uu, dtype=np.float64)
    pp = np.array(pp_guess, dtype=np.float64, order="C")
    n = pp.shape[0]
    ret = np.zeros(n, np.int32)
    iters = np.zeros(n, np.int32)
    guessed = np.zeros(n, np.int32)
    w = None if W0 is None else np.ascontiguousarray(W0, dtype=np.float64)
    lib().oracle_u2p_solver_batch_from(_P(P), n, _d(uu), _d(pp), _d(np.ascontiguousarray(g20)),
                                       _d(np.ascontiguousarray(G20)), etype,
                                       None if w is None else w.ctypes.data_as(ctypes.c_void_p),
                                       _i(ret), _i(iters), _i(guessed))
    return pp, ret, iters, guessed


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


def implicit(P, uu, pp, g20, G20, dt, global_time=0.0, nthreads=1, gx=None, maxiter=0, gxr=None):
    """implicit_lab_rad_source_term on independent cells.
    Returns (uu, pp, flags[n,7], info[n,6]) with info columns
    iters, attempt, whichprim, whicheq, whichframe, total_iters.
    gx [n, 4] (u, u~^1..3), when given, is the start of a first attempt
    (energy equation, lab frame, at most maxiter iterations when maxiter > 0)
    on the gas primitives, or, with gxr [n, 4] (E, F~^1..3) given, on the
    radiation primitives from gxr where the completed gas start asks for
    them; reported as attempt 6 when it succeeds."""
    uu = np.array(uu, dtype=np.float64, order="C")
    pp = np.array(pp, dtype=np.float64, order="C")
    n = pp.shape[0]
    flags = np.zeros((n, 7), np.int32)
    info = np.zeros((n, 6), np.int32)
    dt = np.ascontiguousarray(dt, dtype=np.float64)
    ptrs = []
    for a in (gx, gxr):
        if a is None:
            ptrs.append(None)
            continue
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == (n, 4)
        ptrs.append(a)
    lib().oracle_implicit_batch_guess(_P(P), n, _d(uu), _d(pp), _d(np.ascontiguousarray(g20)),
                                      _d(np.ascontiguousarray(G20)), _d(dt), global_time, _i(flags), _i(info),
                                      nthreads, _d(ptrs[0]) if ptrs[0] is not None else None,
                                      _d(ptrs[1]) if ptrs[1] is not None else None, int(maxiter))
    return uu, pp, flags, info


def inverse44(a):
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(16)
    ia = np.zeros(16)
    r = lib().oracle_inverse44(_d(a), _d(ia))
    return r, ia.reshape(4, 4)


# ---- grid-level operators ----
class OGrid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("ng", ctypes.c_int),
                ("ngx", ctypes.c_int), ("ngy", ctypes.c_int), ("ngz", ctypes.c_int),
                ("sx", ctypes.c_int), ("sy", ctypes.c_int), ("sz", ctypes.c_int),
                ("bc", ctypes.c_int * 6),
                ("xb", _dp * 3),
                ("gc", _dp), ("Gc", _dp), ("kr", _dp), ("gfx", _dp), ("Gfx", _dp), ("gfy", _dp), ("Gfy", _dp)]


class Grid:
    """Wraps a koral_amd.grid.GridSpec for the oracle (keeps arrays alive)."""

    def __init__(self, spec):
        self.spec = spec
        self.t = spec.aos_tables()
        self.xb = [np.ascontiguousarray(x, dtype=np.float64) for x in spec.xb]
        g = OGrid()
        for k in ("nx", "ny", "nz", "ng", "ngx", "ngy", "ngz", "sx", "sy", "sz"):
            setattr(g, k, getattr(spec, k))
        for i in range(6):
            g.bc[i] = spec.bc[i]
        for d in range(3):
            g.xb[d] = _d(self.xb[d])
        for k in ("gc", "Gc", "kr", "gfx", "Gfx", "gfy", "Gfy"):
            setattr(g, k, _d(self.t[k]))
        self.c = g


def _grid_lib():
    L = lib()
    if not hasattr(L, "_grid_sigs"):
        vp = ctypes.c_void_p
        L.oracle_grid_set_bc.argtypes = [vp, vp, _dp, _dp]
        L.oracle_grid_tstepden.argtypes = [vp, vp, _dp]
        L.oracle_grid_tstepden.restype = ctypes.c_double
        L.oracle_grid_explicit.argtypes = [vp, vp, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp]
        L.oracle_grid_explicit_parts.argtypes = [vp, vp, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp, _dp,
                                                 _dp, _dp, _dp]
        L.oracle_grid_u2p.argtypes = [vp, vp, _dp, _dp, _ip, _ip]
        L.oracle_grid_cell_fixup.argtypes = [vp, vp, _dp, _dp, _ip, ctypes.c_int]
        L.oracle_grid_implicit.argtypes = [vp, vp, ctypes.c_double, ctypes.c_double, _dp, _dp, _ip, _ip]
        L.oracle_grid_implicit_guess.argtypes = [vp, vp, ctypes.c_double, ctypes.c_double, _dp, _dp, _ip, _ip,
                                                 _dp, _dp, ctypes.c_int]
        L.oracle_grid_step.argtypes = [vp, vp, ctypes.c_double, ctypes.c_double, _dp, _dp, _ip, _dp, _ip]
        L.oracle_grid_do_correct.argtypes = [vp, vp, _dp, _dp]
        L.oracle_grid_set_threads.argtypes = [ctypes.c_int]
        L._grid_sigs = True
    return L


def grid_set_threads(n):
    """OpenMP threads of the grid operators' independent loops (1: serial;
    results do not depend on it)."""
    _grid_lib().oracle_grid_set_threads(int(n))


class GridState:
    """AoS state u, p [z, y, x, nv] and flags [z, y, x, 7] on an oracle grid."""

    def __init__(self, P, grid, p_aos, u_aos=None, flags=None):
        self.P, self.grid = P, grid
        self.p = np.ascontiguousarray(p_aos, dtype=np.float64).copy()
        self.u = np.zeros_like(self.p) if u_aos is None else np.ascontiguousarray(u_aos, dtype=np.float64).copy()
        self.flags = np.zeros(self.p.shape[:-1] + (7,), np.int32) if flags is None else flags.copy()

    def init_u(self):
        """u = p2u(p) in every cell (as set_initial_profile does)."""
        gc, Gc = self.grid.spec.cell_tables()
        nv = self.P.nv
        self.u[...] = p2u(self.P, self.p.reshape(-1, nv), np.ascontiguousarray(gc).reshape(-1, 20),
                          np.ascontiguousarray(Gc).reshape(-1, 20)).reshape(self.p.shape)

    def set_bc(self):
        _grid_lib().oracle_grid_set_bc(_P(self.P), ctypes.byref(self.grid.c), _d(self.u), _d(self.p))

    def tstepden(self):
        return _grid_lib().oracle_grid_tstepden(_P(self.P), ctypes.byref(self.grid.c), _d(self.p))

    def explicit(self, t, dt):
        tm = np.array([-1.0])
        _grid_lib().oracle_grid_explicit(_P(self.P), ctypes.byref(self.grid.c), t, dt, _d(self.u), _d(self.p),
                                         _d(tm))
        return tm[0]

    def explicit_parts(self, t, dt):
        """op_explicit that also returns the wave speeds [z, y, x, 18] and the
        face fluxes after flux-CT per direction ([z, y, x+1, nv],
        [z, y+1, x, nv], [z+1, y, x, nv]; None for a collapsed direction)."""
        sp = self.grid.spec
        sx, sy, sz, nv = sp.sx, sp.sy, sp.sz, self.P.nv
        ws = np.zeros((sz, sy, sx, 18))
        shapes = [(sz, sy, sx + 1, nv), (sz, sy + 1, sx, nv), (sz + 1, sy, sx, nv)]
        n = [sp.nx, sp.ny, sp.nz]
        fl = [np.zeros(shapes[d]) if n[d] > 1 else None for d in range(3)]
        tm = np.array([-1.0])
        ptr = [_d(f) if f is not None else None for f in fl]
        _grid_lib().oracle_grid_explicit_parts(_P(self.P), ctypes.byref(self.grid.c), t, dt, _d(self.u),
                                               _d(self.p), _d(tm), _d(ws), ptr[0], ptr[1], ptr[2])
        return tm[0], ws, fl

    def u2p(self):
        c = np.zeros(3, np.int32)
        _grid_lib().oracle_grid_u2p(_P(self.P), ctypes.byref(self.grid.c), _d(self.u), _d(self.p),
                                    _i(self.flags), _i(c))
        return c

    def cell_fixup(self, ftype):
        """cell_fixup (finite.c:4468-4640) of flag index ftype (1 RADIMPFIXUP,
        2 HDFIXUP, 3 RADFIXUP) on the current state and flags."""
        _grid_lib().oracle_grid_cell_fixup(_P(self.P), ctypes.byref(self.grid.c), _d(self.u), _d(self.p),
                                           _i(self.flags), int(ftype))

    def do_correct(self):
        """do_correct: the polar-axis correction when P.polar_nc > 0."""
        _grid_lib().oracle_grid_do_correct(_P(self.P), ctypes.byref(self.grid.c), _d(self.u), _d(self.p))

    def implicit(self, t, dt, gxg=None, gxr=None, maxiter=0):
        """op_implicit; with gxg (and gxr) [4][n] learned starts over the
        interior cells in x-fastest order, each cell solved from them as
        o_solve_implicit_lab_guess does (maxiter the learned budget)."""
        c = np.zeros(3, np.int32)
        if gxg is None:
            _grid_lib().oracle_grid_implicit(_P(self.P), ctypes.byref(self.grid.c), t, dt, _d(self.u), _d(self.p),
                                             _i(self.flags), _i(c))
            return c
        gg = np.ascontiguousarray(np.asarray(gxg, np.float64).T)
        gr = None if gxr is None else np.ascontiguousarray(np.asarray(gxr, np.float64).T)
        _grid_lib().oracle_grid_implicit_guess(_P(self.P), ctypes.byref(self.grid.c), t, dt, _d(self.u),
                                               _d(self.p), _i(self.flags), _i(c), _d(gg),
                                               _d(gr) if gr is not None else None, int(maxiter))
        return c

    def step(self, t, dt):
        """One RK2IMEX step; returns the next tstepdenmax and counters
        [implicit successes, implicit iterations, implicit failures,
         hd fixups, rad fixups, entropy inversions]."""
        tm = np.array([-1.0])
        c = np.zeros(6, np.int32)
        _grid_lib().oracle_grid_step(_P(self.P), ctypes.byref(self.grid.c), t, dt, _d(self.u), _d(self.p),
                                     _i(self.flags), _d(tm), _i(c))
        return tm[0], c
