"""Host-side driver of the device grid operators (the Python mirror of ko.c's
outer loop around the C-ABI in include/koral_hip.h).

    dg = DeviceGrid(P, spec)          # uploads metric tables, allocates state
    dg.upload(p_aos)                  # KORAL host layout [z, y, x, nv]
    dg.init()                         # u = p2u(p), set_bc, first dt
    dg.step()                         # one RK2IMEX step, dt from the last one
"""
import ctypes

import numpy as np

from ._lib import lib, check

_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)


class GridTune(ctypes.Structure):
    """koral_grid_tune: the launch choices of a grid (include/koral_hip.h)."""
    _fields_ = [("block", ctypes.c_int * 32), ("fused", ctypes.c_int * 3), ("floor_wave", ctypes.c_int),
                ("update_pre", ctypes.c_int), ("comb_nv", ctypes.c_int), ("u2p_budget", ctypes.c_int), ("u2p_group", ctypes.c_int),
                ("floor_nw", ctypes.c_long),
                ("u2p_nw", ctypes.c_long)]


class GridDesc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("ng", ctypes.c_int),
                ("bc", ctypes.c_int * 6), ("xb", _dp * 3),
                ("gc", _dp), ("Gc", _dp), ("kr", _dp), ("gfx", _dp), ("Gfx", _dp), ("gfy", _dp), ("Gfy", _dp)]


def _sig(L):
    if getattr(L, "_grid_sigs", False):
        return L
    i = ctypes.c_int
    d = ctypes.c_double
    for name, args in {
        "koral_hip_grid_create": [_vp, _vp, ctypes.POINTER(_vp)],
        "koral_hip_grid_destroy": [_vp],
        "koral_hip_grid_set_stream": [_vp, _vp],
        "koral_hip_grid_upload": [_vp, _vp, _vp, _vp],
        "koral_hip_grid_download": [_vp, _vp, _vp, _vp],
        "koral_hip_grid_device_state": [_vp, _vp, _vp, _vp],
        "koral_hip_grid_download_work": [_vp, i, _vp],
        "koral_hip_grid_timing": [_vp, i, _vp, _vp, _vp, i],
        "koral_hip_grid_timing_mask": [_vp, ctypes.c_uint],
        "koral_hip_grid_p2u": [_vp],
        "koral_hip_set_bc": [_vp],
        "koral_hip_tstepden": [_vp, _vp],
        "koral_hip_op_explicit": [_vp, d, d, _vp],
        "koral_hip_calc_u2p": [_vp],
        "koral_hip_do_correct": [_vp],
        "koral_hip_op_implicit": [_vp, d, d],
        "koral_hip_step_rk2imex": [_vp, d, d, _vp, _vp],
        "koral_hip_run_rk2imex": [_vp, i, _vp, _vp, _vp],
        "koral_hip_grid_use_graph": [_vp, i],
        "koral_hip_grid_write_restart": [_vp, ctypes.c_char_p, i, i, d, i],
        "koral_hip_grid_read_restart": [_vp, ctypes.c_char_p, i, _vp, _vp, _vp],
        "koral_hip_grid_get_tune": [_vp, _vp],
        "koral_hip_grid_deferred_counts": [_vp, _vp],
        "koral_hip_grid_set_learned": [_vp, _vp, _vp, _vp, ctypes.c_float, ctypes.c_int],
        "koral_hip_grid_download_guess": [_vp, _vp, _vp],
        "koral_hip_grid_set_tune": [_vp, _vp],
        "koral_hip_grid_autotune": [_vp, d, d, i, ctypes.c_char_p, i],
    }.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = i
    L.koral_hip_grid_op_name.argtypes = [i]
    L.koral_hip_grid_op_name.restype = ctypes.c_char_p
    L._grid_sigs = True
    return L


NKINDS = 32  # at least the library's operator kinds; unnamed kinds are skipped


class DeviceGrid:
    def __init__(self, P, spec, stream=None):
        self.P, self.spec = P, spec
        self.L = _sig(lib())
        t = spec.tables
        self._keep = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in t.items()}
        self._xb = [np.ascontiguousarray(x, dtype=np.float64) for x in spec.xb]
        D = GridDesc()
        D.nx, D.ny, D.nz, D.ng = spec.nx, spec.ny, spec.nz, spec.ng
        for k in range(6):
            D.bc[k] = spec.bc[k]
        for k in range(3):
            D.xb[k] = self._xb[k].ctypes.data_as(_dp)
        for k in ("gc", "Gc", "kr", "gfx", "Gfx", "gfy", "Gfy"):
            setattr(D, k, self._keep[k].ctypes.data_as(_dp))
        self.desc = D
        h = _vp()
        check(self.L.koral_hip_grid_create(ctypes.byref(P), ctypes.byref(D), ctypes.byref(h)), "grid_create")
        self.h = h
        if stream is not None:
            check(self.L.koral_hip_grid_set_stream(self.h, _vp(stream)), "set_stream")
        self.t = 0.0
        self.dt = None
        self.counters = np.zeros(6, np.int32)

    def close(self):
        if self.h:
            self.L.koral_hip_grid_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def shape_aos(self):
        s = self.spec
        return (s.sz, s.sy, s.sx, self.P.nv)

    def upload(self, p_aos, u_aos=None, flags=None):
        p = np.ascontiguousarray(p_aos, dtype=np.float64)
        u = None if u_aos is None else np.ascontiguousarray(u_aos, dtype=np.float64)
        f = None if flags is None else np.ascontiguousarray(flags, dtype=np.int32)
        check(self.L.koral_hip_grid_upload(self.h, p.ctypes.data_as(_vp),
                                           None if u is None else u.ctypes.data_as(_vp),
                                           None if f is None else f.ctypes.data_as(_vp)), "grid_upload")

    def download(self):
        p = np.zeros(self.shape_aos)
        u = np.zeros(self.shape_aos)
        f = np.zeros(self.shape_aos[:-1] + (7,), np.int32)
        check(self.L.koral_hip_grid_download(self.h, p.ctypes.data_as(_vp), u.ctypes.data_as(_vp),
                                             f.ctypes.data_as(_vp)), "grid_download")
        return p, u, f

    def work(self):
        """Wave speeds [z, y, x, 18] and face fluxes after flux-CT per
        direction ([z, y, x+1, nv], [z, y+1, x, nv], [z+1, y, x, nv]; None for
        a collapsed direction) of the last op_explicit."""
        s = self.spec
        sx, sy, sz, nv = s.sx, s.sy, s.sz, self.P.nv
        ws = np.zeros((18, sz, sy, sx))
        check(self.L.koral_hip_grid_download_work(self.h, 0, ws.ctypes.data_as(_vp)), "download_work")
        shapes = [(nv, sz, sy, sx + 1), (nv, sz, sy + 1, sx), (nv, sz + 1, sy, sx)]
        n = [s.nx, s.ny, s.nz]
        fl = []
        for d in range(3):
            if n[d] == 1:
                fl.append(None)
                continue
            f = np.zeros(shapes[d])
            check(self.L.koral_hip_grid_download_work(self.h, d + 1, f.ctypes.data_as(_vp)), "download_work")
            fl.append(np.ascontiguousarray(np.moveaxis(f, 0, -1)))
        return np.ascontiguousarray(np.moveaxis(ws, 0, -1)), fl

    def timing_only(self, names=None):
        """Time only the named operators (all when names is None)."""
        mask = 0xffffffff
        if names is not None:
            mask = 0
            for k in range(NKINDS):
                nm = self.L.koral_hip_grid_op_name(k)
                if nm and nm.decode() in names:
                    mask |= 1 << k
        check(self.L.koral_hip_grid_timing_mask(self.h, mask), "grid_timing_mask")

    def timing(self, enable=True, read=True):
        """Per-operator launch timing.  Returns {name: (ms, threads, launches)}
        for the launches since the previous call (read=True), then records
        the following launches if enable."""
        if not read:
            check(self.L.koral_hip_grid_timing(self.h, int(enable), None, None, None, 0), "grid_timing")
            return {}
        ms = np.zeros(NKINDS)
        units = np.zeros(NKINDS, np.int64)
        nl = np.zeros(NKINDS, np.int32)
        check(self.L.koral_hip_grid_timing(self.h, int(enable), ms.ctypes.data_as(_vp), units.ctypes.data_as(_vp),
                                           nl.ctypes.data_as(_vp), NKINDS), "grid_timing")
        return {self.L.koral_hip_grid_op_name(k).decode(): (float(ms[k]), int(units[k]), int(nl[k]))
                for k in range(NKINDS) if nl[k] and self.L.koral_hip_grid_op_name(k)}

    def p2u(self):
        check(self.L.koral_hip_grid_p2u(self.h), "grid_p2u")

    def set_bc(self):
        check(self.L.koral_hip_set_bc(self.h), "set_bc")

    def tstepden(self):
        v = ctypes.c_double(0.)
        check(self.L.koral_hip_tstepden(self.h, ctypes.byref(v)), "tstepden")
        return v.value

    def explicit(self, t, dt):
        v = ctypes.c_double(0.)
        check(self.L.koral_hip_op_explicit(self.h, t, dt, ctypes.byref(v)), "op_explicit")
        return v.value

    def u2p(self):
        check(self.L.koral_hip_calc_u2p(self.h), "calc_u2p")

    def do_correct(self):
        check(self.L.koral_hip_do_correct(self.h), "do_correct")

    def implicit(self, t, dt):
        check(self.L.koral_hip_op_implicit(self.h, t, dt), "op_implicit")

    def init(self):
        """State as left by KORAL's initialisation: u = p2u(p), ghosts filled,
        first dt from the wave speeds."""
        self.p2u()
        self.set_bc()
        self.dt = 1. / self.tstepden()

    def write_restart(self, folder, nout1, nout2=0, problem=0):
        """KORAL restart files res%04d.head/.dat of the current state at self.t."""
        check(self.L.koral_hip_grid_write_restart(self.h, str(folder).encode(), int(nout1), int(nout2),
                                                  float(self.t), int(problem)), "write_restart")

    def read_restart(self, folder, nout1=-1):
        """Reads res%04d (or reslast for nout1 < 0); returns (found, next file
        number) and sets self.t from the header."""
        t = ctypes.c_double(0.)
        n1, n2 = ctypes.c_int(0), ctypes.c_int(0)
        r = self.L.koral_hip_grid_read_restart(self.h, str(folder).encode(), int(nout1), ctypes.byref(t),
                                               ctypes.byref(n1), ctypes.byref(n2))
        if r == 1:
            return False, 0
        check(r, "read_restart")
        self.t = t.value
        return True, n1.value

    def set_learned(self, gas, rad=None, gate=None, thr=0., maxiter=16):
        """Learned Newton starts for the implicit operator (koral_amd.learn.Net
        networks of kinds gas, radiation, gate; gas None: the plain solve).
        The networks must stay alive while the grid uses them."""
        h = lambda net: net.h if net is not None else None
        check(self.L.koral_hip_grid_set_learned(self.h, h(gas), h(rad), h(gate), float(thr), int(maxiter)),
              "grid_set_learned")
        self._nets = (gas, rad, gate)

    def download_guess(self):
        """The starts of the last learned implicit operator: gas and radiation
        [4][n] over the interior cells in x-fastest order."""
        n = self.spec.nx * self.spec.ny * self.spec.nz
        gxg = np.zeros((4, n))
        gxr = np.zeros((4, n))
        check(self.L.koral_hip_grid_download_guess(self.h, gxg.ctypes.data_as(_vp), gxr.ctypes.data_as(_vp)),
              "grid_download_guess")
        return gxg, gxr

    def deferred_counts(self):
        """[x1, x2, x3 floor faces of the last explicit operator, cells left
        to the full calc_primitives by the last u2p]."""
        c = np.zeros(4, np.uint32)
        check(self.L.koral_hip_grid_deferred_counts(self.h, c.ctypes.data_as(_vp)), "deferred_counts")
        return c.tolist()

    def get_tune(self):
        t = GridTune()
        check(self.L.koral_hip_grid_get_tune(self.h, ctypes.byref(t)), "get_tune")
        return t

    def set_tune(self, t):
        check(self.L.koral_hip_grid_set_tune(self.h, ctypes.byref(t)), "set_tune")

    def autotune(self, steps=3):
        """Searches the launch choices on the current state at (self.t,
        self.dt) and keeps the fastest; the state is left unchanged.  Returns
        the search log as a list of (name, value, ms per step)."""
        buf = ctypes.create_string_buffer(1 << 16)
        check(self.L.koral_hip_grid_autotune(self.h, float(self.t), float(self.dt), int(steps), buf, len(buf)),
              "autotune")
        rows = []
        for line in buf.value.decode().splitlines():
            parts = line.split()
            if len(parts) == 2:
                rows.append((parts[0], None, float(parts[1])))
            elif len(parts) == 3:
                rows.append((parts[0], int(parts[1]), float(parts[2])))
            elif len(parts) == 4:
                rows.append((parts[0] + " " + parts[1], int(parts[2]), float(parts[3])))
        return rows

    def use_graph(self, enable=True):
        """Replay one captured step per step in run() (needs a stream)."""
        check(self.L.koral_hip_grid_use_graph(self.h, int(enable)), "grid_use_graph")

    def run(self, nsteps):
        """nsteps steps with the time kept on the device (dt of each step
        from the previous step's time-step density, as step() does); one
        synchronisation at the end.  self.counters sums the steps."""
        t = ctypes.c_double(self.t)
        dt = ctypes.c_double(self.dt)
        c = np.zeros(6, np.int32)
        check(self.L.koral_hip_run_rk2imex(self.h, int(nsteps), ctypes.byref(t), ctypes.byref(dt),
                                           c.ctypes.data_as(_vp)), "run_rk2imex")
        self.t, self.dt = t.value, dt.value
        self.counters = c
        return self.dt

    def step(self, sync=True):
        """One RK2IMEX step with dt from the previous step (ko.c:399)."""
        tsd = ctypes.c_double(0.)
        c = self.counters
        check(self.L.koral_hip_step_rk2imex(self.h, self.t, self.dt,
                                            ctypes.byref(tsd) if sync else None,
                                            c.ctypes.data_as(_vp) if sync else None), "step_rk2imex")
        self.t += self.dt
        if sync:
            self.dt = 1. / tsd.value
        return self.dt
