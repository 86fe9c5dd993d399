"""Loader for libkoral_hip.so (the C-ABI in include/koral_hip.h).

The product has no CPU fallback: if the library is missing or a call fails,
this module raises.  torch is used only to own device memory and streams.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libkoral_hip.so")

_lib = None
_vp = ctypes.c_void_p


class KoralError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise ImportError("libkoral_hip.so is not built (run python koral_amd/build.py); "
                          "the HIP path has no CPU fallback")
    # torch ships its own HIP runtime under the same soname; loading it first
    # makes this library bind to that one, so a process never holds two HIP
    # runtimes (the second would leave torch without a device).
    import torch  # noqa: F401
    L = ctypes.CDLL(LIBPATH)
    L.koral_hip_last_error.restype = ctypes.c_char_p
    L.koral_hip_version.restype = ctypes.c_int
    L.koral_hip_params_size.restype = ctypes.c_int
    L_long = ctypes.c_long
    sig = {
        "koral_hip_p2u_batch": [_vp, L_long, _vp, _vp, _vp, _vp, _vp],
        "koral_hip_u2p_batch": [_vp, L_long, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
        "koral_hip_u2p_solver_batch": [_vp, L_long, _vp, _vp, _vp, _vp, ctypes.c_int, _vp, _vp],
        "koral_hip_u2p_rad_batch": [_vp, L_long, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
        "koral_hip_Gi_batch": [_vp, L_long, _vp, _vp, _vp, ctypes.c_int, _vp, _vp],
        "koral_hip_implicit_batch": [_vp, L_long, _vp, _vp, _vp, _vp, _vp, ctypes.c_double, _vp, _vp, _vp],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    _lib = L
    return L


def check(rc, what):
    if rc != 0:
        msg = lib().koral_hip_last_error().decode()
        raise KoralError("%s failed (%d): %s" % (what, rc, msg))
