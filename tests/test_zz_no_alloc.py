"""The allocation-free implicit entry point (koral_hip_implicit_batch_noalloc)
against koral_hip_implicit_batch, through the C-ABI: two calls in a row on the
same GPU, the second with a smaller batch, each bit-identical to the
allocating entry point on the same inputs (primitives, conserved state, flags
and per-cell info); a batch above the documented maximum is refused."""
import ctypes
import sys

import numpy as np
import pytest

from koral_amd import params as KP, problems as PR

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _noalloc(L):
    f = L.koral_hip_implicit_batch_noalloc
    v = ctypes.c_void_p
    f.argtypes = [v, ctypes.c_long, v, v, v, v, v, ctypes.c_double, v, v, v]
    f.restype = ctypes.c_int
    return f


def _solve(P, n, seed, entry):
    import torch
    from koral_amd import cell
    from koral_amd._lib import lib, check
    pp, g20, G20, dt = PR.implicit_batch_states(P, n, seed=seed)
    Pp, g, G = cell.to_dev(pp), cell.to_dev(g20), cell.to_dev(G20)
    dtd = torch.from_numpy(dt).cuda()
    U = cell.p2u(P, Pp, g, G)
    flags = torch.zeros((7, n), dtype=torch.int32, device="cuda")
    info = torch.zeros((6, n), dtype=torch.int32, device="cuda")
    fn = _noalloc(lib()) if entry == "noalloc" else lib().koral_hip_implicit_batch
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    rc = fn(ctypes.byref(P), n, p(U), p(Pp), p(g), p(G), p(dtd), 0.0, p(flags), p(info),
            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    check(rc, entry)
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in (Pp, U, flags, info)]


def test_noalloc_matches_allocating_entry():
    import torch
    assert torch.cuda.is_available()
    P = KP.params_implicit_batch()
    cases = ((40000, 71), (9000, 72))
    # the two allocation-free calls back to back (the second, smaller one
    # finds the first one's lists and counters in the shared array)
    got = []
    for n, seed in cases:
        _say("implicit_batch_noalloc: n=%d seed=%d" % (n, seed))
        got.append(_solve(P, n, seed, "noalloc"))
    for (n, seed), a in zip(cases, got):
        _say("implicit_batch (allocating) for comparison: n=%d seed=%d" % (n, seed))
        b = _solve(P, n, seed, "alloc")
        for x, y, what in zip(a, b, ("p", "u", "flags", "info")):
            assert np.array_equal(x, y, equal_nan=True), what
        solved = (a[3][1] >= 0).mean()
        _say("  identical; %.4f of the cells solved" % solved)
        assert solved > 0.9


def test_noalloc_refuses_batch_above_maximum():
    import torch
    from koral_amd._lib import lib
    assert torch.cuda.is_available()
    P = KP.params_implicit_batch()
    n = (1 << 20) + 1
    null = ctypes.c_void_p(0)
    # refused on the host before anything is launched or dereferenced
    rc = _noalloc(lib())(ctypes.byref(P), n, null, null, null, null, null, 0.0, null, null, null)
    assert rc == -1
    assert b"KORAL_HIP_NOALLOC_MAX_CELLS" in lib().koral_hip_last_error()
