"""The device's FP64 square root, division, reciprocal and the portable
exp/log/pow of kd_math.h, compiled as the library is, against the host's
IEEE results bit for bit: the device code and the CPU oracle must round
every operation alike for the parity tests' bit-exact results."""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "tests", "gpu", "libarith_check.so")


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def arith():
    import torch
    assert torch.cuda.is_available()
    assert os.path.exists(LIB), "tests/gpu/libarith_check.so not built (__graft_entry__.build())"
    L = ctypes.CDLL(LIB)
    v = ctypes.c_void_p
    L.arith_check.argtypes = [ctypes.c_int, ctypes.c_long, v, v, v, v]
    L.arith_check.restype = ctypes.c_int

    def run(op, x, y):
        xd = torch.from_numpy(np.ascontiguousarray(x, np.float64)).cuda()
        yd = torch.from_numpy(np.ascontiguousarray(y, np.float64)).cuda()
        out = torch.empty_like(xd)
        assert L.arith_check(op, xd.numel(), v(xd.data_ptr()), v(yd.data_ptr()), v(out.data_ptr()),
                             v(torch.cuda.current_stream().cuda_stream)) == 0
        torch.cuda.synchronize()
        return out.cpu().numpy()
    return run


def _pos(rng, n, lo=-300, hi=300):
    return rng.uniform(1., 2., n) * np.exp2(rng.integers(lo, hi, n).astype(np.float64))


def _same(a, b):
    return (a == b) | (np.isnan(a) & np.isnan(b))


def test_sqrt_div_rcp_match_ieee(arith):
    rng = np.random.default_rng(5)
    n = 1 << 20
    x, y = _pos(rng, n), _pos(rng, n) * np.where(rng.random(n) < .5, -1., 1.)
    # squares of doubles and their neighbours: the hardest cases of rounding
    s = _pos(rng, n // 4, -150, 150)
    sq = np.concatenate([s * s, np.nextafter(s * s, 0.), np.nextafter(s * s, np.inf)])
    bad = {}
    for name, op, a, b, ref in (("sqrt", 0, x, x, np.sqrt(x)), ("sqrt of squares", 0, sq, sq, np.sqrt(sq)),
                                ("div", 1, x, y, x / y), ("rcp", 5, y, y, 1. / y)):
        got = arith(op, a, b)
        nb = int((~_same(got, ref)).sum())
        bad[name] = nb
        _say("%-16s %d of %d differ from IEEE" % (name, nb, len(a)))
    assert all(v == 0 for v in bad.values()), bad


def test_kd_math_matches_host(arith, tmp_path):
    exe = tmp_path / "check_kd_math"
    subprocess.run(["gcc", "-O2", "-std=gnu11", "-ffp-contract=off", "-I", os.path.join(ROOT, "koral_amd", "csrc"),
                    os.path.join(ROOT, "tests", "host", "check_kd_math.c"), "-o", str(exe), "-lm"], check=True)
    rng = np.random.default_rng(6)
    n = 1 << 18
    cases = [(2, 0, rng.uniform(-700., 700., n), np.zeros(n)),
             (3, 1, _pos(rng, n), np.zeros(n)),
             (4, 2, _pos(rng, n, -60, 60), rng.uniform(-8., 8., n))]
    bad = {}
    for op, hop, x, y in cases:
        recs = np.stack([np.full(n, float(hop)), x, y], 1)
        host = np.frombuffer(subprocess.run([str(exe)], input=recs.tobytes(), capture_output=True,
                                            check=True).stdout, np.float64).reshape(-1, 2)[:, 0]
        got = arith(op, x, y)
        bad[op] = int((~_same(got, host)).sum())
        _say("kd op %d: %d of %d differ from the host build" % (op, bad[op], n))
    assert all(v == 0 for v in bad.values()), bad
