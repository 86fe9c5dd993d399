"""The portable exp, log and pow of koral_amd/csrc/kd_math.h (called by both
the device code and the oracle) against 50-digit decimal references: at most
0.5 ulp plus a hair, correctly rounded on all but a few in 10^4 arguments,
and equal to the host libm (the one KORAL's CPU build calls) on all but a
few in 10^3; C99 special cases as the host libm returns them."""
import decimal
import math
import os
import subprocess
from decimal import Decimal as D

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def kdm(tmp_path_factory):
    exe = tmp_path_factory.mktemp("kdm") / "check_kd_math"
    subprocess.run(["gcc", "-O2", "-std=gnu11", "-ffp-contract=off", "-I", os.path.join(ROOT, "koral_amd", "csrc"),
                    os.path.join(ROOT, "tests", "host", "check_kd_math.c"), "-o", str(exe), "-lm"], check=True)

    def run(recs):
        buf = np.ascontiguousarray(recs, dtype=np.float64).tobytes()
        out = subprocess.run([str(exe)], input=buf, capture_output=True, check=True).stdout
        return np.frombuffer(out, np.float64).reshape(-1, 2)
    return run


def _ref(op, x, y):
    X = D(x)
    if op == 0:
        return X.exp()
    if op == 1:
        return X.ln()
    return (D(y) * X.ln()).exp()


def _records(seed, n):
    rng = np.random.default_rng(seed)
    recs = []
    for x in np.concatenate([rng.uniform(-740, 709, n), rng.uniform(-1, 1, n // 2), rng.uniform(-1e-3, 1e-3, n // 2)]):
        recs.append((0., x, 0.))
    for x in np.concatenate([np.exp(rng.uniform(-700, 700, n)), 1 + rng.uniform(-0.01, 0.01, n // 2),
                             1 + rng.uniform(-1e-6, 1e-6, n // 2)]):
        recs.append((1., x, 0.))
    # pow with the hot path's exponents (1/(Gamma-1), 1/(Gamma-1)+1 for
    # Gamma 5/3 and 13/9, 4 for sigma T^4) and random ones
    ys = [1.5, 2.5, 2.25, 3.25, 4.0, 2.0]
    for i, x in enumerate(np.exp(rng.uniform(-120, 120, 2 * n))):
        recs.append((2., x, ys[i % len(ys)] if i % 2 == 0 else rng.uniform(-5, 5)))
    return np.array(recs)


def test_accuracy_against_decimal(kdm):
    decimal.getcontext().prec = 50
    recs = _records(5, 4000)
    res = kdm(recs)
    stats = {}
    for (op, x, y), (a, b) in zip(recs, res):
        op = int(op)
        r = _ref(op, x, y)
        if not (D("2.3e-308") < abs(r) < D("1.7e308")):
            continue
        u = D(math.ulp(float(r)))
        e_kd = float(abs(D(a) - r) / u)
        s = stats.setdefault(op, [0, 0., 0, 0])
        s[0] += 1
        s[1] = max(s[1], e_kd)
        s[2] += a == float(r)
        s[3] += a == b
    for op, (n, worst, cr, same) in stats.items():
        name = ["exp", "log", "pow"][op]
        print("%s: n %d, max %.5f ulp, correctly rounded %.5f, equal to libm %.5f" % (name, n, worst, cr / n, same / n))
        assert worst < 0.5 + 2 ** -8, name
        assert cr / n > 0.9995, name
        assert same / n > 0.998, name


def test_special_cases_match_libm(kdm):
    inf, nan = math.inf, math.nan
    xs = [0., -0., 1., -1., 2., -2., 0.5, -0.5, inf, -inf, nan, 1e-310, 5e-324, 1.7e308]
    ys = [0., -0., 1., -1., 2., -2., 3., -3., 0.5, 2.25, inf, -inf, nan]
    recs = [(2., x, y) for x in xs for y in ys]
    recs += [(0., x, 0.) for x in [0., -0., inf, -inf, nan, 709.7, 709.79, -745.2, -740., 1e-300]]
    recs += [(1., x, 0.) for x in [0., -0., 1., -1., inf, nan, 5e-324, 1e-310, 1.7e308, 0.9999999999999999]]
    res = kdm(np.array(recs))
    for (op, x, y), (a, b) in zip(recs, res):
        same = (a == b and math.copysign(1., a) == math.copysign(1., b)) or (math.isnan(a) and math.isnan(b))
        assert same, (op, x, y, a, b)


def test_tables_regenerate_identically():
    """kd_math_tables.h is what its generator prints."""
    gen = os.path.join(ROOT, "koral_amd", "csrc", "gen_kd_math_tables.py")
    out = subprocess.run(["python3", gen], capture_output=True, text=True, check=True).stdout
    with open(os.path.join(ROOT, "koral_amd", "csrc", "kd_math_tables.h")) as f:
        assert f.read() == out
