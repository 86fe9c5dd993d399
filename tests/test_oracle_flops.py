"""The operation count per implicit solve that the benchmark divides by its
measured time: the counting build of the oracle takes the same iterate
path as the oracle (same solved cells and iterations) and a fresh sample of
the seeded workload lands within a few per cent of the recorded figure."""
import os
import subprocess

import numpy as np
import pytest

from koral_amd import params as KP, problems as PR

HERE = os.path.dirname(os.path.abspath(__file__))
FLOPS = os.path.join(os.path.dirname(HERE), "oracle", "flops")


@pytest.fixture(scope="module")
def flopcount():
    subprocess.run(["make", "-s", "-C", FLOPS], check=True)
    return os.path.join(FLOPS, "flopcount")


def test_implicit_flops_per_solve(oracle, flopcount, tmp_path):
    import json
    import bench
    P = KP.params_implicit_batch()
    n = 3000
    pp, g20, G20, dt = PR.implicit_batch_states(P, n, seed=11)
    uu = oracle.p2u(P, pp, g20, G20)
    path = tmp_path / "batch.bin"
    with open(path, "wb") as f:
        f.write(bytes(P))
        f.write(np.int64(n).tobytes())
        for a in (pp, uu, g20, G20, dt):
            f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    out = subprocess.run([flopcount, "implicit", str(path)], check=True, capture_output=True, text=True).stdout
    r = json.loads(out)
    _, _, _, info = oracle.implicit(P, uu, pp, g20, G20, dt, nthreads=4)
    assert r["solved"] == int((info[:, 1] >= 0).sum())
    assert abs(r["iters"] - info[:, 5].mean()) < 1e-5
    assert abs(r["flops"] / bench.IMPLICIT_FLOPS_PER_SOLVE - 1.) < 0.08
