"""The grid operators of koral_amd/csrc/grid_ops.hpp compiled for the host
(tests/host/grid_check.cpp) with every buffer behind guard pages, index checks
on every grid address, AddressSanitizer and UBSan: no access leaves its
buffer or box, and the results equal the oracle's bit for bit.

The field + radiation explicit operator faulted on the device in an earlier
build of these kernels (illegal address on the radiative torus); these runs
cover that combination (and the earlier build's launch geometry of the ghost
fill and of the fused face kernel) on the host, at the test sizes here and,
with KORAL_HOST_FULL=1, at the 192x192x96 shape of the GRRMHD+M1 step."""
import os

import numpy as np
import pytest

import host_grid as H
from koral_amd import params as KP, problems as PR, grid as GR


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    return H.build(tmp_path_factory.mktemp("hc") / "grid_check", sanitize=True)


def _setup(kind):
    if kind == "tube":
        P = KP.params_tube()
        spec, p = PR.tube_setup(P, nx=64)
    elif kind == "torus_mhd":
        P = KP.params_torus_mhd()
        spec, p = PR.torus_setup(P, n=(16, 16, 8))
    elif kind == "torus_mhd_2d":
        P = KP.params_torus_mhd()
        spec, p = PR.torus_setup(P, n=(16, 12, 1))
    elif kind == "torus_mhd_odd":
        P = KP.params_torus_mhd()
        spec, p = PR.torus_setup(P, n=(13, 10, 6))
    else:
        P = KP.params_torus_rad()
        if kind == "torus_rad_nopolar":
            P.polar_nc = 0
        n = (11, 14, 6) if kind == "torus_rad_odd" else (12, 12, 8)
        spec, p = PR.torus_setup(P, n=n)
    return P, spec, GR.to_aos(p)


def _same(a, b):
    return np.array_equal(a, b, equal_nan=True)


KINDS = ["tube", "torus_mhd", "torus_mhd_odd", "torus_rad", "torus_rad_nopolar", "torus_rad_odd"]


@pytest.mark.parametrize("kind", KINDS + ["torus_mhd_2d"])
def test_host_step_matches_oracle(oracle, exe, tmp_path, kind):
    P, spec, p = _setup(kind)
    nsteps = 4 if kind == "tube" else 2
    S = oracle.GridState(P, oracle.Grid(spec), p)
    S.init_u()
    S.set_bc()
    dt = 1. / S.tstepden()
    t = 0.
    recs_o = []
    for _ in range(nsteps):
        tm, c = S.step(t, dt)
        recs_o.append((tm, c))
        t += dt
        dt = 1. / tm
    inp = tmp_path / "in.bin"
    H.write_input(inp, P, spec, p)
    for mode in (0, 1):
        out = tmp_path / ("out%d.bin" % mode)
        r = H.run(exe, inp, out, mode, 1, "step", nsteps, timeout=900)
        assert r.returncode == 0, r.stderr[-4000:]
        pd, ud, fd, recs = H.read_output(out, spec, P.nv, nsteps)
        assert _same(pd, S.p) and _same(ud, S.u) and _same(fd, S.flags), kind
        for (tm_d, c_d), (tm_o, c_o) in zip(recs, recs_o):
            assert tm_d == tm_o
            assert np.array_equal(c_d, c_o)


@pytest.mark.parametrize("kind", KINDS)
def test_host_explicit_and_legacy_launch_match_oracle(oracle, exe, tmp_path, kind):
    """op_explicit on the initialised state, as the current operators run it
    and as the earlier build launched it (whole-box face fill, one thread per
    edge position, one fused reconstruction + full floors + flux kernel per
    direction).  The radiative torus has faces whose b^2/rho-floor
    re-inversion leaves a negative density and a NaN entropy (the oracle does
    the same); NaNs must sit in the same places."""
    P, spec, p = _setup(kind)
    S = oracle.GridState(P, oracle.Grid(spec), p)
    S.init_u()
    S.set_bc()
    dt = 1. / S.tstepden()
    inp = tmp_path / "in.bin"
    H.write_input(inp, P, spec, S.p, S.u)
    tm = S.explicit(0., dt)
    for cmd in ("explicit", "legacy_explicit"):
        for mode in (0, 1):
            out = tmp_path / ("out_%s_%d.bin" % (cmd, mode))
            r = H.run(exe, inp, out, mode, 1, cmd, repr(dt), timeout=900)
            assert r.returncode == 0, r.stderr[-4000:]
            pd, ud, fd, recs = H.read_output(out, spec, P.nv, 1)
            assert _same(ud, S.u), (kind, cmd, mode)
            assert recs[0][0] == tm


@pytest.mark.parametrize("kind", ["torus_mhd", "torus_rad"])
def test_host_legacy_ghost_fill_matches_oracle(oracle, exe, tmp_path, kind):
    P, spec, p = _setup(kind)
    S = oracle.GridState(P, oracle.Grid(spec), p)
    S.init_u()
    S.set_bc()
    inp = tmp_path / "in.bin"
    H.write_input(inp, P, spec, p)
    for mode in (0, 1):
        out = tmp_path / ("out%d.bin" % mode)
        r = H.run(exe, inp, out, mode, 1, "legacy_bc", timeout=600)
        assert r.returncode == 0, r.stderr[-4000:]
        pd, ud, _, _ = H.read_output(out, spec, P.nv, 0)
        assert _same(pd, S.p) and _same(ud, S.u)


def test_host_check_catches_planted_overrun(tmp_path):
    """The guard pages catch an access one element past a buffer (mode 1)
    and one element before it (mode 0), and the index checks catch a
    coordinate outside the box that stays inside the buffer."""
    src = tmp_path / "plant.cpp"
    src.write_text(r'''
#define main grid_check_main
#include "%s"
#undef main
int main(int argc, char **argv)
{
  g_mode = atoi(argv[1]);
  double *b = (double *)guarded(100 * sizeof(double), "planted");
  struct sigaction sa; memset(&sa, 0, sizeof sa); sa.sa_sigaction = on_segv; sa.sa_flags = SA_SIGINFO;
  sigaction(SIGSEGV, &sa, nullptr);
  const int what = atoi(argv[2]);
  volatile double x = 0.;
  if (what == 0) x = b[-1];
  if (what == 1) x = b[100];
  if (what == 2) {
    koral_grid_desc D{}; D.nx = 8; D.ny = 8; D.nz = 8; D.ng = 4;
    DGrid G; grid_layout(&D, G);
    x = (double)cidx(G, 12, 0, 0);
  }
  (void)x;
  return 0;
}
''' % H.SRC)
    exe = tmp_path / "plant"
    import subprocess
    subprocess.run(["g++", "-O0", "-std=c++17", "-fopenmp", "-I", os.path.join(H.ROOT, "koral_amd", "csrc"), "-I",
                    os.path.join(H.ROOT, "include"), str(src), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe), "0", "0"], capture_output=True, text=True)
    assert r.returncode != 0 and "GUARD FAULT: planted at byte offset -8" in r.stderr
    r = subprocess.run([str(exe), "1", "1"], capture_output=True, text=True)
    assert r.returncode != 0 and "GUARD FAULT: planted at byte offset 800" in r.stderr
    r = subprocess.run([str(exe), "0", "2"], capture_output=True, text=True)
    assert r.returncode != 0 and "INDEX CHECK FAILED: cidx (12, 0, 0)" in r.stderr


@pytest.mark.skipif(os.environ.get("KORAL_HOST_FULL") != "1",
                    reason="full-size host check takes minutes and ~12 GB; set KORAL_HOST_FULL=1")
def test_host_full_size_grrmhd_m1(tmp_path):
    """One GRRMHD+M1 RK2IMEX step of the radiative torus at 192x192x96
    (NV=13, NG=4), every item of every launch, in both guard modes."""
    exe = H.build(tmp_path / "grid_check_fast", sanitize=False)
    P = KP.params_torus_rad()
    spec, p = PR.torus_setup(P, n=(192, 192, 96))
    inp = tmp_path / "in.bin"
    H.write_input(inp, P, spec, GR.to_aos(p))
    del p
    for mode in (0, 1):
        r = H.run(exe, inp, tmp_path / "out.bin", mode, 1, "step", 1, timeout=7200)
        assert r.returncode == 0, r.stderr[-4000:]


@pytest.mark.parametrize("kind", ["tube", "torus_mhd", "torus_rad"])
@pytest.mark.parametrize("tune", ["fused", "fused,lane,pre,combrt", "lane,budget", "pre,budget"])
def test_host_step_launch_choices_match_oracle(oracle, exe, tmp_path, kind, tune):
    """The launch choices the autotuner picks between run different functors
    (reconstruction inside the flux functor, one lane instead of one wave per
    deferred face or cell); each gives the oracle's state bit for bit."""
    P, spec, p = _setup(kind)
    nsteps = 2
    S = oracle.GridState(P, oracle.Grid(spec), p)
    S.init_u()
    S.set_bc()
    dt = 1. / S.tstepden()
    t = 0.
    for _ in range(nsteps):
        tm, c = S.step(t, dt)
        t += dt
        dt = 1. / tm
    inp = tmp_path / "in.bin"
    out = tmp_path / "out.bin"
    H.write_input(inp, P, spec, p)
    r = H.run(exe, inp, out, 0, 1, "step", nsteps, timeout=900, tune=tune)
    assert r.returncode == 0, r.stderr[-4000:]
    pd, ud, fd, recs = H.read_output(out, spec, P.nv, nsteps)
    assert _same(pd, S.p) and _same(ud, S.u) and _same(fd, S.flags), (kind, tune)


@pytest.mark.skipif(not os.path.exists(H.CLANG), reason="clang with MemorySanitizer not installed")
@pytest.mark.parametrize("kind", ["torus_rad", "torus_mhd"])
def test_host_msan_explicit_and_step(oracle, tmp_path, kind):
    """The field + radiation explicit operator (and the GRMHD one) under
    MemorySanitizer: no value that depends on uninitialised memory decides a
    branch, an address or an output, in the current functors or in the
    earlier build's launch geometry; results still the oracle's."""
    exe = H.build_msan(tmp_path / "grid_check_msan")
    P, spec, p = _setup(kind)
    S = oracle.GridState(P, oracle.Grid(spec), p)
    S.init_u()
    S.set_bc()
    dt = 1. / S.tstepden()
    inp = tmp_path / "in.bin"
    H.write_input(inp, P, spec, S.p, S.u)
    for cmd in ("explicit", "legacy_explicit"):
        out = tmp_path / (cmd + ".bin")
        r = H.run(exe, inp, out, 0, 1, cmd, repr(dt), timeout=900)
        assert r.returncode == 0, r.stderr[-4000:]
    S.explicit(0., dt)
    pd, ud, fd, _ = H.read_output(tmp_path / "explicit.bin", spec, P.nv, 1)
    assert _same(ud, S.u)
    H.write_input(inp, P, spec, p)
    r = H.run(exe, inp, tmp_path / "step.bin", 0, 1, "step", 1, timeout=900)
    assert r.returncode == 0, r.stderr[-4000:]


def test_host_implicit_batch_bit_identical_to_oracle(oracle, tmp_path):
    """The device's implicit solve (dev_implicit.hpp: the per-lane Newton step
    of the GPU passes in the sequential fallback chain) built for the host
    gives the oracle's result bit for bit on the seeded batch workload: same
    attempt, iterations and primitives on every cell, so a difference on the
    GPU comes from the GPU's arithmetic, not from the algorithm."""
    import subprocess
    exe = tmp_path / "implicit_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-I",
                    os.path.join(H.ROOT, "koral_amd", "csrc"), "-I", os.path.join(H.ROOT, "include"),
                    os.path.join(H.ROOT, "tests", "host", "implicit_check.cpp"), "-o", str(exe)], check=True)
    P = KP.params_implicit_batch()
    n = 1 << 14
    pp, g20, G20, dt = PR.implicit_batch_states(P, n, seed=90)
    uu = oracle.p2u(P, pp, g20, G20)
    inp, out = tmp_path / "batch.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        f.write(bytes(P))
        f.write(np.int64(n).tobytes())
        for a in (pp, uu, g20, G20, dt):
            f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    subprocess.run([str(exe), str(inp), str(out)], check=True)
    raw = open(out, "rb").read()
    ph = np.frombuffer(raw[:8 * P.nv * n], np.float64).reshape(n, P.nv)
    ih = np.frombuffer(raw[8 * P.nv * n:], np.int32).reshape(n, 6)
    _, po, _, info = oracle.implicit(P, uu, pp, g20, G20, dt, nthreads=4)
    assert np.array_equal(ih, info)
    assert _same(ph, po)
