"""Driver of the host executor of the grid operators (tests/host/grid_check.cpp):
builds it, writes its input from a grid set-up and reads its output."""
import os
import subprocess
import struct

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "host", "grid_check.cpp")
MAGIC = 0x4b4f5241


def build(out, sanitize=True):
    """g++ build of the checker; with sanitize, AddressSanitizer and UBSan
    (float-to-int overflow included, which -fsanitize=undefined leaves out)."""
    cmd = ["g++", "-O1" if sanitize else "-O2", "-g", "-std=c++17", "-fopenmp", "-ffp-contract=off",
           "-fno-fast-math", "-I", os.path.join(ROOT, "koral_amd", "csrc"), "-I", os.path.join(ROOT, "include"),
           SRC, "-o", str(out)]
    if sanitize:
        cmd[1:1] = ["-fsanitize=address,undefined,float-cast-overflow", "-fno-sanitize-recover=all",
                    "-fno-omit-frame-pointer"]
    subprocess.run(cmd, check=True)
    return str(out)


CLANG = "/opt/rocm/lib/llvm/bin/clang++"


def build_msan(out):
    """clang build of the checker with MemorySanitizer (uses of uninitialised
    values, e.g. an index or a branch computed from a variable no path set:
    the kind of error that is harmless on a host stack and not on a GPU),
    serial (libgomp is not instrumented)."""
    cmd = [CLANG, "-fsanitize=memory", "-fsanitize-memory-track-origins=2", "-fno-omit-frame-pointer", "-O1", "-g",
           "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-I", os.path.join(ROOT, "koral_amd", "csrc"),
           "-I", os.path.join(ROOT, "include"), SRC, "-o", str(out)]
    subprocess.run(cmd, check=True)
    return str(out)


def write_input(path, P, spec, p_aos, u_aos=None):
    import ctypes
    hdr = [MAGIC, spec.nx, spec.ny, spec.nz, spec.ng] + list(spec.bc) + [ctypes.sizeof(P)]
    with open(path, "wb") as f:
        f.write(struct.pack("<12i", *hdr))
        f.write(bytes(P))
        for x in spec.xb:
            f.write(np.ascontiguousarray(x, dtype=np.float64).tobytes())
        for k in ("gc", "Gc", "kr", "gfx", "Gfx", "gfy", "Gfy"):
            f.write(np.ascontiguousarray(spec.tables[k], dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(p_aos, dtype=np.float64).tobytes())
        if u_aos is not None:
            f.write(np.ascontiguousarray(u_aos, dtype=np.float64).tobytes())


def read_output(path, spec, nv, nrec):
    shape = (spec.sz, spec.sy, spec.sx)
    n = spec.sz * spec.sy * spec.sx
    with open(path, "rb") as f:
        p = np.frombuffer(f.read(8 * n * nv), np.float64).reshape(shape + (nv,))
        u = np.frombuffer(f.read(8 * n * nv), np.float64).reshape(shape + (nv,))
        fl = np.frombuffer(f.read(4 * n * 7), np.int32).reshape(shape + (7,))
        recs = []
        for _ in range(nrec):
            tsd = struct.unpack("<d", f.read(8))[0]
            c = np.frombuffer(f.read(24), np.int32).copy()
            recs.append((tsd, c))
    return p, u, fl, recs


def run(exe, inp, out, mode, stride, cmd, arg=None, timeout=1800, threads=None, tune=None):
    """tune: launch choices that change which functors run ("fused",
    "lane" or both, comma separated)."""
    env = dict(os.environ)
    if tune:
        env["KORAL_HC_TUNE"] = tune
    env.setdefault("ASAN_OPTIONS", "detect_leaks=0:abort_on_error=1")
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["MSAN_OPTIONS"] = "halt_on_error=1"
    if threads:
        env["OMP_NUM_THREADS"] = str(threads)
    argv = [exe, str(inp), str(out), str(mode), str(stride), cmd] + ([] if arg is None else [str(arg)])
    return subprocess.run(argv, capture_output=True, text=True, timeout=timeout, env=env)
