"""Builds libkoral_hip.so in-tree with hipcc for gfx950.

Every .hip and .cpp under koral_amd/csrc is compiled to an object and linked
into koral_amd/libkoral_hip.so.  -ffp-contract=off keeps each product and sum
individually rounded (the CPU FP64 reference is built without FMA), and
packed-FP32 instructions are disabled for the whole library.
"""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "libkoral_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("KORAL_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
          "-I", os.path.join(ROOT, "include"), "-I", CSRC]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + DEVICE + ["-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC] + COMMON + ["-x", "c++", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-6000:]))
    return obj


def build(verbose=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 2) // 2))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s" % r.stderr[-4000:])
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


def build_binding(verbose=False):
    """The KORAL-signature binding (koral_amd/binding/koral_ops.c, plain C as a
    KORAL build would compile it) and the test driver that plays ko.c
    (tests/host/binding_driver.c), built with gcc against include/koral_hip.h
    and linked with libkoral_hip.so (found next to the driver by rpath)."""
    out = os.path.join(ROOT, "tests", "host", "binding_driver")
    cmd = ["gcc", "-std=gnu11", "-O2", "-Wall", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(HERE, "binding"), os.path.join(HERE, "binding", "koral_ops.c"),
           os.path.join(ROOT, "tests", "host", "binding_driver.c"), "-L", HERE, "-lkoral_hip",
           "-Wl,-rpath,$ORIGIN/../../koral_amd", "-lm", "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("gcc failed for the binding driver:\n%s\n%s" % (" ".join(cmd), r.stderr[-4000:]))
    if verbose:
        print("built", out)
    return out


def build_test_libs(verbose=False):
    """Test libraries the GPU tests load (not part of the product): the
    element-wise FP64 check of tests/gpu/k_arith_check.hip, built with the
    library's own flags."""
    src = os.path.join(ROOT, "tests", "gpu", "k_arith_check.hip")
    out = os.path.join(ROOT, "tests", "gpu", "libarith_check.so")
    cmd = [HIPCC] + COMMON + DEVICE + ["-shared", src, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-4000:]))
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
