"""Build libins2d_hip.so in-tree with hipcc for gfx950.

``python -m ins2d_amd.build`` (or ``build()``) compiles every source under
``csrc/`` and links the shared library next to this file.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libins2d_hip.so")
SOURCES = ["ops1d.cpp", "kernels.hip", "api.hip", "batch.hip", "unet.hip", "fdm.hip", "fdm_host.cpp", "pcg2.hip",
           "noalloc.hip", "ops4.cpp", "w4.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # keep the compiler from emitting packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-Wall", "-Wno-unused-function",
]


def build(verbose=False):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    cmds, objs = [], []
    for s in SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(objdir, os.path.splitext(s)[0] + ".o")
        cmds.append([HIPCC] + FLAGS + ["-x", "hip", "-c", src, "-o", obj])
        objs.append(obj)
    # the translation units compile independently: in parallel, at most
    # MAX_JOBS (or 8) at a time
    from concurrent.futures import ThreadPoolExecutor
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", "8"))))

    def run(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    with ThreadPoolExecutor(jobs) as ex:
        for fut in [ex.submit(run, c) for c in cmds]:
            fut.result()
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_mex_harness(verbose)
    build_w4_harness(verbose)
    return LIB


def build_mex_harness(verbose=False):
    """The MATLAB gateway (matlab/ins2d_mex.c) linked with the test harness
    of tests/native/mex_mock.c, which stands in for MATLAB's MEX API: the
    tests drive the gateway through it (MATLAB itself is not installed)."""
    out = os.path.join(ROOT, "tests", "native", "libmex_mock.so")
    cmd = ["gcc", "-O1", "-std=c11", "-shared", "-fPIC", "-Wall",
           "-I" + os.path.join(ROOT, "tests", "native", "mex_mock"), "-I" + os.path.join(ROOT, "include"),
           os.path.join(ROOT, "matlab", "ins2d_mex.c"), os.path.join(ROOT, "tests", "native", "mex_mock.c"),
           "-L" + HERE, "-lins2d_hip", "-Wl,-rpath,$ORIGIN/../../ins2d_amd", "-o", out + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)
    return out


def build_w4_harness(verbose=False):
    """Host-only build of the fourth-order wall operator assembly
    (csrc/ops4.cpp, csrc/ops1d.cpp) with the test harness of
    tests/native/w4_dump.cpp, which hands the assembled matrices to the CPU
    tests."""
    out = os.path.join(ROOT, "tests", "native", "libw4_dump.so")
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wall",
           os.path.join(CSRC, "ops4.cpp"), os.path.join(CSRC, "ops1d.cpp"),
           os.path.join(ROOT, "tests", "native", "w4_dump.cpp"), "-o", out + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
