"""Build libcanary_dg.so in-tree: host preparation (g++) + HIP kernels for gfx950.

Usage: python canary.jl_amd/build.py [-j JOBS]

Every HIP translation unit is compiled with
``-Xclang -target-feature -Xclang -packed-fp32-ops`` (no packed FP32 ops are
emitted).  The exact-order FP64 kernels and the host metric code are built
with ``-ffp-contract=off`` so that no multiply-add is fused and the
evaluation order of the reference CPU kernels is kept.  The host objects
use the x86-64 FMA instructions (-mfma) only where the code calls fma()
(the correctly rounded pow); no expression is contracted.
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcanary_dg.so")

HOST_SOURCES = ["host/quadrature.cpp", "host/mesh.cpp", "host/metric.cpp", "host/problems.cpp"]
HIP_SOURCES = ["hip/runtime.hip", "hip/euler3d.hip", "hip/euler2d.hip", "hip/swe2d.hip",
               "hip/euler2d_te.hip",
               "hip/batched.hip", "hip/geometry.hip", "hip/aux.hip", "hip/prep.hip",
               "hip/prepared.hip", "hip/driver.hip"]
# kernels allowed to contract a*b+c into FMA
HIP_FAST_SOURCES = ["hip/perturb32.hip", "hip/euler3d_fma.hip", "hip/euler2d_fma.hip",
                    "hip/swe2d_fma.hip", "hip/euler2d_te_fma.hip", "hip/batched_fma.hip",
                    "hip/euler3d_mfma.hip"]

ARCH = os.environ.get("CANARY_OFFLOAD_ARCH", "gfx950")


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found: the HIP kernels cannot be built")


def _cmds():
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
    cmds = []
    for s in HOST_SOURCES:
        o = os.path.join(OBJ, s.replace("/", "_") + ".o")
        cmds.append((o, ["g++", "-std=c++17", "-O2", "-fPIC", "-ffp-contract=off", "-mfma", "-fopenmp",
                         "-fno-fast-math", "-Wall", "-Wno-unused-function", *inc,
                         "-c", os.path.join(CSRC, s), "-o", o]))
    hip = _hipcc()
    hip_common = ["-std=c++17", "-O3", "-fPIC", "--offload-arch=" + ARCH,
                  "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                  *inc]
    for s in HIP_SOURCES + HIP_FAST_SOURCES:
        o = os.path.join(OBJ, s.replace("/", "_") + ".o")
        contract = "-ffp-contract=fast" if s in HIP_FAST_SOURCES else "-ffp-contract=off"
        cmds.append((o, [hip, *hip_common, contract, "-c", os.path.join(CSRC, s), "-o", o]))
    return cmds


def build(jobs=None, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    cmds = _cmds()
    jobs = jobs or min(8, os.cpu_count() or 4)

    def run(item):
        o, cmd = item
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
        return o

    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(run, cmds))
    tmp = LIB + ".tmp"
    link = [_hipcc(), "-shared", "-fPIC", "--offload-arch=" + ARCH, *objs, "-lgomp", "-o", tmp]
    r = subprocess.run(link, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(a.j, a.v))
    sys.exit(0)
