"""Build the HIP engine in-tree: sphere_amd/libsphere_hip.so and sphere_amd/bin/sphere.

Everything is compiled for gfx950 with hipcc.  FP contraction is off
(-ffp-contract=off) so that no multiply-add is fused, matching the
reference's nvcc --fmad=false (src/CMakeLists.txt:25); packed FP32
instructions are disabled.
"""

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsphere_hip.so")
CLI = os.path.join(HERE, "bin", "sphere")
FORCECHAINS = os.path.join(HERE, "bin", "forcechains")

ARCH = os.environ.get("SPHERE_HIP_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]

KERNEL_SOURCES = ["dem_kernels.hip", "adjoint.hip", "analysis.hip"]
HOST_SOURCES = ["engine.cpp"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in l
                    and "(ignoring feature)" not in l)
    if p.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + out)
    if verbose and out.strip():
        print(out)


def build(verbose=False):
    cc = hipcc()
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in KERNEL_SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        _run([cc, "--offload-arch=" + ARCH] + COMMON + ["-c", os.path.join(CSRC, src), "-o", obj],
             verbose)
        objs.append(obj)
    for src in HOST_SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        _run([cc] + COMMON + ["-x", "hip", "--offload-arch=" + ARCH, "-c", os.path.join(CSRC, src),
                              "-o", obj], verbose)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([cc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs, verbose)
    os.replace(tmp, LIB)
    os.makedirs(os.path.dirname(CLI), exist_ok=True)
    for src, exe in (("main.cpp", CLI), ("forcechains.cpp", FORCECHAINS)):
        _run([cc, "-O2", "-std=c++17", os.path.join(CSRC, src), "-o", exe,
              "-L" + HERE, "-lsphere_hip", "-Wl,-rpath,$ORIGIN/.."], verbose)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
