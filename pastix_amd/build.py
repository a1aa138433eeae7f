"""Build the MI355X PaStiX engine: the C host library (API, analysis,
refinement) and the HIP device layer for gfx950, linked into one shared
library ``pastix_amd/lib/libpastix_amd.so``.

Run ``python -m pastix_amd.build`` (or ``__graft_entry__.build()``).  The
build always recompiles; objects go to ``pastix_amd/_build``.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIBNAME = "libpastix_amd.so"

HOST_SOURCES = ["host/order.c", "host/amalgamate.c", "host/symbol.c", "host/split.c", "host/solver.c",
                "host/solver_io.c", "host/factor_io.c", "host/simu.c",
                "host/bcsc.c", "host/api.c", "host/refine.c"]
HIP_SOURCES = ["hip/engine.hip", "hip/gpu_api.hip"]

INCLUDES = ["-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(CSRC, "host"),
            "-I" + os.path.join(CSRC, "hip")]
CFLAGS = ["-O2", "-fPIC", "-std=gnu11", "-Wall", "-Wno-unused-function"]
HIPFLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-result", "-Wno-unused-value"]


def _hipcc():
    for cand in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(cand) or cand == "hipcc":
            return cand
    return "hipcc"


def _run(cmd):
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # the host pass of hipcc does not know the packed-fp32 feature: harmless
    out = "\n".join(l for l in res.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in l)
    if res.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    jobs = []
    objs = []
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        jobs.append(["gcc"] + CFLAGS + INCLUDES + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        jobs.append([_hipcc()] + HIPFLAGS + INCLUDES + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    lib = os.path.join(LIBDIR, LIBNAME)
    tmp = lib + ".tmp"
    _run([_hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lm"])
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
