"""Build libpacosvd.so (HIP for gfx950) and the svd_main driver, in-tree.

Usage: python paco-svd_amd/build.py   (or from __graft_entry__.build()).
Compiles every source every time (the build check must always compile).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpacosvd.so")
DRIVER = os.path.join(HERE, "svd_main")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["bidiag.hip", "bidiag_blk.hip", "bidiag_grid.hip", "bidiag_res.hip", "qp_gemm.hip", "dc.hip", "gemm.hip", "engine.hip", "serve.hip", "tiling.cpp", "autotune.cpp"]
FLAGS = ["-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result",
         # packed FP32 instructions stay off (only the compiler's scalar FP32 forms are used)
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _compile(src):
    obj = os.path.join(BUILD, src + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), err))
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    subprocess.run(cmd, check=True, capture_output=not verbose)
    os.replace(tmp, LIB)
    drv_src = os.path.join(CSRC, "svd_main.cpp")
    if os.path.exists(drv_src):
        cmd = [HIPCC, "-O2", "-std=c++17", "-o", DRIVER, drv_src, "-L" + HERE, "-lpacosvd",
               "-Wl,-rpath,$ORIGIN"]
        subprocess.run(cmd, check=True, capture_output=not verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
