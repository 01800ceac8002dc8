"""Build libxtfem_hip.so (gfx950) in-tree with hipcc.

    python -m xtfem_amd.build            # or xtfem_amd.build.build()

Each translation unit is compiled to an object under xtfem_amd/_build/ in
parallel, then linked into xtfem_amd/libxtfem_hip.so.  Always recompiles.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libxtfem_hip.so")
BUILD = os.path.join(HERE, "_build")
ARCH = os.environ.get("XTFEM_OFFLOAD_ARCH", "gfx950")

SOURCES = ["xt_kron.hip", "xt_bsr.hip", "xt_prec.hip", "xt_band.hip", "xt_gmres.hip",
           "xt_damage.hip", "xt_capi.hip", "xt_prof.hip", "xt_ilut.cpp",
           "xt_delete.cpp", "xt_asm.hip"]
# rocSOLVER / rocBLAS: dense block factorization of the banded preconditioner (setup only).
LIBS = ["-L/opt/rocm/lib", "-lrocsolver", "-lrocblas"]

# Packed FP32 instructions are disabled for all device code (-packed-fp32-ops).
COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-command-line-argument"]


def hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *COMMON, "-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd = [hipcc(), *COMMON, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, *LIBS, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
