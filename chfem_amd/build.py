"""Builds libchfem.so (HIP kernels + C ABI) and the chfem_exec CLI for gfx950
in-tree, with hipcc.  Usage: python -m chfem_amd.build"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libchfem.so")
EXE = os.path.join(HERE, "chfem_exec")
ARCH = os.environ.get("CHFEM_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    # packed FP32 instructions stay off in all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]

SOURCES = ["kernels.hip", "walsh.hip", "stokes.hip", "ebe_mfma.hip", "batch.hip", "guess.hip", "noalloc.hip", "lean.hip",
           "cudapcg.cpp", "femhmg.cpp", "guess.cpp"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out)


def build(verbose=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    cmds, objs = [], []
    for s in SOURCES:
        o = os.path.join(BUILD, s.replace(".", "_") + ".o")
        extra = ["-fopenmp"] if s.endswith(".cpp") else []
        cmds.append([HIPCC, "-x", "hip", "-c", os.path.join(SRC, s), "-o", o] + COMMON + extra)
        objs.append(o)
    # one compiler process per source, at most one per CPU (16 at most)
    jobs = jobs or max(1, min(len(cmds), os.cpu_count() or 1, 16))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for f in [ex.submit(_run, c, verbose) for c in cmds]:
            f.result()
    _run([HIPCC, "-shared", "-o", LIB + ".tmp"] + objs + [f"--offload-arch={ARCH}", "-fopenmp"], verbose)
    os.replace(LIB + ".tmp", LIB)
    mo = os.path.join(BUILD, "main.o")
    _run([HIPCC, "-x", "hip", "-c", os.path.join(SRC, "main.cpp"), "-o", mo] + COMMON, verbose)
    _run([HIPCC, "-o", EXE, mo, f"-L{HERE}", "-lchfem", f"-Wl,-rpath,{HERE}", "-fopenmp",
          f"--offload-arch={ARCH}"], verbose)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
