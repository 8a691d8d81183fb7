"""Build the in-tree HIP library ``meltpooldg_amd/libmpdg_heat.so`` for gfx950.

Usage: python -m meltpooldg_amd.build   (or meltpooldg_amd.build.build())

Every source is compiled by hipcc for gfx950 only, with packed FP32 VALU ops
disabled (-packed-fp32-ops), then linked into one shared library whose
exported symbols are the C ABI of include/mpdg_heat.h.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJDIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmpdg_heat.so")
SOURCES = ["tables.cpp", "heat_kernels.hip", "cell_kernels.hip", "em_kernels.hip", "rhs_kernels.hip",
           "face_kernels.hip",
           "material_kernels.hip", "solver_kernels.hip",
           "batch_kernels.hip", "heat_api.cpp"]
ARCH = "gfx950"


def hipcc():
    for cand in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(cand) or cand == "hipcc":
            return cand


def common_flags():
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-Wall",
        "-Wno-unused-function",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", os.path.join(HERE, "..", "include"),
    ]


def _compile(src):
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc()] + common_flags() + ["-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj, r.stderr


def build(verbose=False, jobs=4):
    os.makedirs(OBJDIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(_compile, SOURCES))
    if verbose:
        for _, err in results:
            if err.strip():
                print(err, file=sys.stderr)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
