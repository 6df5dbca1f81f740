"""Build librbf_amd.so in-tree: host node layer (g++, OpenMP) + gfx950
HIP kernels (hipcc) linked into one shared library beside this file.

    python gpu-assembly_amd/build.py          # full rebuild
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INC = os.path.join(ROOT, "include")
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librbf_amd.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = "gfx950"

# Packed FP32 VALU forms stay off (the compiler would otherwise fuse pairs
# of float ops into v_pk_* instructions); contraction stays off so sums of
# products round exactly as written, FMAs being explicit fma() calls.
# MFMA accumulators stay in ordinary VGPRs (by default LLVM moves them to
# AGPRs, and a kernel's register budget is the sum of both files).
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-Wno-unused-command-line-argument", f"-I{INC}",
]
CXX_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-fopenmp", "-ffp-contract=off", "-march=x86-64-v2",
             f"-I{INC}", f"-I{ROCM}/include", "-D__HIP_PLATFORM_AMD__"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} {os.path.basename(cmd[-1])}")
    return r


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    jobs = []
    objs = []
    for src in sorted(os.listdir(CSRC)):
        path = os.path.join(CSRC, src)
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        if src.endswith(".hip"):
            jobs.append([hipcc, *HIP_FLAGS, "-c", path, "-o", obj])
        elif src.endswith(".cpp"):
            jobs.append(["g++", *CXX_FLAGS, "-c", path, "-o", obj])
        else:
            continue
        objs.append(obj)
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(_run, jobs))
    tmp = LIB + ".tmp"
    _run(["g++", "-shared", "-fopenmp", "-o", tmp, *objs, f"-L{ROCM}/lib", "-lamdhip64",
          f"-Wl,-rpath,{ROCM}/lib"])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_fortran_check():
    """Compile the Fortran bind(c) module (interface check only)."""
    flang = shutil.which("flang") or shutil.which("amdflang")
    if not flang:
        return None
    src = os.path.join(HERE, "fortran", "rbf_assembly.f90")
    if not os.path.exists(src):
        return None
    os.makedirs(BUILD, exist_ok=True)
    obj = os.path.join(BUILD, "rbf_assembly_f.o")
    _run([flang, "-c", "-fPIC", src, "-o", obj, "-module-dir", BUILD])
    return obj


if __name__ == "__main__":
    build()
    build_fortran_check()
