"""Build the Fortran driver beside this file: fortran_driver.f90 against
the bind(c) modules of gpu-assembly_amd/fortran/rbf_assembly.f90, linked to
the in-tree librbf_amd.so (found through an $ORIGIN-relative rpath, so the
executable runs from any copy of the tree) and the HIP runtime.

    python tests/fortran/build.py
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
PKG = os.path.join(ROOT, "gpu-assembly_amd")
EXE = os.path.join(HERE, "fortran_driver")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def compiler():
    return shutil.which("flang") or shutil.which("amdflang")


def build(verbose=True):
    fc = compiler()
    if not fc:
        if verbose:
            print("no Fortran compiler: fortran_driver not built")
        return None
    mods = os.path.join(PKG, "_build", "fortran_driver")
    os.makedirs(mods, exist_ok=True)
    tmp = EXE + ".tmp"
    rel = os.path.relpath(PKG, HERE)
    cmd = [fc, "-O2", "-ffp-contract=off", "-module-dir", mods,
           os.path.join(PKG, "fortran", "rbf_assembly.f90"), os.path.join(HERE, "fortran_driver.f90"),
           "-o", tmp, f"-L{PKG}", "-lrbf_amd", f"-L{ROCM}/lib", "-lamdhip64",
           f"-Wl,-rpath,$ORIGIN/{rel}", f"-Wl,-rpath,{ROCM}/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("build of fortran_driver.f90 failed")
    os.replace(tmp, EXE)
    if verbose:
        print(f"built {EXE}")
    return EXE


if __name__ == "__main__":
    build()
