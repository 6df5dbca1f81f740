"""Build libdropin_user.so beside this file: the drop-in header's template
kernels instantiated by a user program (dropin_user.hip), for the GPU tests.

    python tests/dropin/build.py
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(HERE, "libdropin_user.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def build(verbose=True):
    sys.path.insert(0, os.path.join(ROOT, "gpu-assembly_amd"))
    import build as pkg  # gpu-assembly_amd/build.py: the library's HIP flags

    tmp = LIB + ".tmp"
    cmd = [os.path.join(ROCM, "bin", "hipcc"), *pkg.HIP_FLAGS, "-shared",
           os.path.join(HERE, "dropin_user.hip"), "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("build of dropin_user.hip failed")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
