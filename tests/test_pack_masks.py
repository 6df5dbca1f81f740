"""The packed engine's lane-mask steps on the host (no GPU needed): the
word-parallel choice of each segment's pivot lane and tie test
(WavePack::pivot_lanes / segments_tied, include/rbf_wavepack.h) against a
segment-by-segment brute force, over 8 million random and structured masks
for 16- and 32-lane segments (tests/pack_masks/check_masks.hip, built with
hipcc as host code)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def test_segment_masks_match_brute_force(tmp_path):
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    if not os.path.exists(hipcc):
        hipcc = shutil.which("hipcc")
    if not hipcc:
        pytest.skip("hipcc not available")
    exe = tmp_path / "check_masks"
    src = os.path.join(HERE, "pack_masks", "check_masks.hip")
    r = subprocess.run([hipcc, "--offload-arch=gfx950", "-O2", "-std=c++17",
                        "-I" + os.path.join(ROOT, "include"), src, "-o", str(exe)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and r.stdout.strip().startswith("ok"), r.stdout[-2000:]
