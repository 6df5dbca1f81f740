"""The C ABI library: it loads, exports every entry point include/*.h
declares, validates its arguments, and the Fortran bind(c) module reaches
it. No GPU compute here."""
import ctypes
import os
import re
import shutil
import subprocess

import pytest

import rbf_amd as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            text = open(os.path.join(inc, f)).read()
            names |= set(re.findall(r"RBFA_API\s+[\w\s\*]+?\b(rbfa_\w+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    names = _declared()
    assert len(names) >= 15
    L = ctypes.CDLL(R.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", R.LIB_PATH], capture_output=True,
                         text=True).stdout
    exported = set(re.findall(r"\bT (rbfa_\w+)", out))
    assert names <= exported


def test_measure_peaks_rejects_null_output():
    assert R.lib().rbfa_measure_peaks(None, None) == R.ERR_ARG


def test_version_and_status_strings():
    assert b"gfx950" in R.lib().rbfa_version()
    assert R.status_string(R.ERR_UNSUPPORTED) == "unsupported stencil shape"


def test_plan_argument_checks():
    with pytest.raises(R.RbfError) as e:
        R.Plan(2, 35, 3, 4)                       # even PHS exponent
    assert e.value.status == R.ERR_ARG
    with pytest.raises(R.RbfError) as e:
        R.Plan(2, 35, 3, 5, precision="mixed", max_refine=9)
    assert e.value.status == R.ERR_ARG
    with pytest.raises(R.RbfError) as e:
        R.Plan(2, 17, 3, 5)                       # a shape the library does not ship
    assert e.value.status == R.ERR_UNSUPPORTED


def test_apply_argument_checks():
    L = R.lib()
    tab = (ctypes.c_void_p * 1)(None)
    assert L.rbfa_apply(10, 35, None, None, 5, tab, None, tab, 0, None, None) == R.ERR_ARG
    assert L.rbfa_apply(10, 35, None, None, 1, tab, None, tab, 0, None, None) == R.ERR_ARG
    assert L.rbfa_apply(10, 35, None, None, 1, tab, None, tab, 3, None, None) == R.ERR_ARG
    assert L.rbfa_apply(0, 35, None, None, 1, tab, None, tab, 0, None, None) == R.OK


def test_apply_plan_validates_structure_before_touching_a_device():
    import numpy as np
    with pytest.raises(R.RbfError) as e:
        R.ApplyPlan(np.array([0, 1, 5, 2], np.int32), k=2)      # column 5 of 2 columns
    assert e.value.status == R.ERR_ARG
    with pytest.raises(R.RbfError) as e:
        R.ApplyPlan(np.array([0, 1], np.int32), ia=np.array([0, 2, 1], np.int32))
    assert e.value.status == R.ERR_ARG                           # decreasing row pointers
    with pytest.raises(R.RbfError) as e:
        R.ApplyPlan(np.zeros(3000, np.int32), ia=np.array([0, 3000], np.int32), ncols=1)
    assert e.value.status == R.ERR_ARG                           # a row wider than 2048


def test_device_node_layer_argument_checks():
    L = R.lib()
    # rejected before any device call: k > 128, k > n, bad dim, bad ndiv
    assert L.rbfa_knn_stencils_device(10, 2, 1, 1, None, 129, 1, None) == R.ERR_ARG
    assert L.rbfa_knn_stencils_device(10, 2, 1, 1, None, 11, 1, None) == R.ERR_ARG
    assert L.rbfa_knn_stencils_device(10, 4, 1, 1, None, 5, 1, None) == R.ERR_ARG
    assert L.rbfa_curve_keys_device(10, 2, 1, 1, 1, None, None, 32, 1, None) == R.ERR_ARG
    assert L.rbfa_curve_keys_device(10, 3, 1, 1, 1, 1, None, 8, 1, None) == R.ERR_ARG
    assert L.rbfa_order_by_keys_device(10, 1, 64, 1, None) == R.ERR_ARG
    assert L.rbfa_gather_f64_device(10, 1, 2, 2, None) == R.ERR_ARG


def test_gesv_argument_checks_before_any_launch():
    L = R.lib()
    d = ctypes.c_void_p(16)  # never dereferenced: every call below fails its checks
    assert L.rbfa_gesv_batched(0, 1, 1, d, d, d, d, None) == R.ERR_ARG
    assert L.rbfa_gesv_batched(65, 1, 1, d, d, d, d, None) == R.ERR_ARG
    assert L.rbfa_gesv_batched(15, 17, 1, d, d, d, d, None) == R.ERR_ARG
    assert L.rbfa_gesv_batched(15, -1, 1, d, d, d, d, None) == R.ERR_ARG
    assert L.rbfa_gesv_batched(15, 9, -1, d, d, d, d, None) == R.ERR_ARG
    assert L.rbfa_gesv_batched(15, 9, 1, None, d, d, d, None) == R.ERR_ARG
    assert L.rbfa_gesv_batched(15, 9, 1, d, None, d, d, None) == R.ERR_ARG
    assert L.rbfa_gesv_batched(15, 9, 1, d, d, None, d, None) == R.ERR_ARG
    assert L.rbfa_gesv_batched(15, 9, 0, None, None, None, None, None) == R.OK
    assert L.rbfa_solve_dp(d, d, d, None, None) == R.ERR_ARG


def test_assemble_rejects_null_plan():
    assert R.lib().rbfa_assemble(None, 10, None, None, None, None, None, None, None, None, None,
                                 None, None) == R.ERR_ARG


@pytest.mark.gpu
def test_assemble_argument_checks_before_any_launch():
    """Empty batch: OK with nothing launched (the reference's kernel over
    zero blocks is a no-op); negative counts, missing arrays and CSR sizes
    past int32 are errors, returned before any launch. (Creating the plan
    needs the device.)"""
    L = R.lib()
    plan = R.Plan(2, 35, 3, 5, "diff")
    h = plan._h
    nul = [None] * 10
    assert L.rbfa_assemble(h, 0, *nul, None) == R.OK
    assert L.rbfa_assemble(h, -1, *nul, None) == R.ERR_ARG
    assert L.rbfa_assemble(h, 10, *nul, None) == R.ERR_ARG
    dummy = ctypes.c_void_p(8)
    tab = (ctypes.c_void_p * 3)(8, 8, 8)
    big = 2**31 // 35 + 1   # nstencils * n past INT32_MAX
    assert L.rbfa_assemble(h, big, dummy, dummy, None, dummy, tab, None, None, None, None,
                           None, None) == R.ERR_ARG


@pytest.mark.skipif(not (shutil.which("flang") or shutil.which("amdflang")),
                    reason="no Fortran compiler")
def test_fortran_bind_c_reaches_the_library(tmp_path):
    """A Fortran program computes Hilbert keys and kNN stencils through the
    bind(c) module (the reference reaches its C side from Fortran the same
    way, main.f90:6-14)."""
    flang = shutil.which("flang") or shutil.which("amdflang")
    mod = os.path.join(ROOT, "gpu-assembly_amd", "fortran", "rbf_assembly.f90")
    prog = tmp_path / "t.f90"
    prog.write_text("""
program t
  use, intrinsic :: iso_c_binding
  use rbf_assembly
  implicit none
  integer(c_int32_t), parameter :: m = 16, n = m*m, k = 9
  real(c_double) :: x(n), y(n), z(1)
  integer(c_int64_t) :: keys(n)
  integer(c_int32_t) :: ja(n*k), perm(n), st
  type(rbfa_bbox2) :: bb
  st = rbfa_gen_jittered_grid(2, m, 0.25d0, 1_c_int64_t, x, y, z)
  if (st /= RBFA_OK) error stop 1
  bb = rbfa_bbox2(minval(x), minval(y), maxval(x), maxval(y))
  st = rbfa_hilbert_keys(n, x, y, bb, 16, keys)
  if (st /= RBFA_OK) error stop 2
  st = rbfa_order_by_keys(n, keys, perm)
  st = rbfa_knn_stencils(n, 2, x, y, z, k, ja)
  if (st /= RBFA_OK) error stop 3
  if (ja(1) /= 0 .or. ja(k+1) /= 1) error stop 4
  st = rbfa_knn_stencils(n, 2, x, y, z, n+1, ja)
  if (st /= RBFA_ERR_ARG) error stop 5
  print '(a,i0)', 'ok ', perm(1)
end program
""")
    exe = tmp_path / "t"
    subprocess.run([flang, "-o", str(exe), mod, str(prog), R.LIB_PATH,
                    f"-Wl,-rpath,{os.path.dirname(R.LIB_PATH)}", "-J", str(tmp_path)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    assert out.startswith("ok ")
