"""A Fortran program drives the library on the GPU through the bind(c)
modules only (tests/fortran/fortran_driver.f90, built by `build()`): the
plan-based assembly of n = 35 Laplacian / d/dx / d/dy weights, and the
reference demo's 15 x 15 system with 9 right-hand sides through
rbfa_solve_dp and rbfa_gesv_batched. Its inputs and outputs come back in
fortran_driver.bin and are checked against the oracle here.

CPU: the driver builds against the module. GPU: run it, then weights equal
the oracle's to the FP64 bar, factors / pivots / solutions bit for bit."""
import importlib.util
import os
import subprocess

import numpy as np
import pytest

import rbf_amd as R
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
FP64_TOL = 1e-10


def _builder():
    spec = importlib.util.spec_from_file_location(
        "fortran_build", os.path.join(HERE, "fortran", "build.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_fortran_driver_builds(tmp_path):
    b = _builder()
    if not b.compiler():
        pytest.skip("no Fortran compiler")
    exe = b.build(verbose=False)
    assert os.access(exe, os.X_OK)


class _Reader:
    def __init__(self, path):
        self.buf = open(path, "rb").read()
        self.at = 0

    def take(self, dtype, count):
        a = np.frombuffer(self.buf, dtype=dtype, count=count, offset=self.at)
        self.at += a.nbytes
        return a

    def f64(self, *shape):
        return self.take("<f8", int(np.prod(shape))).reshape(shape, order="F")

    def i32(self, *shape):
        return self.take("<i4", int(np.prod(shape))).reshape(shape, order="F")


@pytest.mark.gpu
def test_fortran_driver_on_gpu(tmp_path):
    exe = os.path.join(HERE, "fortran", "fortran_driver")
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} is missing: build it with tests/fortran/build.py")
    print("running the Fortran driver", flush=True)
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0 and r.stdout.strip().endswith("fortran driver ok")
    f = _Reader(tmp_path / "fortran_driver.bin")
    m, ns, nt, nrhs, nb = 48, 35, 15, 9, 6
    npts = m * m
    x, y = f.f64(npts), f.f64(npts)
    ja = f.i32(npts * ns)
    w = f.f64(npts * ns, 3)
    info = f.i32(npts)
    a0, b0, a, b = f.f64(nt, nt), f.f64(nt, nrhs), f.f64(nt, nt), f.f64(nt, nrhs)
    ipiv, info1 = f.i32(nt), f.i32(1)
    ab0, bb0 = f.f64(nt, nt, nb), f.f64(nt, nrhs, nb)
    ab, bb = f.f64(nt, nt, nb), f.f64(nt, nrhs, nb)
    ipivb, infob = f.i32(nt, nb), f.i32(nb)
    assert f.at == len(f.buf)

    # 1. the host node layer and the assembly, as the Python side does them
    xr, yr = R.jittered_grid(2, m, 0.25, seed=29)
    np.testing.assert_array_equal(x, xr)
    np.testing.assert_array_equal(y, yr)
    np.testing.assert_array_equal(ja, R.knn_stencils(ns, x, y))
    wref, iref = O.assemble(2, ns, 3, 5, x, y, ja)
    assert (info == 0).all() and (iref == 0).all()
    for c in range(3):
        g, r_ = w[:, c].reshape(-1, ns), wref[c].reshape(-1, ns)
        err = (np.abs(g - r_).max(1) / np.abs(r_).max(1)).max()
        assert err <= FP64_TOL, (c, err)

    # 2. the demo system as Fortran filled it is the oracle's fill
    px = np.array([0, 1, 0, -1, 0, 1, 1, -1, -1], float)
    py = np.array([0, 0, 1, 0, -1, 1, -1, -1, 1], float)
    np.testing.assert_array_equal(a0, O.fill(2, 9, 2, 3, px, py))
    np.testing.assert_array_equal(
        b0, O.rhs(2, 9, 2, 3, px, py, op_set=1, centres=(0.5 * px, 0.5 * py)))
    LUr, pr, ir = O.getrf(a0)
    Xr, _ = O.gesv(a0, b0)
    assert info1[0] == ir == 0
    np.testing.assert_array_equal(a, LUr)
    np.testing.assert_array_equal(ipiv, pr)
    np.testing.assert_array_equal(b, Xr)
    for k in range(nb):
        LUr, pr, ir = O.getrf(ab0[:, :, k])
        Xr, _ = O.gesv(ab0[:, :, k], bb0[:, :, k])
        assert infob[k] == ir == 0
        np.testing.assert_array_equal(ab[:, :, k], LUr)
        np.testing.assert_array_equal(ipivb[:, k], pr)
        np.testing.assert_array_equal(bb[:, :, k], Xr)
