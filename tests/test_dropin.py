"""The C++ drop-in (include/rbf_operators.h) as its users call it: the
template kernels assemble_laplace_weights<laplace_config<...>> and
assemble_interp_weights<interp_config<...>> launched with the
configuration's block and LDS sizes (the reference's src/main.cpp:10-45,
src/rbf_operators.h:292-376), from a small user program
(tests/dropin/dropin_user.hip, built into libdropin_user.so by `build()`).

CPU: the library loads and the configurations report the launch contract.
GPU: the weights match the FP64 oracle to 1e-10 (relative, per stencil)
and are bit-identical to the run-time plans of librbf_amd.so, which run the
same engine with the same tuning; the per-stencil info matches the
oracle's on a stencil with duplicated nodes.
"""
import ctypes
import os

import numpy as np
import pytest

import rbf_amd as R
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "dropin", "libdropin_user.so")
FP64_TOL = 1e-10

# which -> (dim, N, P, Q) of dropin_user.hip's laplace configurations (3 and
# 4 have more than 64 unknowns: the four-wave tile engine)
LAPLACE = {0: (2, 21, 2, 3), 1: (2, 35, 3, 5), 2: (3, 30, 2, 3), 3: (2, 56, 4, 7),
           4: (3, 80, 3, 3)}
# which -> (N, P, Q, NRHS) of its interpolation configurations
INTERP = {0: (9, 2, 3, 9), 1: (21, 2, 3, 9)}

_lib = None


def _dropin():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with tests/dropin/build.py")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        L.dropin_laplace.restype = ctypes.c_int
        L.dropin_laplace.argtypes = [ctypes.c_int, ctypes.c_int, P, P, P, P, P, P, P]
        L.dropin_interp.restype = ctypes.c_int
        L.dropin_interp.argtypes = [ctypes.c_int, ctypes.c_int, P, P, P, P, P, P, P, P]
        L.dropin_config.restype = None
        L.dropin_config.argtypes = [ctypes.c_int, ctypes.c_int, P]
        L.dropin_laplace_f32.restype = ctypes.c_int
        L.dropin_laplace_f32.argtypes = [ctypes.c_int, P, P, P, P, P, P]
        L.dropin_interp_f32.restype = ctypes.c_int
        L.dropin_interp_f32.argtypes = [ctypes.c_int, P, P, P, P, P, P, P, P]
        L.dropin_user_interp.restype = ctypes.c_int
        L.dropin_user_interp.argtypes = [ctypes.c_int, ctypes.c_int, P, P, P, P, P, P, P, P]
        for f in ("dropin_gesv15_f64", "dropin_gesv15_f32"):
            getattr(L, f).restype = ctypes.c_int
            getattr(L, f).argtypes = [ctypes.c_int, P, P, P, P, P]
        _lib = L
    return _lib


def _config(interp, which):
    out = (ctypes.c_longlong * 6)()
    _dropin().dropin_config(interp, which, out)
    return dict(zip(("block", "lds", "nt", "N", "NRHS", "waves_per_simd"), list(out)))


def _npoly(dim, P):
    from math import comb
    return comb(P + dim, dim)


@pytest.mark.parametrize("which", sorted(LAPLACE))
def test_laplace_config_launch_contract(which):
    dim, n, P, Q = LAPLACE[which]
    c = _config(0, which)
    assert c["N"] == n and c["NRHS"] == dim + 1
    assert c["nt"] == n + _npoly(dim, P)
    if c["nt"] <= 64:
        assert c["block"] == 64 and c["waves_per_simd"] == 4
        # four waves of this LDS footprint fit one SIMD's share of a 160 KiB CU
        assert 0 < c["lds"] and 16 * c["lds"] <= 160 * 1024
    else:  # one four-wave workgroup per stencil
        assert c["block"] == 256 and c["waves_per_simd"] == 1
        assert 0 < c["lds"] <= 160 * 1024


@pytest.mark.parametrize("which", sorted(INTERP))
def test_interp_config_launch_contract(which):
    n, P, Q, nrhs = INTERP[which]
    c = _config(1, which)
    assert c["block"] == 64 and c["N"] == n and c["NRHS"] == nrhs
    nt = n + _npoly(2, P)
    assert c["nt"] == nt
    assert 0 < c["lds"] and 16 * c["lds"] <= 160 * 1024
    # room for slice(): xs, ys, As, Bs (doubles), ipiv (ints), as the
    # reference's shared_memory_size() counts them (rbf_operators.h:248-258)
    assert c["lds"] >= 8 * (2 * n + nt * nt + nt * nrhs) + 4 * nt


def test_dropin_rejects_unknown_configuration():
    L = _dropin()
    assert L.dropin_laplace(7, 16, None, None, None, None, None, None, None) == -1
    assert L.dropin_laplace(0, 0, None, None, None, None, None, None, None) == -1
    assert L.dropin_laplace(2, 16, None, None, None, None, None, None, None) == -1  # 3-D, no z
    assert L.dropin_laplace(4, 16, None, None, None, None, None, None, None) == -1
    assert L.dropin_interp(5, 16, None, None, None, None, None, None, None, None) == -1


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def _outputs(nops, size):
    import torch
    w = [torch.full((size,), np.nan, dtype=torch.float64, device="cuda:0") for _ in range(nops)]
    tab = torch.tensor([t.data_ptr() for t in w], dtype=torch.int64, device="cuda:0")
    return w, tab


def _laplace(which, pts, ja):
    import torch
    dim, n = LAPLACE[which][:2]
    nst = ja.size // n
    d = [_dev(p) for p in pts]
    jd = _dev(ja)
    w, tab = _outputs(dim + 1, ja.size)
    info = torch.full((nst,), -7, dtype=torch.int32, device="cuda:0")
    z = d[2].data_ptr() if dim == 3 else None
    stream = torch.cuda.current_stream().cuda_stream
    rc = _dropin().dropin_laplace(which, nst, d[0].data_ptr(), d[1].data_ptr(), z,
                                  jd.data_ptr(), tab.data_ptr(), info.data_ptr(), stream)
    assert rc == 0, f"launch failed: hipError {rc}"
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in w], info.cpu().numpy()


def _relerr(w, wref, n):
    return max((np.abs(a.reshape(-1, n) - b.reshape(-1, n)).max(1) /
                np.abs(b.reshape(-1, n)).max(1)).max() for a, b in zip(w, wref))


def _plan_weights(dim, n, P, Q, pts, ja, op="diff", nrhs=0, centres=None):
    import torch
    plan = R.Plan(dim, n, P, Q, op, nrhs=nrhs) if op == "interp" else R.Plan(dim, n, P, Q, op)
    w = [torch.full((ja.size,), np.nan, dtype=torch.float64, device="cuda:0")
         for _ in range(plan.nops)]
    z = _dev(pts[2]) if dim == 3 else None
    cen = None if centres is None else tuple(_dev(c) for c in centres)
    plan.assemble(ja.size // n, _dev(pts[0]), _dev(pts[1]), _dev(ja), w, z=z, centres=cen)
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in w]


@pytest.mark.gpu
@pytest.mark.parametrize("which,m", [(0, 40), (1, 48), (2, 12), (3, 24), (4, 10)])
def test_laplace_weights_match_oracle_and_plan(which, m):
    dim, n, P, Q = LAPLACE[which]
    pts = R.jittered_grid(dim, m, 0.25, seed=31 + which)
    ja = R.knn_stencils(n, *pts)
    print(f"drop-in laplace_config dim={dim} N={n} P={P} Q={Q}: {ja.size // n} stencils",
          flush=True)
    w, info = _laplace(which, pts, ja)
    z = pts[2] if dim == 3 else None
    wref, iref = O.assemble(dim, n, P, Q, pts[0], pts[1], ja, z=z)
    assert (iref == 0).all() and (info == 0).all()
    e = _relerr(w, wref, n)
    print(f"  max rel err vs oracle {e:.3e}")
    assert e <= FP64_TOL
    if dim == 2:  # the run-time plans ship these 2-D shapes
        for a, b in zip(w, _plan_weights(dim, n, P, Q, pts, ja)):
            np.testing.assert_array_equal(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("which", sorted(INTERP))
def test_interp_weights_match_oracle_and_plan(which):
    """The D2Q9 departure points of main.f90:16,55 (centres 0.5 h (cx, cy)),
    in the stencil-local frame as the reference passes them."""
    import torch
    n, P, Q, nrhs = INTERP[which]
    m = 32
    x, y = R.jittered_grid(2, m, 0.25, seed=41 + which)
    ja = R.knn_stencils(n, x, y)
    nst = ja.size // n
    h = 1.0 / m
    cx = 0.5 * h * np.array([0, 1, 0, -1, 0, 1, -1, -1, 1], float)
    cy = 0.5 * h * np.array([0, 0, 1, 0, -1, 1, 1, -1, -1], float)
    print(f"drop-in interp_config N={n} NRHS={nrhs}: {nst} stencils", flush=True)
    xd, yd, jd, cxd, cyd = _dev(x), _dev(y), _dev(ja), _dev(cx), _dev(cy)
    w, tab = _outputs(nrhs, ja.size)
    info = torch.full((nst,), -7, dtype=torch.int32, device="cuda:0")
    rc = _dropin().dropin_interp(which, nst, xd.data_ptr(), yd.data_ptr(), jd.data_ptr(),
                                 tab.data_ptr(), cxd.data_ptr(), cyd.data_ptr(),
                                 info.data_ptr(), torch.cuda.current_stream().cuda_stream)
    assert rc == 0, f"launch failed: hipError {rc}"
    torch.cuda.synchronize()
    w = [t.cpu().numpy() for t in w]
    assert (info.cpu().numpy() == 0).all()
    wref, iref = O.assemble(2, n, P, Q, x, y, ja, op_set=1, centres=(cx, cy))
    assert (iref == 0).all()
    e = _relerr(w, wref, n)
    print(f"  max rel err vs oracle {e:.3e}")
    assert e <= FP64_TOL
    for a, b in zip(w, _plan_weights(2, n, P, Q, (x, y), ja, "interp", nrhs, (cx, cy))):
        np.testing.assert_array_equal(a, b)


@pytest.mark.gpu
def test_laplace_singular_stencil_info():
    """Duplicated nodes: the stencils that hold both get info > 0, the same
    stencils as the oracle's (main.cpp:48-54 counts them and fails)."""
    n, P, Q = 21, 2, 3
    x, y = R.jittered_grid(2, 16, 0.25, seed=5)
    x[10], y[10] = x[11], y[11]
    ja = R.knn_stencils(n, x, y)
    w, info = _laplace(0, (x, y), ja)
    wref, iref = O.assemble(2, n, P, Q, x, y, ja)
    assert (info != 0).sum() > 0
    np.testing.assert_array_equal(info != 0, iref != 0)
    ok = iref == 0
    assert _relerr([a.reshape(-1, n)[ok] for a in w], [b.reshape(-1, n)[ok] for b in wref],
                   n) <= FP64_TOL


def _d2q9(h):
    cx = 0.5 * h * np.array([0, 1, 0, -1, 0, 1, -1, -1, 1], float)
    cy = 0.5 * h * np.array([0, 0, 1, 0, -1, 1, 1, -1, -1], float)
    return cx, cy


@pytest.mark.gpu
def test_float_configurations_round_the_fp64_weights():
    """T = float: coordinates widened on load, the FP64 engine, weights
    rounded on store -- so the float kernels' weights are exactly the FP64
    plan's weights on the widened coordinates, rounded to float."""
    import torch
    m = 32
    x, y = R.jittered_grid(2, m, 0.25, seed=77)
    xf, yf = x.astype(np.float32), y.astype(np.float32)
    x64, y64 = xf.astype(np.float64), yf.astype(np.float64)
    stream = torch.cuda.current_stream().cuda_stream
    # Laplacian / d/dx / d/dy, N = 35
    n = 35
    ja = R.knn_stencils(n, x64, y64)
    nst = ja.size // n
    w = [torch.full((ja.size,), np.nan, dtype=torch.float32, device="cuda:0") for _ in range(3)]
    tab = torch.tensor([t.data_ptr() for t in w], dtype=torch.int64, device="cuda:0")
    info = torch.full((nst,), -7, dtype=torch.int32, device="cuda:0")
    xd, yd, jd = _dev(xf), _dev(yf), _dev(ja)
    print(f"drop-in laplace_config<35,3,5,950,float>: {nst} stencils", flush=True)
    assert _dropin().dropin_laplace_f32(nst, xd.data_ptr(), yd.data_ptr(), jd.data_ptr(),
                                        tab.data_ptr(), info.data_ptr(), stream) == 0
    torch.cuda.synchronize()
    assert (info.cpu().numpy() == 0).all()
    for a, b in zip(w, _plan_weights(2, n, 3, 5, (x64, y64), ja)):
        np.testing.assert_array_equal(a.cpu().numpy(), b.astype(np.float32))
    # interpolation to the 9 D2Q9 centres, N = 9
    n = 9
    ja = R.knn_stencils(n, x64, y64)
    nst = ja.size // n
    cx, cy = (c.astype(np.float32) for c in _d2q9(1.0 / m))
    w = [torch.full((ja.size,), np.nan, dtype=torch.float32, device="cuda:0") for _ in range(9)]
    tab = torch.tensor([t.data_ptr() for t in w], dtype=torch.int64, device="cuda:0")
    info = torch.full((nst,), -7, dtype=torch.int32, device="cuda:0")
    jd, cxd, cyd = _dev(ja), _dev(cx), _dev(cy)
    print(f"drop-in interp_config<9,9,2,3,950,float>: {nst} stencils", flush=True)
    assert _dropin().dropin_interp_f32(nst, xd.data_ptr(), yd.data_ptr(), jd.data_ptr(),
                                       tab.data_ptr(), cxd.data_ptr(), cyd.data_ptr(),
                                       info.data_ptr(), stream) == 0
    torch.cuda.synchronize()
    assert (info.cpu().numpy() == 0).all()
    ref = _plan_weights(2, n, 2, 3, (x64, y64), ja, "interp", 9,
                        (cx.astype(np.float64), cy.astype(np.float64)))
    for a, b in zip(w, ref):
        np.testing.assert_array_equal(a.cpu().numpy(), b.astype(np.float32))


@pytest.mark.gpu
@pytest.mark.parametrize("which", sorted(INTERP))
def test_user_kernel_on_slice_and_solver(which):
    """A user's own kernel in the reference's shape (Config::slice,
    fill_matrix, fill_rhs, solve_system_AB<Config::solver>, scatter; no
    radius scaling): every stencil's weights equal the oracle's fill, rhs and
    GESV of the same shifted coordinates, bit for bit."""
    import torch
    n, P, Q, nrhs = INTERP[which]
    m = 24
    x, y = R.jittered_grid(2, m, 0.25, seed=51 + which)
    ja = R.knn_stencils(n, x, y)
    nst = ja.size // n
    cx, cy = _d2q9(1.0 / m)
    print(f"user kernel on interp_config<{n},{nrhs},{P},{Q}>::slice/solver: {nst} stencils",
          flush=True)
    w, tab = _outputs(nrhs, ja.size)
    info = torch.full((nst,), -7, dtype=torch.int32, device="cuda:0")
    xd, yd, jd, cxd, cyd = _dev(x), _dev(y), _dev(ja), _dev(cx), _dev(cy)
    assert _dropin().dropin_user_interp(which, nst, xd.data_ptr(), yd.data_ptr(), jd.data_ptr(),
                                        tab.data_ptr(), cxd.data_ptr(), cyd.data_ptr(),
                                        info.data_ptr(),
                                        torch.cuda.current_stream().cuda_stream) == 0
    torch.cuda.synchronize()
    w = np.stack([t.cpu().numpy().reshape(nst, n) for t in w])  # (nrhs, nst, n)
    info = info.cpu().numpy()
    J = ja.reshape(nst, n)
    for s in range(nst):
        xs, ys = x[J[s]] - x[s], y[J[s]] - y[s]
        X, inf = O.gesv(O.fill(2, n, P, Q, xs, ys),
                        O.rhs(2, n, P, Q, xs, ys, op_set=1, centres=(cx, cy)))
        assert info[s] == inf == 0
        np.testing.assert_array_equal(w[:, s, :], X[:n].T)


@pytest.mark.gpu
def test_gesv_execute_on_caller_systems_f64_f32():
    """GESV<T, 15, 9, 950>::execute on systems the caller filled (the demo's
    15 x 15, NRHS 9 shape): double -- factors, pivots and solution equal the
    oracle's; float -- the FP64 elimination of the widened system, rounded."""
    import torch
    rng = np.random.default_rng(15)
    nb = 64
    Ms = rng.standard_normal((nb, 15, 15))
    Bs = rng.standard_normal((nb, 15, 9))
    stream = torch.cuda.current_stream().cuda_stream
    for dt, fn in ((np.float64, "dropin_gesv15_f64"), (np.float32, "dropin_gesv15_f32")):
        Mt, Bt = Ms.astype(dt), Bs.astype(dt)
        A = _dev(Mt.transpose(0, 2, 1))
        B = _dev(Bt.transpose(0, 2, 1))
        ipiv = torch.zeros((nb, 15), dtype=torch.int32, device="cuda:0")
        info = torch.full((nb,), -1, dtype=torch.int32, device="cuda:0")
        print(f"{fn}: {nb} systems", flush=True)
        assert getattr(_dropin(), fn)(nb, A.data_ptr(), ipiv.data_ptr(), B.data_ptr(),
                                      info.data_ptr(), stream) == 0
        torch.cuda.synchronize()
        LU = A.cpu().numpy().transpose(0, 2, 1)
        X = B.cpu().numpy().transpose(0, 2, 1)
        ipiv, info = ipiv.cpu().numpy(), info.cpu().numpy()
        for b in range(nb):
            M64, B64 = Mt[b].astype(np.float64), Bt[b].astype(np.float64)
            LUr, pr, ir = O.getrf(M64)
            Xr, _ = O.gesv(M64, B64)
            assert info[b] == ir == 0
            np.testing.assert_array_equal(ipiv[b], pr)
            np.testing.assert_array_equal(LU[b], LUr.astype(dt))
            np.testing.assert_array_equal(X[b], Xr.astype(dt))
