"""GPU parity: the HIP assembly (through the C ABI) against the CPU FP64
oracle on the same seeded nodes and stencils.

Bars: FP64 form, max over stencils and operators of
max_k |w - w_ref| / max_k |w_ref| <= 1e-10; mixed form <= 1e-8.
"""
import numpy as np
import pytest

import rbf_amd as R
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
MIXED_TOL = 1e-8


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def _run(plan, x, y, ja, z=None, centres=None):
    import torch
    nst = ja.size // plan.n
    w = [torch.full((ja.size,), np.nan, dtype=torch.float64, device="cuda:0")
         for _ in range(plan.nops)]
    info = torch.full((nst,), -7, dtype=torch.int32, device="cuda:0")
    status = torch.full((nst,), 255, dtype=torch.uint8, device="cuda:0")
    cen = None if centres is None else tuple(_dev(c) for c in centres)
    plan.assemble(nst, _dev(x), _dev(y), _dev(ja), w, z=None if z is None else _dev(z),
                  centres=cen, info=info, status=status)
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in w], info.cpu().numpy(), status.cpu().numpy()


def _relerr(w, wref, n):
    e = [np.abs(a.reshape(-1, n) - b.reshape(-1, n)).max(1) /
         np.abs(b.reshape(-1, n)).max(1) for a, b in zip(w, wref)]
    return np.stack(e, 1)


SHAPES_2D = [(64, 21, 2, 3), (64, 35, 3, 5), (48, 13, 2, 3), (40, 35, 3, 3), (40, 21, 2, 5),
             (40, 13, 1, 3)]


@pytest.mark.parametrize("m,n,P,Q", SHAPES_2D)
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_diff_ops_match_oracle(m, n, P, Q, prec):
    try:
        plan = R.Plan(2, n, P, Q, "diff", precision=prec)
    except R.RbfError as e:
        if e.status == R.ERR_UNSUPPORTED:
            pytest.skip("shape not shipped in this precision")
        raise
    x, y = R.jittered_grid(2, m, 0.25, seed=7)
    ja = R.knn_stencils(n, x, y)
    w, info, status = _run(plan, x, y, ja)
    wref, iref = O.assemble(2, n, P, Q, x, y, ja)
    assert (info == 0).all() and (iref == 0).all()
    e = _relerr(w, wref, n)
    tol = FP64_TOL if prec == "fp64" else MIXED_TOL
    print(f"{prec} m={m} n={n} P={P} Q={Q}: max rel err {e.max():.3e}")
    assert e.max() <= tol
    if prec == "fp64":
        assert (status == 0).all()
    else:
        assert ((status & 0x3) <= 3).all()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_every_variant_matches_oracle(prec):
    n, P, Q = 35, 3, 5
    plan = R.Plan(2, n, P, Q, "diff", precision=prec)
    x, y = R.jittered_grid(2, 48, 0.25, seed=11)
    ja = R.knn_stencils(n, x, y)
    wref, _ = O.assemble(2, n, P, Q, x, y, ja)
    tol = FP64_TOL if prec == "fp64" else MIXED_TOL
    first = None
    for v in range(plan.num_variants):
        plan.variant = v
        w, info, _ = _run(plan, x, y, ja)
        assert (info == 0).all()
        e = _relerr(w, wref, n)
        print(f"{prec} variant {v} ({plan.variant_name()}): max rel err {e.max():.3e}")
        assert e.max() <= tol
        if prec == "fp64":      # variants differ in speed, not in results
            if first is None:
                first = w
            else:
                for a, b in zip(w, first):
                    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("n,P,Q", [(35, 3, 5), (21, 2, 3)])
def test_mixed_degenerate_stencils_fall_back(n, P, Q):
    """Stencils made singular by duplicated nodes: the mixed form must flag
    them (status FALLBACK) and report the FP64 form's info, which equals the
    oracle's; every other stencil stays within the mixed bar."""
    x, y = R.jittered_grid(2, 32, 0.25, seed=9)
    for k in (100, 400, 777):
        x[k + 1], y[k + 1] = x[k], y[k]
    ja = R.knn_stencils(n, x, y)
    wref, iref = O.assemble(2, n, P, Q, x, y, ja)
    bad = iref != 0
    assert bad.any()
    w64, i64, _ = _run(R.Plan(2, n, P, Q, "diff"), x, y, ja)
    np.testing.assert_array_equal(i64, iref)
    plan = R.Plan(2, n, P, Q, "diff", precision="mixed")
    for v in range(plan.num_variants):
        plan.variant = v
        wm, im, st = _run(plan, x, y, ja)
        miss = np.nonzero(im != iref)[0]
        e = _relerr(wm, wref, n)[~bad]
        print(f"variant {v} ({plan.variant_name()}): {bad.sum()} singular stencils, "
              f"fallback {((st & 4) != 0).sum()}, info mismatches {miss.tolist()[:8]} "
              f"(status {st[miss].tolist()[:8]}, oracle info {iref[miss].tolist()[:8]}), "
              f"max rel err elsewhere {e.max():.3e}")
        np.testing.assert_array_equal(im, iref)
        assert ((st[bad] & 4) != 0).all(), plan.variant_name()
        assert e.max() <= MIXED_TOL


@pytest.mark.parametrize("dim,m,n,P,Q", [(2, 24, 56, 4, 7), (3, 10, 80, 3, 3)])
def test_large_mixed_degenerate_stencils_fall_back(dim, m, n, P, Q):
    """The same for the systems above 64 unknowns (the 2-D n = 56 and the
    3-D n = 80 null-space forms, two nodes per lane in 3-D): stencils made
    singular by duplicated nodes go to the FP64 blocked-LU pass, whose info
    is the oracle's; the rest stay within the mixed bar."""
    pts = [np.array(a) for a in R.jittered_grid(dim, m, 0.25, seed=19)]
    for k in (101, 402, 555):
        for a in pts:
            a[k + 1] = a[k]
    ja = R.knn_stencils(n, *pts)
    z = pts[2] if dim == 3 else None
    wref, iref = O.assemble(dim, n, P, Q, pts[0], pts[1], ja, z=z)
    bad = iref != 0
    assert bad.any()
    plan = R.Plan(dim, n, P, Q, "diff", precision="mixed", max_refine=3)
    for v in range(plan.num_variants):
        plan.variant = v
        wm, im, st = _run(plan, pts[0], pts[1], ja, z=z)
        e = _relerr(wm, wref, n)[~bad]
        print(f"dim={dim} n={n} variant {v} ({plan.variant_name()}): {bad.sum()} singular "
              f"stencils, fallback {((st & 4) != 0).sum()}, max rel err elsewhere {e.max():.3e}")
        np.testing.assert_array_equal(im, iref)
        assert ((st[bad] & 4) != 0).all(), plan.variant_name()
        assert e.max() <= MIXED_TOL


def test_mixed_null_space_converges_in_one_step():
    """The headline shape (n = 35, r^5 + P3) on the null-space variant: no
    stencil falls back and almost every stencil stops after one FP64
    refinement step."""
    n, P, Q = 35, 3, 5
    plan = R.Plan(2, n, P, Q, "diff", precision="mixed")
    assert plan.variant_name(0).startswith("null space")
    x, y = R.jittered_grid(2, 96, 0.25, seed=15)
    ja = R.knn_stencils(n, x, y)
    w, info, st = _run(plan, x, y, ja)
    wref, _ = O.assemble(2, n, P, Q, x, y, ja)
    assert (info == 0).all()
    hist = np.bincount(st & 3, minlength=4)
    print("refinement steps histogram", hist.tolist(), "fallback", int(((st & 4) != 0).sum()))
    assert ((st & 4) == 0).all()
    assert hist[1] >= 0.9 * st.size
    assert _relerr(w, wref, n).max() <= MIXED_TOL


def test_autotune_picks_a_variant():
    import torch
    n, P, Q = 21, 2, 3
    plan = R.Plan(2, n, P, Q, "diff")
    x, y = R.jittered_grid(2, 64, 0.25, seed=12)
    ja = R.knn_stencils(n, x, y)
    w = [torch.empty(ja.size, dtype=torch.float64, device="cuda:0") for _ in range(3)]
    ms = plan.autotune(x.size, _dev(x), _dev(y), _dev(ja), w, reps=2)
    assert len(ms) == plan.num_variants and all(t > 0 for t in ms)
    # the kept variant is within the autotuner's 2% noise margin of the fastest
    assert ms[plan.variant] <= min(ms) / 0.98 * (1 + 1e-6)


def test_interp_d2q9_matches_oracle():
    """Interpolation to the 9 D2Q9 departure points of each node, the
    reference demo's NRHS = 9 (main.f90:16,55: centres 0.5*(x, y))."""
    n, P, Q = 9, 2, 3
    plan = R.Plan(2, n, P, Q, "interp", nrhs=9)
    x, y = R.jittered_grid(2, 32, 0.25, seed=3)
    ja = R.knn_stencils(n, x, y)
    h = 1.0 / 32
    cx = 0.5 * h * np.array([0, 1, 0, -1, 0, 1, -1, -1, 1], float)
    cy = 0.5 * h * np.array([0, 0, 1, 0, -1, 1, 1, -1, -1], float)
    wref, iref = O.assemble(2, n, P, Q, x, y, ja, op_set=1, centres=(cx, cy))
    first = None
    for v in range(plan.num_variants):   # every kernel variant, bit-identical
        plan.variant = v
        w, info, _ = _run(plan, x, y, ja, centres=(cx, cy))
        assert (info == 0).all()
        e = _relerr(w, wref, n)
        print(f"interp variant {v} ({plan.variant_name()}): max rel err {e.max():.3e}")
        assert e.max() <= FP64_TOL
        if first is None:
            first = w
        else:
            for a, b in zip(w, first):
                np.testing.assert_array_equal(a, b)
    # interpolation reproduces the quadratics exactly
    J = ja.reshape(-1, n)
    for c in range(9):
        f = x ** 2 - x * y + 0.5 * y
        v = (w[c].reshape(-1, n) * f[J]).sum(1)
        ex = (x + cx[c]) ** 2 - (x + cx[c]) * (y + cy[c]) + 0.5 * (y + cy[c])
        assert np.abs(v - ex).max() < 1e-11


def test_singular_stencil_reports_info():
    """Exactly duplicated nodes give two identical PHS rows: info > 0, as
    the reference's per-stencil status (main.cpp:48-54)."""
    n, P, Q = 21, 2, 3
    x, y = R.jittered_grid(2, 16, 0.25, seed=5)
    x[10], y[10] = x[11], y[11]
    ja = R.knn_stencils(n, x, y)
    plan = R.Plan(2, n, P, Q, "diff")
    w, info, _ = _run(plan, x, y, ja)
    wref, iref = O.assemble(2, n, P, Q, x, y, ja)
    assert (info != 0).sum() > 0
    np.testing.assert_array_equal(info != 0, iref != 0)
    ok = iref == 0
    assert _relerr([a.reshape(-1, n)[ok] for a in w], [b.reshape(-1, n)[ok] for b in wref],
                   n).max() <= FP64_TOL


@pytest.mark.parametrize("dim,m,n,P,Q", [(2, 32, 56, 4, 7), (3, 10, 80, 3, 3)])
def test_large_systems_match_oracle(dim, m, n, P, Q):
    """Systems larger than a wavefront (71x71 in 2-D, 100x100 in 3-D), every
    engine: the two-wave row-per-lane workgroup and the blocked LU with
    FP64 MFMA trailing updates. Both apply the oracle's operations in the
    oracle's order, so the weights are bit-identical to it."""
    plan = R.Plan(dim, n, P, Q, "diff")
    pts = R.jittered_grid(dim, m, 0.25, seed=13)
    ja = R.knn_stencils(n, *pts)
    z = pts[2] if dim == 3 else None
    wref, iref = O.assemble(dim, n, P, Q, pts[0], pts[1], ja, z=z)
    assert (iref == 0).all()
    assert plan.num_variants >= 2
    for v in range(plan.num_variants):
        plan.variant = v
        w, info, status = _run(plan, pts[0], pts[1], ja, z=z)
        assert (info == 0).all()
        e = _relerr(w, wref, n)
        print(f"dim={dim} n={n} nt={plan.nt} {plan.variant_name()}: max rel err {e.max():.3e}")
        assert e.max() <= FP64_TOL
        for a, b in zip(w, wref):
            np.testing.assert_array_equal(a, b)


def _disk_case(n_int=12000, n_bnd=300, seed=21):
    from rbf_amd import workloads
    return workloads.build_disk(n_int, n_bnd, seed)


@pytest.mark.parametrize("mode", [0, pytest.param(1, marks=pytest.mark.skip(
    reason="persistent launch switched off: a run faulted the GPU (memory aperture "
           "violation) and the cause is not found; set_mode(1) returns UNSUPPORTED"))])
def test_uneven_batch_matches_oracle_and_replays(mode):
    """Graded disk with all four size classes in one ragged graph; the
    captured graph is replayed three times with eager work in between, and
    every replay is checked against the oracle."""
    import torch
    from rbf_amd import workloads
    W = _disk_case()
    ia, ja, x, y = W["ia"], W["ja"], W["x"], W["y"]
    nnz = int(ia[-1])
    wref = [np.zeros(nnz) for _ in range(3)]
    iref = np.zeros(x.size, np.int32)
    for (n, P, Q) in workloads.DISK_CLASSES:
        rows = np.nonzero(W["sizes"] == n)[0].astype(np.int32)
        assert rows.size > 0
        O.assemble_rows(2, n, P, Q, rows, ia, x, y, ja, wref, iref)
    assert (iref == 0).all()
    b = R.Batch(workloads.DISK_CLASSES)
    w = [torch.zeros(nnz, dtype=torch.float64, device="cuda:0") for _ in range(3)]
    info = torch.full((x.size,), -1, dtype=torch.int32, device="cuda:0")
    b.bind(ia, _dev(x), _dev(y), _dev(ia), _dev(ja), w, info=info)
    assert sum(b.class_count(c) for c in range(4)) == x.size
    b.mode = mode
    for rep in range(3):
        for t in w:
            t.zero_()
        info.fill_(-1)
        torch.cuda.synchronize()
        b.launch()
        torch.cuda.synchronize()
        assert (info.cpu().numpy() == 0).all()
        for a, r in zip(w, wref):
            g = a.cpu().numpy()
            err = np.abs(g - r).max() / np.abs(r).max()
            assert err == 0.0 or err <= FP64_TOL, (rep, err)


def test_uneven_batch_autotune():
    import torch
    from rbf_amd import workloads
    W = _disk_case(6000, 200, 22)
    nnz = int(W["ia"][-1])
    b = R.Batch(workloads.DISK_CLASSES)
    w = [torch.zeros(nnz, dtype=torch.float64, device="cuda:0") for _ in range(3)]
    b.bind(W["ia"], _dev(W["x"]), _dev(W["y"]), _dev(W["ia"]), _dev(W["ja"]), w)
    ms = b.autotune(reps=2)
    assert len(ms) == 2 and ms[0] > 0
    assert ms[1] == -1.0 and b.mode == 0  # the persistent mode is switched off
    with pytest.raises(R.RbfError) as e:
        b.mode = 1
    assert e.value.status == R.ERR_UNSUPPORTED and "switched off" in str(e.value)
    assert all(b.class_variant(c) for c in range(4))


def test_uneven_batch_every_class_shape_matches_oracle():
    """Every shape a class's list launch can take (stencils per wave, waves
    per workgroup, engine tuning) gives the oracle's weights bit for bit;
    the classes are numbered in the caller's order, here given largest
    first."""
    import torch
    from rbf_amd import workloads
    W = _disk_case(6000, 200, 24)
    ia, ja, x, y = W["ia"], W["ja"], W["x"], W["y"]
    nnz = int(ia[-1])
    wref = [np.zeros(nnz) for _ in range(3)]
    iref = np.zeros(x.size, np.int32)
    for (n, P, Q) in workloads.DISK_CLASSES:
        rows = np.nonzero(W["sizes"] == n)[0].astype(np.int32)
        O.assemble_rows(2, n, P, Q, rows, ia, x, y, ja, wref, iref)
    classes = list(reversed(workloads.DISK_CLASSES))
    b = R.Batch(classes)
    w = [torch.zeros(nnz, dtype=torch.float64, device="cuda:0") for _ in range(3)]
    info = torch.full((x.size,), -1, dtype=torch.int32, device="cuda:0")
    b.bind(ia, _dev(x), _dev(y), _dev(ia), _dev(ja), w, info=info)
    for c, (n, _, _) in enumerate(classes):
        assert b.class_count(c) == int((W["sizes"] == n).sum())
        nv = b.class_num_variants(c)
        assert nv >= 2 and b.class_variant_name(c, nv) is None
        with pytest.raises(R.RbfError):
            b.set_class_variant(c, nv)
        for v in range(nv):
            b.set_class_variant(c, v)
            assert b.class_variant(c) == b.class_variant_name(c, v)
            for t in w:
                t.zero_()
            info.fill_(-1)
            b.launch()
            torch.cuda.synchronize()
            print(f"class n={n} variant {v}: {b.class_variant(c)}")
            assert (info.cpu().numpy() == 0).all()
            for a, r in zip(w, wref):
                np.testing.assert_array_equal(a.cpu().numpy(), r)
        b.set_class_variant(c, 0)


def test_uneven_batch_failed_bind_leaves_it_unbound():
    """A bind that fails (a row length no class has) leaves the batch
    unbound: launch and autotune refuse it until a bind succeeds."""
    import torch
    from rbf_amd import workloads
    W = _disk_case(3000, 120, 23)
    nnz = int(W["ia"][-1])
    b = R.Batch(workloads.DISK_CLASSES)
    w = [torch.zeros(nnz, dtype=torch.float64, device="cuda:0") for _ in range(3)]
    args = (_dev(W["x"]), _dev(W["y"]), _dev(W["ia"]), _dev(W["ja"]), w)
    b.bind(W["ia"], *args)
    b.launch()
    torch.cuda.synchronize()
    bad = W["ia"].copy()
    bad[1:] += 1  # row 0 one entry longer: no class of that size
    with pytest.raises(R.RbfError) as e:
        b.bind(bad, *args)
    assert e.value.status == R.ERR_ARG
    with pytest.raises(R.RbfError):
        b.launch()
    with pytest.raises(R.RbfError):
        b.autotune(reps=1)
    b.bind(W["ia"], *args)
    for t in w:
        t.zero_()
    b.launch()
    torch.cuda.synchronize()
    wref = [np.zeros(nnz) for _ in range(3)]
    iref = np.zeros(W["x"].size, np.int32)
    for (n, P, Q) in workloads.DISK_CLASSES:
        rows = np.nonzero(W["sizes"] == n)[0].astype(np.int32)
        O.assemble_rows(2, n, P, Q, rows, W["ia"], W["x"], W["y"], W["ja"], wref, iref)
    for a, r in zip(w, wref):
        g = a.cpu().numpy()
        assert np.abs(g - r).max() <= FP64_TOL * np.abs(r).max()


@pytest.mark.parametrize("dim,m,n,P,Q", [(2, 32, 56, 4, 7), (3, 10, 80, 3, 3)])
def test_large_systems_mixed(dim, m, n, P, Q):
    """Mixed precision for systems larger than 64: within 1e-8 of the FP64
    oracle; stencils that do not converge are re-solved in FP64 (and then
    match the oracle exactly)."""
    plan = R.Plan(dim, n, P, Q, "diff", precision="mixed", max_refine=3)
    pts = R.jittered_grid(dim, m, 0.25, seed=14)
    ja = R.knn_stencils(n, *pts)
    z = pts[2] if dim == 3 else None
    w, info, status = _run(plan, pts[0], pts[1], ja, z=z)
    wref, iref = O.assemble(dim, n, P, Q, pts[0], pts[1], ja, z=z)
    assert (info == 0).all()
    e = _relerr(w, wref, n)
    fb = (status & 4) != 0
    print(f"mixed dim={dim} n={n}: max rel err {e.max():.3e}, fallback {fb.mean():.3f}, "
          f"steps {np.bincount(status & 3, minlength=4)}")
    assert e.max() <= MIXED_TOL
    if fb.any():
        assert e[fb].max() <= FP64_TOL
    # the default variants converge: FP64 fallback stays the exception
    assert fb.mean() <= 0.05


# ---------------------------------------------------------------- caller-filled systems

def _demo_system():
    """The reference demo's system (main.f90:22-57): the 9-point stencil
    x, y in {-1, 0, 1}, PHS r^3 + the 6 quadratic monomials (15 x 15,
    rbf_cuda.f90:36-99), interpolation to the 9 centres (0.5 x, 0.5 y)
    (fill_rhs, rbf_cuda.f90:104-136)."""
    x = np.array([0, 1, 0, -1, 0, 1, 1, -1, -1], float)
    y = np.array([0, 0, 1, 0, -1, 1, -1, -1, 1], float)
    M = O.fill(2, 9, 2, 3, x, y)
    B = O.rhs(2, 9, 2, 3, x, y, op_set=1, centres=(0.5 * x, 0.5 * y))
    return M, B


@pytest.mark.parametrize("nt,nrhs,nbatch", [(15, 9, 1), (15, 9, 333), (27, 3, 200), (45, 3, 129),
                                             (64, 16, 65), (7, 0, 10), (33, 5, 17)])
def test_gesv_batched_matches_oracle(nt, nrhs, nbatch):
    """Factors (dgetrf layout), pivots and solution bit-identical to the
    oracle's elimination; random systems plus the reference demo's system."""
    import torch
    rng = np.random.default_rng(nt * 1000 + nrhs)
    Ms = rng.standard_normal((nbatch, nt, nt))
    Bs = rng.standard_normal((nbatch, nt, nrhs))
    if nt == 15 and nrhs == 9:
        Ms[0], Bs[0] = _demo_system()
    A = torch.from_numpy(np.ascontiguousarray(Ms.transpose(0, 2, 1))).to("cuda:0")
    B = torch.from_numpy(np.ascontiguousarray(Bs.transpose(0, 2, 1))).to("cuda:0") if nrhs else None
    ipiv, info = R.gesv_batched(A, B)
    torch.cuda.synchronize()
    LU = A.cpu().numpy().transpose(0, 2, 1)
    X = B.cpu().numpy().transpose(0, 2, 1) if nrhs else None
    ipiv, info = ipiv.cpu().numpy(), info.cpu().numpy()
    for b in range(nbatch):
        LUr, pr, ir = O.getrf(Ms[b])
        np.testing.assert_array_equal(LU[b], LUr)
        np.testing.assert_array_equal(ipiv[b], pr)
        assert info[b] == ir == 0
        if nrhs:
            Xr, _ = O.gesv(Ms[b], Bs[b])
            np.testing.assert_array_equal(X[b], Xr)


def test_gesv_batched_singular_and_args():
    import torch
    M = np.array([[1.0, 2.0, 0.0], [2.0, 4.0, 0.0], [0.0, 0.0, 1.0]])
    A = torch.from_numpy(np.ascontiguousarray(M.T[None])).to("cuda:0")
    B = torch.ones((1, 1, 3), dtype=torch.float64, device="cuda:0")
    ipiv, info = R.gesv_batched(A, B)
    torch.cuda.synchronize()
    assert int(info[0]) == O.gesv(M, np.ones((3, 1)))[1] == 2
    LUr, pr, ir = O.getrf(M)
    np.testing.assert_array_equal(A.cpu().numpy()[0].T, LUr)
    np.testing.assert_array_equal(ipiv.cpu().numpy()[0], pr)
    big = torch.zeros((1, 65, 65), dtype=torch.float64, device="cuda:0")
    with pytest.raises(R.RbfError):
        R.gesv_batched(big)


def test_solve_dp_demo_system():
    """rbfa_solve_dp: the reference's extern "C" solve_dp (15 x 15, NRHS 9)."""
    import ctypes
    import torch
    M, B = _demo_system()
    A = torch.from_numpy(np.ascontiguousarray(M.T)).to("cuda:0")
    Bd = torch.from_numpy(np.ascontiguousarray(B.T)).to("cuda:0")
    ipiv = torch.zeros(15, dtype=torch.int32, device="cuda:0")
    info = torch.full((1,), -1, dtype=torch.int32, device="cuda:0")
    st = R.lib().rbfa_solve_dp(ctypes.c_void_p(A.data_ptr()), ctypes.c_void_p(ipiv.data_ptr()),
                               ctypes.c_void_p(Bd.data_ptr()), ctypes.c_void_p(info.data_ptr()),
                               ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    assert st == 0 and int(info[0]) == 0
    Xr, _ = O.gesv(M, B)
    np.testing.assert_array_equal(Bd.cpu().numpy().T, Xr)
    # interpolation at the stencil's own centre (0, 0) reproduces node 0
    np.testing.assert_allclose(Xr[:9, 0], np.eye(9)[0], atol=1e-12)


@pytest.mark.parametrize("n,P,Q,prec", [(35, 3, 5, "fp64"), (56, 4, 7, "fp64"),
                                        (35, 3, 5, "mixed"), (56, 4, 7, "mixed")])
def test_exact_ties_and_nan_node_match_oracle(n, P, Q, prec):
    """An unjittered grid (exactly tied pivot candidates: equal distances
    give equal PHS values) and one NaN coordinate. FP64: every variant
    bit-identical to the oracle, info included (NaN never wins the pivot;
    with no number among the candidates the first active row is taken and
    info names the step). Mixed: the stencils that hold the NaN node fall
    back to the FP64 solve (status flag) and report the oracle's info; the
    others stay within the mixed bar."""
    m = 24
    x, y = R.jittered_grid(2, m, 0.0, seed=1)
    ja = R.knn_stencils(n, x, y)
    x = x.copy()
    x[m * (m // 2) + m // 2] = np.nan
    wref, iref = O.assemble(2, n, P, Q, x, y, ja)
    nanst = np.any(ja.reshape(-1, n) == m * (m // 2) + m // 2, axis=1)
    assert nanst.any() and (iref[nanst] != 0).all() and (iref[~nanst] == 0).all()
    plan = R.Plan(2, n, P, Q, "diff", precision=prec)
    for v in range(plan.num_variants):
        plan.variant = v
        w, info, st = _run(plan, x, y, ja)
        print(f"{prec} n={n} variant {v} ({plan.variant_name()}): {nanst.sum()} NaN stencils, "
              f"info {np.unique(info[nanst]).tolist()}")
        np.testing.assert_array_equal(info, iref)
        if prec == "fp64":
            for a, b in zip(w, wref):
                np.testing.assert_array_equal(a, b)
        else:
            assert ((st[nanst] & 4) != 0).all()
            assert _relerr(w, wref, n)[~nanst].max() <= MIXED_TOL
