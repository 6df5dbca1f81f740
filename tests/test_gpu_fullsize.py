"""GPU parity at the benchmarked sizes: every stencil of the 1M-node
workloads (jittered 1024 x 1024 unit square, Hilbert order, device-free
host kNN) against the CPU FP64 oracle run over all of them.

- n = 35, r^5 + P3 (45 x 45): FP64 form bit-identical, mixed form within
  1e-8 (max over stencils and operators of max_k |w - w_ref| / max_k |w_ref|).
- n = 56, r^7 + P4 (71 x 71): mixed form within 1e-8, FP64 fallback rare.
- D2Q9 interpolation (n = 9, r^3 + P2, nine centres): bit-identical.
- 3-D n = 80, r^3 + P3 (100 x 100, 2,097,152 stencils of the 128^3 Morton
  cube): the mixed form within 1e-8 and the FP64 form bit-identical, on a
  131,072-stencil contiguous sample of the oracle (the whole cube takes the
  oracle minutes).
- The uneven batch (graded disk, 2,008,000 stencils in four size classes,
  ragged CSR, captured hipGraph): bit-identical over every stencil.
"""
import numpy as np
import pytest

import rbf_amd as R
from oracle import oracle as O

pytestmark = pytest.mark.gpu

MIXED_TOL = 1e-8


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def _assemble(plan, W, centres=None):
    import torch
    nst = W["nstencils"]
    size = nst * plan.n
    w = [torch.full((size,), np.nan, dtype=torch.float64, device="cuda:0")
         for _ in range(plan.nops)]
    info = torch.full((nst,), -7, dtype=torch.int32, device="cuda:0")
    status = torch.full((nst,), 255, dtype=torch.uint8, device="cuda:0")
    cen = None if centres is None else tuple(_dev(c) for c in centres)
    z = None if W.get("z") is None else _dev(W["z"])
    plan.assemble(nst, _dev(W["x"]), _dev(W["y"]), _dev(W["ja"]), w, z=z, centres=cen,
                  info=info, status=status)
    torch.cuda.synchronize()
    out = [t.cpu().numpy() for t in w]
    del w
    return out, info.cpu().numpy(), status.cpu().numpy()


def _relerr(w, wref, n):
    e = [np.abs(a.reshape(-1, n) - b.reshape(-1, n)).max(1) /
         np.abs(b.reshape(-1, n)).max(1) for a, b in zip(w, wref)]
    return np.stack(e, 1)


@pytest.fixture(scope="module")
def square_n35():
    from rbf_amd import workloads
    W = workloads.build("square_1m_n35")
    assert W["nstencils"] == 1 << 20
    ja = W["ja"].reshape(-1, 35)
    assert (ja[:, 0] == np.arange(1 << 20)).all()  # self first, n per row
    wref, iref = O.assemble(2, 35, 3, 5, W["x"], W["y"], W["ja"])
    assert (iref == 0).all()
    return W, wref


def test_headline_fp64_every_stencil_bit_identical(square_n35):
    W, wref = square_n35
    plan = R.Plan(2, 35, 3, 5, "diff")
    w, info, status = _assemble(plan, W)
    print(f"fp64 n=35, 1M stencils, {plan.variant_name()}")
    assert (info == 0).all() and (status == 0).all()
    for a, b in zip(w, wref):
        np.testing.assert_array_equal(a, b)


def test_headline_mixed_every_stencil_within_1e8(square_n35):
    W, wref = square_n35
    plan = R.Plan(2, 35, 3, 5, "diff", precision="mixed", max_refine=3)
    w, info, status = _assemble(plan, W)
    assert (info == 0).all()
    e = _relerr(w, wref, 35)
    fb = (status & 4) != 0
    print(f"mixed n=35, 1M stencils, {plan.variant_name()}: max rel err {e.max():.3e}, "
          f"median {np.median(e):.3e}, fallback {int(fb.sum())}, "
          f"steps {np.bincount(status & 3, minlength=4)}")
    assert e.max() <= MIXED_TOL
    assert fb.mean() <= 1e-3


def test_n56_mixed_every_stencil_within_1e8():
    from rbf_amd import workloads
    W = workloads.build("square_1m_n56")
    plan = R.Plan(2, 56, 4, 7, "diff", precision="mixed", max_refine=3)
    w, info, status = _assemble(plan, W)
    wref, iref = O.assemble(2, 56, 4, 7, W["x"], W["y"], W["ja"])
    assert (iref == 0).all() and (info == 0).all()
    e = _relerr(w, wref, 56)
    fb = (status & 4) != 0
    print(f"mixed n=56, 1M stencils, {plan.variant_name()}: max rel err {e.max():.3e}, "
          f"fallback {int(fb.sum())}, steps {np.bincount(status & 3, minlength=4)}")
    assert e.max() <= MIXED_TOL
    assert fb.mean() <= 1e-3


def test_d2q9_interpolation_every_stencil_bit_identical():
    from rbf_amd import workloads
    W = workloads.build("interp_1m_d2q9")
    cen = workloads.d2q9_centres(1.0 / 1024)
    plan = R.Plan(2, 9, 2, 3, "interp", nrhs=9)
    w, info, status = _assemble(plan, W, centres=cen)
    wref, iref = O.assemble(2, 9, 2, 3, W["x"], W["y"], W["ja"], op_set=1, centres=cen)
    print(f"D2Q9 interpolation, 1M stencils, {plan.variant_name()}")
    assert (iref == 0).all() and (info == 0).all() and (status == 0).all()
    for a, b in zip(w, wref):
        np.testing.assert_array_equal(a, b)


@pytest.fixture(scope="module")
def cube_n80():
    from rbf_amd import workloads
    W = workloads.build("cube_2m_n80")
    assert W["nstencils"] == 1 << 21
    s0, cnt = 1 << 20, 1 << 17  # a contiguous Morton block from the middle of the cube
    wref, iref = O.assemble(3, 80, 3, 3, W["x"], W["y"], W["ja"], z=W["z"], s0=s0, count=cnt)
    assert (iref == 0).all()
    return W, wref, s0, cnt


def test_cube_mixed_within_1e8(cube_n80):
    W, wref, s0, cnt = cube_n80
    plan = R.Plan(3, 80, 3, 3, "diff", precision="mixed", max_refine=3)
    w, info, status = _assemble(plan, W)
    assert (info == 0).all()
    fb = (status & 4) != 0
    sl = slice(s0 * 80, (s0 + cnt) * 80)
    e = _relerr([a[sl] for a in w], wref, 80)
    print(f"mixed 3-D n=80, 2M stencils, {plan.variant_name()}: max rel err {e.max():.3e} "
          f"(sample of {cnt}), fallback {int(fb.sum())}, steps {np.bincount(status & 3, minlength=4)}")
    assert e.max() <= MIXED_TOL
    assert fb.mean() <= 1e-3


def test_cube_fp64_bit_identical(cube_n80):
    W, wref, s0, cnt = cube_n80
    plan = R.Plan(3, 80, 3, 3, "diff")
    w, info, status = _assemble(plan, W)
    print(f"fp64 3-D n=80, 2M stencils, {plan.variant_name()}")
    assert (info == 0).all() and (status == 0).all()
    sl = slice(s0 * 80, (s0 + cnt) * 80)
    for a, b in zip(w, wref):
        np.testing.assert_array_equal(a[sl], b)


def test_disk_uneven_batch_every_stencil_bit_identical():
    """The uneven batch at its benchmarked size: the graded unit disk of
    2,000,000 rejection-sampled interior nodes (spacing 1:4 from the centre
    to the rim) and 8,000 rim nodes, four stencil-size classes in one ragged
    CSR graph, run as the captured hipGraph of per-class launches (each
    class's default shape), against the oracle over every stencil."""
    import torch
    from rbf_amd import workloads
    W = workloads.build_disk()
    ia, ja, x, y = W["ia"], W["ja"], W["x"], W["y"]
    assert x.size == 2_008_000
    nnz = int(ia[-1])
    wref = [np.zeros(nnz) for _ in range(3)]
    iref = np.zeros(x.size, np.int32)
    for (n, P, Q) in workloads.DISK_CLASSES:
        rows = np.nonzero(W["sizes"] == n)[0].astype(np.int32)
        assert rows.size > 0
        O.assemble_rows(2, n, P, Q, rows, ia, x, y, ja, wref, iref)
    assert (iref == 0).all()
    b = R.Batch(workloads.DISK_CLASSES)
    w = [torch.full((nnz,), np.nan, dtype=torch.float64, device="cuda:0") for _ in range(3)]
    info = torch.full((x.size,), -1, dtype=torch.int32, device="cuda:0")
    b.bind(ia, _dev(x), _dev(y), _dev(ia), _dev(ja), w, info=info)
    counts = [b.class_count(c) for c in range(len(workloads.DISK_CLASSES))]
    assert sum(counts) == x.size
    b.launch()
    torch.cuda.synchronize()
    print(f"disk, {x.size} stencils, class counts {counts}, "
          f"{[b.class_variant(c) for c in range(len(counts))]}")
    assert (info.cpu().numpy() == 0).all()
    for a, r in zip(w, wref):
        np.testing.assert_array_equal(a.cpu().numpy(), r)
