"""GPU parity of the packed engine (several stencils per wave64, one per
16- or 32-lane segment) through the C ABI: bit-identical to the CPU FP64
oracle and to the single-wave engine, on stencil counts that leave the
last wave partly empty, on singular stencils (info as the oracle's) and on
batches smaller than one wave's worth of stencils (nothing written past
the last stencil).
"""
import numpy as np
import pytest

import rbf_amd as R
from oracle import oracle as O

pytestmark = pytest.mark.gpu

# (op_set, n, P, Q, nrhs): every shipped shape with a packed variant
PACKED = [("diff", 13, 1, 3, 0), ("diff", 13, 2, 3, 0), ("diff", 21, 2, 3, 0),
          ("diff", 21, 2, 5, 0), ("interp", 9, 2, 3, 9), ("interp", 21, 2, 3, 9)]


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def _centres(h):
    cx = 0.5 * h * np.array([0, 1, 0, -1, 0, 1, -1, -1, 1], float)
    cy = 0.5 * h * np.array([0, 0, 1, 0, -1, 1, 1, -1, -1], float)
    return cx, cy


def _plan(op, n, P, Q, nrhs):
    plan = R.Plan(2, n, P, Q, op, nrhs=nrhs) if op == "interp" else R.Plan(2, n, P, Q, op)
    names = [plan.variant_name(v) for v in range(plan.num_variants)]
    packed = [v for v, s in enumerate(names) if s.startswith("packed")]
    assert packed, names
    return plan, packed


def _run(plan, nst, x, y, ja, centres, pad=0):
    """Weights of stencils 0..nst-1 in arrays with pad extra entries (NaN)
    after the last stencil's row."""
    import torch
    size = nst * plan.n + pad
    w = [torch.full((size,), np.nan, dtype=torch.float64, device="cuda:0")
         for _ in range(plan.nops)]
    info = torch.full((nst + 4,), -7, dtype=torch.int32, device="cuda:0")
    status = torch.full((nst + 4,), 255, dtype=torch.uint8, device="cuda:0")
    cen = None if centres is None else tuple(_dev(c) for c in centres)
    plan.assemble(nst, _dev(x), _dev(y), _dev(ja[:nst * plan.n]), w, centres=cen, info=info,
                  status=status)
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in w], info.cpu().numpy(), status.cpu().numpy()


def _oracle(op, n, P, Q, x, y, ja, centres):
    if op == "interp":
        return O.assemble(2, n, P, Q, x, y, ja, op_set=1, centres=centres)
    return O.assemble(2, n, P, Q, x, y, ja)


@pytest.mark.parametrize("op,n,P,Q,nrhs", PACKED)
def test_packed_bit_identical_to_oracle_and_single_wave(op, n, P, Q, nrhs):
    m = 33  # 1089 stencils: the last wave holds 1 of its 4 (or 2) stencils
    x, y = R.jittered_grid(2, m, 0.25, seed=21)
    ja = R.knn_stencils(n, x, y)
    nst = ja.size // n
    cen = _centres(1.0 / m) if op == "interp" else None
    wref, iref = _oracle(op, n, P, Q, x, y, ja, cen)
    assert (iref == 0).all()
    plan, packed = _plan(op, n, P, Q, nrhs)
    plan.variant = 0
    w0, i0, _ = _run(plan, nst, x, y, ja, cen)
    for vp in packed:
        plan.variant = vp
        wp, ip, sp = _run(plan, nst, x, y, ja, cen)
        print(f"{op} n={n} P={P} Q={Q}: {plan.variant_name()}")
        np.testing.assert_array_equal(ip[:nst], iref)
        assert (sp[:nst] == 0).all()
        for c in range(plan.nops):
            np.testing.assert_array_equal(wp[c], wref[c])
            np.testing.assert_array_equal(wp[c], w0[c])


@pytest.mark.parametrize("op,n,P,Q,nrhs", [PACKED[0], PACKED[2], PACKED[4]])
def test_packed_singular_stencils_report_the_oracle_info(op, n, P, Q, nrhs):
    """Duplicated nodes make some systems singular: info as the oracle's
    (first zero pivot), every other stencil still bit-identical."""
    m = 24
    x, y = R.jittered_grid(2, m, 0.25, seed=9)
    for k in (37, 100, 301, 302):
        x[k + 1], y[k + 1] = x[k], y[k]
    ja = R.knn_stencils(n, x, y)
    nst = ja.size // n
    cen = _centres(1.0 / m) if op == "interp" else None
    wref, iref = _oracle(op, n, P, Q, x, y, ja, cen)
    bad = iref != 0
    assert bad.any()
    plan, packed = _plan(op, n, P, Q, nrhs)
    for vp in packed:
        plan.variant = vp
        w, info, _ = _run(plan, nst, x, y, ja, cen)
        print(f"{op} n={n} {plan.variant_name()}: {bad.sum()} singular stencils")
        np.testing.assert_array_equal(info[:nst], iref)
        for c in range(plan.nops):
            a = w[c].reshape(-1, n)[~bad]
            b = wref[c].reshape(-1, n)[~bad]
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("op,n,P,Q,nrhs", [PACKED[0], PACKED[4], PACKED[5]])
@pytest.mark.parametrize("nst", [1, 2, 3, 5])
def test_packed_small_batches_write_nothing_past_the_last_stencil(op, n, P, Q, nrhs, nst):
    m = 16
    x, y = R.jittered_grid(2, m, 0.25, seed=4)
    ja = R.knn_stencils(n, x, y)
    cen = _centres(1.0 / m) if op == "interp" else None
    wref, iref = _oracle(op, n, P, Q, x, y, ja, cen)
    plan, packed = _plan(op, n, P, Q, nrhs)
    for vp in packed:
        plan.variant = vp
        w, info, status = _run(plan, nst, x, y, ja, cen, pad=4 * n)
        np.testing.assert_array_equal(info[:nst], iref[:nst])
        assert (info[nst:] == -7).all() and (status[nst:] == 255).all()
        for c in range(plan.nops):
            np.testing.assert_array_equal(w[c][:nst * n], wref[c][:nst * n])
            assert np.isnan(w[c][nst * n:]).all()


@pytest.mark.parametrize("op,n,P,Q,nrhs", [PACKED[0], PACKED[2], PACKED[4]])
def test_packed_exact_ties_and_nan_stencils(op, n, P, Q, nrhs):
    """An unjittered grid gives exactly tied pivot candidates (equal
    distances, equal PHS values): the first row among them must win, as in
    the oracle. A NaN coordinate makes every candidate of its stencils NaN:
    the first active row is taken and info is the step, as in the oracle."""
    m = 20
    x, y = R.jittered_grid(2, m, 0.0, seed=1)
    ja = R.knn_stencils(n, x, y)
    x = x.copy()
    x[123] = np.nan
    nst = ja.size // n
    cen = _centres(1.0 / m) if op == "interp" else None
    wref, iref = _oracle(op, n, P, Q, x, y, ja, cen)
    nanst = np.any(ja.reshape(-1, n) == 123, axis=1)
    assert nanst.any() and (~nanst).sum() > 300
    plan, packed = _plan(op, n, P, Q, nrhs)
    plan.variant = [v for v in range(plan.num_variants) if v not in packed][0]
    w0, i0, _ = _run(plan, nst, x, y, ja, cen)
    np.testing.assert_array_equal(i0[:nst], iref)
    for vp in packed:
        plan.variant = vp
        w, info, _ = _run(plan, nst, x, y, ja, cen)
        print(f"{op} n={n} {plan.variant_name()}: {nanst.sum()} stencils with a NaN node, "
              f"info {np.unique(info[:nst][nanst]).tolist()}")
        np.testing.assert_array_equal(info[:nst], iref)
        for c in range(plan.nops):
            np.testing.assert_array_equal(w[c], wref[c])
            np.testing.assert_array_equal(w[c], w0[c])
