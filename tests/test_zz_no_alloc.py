"""rbfa_assemble_noalloc against rbfa_assemble: the same kernels, with the
fallback list of the mixed-form shapes in the library's fixed per-device
buffer instead of the plan's workspace. Outputs must be identical, on two
calls in a row (the second on fewer stencils, so the buffer still holds the
first call's list), for shapes with and without a fallback list; calls past
the buffer's limit are refused before any launch."""
import ctypes

import numpy as np
import pytest

import rbf_amd as R

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def _case(dim, m, n, seed, dup):
    pts = list(R.jittered_grid(dim, m, 0.25, seed=seed))
    # duplicated nodes make the stencils holding both singular: the mixed
    # form lists them for the FP64 pass, which reports info > 0
    for k in dup:
        for c in pts:
            c[k] = c[k + 1]
    ja = R.knn_stencils(n, *pts)
    return pts, ja


def _assemble(plan, pts, ja, noalloc):
    import torch
    nst = ja.size // plan.n
    w = [torch.full((ja.size,), -3.0, dtype=torch.float64, device="cuda:0")
         for _ in range(plan.nops)]
    info = torch.full((nst,), -7, dtype=torch.int32, device="cuda:0")
    status = torch.full((nst,), 255, dtype=torch.uint8, device="cuda:0")
    z = _dev(pts[2]) if len(pts) == 3 else None
    plan.assemble(nst, _dev(pts[0]), _dev(pts[1]), _dev(ja), w, z=z, info=info, status=status,
                  noalloc=noalloc)
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in w], info.cpu().numpy(), status.cpu().numpy()


@pytest.mark.parametrize("dim,n,P,Q,precision,m1,m2", [
    (2, 56, 4, 7, "mixed", 40, 24),
    (3, 80, 3, 3, "mixed", 11, 8),
    (2, 35, 3, 5, "mixed", 64, 40),
    (2, 35, 3, 5, "fp64", 64, 40),
])
def test_noalloc_matches_assemble_on_two_calls(dim, n, P, Q, precision, m1, m2):
    plan = R.Plan(dim, n, P, Q, "diff", precision=precision, max_refine=3)
    for call, m in enumerate((m1, m2)):
        dup = (100, 400, 777) if call == 0 else (50, 333)
        pts, ja = _case(dim, m, n, 60 + call, dup)
        nst = ja.size // n
        print(f"noalloc dim={dim} n={n} {precision} call {call}: {nst} stencils", flush=True)
        w0, i0, s0 = _assemble(plan, pts, ja, noalloc=False)
        w1, i1, s1 = _assemble(plan, pts, ja, noalloc=True)
        np.testing.assert_array_equal(i1, i0)
        np.testing.assert_array_equal(s1, s0)
        for a, b in zip(w1, w0):
            np.testing.assert_array_equal(a, b)
        assert (i0 != 0).any()  # the duplicated nodes were seen
        if precision == "mixed" and n > 35:
            assert ((s0 & 4) != 0).any()  # and went through the fallback list


def test_noalloc_refuses_calls_past_its_buffer():
    plan = R.Plan(2, 56, 4, 7, "diff", precision="mixed", max_refine=3)
    L = R.lib()
    d = ctypes.c_void_p(64)  # never dereferenced: the size check comes first
    tab = (ctypes.c_void_p * 3)(64, 64, 64)
    big = (1 << 22) + 1
    assert L.rbfa_assemble_noalloc(plan._h, big, d, d, None, d, tab, None, None, None, None,
                                   None, None) == R.ERR_ARG
    assert b"4194304" in L.rbfa_last_error()
    assert L.rbfa_assemble_noalloc(plan._h, 0, None, None, None, None, None, None, None, None,
                                   None, None, None) == R.OK
