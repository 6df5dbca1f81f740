"""Pinning the CPU oracle (oracle/rbf_oracle.c) before anything is checked
against it.

Each oracle function is checked against something independent of its own
formula: the reference's compiled output (golden vectors from oracle/_ref),
closed forms, numerical derivatives, exact rational arithmetic on tiny
systems, and the invariants of the saddle-point system. The LU itself has
no reference counterpart to run (cuSolverDx is closed and absent): it is
pinned by backward error, an exact rational solve and polynomial
reproduction, not by the reference's bits.
"""
import os
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _cloud(n, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    xs = rng.uniform(-1, 1, n) * scale
    ys = rng.uniform(-1, 1, n) * scale
    xs[0] = ys[0] = 0.0
    r = np.sqrt(xs ** 2 + ys ** 2).max()
    return xs / r, ys / r


# ---------------------------------------------------------------- basis / fill

@pytest.mark.parametrize("Q", [3, 5, 7, 9])
def test_phs_closed_form(Q):
    # two nodes at distance r: M = [[0, r^Q, 1], [r^Q, 0, 1], [1, 1, 0]]
    for r in (0.3, 1.0, 0.7071):
        M = O.fill(2, 2, 0, Q, [0.0, r * 0.6], [0.0, r * 0.8])
        assert M[0, 0] == 0.0 and M[1, 1] == 0.0
        assert M[0, 1] == pytest.approx(r ** Q, rel=4e-16 * Q)
        np.testing.assert_array_equal(M[:2, 2], [1.0, 1.0])
        np.testing.assert_array_equal(M[2, :2], [1.0, 1.0])
        assert M[2, 2] == 0.0


def test_fill_structure_2d():
    n, P, Q = 35, 3, 5
    xs, ys = _cloud(n, 1)
    M = O.fill(2, n, P, Q, xs, ys)
    A = M[:n, :n]
    np.testing.assert_array_equal(A, A.T)                     # symmetric, bitwise
    np.testing.assert_array_equal(np.diag(A), 0.0)            # phi(0) = 0
    np.testing.assert_array_equal(M[:n, n:], M[n:, :n].T)     # P and P^T blocks
    np.testing.assert_array_equal(M[n:, n:], 0.0)             # zero block
    # degree-major order of rbf_operators.h:93-98: 1, x, y, x^2, xy, y^2, x^3, x^2y, xy^2, y^3
    expo = [(0, 0), (1, 0), (0, 1), (2, 0), (1, 1), (0, 2), (3, 0), (2, 1), (1, 2), (0, 3)]
    for t, (a, b) in enumerate(expo):
        np.testing.assert_allclose(M[:n, n + t], xs ** a * ys ** b, rtol=4e-16, atol=0)
    # entries are r^Q of the pairwise distances
    d = np.hypot(xs[:, None] - xs[None, :], ys[:, None] - ys[None, :])
    np.testing.assert_allclose(A, d ** Q, rtol=2e-15, atol=1e-300)


def test_fill_structure_3d():
    n, P, Q = 30, 2, 3
    rng = np.random.default_rng(4)
    xs, ys, zs = rng.uniform(-0.5, 0.5, (3, n))
    M = O.fill(3, n, P, Q, xs, ys, zs)
    assert M.shape == (n + 10, n + 10)
    expo = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 0, 0), (1, 1, 0), (1, 0, 1),
            (0, 2, 0), (0, 1, 1), (0, 0, 2)]
    for t, (a, b, c) in enumerate(expo):
        np.testing.assert_allclose(M[:n, n + t], xs ** a * ys ** b * zs ** c, rtol=4e-16, atol=0)
    d = np.sqrt((xs[:, None] - xs) ** 2 + (ys[:, None] - ys) ** 2 + (zs[:, None] - zs) ** 2)
    np.testing.assert_allclose(M[:n, :n], d ** 3, rtol=2e-15, atol=1e-300)


def _num_derivs(f, h=2e-4):
    """Central differences at the origin: Laplacian and gradient of f(x, y, z)."""
    e = np.eye(3) * h
    f0 = f(np.zeros(3))
    lap = sum((f(e[i]) - 2 * f0 + f(-e[i])) / h ** 2 for i in range(3))
    grad = [(f(e[i]) - f(-e[i])) / (2 * h) for i in range(3)]
    return lap, grad


@pytest.mark.parametrize("dim,Q,P", [(2, 3, 2), (2, 5, 3), (2, 7, 4), (3, 3, 3), (3, 5, 2)])
def test_derivative_rhs_against_numerical_derivatives(dim, Q, P):
    n = 12
    rng = np.random.default_rng(Q + 10 * dim)
    pts = rng.uniform(-1, 1, (3, n))
    if dim == 2:
        pts[2] = 0
    pts /= np.sqrt((pts ** 2).sum(0)).max()
    B = O.rhs(dim, n, P, Q, pts[0], pts[1], pts[2] if dim == 3 else None)
    for k in range(n):
        pk = pts[:, k]

        def phi(p, pk=pk):
            return np.sqrt(((p[:dim] - pk[:dim]) ** 2).sum()) ** Q

        lap, grad = _num_derivs(phi)
        assert B[k, 0] == pytest.approx(lap, rel=1e-5, abs=1e-6)
        for a in range(dim):
            assert B[k, 1 + a] == pytest.approx(grad[a], rel=1e-5, abs=1e-7)
    m = O.npoly(dim, P)
    for t in range(m):
        e = np.zeros(m)
        e[t] = 1.0

        def poly(p, e=e):
            M1 = O.fill(dim, 1, P, 3, [p[0]], [p[1]], [p[2]] if dim == 3 else None)
            return float(M1[0, 1:] @ e)

        lap, _ = _num_derivs(poly, h=1e-4)
        _, grad = _num_derivs(poly, h=1e-5)
        assert B[n + t, 0] == pytest.approx(lap, abs=1e-6)
        for a in range(dim):
            assert B[n + t, 1 + a] == pytest.approx(grad[a], abs=1e-9)


def test_interp_rhs_is_fill_column_at_the_centre():
    # interpolating at a stencil node reproduces that node's matrix column
    n, P, Q = 9, 2, 3
    xs, ys = _cloud(n, 2)
    M = O.fill(2, n, P, Q, xs, ys)
    B = O.rhs(2, n, P, Q, xs, ys, op_set=1, centres=(xs[[3, 5]], ys[[3, 5]]))
    np.testing.assert_array_equal(B[:, 0], M[:, 3])
    np.testing.assert_array_equal(B[:, 1], M[:, 5])


# ---------------------------------------------------------------- LU solve

def _exact_solve(M, B):
    """Gauss-Jordan in exact rationals of the FP64 entries."""
    nt, r = B.shape
    A = [[Fraction(float(M[i, j])) for j in range(nt)] + [Fraction(float(B[i, c])) for c in range(r)]
         for i in range(nt)]
    for k in range(nt):
        p = next(i for i in range(k, nt) if A[i][k] != 0)
        A[k], A[p] = A[p], A[k]
        for i in range(nt):
            if i != k and A[i][k] != 0:
                f = A[i][k] / A[k][k]
                A[i] = [a - f * b for a, b in zip(A[i], A[k])]
    return np.array([[float(A[i][nt + c] / A[i][i]) for c in range(r)] for i in range(nt)])


@pytest.mark.parametrize("n,P,Q", [(6, 1, 3), (8, 1, 5), (7, 2, 3)])
def test_gesv_against_exact_rational_solve(n, P, Q):
    xs, ys = _cloud(n, n + Q)
    M = O.fill(2, n, P, Q, xs, ys)
    B = O.rhs(2, n, P, Q, xs, ys)
    X, info = O.gesv(M, B)
    assert info == 0
    Xe = _exact_solve(M, B)
    kappa = np.linalg.cond(M)
    err = np.abs(X - Xe).max(0) / np.abs(Xe).max(0)
    assert (err <= 50 * kappa * 2.2e-16).all(), (err, kappa)


@pytest.mark.parametrize("n,P,Q", [(21, 2, 3), (35, 3, 5), (56, 4, 7)])
def test_gesv_backward_error(n, P, Q):
    xs, ys = _cloud(n, 3)
    M = O.fill(2, n, P, Q, xs, ys)
    B = O.rhs(2, n, P, Q, xs, ys)
    X, info = O.gesv(M, B)
    assert info == 0
    nt = M.shape[0]
    res = np.abs(M @ X - B).max(0)
    bound = 4 * nt * 2.2e-16 * (np.abs(M).sum(1).max() * np.abs(X).max(0) + np.abs(B).max(0))
    assert (res <= bound).all()
    # a transposed operand or a dropped RHS term does not pass
    Xt, _ = O.gesv(M.T.copy() + np.triu(np.ones_like(M), 1) * 1e-3, B)
    assert np.abs(M @ Xt - B).max() > 1e3 * bound.max()


def test_gesv_partial_pivoting_handles_zero_diagonal():
    # PHS matrices have phi(0) = 0 on the whole diagonal: no pivoting would fail
    M = np.array([[0.0, 2.0], [3.0, 0.0]])
    X, info = O.gesv(M, np.array([[4.0], [9.0]]))
    assert info == 0
    np.testing.assert_array_equal(X[:, 0], [3.0, 2.0])


@pytest.mark.parametrize("nt", [5, 15, 45, 64])
def test_getrf_layout_matches_lapack(nt):
    """The factors and pivots GESV leaves in A and ipiv, in dgetrf's layout:
    pivots equal scipy's dgetrf (no ties in a random matrix), factors agree
    to rounding, and P A = L U holds with the 1-based sequential swaps."""
    import scipy.linalg as sl
    A = np.random.default_rng(nt).standard_normal((nt, nt))
    LU, ipiv, info = O.getrf(A)
    lu, piv = sl.lu_factor(A)
    assert info == 0
    np.testing.assert_array_equal(ipiv - 1, piv)
    assert np.abs(LU - lu).max() <= 1e-12 * np.abs(lu).max()
    L = np.tril(LU, -1) + np.eye(nt)
    U = np.triu(LU)
    PA = A.copy()
    for k, q in enumerate(ipiv - 1):
        PA[[k, q]] = PA[[q, k]]
    assert np.abs(PA - L @ U).max() <= 64 * nt * np.finfo(float).eps * np.abs(A).max()
    # a transposed operand gives other pivots
    _, ipiv_t, _ = O.getrf(A.T)
    assert (ipiv_t != ipiv).any()
    # the same elimination as the solve: X from the factors equals oracle_gesv's
    B = np.random.default_rng(1).standard_normal((nt, 3))
    X, _ = O.gesv(A, B)
    np.testing.assert_allclose(sl.lu_solve((LU, ipiv - 1), B), X, rtol=0, atol=1e-10 * np.abs(X).max())


def test_gesv_singular_info():
    M = np.array([[1.0, 2.0, 0.0], [2.0, 4.0, 0.0], [0.0, 0.0, 1.0]])
    _, info = O.gesv(M, np.ones((3, 1)))
    assert info == 2                     # second pivot is exactly zero


# ---------------------------------------------------------------- assembled weights

@pytest.mark.parametrize("m,n,P,Q", [(32, 21, 2, 3), (32, 35, 3, 5), (24, 56, 4, 7)])
def test_assembled_weights_reproduce_polynomials(m, n, P, Q):
    import rbf_amd as R
    x, y = R.jittered_grid(2, m, 0.25, seed=9)
    ja = R.knn_stencils(n, x, y)
    w, info = O.assemble(2, n, P, Q, x, y, ja)
    assert (info == 0).all()
    J = ja.reshape(-1, n)
    dx = x[J] - x[:, None]
    dy = y[J] - y[:, None]
    for a in range(P + 1):
        for b in range(P + 1 - a):
            p = dx ** a * dy ** b
            lap = 2.0 if (a, b) in ((2, 0), (0, 2)) else 0.0
            gx = 1.0 if (a, b) == (1, 0) else 0.0
            gy = 1.0 if (a, b) == (0, 1) else 0.0
            h = 1.0 / m
            for wi, ex, s in ((w[0], lap, h ** -2), (w[1], gx, h ** -1), (w[2], gy, h ** -1)):
                v = (wi.reshape(-1, n) * p).sum(1)
                assert np.abs(v - ex).max() <= 1e-9 * s * h ** (a + b) + 1e-12


def test_assembled_laplacian_converges():
    import rbf_amd as R
    errs = []
    for m in (32, 64):
        x, y = R.jittered_grid(2, m, 0.25, seed=1)
        ja = R.knn_stencils(35, x, y)
        w, _ = O.assemble(2, 35, 3, 5, x, y, ja)
        f = np.exp(x + y ** 2)
        v = (w[0].reshape(-1, 35) * f[ja.reshape(-1, 35)]).sum(1)
        ex = (3 + 4 * y ** 2) * f
        inner = (x > 0.25) & (x < 0.75) & (y > 0.25) & (y < 0.75)
        errs.append(np.abs(v - ex)[inner].max() / np.abs(ex).max())
    assert errs[1] < errs[0] / 2.5          # P = 3: second order in h
    assert errs[1] < 1e-3


# ---------------------------------------------------------------- keys and kNN

def test_keys_match_reference_fortran():
    g = np.load(os.path.join(GOLD, "keys_4096.npz"))
    x, y = g["x"], g["y"]
    bb = [x.min(), y.min(), x.max(), y.max()]
    for i in range(0, x.size, 3):
        assert O.zidx(x[i], y[i], bb, 16) == g["morton"][i]
        assert O.hidx(x[i], y[i], bb, 16) == g["hilbert"][i]


def _interleave(ix, iy, ndiv):
    z = 0
    for b in range(ndiv - 1, -1, -1):
        z = (z << 2) | (((iy >> b) & 1) << 1) | ((ix >> b) & 1)
    return z


def test_morton_known_answer_and_hilbert_adjacency():
    # test_reorder.cpp:72-102: 4x4 cell centres in [0,4]^2, ndiv 2
    bb = [0, 0, 4, 4]
    hk = {}
    for j in range(4):
        for i in range(4):
            assert O.zidx(i + 0.5, j + 0.5, bb, 2) == _interleave(i, j, 2)
            hk[O.hidx(i + 0.5, j + 0.5, bb, 2)] = (i, j)
    assert sorted(hk) == list(range(16))
    for k in range(1, 16):
        (a, b), (c, d) = hk[k - 1], hk[k]
        assert abs(a - c) + abs(b - d) == 1


def test_knn_brute_matches_reference_nanoflann():
    g = np.load(os.path.join(GOLD, "ja_4096_n21.npz"))
    np.testing.assert_array_equal(O.knn_brute(21, g["x"], g["y"]), g["ja"])


def test_knn_brute_matches_reference_ckdtree_3d():
    g = np.load(os.path.join(GOLD, "ja_cube_1000_k80.npz"))
    np.testing.assert_array_equal(O.knn_brute(80, g["x"], g["y"], g["z"]), g["ja"])


def test_apply_against_scipy_csr_and_closed_form():
    """The oracle's mv_op (main.cpp:59-66) against scipy's CSR product on a
    ragged random matrix, the boundary rows against the identity, and, on
    assembled weights, the Laplacian of x^2 + y^2 (= 4) as a closed form."""
    import scipy.sparse as sp
    rng = np.random.default_rng(5)
    nr = 300
    lens = rng.integers(1, 40, nr)
    ia = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ja = rng.integers(0, nr, ia[-1]).astype(np.int32)
    w = rng.standard_normal(ia[-1])
    u = rng.standard_normal(nr)
    A = sp.csr_matrix((w, ja, ia), shape=(nr, nr))
    y, scale = O.apply(ia, ja, w, u)
    np.testing.assert_allclose(y, A @ u, rtol=0, atol=1e-13 * scale.max())
    absA = sp.csr_matrix((np.abs(w), ja, ia), shape=(nr, nr))
    np.testing.assert_allclose(scale, absA @ np.abs(u), rtol=1e-14)
    bnd = np.array([3, 77, 299], np.int32)
    y2, _ = O.apply(ia, ja, w, u, bnd)
    np.testing.assert_array_equal(y2[bnd], u[bnd])
    keep = np.setdiff1d(np.arange(nr), bnd)
    np.testing.assert_array_equal(y2[keep], y[keep])
    x, yy = np.meshgrid((np.arange(24) + 0.5) / 24, (np.arange(24) + 0.5) / 24)
    x, yy = x.ravel(), yy.ravel()
    jn = O.knn_brute(21, x, yy)
    wl, info = O.assemble(2, 21, 2, 3, x, yy, jn)
    lap, _ = O.apply(np.arange(0, jn.size + 1, 21), jn, wl[0], x * x + yy * yy)
    np.testing.assert_allclose(lap, 4.0, atol=1e-8)
