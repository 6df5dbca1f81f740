"""Host node layer of librbf_amd (the producer of the CSR structure) against
the reference: golden vectors from the reference's compiled host code, the
brute-force search of test_spatial.cpp, and the invariants of
test_reorder.cpp. CSR structure must be bit-exact."""
import hashlib
import json
import os

import numpy as np
import pytest

import rbf_amd as R
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _gold(name):
    return np.load(os.path.join(GOLD, name))


# ---------------------------------------------------------------- generators

def test_jittered_grid_is_seeded_and_in_cells():
    x, y = R.jittered_grid(2, 64, 0.25, seed=1)
    x2, y2 = R.jittered_grid(2, 64, 0.25, seed=1)
    np.testing.assert_array_equal(x, x2)
    g = _gold("keys_4096.npz")
    np.testing.assert_array_equal(x, g["x"])      # the fixtures' own inputs
    i = np.arange(4096) % 64
    j = np.arange(4096) // 64
    assert np.abs(x - (i + 0.5) / 64).max() <= 0.25 / 64
    assert np.abs(y - (j + 0.5) / 64).max() <= 0.25 / 64
    x3, y3, z3 = R.jittered_grid(3, 8, 0.25, seed=4)
    assert x3.size == 512 and np.abs(z3 - (np.arange(512) // 64 + 0.5) / 8).max() <= 0.25 / 8


# ---------------------------------------------------------------- keys and orders

def test_keys_match_reference_fortran():
    g = _gold("keys_4096.npz")
    np.testing.assert_array_equal(R.morton_keys(g["x"], g["y"], 16), g["morton"])
    np.testing.assert_array_equal(R.hilbert_keys(g["x"], g["y"], 16), g["hilbert"])
    np.testing.assert_array_equal(R.hilbert_order(g["x"], g["y"], 16).map(), g["hilbert_order"])


def _interleave(ix, iy, ndiv):
    z = 0
    for b in range(ndiv - 1, -1, -1):
        z = (z << 2) | (((iy >> b) & 1) << 1) | ((ix >> b) & 1)
    return z


def _grid4():
    i = np.tile(np.arange(4), 4)
    j = np.repeat(np.arange(4), 4)
    return i + 0.5, j + 0.5


def test_morton_known_answer():
    # test_reorder.cpp:72-80
    x, y = _grid4()
    keys = R.morton_keys(x, y, 2, (0, 0, 4, 4))
    for s in range(16):
        assert keys[s] == _interleave(s % 4, s // 4, 2)


def test_hilbert_property():
    # test_reorder.cpp:82-102: a permutation of 0..15, consecutive keys adjacent
    x, y = _grid4()
    keys = R.hilbert_keys(x, y, 2, (0, 0, 4, 4))
    assert sorted(keys.tolist()) == list(range(16))
    cell = np.argsort(keys)
    for k in range(1, 16):
        a, b = cell[k - 1], cell[k]
        assert abs(a % 4 - b % 4) + abs(a // 4 - b // 4) == 1


def test_morton_3d_known_answer():
    c = np.arange(8) + 0.5
    i, j, k = np.meshgrid(c, c, c, indexing="ij")
    x, y, z = i.ravel(), j.ravel(), k.ravel()
    keys = R.morton_keys_3d(x, y, z, 3, (0, 0, 0, 8, 8, 8))
    for s in range(x.size):
        ix, iy, iz = int(x[s]), int(y[s]), int(z[s])
        want = 0
        for b in range(2, -1, -1):
            want = (want << 3) | (((iz >> b) & 1) << 2) | (((iy >> b) & 1) << 1) | ((ix >> b) & 1)
        assert keys[s] == want


def test_order_is_stable():
    keys = np.array([3, 1, 3, 0, 1, 3], np.int64)
    np.testing.assert_array_equal(R.order_by_keys(keys).map(), [3, 1, 4, 0, 2, 5])


@pytest.mark.parametrize("bad", [0, 32])
def test_ndiv_out_of_range_is_an_error(bad):
    with pytest.raises(R.RbfError) as e:
        R.hilbert_keys([0.1, 0.2], [0.1, 0.2], bad)
    assert e.value.status == R.ERR_ARG


# ---------------------------------------------------------------- permutations

def test_permutation_algebra():
    # test_reorder.cpp:30-59
    a = np.array([10, 11, 12, 13.0])
    p = R.Permutation([2, 0, 3, 1])
    np.testing.assert_array_equal(p.permute(a), [12, 10, 13, 11])
    np.testing.assert_array_equal(p.unpermute(p.permute(a)), a)
    for i in range(4):
        assert p.inv()[p.map()[i]] == i
    q = R.Permutation([1, 3, 0, 2])
    np.testing.assert_array_equal(q.permute(p.permute(a)), p.then(q).permute(a))
    np.testing.assert_array_equal(p.then(p.inverse()).map(), np.arange(4))
    with pytest.raises(ValueError):
        R.Permutation([0, 0, 1])


def test_renumber_stencils():
    # test_reorder.cpp:119-140
    ja = np.array([0, 1, 1, 2, 2, 3, 3, 0], np.int32)
    p = R.Permutation([2, 0, 3, 1])
    out = R.renumber_stencils(ja, 2, p)
    for s in range(4):
        assert out[2 * s] == s
        assert out[2 * s + 1] == p.inv()[(p.map()[s] + 1) % 4]
    np.testing.assert_array_equal(R.make_row_ptr(4, 2), [0, 2, 4, 6, 8])


def test_renumber_commutes_with_knn():
    # stencils of the renumbered cloud == renumbered stencils of the cloud
    g = _gold("ja_4096_n21.npz")
    p = R.hilbert_order(g["x"], g["y"])
    a = R.knn_stencils(21, p.permute(g["x"]), p.permute(g["y"]))
    b = R.renumber_stencils(g["ja"], 21, p)
    np.testing.assert_array_equal(a, b)


def test_nodeset_chain():
    # test_reorder.cpp:142-194
    x, y = _grid4()
    flag = ((x < 1) | (x > 3) | (y < 1) | (y > 3)).astype(np.int32)
    ns = R.NodeSet(x, y, flag)
    assert ns.num_boundary() == 12 and ns.num_interior() == 4
    fx = ns.x.copy()
    ns.renumber(R.morton_order(ns.x, ns.y, 2, (0, 0, 4, 4)))
    ns.renumber(R.partition_last(ns.flag != 0))
    assert (ns.flag[:4] == 0).all() and (ns.flag[4:] != 0).all()
    assert ns.bnd[0] == 4
    np.testing.assert_array_equal(ns.x, fx[ns.file_order().map()])
    ja = ns.stencils(4)
    assert (ja[::4] == np.arange(16)).all()


# ---------------------------------------------------------------- kNN

def test_knn_matches_reference_nanoflann():
    g = _gold("ja_4096_n21.npz")
    np.testing.assert_array_equal(R.knn_stencils(21, g["x"], g["y"]), g["ja"])


def test_knn_matches_reference_ckdtree_3d():
    g = _gold("ja_cube_1000_k80.npz")
    np.testing.assert_array_equal(R.knn_stencils(80, g["x"], g["y"], g["z"]), g["ja"])


@pytest.mark.parametrize("dim,k", [(2, 1), (2, 7), (2, 60), (3, 13), (3, 99)])
def test_knn_matches_brute_force(dim, k):
    rng = np.random.default_rng(dim * 100 + k)
    pts = rng.random((dim, 700))
    if dim == 2:
        pts[0] *= 4.0          # anisotropic box
    ja = R.knn_stencils(k, *pts)
    np.testing.assert_array_equal(ja, O.knn_brute(k, *pts))
    assert (ja[::k] == np.arange(700)).all()


def test_knn_degenerate_clouds():
    # test_spatial.cpp:246-272: one point, coincident points, collinear points
    np.testing.assert_array_equal(R.knn_stencils(1, [0.5], [0.5]), [0])
    x = np.full(16, 0.25)
    ja = R.knn_stencils(16, x, x).reshape(16, 16)
    assert (ja == np.arange(16)).all()      # all distances 0: by index
    t = np.linspace(0, 1, 50)
    ja = R.knn_stencils(5, t, 2 * t)
    np.testing.assert_array_equal(ja, O.knn_brute(5, t, 2 * t))
    ja = R.knn_stencils(50, t, t)            # k = n
    assert (np.sort(ja.reshape(50, 50), 1) == np.arange(50)).all()


def test_knn_errors():
    with pytest.raises(R.RbfError):
        R.knn_stencils(5, [0.0, 1.0], [0.0, 1.0])        # k > n


def test_benchmark_structure_matches_reference_digests():
    """The 1M-node CSR structure of the benchmark, bit-exact against the
    reference's hilbert_order + NodeSet::stencils(35) (digests written by
    tests/golden/make_golden.py)."""
    with open(os.path.join(GOLD, "digests.json")) as f:
        d = json.load(f)
    x, y = R.jittered_grid(2, 1024, 0.25, seed=2)
    p = R.hilbert_order(x, y, 16)
    assert hashlib.sha256(p.map().tobytes()).hexdigest() == d["hilbert_order_sha256"]
    x, y = p.permute(x), p.permute(y)
    assert hashlib.sha256(np.ascontiguousarray(np.stack([x, y])).tobytes()).hexdigest() == \
        d["xy_sha256"]
    ja = R.knn_stencils(35, x, y)
    assert hashlib.sha256(ja.tobytes()).hexdigest() == d["ja_n35_sha256"]


@pytest.mark.skipif(not O.ref_available(), reason="reference build absent")
def test_renumber_matches_reference_build():
    g = _gold("ja_4096_n21.npz")
    p = R.hilbert_order(g["x"], g["y"])
    ja = g["ja"].copy()
    O.ref().ref_renumber_stencils(4096, 21, O._p(p.map()), O._p(ja))
    np.testing.assert_array_equal(R.renumber_stencils(g["ja"], 21, p), ja)


def test_graded_disk_and_ragged_stencils():
    from rbf_amd import workloads
    x, y, flag = R.graded_disk(20000, 500, seed=3)
    assert x.size == 20500 and (flag[:20000] == 0).all() and (flag[20000:] == 1).all()
    r = np.hypot(x, y)
    assert r[:20000].max() < 1.0 and np.allclose(r[20000:], 1.0)
    x2, y2, _ = R.graded_disk(20000, 500, seed=3)
    np.testing.assert_array_equal(x, x2)
    # spacing grades from hc at the centre to hc/4 at the rim
    ja = R.knn_stencils(2, x, y)[1::2]
    d = np.hypot(x - x[ja], y - y[ja])
    inner, outer = r < 0.2, (r > 0.85) & (flag == 0)
    assert 2.5 < np.median(d[inner]) / np.median(d[outer]) < 6.0
    sizes = workloads.disk_sizes(x, y, flag)
    assert set(np.unique(sizes)) == {13, 21, 35, 56}
    ia, jr = R.ragged_stencils(sizes, x, y)
    full = R.knn_stencils(56, x, y).reshape(-1, 56)
    for s in range(0, x.size, 97):
        np.testing.assert_array_equal(jr[ia[s]:ia[s + 1]], full[s, :sizes[s]])
        assert jr[ia[s]] == s


def test_disk_darts_rejection_sampling():
    """Seeded rejection sampling on the unit disk: spacing h(r) = hc (1 -
    0.75 r) (1:4 centre to rim), one node per h^2, a hard core of
    0.7136 (h(p) + h(q)) / 2 between every pair, equispaced rim nodes."""
    from scipy.spatial import cKDTree
    from rbf_amd import workloads
    n, nb = 40000, 800
    x, y, flag = R.disk_darts(n, nb, seed=9)
    assert x.size == n + nb and (flag[:n] == 0).all() and (flag[n:] == 1).all()
    r = np.hypot(x, y)
    assert r[:n].max() < 1.0
    th = np.arctan2(y[n:], x[n:]) % (2 * np.pi)
    np.testing.assert_allclose(r[n:], 1.0, rtol=0, atol=1e-15)
    np.testing.assert_allclose(np.sort(th), 2 * np.pi * np.arange(nb) / nb, atol=1e-12)
    x2, y2, _ = R.disk_darts(n, nb, seed=9)
    np.testing.assert_array_equal(x, x2)
    np.testing.assert_array_equal(y, y2)
    assert not np.array_equal(R.disk_darts(n, nb, seed=10)[0], x)

    a = 0.75

    def mass(rr):  # integral_0^r s / (1 - a s)^2 ds
        u = 1 - a * rr
        return (1 / u - 1 + np.log(u)) / a ** 2

    hc = np.sqrt(2 * np.pi * mass(1.0) / n)
    h = hc * (1 - a * np.minimum(r, 1.0))
    # the hard core, over every pair that could violate it
    pairs = np.array(sorted(cKDTree(np.c_[x, y]).query_pairs(0.7136 * hc)))
    d = np.hypot(x[pairs[:, 0]] - x[pairs[:, 1]], y[pairs[:, 0]] - y[pairs[:, 1]])
    assert (d >= 0.7136 * 0.5 * (h[pairs[:, 0]] + h[pairs[:, 1]]) * (1 - 1e-12)).all()
    # the radial distribution is the target density's: fraction inside rho
    # is G(rho) / G(1)
    ri = r[:n]
    for rho in (0.25, 0.5, 0.75, 0.9):
        assert abs((ri < rho).mean() - mass(rho) / mass(1.0)) < 0.01, rho
    # nearest-neighbour spacing grades with h: centre / rim ~ 0.9625 / 0.2875
    ja = R.knn_stencils(2, x, y)[1::2]
    nn = np.hypot(x - x[ja], y - y[ja])
    ratio = np.median(nn[r < 0.1]) / np.median(nn[(r > 0.9) & (flag == 0)])
    assert 0.9 * 3.35 < ratio < 1.1 * 3.35
    sizes = workloads.disk_sizes(x, y, flag)
    assert set(np.unique(sizes)) == {13, 21, 35, 56}


def test_disk_darts_rejects_bad_arguments():
    with pytest.raises(R.RbfError):
        R.disk_darts(0, 10)
    # a boundary far denser than the rim spacing jams the darts: an error,
    # not a short node set
    with pytest.raises(R.RbfError):
        R.disk_darts(100, 20000)


def test_oracle_ragged_rows_equal_fixed_width():
    x, y = R.jittered_grid(2, 24, 0.25, seed=4)
    ja = R.knn_stencils(21, x, y)
    ia = R.make_row_ptr(x.size, 21)
    w, info = O.assemble(2, 21, 2, 3, x, y, ja)
    w2 = [np.zeros(ja.size) for _ in range(3)]
    info2 = np.zeros(x.size, np.int32)
    O.assemble_rows(2, 21, 2, 3, np.arange(x.size, dtype=np.int32), ia, x, y, ja, w2, info2)
    for a, b in zip(w, w2):
        np.testing.assert_array_equal(a, b)
