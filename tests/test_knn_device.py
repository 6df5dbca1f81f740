"""The node layer on the GPU against the host node layer and the reference's
fixtures: curve keys, stable order, gathers and kNN stencils must be
identical, bit for bit (integer results)."""
import hashlib
import json
import os

import numpy as np
import pytest

import rbf_amd as R
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def _knn_dev(k, *pts):
    d = [_dev(p) for p in pts]
    return R.knn_stencils_device(k, *d).cpu().numpy()


def test_device_knn_matches_reference_fixtures():
    g = np.load(os.path.join(GOLD, "ja_4096_n21.npz"))
    np.testing.assert_array_equal(_knn_dev(21, g["x"], g["y"]), g["ja"])   # nanoflann
    g3 = np.load(os.path.join(GOLD, "ja_cube_1000_k80.npz"))
    np.testing.assert_array_equal(_knn_dev(80, g3["x"], g3["y"], g3["z"]), g3["ja"])  # ckdtree


def test_device_pipeline_reproduces_the_benchmark_structure():
    """Keys, stable order, gather and kNN on the device for the 1M-node
    benchmark cloud: the same SHA-256 digests as the reference's
    hilbert_order + NodeSet::stencils(35)."""
    import torch
    with open(os.path.join(GOLD, "digests.json")) as f:
        dig = json.load(f)
    x, y = R.jittered_grid(2, 1024, 0.25, seed=2)
    xd, yd = _dev(x), _dev(y)
    keys = R.curve_keys_device(xd, yd, curve="hilbert", ndiv=16)
    np.testing.assert_array_equal(keys.cpu().numpy(), R.hilbert_keys(x, y, 16))
    perm = R.order_by_keys_device(keys, key_bits=32)
    assert hashlib.sha256(perm.cpu().numpy().tobytes()).hexdigest() == dig["hilbert_order_sha256"]
    xs, ys = R.gather_device(perm, xd), R.gather_device(perm, yd)
    xy = torch.stack([xs, ys]).cpu().numpy()
    assert hashlib.sha256(np.ascontiguousarray(xy).tobytes()).hexdigest() == dig["xy_sha256"]
    ja = R.knn_stencils_device(35, xs, ys).cpu().numpy()
    assert hashlib.sha256(ja.tobytes()).hexdigest() == dig["ja_n35_sha256"]


@pytest.mark.parametrize("curve", ["morton", "hilbert"])
def test_device_keys_and_order_match_host(curve):
    rng = np.random.default_rng(4)
    x, y = rng.random(5000) * 3 - 1, rng.random(5000)
    x[:50] = x[50:100]                         # repeated points: equal keys, stable order
    y[:50] = y[50:100]
    fn = R.hilbert_keys if curve == "hilbert" else R.morton_keys
    for nd in (4, 16, 31):
        kd = R.curve_keys_device(_dev(x), _dev(y), curve=curve, ndiv=nd).cpu().numpy()
        kh = fn(x, y, nd)
        np.testing.assert_array_equal(kd, kh)
        perm = R.order_by_keys_device(_dev(kh), key_bits=2 * nd).cpu().numpy()
        np.testing.assert_array_equal(perm, R.order_by_keys(kh).map())
    bb = (-2.0, -1.0, 3.0, 2.0)
    np.testing.assert_array_equal(
        R.curve_keys_device(_dev(x), _dev(y), curve=curve, ndiv=10, bbox=bb).cpu().numpy(),
        fn(x, y, 10, bb))
    x3, y3, z3 = R.jittered_grid(3, 12, 0.25, seed=5)
    k3 = R.curve_keys_device(_dev(x3), _dev(y3), _dev(z3), curve="morton", ndiv=16).cpu().numpy()
    np.testing.assert_array_equal(k3, R.morton_keys_3d(x3, y3, z3, 16))


def test_device_knn_on_graded_disk_matches_host():
    """Density varying 16x across the disk: rings of different widths."""
    x, y, _ = R.graded_disk(60000, 1500, seed=8)
    for k in (13, 56, 128):
        np.testing.assert_array_equal(_knn_dev(k, x, y), R.knn_stencils(k, x, y))


def test_device_knn_3d_and_anisotropic_match_host():
    x, y, z = R.jittered_grid(3, 24, 0.25, seed=9)
    np.testing.assert_array_equal(_knn_dev(80, x, y, z), R.knn_stencils(80, x, y, z))
    rng = np.random.default_rng(10)
    a = rng.random((2, 3000))
    a[0] *= 40.0                               # thin box
    np.testing.assert_array_equal(_knn_dev(35, *a), R.knn_stencils(35, *a))


def test_device_knn_degenerate_clouds():
    # test_spatial.cpp:246-272: one point, coincident points, collinear points, k = n
    np.testing.assert_array_equal(_knn_dev(1, [0.5], [0.5]), [0])
    x = np.full(16, 0.25)
    np.testing.assert_array_equal(_knn_dev(16, x, x), R.knn_stencils(16, x, x))
    t = np.linspace(0, 1, 50)
    np.testing.assert_array_equal(_knn_dev(5, t, 2 * t), O.knn_brute(5, t, 2 * t))
    np.testing.assert_array_equal(_knn_dev(50, t, t), R.knn_stencils(50, t, t))
    g = (np.arange(10) + 0.5) / 10                # exact ties on a Cartesian grid
    gx, gy = np.meshgrid(g, g)
    np.testing.assert_array_equal(_knn_dev(9, gx.ravel(), gy.ravel()),
                                  R.knn_stencils(9, gx.ravel(), gy.ravel()))
