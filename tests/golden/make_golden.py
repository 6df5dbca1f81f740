"""Write the golden vectors under tests/golden/ from the reference's own host
code, compiled from its sources by oracle/Makefile (oracle/_ref/libref.so):

  keys_4096.npz      Morton and Hilbert keys (rbf_ordering.F90 zidx / hidx,
                     ndiv 16, bbox of the points) of the 4,096-node jittered
                     unit-square cloud (64x64, +-0.25h, seed 1), and the
                     Hilbert order of rbf_reorder.h:222-229
  ja_4096_n21.npz    NodeSet::stencils(21) (nanoflann) of that cloud
  ja_cube_1000_k80.npz  KdTree::knn_stencils(80) (SciPy ckdtree) of a
                     10^3 jittered unit-cube cloud (seed 3)
  digests.json       SHA-256 of the Hilbert order and NodeSet::stencils(35)
                     of the 1,048,576-node cloud of the benchmark (1024^2,
                     seed 2, renumbered by hilbert_order first)
  mm_small.npz, mm_small.mtx, mm_small_pattern.mtx
                     a small ragged CSR operator (values chosen to exercise
                     the shortest round-trip number text: 0.1, 1/3, the
                     smallest subnormal, 1e300, ...) and the reference's
                     write_matrix_market output for it (rbf_io.h:782-823)

cartesian_32_21.{points,graph} and poisson_32_21.{points,graph} are copies
of the reference's data files (data/), kept as input fixtures.

The node clouds themselves come from rbfa_gen_jittered_grid; the fixture
files store the inputs they were made from, so the tests check the inputs
too.

    python tests/golden/make_golden.py
"""
import ctypes
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "gpu-assembly_amd"))
sys.path.insert(0, ROOT)

import rbf_amd as R  # noqa: E402
from oracle import oracle as O  # noqa: E402


def fortran_keys(fn, x, y, ndiv):
    n = ctypes.c_int32(x.size)
    nd = ctypes.c_int32(ndiv)
    bb = np.array([x.min(), y.min(), x.max(), y.max()])
    keys = np.empty(x.size, np.int64)
    fn(ctypes.byref(n), O._p(x), O._p(y), O._p(bb), ctypes.byref(nd), O._p(keys))
    return keys


def mm_small_inputs():
    """12 ragged rows over 40 columns, seeded; special values first."""
    rng = np.random.default_rng(11)
    lens = rng.integers(1, 9, 12)
    ia = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ja = np.concatenate([np.sort(rng.choice(40, l, replace=False)) for l in lens]).astype(np.int32)
    a = rng.standard_normal(ia[-1]) * 10.0 ** rng.integers(-12, 12, ia[-1])
    special = [0.1, 1.0 / 3.0, 5e-324, 1e300, -2.5e-300, 123456789.0, 2.0, 1.0000000000000002,
               np.pi, -0.0, 1e-5, 6.02214076e23]
    a[:len(special)] = special
    return ia, ja, a


def main():
    if not O.ref_available():
        O.build(ref=True)
    ref = O.ref()
    x, y = R.jittered_grid(2, 64, 0.25, seed=1)
    mk = fortran_keys(ref.rbf_morton_keys, x, y, 16)
    hk = fortran_keys(ref.rbf_hilbert_keys, x, y, 16)
    horder = np.empty(x.size, np.int32)
    ref.ref_hilbert_order(x.size, O._p(x), O._p(y), 16, O._p(horder))
    np.savez_compressed(os.path.join(HERE, "keys_4096.npz"), x=x, y=y, morton=mk, hilbert=hk,
                        hilbert_order=horder)
    ja = np.empty(x.size * 21, np.int32)
    ref.ref_stencils_nanoflann(x.size, O._p(x), O._p(y), 21, O._p(ja))
    np.savez_compressed(os.path.join(HERE, "ja_4096_n21.npz"), x=x, y=y, ja=ja)

    xc, yc, zc = R.jittered_grid(3, 10, 0.25, seed=3)
    pts = np.ascontiguousarray(np.stack([xc, yc, zc], 1).ravel())
    ja3 = np.empty(xc.size * 80, np.int32)
    ref.ref_stencils_ckdtree(xc.size, 3, O._p(pts), 80, O._p(ja3))
    np.savez_compressed(os.path.join(HERE, "ja_cube_1000_k80.npz"), x=xc, y=yc, z=zc, ja=ja3)

    X, Y = R.jittered_grid(2, 1024, 0.25, seed=2)
    order = np.empty(X.size, np.int32)
    ref.ref_hilbert_order(X.size, O._p(X), O._p(Y), 16, O._p(order))
    X, Y = np.ascontiguousarray(X[order]), np.ascontiguousarray(Y[order])
    ja35 = np.empty(X.size * 35, np.int32)
    ref.ref_stencils_nanoflann(X.size, O._p(X), O._p(Y), 35, O._p(ja35))
    dig = {
        "cloud": "jittered 1024x1024 unit square, jitter 0.25h, seed 2",
        "xy_sha256": hashlib.sha256(np.ascontiguousarray(np.stack([X, Y])).tobytes()).hexdigest(),
        "hilbert_order_sha256": hashlib.sha256(order.tobytes()).hexdigest(),
        "ja_n35_sha256": hashlib.sha256(ja35.tobytes()).hexdigest(),
        "source": "oracle/_ref/libref.so: rbf::hilbert_order + rbf::NodeSet::stencils(35)",
    }
    with open(os.path.join(HERE, "digests.json"), "w") as f:
        json.dump(dig, f, indent=1)

    ia, ja, a = mm_small_inputs()
    np.savez_compressed(os.path.join(HERE, "mm_small.npz"), ia=ia, ja=ja, a=a)
    for name, vals in (("mm_small.mtx", a), ("mm_small_pattern.mtx", None)):
        rc = ref.ref_write_mm(os.fsencode(os.path.join(HERE, name)), ia.size - 1, 40, O._p(ia),
                              O._p(ja), None if vals is None else O._p(vals))
        assert rc == 0
    print("golden vectors written")


if __name__ == "__main__":
    main()
