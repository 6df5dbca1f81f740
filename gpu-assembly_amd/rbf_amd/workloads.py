"""Seeded synthetic node sets and stencil graphs of the benchmarked shapes.

Each builder returns a dict with the host arrays (x, y[, z]) in the final
numbering, the stencil graph ja (int32, n per node, self first), and the
basis (dim, n, P, Q). All of it comes from the host node layer of
librbf_amd (jittered grid, curve renumbering, kNN).
"""
import time

import numpy as np

from . import (disk_darts, graded_disk, hilbert_order, jittered_grid, knn_stencils, morton_order_3d,
               ragged_stencils)

SHAPES = {
    # name: (dim, points per axis, n, P, Q, renumbering)
    "square_4k_n21": (2, 64, 21, 2, 3, None),
    "square_1m_n35": (2, 1024, 35, 3, 5, "hilbert"),
    "square_1m_n56": (2, 1024, 56, 4, 7, "hilbert"),
    "cube_2m_n80": (3, 128, 80, 3, 3, "morton"),
    "interp_1m_d2q9": (2, 1024, 9, 2, 3, "hilbert"),
}

# Interpolation workloads: number of centres. D2Q9 semi-Lagrangian streaming
# (the reference's CUDA-Fortran demo, main.f90:16,55: n = 9, P = 2, NRHS = 9,
# centres half-way along each of the nine lattice directions).
INTERP_NRHS = {"interp_1m_d2q9": 9}


def d2q9_centres(h):
    """The nine interpolation centres in the stencil-local frame: 0.5 h c_i
    for the D2Q9 lattice velocities c_i (rest, 4 axis, 4 diagonal)."""
    cx = 0.5 * h * np.array([0, 1, 0, -1, 0, 1, -1, -1, 1], float)
    cy = 0.5 * h * np.array([0, 0, 1, 0, -1, 1, 1, -1, -1], float)
    return cx, cy


def build(name=None, dim=2, m=64, n=21, P=2, Q=3, order=None, seed=1, jitter=0.25):
    if name is not None:
        dim, m, n, P, Q, order = SHAPES[name]
    t0 = time.perf_counter()
    pts = jittered_grid(dim, m, jitter, seed)
    t1 = time.perf_counter()
    if order == "hilbert":
        p = hilbert_order(pts[0], pts[1], 16)
        pts = tuple(p.permute(a) for a in pts)
    elif order == "morton":
        p = morton_order_3d(*pts, ndiv=16)
        pts = tuple(p.permute(a) for a in pts)
    t2 = time.perf_counter()
    ja = knn_stencils(n, *pts)
    t3 = time.perf_counter()
    out = {"dim": dim, "n": n, "P": P, "Q": Q, "x": pts[0], "y": pts[1],
           "z": pts[2] if dim == 3 else None, "ja": ja, "nstencils": pts[0].size,
           "host_s": {"nodes": t1 - t0, "renumber": t2 - t1, "knn": t3 - t2}}
    return out


def flops_per_stencil(nt, r):
    """LAPACK operation counts: getrf(nt) + getrs(nt, r)."""
    getrf = 2 * nt ** 3 / 3 - nt ** 2 / 2 - nt / 6
    getrs = r * (2 * nt ** 2 - nt)
    return getrf + getrs


def bytes_per_stencil(n, r, dim=2):
    """Algorithmic HBM bytes: ja (4n), the node's own coordinates (8 dim),
    the weights (8 r n) and the solver status (4)."""
    return 4 * n + 8 * dim + 8 * r * n + 4


# Uneven batch on the unit disk: size class by radius for interior nodes,
# n = 13 (degree 2, r^3) on the boundary.
DISK_CLASSES = [(13, 2, 3), (21, 2, 3), (35, 3, 5), (56, 4, 7)]


def disk_sizes(x, y, flag):
    r = np.hypot(x, y)
    n = np.where(r < 1.0 / 3.0, 21, np.where(r < 2.0 / 3.0, 35, 56)).astype(np.int32)
    n[flag != 0] = 13
    return n


def build_disk(n_interior=2_000_000, n_boundary=8000, seed=5, method="darts"):
    """Graded unit disk, Hilbert-renumbered, ragged kNN stencils by class.
    method "darts": seeded rejection sampling with the 1:4 radial spacing
    (rbfa_gen_disk_darts); "rings": jittered concentric rings of the same
    grading (rbfa_gen_graded_disk)."""
    t0 = time.perf_counter()
    if method == "darts":
        x, y, flag = disk_darts(n_interior, n_boundary, seed)
    elif method == "rings":
        x, y, flag = graded_disk(n_interior, n_boundary, seed)
    else:
        raise ValueError(f"unknown disk method {method!r}")
    t1 = time.perf_counter()
    p = hilbert_order(x, y, 16)
    x, y, flag = p.permute(x), p.permute(y), p.permute(flag)
    t2 = time.perf_counter()
    sizes = disk_sizes(x, y, flag)
    ia, ja = ragged_stencils(sizes, x, y)
    t3 = time.perf_counter()
    return {"x": x, "y": y, "flag": flag, "sizes": sizes, "ia": ia, "ja": ja,
            "nstencils": x.size,
            "host_s": {"nodes": t1 - t0, "renumber": t2 - t1, "knn": t3 - t2}}
