"""ctypes wrapper of the CPU oracle (liboracle.so, from rbf_oracle.c) and of
the reference build (oracle/_ref/libref.so, when present).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the CPU-baseline leg of bench.py, always as the checker, never as the
product path.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref.so")

_o = None
_r = None


def build(ref=True):
    """make -C oracle (oracle always; _ref when the reference is present)."""
    targets = ["oracle"]
    if ref and os.path.isdir("/root/reference"):
        targets.append("ref")
    subprocess.run(["make", "-s", "-B", "-C", HERE, *targets], check=True)


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def lib():
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_SO):
            build(ref=False)
        L = ctypes.CDLL(ORACLE_SO)
        P, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
        L.oracle_npoly.restype = ctypes.c_int
        L.oracle_npoly.argtypes = [ctypes.c_int, ctypes.c_int]
        L.oracle_fill.argtypes = [ctypes.c_int] * 4 + [P, P, P, P]
        L.oracle_rhs.argtypes = [ctypes.c_int] * 6 + [P] * 7
        L.oracle_gesv.restype = ctypes.c_int
        L.oracle_gesv.argtypes = [ctypes.c_int, ctypes.c_int, P, P, P]
        L.oracle_getrf.restype = ctypes.c_int
        L.oracle_getrf.argtypes = [ctypes.c_int, P, P, ctypes.POINTER(ctypes.c_int)]
        L.oracle_assemble.restype = ctypes.c_int
        L.oracle_assemble.argtypes = [ctypes.c_int] * 6 + [i32, i32] + [P] * 7 + [P, P, ctypes.c_int]
        L.oracle_assemble_rows.restype = ctypes.c_int
        L.oracle_assemble_rows.argtypes = [ctypes.c_int] * 4 + [i32] + [P] * 8 + [ctypes.c_int]
        L.oracle_zidx.restype = i64
        L.oracle_zidx.argtypes = [ctypes.c_double, ctypes.c_double, P, ctypes.c_int]
        L.oracle_hidx.restype = i64
        L.oracle_hidx.argtypes = [ctypes.c_double, ctypes.c_double, P, ctypes.c_int]
        L.oracle_knn_brute.argtypes = [i32, ctypes.c_int, P, P, P, ctypes.c_int, P]
        L.oracle_max_threads.restype = ctypes.c_int
        _o = L
    return _o


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _r
    if _r is None:
        L = ctypes.CDLL(REF_SO)
        P, i32 = ctypes.c_void_p, ctypes.c_int32
        L.ref_stencils_nanoflann.argtypes = [i32, P, P, i32, P]
        L.ref_stencils_ckdtree.argtypes = [i32, i32, P, i32, P]
        L.ref_hilbert_order.argtypes = [i32, P, P, i32, P]
        L.ref_morton_order.argtypes = [i32, P, P, i32, P]
        L.ref_renumber_stencils.argtypes = [i32, i32, P, P]
        L.rbf_hilbert_keys.argtypes = [P, P, P, P, P, P]
        L.rbf_morton_keys.argtypes = [P, P, P, P, P, P]
        L.ref_read_graph.argtypes = [ctypes.c_char_p, i32, P, P, P, P]
        L.ref_write_mm.argtypes = [ctypes.c_char_p, ctypes.c_int64, ctypes.c_int64, P, P, P]
        _r = L
    return _r


def npoly(dim, P):
    return lib().oracle_npoly(dim, P)


def fill(dim, n, P, Q, xs, ys, zs=None):
    """Collocation matrix (nt x nt) of one stencil in its local frame."""
    nt = n + npoly(dim, P)
    M = np.zeros(nt * nt)
    xs, ys = np.ascontiguousarray(xs, float), np.ascontiguousarray(ys, float)
    zs = np.zeros(n) if zs is None else np.ascontiguousarray(zs, float)
    lib().oracle_fill(dim, n, P, Q, _p(xs), _p(ys), _p(zs), _p(M))
    return M.reshape(nt, nt, order="F")


def rhs(dim, n, P, Q, xs, ys, zs=None, op_set=0, centres=None):
    nt = n + npoly(dim, P)
    if op_set == 0:
        r = dim + 1
        cx = cy = cz = None
    else:
        cx, cy = (np.ascontiguousarray(c, float) for c in centres[:2])
        cz = np.ascontiguousarray(centres[2], float) if dim == 3 else np.zeros(cx.size)
        r = cx.size
    B = np.zeros(nt * r)
    xs, ys = np.ascontiguousarray(xs, float), np.ascontiguousarray(ys, float)
    zs = np.zeros(n) if zs is None else np.ascontiguousarray(zs, float)
    lib().oracle_rhs(dim, n, P, Q, op_set, r, _p(xs), _p(ys), _p(zs), _p(cx), _p(cy), _p(cz),
                     _p(B))
    return B.reshape(nt, r, order="F")


def gesv(M, B):
    M = np.asfortranarray(M, float)
    B = np.asfortranarray(B, float)
    nt, r = B.shape
    X = np.zeros((nt, r), order="F")
    info = lib().oracle_gesv(nt, r, _p(M.ravel(order="F")), _p(B.ravel(order="F")), _p(X))
    return X, info


def getrf(M):
    """LU factors (dgetrf layout, column-major nt x nt), 1-based pivots and
    info of the oracle's elimination."""
    M = np.asfortranarray(M, float)
    nt = M.shape[0]
    LU = np.zeros((nt, nt), order="F")
    ipiv = np.zeros(nt, np.int32)
    info = lib().oracle_getrf(nt, _p(M.ravel(order="F")), _p(LU),
                              ipiv.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return LU, ipiv, info


def assemble(dim, n, P, Q, x, y, ja, z=None, op_set=0, centres=None, s0=0, count=None,
             nthreads=0):
    """FP64 weights of stencils s0 .. s0+count-1: a list of nops arrays of
    length count*n, and their info."""
    x, y = np.ascontiguousarray(x, float), np.ascontiguousarray(y, float)
    z = None if z is None else np.ascontiguousarray(z, float)
    ja = np.ascontiguousarray(ja, np.int32)
    nst = ja.size // n
    count = nst - s0 if count is None else count
    if op_set == 0:
        r, nrhs = dim + 1, 0
        cx = cy = cz = None
    else:
        cx, cy = (np.ascontiguousarray(c, float) for c in centres[:2])
        cz = np.ascontiguousarray(centres[2], float) if dim == 3 else None
        r = nrhs = cx.size
    w = [np.zeros(count * n) for _ in range(r)]
    tab = (ctypes.c_void_p * r)(*[a.ctypes.data for a in w])
    info = np.zeros(count, np.int32)
    lib().oracle_assemble(dim, n, P, Q, op_set, nrhs, s0, count, _p(x), _p(y), _p(z), _p(ja),
                          _p(cx), _p(cy), _p(cz), tab, _p(info), int(nthreads))
    return w, info


def assemble_rows(dim, n, P, Q, rows, ia, x, y, ja, w, info, z=None, nthreads=0):
    """FP64 weights of the listed stencils (one size class) of a ragged graph,
    written into the caller's arrays w[op][ia[s] + k] and info[s]."""
    rows = np.ascontiguousarray(rows, np.int32)
    tab = (ctypes.c_void_p * len(w))(*[a.ctypes.data for a in w])
    return lib().oracle_assemble_rows(dim, n, P, Q, rows.size, _p(rows), _p(ia), _p(x), _p(y),
                                      _p(z), _p(ja), tab, _p(info), int(nthreads))


def zidx(px, py, bbox, ndiv):
    bb = np.ascontiguousarray(bbox, float)
    return lib().oracle_zidx(float(px), float(py), _p(bb), int(ndiv))


def hidx(px, py, bbox, ndiv):
    bb = np.ascontiguousarray(bbox, float)
    return lib().oracle_hidx(float(px), float(py), _p(bb), int(ndiv))


def knn_brute(k, x, y, z=None):
    x, y = np.ascontiguousarray(x, float), np.ascontiguousarray(y, float)
    zz = None if z is None else np.ascontiguousarray(z, float)
    ja = np.empty(x.size * k, np.int32)
    lib().oracle_knn_brute(x.size, 2 if z is None else 3, _p(x), _p(y), _p(zz), k, _p(ja))
    return ja


def max_threads():
    return lib().oracle_max_threads()


def apply(ia, ja, w, u, bnd=None):
    """The reference's mv_op (src/main.cpp:59-66): CSR product y = W u with
    each row summed left to right, then y[b] = u[b] on the boundary rows b.
    Also returns sum_j |w_j u_j| per row, the scale of the rounding bound
    for any reassociated sum. Numpy; w is one operator's value array."""
    ia = np.asarray(ia, np.int64)
    rows = np.repeat(np.arange(ia.size - 1), np.diff(ia))
    pos = np.arange(ia[-1]) - ia[rows]
    width = int(np.diff(ia).max()) if ia.size > 1 else 0
    prod = np.asarray(w[:ia[-1]], float) * np.asarray(u, float)[np.asarray(ja[:ia[-1]])]
    y = np.zeros(ia.size - 1)
    for j in range(width):                 # column j of every row, in order
        sel = pos == j
        y[rows[sel]] += prod[sel]
    scale = np.zeros(ia.size - 1)
    np.add.at(scale, rows, np.abs(prod))
    if bnd is not None:
        b = np.asarray(bnd, np.int64)
        y[b] = np.asarray(u, float)[b]
        scale[b] = np.abs(np.asarray(u, float)[b])
    return y, scale
