"""Python host layer over librbf_amd.so (the C ABI of include/rbf_assembly.h).

Mirrors the reference's host interface (ivan-pi/gpu-assembly) so code and
tests read the same:

    Permutation            rbf_reorder.h:69-157   (gather form, map/inv/then)
    morton_keys, hilbert_keys, morton_order, hilbert_order
                           rbf_reorder.h:193-229
    renumber_stencils      rbf_reorder.h:278-288
    make_row_ptr           rbf_reorder.h:316-322
    NodeSet                rbf_nodeset.h:46-235   (renumber chain, file_order,
                                                   stencils(k))
    Plan / assemble        interp_config + assemble_interp_weights /
                           assemble_laplace_weights (rbf_operators.h:187-376,
                           main.cpp:15-16,44-54)

Host arrays are numpy; device arrays are torch tensors on the GPU (torch is
only the allocator and stream provider here). Every compute call goes
through the HIP library; there is no CPU fallback: if the library or a GPU
is missing, the call raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "librbf_amd.so")

OK, ERR_ARG, ERR_UNSUPPORTED, ERR_HIP, ERR_NOMEM, ERR_NODEVICE, ERR_FORMAT, ERR_IO = range(8)
OPS_DIFF, OPS_INTERP = 0, 1
PREC_FP64, PREC_MIXED = 0, 1
STAT_STEPS_MASK, STAT_FALLBACK = 0x03, 0x04

_lib = None


class RbfError(RuntimeError):
    def __init__(self, status, what):
        self.status = status
        super().__init__(f"{what}: {status_string(status)} (status {status})")


class _BBox2(ctypes.Structure):
    _fields_ = [("xmin", ctypes.c_double), ("ymin", ctypes.c_double),
                ("xmax", ctypes.c_double), ("ymax", ctypes.c_double)]


class _PlanDesc(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in
                ("dim", "n", "degree", "phs", "op_set", "nrhs", "precision", "max_refine")]


def lib():
    """Load librbf_amd.so (built in-tree by gpu-assembly_amd/build.py)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with gpu-assembly_amd/build.py")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    i32, i64, u64, dbl = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_double
    sig = {
        "rbfa_version": (ctypes.c_char_p, []),
        "rbfa_measure_peaks": (i32, [P, P]),
        "rbfa_status_string": (ctypes.c_char_p, [i32]),
        "rbfa_morton_keys": (i32, [i32, P, P, P, i32, P]),
        "rbfa_hilbert_keys": (i32, [i32, P, P, P, i32, P]),
        "rbfa_morton_keys_3d": (i32, [i32, P, P, P, P, i32, P]),
        "rbfa_order_by_keys": (i32, [i32, P, P]),
        "rbfa_permute_f64": (i32, [i32, P, P]),
        "rbfa_permute_i32": (i32, [i32, P, P]),
        "rbfa_renumber_stencils": (i32, [i32, i32, P, P]),
        "rbfa_make_row_ptr": (i32, [i32, i32, P]),
        "rbfa_knn_stencils": (i32, [i32, i32, P, P, P, i32, P]),
        "rbfa_gen_jittered_grid": (i32, [i32, i32, dbl, u64, P, P, P]),
        "rbfa_plan_create": (i32, [P, P]),
        "rbfa_plan_destroy": (i32, [P]),
        "rbfa_plan_nops": (i32, [P]),
        "rbfa_plan_system_size": (i32, [P]),
        "rbfa_assemble": (i32, [P, i32, P, P, P, P, P, P, P, P, P, P, P]),
        "rbfa_assemble_noalloc": (i32, [P, i32, P, P, P, P, P, P, P, P, P, P, P]),
        "rbfa_plan_num_variants": (i32, [P]),
        "rbfa_plan_variant_name": (ctypes.c_char_p, [P, i32]),
        "rbfa_plan_get_variant": (i32, [P]),
        "rbfa_plan_set_variant": (i32, [P, i32]),
        "rbfa_plan_autotune": (i32, [P, i32, P, P, P, P, P, P, P, P, P, P, P, i32, P]),
        "rbfa_gen_graded_disk": (i32, [i32, i32, u64, P, P, P]),
        "rbfa_gen_disk_darts": (i32, [i32, i32, u64, P, P, P]),
        "rbfa_batch_create": (i32, [i32, P, P]),
        "rbfa_batch_destroy": (i32, [P]),
        "rbfa_batch_bind": (i32, [P, i32, P, P, P, P, P, P, P, P, P, P]),
        "rbfa_batch_class_count": (i32, [P, i32]),
        "rbfa_batch_class_variant": (ctypes.c_char_p, [P, i32]),
        "rbfa_batch_class_num_variants": (i32, [P, i32]),
        "rbfa_batch_class_variant_name": (ctypes.c_char_p, [P, i32, i32]),
        "rbfa_batch_set_class_variant": (i32, [P, i32, i32]),
        "rbfa_batch_set_mode": (i32, [P, i32]),
        "rbfa_batch_get_mode": (i32, [P]),
        "rbfa_batch_launch": (i32, [P, P]),
        "rbfa_batch_autotune": (i32, [P, P, i32, P]),
        "rbfa_apply": (i32, [i32, i32, P, P, i32, P, P, P, i32, P, P]),
        "rbfa_apply_plan_create": (i32, [i32, i32, i32, P, P, i32, P, P]),
        "rbfa_apply_plan_destroy": (i32, [P]),
        "rbfa_apply_plan_stats": (i32, [P, P]),
        "rbfa_apply_plan_run": (i32, [P, i32, P, P, P, P]),
        "rbfa_last_error": (ctypes.c_char_p, []),
        "rbfa_knn_stencils_device": (i32, [i32, i32, P, P, P, i32, P, P]),
        "rbfa_curve_keys_device": (i32, [i32, i32, i32, P, P, P, P, i32, P, P]),
        "rbfa_order_by_keys_device": (i32, [i32, P, i32, P, P]),
        "rbfa_gather_f64_device": (i32, [i32, P, P, P, P]),
        "rbfa_read_points": (i32, [ctypes.c_char_p, i32, P, P, P]),
        "rbfa_read_graph": (i32, [ctypes.c_char_p, i32, P, P, P, P, i64]),
        "rbfa_write_matrix_market": (i32, [ctypes.c_char_p, i64, i64, P, P, P, i32]),
        "rbfa_gesv_batched": (i32, [i32, i32, i32, P, P, P, P, P]),
        "rbfa_solve_dp": (i32, [P, P, P, P, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def status_string(status):
    return lib().rbfa_status_string(int(status)).decode()


def _check(status, what):
    if status != OK:
        raise RbfError(status, what)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ---------------------------------------------------------------------------
# Renumbering (rbf_reorder.h)
# ---------------------------------------------------------------------------

class Permutation:
    """Gather-form permutation: map()[i] is the OLD index placed at NEW i."""

    def __init__(self, new_to_old):
        self._p = _i32(new_to_old)
        n = self._p.size
        if n and (self._p.min() < 0 or self._p.max() >= n or np.unique(self._p).size != n):
            raise ValueError("not a permutation")
        self._ip = np.empty_like(self._p)
        self._ip[self._p] = np.arange(n, dtype=np.int32)

    @classmethod
    def identity(cls, n):
        return cls(np.arange(n, dtype=np.int32))

    @classmethod
    def from_inverse(cls, old_to_new):
        ip = _i32(old_to_new)
        p = np.empty_like(ip)
        p[ip] = np.arange(ip.size, dtype=np.int32)
        return cls(p)

    def size(self):
        return self._p.size

    def map(self):
        return self._p

    def inv(self):
        return self._ip

    def permute(self, a):
        """a_new[i] = a_old[map()[i]] (returns a new array)."""
        return np.asarray(a)[self._p]

    def unpermute(self, a):
        out = np.empty_like(np.asarray(a))
        out[self._p] = a
        return out

    def then(self, q):
        return Permutation(self._p[q.map()])

    def inverse(self):
        return Permutation(self._ip)


def _keys(fn, x, y, ndiv, bbox):
    x, y = _f64(x), _f64(y)
    keys = np.empty(x.size, dtype=np.int64)
    bb = None if bbox is None else ctypes.byref(_BBox2(*[float(v) for v in bbox]))
    _check(fn(x.size, _ptr(x), _ptr(y), bb, int(ndiv), _ptr(keys)), fn.__name__)
    return keys


def morton_keys(x, y, ndiv=16, bbox=None):
    """bbox = (xmin, ymin, xmax, ymax) or None for the points' own box."""
    return _keys(lib().rbfa_morton_keys, x, y, ndiv, bbox)


def hilbert_keys(x, y, ndiv=16, bbox=None):
    return _keys(lib().rbfa_hilbert_keys, x, y, ndiv, bbox)


def morton_keys_3d(x, y, z, ndiv=16, bbox=None):
    x, y, z = _f64(x), _f64(y), _f64(z)
    keys = np.empty(x.size, dtype=np.int64)
    bb = None if bbox is None else _f64(bbox)
    _check(lib().rbfa_morton_keys_3d(x.size, _ptr(x), _ptr(y), _ptr(z), _ptr(bb), int(ndiv),
                                     _ptr(keys)), "rbfa_morton_keys_3d")
    return keys


def order_by_keys(keys):
    keys = np.ascontiguousarray(keys, dtype=np.int64)
    perm = np.empty(keys.size, dtype=np.int32)
    _check(lib().rbfa_order_by_keys(keys.size, _ptr(keys), _ptr(perm)), "rbfa_order_by_keys")
    return Permutation(perm)


def morton_order(x, y, ndiv=16, bbox=None):
    return order_by_keys(morton_keys(x, y, ndiv, bbox))


def hilbert_order(x, y, ndiv=16, bbox=None):
    return order_by_keys(hilbert_keys(x, y, ndiv, bbox))


def morton_order_3d(x, y, z, ndiv=16, bbox=None):
    return order_by_keys(morton_keys_3d(x, y, z, ndiv, bbox))


def partition_last(pred):
    """Stable partition: items with pred False first, then True
    (partition_last, rbf_reorder.h:231-239)."""
    pred = np.asarray(pred, bool)
    return Permutation(np.argsort(pred, kind="stable").astype(np.int32))


def boundary_last_by_flag(flag):
    return partition_last(np.asarray(flag) != 0)


def boundary_last_by_index(n, bnd):
    mask = np.zeros(int(n), bool)
    mask[np.asarray(bnd, np.int64)] = True
    return partition_last(mask)


def renumber_stencils(ja, k, p):
    """A' = P A P^T on a fixed-width graph (returns a new array)."""
    ja = _i32(ja).copy()
    n = p.size()
    _check(lib().rbfa_renumber_stencils(n, int(k), _ptr(p.map()), _ptr(ja)),
           "rbfa_renumber_stencils")
    return ja


def make_row_ptr(n, k):
    ia = np.empty(n + 1, dtype=np.int32)
    _check(lib().rbfa_make_row_ptr(int(n), int(k), _ptr(ia)), "rbfa_make_row_ptr")
    return ia


def knn_stencils(k, x, y, z=None):
    """ja of shape (n*k,): the k nearest of node s at ja[s*k], self first."""
    x, y = _f64(x), _f64(y)
    z = None if z is None else _f64(z)
    ja = np.empty(x.size * int(k), dtype=np.int32)
    _check(lib().rbfa_knn_stencils(x.size, 2 if z is None else 3, _ptr(x), _ptr(y), _ptr(z),
                                   int(k), _ptr(ja)), "rbfa_knn_stencils")
    return ja


def jittered_grid(dim, m, jitter=0.25, seed=1):
    """Jittered Cartesian nodes on the unit square (dim 2) or cube (dim 3)."""
    n = m ** dim
    x, y = np.empty(n), np.empty(n)
    z = np.empty(n) if dim == 3 else None
    _check(lib().rbfa_gen_jittered_grid(int(dim), int(m), float(jitter), int(seed), _ptr(x),
                                        _ptr(y), _ptr(z)), "rbfa_gen_jittered_grid")
    return (x, y) if dim == 2 else (x, y, z)


def graded_disk(n_interior, n_boundary, seed=1):
    """Graded unit disk (ring spacing hc at the centre to hc/4 at the rim,
    jittered) plus n_boundary equispaced rim nodes; returns x, y, flag."""
    n = int(n_interior) + int(n_boundary)
    x, y = np.empty(n), np.empty(n)
    flag = np.empty(n, np.int32)
    _check(lib().rbfa_gen_graded_disk(int(n_interior), int(n_boundary), int(seed), _ptr(x),
                                      _ptr(y), _ptr(flag)), "rbfa_gen_graded_disk")
    return x, y, flag


def disk_darts(n_interior, n_boundary, seed=1):
    """Graded unit disk by seeded rejection sampling (spacing hc at the
    centre to hc/4 at the rim, Poisson-disk hard core) plus n_boundary
    equispaced rim nodes; returns x, y, flag."""
    n = int(n_interior) + int(n_boundary)
    x, y = np.empty(n), np.empty(n)
    flag = np.empty(n, np.int32)
    _check(lib().rbfa_gen_disk_darts(int(n_interior), int(n_boundary), int(seed), _ptr(x),
                                     _ptr(y), _ptr(flag)), "rbfa_gen_disk_darts")
    return x, y, flag


def ragged_stencils(sizes, x, y, z=None):
    """kNN stencils of per-node size sizes[s] as a ragged CSR graph (ia, ja):
    the first sizes[s] of each node's nearest neighbours (self first)."""
    sizes = np.ascontiguousarray(sizes, np.int32)
    kmax = int(sizes.max())
    full = knn_stencils(kmax, x, y, z).reshape(-1, kmax)
    ia = np.zeros(sizes.size + 1, np.int32)
    np.cumsum(sizes, out=ia[1:])
    mask = np.arange(kmax)[None, :] < sizes[:, None]
    return ia, np.ascontiguousarray(full[mask])


class NodeSet:
    """Node set with a renumbering chain (rbf_nodeset.h:46-235)."""

    def __init__(self, x, y, flag=None, z=None):
        self.x, self.y = _f64(x).copy(), _f64(y).copy()
        self.z = None if z is None else _f64(z).copy()
        self.flag = np.zeros(self.x.size, np.int32) if flag is None else _i32(flag).copy()
        self._file_order = Permutation.identity(self.x.size)
        self._rebuild_bnd()

    def _rebuild_bnd(self):
        self.bnd = np.nonzero(self.flag)[0].astype(np.int32)

    def num_points(self):
        return self.x.size

    def num_boundary(self):
        return self.bnd.size

    def num_interior(self):
        return self.x.size - self.bnd.size

    def renumber(self, p):
        self.x, self.y, self.flag = p.permute(self.x), p.permute(self.y), p.permute(self.flag)
        if self.z is not None:
            self.z = p.permute(self.z)
        self._rebuild_bnd()
        self._file_order = self._file_order.then(p)
        return self

    def file_order(self):
        return self._file_order

    def stencils(self, k):
        if k < 1 or k > self.x.size:
            raise ValueError(f"NodeSet.stencils: k={k} with {self.x.size} nodes")
        return knn_stencils(k, self.x, self.y, self.z)


# ---------------------------------------------------------------------------
# Device assembly
# ---------------------------------------------------------------------------

class Batch:
    """An uneven batch over size classes [(n, P, Q), ...] (2-D, FP64,
    Laplacian / d/dx / d/dy), bound to one ragged graph and captured into a
    hipGraph; launch() replays it."""

    def __init__(self, classes):
        arr = (_PlanDesc * len(classes))(*[_PlanDesc(2, n, P, Q, OPS_DIFF, 0, PREC_FP64, 3)
                                           for n, P, Q in classes])
        self._h = ctypes.c_void_p()
        _check(lib().rbfa_batch_create(len(classes), arr, ctypes.byref(self._h)),
               "rbfa_batch_create")
        self.classes = sorted(classes)
        self._keep = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.rbfa_batch_destroy(h)
            self._h = None

    def bind(self, ia_host, x, y, ia, ja, w, info=None, status=None, stream=None):
        import torch

        def dp(t):
            return None if t is None else ctypes.c_void_p(t.data_ptr())

        ia_host = _i32(ia_host)
        tab = (ctypes.c_void_p * 3)(*[t.data_ptr() for t in w])
        self._keep = (ia_host, tab, x, y, ia, ja, w, info, status)
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        _check(lib().rbfa_batch_bind(self._h, ia_host.size - 1, _ptr(ia_host), dp(x), dp(y), None,
                                     dp(ia), dp(ja), tab, dp(info), dp(status),
                                     ctypes.c_void_p(stream)), "rbfa_batch_bind")

    def class_count(self, c):
        return lib().rbfa_batch_class_count(self._h, int(c))

    def class_variant(self, c):
        return lib().rbfa_batch_class_variant(self._h, int(c)).decode()

    def class_num_variants(self, c):
        return lib().rbfa_batch_class_num_variants(self._h, int(c))

    def class_variant_name(self, c, v):
        r = lib().rbfa_batch_class_variant_name(self._h, int(c), int(v))
        return None if r is None else r.decode()

    def set_class_variant(self, c, v):
        _io_check(lib().rbfa_batch_set_class_variant(self._h, int(c), int(v)),
                  "rbfa_batch_set_class_variant")

    @property
    def mode(self):
        return lib().rbfa_batch_get_mode(self._h)

    @mode.setter
    def mode(self, m):
        _io_check(lib().rbfa_batch_set_mode(self._h, int(m)), "rbfa_batch_set_mode")

    def launch(self, stream=None):
        import torch
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        _check(lib().rbfa_batch_launch(self._h, ctypes.c_void_p(stream)), "rbfa_batch_launch")

    def autotune(self, reps=3, stream=None):
        import torch
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        ms = (ctypes.c_float * 2)()
        _check(lib().rbfa_batch_autotune(self._h, ctypes.c_void_p(stream), int(reps), ms),
               "rbfa_batch_autotune")
        return list(ms)


def _stream(stream):
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream if stream is None else stream)


def _dptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def knn_stencils_device(k, x, y, z=None, stream=None):
    """rbfa_knn_stencils on the GPU: x, y (, z) float64 CUDA tensors; returns
    the int32 CUDA tensor ja (n*k), identical to knn_stencils."""
    import torch
    n = x.numel()
    ja = torch.empty(n * int(k), dtype=torch.int32, device=x.device)
    _check(lib().rbfa_knn_stencils_device(n, 2 if z is None else 3, _dptr(x), _dptr(y), _dptr(z),
                                          int(k), _dptr(ja), _stream(stream)),
           "rbfa_knn_stencils_device")
    return ja


def curve_keys_device(x, y, z=None, curve="hilbert", ndiv=16, bbox=None, stream=None):
    """Morton / Hilbert keys (2-D) or octree Morton keys (3-D) on the GPU."""
    import torch
    n = x.numel()
    keys = torch.empty(n, dtype=torch.int64, device=x.device)
    bb = None if bbox is None else np.ascontiguousarray(bbox, dtype=np.float64)
    _check(lib().rbfa_curve_keys_device(n, 2 if z is None else 3, 1 if curve == "hilbert" else 0,
                                        _dptr(x), _dptr(y), _dptr(z),
                                        None if bb is None else _ptr(bb), int(ndiv), _dptr(keys),
                                        _stream(stream)), "rbfa_curve_keys_device")
    return keys


def order_by_keys_device(keys, key_bits=62, stream=None):
    """Stable order of non-negative int64 keys on the GPU (gather form)."""
    import torch
    perm = torch.empty(keys.numel(), dtype=torch.int32, device=keys.device)
    _check(lib().rbfa_order_by_keys_device(keys.numel(), _dptr(keys), int(key_bits), _dptr(perm),
                                           _stream(stream)), "rbfa_order_by_keys_device")
    return perm


def gather_device(perm, src, stream=None):
    """dst[i] = src[perm[i]] on the GPU."""
    import torch
    dst = torch.empty_like(src)
    _check(lib().rbfa_gather_f64_device(src.numel(), _dptr(perm), _dptr(src), _dptr(dst),
                                        _stream(stream)), "rbfa_gather_f64_device")
    return dst


def _io_check(status, what):
    if status != OK:
        raise RbfError(status, f"{what}: {lib().rbfa_last_error().decode()}")


def read_points(path):
    """x, y of a .points file (the reference's read_points, rbf_io.h:223-236)."""
    n = ctypes.c_int32()
    p = os.fsencode(path)
    _io_check(lib().rbfa_read_points(p, 0, ctypes.byref(n), None, None), "rbfa_read_points")
    x, y = np.empty(n.value), np.empty(n.value)
    _io_check(lib().rbfa_read_points(p, n.value, ctypes.byref(n), _ptr(x), _ptr(y)),
              "rbfa_read_points")
    return x, y


def read_graph(path, k=0):
    """(ia, ja) of a .graph file, 0-based (the reference's read_graph_csr,
    rbf_io.h:649-707); k > 0 demands rows of exactly k entries."""
    n, nnz = ctypes.c_int32(), ctypes.c_int64()
    p = os.fsencode(path)
    _io_check(lib().rbfa_read_graph(p, int(k), ctypes.byref(n), ctypes.byref(nnz), None, None, 0),
              "rbfa_read_graph")
    ia = np.empty(n.value + 1, np.int32)
    ja = np.empty(nnz.value, np.int32)
    _io_check(lib().rbfa_read_graph(p, int(k), ctypes.byref(n), ctypes.byref(nnz), _ptr(ia),
                                    _ptr(ja), ja.size), "rbfa_read_graph")
    return ia, ja


def write_matrix_market(path, ia, ja, a=None, cols=None, csr_base=0):
    """Matrix Market dump of a CSR operator (the reference's
    write_matrix_market, rbf_io.h:782-812); a=None writes the pattern."""
    ia, ja = _i32(ia), _i32(ja)
    rows = ia.size - 1
    a = None if a is None else _f64(a)
    _io_check(lib().rbfa_write_matrix_market(os.fsencode(path), rows,
                                             rows if cols is None else int(cols), _ptr(ia),
                                             _ptr(ja), None if a is None else _ptr(a),
                                             int(csr_base)), "rbfa_write_matrix_market")


def measure_peaks(stream=None):
    """Peaks of the GPU at hand, by microbenchmark: FP64 vector FMA and FP64
    MFMA TFLOP/s, HBM copy GB/s (the roofline peaks, re-measured)."""
    out = (ctypes.c_double * 3)()
    _check(lib().rbfa_measure_peaks(out, _stream(stream)), "rbfa_measure_peaks")
    return {"fp64_valu_tflops": out[0], "fp64_mfma_tflops": out[1], "hbm_copy_gbs": out[2]}


def gesv_batched(A, B=None, ipiv=None, info=None, stream=None):
    """Solve a batch of caller-filled systems in place (the reference's
    execute(A, ipiv, B, info), rbf_operators.h:29-37). A: float64 CUDA
    tensor (nbatch, nt, nt) holding each system column-major (A[b, j, i] is
    row i, column j); B: (nbatch, nrhs, nt), column-major likewise, or None.
    On return A holds the dgetrf-layout LU factors, B the solution; returns
    (ipiv (nbatch, nt) int32, 1-based; info (nbatch,) int32)."""
    import torch
    nbatch, nt = int(A.shape[0]), int(A.shape[1])
    nrhs = 0 if B is None else int(B.shape[1])
    if ipiv is None:
        ipiv = torch.empty((nbatch, nt), dtype=torch.int32, device=A.device)
    if info is None:
        info = torch.empty(nbatch, dtype=torch.int32, device=A.device)
    _check(lib().rbfa_gesv_batched(nt, nrhs, nbatch, _dptr(A), _dptr(ipiv), _dptr(B),
                                   _dptr(info), _stream(stream)), "rbfa_gesv_batched")
    return ipiv, info


def apply(ja, w, u, y, k=None, ia=None, bnd=None, stream=None):
    """Operator application (the mv_op of the reference's main.cpp:59-66):
    y[c] = W[c] u for each assembled operator c, then y[c][b] = u[b] for the
    rows b in bnd. ja, w[*], u, y[*], ia, bnd are torch CUDA tensors
    (int32 / float64); give k for fixed-width rows or ia for ragged rows."""
    import torch

    def dp(t):
        return None if t is None else ctypes.c_void_p(t.data_ptr())

    if (k is None) == (ia is None):
        raise ValueError("give exactly one of k (fixed width) or ia (row pointers)")
    if len(w) != len(y) or not 1 <= len(w) <= 4:
        raise ValueError("w and y must hold the same number (1..4) of operators")
    nrows = y[0].numel()
    nnz = int(ia[-1].item()) if ia is not None else nrows * int(k)
    for t in (ja, u, *w, *y) + ((ia,) if ia is not None else ()) + ((bnd,) if bnd is not None else ()):
        if not t.is_cuda or not t.is_contiguous():
            raise ValueError("device tensors must be contiguous CUDA tensors")
    if ja.dtype != torch.int32 or ja.numel() < nnz or any(
            t.dtype != torch.float64 or t.numel() < nnz for t in w):
        raise ValueError("ja must be int32 and w float64 with nnz entries")
    if u.dtype != torch.float64 or any(t.dtype != torch.float64 or t.numel() != nrows for t in y):
        raise ValueError("u and y must be float64, y with one entry per row")
    if ia is not None and (ia.dtype != torch.int32 or ia.numel() != nrows + 1):
        raise ValueError("ia must be int32 with nrows + 1 entries")
    if bnd is not None and bnd.dtype != torch.int32:
        raise ValueError("bnd must be int32")
    wt = (ctypes.c_void_p * len(w))(*[t.data_ptr() for t in w])
    yt = (ctypes.c_void_p * len(y))(*[t.data_ptr() for t in y])
    if stream is None:
        stream = torch.cuda.current_stream().cuda_stream
    nb = 0 if bnd is None else bnd.numel()
    _check(lib().rbfa_apply(nrows, 0 if k is None else int(k), dp(ia), dp(ja), len(w), wt, dp(u),
                            yt, nb, dp(bnd) if nb else None, ctypes.c_void_p(stream)),
           "rbfa_apply")


class ApplyPlan:
    """Planned operator application over one CSR structure (host arrays:
    ja, and ia or a fixed width k; bnd = identity rows): run(w, u, y) sets
    y[c] = W[c] u, equal bit for bit to the sequential CSR product."""

    def __init__(self, ja, k=None, ia=None, bnd=None, ncols=None):
        if (k is None) == (ia is None):
            raise ValueError("give exactly one of k (fixed width) or ia (row pointers)")
        ja = _i32(ja)
        ia = None if ia is None else _i32(ia)
        nrows = ia.size - 1 if ia is not None else ja.size // int(k)
        bnd = None if bnd is None else _i32(bnd)
        self._h = ctypes.c_void_p()
        _check(lib().rbfa_apply_plan_create(nrows, int(ncols or nrows), 0 if k is None else int(k),
                                            None if ia is None else _ptr(ia), _ptr(ja),
                                            0 if bnd is None else bnd.size,
                                            None if bnd is None else _ptr(bnd),
                                            ctypes.byref(self._h)), "rbfa_apply_plan_create")
        self.nrows = nrows
        self.nnz = int(ia[-1]) if ia is not None else nrows * int(k)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.rbfa_apply_plan_destroy(h)
            self._h = None

    def stats(self):
        out = np.zeros(5, np.int64)
        _check(lib().rbfa_apply_plan_stats(self._h, _ptr(out)), "rbfa_apply_plan_stats")
        return dict(zip(["blocks", "max_distinct", "distinct_total", "nnz", "device_bytes"],
                        out.tolist()))

    def run(self, w, u, y, stream=None):
        import torch
        if len(w) != len(y) or not 1 <= len(w) <= 4:
            raise ValueError("w and y must hold the same number (1..4) of operators")
        for t in (u, *w, *y):
            if not t.is_cuda or not t.is_contiguous() or t.dtype != torch.float64:
                raise ValueError("w, u, y must be contiguous float64 CUDA tensors")
        if any(t.numel() < self.nnz for t in w) or any(t.numel() != self.nrows for t in y):
            raise ValueError("w needs nnz entries, y one entry per row")
        wt = (ctypes.c_void_p * len(w))(*[t.data_ptr() for t in w])
        yt = (ctypes.c_void_p * len(y))(*[t.data_ptr() for t in y])
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        _check(lib().rbfa_apply_plan_run(self._h, len(w), wt, ctypes.c_void_p(u.data_ptr()), yt,
                                         ctypes.c_void_p(stream)), "rbfa_apply_plan_run")


class Plan:
    """A compiled stencil shape: dim, stencil size n, degree P, PHS r^Q,
    operator set ('diff' = Laplacian, d/dx, d/dy (, d/dz); 'interp' = nrhs
    centres), precision ('fp64' or 'mixed')."""

    def __init__(self, dim, n, degree, phs, op_set="diff", nrhs=0, precision="fp64",
                 max_refine=3):
        d = _PlanDesc(int(dim), int(n), int(degree), int(phs),
                      OPS_DIFF if op_set == "diff" else OPS_INTERP, int(nrhs),
                      PREC_FP64 if precision == "fp64" else PREC_MIXED, int(max_refine))
        self._h = ctypes.c_void_p()
        _check(lib().rbfa_plan_create(ctypes.byref(d), ctypes.byref(self._h)), "rbfa_plan_create")
        self.dim, self.n, self.degree, self.phs = dim, n, degree, phs
        self.op_set, self.precision = op_set, precision
        self.nops = lib().rbfa_plan_nops(self._h)
        self.nt = lib().rbfa_plan_system_size(self._h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.rbfa_plan_destroy(h)
            self._h = None

    @property
    def num_variants(self):
        return lib().rbfa_plan_num_variants(self._h)

    def variant_name(self, v=None):
        v = self.variant if v is None else v
        return lib().rbfa_plan_variant_name(self._h, int(v)).decode()

    @property
    def variant(self):
        return lib().rbfa_plan_get_variant(self._h)

    @variant.setter
    def variant(self, v):
        _check(lib().rbfa_plan_set_variant(self._h, int(v)), "rbfa_plan_set_variant")

    def autotune(self, nstencils, x, y, ja, w, z=None, centres=None, info=None, status=None,
                 reps=3, stream=None):
        """Time every kernel variant on these inputs, keep the fastest; returns
        the per-variant milliseconds."""
        args = self._args(nstencils, x, y, ja, w, z, centres, info, status, stream)
        ms = (ctypes.c_float * self.num_variants)()
        _check(lib().rbfa_plan_autotune(self._h, *args, int(reps), ms), "rbfa_plan_autotune")
        return list(ms)

    def assemble(self, nstencils, x, y, ja, w, z=None, centres=None, info=None, status=None,
                 stream=None, noalloc=False):
        """Enqueue the assembly. x, y, z, ja, w[*], info, status are torch CUDA
        tensors (float64 / int32 / uint8); centres = (xc, yc[, zc]) tensors.
        noalloc: call rbfa_assemble_noalloc (no device allocation; at most
        4194304 stencils; calls on one device must not overlap)."""
        args = self._args(nstencils, x, y, ja, w, z, centres, info, status, stream)
        fn = "rbfa_assemble_noalloc" if noalloc else "rbfa_assemble"
        _check(getattr(lib(), fn)(self._h, *args), fn)

    def _args(self, nstencils, x, y, ja, w, z, centres, info, status, stream):
        import torch

        def dp(t):
            return None if t is None else ctypes.c_void_p(t.data_ptr())

        if len(w) != self.nops:
            raise ValueError(f"expected {self.nops} weight arrays")
        nnz = int(nstencils) * self.n
        for t in (x, y, ja, *w):
            if not t.is_cuda or not t.is_contiguous():
                raise ValueError("device tensors must be contiguous CUDA tensors")
        if ja.dtype != torch.int32 or ja.numel() < nnz or any(
                t.dtype != torch.float64 or t.numel() < nnz for t in w):
            raise ValueError("ja must be int32 and w float64 with nstencils*n entries")
        tab = (ctypes.c_void_p * self.nops)(*[t.data_ptr() for t in w])
        cx = cy = cz = None
        if centres is not None:
            cx, cy = centres[0], centres[1]
            cz = centres[2] if len(centres) > 2 else None
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        self._keep = tab
        return (int(nstencils), dp(x), dp(y), dp(z), dp(ja), tab, dp(cx), dp(cy), dp(cz),
                dp(info), dp(status), ctypes.c_void_p(stream))
