/*
 * rbf_assembly.h -- C ABI of the MI355X (gfx950) RBF-FD operator-assembly
 * library (librbf_amd.so).
 *
 * Two layers sit behind this header:
 *
 *   1. The host node layer that produces the CSR structure: curve keys,
 *      stable renumbering and k-nearest-neighbour stencils. Plain host
 *      memory, OpenMP inside.
 *   2. Batched operator assembly on the GPU: per stencil, gather, build the
 *      polyharmonic-spline (PHS) + polynomial saddle-point system, factor it
 *      with partial-pivoting LU and scatter the weights into CSR value
 *      arrays that share the stencil column indices ja. All pointers given
 *      to the assembly entry points are DEVICE pointers unless stated.
 *
 * Conventions (mirroring the reference's bind(c) layer):
 *   - every entry point returns an int32 status, RBFA_OK == 0; where the
 *     reference host library prints to stderr and calls exit(1)
 *     (rbf_spatial.cpp:38-41, rbf_io.h:49-52), these return a code;
 *   - scalars by value, arrays as plain pointers, int32 indices, int64
 *     keys, double coordinates; 0-based indices throughout;
 *   - per-stencil solver status info[s]: 0 on success, k > 0 when the k-th
 *     pivot (1-based) was exactly zero, as the reference's per-block
 *     solver status (rbf_operators.h:315,365); the host policy "count the
 *     nonzero entries and fail" (main.cpp:48-54) is the caller's.
 *
 * Reference interfaces each entry point replaces are cited beside it as
 * file:line of ivan-pi/gpu-assembly.
 */
#ifndef RBF_ASSEMBLY_H
#define RBF_ASSEMBLY_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RBFA_API __attribute__((visibility("default")))

/* ---- status codes ------------------------------------------------------ */
enum {
    RBFA_OK = 0,
    RBFA_ERR_ARG = 1,         /* invalid argument (size, null pointer, range) */
    RBFA_ERR_UNSUPPORTED = 2, /* no kernel instantiated for this shape       */
    RBFA_ERR_HIP = 3,         /* a HIP runtime call failed                   */
    RBFA_ERR_NOMEM = 4,       /* host or device allocation failed            */
    RBFA_ERR_NODEVICE = 5,    /* no usable gfx950 device                     */
    RBFA_ERR_FORMAT = 6,      /* malformed input file (rbfa_last_error says where) */
    RBFA_ERR_IO = 7           /* a file could not be opened, read or written  */
};

/* ---- operator sets and precision modes --------------------------------- */
enum {
    /* Laplacian, d/dx, d/dy (and d/dz in 3-D), in this order: weights
     * written to w[0], w[1], w[2] (, w[3]). The Laplacian entry point of
     * the reference's pipeline sketch (main.cpp:15-16,44-45). */
    RBFA_OPS_DIFF = 0,
    /* Interpolation to nrhs centres given in the stencil-local frame
     * (offsets from the stencil's own node), as assemble_interp_weights
     * (rbf_operators.h:320-376, fill_rhs :158-185). */
    RBFA_OPS_INTERP = 1
};

enum {
    RBFA_PREC_FP64 = 0,  /* plain FP64 LU and solve                         */
    RBFA_PREC_MIXED = 1  /* FP32 LU + FP64 iterative refinement, FP64 fallback */
};

/* Bits of the optional per-stencil status byte written by the mixed form. */
enum {
    RBFA_STAT_STEPS_MASK = 0x03, /* refinement steps taken (0..3)            */
    RBFA_STAT_FALLBACK = 0x04    /* did not converge: solved again in FP64   */
};

/* Layout-compatible with rbf::BBox2<double> (rbf_reorder.h:30-40). */
typedef struct rbfa_bbox2 {
    double xmin, ymin, xmax, ymax;
} rbfa_bbox2;

RBFA_API const char* rbfa_version(void);

/* Re-measured peaks of the GPU at hand (no counterpart in the reference;
 * the benchmark prices its roofline against them beside the datasheet
 * figures): out[0] FP64 vector FMA TFLOP/s, out[1] FP64 MFMA
 * (v_mfma_f64_16x16x4f64) TFLOP/s, out[2] HBM copy GB/s (1 GiB each way).
 * Allocates and frees its own buffers; synchronises the stream. */
RBFA_API int32_t rbfa_measure_peaks(double* out, void* stream);
RBFA_API const char* rbfa_status_string(int32_t status);

/* ======================================================================
 * Host node layer (all pointers are host pointers)
 * ====================================================================== */

/* Morton / Hilbert keys of n points inside bbox, ndiv quadtree levels
 * (1..31), 2 bits per level. bbox == NULL: the bounding box of the points.
 * Replaces rbf_morton_keys / rbf_hilbert_keys (rbf_ordering.F90:31-67,
 * zidx :82-141, hidx :144-226) behind morton_keys / hilbert_keys
 * (rbf_reorder.h:198-210). */
RBFA_API int32_t rbfa_morton_keys(int32_t n, const double* x, const double* y,
                                  const rbfa_bbox2* bbox, int32_t ndiv, int64_t* keys);
RBFA_API int32_t rbfa_hilbert_keys(int32_t n, const double* x, const double* y,
                                   const rbfa_bbox2* bbox, int32_t ndiv, int64_t* keys);

/* 3-D Morton keys, 3 bits per level, ndiv 1..21; bbox = {xmin, ymin, zmin,
 * xmax, ymax, zmax} or NULL. The reference's curves are 2-D only
 * (rbf_ordering.F90:8); this extends zidx's midpoint rule to an octree. */
RBFA_API int32_t rbfa_morton_keys_3d(int32_t n, const double* x, const double* y,
                                     const double* z, const double* bbox6, int32_t ndiv,
                                     int64_t* keys);

/* Stable order of the keys: perm[i] is the old index of the item placed
 * at new position i (gather form). Replaces detail::order_by_keys
 * (rbf_reorder.h:183-189) and hence morton_order / hilbert_order
 * (:214-229). */
RBFA_API int32_t rbfa_order_by_keys(int32_t n, const int64_t* keys, int32_t* perm);

/* a_new[i] = a_old[perm[i]], in place. Permutation::permute
 * (rbf_reorder.h:108-115). */
RBFA_API int32_t rbfa_permute_f64(int32_t n, const int32_t* perm, double* a);
RBFA_API int32_t rbfa_permute_i32(int32_t n, const int32_t* perm, int32_t* a);

/* Fixed-width stencil graph renumbered in place, A' = P A P^T
 * (renumber_stencils, rbf_reorder.h:278-288). */
RBFA_API int32_t rbfa_renumber_stencils(int32_t n, int32_t k, const int32_t* perm,
                                        int32_t* ja);

/* Row pointer {0, k, 2k, ...} of length n+1 (make_row_ptr,
 * rbf_reorder.h:316-322). */
RBFA_API int32_t rbfa_make_row_ptr(int32_t n, int32_t k, int32_t* ia);

/* k nearest neighbours of every node, ja[s*k + j], sorted by squared
 * distance (dx*dx + dy*dy (+ dz*dz), query minus point), ties by the
 * lower index, so ja[s*k] == s for distinct points. dim is 2 or 3 (z
 * ignored for 2). Replaces NodeSet::stencils (rbf_nodeset.h:155-181,
 * nanoflann knnSearch) and KdTree::knn_stencils (rbf_spatial.cpp:251-260). */
RBFA_API int32_t rbfa_knn_stencils(int32_t n, int32_t dim, const double* x, const double* y,
                                   const double* z, int32_t k, int32_t* ja);

/* Synthetic node sets (seeded xorshift64, the generator family of the
 * reference's test_spatial.cpp:39-47).
 *
 * Jittered Cartesian grid on the unit square / cube: m points per axis at
 * ((i + 1/2) h) + jitter * h * (2u - 1), h = 1/m, row-major with x
 * fastest. dim 2 writes m^2 points to x, y; dim 3 writes m^3 to x, y, z. */
RBFA_API int32_t rbfa_gen_jittered_grid(int32_t dim, int32_t m, double jitter, uint64_t seed,
                                        double* x, double* y, double* z);

/* Graded unit disk: n_interior nodes on concentric rings whose spacing
 * falls linearly from hc at the centre to hc/4 at the rim (seeded phase and
 * +-0.15h jitter per node, a Poisson-disk-like set; hc solved for the
 * count), then n_boundary equispaced nodes on the unit circle. flag is 0
 * for interior, 1 for boundary nodes. Interior nodes come first. */
RBFA_API int32_t rbfa_gen_graded_disk(int32_t n_interior, int32_t n_boundary, uint64_t seed,
                                      double* x, double* y, int32_t* flag);

/* Graded unit disk by seeded rejection sampling (dart throwing): target
 * spacing h(r) = hc (1 - 0.75 r), spacing ratio 1:4 from centre to rim, hc
 * solved for n_interior nodes at one node per h^2. Candidates are drawn
 * from that density and kept when every kept node q lies at least
 * 0.7136 (h(p) + h(q)) / 2 away (a Poisson-disk set at packing fraction
 * 0.40); the n_boundary equispaced nodes on the unit circle are placed
 * first, so the interior keeps its distance from them too. Same output
 * layout as rbfa_gen_graded_disk. RBFA_ERR_ARG if the darts jam before
 * n_interior nodes (a boundary far denser than the rim spacing). */
RBFA_API int32_t rbfa_gen_disk_darts(int32_t n_interior, int32_t n_boundary, uint64_t seed,
                                     double* x, double* y, int32_t* flag);

/* ---- the node layer on the device (device pointers, stream-ordered) -----
 * Same results as the host functions above, bit for bit, for the same
 * points: the same keys, the same stable order, the same stencils (the
 * device search uses the host search's bucket grid, ring order and
 * stopping rule). Each call synchronises `stream` before it returns and
 * allocates its temporaries itself (hipMalloc / hipFree).
 *
 * rbfa_knn_stencils_device: rbfa_knn_stencils (NodeSet::stencils,
 * rbf_nodeset.h:155-181) for k <= 128, one wave per node.
 * rbfa_curve_keys_device: dim 2, curve 0 = Morton / 1 = Hilbert
 * (rbf_ordering.F90 zidx / hidx, rbf_reorder.h:198-210), bbox =
 * {xmin, ymin, xmax, ymax}; dim 3, curve 0 = octree Morton, bbox =
 * {xmin, ymin, zmin, xmax, ymax, zmax}; bbox NULL: the points' own box.
 * rbfa_order_by_keys_device: stable order of keys in [0, 2^key_bits)
 * (detail::order_by_keys, rbf_reorder.h:183-189); perm is the gather form.
 * rbfa_gather_f64_device: dst[i] = src[perm[i]] (Permutation::permute,
 * rbf_reorder.h:108-115), src != dst. */
RBFA_API int32_t rbfa_knn_stencils_device(int32_t n, int32_t dim, const double* x,
                                          const double* y, const double* z, int32_t k,
                                          int32_t* ja, void* stream);
RBFA_API int32_t rbfa_curve_keys_device(int32_t n, int32_t dim, int32_t curve, const double* x,
                                        const double* y, const double* z, const double* bbox,
                                        int32_t ndiv, int64_t* keys, void* stream);
RBFA_API int32_t rbfa_order_by_keys_device(int32_t n, const int64_t* keys, int32_t key_bits,
                                           int32_t* perm, void* stream);
RBFA_API int32_t rbfa_gather_f64_device(int32_t n, const int32_t* perm, const double* src,
                                        double* dst, void* stream);

/* ======================================================================
 * Device assembly
 * ====================================================================== */

typedef struct rbfa_plan_desc {
    int32_t dim;        /* 2 or 3                                            */
    int32_t n;          /* stencil size (neighbours per node, self first)   */
    int32_t degree;     /* polynomial augmentation degree P (0..4)          */
    int32_t phs;        /* PHS exponent Q: phi(r) = r^Q, odd, 3..7          */
    int32_t op_set;     /* RBFA_OPS_DIFF or RBFA_OPS_INTERP                 */
    int32_t nrhs;       /* RBFA_OPS_INTERP: number of centres; else ignored */
    int32_t precision;  /* RBFA_PREC_FP64 or RBFA_PREC_MIXED                */
    int32_t max_refine; /* mixed form: refinement steps before fallback, 1..3 */
} rbfa_plan_desc;

typedef struct rbfa_plan rbfa_plan;

/* The plan replaces the compile-time interp_config<N,NRHS,P,Q,ARCH,T>
 * (rbf_operators.h:187-290): it picks the kernel instantiation and launch
 * shape for the gfx950 device current at creation time. */
RBFA_API int32_t rbfa_plan_create(const rbfa_plan_desc* desc, rbfa_plan** plan);
RBFA_API int32_t rbfa_plan_destroy(rbfa_plan* plan);
/* Number of operators (value arrays) the plan writes. */
RBFA_API int32_t rbfa_plan_nops(const rbfa_plan* plan);
/* nt = n + npoly, the order of each saddle-point system. */
RBFA_API int32_t rbfa_plan_system_size(const rbfa_plan* plan);

/* Assemble nstencils stencils on `stream` (a hipStream_t; NULL = default).
 *
 *   x, y, z   node coordinates (device; z only for dim 3), indexed by ja
 *   ja        nstencils*n column indices, stencil s at ja[s*n], ja[s*n] == s
 *   w         HOST array of nops DEVICE pointers, each to nstencils*n
 *             doubles laid out like ja: w[op][s*n + k] is the weight of
 *             node ja[s*n + k] in operator op of row s
 *   xc,yc,zc  RBFA_OPS_INTERP only: nrhs centres (device), local frame
 *   info      nstencils int32 solver statuses (device), may be NULL
 *   status    nstencils status bytes (device, RBFA_STAT_*), may be NULL
 *
 * Replaces the launch of assemble_interp_weights<Config>
 * (rbf_operators.h:317-328) and the sketched assemble_laplace_weights
 * (main.cpp:44-45). Asynchronous: returns once the work is enqueued.
 * Mixed-form plans of systems above 64 unknowns (2-D n = 56, 3-D n = 80)
 * keep their FP64-fallback list in a workspace of the plan (grown here,
 * with a synchronisation of `stream` when it grows): calls with one such
 * plan must not overlap (one stream, one host thread at a time), or a
 * call can clear the other's list and leave its listed stencils unsolved.
 * Every other plan keeps no state between calls. */
RBFA_API int32_t rbfa_assemble(const rbfa_plan* plan, int32_t nstencils, const double* x,
                               const double* y, const double* z, const int32_t* ja,
                               double* const* w, const double* xc, const double* yc,
                               const double* zc, int32_t* info, uint8_t* status, void* stream);

/* rbfa_assemble without any allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * arguments, checks and results as rbfa_assemble; the mixed-form shapes
 * whose stencils may be listed for an FP64 pass (2-D n = 56, 3-D n = 80)
 * keep that list in a fixed per-device buffer inside the library (16 MiB,
 * held until the process ends) instead of the plan's own workspace.
 * Limit: nstencils <= 4194304 (1 << 22) per call; larger calls return
 * RBFA_ERR_ARG and rbfa_last_error says why. Calls that overlap on one
 * device (other streams, other host threads) share that buffer and race:
 * order them on one stream. Call with the plan's device current. */
RBFA_API int32_t rbfa_assemble_noalloc(const rbfa_plan* plan, int32_t nstencils, const double* x,
                                       const double* y, const double* z, const int32_t* ja,
                                       double* const* w, const double* xc, const double* yc,
                                       const double* zc, int32_t* info, uint8_t* status,
                                       void* stream);

/* ---- kernel variants and the autotuner --------------------------------
 * Each shape ships several kernel variants (pivot-row broadcast through
 * SGPRs or LDS, pivot search by shuffles or DPP, waves per workgroup). All
 * give identical results; they differ in speed. These have no counterpart
 * in the reference, where cuSolverDx fixes block_dim at compile time
 * (rbf_operators.h:280-281). */
RBFA_API int32_t rbfa_plan_num_variants(const rbfa_plan* plan);
RBFA_API const char* rbfa_plan_variant_name(const rbfa_plan* plan, int32_t variant);
RBFA_API int32_t rbfa_plan_get_variant(const rbfa_plan* plan);
RBFA_API int32_t rbfa_plan_set_variant(rbfa_plan* plan, int32_t variant);
/* Time every variant on the given inputs (reps launches each, HIP events on
 * `stream`, synchronising it), keep the fastest in the plan (a later
 * variant must beat the best earlier one by more than 2%, so run-to-run
 * noise keeps the lower-numbered variant; the current variant runs reps+1
 * times untimed first), and write the
 * per-variant milliseconds to ms_out[num_variants] if non-NULL. The
 * outputs w/info/status are overwritten. */
RBFA_API int32_t rbfa_plan_autotune(rbfa_plan* plan, int32_t nstencils, const double* x,
                                    const double* y, const double* z, const int32_t* ja,
                                    double* const* w, const double* xc, const double* yc,
                                    const double* zc, int32_t* info, uint8_t* status,
                                    void* stream, int32_t reps, float* ms_out);

/* ---- uneven batches ------------------------------------------------------
 * Stencils of several sizes in one ragged CSR graph (row s at ia[s], of
 * length ia[s+1] - ia[s]; the length names the size class). Replaces the
 * fixed-width launch of assemble_interp_weights (rbf_operators.h:307-310,
 * implicit ia[s] = s*N) for graphs like the reference's ragged .graph files
 * (read_graph_csr, rbf_io.h:649-707; rbf_cuda.f90:52-56). Shipped classes
 * (2-D, FP64, Laplacian/d/dx/d/dy): n/P/Q = 13/2/3, 21/2/3, 35/3/5, 56/4/7.
 *
 * bind captures the pipeline into a hipGraph with the given device arrays
 * (w: host array of 3 device pointers laid out like ja); launch replays
 * it. Mode 0: one launch per class; mode 1: one persistent launch over
 * size-bucketed work queues, currently switched off (set_mode(1) returns
 * RBFA_ERR_UNSUPPORTED and leaves the reason in rbfa_last_error; so does
 * set_class_variant, with RBFA_ERR_ARG, for a class or variant past the
 * lists below). autotune picks each class's kernel variant and
 * the faster enabled mode (ms_out[2] = mode 0, mode 1 milliseconds; -1 for
 * a mode that is switched off). Calls that overlap on one batch race on
 * its work-queue counters.
 *
 * Each class (numbered in the order of `classes`) lists the shapes its
 * list launch can take: several stencils per wave (16- or 32-lane
 * segments, where the system has at most 32 unknowns), one stencil per
 * wave with 1, 2 or 4 waves per workgroup, a blocked-LU workgroup per
 * stencil above 64 unknowns. class_num_variants / class_variant_name list
 * them, set_class_variant picks one (autotune picks the fastest);
 * class_variant names the current one. All give the same weights. */
typedef struct rbfa_batch rbfa_batch;
RBFA_API int32_t rbfa_batch_create(int32_t nclasses, const rbfa_plan_desc* classes,
                                   rbfa_batch** batch);
RBFA_API int32_t rbfa_batch_destroy(rbfa_batch* batch);
RBFA_API int32_t rbfa_batch_bind(rbfa_batch* batch, int32_t nstencils, const int32_t* ia_host,
                                 const double* x, const double* y, const double* z,
                                 const int32_t* ia, const int32_t* ja, double* const* w,
                                 int32_t* info, uint8_t* status, void* stream);
RBFA_API int32_t rbfa_batch_class_count(const rbfa_batch* batch, int32_t cls);
RBFA_API const char* rbfa_batch_class_variant(const rbfa_batch* batch, int32_t cls);
RBFA_API int32_t rbfa_batch_class_num_variants(const rbfa_batch* batch, int32_t cls);
RBFA_API const char* rbfa_batch_class_variant_name(const rbfa_batch* batch, int32_t cls,
                                                   int32_t v);
RBFA_API int32_t rbfa_batch_set_class_variant(rbfa_batch* batch, int32_t cls, int32_t v);
RBFA_API int32_t rbfa_batch_set_mode(rbfa_batch* batch, int32_t mode);
RBFA_API int32_t rbfa_batch_get_mode(const rbfa_batch* batch);
RBFA_API int32_t rbfa_batch_launch(rbfa_batch* batch, void* stream);
RBFA_API int32_t rbfa_batch_autotune(rbfa_batch* batch, void* stream, int32_t reps, float* ms_out);

/* ------------------------------------------------------------ caller-filled systems
 * The reference's solve of a system its caller filled: execute(A, ipiv, B,
 * info) of gesv_partial_pivot (rbf_operators.h:29-37), exported to Fortran
 * as extern "C" solve_dp (rbf_operators.cpp:7-11, a 15 x 15 system with
 * NRHS = 9, launched from main.f90:57). Device pointers, stream-ordered.
 *
 * rbfa_gesv_batched: nbatch independent systems, one after another in
 * memory: system b is A + b*nt*nt (column-major nt x nt, leading dimension
 * nt), right-hand sides B + b*nt*nrhs (column-major nt x nrhs), pivots
 * ipiv + b*nt, status info[b] (info may be NULL). 1 <= nt <= 64,
 * 0 <= nrhs <= 16 (0: factor only, B may be NULL). On return A holds the
 * LU factors in LAPACK dgetrf layout (L unit lower, U upper, the row that
 * was pivot at step k in row k), ipiv the 1-based dgetrf interchanges, B
 * the solution, info[b] 0 or the 1-based index of the first exactly zero
 * pivot. The pivot is the first largest |a| in row order; the factors and
 * the solution are bit-identical to the textbook FP64 elimination.
 *
 * rbfa_solve_dp: one 15 x 15 system with 9 right-hand sides (solve_dp). */
RBFA_API int32_t rbfa_gesv_batched(int32_t nt, int32_t nrhs, int32_t nbatch, double* A,
                                   int32_t* ipiv, double* B, int32_t* info, void* stream);
RBFA_API int32_t rbfa_solve_dp(double* A, int32_t* ipiv, double* B, int32_t* info, void* stream);

/* ------------------------------------------------------------ operator application
 * Replaces the mv_op of src/main.cpp:59-66 (A.spmv(x, y), then
 * y[bnd[k]] = x[bnd[k]]): y_c = W_c u for nops (1..4) assembled operators
 * sharing one CSR structure (ia: nrows+1 device row pointers, or NULL for
 * fixed width k), then the identity on the nbnd listed rows of every y_c.
 * All pointers are device pointers except the host arrays w and y of nops
 * device pointers each. y_c must not alias u. Row sums are reassociated
 * (eight partial sums per row), so results agree with a sequential sum to
 * within k * eps * sum_j |w_j u_j|. */
RBFA_API int32_t rbfa_apply(int32_t nrows, int32_t k, const int32_t* ia, const int32_t* ja,
                            int32_t nops, double* const* w, const double* u, double* const* y,
                            int32_t nbnd, const int32_t* bnd, void* stream);

/* ------------------------------------------------------------ file I/O
 * The reference's file formats (rbf_io.h), host arrays only.
 *
 * rbfa_read_points replaces read_points (rbf_io.h:223-236): the point count
 * on the first line, then "x y" per line. Call with cap = 0 for the count
 * alone, then with cap >= n.
 *
 * rbfa_read_graph replaces read_graph_csr (rbf_io.h:649-707): header
 * "n nnz", then one line of 0-based stencil indices per node; k > 0
 * demands exactly k entries per row, k = 0 accepts ragged rows. Every row
 * needs an entry, indices lie in [0, n) and appear once per row, the entry
 * count matches nnz, nothing follows the n-th row. Call with ia = ja = NULL
 * for n and nnz, then with ia (n+1) and ja (cap >= nnz).
 *
 * rbfa_write_matrix_market replaces write_matrix_market (rbf_io.h:782-812):
 * "%%MatrixMarket matrix coordinate real general" (or "pattern" when
 * a == NULL), "rows cols nnz", then 1-based "i j a_ij" lines in CSR order,
 * each value the shortest text that reads back exactly; csr_base is the
 * base of ia / ja (0 or 1).
 *
 * On RBFA_ERR_FORMAT / RBFA_ERR_IO, rbfa_last_error (per thread) names the
 * file and the problem, as the reference's exception text does. */
RBFA_API const char* rbfa_last_error(void);
RBFA_API int32_t rbfa_read_points(const char* path, int32_t cap, int32_t* n, double* x, double* y);
RBFA_API int32_t rbfa_read_graph(const char* path, int32_t k, int32_t* n, int64_t* nnz,
                                 int32_t* ia, int32_t* ja, int64_t cap);
RBFA_API int32_t rbfa_write_matrix_market(const char* path, int64_t rows, int64_t cols,
                                          const int32_t* ia, const int32_t* ja, const double* a,
                                          int32_t csr_base);

/* Planned operator application: the same mv_op (src/main.cpp:59-66) for a
 * structure applied many times (the reference applies it once per solver
 * iteration). create takes HOST arrays (ia: nrows+1 row pointers or NULL
 * for fixed width k; ja: column indices in [0, ncols); bnd: the nbnd
 * identity rows) and builds, on the current device, a blocked copy of the
 * structure: consecutive row blocks (<= 256 rows, <= 2048 nonzeros), each
 * with its sorted distinct columns and a 16-bit local column index per
 * nonzero. run computes y_c = W_c u (identity on the bnd rows) for nops
 * (1..4) weight arrays in the original CSR entry order; every row is summed
 * left to right, so y equals the sequential CSR product bit for bit.
 * stats: out[0..4] = blocks, most distinct columns in a block, total
 * distinct-column entries, nonzeros, device bytes held by the plan. */
typedef struct rbfa_apply_plan rbfa_apply_plan;
RBFA_API int32_t rbfa_apply_plan_create(int32_t nrows, int32_t ncols, int32_t k,
                                        const int32_t* ia, const int32_t* ja, int32_t nbnd,
                                        const int32_t* bnd, rbfa_apply_plan** plan);
RBFA_API int32_t rbfa_apply_plan_destroy(rbfa_apply_plan* plan);
RBFA_API int32_t rbfa_apply_plan_stats(const rbfa_apply_plan* plan, int64_t* out);
RBFA_API int32_t rbfa_apply_plan_run(const rbfa_apply_plan* plan, int32_t nops, double* const* w,
                                     const double* u, double* const* y, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* RBF_ASSEMBLY_H */
