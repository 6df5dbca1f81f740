/*
 * rbf_oracle.c -- CPU FP64 oracle for batched RBF-FD operator assembly.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product library links, loads or
 * calls this file; only the tests, the smoke check and the CPU-baseline leg
 * of the benchmark use it, as the checker. It restates, in plain C, the
 * reference's (ivan-pi/gpu-assembly) algorithm for this path:
 *
 *   phs()          PHS<Q>::operator()            rbf_operators.h:39-76
 *   poly_basis()   PolyBasis<P> (degree-major)   rbf_operators.h:78-114
 *   oracle_fill()  fill_matrix<N,P,Q>            rbf_operators.h:116-154
 *   interp rhs     fill_rhs<N,P,NRHS,Q>          rbf_operators.h:156-185
 *   derivative rhs Laplacian / d/dx / d/dy / d/dz of the basis at the
 *                  stencil centre (no reference code; the Laplacian entry
 *                  point is sketched at main.cpp:15-16 and packed as
 *                  W(:,1)+W(:,2) at rbf_cuda.f90:139-157)
 *   gather/scatter assemble_interp_weights       rbf_operators.h:340-375
 *   oracle_getrf() the factors and pivots that GESV leaves in A and ipiv
 *                  (execute(A, ipiv, B, info), rbf_operators.h:29-37): the
 *                  same elimination as oracle_gesv, written out in LAPACK
 *                  dgetrf layout (row k of the output = the row that was
 *                  pivot at step k; 1-based sequential-swap ipiv); pinned
 *                  against scipy's dgetrf (tests/test_oracle.py)
 *   oracle_gesv()  the block GESV of cuSolverDx (rbf_operators.h:365),
 *                  closed and absent: restated as textbook partial-pivot
 *                  LU (Golub & Van Loan Alg. 3.4.1) -- parity unpinned
 *                  against cuSolverDx itself; pinned by backward error,
 *                  polynomial reproduction and a long-double brute force
 *   keys           zidx / hidx                    rbf_ordering.F90:82-226
 *   knn            brute force by (d2, index), the check of
 *                  test_spatial.cpp:70-101
 *
 * Local frame. Nodes are shifted to the stencil's own node (as the
 * reference does, rbf_operators.h:354-358) and scaled by the stencil
 * radius rho = sqrt(max_k |p_k - p_s|^2); the fill takes any local frame
 * (rbf_operators.h:123-124). Derivative weights are unscaled by rho^-2
 * (Laplacian) and rho^-1 (first derivatives) afterwards.
 *
 * Elimination order (shared with the GPU FP64 kernel, so the two round the
 * same way): rows are never moved; at step k the pivot is the first row,
 * in original order, of largest |a_ik| among rows not yet pivotal;
 * rcp = 1/pivot; each remaining row takes l = a_ik * rcp and
 * a_ij = fma(-l, a_pj, a_ij) for j > k, right-hand sides included, so the
 * forward substitution happens during the elimination (dgetf2 + dgetrs
 * order). Back substitution runs k = nt-1 .. 0 with t = b_p * rcp_p and
 * b_i = fma(-u_ik, t, b_i). info = first k+1 whose pivot is exactly 0.
 *
 * Built with -ffp-contract=off: every fused multiply-add is an explicit
 * fma() call, every other product and sum rounds on its own.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_API __attribute__((visibility("default")))

enum { OR_OPS_DIFF = 0, OR_OPS_INTERP = 1 };

/* PHS<Q> (rbf_operators.h:39-76); q = 1 is the plain radius, used by the
 * derivative right-hand sides (r^(Q-2) with Q = 3). */
static double phs(int q, double r2) {
    switch (q) {
        case 1: return sqrt(r2);
        case 3: return r2 * sqrt(r2);
        case 5: return r2 * r2 * sqrt(r2);
        case 7: {
            const double r4 = r2 * r2;
            return r4 * r2 * sqrt(r2);
        }
        default: {
            const double r = sqrt(r2);
            double p = r;
            for (int i = 0; i < (q - 1) / 2; ++i) p *= r2;
            return p;
        }
    }
}

OR_API int oracle_npoly(int dim, int P) {
    if (P < 0) return 0;
    if (dim == 2) return (P + 1) * (P + 2) / 2;
    return (P + 1) * (P + 2) * (P + 3) / 6;
}

/* Monomials of degree <= P in degree-major order. 2-D: for d = 0..P, for
 * j = 0..d: x^(d-j) y^j (rbf_operators.h:93-98), powers by repeated
 * products X[i] = X[i-1]*x. 3-D (no reference code): for d, for j = 0..d
 * (degree in y and z together), for l = 0..j: x^(d-j) y^(j-l) z^l,
 * formed as (X[d-j] * Y[j-l]) * Z[l]; with z absent it is the 2-D order. */
static void poly_basis(int dim, int P, double x, double y, double z, double* b, int inc) {
    double X[8], Y[8], Z[8];
    X[0] = Y[0] = Z[0] = 1.0;
    for (int i = 1; i <= P; ++i) {
        X[i] = X[i - 1] * x;
        Y[i] = Y[i - 1] * y;
        Z[i] = Z[i - 1] * z;
    }
    int k = 0;
    for (int d = 0; d <= P; ++d) {
        if (dim == 2) {
            for (int j = 0; j <= d; ++j) b[(k++) * inc] = X[d - j] * Y[j];
        } else {
            for (int j = 0; j <= d; ++j)
                for (int l = 0; l <= j; ++l) b[(k++) * inc] = (X[d - j] * Y[j - l]) * Z[l];
        }
    }
}

/* Exponents (ex, ey, ez) of monomial t in the order above. */
static void poly_exponents(int dim, int P, int t, int* ex, int* ey, int* ez) {
    int k = 0;
    for (int d = 0; d <= P; ++d) {
        if (dim == 2) {
            for (int j = 0; j <= d; ++j, ++k)
                if (k == t) { *ex = d - j; *ey = j; *ez = 0; return; }
        } else {
            for (int j = 0; j <= d; ++j)
                for (int l = 0; l <= j; ++l, ++k)
                    if (k == t) { *ex = d - j; *ey = j - l; *ez = l; return; }
        }
    }
    *ex = *ey = *ez = -1;
}

static double sqdist(int dim, double dx, double dy, double dz) {
    double s = dx * dx + dy * dy;
    if (dim == 3) s = s + dz * dz;
    return s;
}

/* fill_matrix (rbf_operators.h:123-154) for one stencil in its local
 * frame: M is nt x nt, column-major, ldm = nt. */
OR_API void oracle_fill(int dim, int n, int P, int Q, const double* xs, const double* ys,
                        const double* zs, double* M) {
    const int m = oracle_npoly(dim, P), nt = n + m;
    for (int col = 0; col < n; ++col) {
        const double xc = xs[col], yc = ys[col], zc = dim == 3 ? zs[col] : 0.0;
        for (int k = 0; k < n; ++k) {
            const double dx = xc - xs[k], dy = yc - ys[k], dz = dim == 3 ? zc - zs[k] : 0.0;
            M[k + col * nt] = phs(Q, sqdist(dim, dx, dy, dz));
        }
    }
    for (int k = 0; k < n; ++k) {
        const double zk = dim == 3 ? zs[k] : 0.0;
        poly_basis(dim, P, xs[k], ys[k], zk, &M[k + n * nt], nt); /* upper right */
        poly_basis(dim, P, xs[k], ys[k], zk, &M[n + k * nt], 1);  /* lower left  */
    }
    for (int c = 0; c < m; ++c)
        for (int r = 0; r < m; ++r) M[(n + r) + (n + c) * nt] = 0.0;
}

/* Right-hand sides, nt x r column-major (ldb = nt).
 * DIFF: columns Laplacian, d/dx, d/dy (, d/dz) of phi(|x - p_k|) and of the
 * monomials, evaluated at the origin of the local frame:
 *   Lap phi  = Q (Q + dim - 2) r^(Q-2)
 *   d/dx phi = -(Q r^(Q-2)) x_k
 *   Lap x^2 = Lap y^2 = Lap z^2 = 2, d/dx x = d/dy y = d/dz z = 1, else 0.
 * INTERP: column c is the reference's interpolation RHS at centre c
 * (fill_rhs, rbf_operators.h:172-184). */
OR_API void oracle_rhs(int dim, int n, int P, int Q, int op_set, int nrhs, const double* xs,
                       const double* ys, const double* zs, const double* cx, const double* cy,
                       const double* cz, double* B) {
    const int m = oracle_npoly(dim, P), nt = n + m;
    if (op_set == OR_OPS_DIFF) {
        const double clap = (double)(Q * (Q + dim - 2));
        for (int k = 0; k < n; ++k) {
            const double zk = dim == 3 ? zs[k] : 0.0;
            const double t = phs(Q - 2, sqdist(dim, xs[k], ys[k], zk));
            const double g = (double)Q * t;
            B[k + 0 * nt] = clap * t;
            B[k + 1 * nt] = -g * xs[k];
            B[k + 2 * nt] = -g * ys[k];
            if (dim == 3) B[k + 3 * nt] = -g * zk;
        }
        for (int t = 0; t < m; ++t) {
            int ex, ey, ez;
            poly_exponents(dim, P, t, &ex, &ey, &ez);
            const int deg = ex + ey + ez;
            B[n + t + 0 * nt] = (deg == 2 && (ex == 2 || ey == 2 || ez == 2)) ? 2.0 : 0.0;
            B[n + t + 1 * nt] = (deg == 1 && ex == 1) ? 1.0 : 0.0;
            B[n + t + 2 * nt] = (deg == 1 && ey == 1) ? 1.0 : 0.0;
            if (dim == 3) B[n + t + 3 * nt] = (deg == 1 && ez == 1) ? 1.0 : 0.0;
        }
    } else {
        for (int c = 0; c < nrhs; ++c) {
            const double xe = cx[c], ye = cy[c], ze = dim == 3 ? cz[c] : 0.0;
            for (int k = 0; k < n; ++k) {
                const double dx = xe - xs[k], dy = ye - ys[k], dz = dim == 3 ? ze - zs[k] : 0.0;
                B[k + c * nt] = phs(Q, sqdist(dim, dx, dy, dz));
            }
            poly_basis(dim, P, xe, ye, ze, &B[n + c * nt], 1);
        }
    }
}

/* Partial-pivot LU solve in the order stated in the header. M is nt x nt
 * and B nt x r, both column-major; X (nt x r, column-major) receives the
 * solution. work holds nt*(nt+r) doubles + 3*nt ints' worth of space.
 * Returns info. */
/* LU factors and pivots in dgetrf layout, from the elimination of
 * oracle_gesv (rows never moved, first largest |a| among rows not yet
 * pivotal). Row p_k, pivot at step k, becomes row k of LU: its multipliers
 * of steps j < k (stored when it was eliminated) and its U entries j >= k.
 * ipiv[k] (1-based) is the position, in the order left by the swaps of
 * steps 0..k-1, of row p_k -- dgetrf's sequential row interchanges. */
OR_API int oracle_getrf(int nt, const double* M, double* LU, int* ipiv) {
    double* a = (double*)malloc(sizeof(double) * (size_t)nt * nt);
    int* done = (int*)malloc(sizeof(int) * 3 * (size_t)nt);
    int* step = done + nt;  /* step at which each row was pivot */
    int* pos = step + nt;   /* current position of each row */
    int info = 0;
    for (int i = 0; i < nt; ++i) {
        for (int j = 0; j < nt; ++j) a[i * nt + j] = M[i + j * nt];
        done[i] = 0;
        pos[i] = i;
    }
    for (int k = 0; k < nt; ++k) {
        double best = -1.0;
        int p = -1;
        for (int i = 0; i < nt; ++i) {
            if (done[i]) continue;
            const double v = fabs(a[i * nt + k]);
            if (v > best) {
                best = v;
                p = i;
            }
        }
        if (p < 0) {
            for (int i = 0; i < nt; ++i)
                if (!done[i]) { p = i; break; }
            if (!info) info = k + 1;
        }
        const double piv = a[p * nt + k];
        if (piv == 0.0 && !info) info = k + 1;
        const double rcp = 1.0 / piv;
        /* the swap of positions k and pos[p] */
        const int q = pos[p];
        ipiv[k] = q + 1;
        for (int i = 0; i < nt; ++i)
            if (pos[i] == k) pos[i] = q;
        pos[p] = k;
        done[p] = 1;
        step[p] = k;
        const double* up = &a[p * nt];
        for (int i = 0; i < nt; ++i) {
            if (done[i]) continue;
            double* ai = &a[i * nt];
            const double l = ai[k] * rcp;
            ai[k] = l;
            for (int j = k + 1; j < nt; ++j) ai[j] = fma(-l, up[j], ai[j]);
        }
    }
    for (int i = 0; i < nt; ++i)
        for (int j = 0; j < nt; ++j) LU[step[i] + j * nt] = a[i * nt + j];
    free(a);
    free(done);
    return info;
}

OR_API int oracle_gesv(int nt, int r, const double* M, const double* B, double* X) {
    const int nc = nt + r;
    double* aug = (double*)malloc(sizeof(double) * (size_t)nt * nc + sizeof(double) * nt);
    int* perm = (int*)malloc(sizeof(int) * 2 * (size_t)nt);
    double* rcp_of = aug + (size_t)nt * nc;
    int* done = perm + nt;
    int info = 0;
    for (int i = 0; i < nt; ++i) {
        for (int j = 0; j < nt; ++j) aug[i * nc + j] = M[i + j * nt];
        for (int c = 0; c < r; ++c) aug[i * nc + nt + c] = B[i + c * nt];
        done[i] = 0;
    }
    for (int k = 0; k < nt; ++k) {
        double best = -1.0;
        int p = -1;
        for (int i = 0; i < nt; ++i) {
            if (done[i]) continue;
            const double a = fabs(aug[i * nc + k]);
            if (a > best) {
                best = a;
                p = i;
            }
        }
        if (p < 0) { /* every candidate NaN: take the first remaining row */
            for (int i = 0; i < nt; ++i)
                if (!done[i]) { p = i; break; }
            if (!info) info = k + 1;
        }
        done[p] = 1;
        perm[k] = p;
        const double piv = aug[p * nc + k];
        if (piv == 0.0 && !info) info = k + 1;
        const double rcp = 1.0 / piv;
        rcp_of[p] = rcp;
        const double* up = &aug[p * nc];
        for (int i = 0; i < nt; ++i) {
            if (done[i]) continue;
            double* ai = &aug[i * nc];
            const double l = ai[k] * rcp;
            ai[k] = 0.0;
            for (int j = k + 1; j < nc; ++j) ai[j] = fma(-l, up[j], ai[j]);
        }
        aug[p * nc + k] = 0.0;
    }
    for (int k = nt - 1; k >= 0; --k) {
        const int p = perm[k];
        for (int c = 0; c < r; ++c) {
            const double t = aug[p * nc + nt + c] * rcp_of[p];
            for (int i = 0; i < nt; ++i) {
                double* ai = &aug[i * nc];
                ai[nt + c] = fma(-ai[k], t, ai[nt + c]);
            }
        }
    }
    for (int k = 0; k < nt; ++k) {
        const int p = perm[k];
        for (int c = 0; c < r; ++c) X[k + c * nt] = aug[p * nc + nt + c] * rcp_of[p];
    }
    free(aug);
    free(perm);
    return info;
}

/* Gather, shift to node s, scale by the stencil radius (rbf_operators.h:
 * 354-358 plus the scaling). Returns rho2 = max squared offset. */
static double gather_scaled(int dim, int n, int32_t s, const double* x, const double* y,
                            const double* z, const int32_t* row, double* xs, double* ys,
                            double* zs) {
    const double x0 = x[s], y0 = y[s], z0 = dim == 3 ? z[s] : 0.0;
    double rho2 = 0.0;
    for (int k = 0; k < n; ++k) {
        const int32_t j = row[k];
        xs[k] = x[j] - x0;
        ys[k] = y[j] - y0;
        zs[k] = dim == 3 ? z[j] - z0 : 0.0;
        const double d2 = sqdist(dim, xs[k], ys[k], zs[k]);
        if (d2 > rho2) rho2 = d2;
    }
    const double rho = rho2 > 0.0 ? sqrt(rho2) : 1.0;
    for (int k = 0; k < n; ++k) {
        xs[k] = xs[k] / rho;
        ys[k] = ys[k] / rho;
        zs[k] = zs[k] / rho;
    }
    return rho2 > 0.0 ? rho2 : 1.0;
}

OR_API int oracle_nops(int dim, int op_set, int nrhs) {
    return op_set == OR_OPS_DIFF ? dim + 1 : nrhs;
}

/* Whole-batch FP64 assembly: stencils s0 .. s0+count-1 of a fixed-width
 * graph (ja row s at s*n), weights to w[op][(s - s0)*n + k] and status to
 * info[s - s0]. nthreads <= 0: the OpenMP default. Returns the number of
 * stencils with info != 0. */
OR_API int oracle_assemble(int dim, int n, int P, int Q, int op_set, int nrhs, int32_t s0,
                           int32_t count, const double* x, const double* y, const double* z,
                           const int32_t* ja, const double* cx, const double* cy,
                           const double* cz, double* const* w, int32_t* info, int nthreads) {
    const int m = oracle_npoly(dim, P), nt = n + m;
    const int r = oracle_nops(dim, op_set, nrhs);
    int nbad = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads) reduction(+ : nbad)
#endif
    {
        double* buf = (double*)malloc(sizeof(double) * ((size_t)3 * n + (size_t)nt * nt +
                                                        (size_t)2 * nt * r + (size_t)3 * r));
        double *xs = buf, *ys = xs + n, *zs = ys + n, *M = zs + n, *B = M + (size_t)nt * nt;
        double *X = B + (size_t)nt * r, *ccx = X + (size_t)nt * r, *ccy = ccx + r,
               *ccz = ccy + r;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int32_t s = s0; s < s0 + count; ++s) {
            const int32_t* row = &ja[(size_t)s * n];
            const double rho2 = gather_scaled(dim, n, s, x, y, z, row, xs, ys, zs);
            const double rho = sqrt(rho2);
            oracle_fill(dim, n, P, Q, xs, ys, zs, M);
            if (op_set == OR_OPS_INTERP) {
                for (int c = 0; c < r; ++c) {
                    ccx[c] = cx[c] / rho;
                    ccy[c] = cy[c] / rho;
                    ccz[c] = dim == 3 ? cz[c] / rho : 0.0;
                }
            }
            oracle_rhs(dim, n, P, Q, op_set, nrhs, xs, ys, zs, ccx, ccy, ccz, B);
            const int inf = oracle_gesv(nt, r, M, B, X);
            if (info) info[s - s0] = inf;
            if (inf) ++nbad;
            for (int op = 0; op < r; ++op) {
                double* wo = w[op];
                for (int k = 0; k < n; ++k) {
                    double v = X[k + (size_t)op * nt];
                    if (op_set == OR_OPS_DIFF) v = op == 0 ? v / rho2 : v / rho;
                    wo[(size_t)(s - s0) * n + k] = v;
                }
            }
        }
        free(buf);
    }
    return nbad;
}

/* Ragged graph: the listed stencils rows[0..count) of size n (all of one
 * size class), row s at ia[s]; weights to w[op][ia[s] + k], info[s]. */
OR_API int oracle_assemble_rows(int dim, int n, int P, int Q, int32_t count,
                                const int32_t* rows, const int32_t* ia, const double* x,
                                const double* y, const double* z, const int32_t* ja,
                                double* const* w, int32_t* info, int nthreads) {
    const int m = oracle_npoly(dim, P), nt = n + m, r = dim + 1;
    int nbad = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads) reduction(+ : nbad)
#endif
    {
        double* buf = (double*)malloc(sizeof(double) * ((size_t)3 * n + (size_t)nt * nt +
                                                        (size_t)2 * nt * r));
        double *xs = buf, *ys = xs + n, *zs = ys + n, *M = zs + n, *B = M + (size_t)nt * nt;
        double* X = B + (size_t)nt * r;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int32_t i = 0; i < count; ++i) {
            const int32_t s = rows[i];
            const double rho2 = gather_scaled(dim, n, s, x, y, z, &ja[ia[s]], xs, ys, zs);
            const double rho = sqrt(rho2);
            oracle_fill(dim, n, P, Q, xs, ys, zs, M);
            oracle_rhs(dim, n, P, Q, OR_OPS_DIFF, 0, xs, ys, zs, NULL, NULL, NULL, B);
            const int inf = oracle_gesv(nt, r, M, B, X);
            if (info) info[s] = inf;
            if (inf) ++nbad;
            for (int op = 0; op < r; ++op)
                for (int k = 0; k < n; ++k) {
                    const double v = X[k + (size_t)op * nt];
                    w[op][(size_t)ia[s] + k] = op == 0 ? v / rho2 : v / rho;
                }
        }
        free(buf);
    }
    return nbad;
}

/* ---- curve keys (rbf_ordering.F90:82-226) ---------------------------- */

OR_API int64_t oracle_zidx(double px, double py, const double* bb, int ndiv) {
    double x0 = bb[0], y0 = bb[1], x1 = bb[2], y1 = bb[3];
    int64_t z = 0;
    for (int i = 1; i <= ndiv; ++i) {
        z = z << 2;
        const double cx = x0 + 0.5 * (x1 - x0);
        const double cy = y0 + 0.5 * (y1 - y0);
        if (px > cx) z = z + 1;
        if (py > cy) z = z + 2;
        if (px > cx) x0 = cx; else x1 = cx;
        if (py > cy) y0 = cy; else y1 = cy;
    }
    return z;
}

OR_API int64_t oracle_hidx(double px, double py, const double* bb, int ndiv) {
    /* table(q, rot) and tr(q, rot) of hidx, Fortran column-major data */
    static const int table[16] = {0, 3, 1, 2, 2, 3, 1, 0, 2, 1, 3, 0, 0, 1, 3, 2};
    static const int tr[16] = {3, 1, 0, 0, 1, 0, 1, 2, 2, 2, 3, 1, 0, 3, 2, 3};
    double x0 = bb[0], y0 = bb[1], x1 = bb[2], y1 = bb[3];
    int64_t h = 0;
    int rot = 0;
    for (int i = 1; i <= ndiv; ++i) {
        h = h << 2;
        const double cx = x0 + 0.5 * (x1 - x0);
        const double cy = y0 + 0.5 * (y1 - y0);
        int q = 0;
        if (px > cx) q = q + 1;
        if (py > cy) q = q + 2;
        h = h + table[q + 4 * rot];
        rot = tr[q + 4 * rot];
        if (px > cx) x0 = cx; else x1 = cx;
        if (py > cy) y0 = cy; else y1 = cy;
    }
    return h;
}

/* ---- brute-force kNN (test_spatial.cpp:70-101 pattern) --------------- */

typedef struct {
    double d2;
    int32_t i;
} cand_t;

static int cand_cmp(const void* a, const void* b) {
    const cand_t* p = (const cand_t*)a;
    const cand_t* q = (const cand_t*)b;
    if (p->d2 < q->d2) return -1;
    if (p->d2 > q->d2) return 1;
    return (p->i > q->i) - (p->i < q->i);
}

OR_API void oracle_knn_brute(int32_t n, int dim, const double* x, const double* y,
                             const double* z, int k, int32_t* ja) {
#ifdef _OPENMP
#pragma omp parallel
#endif
    {
        cand_t* c = (cand_t*)malloc(sizeof(cand_t) * (size_t)n);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int32_t s = 0; s < n; ++s) {
            for (int32_t i = 0; i < n; ++i) {
                const double dx = x[s] - x[i], dy = y[s] - y[i];
                double acc = 0.0;
                acc += dx * dx;
                acc += dy * dy;
                if (dim == 3) {
                    const double dz = z[s] - z[i];
                    acc += dz * dz;
                }
                c[i].d2 = acc;
                c[i].i = i;
            }
            qsort(c, (size_t)n, sizeof(cand_t), cand_cmp);
            for (int j = 0; j < k; ++j) ja[(size_t)s * k + j] = c[j].i;
        }
        free(c);
    }
}

OR_API int oracle_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
