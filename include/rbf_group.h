// rbf_group.h -- stencil systems larger than a wavefront (64 < nt <= 64*NW)
// for gfx950: one workgroup of NW wave64s per stencil, one system row per
// lane across the waves. Companion of rbf_operators.h (same fill, same
// elimination order, same results as the CPU oracle).
//
// Per elimination step every wave finds its own pivot candidate (wave max
// + ballot) and publishes, through LDS, the candidate's value, global row
// and the candidate row itself; one barrier later every wave picks the
// same winner (largest |a|, ties to the lower global row, as a serial scan
// would) and reads the winner's row back from LDS as the broadcast for its
// rank-1 update. The LDS buffers alternate by step parity, so a wave that
// runs ahead never overwrites data another wave still reads: one barrier
// per step. Back substitution broadcasts each solved unknown the same way.
#ifndef RBF_GROUP_H
#define RBF_GROUP_H

#include "rbf_wave.h"

namespace rbf_operators {
namespace detail {

template <int DIM, int N, int P, int Q, int R, bool INTERP, int NW>
struct GroupAssembly {
    static constexpr int M = npoly_dim(DIM, P);
    static constexpr int NT = N + M;
    static constexpr int NC = NT + R;
    static constexpr int NCP = (NC + 1) / 2 * 2;  // row buffer stride (16-byte multiple)
    static constexpr int nrhs = R;
    static constexpr int NCPF = (NC + 3) / 4 * 4;  // float row buffer stride
    static constexpr int nsten = N;  // stencil size
    static_assert(NT > 0 && NT <= 64 * NW, "one system row per lane of the workgroup");
    static_assert(N <= 64 * NW, "one stencil node per lane");
    static_assert(!INTERP, "interpolation operators use the single-wave engine");
    static_assert(R == DIM + 1, "derivative set: Laplacian + one per axis");

    // LDS layout (doubles): candidate rows [2][NW][NCP], candidate values
    // [2][NW][2] (|a| key, a), candidate rows' indices [2][NW] (as double),
    // solved unknowns [2][R], scaled points [DIM][N], monomials [M][N],
    // per-wave maxima [NW].
    static constexpr int off_rows = 0;
    static constexpr int off_cval = off_rows + 2 * NW * NCP;
    static constexpr int off_cidx = off_cval + 2 * NW * 2;
    static constexpr int off_x = off_cidx + 2 * NW;
    static constexpr int off_pts = off_x + 2 * R;
    static constexpr int off_poly = off_pts + DIM * N;
    static constexpr int off_wmax = off_poly + M * N;
    static constexpr int lds_doubles = (off_wmax + NW + 1) / 2 * 2;

    // Elimination step K of rows of type T held one per lane across the
    // workgroup (see the file comment). STORE_L keeps the multiplier in
    // row[K]; otherwise row[K] of every row still active becomes 0.
    template <int K, bool STORE_L, typename T>
    __device__ __forceinline__ static void elim_from(T (&row)[NC], ElimState<NT, T>& st,
                                                     double* lds, int g, int w) {
        if constexpr (K < NT) {
            constexpr int stride = sizeof(T) == 8 ? NCP : NCPF;
            const int lane = lane_id();
            const int par = K & 1;
            T* rows = reinterpret_cast<T*>(lds + off_rows) + par * NW * stride;
            double* cval = lds + off_cval + par * NW * 2;
            double* cidx = lds + off_cidx + par * NW;
            const bool active = g < NT && !st.done;
            const T a = row[K];
            int pl = pivot_lane<1>(a, active);
            double key = -1.0;
            if (pl >= 0) {
                key = static_cast<double>(fabs(bcast(a, pl)));
            } else {
                const unsigned long long act = __ballot(active);
                pl = act ? __ffsll(act) - 1 : -1;
            }
            if (pl >= 0) {
                pl = __builtin_amdgcn_readfirstlane(pl);
                if (lane == pl) {
                    publish_from<1, NC, K + 1, 0>(rows + w * stride, row);
                    cval[2 * w] = key;
                    cval[2 * w + 1] = static_cast<double>(a);
                    cidx[w] = static_cast<double>(w * 64 + pl);
                }
            } else if (lane == 0) {
                cval[2 * w] = -2.0;  // no active row in this wave
                cidx[w] = -1.0;
            }
            __syncthreads();
            // same winner in every wave: largest key, ties to the lower
            // global row (= lower wave); key -1 marks "first active row"
            int ww = -1;
            double wkey = -3.0;
#pragma unroll
            for (int v = 0; v < NW; ++v) {
                const double kv = cval[2 * v];
                if (kv > wkey) {
                    wkey = kv;
                    ww = v;
                }
            }
            ww = __builtin_amdgcn_readfirstlane(ww);
            const T piv = static_cast<T>(cval[2 * ww + 1]);
            // wave-uniform: keep it (and the NT-entry permutation) in SGPRs
            const int p = __builtin_amdgcn_readfirstlane(static_cast<int>(cidx[ww]));
            if ((wkey < 0.0 || piv == T(0)) && !st.info) st.info = K + 1;
            const T rcp = T(1) / piv;
            const bool is_piv = g == p;
            const bool elim = active && !is_piv;
            const T l = elim ? a * rcp : T(0);
            if constexpr (STORE_L) {
                row[K] = elim ? l : a;
            } else {
                row[K] = active ? T(0) : a;
            }
            if (is_piv) {
                st.done = true;
                st.mystep = K;
                st.myrcp = rcp;
            }
            asm volatile("" : "+v"(st.myrcp), "+v"(st.mystep));
            st.perm[K] = p;
            update_from<1, NC, K + 1, (K + 1) / 8 * 8>(row, l, 0, rows + ww * stride);
            elim_from<K + 1, STORE_L>(row, st, lds, g, w);
        }
    }

    // Back substitution, unknown K down to 0: the lane that pivoted at step
    // K publishes t = b * rcp, every lane updates with its row[K] (u_iK for
    // rows pivotal before K, 0 otherwise).
    template <int K>
    __device__ __forceinline__ static void back_from(double (&row)[NC],
                                                     const ElimState<NT, double>& st, double* xb,
                                                     int g) {
        if constexpr (K >= 0) {
            double* xk = xb + (K & 1) * R;
            if (g == st.perm[K]) {
#pragma unroll
                for (int c = 0; c < R; ++c) xk[c] = row[NT + c] * st.myrcp;
            }
            __syncthreads();
#pragma unroll
            for (int c = 0; c < R; ++c) row[NT + c] = fma(-row[K], xk[c], row[NT + c]);
            back_from<K - 1>(row, st, xb, g);
        }
    }

    // Gather, shift and scale stencil s (CSR row at base); this lane's node
    // (scaled) in xs/ys/zs, the points and transposed monomials in LDS.
    // Returns rho^2.
    __device__ __forceinline__ static double gather(const int s, const size_t base,
                                                    const double* __restrict__ x,
                                                    const double* __restrict__ y,
                                                    const double* __restrict__ z,
                                                    const int* __restrict__ ja, double* lds,
                                                    int g, int w, double& xs, double& ys,
                                                    double& zs) {
#pragma clang fp contract(off)
        const int lane = lane_id();
        const int* ja_row = ja + base;
        double* pts = lds + off_pts;
        double* polyT = lds + off_poly;
        const int jn = g < N ? ja_row[g] : s;
        const double px = x[jn] - x[s];
        const double py = y[jn] - y[s];
        const double pz = DIM == 3 ? z[jn] - z[s] : 0.0;
        double d2 = px * px + py * py;
        if constexpr (DIM == 3) d2 = d2 + pz * pz;
        const double wm = wave_max(g < N ? d2 : 0.0);
        if (lane == 0) lds[off_wmax + w] = wm;
        __syncthreads();
        double rho2 = 0.0;
#pragma unroll
        for (int v = 0; v < NW; ++v) rho2 = fmax(rho2, lds[off_wmax + v]);
        if (!(rho2 > 0.0)) rho2 = 1.0;
        const double rho = sqrt(rho2);
        xs = px / rho;
        ys = py / rho;
        zs = pz / rho;
        constexpr PolyBasis<P, DIM> basis{};
        double mono[M];
        basis(xs, ys, zs, mono, 1);
        if (g < N) {
            pts[g] = xs;
            pts[N + g] = ys;
            if constexpr (DIM == 3) pts[2 * N + g] = zs;
#pragma unroll
            for (int t = 0; t < M; ++t) polyT[t * N + g] = mono[t];
        }
        __syncthreads();
        return rho2;
    }

    // Matrix entry (g, col) of this lane's row, col < NT, in FP64.
    __device__ __forceinline__ static double entry(int g, int col, double xs, double ys,
                                                   double zs, const double* pts,
                                                   const double* polyT, const double* mono) {
#pragma clang fp contract(off)
        constexpr PHS<Q> phi{};
        if (g < N) {
            if (col < N) {
                const double dx = pts[col] - xs;
                const double dy = pts[N + col] - ys;
                double r2 = dx * dx + dy * dy;
                if constexpr (DIM == 3) {
                    const double dz = pts[2 * N + col] - zs;
                    r2 = r2 + dz * dz;
                }
                return phi(r2);
            }
            return mono[col - N];
        }
        if (g < NT) return col < N ? polyT[(g - N) * N + col] : 0.0;
        return 0.0;
    }

    // Right-hand side c of this lane's row, in FP64.
    __device__ __forceinline__ static double rhs(int g, int c, double xs, double ys, double zs) {
#pragma clang fp contract(off)
        if (g < N) {
            double r2 = xs * xs + ys * ys;
            if constexpr (DIM == 3) r2 = r2 + zs * zs;
            const double tq = PHS<Q - 2>{}(r2);
            const double gq = static_cast<double>(Q) * tq;
            if (c == 0) return static_cast<double>(Q * (Q + DIM - 2)) * tq;
            return -gq * (c == 1 ? xs : (c == 2 ? ys : zs));
        }
        if (g < NT) {
            using Mk = PolyDerivMasks<DIM, P>;
            const int t = g - N;
            const uint64_t mk = c == 0 ? Mk::lap : (c == 1 ? Mk::dx : (c == 2 ? Mk::dy : Mk::dz));
            return ((mk >> t) & 1u) ? (c == 0 ? 2.0 : 1.0) : 0.0;
        }
        return 0.0;
    }

    // This lane's row of [M | B], each entry computed in FP64 and stored as T.
    template <typename T>
    __device__ __forceinline__ static void fill(T (&row)[NC], int g, double xs, double ys,
                                                double zs, const double* lds) {
        const double* pts = lds + off_pts;
        const double* polyT = lds + off_poly;
        constexpr PolyBasis<P, DIM> basis{};
        double mono[M];
        basis(xs, ys, zs, mono, 1);
#pragma unroll
        for (int col = 0; col < NT; ++col) {
            row[col] = static_cast<T>(entry(g, col, xs, ys, zs, pts, polyT, mono));
            if ((col & 3) == 3) __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int c = 0; c < R; ++c) row[NT + c] = static_cast<T>(rhs(g, c, xs, ys, zs));
    }

    // Stencil s, whose CSR row starts at base (s*N for a fixed-width graph).
    template <class Out>
    __device__ __forceinline__ static void run(const int s, const size_t base,
                                               const double* __restrict__ x,
                                               const double* __restrict__ y,
                                               const double* __restrict__ z,
                                               const int* __restrict__ ja, Out out, double* lds,
                                               int* __restrict__ info) {
#pragma clang fp contract(off)
        const int w = static_cast<int>(threadIdx.x >> 6);
        const int g = w * 64 + lane_id();  // this lane's row / stencil node
        double xs, ys, zs;
        const double rho2 = gather(s, base, x, y, z, ja, lds, g, w, xs, ys, zs);
        const double rho = sqrt(rho2);
        double row[NC];
        fill(row, g, xs, ys, zs, lds);
        ElimState<NT, double> st;
        elim_from<0, false>(row, st, lds, g, w);
        back_from<NT - 1>(row, st, lds + off_x, g);
        if (st.mystep < N) {
            const size_t at = base + st.mystep;
#pragma unroll
            for (int c = 0; c < R; ++c) {
                const double v = row[NT + c] * st.myrcp;
                out(c)[at] = c == 0 ? v / rho2 : v / rho;
            }
        }
        if (info && g == 0) info[s] = st.info;
    }
};

// Mixed precision for systems larger than 64: FP32 rows and factors in
// registers, FP64 residuals computed column by column from the stencil
// points in LDS (no FP64 copy of the matrix is kept), FP32 triangular
// solves with the stored factors, FP64 iterate. A stencil that does not
// converge within max_refine steps is appended to a fallback list for the
// FP64 engine (GroupAssembly) to solve again in a second launch.
template <int DIM, int N, int P, int Q, int NW>
struct GroupMixed {
    using Base = GroupAssembly<DIM, N, P, Q, DIM + 1, false, NW>;
    static constexpr int M = Base::M, NT = Base::NT, NC = Base::NC, R = DIM + 1;
    static constexpr int nsten = N, nrhs = R;
    // LDS: the FP64 engine's layout, then the iterate [R][NT], the
    // triangular-solve broadcast slots [2][R] and the convergence maxima
    // [2][NW][2R].
    static constexpr int off_xl = Base::lds_doubles;
    static constexpr int off_bc = off_xl + R * NT;
    static constexpr int off_red = off_bc + 2 * R;
    static constexpr int off_guard = off_red + 2 * NW * 2 * R;  // [2][NW]: max |entry|, tiny pivot
    static constexpr int lds_doubles = (off_guard + 2 * NW + 1) / 2 * 2;
    static constexpr double kTol = 1e-10;

    // forward substitution with L (unit lower; rows pivotal after step K)
    template <int K>
    __device__ __forceinline__ static void fwd_from(const float (&a)[NC],
                                                    const ElimState<NT, float>& st, float (&r)[R],
                                                    double* bc, int g) {
        if constexpr (K < NT) {
            double* slot = bc + (K & 1) * R;
            if (g == st.perm[K]) {
#pragma unroll
                for (int c = 0; c < R; ++c) slot[c] = static_cast<double>(r[c]);
            }
            __syncthreads();
            const float lk = st.mystep > K ? a[K] : 0.f;
#pragma unroll
            for (int c = 0; c < R; ++c) r[c] = fmaf(-lk, static_cast<float>(slot[c]), r[c]);
            fwd_from<K + 1>(a, st, r, bc, g);
        }
    }
    // back substitution with U (rows pivotal before step K); slot parity
    // offset by one from the forward sweep so the two never share a slot
    // across the sweep boundary
    template <int K>
    __device__ __forceinline__ static void bwd_from(const float (&a)[NC],
                                                    const ElimState<NT, float>& st, float (&r)[R],
                                                    double* bc, int g) {
        if constexpr (K >= 0) {
            double* slot = bc + ((K + 1) & 1) * R;
            if (g == st.perm[K]) {
#pragma unroll
                for (int c = 0; c < R; ++c) slot[c] = static_cast<double>(r[c] * st.myrcp);
            }
            __syncthreads();
            const float uk = st.mystep < K ? a[K] : 0.f;
#pragma unroll
            for (int c = 0; c < R; ++c) r[c] = fmaf(-uk, static_cast<float>(slot[c]), r[c]);
            bwd_from<K - 1>(a, st, r, bc, g);
        }
    }

    // A stencil that converges gets its weights here; one that does not is
    // appended to fb_list (its weights are left for the FP64 fallback).
    template <class Out>
    __device__ __forceinline__ static void run(const int s, const size_t base,
                                               const double* __restrict__ x,
                                               const double* __restrict__ y,
                                               const double* __restrict__ z,
                                               const int* __restrict__ ja, Out out, double* lds,
                                               int* __restrict__ info, uint8_t* __restrict__ status,
                                               int max_refine, int* __restrict__ fb_count,
                                               int* __restrict__ fb_list) {
#pragma clang fp contract(off)
        const int lane = lane_id();
        const int w = static_cast<int>(threadIdx.x >> 6);
        const int g = w * 64 + lane;
        double xs, ys, zs;
        const double rho2 = Base::gather(s, base, x, y, z, ja, lds, g, w, xs, ys, zs);
        const double rho = sqrt(rho2);
        float a[NC];
        Base::fill(a, g, xs, ys, zs, lds);
        float rowmax = 0.f;
        if (g < NT) {
#pragma unroll
            for (int j = 0; j < NT; ++j) rowmax = fmaxf(rowmax, fabsf(a[j]));
        }
        ElimState<NT, float> st;
        Base::template elim_from<0, true>(a, st, lds, g, w);
        // a pivot below 1e-6 of the largest matrix entry (duplicated nodes,
        // nearly singular stencils): FP32 factors that small carry no
        // information, and refinement can appear to converge on them (as
        // in the one-wave whole-system form); the workgroup's verdict
        // through LDS
        double* guard = lds + off_guard;
        bool tiny_pivot = false;
        {
            const unsigned wm = wave_umax(__float_as_uint(rowmax));
            if (lane == 0) guard[w] = static_cast<double>(__uint_as_float(wm));
            __syncthreads();
            float smax = 0.f;
#pragma unroll
            for (int v = 0; v < NW; ++v) smax = fmaxf(smax, static_cast<float>(guard[v]));
            const bool tiny = g < NT && !(fabsf(st.myrcp) * (1e-6f * smax) < 1.f);
            if (lane == 0) guard[NW + w] = wave_ballot(tiny) ? 1.0 : 0.0;
            __syncthreads();
#pragma unroll
            for (int v = 0; v < NW; ++v) tiny_pivot = tiny_pivot || guard[NW + v] != 0.0;
        }
        double* bc = lds + off_bc;
        double* xl = lds + off_xl;
        float r32[R];
#pragma unroll
        for (int c = 0; c < R; ++c) r32[c] = a[NT + c];
        bwd_from<NT - 1>(a, st, r32, bc, g);
        double xd[R];
#pragma unroll
        for (int c = 0; c < R; ++c) xd[c] = static_cast<double>(r32[c] * st.myrcp);

        const bool is_weight = st.mystep < N;
        const double* pts = lds + Base::off_pts;
        const double* polyT = lds + Base::off_poly;
        constexpr PolyBasis<P, DIM> basis{};
        double mono[M];
        basis(xs, ys, zs, mono, 1);
        int steps = 0;
        bool converged = false;
        float dprev[R];
#pragma unroll
        for (int c = 0; c < R; ++c) dprev[c] = 0.f;
        if (!st.info && !tiny_pivot) {
            for (int it = 0; it < max_refine; ++it) {
                if (st.mystep < NT) {
#pragma unroll
                    for (int c = 0; c < R; ++c) xl[c * NT + st.mystep] = xd[c];
                }
                __syncthreads();
                double rr[R];
#pragma unroll
                for (int c = 0; c < R; ++c) rr[c] = Base::rhs(g, c, xs, ys, zs);
                for (int col = 0; col < N; ++col) {
                    const double m = Base::entry(g, col, xs, ys, zs, pts, polyT, mono);
#pragma unroll
                    for (int c = 0; c < R; ++c) rr[c] = fma(-m, xl[c * NT + col], rr[c]);
                }
#pragma unroll
                for (int t = 0; t < M; ++t) {
                    const double m = g < N ? mono[t] : 0.0;  // P block; zero block below
#pragma unroll
                    for (int c = 0; c < R; ++c) rr[c] = fma(-m, xl[c * NT + N + t], rr[c]);
                }
                __syncthreads();
#pragma unroll
                for (int c = 0; c < R; ++c) r32[c] = static_cast<float>(rr[c]);
                fwd_from<0>(a, st, r32, bc, g);
                bwd_from<NT - 1>(a, st, r32, bc, g);
                ++steps;
                // workgroup maxima of |d| and |x| over the weight unknowns
                double* red = lds + off_red + (it & 1) * NW * 2 * R;
#pragma unroll
                for (int c = 0; c < R; ++c) {
                    const float d = r32[c] * st.myrcp;
                    xd[c] = xd[c] + static_cast<double>(d);
                    const unsigned dm = wave_umax(__float_as_uint(is_weight ? fabsf(d) : 0.f));
                    const unsigned xm = wave_umax(
                        __float_as_uint(is_weight ? static_cast<float>(fabs(xd[c])) : 0.f));
                    if (lane == 0) {
                        red[w * 2 * R + 2 * c] = static_cast<double>(__uint_as_float(dm));
                        red[w * 2 * R + 2 * c + 1] = static_cast<double>(__uint_as_float(xm));
                    }
                }
                __syncthreads();
                bool ok = true;
#pragma unroll
                for (int c = 0; c < R; ++c) {
                    double dm = 0.0, xm = 0.0;
#pragma unroll
                    for (int v = 0; v < NW; ++v) {
                        dm = fmax(dm, red[v * 2 * R + 2 * c]);
                        xm = fmax(xm, red[v * 2 * R + 2 * c + 1]);
                    }
                    const double ratio = (it > 0 && dprev[c] > 0.f)
                                             ? fmin(1.0, dm / static_cast<double>(dprev[c]))
                                             : 1.0;
                    if (!(dm * ratio <= kTol * xm)) ok = false;
                    dprev[c] = static_cast<float>(dm);
                }
                if (ok) {
                    converged = true;
                    break;
                }
            }
        }
        if (converged) {
            if (st.mystep < N) {
                const size_t at = base + st.mystep;
#pragma unroll
                for (int c = 0; c < R; ++c) out(c)[at] = c == 0 ? xd[c] / rho2 : xd[c] / rho;
            }
            if (g == 0) {
                if (info) info[s] = 0;
                if (status) status[s] = static_cast<uint8_t>(steps & 3);
            }
        } else if (g == 0) {
            fb_list[atomicAdd(fb_count, 1)] = s;
            if (status) status[s] = static_cast<uint8_t>((steps & 3) | 4);
        }
    }
};

}  // namespace detail
}  // namespace rbf_operators

#endif  // RBF_GROUP_H
