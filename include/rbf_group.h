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

#include "rbf_operators.h"

namespace rbf_operators {
namespace detail {

template <int DIM, int N, int P, int Q, int R, bool INTERP, int NW>
struct GroupAssembly {
    static constexpr int M = npoly_dim(DIM, P);
    static constexpr int NT = N + M;
    static constexpr int NC = NT + R;
    static constexpr int NCP = (NC + 1) / 2 * 2;  // row buffer stride (16-byte multiple)
    static constexpr int nrhs = R;
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

    template <int K>
    __device__ __forceinline__ static void elim_from(double (&row)[NC], ElimState<NT, double>& st,
                                                     double* lds, int g, int w) {
        if constexpr (K < NT) {
            const int lane = lane_id();
            const int par = K & 1;
            double* rows = lds + off_rows + par * NW * NCP;
            double* cval = lds + off_cval + par * NW * 2;
            double* cidx = lds + off_cidx + par * NW;
            const bool active = g < NT && !st.done;
            const double a = row[K];
            int pl = pivot_lane<1>(a, active);
            double key = -1.0;
            if (pl >= 0) {
                key = fabs(bcast(a, pl));
            } else {
                const unsigned long long act = __ballot(active);
                pl = act ? __ffsll(act) - 1 : -1;
            }
            if (pl >= 0) {
                pl = __builtin_amdgcn_readfirstlane(pl);
                if (lane == pl) {
                    publish_from<1, NC, K + 1, 0>(rows + w * NCP, row);
                    cval[2 * w] = key;
                    cval[2 * w + 1] = a;
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
            const double piv = cval[2 * ww + 1];
            // wave-uniform: keep it (and the NT-entry permutation) in SGPRs
            const int p = __builtin_amdgcn_readfirstlane(static_cast<int>(cidx[ww]));
            if ((wkey < 0.0 || piv == 0.0) && !st.info) st.info = K + 1;
            const double rcp = 1.0 / piv;
            const bool is_piv = g == p;
            const bool elim = active && !is_piv;
            const double l = elim ? a * rcp : 0.0;
            row[K] = active ? 0.0 : a;
            if (is_piv) {
                st.done = true;
                st.mystep = K;
                st.myrcp = rcp;
            }
            asm volatile("" : "+v"(st.myrcp), "+v"(st.mystep));
            st.perm[K] = p;
            update_from<1, NC, K + 1, (K + 1) / 8 * 8>(row, l, 0, rows + ww * NCP);
            elim_from<K + 1>(row, st, lds, g, w);
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

    // Stencil s, whose CSR row starts at base (s*N for a fixed-width graph).
    template <class Out>
    __device__ __forceinline__ static void run(const int s, const size_t base,
                                               const double* __restrict__ x,
                                               const double* __restrict__ y,
                                               const double* __restrict__ z,
                                               const int* __restrict__ ja, Out out, double* lds,
                                               int* __restrict__ info) {
#pragma clang fp contract(off)
        const int lane = lane_id();
        const int w = static_cast<int>(threadIdx.x >> 6);
        const int g = w * 64 + lane;  // this lane's row / stencil node
        const int* ja_row = ja + base;
        double* pts = lds + off_pts;
        double* polyT = lds + off_poly;

        // gather, shift, radius over the whole stencil
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
        const double xs = px / rho, ys = py / rho, zs = pz / rho;
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

        double row[NC];
        constexpr PHS<Q> phi{};
        if (g < N) {
#pragma unroll
            for (int col = 0; col < N; ++col) {
                const double dx = pts[col] - xs;
                const double dy = pts[N + col] - ys;
                double r2 = dx * dx + dy * dy;
                if constexpr (DIM == 3) {
                    const double dz = pts[2 * N + col] - zs;
                    r2 = r2 + dz * dz;
                }
                row[col] = phi(r2);
                if ((col & 3) == 3) __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll
            for (int t = 0; t < M; ++t) row[N + t] = mono[t];
            double r2 = xs * xs + ys * ys;
            if constexpr (DIM == 3) r2 = r2 + zs * zs;
            const double tq = PHS<Q - 2>{}(r2);
            const double gq = static_cast<double>(Q) * tq;
            row[NT + 0] = static_cast<double>(Q * (Q + DIM - 2)) * tq;
            row[NT + 1] = -gq * xs;
            row[NT + 2] = -gq * ys;
            if constexpr (DIM == 3) row[NT + 3] = -gq * zs;
        } else if (g < NT) {
            const int t = g - N;
#pragma unroll
            for (int col = 0; col < N; ++col) row[col] = polyT[t * N + col];
#pragma unroll
            for (int c = N; c < NT; ++c) row[c] = 0.0;
            using Mk = PolyDerivMasks<DIM, P>;
            row[NT + 0] = ((Mk::lap >> t) & 1u) ? 2.0 : 0.0;
            row[NT + 1] = ((Mk::dx >> t) & 1u) ? 1.0 : 0.0;
            row[NT + 2] = ((Mk::dy >> t) & 1u) ? 1.0 : 0.0;
            if constexpr (DIM == 3) row[NT + 3] = ((Mk::dz >> t) & 1u) ? 1.0 : 0.0;
        } else {
#pragma unroll
            for (int c = 0; c < NC; ++c) row[c] = 0.0;
        }

        ElimState<NT, double> st;
        elim_from<0>(row, st, lds, g, w);

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

}  // namespace detail
}  // namespace rbf_operators

#endif  // RBF_GROUP_H
