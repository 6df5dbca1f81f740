// rbf_mixed.h -- mixed-precision stencil solve for gfx950: FP32 LU and
// triangular solves, FP64 residuals and iterative refinement, FP64
// fallback. Companion of rbf_operators.h (same wave64 row-per-lane layout).
//
// Per stencil (one wave):
//   1. build the FP64 row of [M | B] (gather, shift, scale, fill: the same
//      code and rounding as the FP64 path) and keep it in registers;
//   2. round to FP32, factor with partial pivoting keeping the multipliers
//      (L) in place, and solve for x0;
//   3. up to max_refine times: r = B - M x in FP64 (x broadcast through
//      LDS), d = U^-1 L^-1 r in FP32, x += d in FP64; stop when the
//      estimated remaining error of the weights, |d| min(1, |d|/|d_prev|),
//      is below 1e-10 of |x| for every operator;
//   4. a stencil that has not converged (or met an FP32 pivot below 1e-6
//      of the largest matrix entry) is
//      solved again by the FP64 path from the row kept in step 1 and
//      flagged RBFA_STAT_FALLBACK in its status byte.
// The polynomial block keeps FP64 where it decides anything: its entries,
// the residual and the fallback are FP64; only the factors are FP32.
#ifndef RBF_MIXED_H
#define RBF_MIXED_H

#include "rbf_wave.h"

namespace rbf_operators {
namespace detail {

template <int DIM, int N, int P, int Q, class Tune = DefaultTune>
struct WaveMixed {
    using Base = WaveAssembly<DIM, N, P, Q, DIM + 1, false, Tune>;
    using tune = Tune;
    static constexpr int M = Base::M, NT = Base::NT, NC = Base::NC, R = DIM + 1;
    static constexpr int nsten = N;  // stencil size
    // Base scratch + the iterate x, [R][64] doubles, for the residual.
    static constexpr int lds_doubles = Base::lds_doubles + R * 64;
    static constexpr double kTol = 1e-10;

    // FP32 LU of a32 in place: multipliers stay below the pivots. Returns
    // nonzero if a pivot was exactly zero (or every candidate NaN).
    __device__ __forceinline__ static int lu32(float (&a)[NC], int (&perm)[NT], int& mystep, float& myrcp,
                               float* pb) {
        ElimState<NT, float> st;
        elim_from<0, NT, NC, Tune, true>(a, st, pb, lane_id());
#pragma unroll
        for (int k = 0; k < NT; ++k) perm[k] = st.perm[k];
        mystep = st.mystep;
        myrcp = st.myrcp;
        return st.info;
    }

    // Solve L U d = r for the R columns held one equation per lane;
    // on return lane i holds the component of unknown mystep.
    __device__ __forceinline__ static void solve32(const float (&a)[NC], const int (&perm)[NT], int mystep,
                                   float myrcp, float (&r)[R]) {
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            const int p = perm[k];
            const float lk = mystep > k ? a[k] : 0.f;
#pragma unroll
            for (int c = 0; c < R; ++c) r[c] = fmaf(-lk, bcast(r[c], p), r[c]);
        }
#pragma unroll
        for (int k = NT - 1; k >= 0; --k) {
            const int p = perm[k];
            const float uk = mystep < k ? a[k] : 0.f;
#pragma unroll
            for (int c = 0; c < R; ++c) r[c] = fmaf(-uk, bcast(r[c] * myrcp, p), r[c]);
        }
#pragma unroll
        for (int c = 0; c < R; ++c) r[c] = r[c] * myrcp;
    }

    // Stencil s, whose CSR row starts at base (s*N for a fixed-width graph).
    template <class Out>
    __device__ __forceinline__ static void run(const int s, const size_t base,
                               const double* __restrict__ x,
                               const double* __restrict__ y, const double* __restrict__ z,
                               const int* __restrict__ ja, Out out, double* lds,
                               int* __restrict__ info, uint8_t* __restrict__ status,
                               int max_refine) {
#pragma clang fp contract(off)
        const int lane = lane_id();
        const int* ja_row = ja + base;
        double row[NC];
        double rho2;
        Base::template build_row<double>(s, x, y, z, ja_row, nullptr, nullptr, nullptr, lds, row,
                                         rho2);
        double* xl = lds + Base::lds_doubles;  // [R][64]

        float a[NC];
        float rowmax = 0.f;
#pragma unroll
        for (int j = 0; j < NC; ++j) {
            a[j] = static_cast<float>(row[j]);
            if (j < NT) rowmax = fmaxf(rowmax, fabsf(a[j]));
        }
        int perm[NT];
        int mystep;
        float myrcp;
        int bad = lu32(a, perm, mystep, myrcp, reinterpret_cast<float*>(lds));
        // a pivot below 1e-6 of the largest matrix entry (duplicated nodes,
        // nearly singular stencils): FP32 factors that small carry no
        // information, and refinement can appear to converge on them
        const float smax = __uint_as_float(wave_umax(__float_as_uint(rowmax)));
        if (wave_ballot(lane < NT && !(fabsf(myrcp) * (1e-6f * smax) < 1.f))) bad = 1;

        // initial solve from the eliminated right-hand sides
        float r32[R];
#pragma unroll
        for (int c = 0; c < R; ++c) r32[c] = a[NT + c];
#pragma unroll
        for (int k = NT - 1; k >= 0; --k) {
            const int p = perm[k];
            const float uk = mystep < k ? a[k] : 0.f;
#pragma unroll
            for (int c = 0; c < R; ++c) r32[c] = fmaf(-uk, bcast(r32[c] * myrcp, p), r32[c]);
        }
        double xd[R];
#pragma unroll
        for (int c = 0; c < R; ++c) xd[c] = static_cast<double>(r32[c] * myrcp);

        const bool is_weight = mystep < N;
        int steps = 0;
        bool converged = false;
        float dprev[R];
#pragma unroll
        for (int c = 0; c < R; ++c) dprev[c] = 0.f;
        if (!bad) {
            for (int it = 0; it < max_refine; ++it) {
                if (mystep < NT) {
#pragma unroll
                    for (int c = 0; c < R; ++c) xl[c * 64 + mystep] = xd[c];
                }
                wave_lds_sync();
                double rr[R];
#pragma unroll
                for (int c = 0; c < R; ++c) rr[c] = row[NT + c];
#pragma unroll
                for (int j = 0; j < NT; ++j) {
#pragma unroll
                    for (int c = 0; c < R; ++c) rr[c] = fma(-row[j], xl[c * 64 + j], rr[c]);
                }
                wave_lds_sync();
#pragma unroll
                for (int c = 0; c < R; ++c) r32[c] = static_cast<float>(rr[c]);
                solve32(a, perm, mystep, myrcp, r32);
                ++steps;
                bool ok = true;
#pragma unroll
                for (int c = 0; c < R; ++c) {
                    xd[c] = xd[c] + static_cast<double>(r32[c]);
                    const float dm = __uint_as_float(
                        wave_umax(__float_as_uint(is_weight ? fabsf(r32[c]) : 0.f)));
                    const float xm = __uint_as_float(wave_umax(
                        __float_as_uint(is_weight ? static_cast<float>(fabs(xd[c])) : 0.f)));
                    const float ratio = (it > 0 && dprev[c] > 0.f) ? fminf(1.f, dm / dprev[c]) : 1.f;
                    if (!(dm * ratio <= static_cast<float>(kTol) * xm)) ok = false;
                    dprev[c] = dm;
                }
                if (ok) {
                    converged = true;
                    break;
                }
            }
        }

        int inf = 0;
        const bool fallback = !converged;
        if (fallback) {  // FP64 solve of the kept FP64 row
            double xo[R];
            int ms;
            inf = Base::lu_solve(row, ms, xo, lds);
            mystep = ms;
#pragma unroll
            for (int c = 0; c < R; ++c) xd[c] = xo[c];
        }
        if (mystep < N) {
            const double rho = sqrt(rho2);
            const size_t at = base + mystep;
#pragma unroll
            for (int c = 0; c < R; ++c) {
                const double v = c == 0 ? xd[c] / rho2 : xd[c] / rho;
                out(c)[at] = v;
            }
        }
        if (lane == 0) {
            if (info) info[s] = inf;
            if (status) status[s] = static_cast<uint8_t>((steps & 3) | (fallback ? 4 : 0));
        }
    }
};

}  // namespace detail
}  // namespace rbf_operators

#endif  // RBF_MIXED_H
