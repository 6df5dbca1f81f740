// rbf_nullspace.h -- mixed-precision stencil solve on the null space of the
// polynomial constraints (gfx950, one wave64 per stencil).
//
// The saddle-point system of a stencil (scaled frame, n nodes, m monomials)
//
//     [ A   P ] [w]   [b]        A = phi(|q_i - q_j|)    n x n
//     [ P^T 0 ] [l] = [g]        P = monomials at q_i    n x m
//
// is ill-conditioned as a whole (condition numbers 2e5 at n = 35, r^5 + P3,
// 2e7 at n = 56, r^7 + P4), but almost all of that comes from the
// polynomial block. Restricted to the null space of P^T the kernel block is
// well conditioned (3e3 and 4e4 for the same stencils), so FP32 factors of
// the restriction converge under FP64 refinement in one step where FP32
// factors of the whole system take two or three, or never converge.
//
// Per stencil:
//   A  FP64 LU with partial pivoting of P (the polynomial block's pivoting):
//      P_pi = L1 U, pivot nodes pi (m of them), free nodes F (the other
//      f = n - m, in lane order); G = L2 L1^-1 (f x m) so that P_F = G P_pi.
//      Z = [-G^T ; I] (rows pi ; F) spans the null space of P^T.
//   B  Schur complement S = Z^T A Z = A_FF - G H^T - H G^T (f x f, FP64)
//      with H = A_Fpi - G A_pipi / 2: one FP64 MFMA product of inner
//      dimension 2m, [G | H] times [H | G]^T, accumulated onto A_FF.
//   D  right-hand sides: a particular solution w_p = [P_pi^-T g ; 0] of
//      the constraints, t = b - A_:pi w_p, y = t_F - G t_pi (f x R, FP64).
//   C  S rounded to FP32 and factored without pivoting (S is definite for
//      the shipped kernel / degree pairs: r^Q with polynomial degree at least
//      (Q - 1) / 2 makes Z^T A Z (-1)^((Q+1)/2)-definite), row k on lane k,
//      the pivot row taken from the published column (S stays symmetric);
//      v0 = S^-1 y in FP32.
//   E  FP64 iterative refinement on the scaled reduced system:
//      r = y - S v (S and y in FP64), v += S32^-1 r, until the estimated
//      remaining error of the weights is below tol of their size.
//   F  w_F = v, w_pi = w_p - G^T v; unscaled and scattered into CSR.
// A stencil whose refinement does not converge, whose FP32 factors meet a
// pivot below 1e-6 of the largest entry of S (duplicated nodes), or whose
// polynomial block is rank deficient is solved again from scratch in FP64
// and flagged RBFA_STAT_FALLBACK; its weights and info are then those of
// the FP64 form. Systems of up to 64 unknowns take the wave LU in the same
// wave; larger ones are listed (fb_count, fb_list) for the blocked LU.
//
// Stencils of more than 64 nodes (up to 128; at most 64 free nodes) keep
// two nodes per lane, node l and node 64 + l, through the node-wise phases
// (gather, A, the kernel entries against the pivot nodes, G, t, y); every
// phase from the Schur complement on works on free ranks, one per lane.
#ifndef RBF_NULLSPACE_H
#define RBF_NULLSPACE_H

#include "rbf_wave.h"

namespace rbf_operators {
namespace detail {

typedef double ns_dbl4 __attribute__((ext_vector_type(4)));

// Knob of the reduced FP32 LU: its first SYM elimination steps take the
// pivot row from the published column (symmetry of S, LDS), the rest by
// v_readlane from the pivot lane.
template <int SYM_>
struct NsTune {
    static constexpr int SYM = SYM_;
};

template <int DIM, int N, int P, int Q, class Tune = NsTune<0>>
struct NullSpaceMixed {
    static constexpr int M = npoly_dim(DIM, P);
    static constexpr int F = N - M;  // free nodes: order of the reduced system
    static constexpr int R = DIM + 1;
    static constexpr int NT = N + M;
    static constexpr int NC = F + R;  // FP32 row: S plus the right-hand sides
    static constexpr int nsten = N;
    static constexpr int T = (F + 15) / 16;  // 16 x 16 tiles per side of S
    static constexpr int TP = 16 * T;        // padded order of S
    static constexpr int NTILE = T * (T + 1) / 2;  // upper-triangle tiles (S is symmetric)
    static constexpr int K2 = 2 * M;  // inner dimension of the Schur update
    static constexpr int KS = (K2 + 3) / 4;  // MFMA k-steps
    static constexpr int NSETS = N > 64 ? 2 : 1;  // nodes per lane
    static_assert(N <= 128 && F >= 1 && F <= 64, "two nodes per lane at most, one free rank per lane");
    static_assert(Q >= 3 && Q % 2 == 1, "odd PHS exponent >= 3");
    // Fallback: systems of up to 64 unknowns are solved again in the same
    // wave by the FP64 wave LU; larger ones are appended to a list that a
    // second launch solves with the FP64 blocked LU.
    static constexpr bool kInWaveFallback = NT <= 64;
    using Fb = WaveAssembly<DIM, N, P, Q, R, false, LuTune<4, 2, 2, 1, 1>>;
    static constexpr int fb_lds = [] {
        if constexpr (kInWaveFallback) return Fb::lds_doubles;
        else return 0;
    }();
    using tune = Tune;
    // estimated remaining error of the weights, relative to their size, at
    // which refinement stops (the weights are held to 1e-8 of the FP64
    // solution; larger systems keep more margin)
    static constexpr double kTol = NT > 48 ? 1e-10 : 3e-10;

    // LDS (doubles). Region 0 lives through the whole stencil and ends with
    // the f rows of G; region 1 follows it. In phases A-D region 1 holds the
    // pivot rows of P's LU, A_pipi, t_pi and H; the padding rows of G
    // (ranks f .. TP-1, zero, read only by the MFMA products) take the
    // place of P's pivot rows once the particular solution has used them;
    // after the Schur product region 1 holds S (FP64): row-major, or, for
    // the systems that list their fallbacks (where the LDS of a wave sets
    // the waves per CU), its upper triangle packed by rows.
    static constexpr int GS = M + 1;  // row stride of G and H (spreads the banks)
    static constexpr int SS = F;      // row stride of S
    static constexpr bool kPackS = !kInWaveFallback;
    static constexpr int CB = 64;  // floats per published column (every lane stores)
    static constexpr int o_qf = 0;                   // [DIM][TP] free points by rank (0-padded)
    static constexpr int o_y = o_qf + DIM * TP;      // [F][R] reduced right-hand sides
    static constexpr int o_v = o_y + F * R;          // [F][R] iterate, by unknown
    static constexpr int o_wp = o_v + F * R;         // [M][R] particular solution
    static constexpr int o_ru = o_wp + M * R;        // [M] 1 / U_tt of P's LU
    static constexpr int o_int = o_ru + M;           // ints: [F] node of each free rank,
                                                     //       [M] node of each pivot
    // FP32 pivot row (elim_np) or two published columns (elim_sym)
    static constexpr int o_pb = (o_int + (F + M + 1) / 2 + 1) / 2 * 2;
    static constexpr int pb_doubles = (NC + 3) / 4 * 2 > CB ? (NC + 3) / 4 * 2 : CB;
    static constexpr int o_g = o_pb + pb_doubles;    // [TP][GS] G by rank (rows >= F in region 1)
    static constexpr int o_r1 = o_g + F * GS;
    static constexpr int o_l = o_r1;                 // [M][M] pivot rows of P's LU
    static constexpr int l_span = M * M > (TP - F) * GS ? M * M : (TP - F) * GS;
    static constexpr int o_app = o_l + l_span;       // [M][M] A_pipi
    static constexpr int o_t = o_app + M * M;        // [M][R] t at the pivot nodes
    static constexpr int o_h = o_t + M * R;          // [TP][GS] H by rank (0-padded)
    static constexpr int r1_a = o_h + TP * GS - o_r1;
    static constexpr int r1_b = kPackS ? F * (F + 1) / 2 : F * SS;
    static constexpr int o_s = o_r1;
    static constexpr int ns_doubles = o_r1 + (r1_a > r1_b ? r1_a : r1_b);
    static constexpr int lds_doubles = ((ns_doubles > fb_lds ? ns_doubles : fb_lds) + 1) / 2 * 2;

    // Offset of S(i, j) in region 1 (packed: row i of the upper triangle
    // starts at i F - i (i - 1) / 2; S(i, j) = S(j, i))
    __device__ __forceinline__ static int s_at(int i, int j) {
        if constexpr (kPackS) {
            const int r = i < j ? i : j, c = i < j ? j : i;
            return r * F - (r * (r - 1)) / 2 + (c - r);
        } else {
            return i * SS + j;
        }
    }

    // r^E from r^2 for odd E: the v_rsq_f64 seed refined to the square
    // root as the compiler's own expansion does, without its rescaling of
    // denormal inputs and its special cases (stencil distances are normal
    // or exactly zero; zero gives zero)
    template <int E>
    __device__ __forceinline__ static double rpow(double r2) {
#pragma clang fp contract(off)
        const double y = __builtin_amdgcn_rsq(r2);
        double g = r2 * y, h = 0.5 * y;
        const double e = fma(-h, g, 0.5);
        g = fma(g, e, g);
        h = fma(h, e, h);
        double d = fma(-g, g, r2);
        g = fma(d, h, g);
        d = fma(-g, g, r2);
        g = fma(d, h, g);
        double p = g;
#pragma unroll
        for (int i = 0; i < (E - 1) / 2; ++i) p = p * r2;
        return r2 == 0.0 ? 0.0 : p;
    }

    // phi(|a - b|) = |a - b|^Q
    __device__ __forceinline__ static double phi_pair(double ax, double ay, double az, double bx,
                                                      double by, double bz) {
#pragma clang fp contract(off)
        const double dx = ax - bx, dy = ay - by;
        double r2 = dx * dx + dy * dy;
        if constexpr (DIM == 3) {
            const double dz = az - bz;
            r2 = r2 + dz * dz;
        }
        return rpow<Q>(r2);
    }

    // Elimination step K of the reduced system without pivoting: row K is
    // lane K. Rows below K store their multiplier in a[K]; rows K and above
    // keep their U entries.
    template <int K>
    __device__ __forceinline__ static void elim_np(float (&a)[NC], float& myrcp, float* pb) {
        if constexpr (K < F) {
            constexpr unsigned long long rows = F >= 64 ? ~0ull : ((1ull << F) - 1);
            constexpr unsigned long long below = rows & ~((2ull << K) - 1);
            const float piv = bcast(a[K], K);
            const float rcp = __builtin_amdgcn_rcpf(piv);
            const bool elim = lane_in(below);
            const bool me = lane_in(1ull << K);
            const float l = elim ? a[K] * rcp : 0.f;
            a[K] = elim ? l : a[K];
            myrcp = me ? rcp : myrcp;
            asm volatile("" : "+v"(myrcp), "+v"(a[K]));
            update_cols<K + 1, 0>(a, l, K, pb, me);
            elim_np<K + 1>(a, myrcp, pb);
        }
    }

    // The same step using the symmetry of S: the trailing matrix of a
    // symmetric matrix stays symmetric under elimination without pivoting,
    // so row K of it equals column K, which every lane publishes at once
    // (one LDS store per lane) and reads back at uniform addresses, instead
    // of a v_readlane per pivot-row entry. The right-hand sides still come
    // from lane K. Rows K and above keep their own (row) values as U.
    template <int K>
    __device__ __forceinline__ static void elim_sym(float (&a)[NC], float& myrcp, float* cb) {
        if constexpr (K >= Tune::SYM) {
            wave_lds_sync();  // column reads done before the buffer is reused
            elim_np<K>(a, myrcp, cb);
        } else if constexpr (K < F) {
            constexpr unsigned long long rows = F >= 64 ? ~0ull : ((1ull << F) - 1);
            constexpr unsigned long long below = rows & ~((2ull << K) - 1);
            float* col = cb + (K & 1) * CB;  // two buffers taken in turn
            col[lane_id()] = a[K];
            const float piv = bcast(a[K], K);
            const float rcp = __builtin_amdgcn_rcpf(piv);
            const bool elim = lane_in(below);
            const bool me = lane_in(1ull << K);
            const float l = elim ? a[K] * rcp : 0.f;
            a[K] = elim ? l : a[K];
            myrcp = me ? rcp : myrcp;
            asm volatile("" : "+v"(myrcp), "+v"(a[K]));
            float u[R];
#pragma unroll
            for (int c = 0; c < R; ++c) u[c] = bcast(a[F + c], K);
            wave_lds_sync();
#pragma unroll
            for (int j = K + 1; j < F; ++j) a[j] = fmaf(-l, col[j], a[j]);
#pragma unroll
            for (int c = 0; c < R; ++c) a[F + c] = fmaf(-l, u[c], a[F + c]);
            elim_sym<K + 1>(a, myrcp, cb);
        }
    }

    // Rows i < K of the factors divided by their pivot at column K, so U
    // has a unit diagonal (its reciprocal pivots are applied to the
    // right-hand sides instead): row i's U entries sit at columns j > i.
    template <int K>
    __device__ __forceinline__ static void unit_u(float (&a)[NC], float myrcp) {
        if constexpr (K < F) {
            constexpr unsigned long long lt = (1ull << K) - 1;  // lanes (rows) < K
            if (lane_in(lt)) a[K] = a[K] * myrcp;
            unit_u<K + 1>(a, myrcp);
        }
    }

    // U x = r with the unit-diagonal factors, r already scaled by the row's
    // reciprocal pivot: x_k is row k's value once the rows below have been
    // folded in; rows above take -u_ik x_k. Each row ends holding its x.
    __device__ __forceinline__ static void back_unit(const float (&a)[NC], float (&r)[R]) {
#pragma unroll
        for (int k = F - 1; k > 0; --k) {
            float xk[R];
#pragma unroll
            for (int c = 0; c < R; ++c) xk[c] = bcast(r[c], k);
            if (lane_in((1ull << k) - 1)) {
#pragma unroll
                for (int c = 0; c < R; ++c) r[c] = fmaf(-a[k], xk[c], r[c]);
            }
        }
    }

    // L y = r with the stored multipliers: rows below k take -l_ik y_k.
    __device__ __forceinline__ static void fwd_unit(const float (&a)[NC], float (&r)[R]) {
        constexpr unsigned long long rows = F >= 64 ? ~0ull : ((1ull << F) - 1);
#pragma unroll
        for (int k = 0; k < F - 1; ++k) {
            float yk[R];
#pragma unroll
            for (int c = 0; c < R; ++c) yk[c] = bcast(r[c], k);
            if (lane_in(rows & ~((2ull << k) - 1))) {
#pragma unroll
                for (int c = 0; c < R; ++c) r[c] = fmaf(-a[k], yk[c], r[c]);
            }
        }
    }

    // Phases A and D at two nodes per lane (node l and node 64 + l; N > 64):
    // gather, shift and scale; the polynomial block's LU with partial
    // pivoting over both node sets (the largest high word wins, the lowest
    // node among equal ones); G; the kernel entries against the pivot
    // nodes; the particular solution, t and H; the reduced right-hand sides
    // y. Free nodes take ranks in node order. The phases run in the order
    // that keeps the fewest values live (G goes to LDS before the kernel
    // entries are formed, t is formed before the H product and y reads G
    // back), and leave the LDS as the one-node-per-lane phases do.
    __device__ __forceinline__ static void node_phases_two(
        const int s, const int* __restrict__ ja_row, const double* __restrict__ x,
        const double* __restrict__ y, const double* __restrict__ z, const int lane, double* lds,
        double& irho, double& irho2, bool& bad) {
#pragma clang fp contract(off)
        constexpr int N1 = N - 64;  // nodes of the second set
        constexpr unsigned long long kRows1 = N1 >= 64 ? ~0ull : ((1ull << N1) - 1);
        double* qf = lds + o_qf;
        double* yb = lds + o_y;
        double* gb = lds + o_g;
        double* wpb = lds + o_wp;
        double* rub = lds + o_ru;
        int* fidx = reinterpret_cast<int*>(lds + o_int);
        int* pvn = fidx + F;
        double* lb = lds + o_l;
        double* app = lds + o_app;
        double* tb = lds + o_t;
        double* hb = lds + o_h;
        const bool has1 = lane_in(kRows1);

        // gather, shift, scale; the radius from the largest squared distance
        // by its high word, as with one node per lane
        double qx[2], qy[2], qz[2], d2[2];
        {
            const int jn[2] = {ja_row[lane], has1 ? ja_row[64 + lane] : s};
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                qx[k] = x[jn[k]] - x[s];
                qy[k] = y[jn[k]] - y[s];
                qz[k] = DIM == 3 ? z[jn[k]] - z[s] : 0.0;
                d2[k] = qx[k] * qx[k] + qy[k] * qy[k];
                if constexpr (DIM == 3) d2[k] = d2[k] + qz[k] * qz[k];
            }
        }
        const unsigned dkey0 = static_cast<unsigned>(__double2hiint(d2[0]));
        const unsigned dkey1 = has1 ? static_cast<unsigned>(__double2hiint(d2[1])) : 0u;
        const unsigned dmax = wave_umax_bcast(dkey0 > dkey1 ? dkey0 : dkey1);
        const unsigned long long at0 = wave_ballot(dkey0 == dmax);
        double rho2 = at0 ? bcast(d2[0], __ffsll(at0) - 1)
                          : bcast(d2[1], __ffsll(wave_ballot(dkey1 == dmax)) - 1);
        if (!(rho2 > 0.0)) rho2 = 1.0;
        const double rho = sqrt(rho2);
        irho = wave_uniform(1.0 / rho);
        irho2 = irho * irho;
        double mono[2][M];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            qx[k] = qx[k] * irho;
            qy[k] = qy[k] * irho;
            qz[k] = qz[k] * irho;
            PolyBasis<P, DIM>{}(qx[k], qy[k], qz[k], mono[k], 1);
        }

        // A: LU with partial pivoting of P over both node sets
        unsigned long long act[2] = {~0ull, kRows1};
        int mypt[2] = {M, M};
#pragma unroll
        for (int t = 0; t < M; ++t) {
            unsigned key[2];
#pragma unroll
            for (int k = 0; k < 2; ++k)
                key[k] = lane_in(act[k])
                             ? (static_cast<unsigned>(__double2hiint(mono[k][t])) & 0x7fffffffu)
                             : 0u;
            const unsigned mk = wave_umax_bcast(key[0] > key[1] ? key[0] : key[1]);
            const unsigned long long c0 = act[0] & wave_ballot(key[0] == mk);
            const int ps = c0 ? 0 : 1;  // set of the pivot node
            const int p = __ffsll(c0 ? c0 : act[1] & wave_ballot(key[1] == mk)) - 1;
            // (selects of values: a select of the two elements' addresses
            // would keep mono in scratch memory)
            const double a0 = mono[0][t], a1 = mono[1][t];
            const double piv = bcast(ps ? a1 : a0, p);
            if (!(piv != 0.0)) bad = true;
            double rcp = __builtin_amdgcn_rcp(piv);  // two Newton steps from the seed
            rcp = fma(rcp, fma(-piv, rcp, 1.0), rcp);
            rcp = fma(rcp, fma(-piv, rcp, 1.0), rcp);
            // (inner loops over all M columns, guarded, so that they unroll
            // fully and mono stays in registers)
            double pv[M];  // the pivot row right of column t
#pragma unroll
            for (int u = 0; u < M; ++u) {
                const double u0 = mono[0][u], u1 = mono[1][u];
                if (u > t) pv[u] = bcast(ps ? u1 : u0, p);
            }
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const unsigned long long pm = ps == k ? 1ull << p : 0ull;
                if (lane_in(act[k] & ~pm)) {
                    const double l = mono[k][t] * rcp;
#pragma unroll
                    for (int u = 0; u < M; ++u)
                        if (u > t) mono[k][u] = fma(-l, pv[u], mono[k][u]);
                    mono[k][t] = l;
                }
                mypt[k] = lane_in(pm) ? t : mypt[k];
                act[k] &= ~pm;
            }
            if (lane == 0) rub[t] = rcp;
        }
        const bool isfree[2] = {lane_in(act[0]), lane_in(act[1])};
        int rank[2];  // free nodes of the lanes below, then the whole first set
#pragma unroll
        for (int k = 0; k < 2; ++k)
            rank[k] = static_cast<int>(__builtin_amdgcn_mbcnt_hi(
                static_cast<unsigned>(act[k] >> 32),
                __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(act[k]), 0u)));
        rank[1] += static_cast<int>(__popcll(act[0]));
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            if (mypt[k] < M) {
#pragma unroll
                for (int u = 0; u < M; ++u) lb[mypt[k] * M + u] = mono[k][u];
                pvn[mypt[k]] = lane + 64 * k;
                tb[mypt[k] * R] = qx[k];
                tb[mypt[k] * R + 1] = qy[k];
                if constexpr (DIM == 3) tb[mypt[k] * R + 2] = qz[k];
            }
        }
        wave_lds_sync();
        // G = L2 L1^-1 for both sets, to LDS by free rank
#pragma unroll
        for (int t = M - 2; t >= 0; --t) {
            double acc[2] = {mono[0][t], mono[1][t]};
#pragma unroll
            for (int u = 0; u < M; ++u) {
                if (u > t) {
                    const double l1 = lb[u * M + t];
#pragma unroll
                    for (int k = 0; k < 2; ++k) acc[k] = fma(-mono[k][u], l1, acc[k]);
                }
            }
            mono[0][t] = acc[0];
            mono[1][t] = acc[1];
        }
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            if (isfree[k]) {
                qf[rank[k]] = qx[k];
                qf[TP + rank[k]] = qy[k];
                if constexpr (DIM == 3) qf[2 * TP + rank[k]] = qz[k];
                fidx[rank[k]] = lane + 64 * k;
#pragma unroll
                for (int t = 0; t < M; ++t) gb[rank[k] * GS + t] = mono[k][t];
            }
        }
        // kernel entries against the pivot nodes: A_pipi, H's initial rows
        // (A_Fpi by free rank)
        double akp[2][M];
#pragma unroll
        for (int t = 0; t < M; ++t) {
            const double bx = tb[t * R], by = tb[t * R + 1], bz = DIM == 3 ? tb[t * R + 2] : 0.0;
#pragma unroll
            for (int k = 0; k < 2; ++k) akp[k][t] = phi_pair(qx[k], qy[k], qz[k], bx, by, bz);
        }
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            if (mypt[k] < M) {
#pragma unroll
                for (int t = 0; t < M; ++t) app[mypt[k] * M + t] = akp[k][t];
            }
            if (isfree[k]) {
#pragma unroll
                for (int t = 0; t < M; ++t) hb[rank[k] * GS + t] = akp[k][t];
            }
        }
        if (lane < TP - F) {
#pragma unroll
            for (int a = 0; a < DIM; ++a) qf[a * TP + F + lane] = 0.0;
#pragma unroll
            for (int t = 0; t < M; ++t) hb[(F + lane) * GS + t] = 0.0;
        }
        wave_lds_sync();

        // D: the particular solution; t = b - A_:pi w_p at every node (the
        // pivot nodes' values to LDS); H; y = t_F - G t_pi
        particular_solution(lane, lb, rub, wpb);
        double tk[2][R];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            double r2 = qx[k] * qx[k] + qy[k] * qy[k];
            if constexpr (DIM == 3) r2 = r2 + qz[k] * qz[k];
            const double tq = rpow<Q - 2>(r2);
            const double gq = static_cast<double>(Q) * tq;
            tk[k][0] = static_cast<double>(Q * (Q + DIM - 2)) * tq;
            tk[k][1] = -gq * qx[k];
            tk[k][2] = -gq * qy[k];
            if constexpr (DIM == 3) tk[k][3] = -gq * qz[k];
        }
#pragma unroll
        for (int c = 0; c < R; ++c) {
#pragma unroll
            for (int u = 0; u < M; ++u) {
                const double wp = wpb[u * R + c];
#pragma unroll
                for (int k = 0; k < 2; ++k) tk[k][c] = fma(-akp[k][u], wp, tk[k][c]);
            }
        }
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            if (mypt[k] < M) {
#pragma unroll
                for (int c = 0; c < R; ++c) tb[mypt[k] * R + c] = tk[k][c];
            }
        }
        h_product(lane, gb, hb, app);  // ends with a wave barrier: t_pi is in LDS
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            if (isfree[k]) {
#pragma unroll
                for (int c = 0; c < R; ++c) {
                    double acc = tk[k][c];
#pragma unroll
                    for (int u = 0; u < M; ++u)
                        acc = fma(-gb[rank[k] * GS + u], tb[u * R + c], acc);
                    yb[rank[k] * R + c] = acc;
                }
            }
        }
    }

    // The rank-wise parts of phase D. particular_solution: lane c < R
    // solves P_pi^T w = g_c (U^T z = g_c forward, L1^T w = z back) into
    // wpb. h_product: G's padding rows take the place of P's pivot rows,
    // and H = A_Fpi - G A_pipi / 2 is formed on FP64 MFMA: tile I of free
    // ranks, column tile J of the m pivot columns (C = A_Fpi from hb,
    // A = -G/2, B = A_pipi).
    __device__ __forceinline__ static void particular_solution(const int lane, const double* lb,
                                                               const double* rub, double* wpb) {
#pragma clang fp contract(off)
        if (lane < R) {
            using Mk = PolyDerivMasks<DIM, P>;
            const uint64_t gm = lane == 0 ? Mk::lap : (lane == 1 ? Mk::dx : (lane == 2 ? Mk::dy : Mk::dz));
            const double gv = lane == 0 ? 2.0 : 1.0;
            double w[M];
#pragma unroll
            for (int t = 0; t < M; ++t) {
                double acc = ((gm >> t) & 1u) ? gv : 0.0;
#pragma unroll
                for (int u = 0; u < t; ++u) acc = fma(-lb[u * M + t], w[u], acc);
                w[t] = acc * rub[t];
            }
#pragma unroll
            for (int t = M - 2; t >= 0; --t) {
                double acc = w[t];
#pragma unroll
                for (int u = t + 1; u < M; ++u) acc = fma(-lb[u * M + t], w[u], acc);
                w[t] = acc;
            }
#pragma unroll
            for (int t = 0; t < M; ++t) wpb[t * R + lane] = w[t];
        }
        wave_lds_sync();  // P's pivot rows read: G's padding rows take their place
    }
    __device__ __forceinline__ static void h_product(const int lane, double* gb, double* hb,
                                                     const double* app) {
        if (lane < TP - F) {
#pragma unroll
            for (int t = 0; t < M; ++t) gb[(F + lane) * GS + t] = 0.0;
        }
        wave_lds_sync();
        {
            const int li = lane & 15, lg = lane >> 4;
            constexpr int KH = (M + 3) / 4;
            constexpr int TH = (M + 15) / 16;  // column tiles of H (pivot columns)
#pragma unroll
            for (int I = 0; I < T; ++I) {
#pragma unroll
                for (int J = 0; J < TH; ++J) {
                    const int col = 16 * J + li;
                    ns_dbl4 c;
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                        c[v] = col < M ? hb[(16 * I + lg + 4 * v) * GS + col] : 0.0;
#pragma unroll
                    for (int k = 0; k < KH; ++k) {
                        const int kk = 4 * k + lg;
                        const double av = kk < M ? -0.5 * gb[(16 * I + li) * GS + kk] : 0.0;
                        const double bv = (kk < M && col < M) ? app[kk * M + col] : 0.0;
                        c = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, c, 0, 0, 0);
                    }
                    if (col < M) {
#pragma unroll
                        for (int v = 0; v < 4; ++v) hb[(16 * I + lg + 4 * v) * GS + col] = c[v];
                    }
                }
            }
        }
        wave_lds_sync();
    }

    template <class Out>
    __device__ __forceinline__ static void run(const int s_arg, const size_t base_arg,
                                               const double* __restrict__ x,
                                               const double* __restrict__ y,
                                               const double* __restrict__ z,
                                               const int* __restrict__ ja, Out out, double* lds,
                                               int* __restrict__ info,
                                               uint8_t* __restrict__ status, int max_refine,
                                               int* __restrict__ fb_count = nullptr,
                                               int* __restrict__ fb_list = nullptr) {
#pragma clang fp contract(off)
        const int s = __builtin_amdgcn_readfirstlane(s_arg);
        const size_t base = wave_uniform(base_arg);
        const int lane = lane_id();
        const int* ja_row = ja + base;
        double* qf = lds + o_qf;
        double* yb = lds + o_y;
        double* vb = lds + o_v;
        double* gb = lds + o_g;
        double* wpb = lds + o_wp;
        double* rub = lds + o_ru;
        int* fidx = reinterpret_cast<int*>(lds + o_int);
        int* pvn = fidx + F;
        float* pb = reinterpret_cast<float*>(lds + o_pb);
        double* lb = lds + o_l;
        double* app = lds + o_app;
        double* tb = lds + o_t;
        double* hb = lds + o_h;
        double* sb = lds + o_s;

        double irho, irho2;
        bool bad = false;
        if constexpr (NSETS == 2) {
            node_phases_two(s, ja_row, x, y, z, lane, lds, irho, irho2, bad);
        } else {
            // ---- gather, shift, scale (as the FP64 path) ---------------------
            const int jn = lane < N ? ja_row[lane] : s;
            const double px = x[jn] - x[s];
            const double py = y[jn] - y[s];
            const double pz = DIM == 3 ? z[jn] - z[s] : 0.0;
            double d2 = px * px + py * py;
            if constexpr (DIM == 3) d2 = d2 + pz * pz;
            // the stencil radius: the node with the largest squared distance by
            // its high word (the scaling only needs a radius of the right size)
            const unsigned dkey = lane < N ? static_cast<unsigned>(__double2hiint(d2)) : 0u;
            const unsigned dmax = wave_umax_bcast(dkey);
            double rho2 = bcast(d2, __ffsll(wave_ballot(dkey == dmax)) - 1);
            if (!(rho2 > 0.0)) rho2 = 1.0;
            const double rho = sqrt(rho2);
            irho = wave_uniform(1.0 / rho);
            irho2 = irho * irho;
            const double qx = px * irho, qy = py * irho, qz = pz * irho;
            double mono[M];
            PolyBasis<P, DIM>{}(qx, qy, qz, mono, 1);
            // right-hand sides of the kernel rows (the Laplacian and gradient of
            // phi(|q - q_k|) at the origin)
            double bk[R];
            {
                double r2 = qx * qx + qy * qy;
                if constexpr (DIM == 3) r2 = r2 + qz * qz;
                const double tq = rpow<Q - 2>(r2);
                const double gq = static_cast<double>(Q) * tq;
                bk[0] = static_cast<double>(Q * (Q + DIM - 2)) * tq;
                bk[1] = -gq * qx;
                bk[2] = -gq * qy;
                if constexpr (DIM == 3) bk[3] = -gq * qz;
            }

            // ---- A: LU with partial pivoting of P, FP64 ------------------------
            constexpr unsigned long long kRows = N >= 64 ? ~0ull : ((1ull << N) - 1);
            unsigned long long act = kRows;
            int mypt = M;  // pivot step of this lane's node (M: a free node)
#pragma unroll
            for (int t = 0; t < M; ++t) {
                const double a = mono[t];
                // pivot: the largest |a| among the rows still active, compared by
                // its high word (lowest lane among equal high words: any of them
                // serves the conditioning); NaN wins and sends the stencil to the
                // FP64 solve through the checks below
                const unsigned key =
                    lane_in(act) ? (static_cast<unsigned>(__double2hiint(a)) & 0x7fffffffu) : 0u;
                const unsigned mk = wave_umax_bcast(key);
                const int p = __ffsll(act & wave_ballot(key == mk)) - 1;
                const double piv = bcast(a, p);
                if (!(piv != 0.0)) bad = true;
                double rcp = __builtin_amdgcn_rcp(piv);  // two Newton steps from the seed
                rcp = fma(rcp, fma(-piv, rcp, 1.0), rcp);
                rcp = fma(rcp, fma(-piv, rcp, 1.0), rcp);
                const unsigned long long pm = 1ull << p;
                if (lane_in(act & ~pm)) {
                    const double l = a * rcp;
#pragma unroll
                    for (int u = t + 1; u < M; ++u) mono[u] = fma(-l, bcast(mono[u], p), mono[u]);
                    mono[t] = l;
                }
                mypt = lane_in(pm) ? t : mypt;
                if (lane == 0) rub[t] = rcp;
                act &= ~pm;
            }
            const unsigned long long freem = act;  // free nodes, in lane order
            const bool isfree = lane_in(freem);
            const int rank = __builtin_amdgcn_mbcnt_hi(
                static_cast<unsigned>(freem >> 32),
                __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(freem), 0u));
            // pivot rows: L1 below the diagonal, U on and above it; the pivot
            // nodes' points (in t_pi's place, which is written later)
            if (mypt < M) {
#pragma unroll
                for (int u = 0; u < M; ++u) lb[mypt * M + u] = mono[u];
                pvn[mypt] = lane;
                tb[mypt * R] = qx;
                tb[mypt * R + 1] = qy;
                if constexpr (DIM == 3) tb[mypt * R + 2] = qz;
            }
            wave_lds_sync();
            // kernel entries against the pivot nodes (the pivot lanes' values
            // form A_pipi)
            double akp[M];
#pragma unroll
            for (int t = 0; t < M; ++t)
                akp[t] = phi_pair(qx, qy, qz, tb[t * R], tb[t * R + 1], DIM == 3 ? tb[t * R + 2] : 0.0);
            if (mypt < M) {
#pragma unroll
                for (int t = 0; t < M; ++t) app[mypt * M + t] = akp[t];
            }
            // padding ranks F .. TP-1: zero points and H rows (G's below)
            if (lane < TP - F) {
#pragma unroll
                for (int a = 0; a < DIM; ++a) qf[a * TP + F + lane] = 0.0;
#pragma unroll
                for (int t = 0; t < M; ++t) hb[(F + lane) * GS + t] = 0.0;
            }
            wave_lds_sync();
            // G = L2 L1^-1: g L1 = l, from the last column (L1 read from LDS; the
            // pivot lanes compute values nobody reads)
#pragma unroll
            for (int t = M - 2; t >= 0; --t) {
                double acc = mono[t];
#pragma unroll
                for (int u = t + 1; u < M; ++u) acc = fma(-mono[u], lb[u * M + t], acc);
                mono[t] = acc;
            }
            if (isfree) {
                qf[rank] = qx;
                qf[TP + rank] = qy;
                if constexpr (DIM == 3) qf[2 * TP + rank] = qz;
                fidx[rank] = lane;
#pragma unroll
                for (int t = 0; t < M; ++t) {
                    gb[rank * GS + t] = mono[t];
                    hb[rank * GS + t] = akp[t];
                }
            }

            // ---- D: particular solution, t, y -----------------------------------
            particular_solution(lane, lb, rub, wpb);
            h_product(lane, gb, hb, app);
            // t = b - A_:pi w_p at every node; the pivot nodes' values to LDS
            double tk[R];
#pragma unroll
            for (int c = 0; c < R; ++c) {
                double acc = bk[c];
#pragma unroll
                for (int u = 0; u < M; ++u) acc = fma(-akp[u], wpb[u * R + c], acc);
                tk[c] = acc;
            }
            if (mypt < M) {
#pragma unroll
                for (int c = 0; c < R; ++c) tb[mypt * R + c] = tk[c];
            }
            wave_lds_sync();
            if (isfree) {
#pragma unroll
                for (int c = 0; c < R; ++c) {
                    double acc = tk[c];
#pragma unroll
                    for (int u = 0; u < M; ++u) acc = fma(-mono[u], tb[u * R + c], acc);
                    yb[rank * R + c] = acc;
                }
            }
        }

        // ---- B: S = A_FF - [G | H] [H | G]^T, FP64 MFMA ---------------------
        // (padding ranks have zero G and H rows; their entries are ignored)
        const int li = lane & 15, lg = lane >> 4;
        ns_dbl4 acc[NTILE];
        {
            int ti = 0;
#pragma unroll
            for (int I = 0; I < T; ++I) {
#pragma unroll
                for (int J = I; J < T; ++J, ++ti) {
                    const int col = 16 * J + li;
                    const double cx = qf[col], cy = qf[TP + col];
                    const double cz = DIM == 3 ? qf[2 * TP + col] : 0.0;
                    ns_dbl4 c;
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int row = 16 * I + lg + 4 * v;
                        c[v] = phi_pair(qf[row], qf[TP + row], DIM == 3 ? qf[2 * TP + row] : 0.0,
                                        cx, cy, cz);
                    }
                    const int ra = 16 * I + li;  // A operand row (free rank)
#pragma unroll
                    for (int k = 0; k < KS; ++k) {
                        const int kk = 4 * k + lg;
                        // A = -[G | H] rows of tile I; B = [H | G]^T columns of tile J
                        double av, bv;
                        if (4 * k + 3 < M) {
                            av = -gb[ra * GS + kk];
                            bv = hb[col * GS + kk];
                        } else if (4 * k >= M && 4 * k + 3 < K2) {
                            av = -hb[ra * GS + kk - M];
                            bv = gb[col * GS + kk - M];
                        } else {
                            const bool lo = kk < M;
                            const int kq = lo ? kk : kk - M;
                            const double* ga = lo ? gb : hb;
                            const double* gbb = lo ? hb : gb;
                            av = kk < K2 ? -ga[ra * GS + kq] : 0.0;
                            bv = kk < K2 ? gbb[col * GS + kq] : 0.0;
                        }
                        c = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, c, 0, 0, 0);
                    }
                    acc[ti] = c;
                }
            }
        }
        wave_lds_sync();  // G/H reads done before S overwrites region 1
        {
            int ti = 0;
#pragma unroll
            for (int I = 0; I < T; ++I) {
#pragma unroll
                for (int J = I; J < T; ++J, ++ti) {
                    const int col = 16 * J + li;
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int row = 16 * I + lg + 4 * v;
                        if constexpr (kPackS) {
                            if (row <= col && col < F) sb[s_at(row, col)] = acc[ti][v];
                        } else if (row < F && col < F) {
                            sb[row * SS + col] = acc[ti][v];
                            if (I != J) sb[col * SS + row] = acc[ti][v];
                        }
                    }
                }
            }
        }
        wave_lds_sync();

        // ---- C: FP32 LU of S (no pivoting) with the reduced right-hand sides
        const bool is_row = lane < F;
        float a32[NC];
        double yr[R];
        float rowmax = 0.f;
#pragma unroll
        for (int j = 0; j < F; ++j) {
            a32[j] = is_row ? static_cast<float>(sb[s_at(lane, j)]) : 0.f;
            rowmax = fmaxf(rowmax, fabsf(a32[j]));
        }
#pragma unroll
        for (int c = 0; c < R; ++c) {
            yr[c] = is_row ? yb[lane * R + c] : 0.0;
            a32[F + c] = static_cast<float>(yr[c]);
        }
        float myrcp = 0.f;
        elim_sym<0>(a32, myrcp, pb);
        // a pivot below 1e-6 of the largest entry of S (duplicated nodes,
        // nearly singular stencils): the FP32 factors carry no information
        // and the stencil goes to the FP64 solve
        const float smax = __uint_as_float(wave_umax(__float_as_uint(rowmax)));
        if (wave_ballot(is_row && !(fabsf(myrcp) * (1e-6f * smax) < 1.f))) bad = true;
        unit_u<1>(a32, myrcp);
        // initial solve from the eliminated right-hand sides
        float r32[R];
#pragma unroll
        for (int c = 0; c < R; ++c) r32[c] = a32[F + c] * myrcp;
        back_unit(a32, r32);
        double vd[R];
        float xm[R];  // size of each operator's weights, for the stopping test
#pragma unroll
        for (int c = 0; c < R; ++c) {
            vd[c] = is_row ? static_cast<double>(r32[c]) : 0.0;
            xm[c] = __uint_as_float(wave_umax(__float_as_uint(is_row ? fabsf(r32[c]) : 0.f)));
        }

        // ---- E: FP64 refinement on S v = y ----------------------------------
        int steps = 0;
        bool converged = false;
        float dprev[R];
#pragma unroll
        for (int c = 0; c < R; ++c) dprev[c] = 0.f;
        if (!bad) {
            for (int it = 0; it < max_refine; ++it) {
                if (is_row) {
#pragma unroll
                    for (int c = 0; c < R; ++c) vb[lane * R + c] = vd[c];
                }
                wave_lds_sync();
                double rr[R];
#pragma unroll
                for (int c = 0; c < R; ++c) rr[c] = yr[c];
                if (is_row) {
#pragma unroll
                    for (int j = 0; j < F; ++j) {
                        const double sij = sb[s_at(lane, j)];
#pragma unroll
                        for (int c = 0; c < R; ++c) rr[c] = fma(-sij, vb[j * R + c], rr[c]);
                    }
                }
#pragma unroll
                for (int c = 0; c < R; ++c) r32[c] = static_cast<float>(rr[c]);
                fwd_unit(a32, r32);
#pragma unroll
                for (int c = 0; c < R; ++c) r32[c] = r32[c] * myrcp;
                back_unit(a32, r32);
                ++steps;
                bool ok = true;
#pragma unroll
                for (int c = 0; c < R; ++c) {
                    const float d = is_row ? r32[c] : 0.f;
                    vd[c] = vd[c] + static_cast<double>(d);
                    const float dm = __uint_as_float(wave_umax(__float_as_uint(fabsf(d))));
                    // remaining error ~ |d| x contraction; the contraction is
                    // |d| / |v| after the first step (the initial solve's
                    // relative error), |d_k| / |d_k-1| after later ones
                    const float rho_c = it == 0 ? (xm[c] > 0.f ? dm / xm[c] : 1.f)
                                                : (dprev[c] > 0.f ? dm / dprev[c] : 1.f);
                    if (!(dm * fminf(1.f, rho_c) <= static_cast<float>(kTol) * xm[c])) ok = false;
                    dprev[c] = dm;
                }
                if (ok) {
                    converged = true;
                    break;
                }
            }
        }

        if (!converged) {
            if constexpr (kInWaveFallback) {
                // the plain FP64 solve of the whole stencil (its own fill,
                // LU, scatter and info); regions 0 and 1 are no longer needed
                wave_lds_sync();
                Fb::run(s, base, x, y, z, ja, out, nullptr, nullptr, nullptr, lds, info);
            } else if (lane == 0) {
                fb_list[atomicAdd(fb_count, 1)] = s;  // solved by the FP64 pass
            }
            if (lane == 0 && status) status[s] = static_cast<uint8_t>((steps & 3) | 4);
            return;
        }

        // ---- F: weights --------------------------------------------------------
        if (is_row) {
#pragma unroll
            for (int c = 0; c < R; ++c) vb[lane * R + c] = vd[c];
            const size_t at = base + static_cast<size_t>(fidx[lane]);
#pragma unroll
            for (int c = 0; c < R; ++c) out(c)[at] = vd[c] * (c == 0 ? irho2 : irho);
        }
        wave_lds_sync();
        if (lane < M) {
            // w_pi = w_p - G^T v for pivot step t = lane
            double wsum[R];
#pragma unroll
            for (int c = 0; c < R; ++c) wsum[c] = wpb[lane * R + c];
#pragma unroll
            for (int j = 0; j < F; ++j) {
                const double g = gb[j * GS + lane];
#pragma unroll
                for (int c = 0; c < R; ++c) wsum[c] = fma(-g, vb[j * R + c], wsum[c]);
            }
            const size_t at = base + static_cast<size_t>(pvn[lane]);
#pragma unroll
            for (int c = 0; c < R; ++c) out(c)[at] = wsum[c] * (c == 0 ? irho2 : irho);
        }
        if (lane == 0) {
            if (info) info[s] = 0;
            if (status) status[s] = static_cast<uint8_t>(steps & 3);
        }
    }
};

}  // namespace detail
}  // namespace rbf_operators

#endif  // RBF_NULLSPACE_H
