// rbf_tile.h -- blocked LU with FP64 MFMA trailing updates for stencil
// systems of 64 unknowns and more (gfx950). Companion of rbf_operators.h
// and rbf_group.h: same fill, same pivot choices, and the same rounding as
// the right-looking oracle, operation for operation.
//
// Layout. The system [M | B] (NT x NT plus R right-hand sides) is padded to
// 16 x 16 tiles and held in registers in the MFMA accumulator layout of
// v_mfma_f64_16x16x4f64: lane l, component v holds entry
// (16T + l/16 + 4v, 16C + l%16) of tile (T, C). The workgroup has four
// waves (a multiple of the CU's four SIMDs: a workgroup of 5 or 7 waves
// reserves two wave slots on every SIMD and halves the workgroups a CU can
// hold); wave w holds column tiles w, w + 4, ... The right-hand sides share
// the last column tile with the last matrix columns.
//
// Panels. Panel q is column tile q (elimination steps 16q .. 16q+15). Its
// owner transposes it through LDS to one row per lane (two rows per lane
// above 64) and eliminates it with the wave-LU step of rbf_operators.h,
// restricted to the panel's 16 columns: pivot = largest |a| among rows not
// yet pivotal (first row on ties), multiplier l = a * (1 / pivot), column
// zeroed, panel columns updated by fma(-l, u, a), the pivot row's values
// read from its lane into SGPRs. It publishes -l for every row, the pivot
// rows and their reciprocals. After one barrier every
// wave updates its own column tiles to the right of the panel:
//   U block: the panel's 16 pivot rows at those columns, by the same
//            chain of fma the oracle applies to them while still active;
//   trailing: four chained MFMAs per tile with A = -L (rows x 4 steps) and
//            B = U (4 steps x columns).
// v_mfma_f64_16x16x4f64 computes each entry as fma(a3, b3, fma(a2, b2,
// fma(a1, b1, fma(a0, b0, c)))), so a tile receives exactly the oracle's
// per-step updates, in step order. Rows never move; a row's multipliers are
// 0 outside the steps where it was active.
//
// Back substitution runs in the wave that owns the right-hand sides, one
// row per lane, block by block from the last column tile down (each block
// passed through LDS by its owner, one barrier per block, its columns read
// one step at a time), in the oracle's order: t = b_p * rcp_p,
// b_i = fma(-u_ik, t, b_i).
#ifndef RBF_TILE_H
#define RBF_TILE_H

#include "rbf_group.h"

namespace rbf_operators {
namespace detail {

typedef double dbl4 __attribute__((ext_vector_type(4)));

template <int DIM, int N, int P, int Q, int W = 4, bool PANEL_PRIO = true>
struct TileAssembly {
    static constexpr int M = npoly_dim(DIM, P);
    static constexpr int NT = N + M;
    static constexpr int R = DIM + 1;
    static constexpr int NC = NT + R;
    static constexpr int nrhs = R;
    static constexpr int nsten = N;
    static constexpr int RT = (NT + 15) / 16;  // row tiles = panels
    static constexpr int MP = 16 * RT;         // padded rows
    static constexpr int CT = (NC + 15) / 16;  // column tiles
    static constexpr int LC = (CT + W - 1) / W;  // column-tile slots per wave
    static constexpr int S = 17;  // row stride of every LDS row block (spreads the banks)
    static constexpr int NONE = 1 << 20;
    static_assert(CT == RT, "the right-hand sides share the last column tile");
    static_assert(MP <= 128, "at most two rows per lane in a panel");
    static_assert(N <= 64 * W, "one stencil node per lane in the gather");
    static_assert(NT >= 64, "smaller systems use the single-wave engine");
    using Helpers = GroupAssembly<DIM, N, P, Q, R, false, (NT + 63) / 64>;

    // LDS (doubles): [fill: points, monomials | panel row block [MP][S]],
    // multiplier panels [2][MP][S] (reused as the back-substitution block
    // buffers), U staging [W][16][16], the multipliers of each panel's
    // pivot rows in step order [2][16][16], pivot reciprocals [NT], per-wave
    // maxima [W], then ints: pivot row of each step [NT], step of each row
    // [MP], info.
    static constexpr int FILL = DIM * N + M * N;
    static constexpr int XB = (MP * S + 1) / 2 * 2;
    static constexpr int R0 = ((FILL > XB ? FILL : XB) + 1) / 2 * 2;
    static constexpr int off_pts = 0;
    static constexpr int off_poly = DIM * N;
    static constexpr int off_x = 0;
    static constexpr int off_l = R0;
    static constexpr int off_u = off_l + 2 * XB;
    static constexpr int off_l11 = off_u + W * 256;
    static constexpr int off_rcp = off_l11 + 2 * 256;
    static constexpr int off_wmax = off_rcp + (NT + 1) / 2 * 2;
    static constexpr int off_int = off_wmax + (W + 1) / 2 * 2;
    static constexpr int int_perm = 0, int_step = NT, int_info = NT + MP;
    static constexpr int lds_doubles = off_int + (NT + MP + 2) / 2;
    static constexpr int threads = 64 * W;
    static_assert(lds_doubles * 8 <= 65536, "static LDS limit");
    // Waves per SIMD to ask the register allocator for (the kernels'
    // __launch_bounds__): as many four-wave workgroups as the CU's 160 KB
    // of LDS holds, at most three (168 registers). Three workgroups per CU
    // beat two even where 168 registers spill a few values to scratch
    // (71x71: 115.7 ms against 147.0 ms for 1M stencils).
    static constexpr int occupancy = (160 * 1024) / (lds_doubles * 8) >= 3 ? 3 : 2;

    // ------------------------------------------------------------------ fill
    // Entry (i, j) of [M | B | 0] from the scaled points and monomials in
    // LDS, with the arithmetic of GroupAssembly::entry / rhs.
    __device__ __forceinline__ static double elem(int i, int j, const double* lds) {
#pragma clang fp contract(off)
        const double* pts = lds + off_pts;
        const double* polyT = lds + off_poly;
        if (i < N) {
            const double xi = pts[i];
            const double yi = pts[N + i];
            const double zi = DIM == 3 ? pts[2 * N + i] : 0.0;
            if (j < N) {
                const double dx = pts[j] - xi;
                const double dy = pts[N + j] - yi;
                double r2 = dx * dx + dy * dy;
                if constexpr (DIM == 3) {
                    const double dz = pts[2 * N + j] - zi;
                    r2 = r2 + dz * dz;
                }
                return PHS<Q>{}(r2);
            }
            if (j < NT) return polyT[(j - N) * N + i];
            if (j < NC) return Helpers::rhs(i, j - NT, xi, yi, zi);
            return 0.0;
        }
        if (i < NT) {
            if (j < N) return polyT[(i - N) * N + j];
            if (j < NT) return 0.0;
            if (j < NC) return Helpers::rhs(i, j - NT, 0.0, 0.0, 0.0);
        }
        return 0.0;
    }

    // Gather, shift and scale stencil s into LDS (points, monomials), as
    // GroupAssembly::gather does; reset the pivot bookkeeping. Returns rho^2.
    template <typename Coord>
    __device__ __forceinline__ static double gather(const int s, const size_t base,
                                                    const Coord* __restrict__ x,
                                                    const Coord* __restrict__ y,
                                                    nodeduce<const Coord*> __restrict__ z,
                                                    const int* __restrict__ ja, double* lds, int g,
                                                    int w) {
#pragma clang fp contract(off)
        const int lane = lane_id();
        const int jn = g < N ? ja[base + g] : s;
        const double px = static_cast<double>(x[jn]) - static_cast<double>(x[s]);
        const double py = static_cast<double>(y[jn]) - static_cast<double>(y[s]);
        const double pz = DIM == 3 ? static_cast<double>(z[jn]) - static_cast<double>(z[s]) : 0.0;
        double d2 = px * px + py * py;
        if constexpr (DIM == 3) d2 = d2 + pz * pz;
        const double wm = wave_max(g < N ? d2 : 0.0);
        if (lane == 0) lds[off_wmax + w] = wm;
        __syncthreads();
        double rho2 = 0.0;
#pragma unroll
        for (int v = 0; v < W; ++v) rho2 = fmax(rho2, lds[off_wmax + v]);
        if (!(rho2 > 0.0)) rho2 = 1.0;
        const double rho = sqrt(rho2);
        const double xs = px / rho;
        const double ys = py / rho;
        const double zs = pz / rho;
        if (g < N) {
            constexpr PolyBasis<P, DIM> basis{};
            double mono[M];
            basis(xs, ys, zs, mono, 1);
            lds[off_pts + g] = xs;
            lds[off_pts + N + g] = ys;
            if constexpr (DIM == 3) lds[off_pts + 2 * N + g] = zs;
#pragma unroll
            for (int t = 0; t < M; ++t) lds[off_poly + t * N + g] = mono[t];
        }
        int* ints = reinterpret_cast<int*>(lds + off_int);
        for (int i = g; i < MP; i += threads) ints[int_step + i] = NONE;
        if (g == 0) ints[int_info] = 0;
        __syncthreads();
        return rho2;
    }

    // ------------------------------------------------------------------ panel
    // Pivot of step k from its column (a0: rows 0..63, a1: rows 64..): the
    // largest |a| among active rows (NaN never wins), the lowest row on
    // ties, else the first active row with info = k + 1; info = k + 1 also
    // for a zero pivot. Each lane first keeps the better of its two rows
    // (row lane before row 64 + lane on ties), then one wave maximum on the
    // high word of |a| (low word only when high words tie). Returns the
    // global row and 1 / pivot.
    __device__ __forceinline__ static int pivot_of(double a0, double a1, bool act0, bool act1,
                                                   int k, int& info, double& rcp) {
        const bool c0 = act0 && !__builtin_isnan(a0);
        const bool c1 = act1 && !__builtin_isnan(a1);
        const unsigned long long k0 = static_cast<unsigned long long>(__double_as_longlong(fabs(a0)));
        const unsigned long long k1 = static_cast<unsigned long long>(__double_as_longlong(fabs(a1)));
        const bool use1 = c1 && (!c0 || k1 > k0);
        const bool cb = c0 || c1;
        const unsigned long long kb = use1 ? k1 : k0;
        const unsigned hi = cb ? static_cast<unsigned>(kb >> 32) : 0u;
        const unsigned mh = wave_umax_bcast(hi);
        // lane masks combined in SGPRs: a ballot of a plain compare needs no
        // VALU select (71x71: 91.0 against 93.6 ms with the candidate flags
        // balloted as combined predicates; 100x100: 2.4%)
        const unsigned long long mcb = __ballot(cb);
        const unsigned long long mu1 = __ballot(use1);
        unsigned long long top = wave_ballot(hi == mh) & mcb;
        if (__popcll(top) > 1) {
            const unsigned lo = lane_in(top) ? static_cast<unsigned>(kb) : 0u;
            const unsigned ml = wave_umax_bcast(lo);
            top &= wave_ballot(lo == ml);
        }
        const unsigned long long h0 = top & ~mu1;
        const unsigned long long h1 = top & mu1;
        int win;
        if (h0) {
            win = __ffsll(h0) - 1;
        } else if (h1) {
            win = 64 + __ffsll(h1) - 1;
        } else {  // no candidate that is a number: first row still active
            const unsigned long long b0 = __ballot(act0);
            const unsigned long long b1 = __ballot(act1);
            win = b0 ? __ffsll(b0) - 1 : (b1 ? 64 + __ffsll(b1) - 1 : 0);
            if (!info) info = k + 1;
        }
        win = __builtin_amdgcn_readfirstlane(win);
        const double piv = bcast(win >= 64 ? a1 : a0, win & 63);
        if (piv == 0.0 && !info) info = k + 1;
        rcp = 1.0 / piv;
        return win;
    }

    // Update panel columns J.. of both row sets with the pivot row's values
    // u[J..], read from the pivot lane into SGPRs (v_readlane): the panel
    // is latency-bound, and this is shorter than a round trip through LDS
    // (71x71: 93.9 against 101.9 ms; 100x100: 455.8 against 470.9 ms).
    template <int J>
    __device__ __forceinline__ static void panel_update(double (&p0)[16], double (&p1)[16],
                                                        double l0, double l1, bool ws, int wl) {
        if constexpr (J < 16) {
            const double u = bcast(ws ? p1[J] : p0[J], wl);
            p0[J] = fma(-l0, u, p0[J]);
            p1[J] = fma(-l1, u, p1[J]);
            panel_update<J + 1>(p0, p1, l0, l1, ws, wl);
        }
    }

    // Elimination step C (compile time) of the panel starting at step k0,
    // whose pivot (win, rcp) is known. Lookahead: column C+1 is updated
    // first and the pivot of step C+1 found before the remaining columns
    // are updated, so the pivot search and the division overlap them.
    template <int C, int KQ, bool STORE_L>
    __device__ __forceinline__ static void panel_step(double (&p0)[16], double (&p1)[16],
                                                      bool& act0, bool& act1, int k0, double* lb,
                                                      double* lds, int& info, int lane, int win,
                                                      double rcp, int& sw, double& sr) {
        if constexpr (C < KQ) {
            int* ints = reinterpret_cast<int*>(lds + off_int);
            const int k = k0 + C;
            const double a0 = p0[C];
            const double a1 = p1[C];
            const int wl = win & 63;
            const bool ws = win >= 64;
            const bool ip0 = !ws && lane == wl;
            const bool ip1 = ws && lane == wl;
            const double l0 = (act0 && !ip0) ? a0 * rcp : 0.0;
            const double l1 = (act1 && !ip1) ? a1 * rcp : 0.0;
            p0[C] = act0 ? 0.0 : a0;
            p1[C] = act1 ? 0.0 : a1;
            if (ip0) act0 = false;
            if (ip1) act1 = false;
            if constexpr (STORE_L) {
                lb[lane * S + C] = -l0;
                if (lane + 64 < MP) lb[(lane + 64) * S + C] = -l1;
            }
            // step C's pivot row and reciprocal stay in lane C until the
            // panel ends (nothing reads them before)
            if (lane == C) {
                sw = win;
                sr = rcp;
            }
            if constexpr (C + 1 < 16) {
                const double u = bcast(ws ? p1[C + 1] : p0[C + 1], wl);
                p0[C + 1] = fma(-l0, u, p0[C + 1]);
                p1[C + 1] = fma(-l1, u, p1[C + 1]);
            }
            if constexpr (C + 1 < KQ) {
                double nrcp;
                const int nwin = pivot_of(p0[C + 1], p1[C + 1], act0, act1, k + 1, info, nrcp);
                panel_update<C + 2>(p0, p1, l0, l1, ws, wl);
                panel_step<C + 1, KQ, STORE_L>(p0, p1, act0, act1, k0, lb, lds, info, lane, nwin,
                                               nrcp, sw, sr);
            } else {
                panel_update<C + 2>(p0, p1, l0, l1, ws, wl);
            }
        }
    }

    // This wave's column tile (all row tiles) to the row block X, and back.
    __device__ __forceinline__ static void tile_to_rows(const dbl4 (&t)[RT], double* X, int lane) {
        const int col = lane & 15, r = lane >> 4;
#pragma unroll
        for (int T = 0; T < RT; ++T)
#pragma unroll
            for (int v = 0; v < 4; ++v) X[(16 * T + r + 4 * v) * S + col] = t[T][v];
    }
    __device__ __forceinline__ static void rows_to_tile(dbl4 (&t)[RT], const double* X, int lane) {
        const int col = lane & 15, r = lane >> 4;
#pragma unroll
        for (int T = 0; T < RT; ++T)
#pragma unroll
            for (int v = 0; v < 4; ++v) t[T][v] = X[(16 * T + r + 4 * v) * S + col];
    }
    __device__ __forceinline__ static void load_rows(double (&p0)[16], double (&p1)[16],
                                                     const double* X, int lane) {
#pragma unroll
        for (int j = 0; j < 16; ++j) p0[j] = X[lane * S + j];
#pragma unroll
        for (int j = 0; j < 16; ++j) p1[j] = 0.0;
        if (lane + 64 < MP) {
#pragma unroll
            for (int j = 0; j < 16; ++j) p1[j] = X[(lane + 64) * S + j];
        }
    }
    __device__ __forceinline__ static void store_rows(const double (&p0)[16],
                                                      const double (&p1)[16], double* X,
                                                      int lane) {
#pragma unroll
        for (int j = 0; j < 16; ++j) X[lane * S + j] = p0[j];
        if (lane + 64 < MP) {
#pragma unroll
            for (int j = 0; j < 16; ++j) X[(lane + 64) * S + j] = p1[j];
        }
    }

    // Factorisation of panel q by its owner (this wave): KQ steps; STORE_L
    // publishes -l. The factored panel goes back into the tile, and stays in
    // the row block X (the last panel's is read from there by the back
    // substitution).
    template <int KQ, bool STORE_L>
    __device__ __forceinline__ static void factor(dbl4 (&tile)[RT], int q, double* lds, int lane) {
        int* ints = reinterpret_cast<int*>(lds + off_int);
        double* X = lds + off_x;
        // The panel is the stencil's critical path: while it runs, this wave
        // takes issue priority on its SIMD over the trailing updates of the
        // other waves there (wave priority only; the stream's is untouched).
        if constexpr (PANEL_PRIO) __builtin_amdgcn_s_setprio(3);
        tile_to_rows(tile, X, lane);
        wave_lds_sync();
        double p0[16], p1[16];
        load_rows(p0, p1, X, lane);
        bool act0 = lane < NT && ints[int_step + lane] == NONE;
        bool act1 = lane + 64 < NT && ints[int_step + (lane + 64 < MP ? lane + 64 : 0)] == NONE;
        int info = 0;
        double rcp0;
        const int win0 = pivot_of(p0[0], p1[0], act0, act1, 16 * q, info, rcp0);
        int sw = 0;
        double sr = 0.0;
        panel_step<0, KQ, STORE_L>(p0, p1, act0, act1, 16 * q, lds + off_l + (q & 1) * XB, lds,
                                   info, lane, win0, rcp0, sw, sr);
        if (lane < KQ) {
            ints[int_perm + 16 * q + lane] = sw;
            ints[int_step + sw] = 16 * q + lane;
            lds[off_rcp + 16 * q + lane] = sr;
        }
        wave_lds_sync();
        store_rows(p0, p1, X, lane);
        if constexpr (STORE_L) {
            // the pivot rows' multipliers, row kk = pivot of step 16q + kk
            const double* lb = lds + off_l + (q & 1) * XB;
            double* l11 = lds + off_l11 + (q & 1) * 256;
            if (lane < 16) {
                const int prow = ints[int_perm + 16 * q + lane];
#pragma unroll
                for (int t = 0; t < 16; ++t) l11[lane * 16 + t] = lb[prow * S + t];
            }
        }
        wave_lds_sync();
        if constexpr (STORE_L) rows_to_tile(tile, X, lane);
        if (lane == 0 && info && ints[int_info] == 0) ints[int_info] = info;
        if constexpr (PANEL_PRIO) __builtin_amdgcn_s_setprio(0);
    }

    // ------------------------------------------------------------ trailing
    // U block of panel q at this wave's column tile: lane (col = l%16) gets
    // u[kk] of the panel's 16 pivot rows, by the oracle's chain of fma.
    __device__ __forceinline__ static void u_block(const dbl4 (&t)[RT], int q, double* lds,
                                                   double* ub, const double* lb, int lane,
                                                   double (&u)[16]) {
        const int* ints = reinterpret_cast<const int*>(lds + off_int);
        const int col = lane & 15, r = lane >> 4;
        // scatter this tile's entries of the panel's pivot rows
#pragma unroll
        for (int T = 0; T < RT; ++T)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int kk = ints[int_step + 16 * T + r + 4 * v] - 16 * q;
                if (kk >= 0 && kk < 16) ub[kk * 16 + col] = t[T][v];
            }
        wave_lds_sync();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) u[kk] = ub[kk * 16 + col];
        wave_lds_sync();  // reads done before this buffer is reused
        const double* l11 = lds + off_l11 + (q & 1) * 256;
#pragma unroll
        for (int kk = 1; kk < 16; ++kk) {
            const double* lr = l11 + kk * 16;
#pragma unroll
            for (int s = 0; s < kk; ++s) u[kk] = fma(lr[s], u[s], u[kk]);
            // computed here: left alone, the (pure) chains sink to their use
            // and every multiplier load stays live until then
            asm volatile("" : "+v"(u[kk]));
        }
    }

    // Trailing update of this wave's column tiles right of panel q, one
    // slot at a time (U block, then four chained MFMAs per row tile). The
    // owner of panel q + 1 (when it is not the last) factors it as soon as
    // that tile is updated, before its other tiles: the next panel's
    // factorisation is the critical path.
    __device__ __forceinline__ static void trailing(dbl4 (&tile)[LC][RT], int q, double* lds,
                                                    int w, int lane) {
        const double* lb = lds + off_l + (q & 1) * XB;
        double* ub = lds + off_u + w * 256;
        const int r = lane >> 4, col = lane & 15;
        const int jn = (q + 1) / W;  // slot of tile q + 1 in its owner
#pragma unroll
        for (int pass = 0; pass < 2; ++pass) {
#pragma unroll
            for (int j = 0; j < LC; ++j) {
                const int C = w + W * j;
                // pass 0: the next panel's tile (its owner only); pass 1: the rest
                const bool next = C == q + 1;
                if (C > q && C < CT && next == (pass == 0)) {
                    double bop[4];
                    {
                        double u[16];
                        u_block(tile[j], q, lds, ub, lb, lane, u);
#pragma unroll
                        for (int c = 0; c < 4; ++c)
                            bop[c] = r == 0 ? u[4 * c]
                                            : (r == 1 ? u[4 * c + 1]
                                                      : (r == 2 ? u[4 * c + 2] : u[4 * c + 3]));
                    }
#pragma unroll
                    for (int T = 0; T < RT; ++T) {
                        double a[4];
#pragma unroll
                        for (int c = 0; c < 4; ++c) a[c] = lb[(16 * T + col) * S + 4 * c + r];
#pragma unroll
                        for (int c = 0; c < 4; ++c)
                            tile[j][T] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[c], bop[c],
                                                                               tile[j][T], 0, 0, 0);
                    }
                    if (pass == 0 && q + 1 < RT - 1 && j == jn)
                        factor<16, true>(tile[j], q + 1, lds, lane);
                }
            }
        }
    }

    // ----------------------------------------------------- back substitution
    // Steps K0 + C, C = CN-1 .. 0, of one block whose rows are in X (one row
    // per lane, two above 64), updating the right-hand sides b0/b1.
    __device__ __forceinline__ static void back_block(const double* X, int K0, int CN,
                                                      double (&b0)[R], double (&b1)[R],
                                                      const double* lds, int lane) {
        const int* ints = reinterpret_cast<const int*>(lds + off_int);
        const bool two = lane + 64 < MP;
        for (int C = CN - 1; C >= 0; --C) {
            const int p = __builtin_amdgcn_readfirstlane(ints[int_perm + K0 + C]);
            const double rk = lds[off_rcp + K0 + C];
            const double v0 = X[lane * S + C];
            const double v1 = two ? X[(lane + 64) * S + C] : 0.0;
            const int pl = p & 63;
            const bool ps = p >= 64;
#pragma unroll
            for (int c = 0; c < R; ++c) {
                const double t = bcast(ps ? b1[c] : b0[c], pl) * rk;
                b0[c] = fma(-v0, t, b0[c]);
                b1[c] = fma(-v1, t, b1[c]);
            }
        }
    }

    // Back substitution of blocks (J*W + o, o = W-1 .. 0, only those below
    // the last) in wave BW; each block is passed through LDS by its owner
    // (alternating buffers, one barrier per block).
    template <int J, int BW>
    __device__ __forceinline__ static void back_slots(dbl4 (&tile)[LC][RT], double (&b0)[R],
                                                      double (&b1)[R], double* lds, int w,
                                                      int lane) {
        if constexpr (J >= 0) {
            for (int o = W - 1; o >= 0; --o) {
                const int qb = J * W + o;
                if (qb > RT - 2) continue;
                double* X = lds + off_l + (qb & 1) * XB;
                if (w == o) tile_to_rows(tile[J], X, lane);
                __syncthreads();
                if (w == BW) back_block(X, 16 * qb, 16, b0, b1, lds, lane);
            }
            back_slots<J - 1, BW>(tile, b0, b1, lds, w, lane);
        }
    }

    // Stencil s, whose CSR row starts at base (s*N for a fixed-width graph).
    // Coordinates of type Coord (double or float) are widened on load;
    // out(c) points at Coord.
    template <class Out, typename Coord>
    __device__ __forceinline__ static void run(const int s, const size_t base,
                                               const Coord* __restrict__ x,
                                               const Coord* __restrict__ y,
                                               nodeduce<const Coord*> __restrict__ z,
                                               const int* __restrict__ ja, Out out, double* lds,
                                               int* __restrict__ info) {
#pragma clang fp contract(off)
        const int w = static_cast<int>(threadIdx.x >> 6);
        const int lane = lane_id();
        const int g = w * 64 + lane;
        const double rho2 = gather<Coord>(s, base, x, y, z, ja, lds, g, w);
        const double rho = sqrt(rho2);
        dbl4 tile[LC][RT];
        {
            const int col = lane & 15, r = lane >> 4;
#pragma unroll
            for (int j = 0; j < LC; ++j) {
                const int C = w + W * j;
#pragma unroll
                for (int T = 0; T < RT; ++T)
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                        tile[j][T][v] =
                            C < CT ? elem(16 * T + r + 4 * v, 16 * C + col, lds) : 0.0;
            }
        }
        __syncthreads();  // points and monomials no longer needed
        if (w == 0) factor<16, true>(tile[0], 0, lds, lane);
        for (int q = 0; q < RT - 1; ++q) {
            __syncthreads();  // panel q published
            trailing(tile, q, lds, w, lane);
        }
        constexpr int BW = (RT - 1) % W;  // owner of the last panel and of the right-hand sides
        constexpr int JB = (RT - 1) / W;
        constexpr int K0 = 16 * (RT - 1);
        constexpr int cb = NT - K0;  // first right-hand side in the last tile
        if (w == BW) factor<cb, false>(tile[JB], RT - 1, lds, lane);
        __syncthreads();
        double b0[R], b1[R];
        const int* ints = reinterpret_cast<const int*>(lds + off_int);
        if (w == BW) {
            const double* X = lds + off_x;
            const bool two = lane + 64 < MP;
#pragma unroll
            for (int c = 0; c < R; ++c) {
                b0[c] = X[lane * S + cb + c];
                b1[c] = two ? X[(lane + 64) * S + cb + c] : 0.0;
            }
            back_block(X, K0, cb, b0, b1, lds, lane);
        }
        back_slots<LC - 1, BW>(tile, b0, b1, lds, w, lane);
        if (w == BW) {
#pragma unroll
            for (int set = 0; set < 2; ++set) {
                const int row = lane + 64 * set;
                if (row < NT) {
                    const int st = ints[int_step + row];
                    if (st < N) {
                        const double myrcp = lds[off_rcp + st];
                        const size_t at = base + st;
#pragma unroll
                        for (int c = 0; c < R; ++c) {
                            const double v = (set ? b1[c] : b0[c]) * myrcp;
                            out(c)[at] = static_cast<Coord>(c == 0 ? v / rho2 : v / rho);
                        }
                    }
                }
            }
        }
        if (info && g == 0) info[s] = ints[int_info];
    }
};

}  // namespace detail
}  // namespace rbf_operators

#endif  // RBF_TILE_H
