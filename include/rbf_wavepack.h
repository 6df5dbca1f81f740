// rbf_wavepack.h -- several small stencils per wave64 (gfx950): the wave is
// split into segments of SL = 16 or 32 lanes, one stencil per segment, one
// system row per lane (so 4 or 2 stencils per wave where the single-wave
// engine of rbf_wave.h leaves 49 or 37 of 64 lanes idle for a 15 x 15 or a
// 27 x 27 system).
//
// The arithmetic is that of WaveAssembly and of the oracle, operation for
// operation: the same fill, the same pivot (first row of largest |a| among
// rows not yet pivotal, NaN never chosen), rcp = 1 / pivot, l = a * rcp,
// a_ij = fma(-l, a_pj, a_ij) on every column to the right, rows never moved,
// back substitution t = b_p * rcp_p, b_i = fma(-u_ik, t, b_i). What changes
// is the bookkeeping, which has to be per segment:
//   - lane masks carry one bit per (segment, row); each segment's pivot is
//     found by a maximum over its own lanes (DPP within 16-lane rows, plus a
//     cross-row exchange for 32-lane segments) and a ballot, and its first
//     candidate lane by scalar bit scans of the ballot, segment by segment;
//   - the pivot row of each segment reaches the segment's lanes through LDS
//     (one slot per segment, two slots taken in turn), where the single-wave
//     engine reads it into scalar registers (v_readlane has one source lane
//     for the whole wave);
//   - the back substitution passes t_c through the same slots.
// Every lane computes its own pivot reciprocal (the reciprocal is no longer
// wave-uniform), which is the VALU cost the single-wave engine pays too.
#ifndef RBF_WAVEPACK_H
#define RBF_WAVEPACK_H

#include "rbf_wave.h"

namespace rbf_operators {
namespace detail {

// Maximum of v over the lanes of each segment of SL lanes, in every lane.
// Within a 16-lane row: quad swaps, half mirror, mirror (DPP); 32-lane
// segments then exchange rows 2i and 2i+1 by a lane-xor shuffle.
template <int SL>
__device__ __forceinline__ unsigned seg_umax(unsigned v) {
    v = dpp_max<0xB1, 0xF>(v);   // quad_perm [1,0,3,2]
    v = dpp_max<0x4E, 0xF>(v);   // quad_perm [2,3,0,1]
    v = dpp_max<0x141, 0xF>(v);  // row_half_mirror
    v = dpp_max<0x140, 0xF>(v);  // row_mirror
    if constexpr (SL == 32) v = max(v, static_cast<unsigned>(__shfl_xor(static_cast<int>(v), 16, 64)));
    return v;
}

// Maximum of a non-negative double over each segment, in every lane: the
// bit patterns of non-negative doubles order as the values do, so it is a
// maximum of the high words, then of the low words among the lanes that
// hold the top high word.
template <int SL>
__device__ __forceinline__ double seg_max_nonneg(double v) {
    const unsigned hi = static_cast<unsigned>(__double2hiint(v));
    const unsigned mh = seg_umax<SL>(hi);
    const unsigned lo = hi == mh ? static_cast<unsigned>(__double2loint(v)) : 0u;
    const unsigned ml = seg_umax<SL>(lo);
    return __hiloint2double(static_cast<int>(mh), static_cast<int>(ml));
}

template <int DIM, int N, int P, int Q, int R, bool INTERP, int SL, int CH = 0, int BCM = 0>
struct WavePack {
    static constexpr int M = npoly_dim(DIM, P);
    static constexpr int NT = N + M;
    static constexpr int NC = NT + R;
    static constexpr int S = 64 / SL;  // stencils per wave
    static constexpr int nsten = N;
    static_assert(SL == 16 || SL == 32, "segments of 16 or 32 lanes");
    static_assert(NT <= SL, "one system row per lane of a segment");
    static_assert(DIM == 2 || DIM == 3, "2-D or 3-D");
    static_assert(INTERP || R == DIM + 1, "derivative set: Laplacian + one per axis");
    static_assert(Q >= 3 && Q % 2 == 1, "odd PHS exponent >= 3");
    static constexpr unsigned long long SEGM = SL == 64 ? ~0ull : ((1ull << SL) - 1);
    // rows of every segment, as one wave-wide lane mask
    static constexpr unsigned long long ROWS = [] {
        unsigned long long m = 0;
        for (int g = 0; g < S; ++g) m |= ((1ull << NT) - 1) << (g * SL);
        return m;
    }();

    // LDS per segment (doubles): the scaled points interleaved [N][PD]
    // (x, y, z, pad), the monomials [M][N], the interpolation centres
    // [DIM][R] and their monomials [R][M]; once the rows are built the
    // same region holds two pivot-row slots of PS doubles (16-byte
    // aligned, one column every other double), which the back
    // substitution reuses for t.
    static constexpr int PD = DIM == 2 ? 2 : 4;
    static constexpr int o_pts = 0;
    static constexpr int o_poly = N * PD;
    static constexpr int o_cen = o_poly + M * N;
    static constexpr int o_cmono = o_cen + (INTERP ? DIM * R : 0);
    static constexpr int fill_doubles = o_cmono + (INTERP ? R * M : 0);
    // BCM: how the segment's lanes read the published pivot row.
    //   0: every lane reads every column at the segment's address (LDS
    //      broadcast: 64 lanes x 8 B per column and instruction);
    //   1: lane j of each 16-lane row reads column 16c + j of chunk c (one
    //      read per chunk of 16 columns), and column j reaches the row's
    //      lanes by DPP row_newbcast:j from lane j (v_mov_b64_dpp, one VALU
    //      instruction) -- far fewer LDS cycles for one VALU per column.
    static_assert(BCM == 0 || BCM == 1, "broadcast mode");
    static_assert(BCM == 0 || R <= 16, "t of the back substitution in one chunk");
    static constexpr int NCH = (NC + 15) / 16;  // chunks of 16 columns
    static constexpr int PS = BCM == 1 ? 32 * NCH : 2 * NC;  // a pivot-row slot (see slot_at)
    static_assert((R + 1) / 2 * 2 <= PS, "back-substitution slots fit the pivot slots");
    static constexpr int seg_doubles = (cmax(fill_doubles, 2 * PS) + 1) / 2 * 2;
    static constexpr int lds_doubles = S * seg_doubles;

    // ---------------------------------------------------------------- fill
    // This lane's row of [M | B]: lane li < N of the segment holds point row
    // li, lane N + t polynomial row t, lanes past NT zeros.
    template <typename C>
    __device__ __forceinline__ static void build_row(const int s, const int li, const C* __restrict__ x,
                                                     const C* __restrict__ y,
                                                     nodeduce<const C*> __restrict__ z,
                                                     const int* __restrict__ ja_row,
                                                     nodeduce<const C*> __restrict__ cx,
                                                     nodeduce<const C*> __restrict__ cy,
                                                     nodeduce<const C*> __restrict__ cz,
                                                     double* seg, double (&row)[NC], double& rho2_out) {
#pragma clang fp contract(off)
        double* pts = seg + o_pts;
        double* polyT = seg + o_poly;
        double* cen = seg + o_cen;
        double* cmono = seg + o_cmono;

        const int jn = li < N ? ja_row[li] : s;
        const double px = static_cast<double>(x[jn]) - static_cast<double>(x[s]);
        const double py = static_cast<double>(y[jn]) - static_cast<double>(y[s]);
        const double pz = DIM == 3 ? static_cast<double>(z[jn]) - static_cast<double>(z[s]) : 0.0;
        double d2 = px * px + py * py;
        if constexpr (DIM == 3) d2 = d2 + pz * pz;
        // the largest squared distance, NaN ignored (as the oracle's
        // "d2 > rho2" scan and the single-wave engine's fmax do)
        double rho2 = seg_max_nonneg<SL>(li < N && d2 == d2 ? d2 : 0.0);
        if (!(rho2 > 0.0)) rho2 = 1.0;
        const double rho = sqrt(rho2);
        rho2_out = rho2;
        const double xs = px / rho, ys = py / rho, zs = pz / rho;

        constexpr PolyBasis<P, DIM> basis{};
        double mono[M];
        basis(xs, ys, zs, mono, 1);
        if (li < N) {
            if constexpr (DIM == 2) {
                *reinterpret_cast<double2*>(pts + li * PD) = make_double2(xs, ys);
            } else {
                *reinterpret_cast<double2*>(pts + li * PD) = make_double2(xs, ys);
                pts[li * PD + 2] = zs;
            }
#pragma unroll
            for (int t = 0; t < M; ++t) polyT[t * N + li] = mono[t];
        }
        if constexpr (INTERP) {
            if (li < R) {
                const double ex = static_cast<double>(cx[li]) / rho;
                const double ey = static_cast<double>(cy[li]) / rho;
                const double ez = DIM == 3 ? static_cast<double>(cz[li]) / rho : 0.0;
                cen[li] = ex;
                cen[R + li] = ey;
                if constexpr (DIM == 3) cen[2 * R + li] = ez;
                double cm[M];
                basis(ex, ey, ez, cm, 1);
#pragma unroll
                for (int t = 0; t < M; ++t) cmono[li * M + t] = cm[t];
            }
        }
        wave_lds_sync();

        constexpr PHS<Q> phi{};
        if (li < N) {
#pragma unroll
            for (int col = 0; col < N; ++col) {
                const double2 q = *reinterpret_cast<const double2*>(pts + col * PD);
                const double dx = q.x - xs;
                const double dy = q.y - ys;
                double r2 = dx * dx + dy * dy;
                if constexpr (DIM == 3) {
                    const double dz = pts[col * PD + 2] - zs;
                    r2 = r2 + dz * dz;
                }
                row[col] = phi(r2);
                if ((col & 3) == 3) __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll
            for (int t = 0; t < M; ++t) row[N + t] = mono[t];
            if constexpr (!INTERP) {
                double r2 = xs * xs + ys * ys;
                if constexpr (DIM == 3) r2 = r2 + zs * zs;
                const double tq = PHS<Q - 2>{}(r2);
                const double g = static_cast<double>(Q) * tq;
                row[NT + 0] = static_cast<double>(Q * (Q + DIM - 2)) * tq;
                row[NT + 1] = -g * xs;
                row[NT + 2] = -g * ys;
                if constexpr (DIM == 3) row[NT + 3] = -g * zs;
            } else {
#pragma unroll
                for (int c = 0; c < R; ++c) {
                    const double dx = cen[c] - xs;
                    const double dy = cen[R + c] - ys;
                    double r2 = dx * dx + dy * dy;
                    if constexpr (DIM == 3) {
                        const double dz = cen[2 * R + c] - zs;
                        r2 = r2 + dz * dz;
                    }
                    row[NT + c] = phi(r2);
                }
            }
        } else if (li < NT) {
            const int t = li - N;
#pragma unroll
            for (int col = 0; col < N; ++col) row[col] = polyT[t * N + col];
#pragma unroll
            for (int c = N; c < NT; ++c) row[c] = 0.0;
            if constexpr (!INTERP) {
                using Mk = PolyDerivMasks<DIM, P>;
                row[NT + 0] = ((Mk::lap >> t) & 1u) ? 2.0 : 0.0;
                row[NT + 1] = ((Mk::dx >> t) & 1u) ? 1.0 : 0.0;
                row[NT + 2] = ((Mk::dy >> t) & 1u) ? 1.0 : 0.0;
                if constexpr (DIM == 3) row[NT + 3] = ((Mk::dz >> t) & 1u) ? 1.0 : 0.0;
            } else {
#pragma unroll
                for (int c = 0; c < R; ++c) row[NT + c] = cmono[c * M + t];
            }
        } else {
#pragma unroll
            for (int c = 0; c < NC; ++c) row[c] = 0.0;
        }
        wave_lds_sync();  // points and monomials read: the pivot slots overwrite them
    }

    // -------------------------------------------------------- elimination
    // Lowest and highest bit of every segment, for word-parallel (SWAR)
    // operations on the 64-bit lane masks.
    static constexpr unsigned long long LOWB = [] {
        unsigned long long m = 0;
        for (int g = 0; g < S; ++g) m |= 1ull << (g * SL);
        return m;
    }();
    static constexpr unsigned long long HIGHB = LOWB << (SL - 1);

    // some segment of x holds no set bit (x - LOWB borrows into a segment's
    // top bit exactly where the segment is zero: the classic zero-lane test)
    __host__ __device__ __forceinline__ static bool has_empty_segment(unsigned long long x) {
        return ((x - LOWB) & ~x & HIGHB) != 0;
    }

    // Per-segment pivot lane mask of step K: the first lane of each
    // segment among the hit lanes, or, for a segment with no candidate
    // (every active entry NaN), its first active lane, that segment then
    // marked in nocand. (With every segment nonempty, hit & (~hit + LOWB)
    // gives the same mask: the carries stop inside each segment.)
    __host__ __device__ __forceinline__ static unsigned long long first_per_segment(
        unsigned long long hit, unsigned long long act, unsigned long long& nocand) {
        unsigned long long pm = 0;
        nocand = 0;
#pragma unroll
        for (int g = 0; g < S; ++g) {
            const unsigned long long hg = (hit >> (g * SL)) & SEGM;
            const unsigned long long ag = (act >> (g * SL)) & SEGM;
            const unsigned long long pick = hg ? hg : ag;
            pm |= (pick & (~pick + 1)) << (g * SL);
            if (!hg) nocand |= SEGM << (g * SL);
        }
        return pm;
    }

    __host__ __device__ __forceinline__ static bool some_segment_tied(unsigned long long hit) {
        bool tied = false;
#pragma unroll
        for (int g = 0; g < S; ++g) {
            const unsigned long long hg = (hit >> (g * SL)) & SEGM;
            tied = tied || (hg & (hg - 1)) != 0;
        }
        return tied;
    }

    // The two mask steps of a pivot search, from the ballot `hit` of the
    // lanes holding their segment's top key: whether some segment has more
    // than one such lane (then the low words decide), and the pivot lane of
    // every segment (nocand: the segments with no candidate). The usual
    // case, a candidate in every segment, takes a few scalar operations on
    // the whole mask; the segment-by-segment loops only run otherwise.
    __host__ __device__ __forceinline__ static bool segments_tied(unsigned long long hit) {
        return !has_empty_segment(hit) ? (hit & (hit - LOWB)) != 0 : some_segment_tied(hit);
    }
    __host__ __device__ __forceinline__ static unsigned long long pivot_lanes(
        unsigned long long hit, unsigned long long act, unsigned long long& nocand) {
        nocand = 0;
        return !has_empty_segment(hit) ? hit & (~hit + LOWB) : first_per_segment(hit, act, nocand);
    }

    // Slot position of pivot-row column j: every other double, so no two
    // columns are adjacent in LDS and the compiler pairs them as two 8-byte
    // accesses (ds_write2_b64 / ds_read2_b64) from any two registers, not
    // as one 16-byte access that needs the pair in adjacent registers (which
    // costs register copies on the pivot lanes' stores).
    static constexpr int slot_at(int j) { return 2 * j; }

    // pivot row (columns K .. NC-1) into this segment's slot; every lane of
    // the wave executes the stores, the pivot lanes' are the ones kept
    template <int K, int J>
    __device__ __forceinline__ static void publish(double* pb, const double (&row)[NC]) {
        if constexpr (J < NC) {
            pb[slot_at(J)] = row[J];
            publish<K, J + 1>(pb, row);
        }
    }

    // row[j] = fma(-l, u_j, row[j]) for j = J .. NC-1 (J > K), the pivot
    // row's u_j read from the slot; a scheduling barrier after every CH
    // columns (CH > 0) keeps the compiler from fetching the whole pivot row
    // ahead of its FMAs
    template <int K, int J>
    __device__ __forceinline__ static void update(double (&row)[NC], double l, const double* pb) {
        if constexpr (J < NC) {
            row[J] = fma(-l, pb[slot_at(J)], row[J]);
            if constexpr (CH > 0 && (J + 1) % CH == 0) __builtin_amdgcn_sched_barrier(0);
            update<K, J + 1>(row, l, pb);
        }
    }

    // value of v in lane n of this lane's 16-lane row (DPP row_newbcast)
    template <int n>
    __device__ __forceinline__ static double row_bcast(double v) {
        return __builtin_amdgcn_update_dpp(0.0, v, 0x150 + n, 0xF, 0xF, false);
    }

    // BCM = 1: this lane's columns of every chunk that still holds a column
    // >= K0 (lane j of a row reads column 16c + j)
    template <int K0, int C>
    __device__ __forceinline__ static void read_chunks(double (&uc)[NCH], const double* pb, int lj) {
        if constexpr (C < NCH) {
            if constexpr (16 * C + 15 >= K0) uc[C] = pb[slot_at(16 * C + lj)];
            read_chunks<K0, C + 1>(uc, pb, lj);
        }
    }

    // row[j] = fma(-l, u_j, row[j]) for j = J .. NC-1, u_j by row_newbcast
    template <int J>
    __device__ __forceinline__ static void update_dpp(double (&row)[NC], double l,
                                                      const double (&uc)[NCH]) {
        if constexpr (J < NC) {
            row[J] = fma(-l, row_bcast<J % 16>(uc[J / 16]), row[J]);
            if constexpr (CH > 0 && (J + 1) % CH == 0) __builtin_amdgcn_sched_barrier(0);
            update_dpp<J + 1>(row, l, uc);
        }
    }

    // BCM = 1 back substitution: b_c = fma(-u, t_c, b_c), t_c from lane c
    template <int C>
    __device__ __forceinline__ static void back_dpp(double (&row)[NC], double u, double tv) {
        if constexpr (C < R) {
            row[NT + C] = fma(-u, row_bcast<C>(tv), row[NT + C]);
            back_dpp<C + 1>(row, u, tv);
        }
    }

    template <int K>
    __device__ __forceinline__ static void elim(double (&row)[NC], unsigned long long& done,
                                                int& mystep, double& myrcp, int& info,
                                                double* slots) {
        if constexpr (K < NT) {
            const unsigned long long act = ROWS & ~done;
            const double a = row[K];
            // pivot search, every segment at once: largest |a| by its high
            // word, then (only where a segment has a tie) by its low word
            const unsigned long long cm = act & wave_ballot(a == a);  // not NaN
            const double fa = fabs(a);
            const unsigned hi = lane_in(cm) ? static_cast<unsigned>(__double2hiint(fa)) : 0u;
            const unsigned mh = seg_umax<SL>(hi);
            unsigned long long hit = cm & wave_ballot(hi == mh);
            if (segments_tied(hit)) {
                const unsigned lo = lane_in(hit) ? static_cast<unsigned>(__double2loint(fa)) : 0u;
                const unsigned ml = seg_umax<SL>(lo);
                hit &= wave_ballot(lo == ml);
            }
            unsigned long long nocand;
            const unsigned long long pm = pivot_lanes(hit, act, nocand);
            if (lane_in(nocand) && !info) info = K + 1;
            const bool is_piv = lane_in(pm);
            double* pb = slots + (K & 1) * PS;
            if (is_piv) publish<K, K>(pb, row);
            wave_lds_sync();
            double uc[NCH];
            double piv;
            if constexpr (BCM == 1) {
                read_chunks<K, 0>(uc, pb, lane_id() & 15);
                piv = row_bcast<K % 16>(uc[K / 16]);
            } else {
                piv = pb[slot_at(K)];
            }
            if (piv == 0.0 && !info) info = K + 1;
            const double rcp = 1.0 / piv;
            const bool elim_row = lane_in(act & ~pm);
            const double l = elim_row ? a * rcp : 0.0;
            row[K] = lane_in(act) ? 0.0 : a;
            mystep = is_piv ? K : mystep;
            myrcp = is_piv ? rcp : myrcp;
            done |= pm;
            if constexpr (BCM == 1)
                update_dpp<K + 1>(row, l, uc);
            else
                update<K, K + 1>(row, l, pb);
            elim<K + 1>(row, done, mystep, myrcp, info, slots);
        }
    }

    // Elimination, then back substitution: lane i ends with the solution
    // of unknown mystep in xout. Returns this segment's info.
    __device__ __forceinline__ static int lu_solve(double (&row)[NC], int& mystep, double (&xout)[R],
                                                   double* slots) {
#pragma clang fp contract(off)
        unsigned long long done = 0;
        int info = 0;
        double myrcp = 0.0;
        mystep = 1 << 20;
        elim<0>(row, done, mystep, myrcp, info, slots);
        wave_lds_sync();  // the last pivot row read before the slots take t
        constexpr int TS = (R + 1) / 2 * 2;
#pragma unroll
        for (int k = NT - 1; k >= 0; --k) {
            double* tb = slots + (k & 1) * PS;
            if (mystep == k) {
#pragma unroll
                for (int c = 0; c < R; ++c) tb[c] = row[NT + c] * myrcp;
            }
            wave_lds_sync();
            if constexpr (BCM == 1) {
                back_dpp<0>(row, row[k], tb[lane_id() & 15]);
            } else {
#pragma unroll
                for (int c = 0; c < R; c += 2) {
                    if (c + 1 < R) {
                        const double2 t = *reinterpret_cast<const double2*>(tb + c);
                        row[NT + c] = fma(-row[k], t.x, row[NT + c]);
                        row[NT + c + 1] = fma(-row[k], t.y, row[NT + c + 1]);
                    } else {
                        row[NT + c] = fma(-row[k], tb[c], row[NT + c]);
                    }
                }
            }
            static_assert(TS <= PS, "slots");
        }
        wave_lds_sync();  // slots read before the caller reuses LDS
#pragma unroll
        for (int c = 0; c < R; ++c) xout[c] = row[NT + c] * myrcp;
        return info;
    }

    // Stencil s of this lane's segment, whose CSR row starts at base; a
    // segment with valid false repeats a real stencil and stores nothing.
    template <class Out, typename C>
    __device__ __forceinline__ static void run_one(const int s, const size_t base, const bool valid,
                                                   const C* __restrict__ x, const C* __restrict__ y,
                                                   nodeduce<const C*> __restrict__ z,
                                                   const int* __restrict__ ja, Out out,
                                                   nodeduce<const C*> __restrict__ cx,
                                                   nodeduce<const C*> __restrict__ cy,
                                                   nodeduce<const C*> __restrict__ cz, double* lds,
                                                   int* __restrict__ info,
                                                   uint8_t* __restrict__ status) {
#pragma clang fp contract(off)
        const int lane = lane_id();
        const int seg = lane / SL;
        const int li = lane & (SL - 1);
        double* segl = lds + seg * seg_doubles;
        double row[NC];
        double rho2;
        build_row<C>(s, li, x, y, z, ja + base, cx, cy, cz, segl, row, rho2);
        int mystep;
        double xo[R];
        const int inf = lu_solve(row, mystep, xo, segl);
        if (valid && mystep < N) {
            const double rho = sqrt(rho2);
            const size_t at = base + mystep;
#pragma unroll
            for (int c = 0; c < R; ++c) {
                double v = xo[c];
                if constexpr (!INTERP) v = c == 0 ? v / rho2 : v / rho;
                out(c)[at] = static_cast<C>(v);
            }
        }
        if (valid && li == 0) {
            if (info) info[s] = inf;
            if (status) status[s] = 0;
        }
    }

    // Stencils s0 .. s0 + S - 1 (one per segment; segments past the last
    // stencil repeat it and store nothing), fixed-width rows of N.
    template <class Out, typename C>
    __device__ __forceinline__ static void run(const int s0, const int nstencils,
                                               const C* __restrict__ x, const C* __restrict__ y,
                                               nodeduce<const C*> __restrict__ z,
                                               const int* __restrict__ ja, Out out,
                                               nodeduce<const C*> __restrict__ cx,
                                               nodeduce<const C*> __restrict__ cy,
                                               nodeduce<const C*> __restrict__ cz, double* lds,
                                               int* __restrict__ info, uint8_t* __restrict__ status) {
        const int sw = s0 + lane_id() / SL;
        const bool valid = sw < nstencils;
        const int s = valid ? sw : nstencils - 1;
        run_one(s, static_cast<size_t>(s) * N, valid, x, y, z, ja, out, cx, cy, cz, lds, info,
                status);
    }

    // Entries i0 .. i0 + S - 1 of a stencil list (order[i] names the
    // stencil, ia[s] its CSR row start; entries past count repeat the last
    // and store nothing): the per-class lists of a ragged batch.
    template <class Out, typename C>
    __device__ __forceinline__ static void run_list(const int i0, const int count,
                                                    const int* __restrict__ order,
                                                    const int* __restrict__ ia,
                                                    const C* __restrict__ x, const C* __restrict__ y,
                                                    nodeduce<const C*> __restrict__ z,
                                                    const int* __restrict__ ja, Out out,
                                                    double* lds, int* __restrict__ info,
                                                    uint8_t* __restrict__ status) {
        const int i = i0 + lane_id() / SL;
        const bool valid = i < count;
        const int s = order[valid ? i : count - 1];
        run_one(s, static_cast<size_t>(ia[s]), valid, x, y, z, ja, out, nullptr, nullptr, nullptr,
                lds, info, status);
    }
};

}  // namespace detail
}  // namespace rbf_operators

#endif  // RBF_WAVEPACK_H
