// rbf_wave.h -- the single-wave RBF-FD stencil engine for AMD gfx950
// (MI355X), and the basis and fill routines the drop-in header
// rbf_operators.h exports under the reference's names:
//
//   PHS<Q>, npoly, PolyBasis<P>               rbf_operators.h:39-114
//   fill_matrix<N,P,Q>, fill_rhs<N,P,NRHS,Q>  rbf_operators.h:116-185
//
// The reference solves each stencil with cuSolverDx's block GESV; here a
// wavefront LU written for CDNA4 does it: one wave64 per stencil, one
// matrix row per lane held in VGPRs together with its right-hand sides,
// pivot search by a wave max-reduction plus ballot, the pivot row broadcast
// lane-to-SGPR (v_readlane) straight into the FMA operands, and rows never
// moved (a lane that supplied a pivot is retired; the permutation lives in
// scalar registers).
//
// Differences from the reference, on purpose:
//   - the system is built in coordinates shifted to the stencil's node AND
//     scaled by the stencil radius rho = max_k |p_k - p_s|; interpolation
//     weights are scale-free, derivative weights are unscaled by rho^-1
//     (first derivatives) and rho^-2 (Laplacian). Without the scaling the
//     PHS systems at h = 1/1024 are numerically singular in FP64;
//   - fills round each product and sum on its own (no FMA contraction) and
//     the elimination uses explicit fma(), in the same order as the CPU
//     oracle, so GPU and CPU results are bit-identical;
//   - the engine computes in FP64 whatever the coordinate and weight type
//     (float inputs are widened on load, weights rounded on store).
#ifndef RBF_WAVE_H
#define RBF_WAVE_H

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <type_traits>

#include "gfx_arch.h"

namespace rbf_operators {

// ---------------------------------------------------------------------------
// Basis functions
// ---------------------------------------------------------------------------

// phi(r) = r^Q from r^2, Q odd. The reference accepts Q >= 3
// (rbf_operators.h:41); Q = 1 (plain r) is used here for the derivative
// right-hand sides, which need r^(Q-2).
template <int Q>
struct PHS {
    static_assert(Q >= 1 && Q % 2 == 1, "polyharmonic splines need an odd exponent");
    template <typename T>
    __host__ __device__ inline T operator()(T r2) const {
        const T r = sqrt(r2);
        T p = r;
        for (int i = 0; i < (Q - 1) / 2; ++i) p *= r2;
        return p;
    }
};
template <>
struct PHS<1> {
    template <typename T>
    __host__ __device__ inline T operator()(T r2) const { return sqrt(r2); }
};
template <>
struct PHS<3> {
    template <typename T>
    __host__ __device__ inline T operator()(T r2) const { return r2 * sqrt(r2); }
};
template <>
struct PHS<5> {
    template <typename T>
    __host__ __device__ inline T operator()(T r2) const { return r2 * r2 * sqrt(r2); }
};
template <>
struct PHS<7> {
    template <typename T>
    __host__ __device__ inline T operator()(T r2) const {
        const T r4 = r2 * r2;
        return r4 * r2 * sqrt(r2);
    }
};

// number of 2-D monomials of degree <= P
constexpr int npoly(int P) { return (P + 1) * (P + 2) / 2; }
// number of DIM-D monomials of degree <= P (DIM = 2 or 3)
constexpr int npoly_dim(int DIM, int P) {
    return DIM == 2 ? npoly(P) : (P + 1) * (P + 2) * (P + 3) / 6;
}

// Monomials in degree-major order: 2-D x^(d-j) y^j for j = 0..d
// (rbf_operators.h:93-98); 3-D x^(d-j) y^(j-l) z^l for j = 0..d, l = 0..j,
// formed as (X[d-j] * Y[j-l]) * Z[l] -- the 2-D order when z is absent.
template <int P, int DIM = 2>
struct PolyBasis {
    static constexpr int np = npoly_dim(DIM, P);
    template <typename T>
    __host__ __device__ inline void operator()(T x, T y, T* b, int inc = 1) const {
        static_assert(DIM == 2, "the 2-D basis takes (x, y)");
        T X[P + 1], Y[P + 1];
        X[0] = Y[0] = T(1);
        for (int i = 1; i <= P; ++i) {
            X[i] = X[i - 1] * x;
            Y[i] = Y[i - 1] * y;
        }
        int k = 0;
        for (int d = 0; d <= P; ++d)
            for (int j = 0; j <= d; ++j) b[(k++) * inc] = X[d - j] * Y[j];
    }
    template <typename T>
    __host__ __device__ inline void operator()(T x, T y, T z, T* b, int inc = 1) const {
        if constexpr (DIM == 2) {
            (*this)(x, y, b, inc);
        } else {
            T X[P + 1], Y[P + 1], Z[P + 1];
            X[0] = Y[0] = Z[0] = T(1);
            for (int i = 1; i <= P; ++i) {
                X[i] = X[i - 1] * x;
                Y[i] = Y[i - 1] * y;
                Z[i] = Z[i - 1] * z;
            }
            int k = 0;
            for (int d = 0; d <= P; ++d)
                for (int j = 0; j <= d; ++j)
                    for (int l = 0; l <= j; ++l) b[(k++) * inc] = (X[d - j] * Y[j - l]) * Z[l];
        }
    }
};

// Which monomials have a nonzero derivative at the origin: bit t set when
// monomial t is x^2, y^2 or z^2 (Laplacian 2), x, y or z (first derivative 1).
template <int DIM, int P>
struct PolyDerivMasks {
    static constexpr uint64_t bit_of(int ex, int ey, int ez) {
        int k = 0;
        for (int d = 0; d <= P; ++d) {
            if (DIM == 2) {
                for (int j = 0; j <= d; ++j, ++k)
                    if (ex == d - j && ey == j && ez == 0) return uint64_t(1) << k;
            } else {
                for (int j = 0; j <= d; ++j)
                    for (int l = 0; l <= j; ++l, ++k)
                        if (ex == d - j && ey == j - l && ez == l) return uint64_t(1) << k;
            }
        }
        return 0;
    }
    static constexpr uint64_t lap = bit_of(2, 0, 0) | bit_of(0, 2, 0) | (DIM == 3 ? bit_of(0, 0, 2) : 0);
    static constexpr uint64_t dx = bit_of(1, 0, 0);
    static constexpr uint64_t dy = bit_of(0, 1, 0);
    static constexpr uint64_t dz = DIM == 3 ? bit_of(0, 0, 1) : 0;
    static_assert(npoly_dim(DIM, P) <= 64, "masks hold at most 64 monomials");
};

// ---------------------------------------------------------------------------
// Reference fill routines, same signatures and layout as rbf_operators.h
// (column-major M, ldm = nt), for callers that fill their own systems.
// ---------------------------------------------------------------------------

template <int N, int P, int Q = 3, typename T>
__host__ __device__ void fill_matrix(T* M, const T* xs, const T* ys, int tid, int nthreads) {
#pragma clang fp contract(off)
    constexpr PolyBasis<P> monomials{};
    constexpr int NP = decltype(monomials)::np;
    constexpr int ldm = N + NP;
    constexpr PHS<Q> phi{};
    for (int col = 0; col < N; ++col) {
        const T xc = xs[col], yc = ys[col];
        for (int k = tid; k < N; k += nthreads) {
            const T dx = xc - xs[k];
            const T dy = yc - ys[k];
            M[k + col * ldm] = phi(dx * dx + dy * dy);
        }
    }
    for (int k = tid; k < N; k += nthreads) {
        monomials(xs[k], ys[k], &M[k + N * ldm], ldm);
        monomials(xs[k], ys[k], &M[N + k * ldm], 1);
    }
    for (int k = tid; k < NP * NP; k += nthreads) M[(N + k % NP) + (N + k / NP) * ldm] = T(0);
}

template <int N, int P, int NRHS, int Q = 3, typename T>
__host__ __device__ void fill_rhs(T* B, const T* xs, const T* ys, const T* xc, const T* yc,
                                  int tid, int nthreads) {
#pragma clang fp contract(off)
    constexpr PolyBasis<P> monomials{};
    constexpr int NP = decltype(monomials)::np;
    constexpr int ldb = N + NP;
    constexpr PHS<Q> phi{};
    for (int col = 0; col < NRHS; ++col) {
        const T xe = xc[col], ye = yc[col];
        for (int k = tid; k < N; k += nthreads) {
            const T dx = xe - xs[k], dy = ye - ys[k];
            B[k + col * ldb] = phi(dx * dx + dy * dy);
        }
    }
    for (int col = tid; col < NRHS; col += nthreads) monomials(xc[col], yc[col], &B[N + col * ldb], 1);
}

namespace detail {

// ---------------------------------------------------------------------------
// wave64 primitives
// ---------------------------------------------------------------------------

__device__ __forceinline__ int lane_id() { return static_cast<int>(threadIdx.x & 63u); }

constexpr int cmax(int a, int b) { return a > b ? a : b; }

// keeps a parameter out of template argument deduction (C++20's
// std::type_identity): the coordinate type is deduced from x alone, so z and
// the centres may be passed as nullptr
template <class T>
struct same_as {
    using type = T;
};
template <class T>
using nodeduce = typename same_as<T>::type;

// value of v in lane l (l wave-uniform): two v_readlane_b32 into an SGPR pair
__device__ __forceinline__ double bcast(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), l);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), l);
    return __longlong_as_double((static_cast<long long>(hi) << 32) |
                                static_cast<unsigned int>(lo));
}
__device__ __forceinline__ float bcast(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// A value equal in every lane, moved to scalar registers: it then costs no
// vector register while it waits (the stencil's index, CSR base and
// radius live from the gather to the scatter).
__device__ __forceinline__ double wave_uniform(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane(static_cast<int>(b));
    const int hi = __builtin_amdgcn_readfirstlane(static_cast<int>(b >> 32));
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}
__device__ __forceinline__ size_t wave_uniform(size_t v) {
    const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<int>(v)));
    const unsigned hi =
        static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<int>(v >> 32)));
    return (static_cast<size_t>(hi) << 32) | lo;
}

// 1 / b for a wave-uniform FP64 b, bit-identical to the IEEE division
// 1.0 / b. hipcc expands that division as v_div_scale (operand scaling), a
// v_rcp_f64 seed r, two Newton-Raphson steps r = r + r (1 - b r), the
// correction q = r + r (1 - b r) (v_div_fmas) and v_div_fixup (special
// operands). For 2^-700 <= |b| < 2^701 the scaling is the identity (the
// exponent gap to 1.0 stays below 768, b and 1/b are normal) and the fixup
// returns q, so the same seed and the same six FMAs in the same order give
// the same bits, four VALU instructions fewer. The exponent test runs on
// scalar registers; any other b (0, inf, NaN, extreme exponents) takes the
// division itself.
__device__ __forceinline__ double recip_uniform(double b) {
#pragma clang fp contract(off)
    const unsigned hi = static_cast<unsigned>(
        __builtin_amdgcn_readfirstlane(static_cast<int>(__double_as_longlong(b) >> 32)));
    const unsigned ex = (hi >> 20) & 0x7ffu;
    if (ex - (1023u - 700u) <= 1400u) {
        double r = __builtin_amdgcn_rcp(b);
        double e = fma(-b, r, 1.0);
        r = fma(r, e, r);
        e = fma(-b, r, 1.0);
        r = fma(r, e, r);
        e = fma(-b, r, 1.0);
        return fma(e, r, r);
    }
    return 1.0 / b;
}
__device__ __forceinline__ float recip_uniform(float b) { return 1.0f / b; }

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// Unsigned max over the wave: DPP within each 16-lane row (quad swaps,
// half mirror, mirror), then the four row maxima through SGPRs
// (v_readlane + s_max_u32). The result is wave-uniform.
__device__ __forceinline__ unsigned wave_umax(unsigned v) {
    v = max(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0xB1, 0xF, 0xF, false)));
    v = max(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x4E, 0xF, 0xF, false)));
    v = max(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x141, 0xF, 0xF, false)));
    v = max(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x140, 0xF, 0xF, false)));
    const unsigned a = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(v), 0));
    const unsigned b = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(v), 16));
    const unsigned c = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(v), 32));
    const unsigned d = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(v), 48));
    return max(max(a, b), max(c, d));
}

// The same maximum all in DPP: after the in-row steps, row 0 into row 1
// and row 2 into row 3 (row_bcast:15), then rows 0-1 into rows 2-3
// (row_bcast:31); lane 63 ends with the maximum, read into an SGPR.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ unsigned dpp_max(unsigned v) {
    // rows outside ROW_MASK read 0, the identity of the maximum
    return max(v, static_cast<unsigned>(
                      __builtin_amdgcn_update_dpp(0, static_cast<int>(v), CTRL, ROW_MASK, 0xF, false)));
}
__device__ __forceinline__ unsigned wave_umax_bcast(unsigned v) {
    v = dpp_max<0xB1, 0xF>(v);   // quad_perm [1,0,3,2]
    v = dpp_max<0x4E, 0xF>(v);   // quad_perm [2,3,0,1]
    v = dpp_max<0x141, 0xF>(v);  // row_half_mirror
    v = dpp_max<0x140, 0xF>(v);  // row_mirror
    v = dpp_max<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
    v = dpp_max<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
    return static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(v), 63));
}

// Lanes where pred holds, as a wave-uniform 64-bit mask.
__device__ __forceinline__ unsigned long long wave_ballot(bool pred) {
    return __builtin_amdgcn_ballot_w64(pred);
}

// A wave-uniform 64-bit lane mask as this lane's predicate: the mask itself
// becomes the condition register, no vector instruction.
__device__ __forceinline__ bool lane_in(unsigned long long m) {
    return __builtin_amdgcn_inverse_ballot_w64(m);
}

// Lowest lane, among the candidates, of largest |a| (NaN never wins):
// the same choice as a serial "first strict maximum" scan in lane order.
// Returns -1 when there is no candidate. PV selects the reduction:
// 0 = shuffle max of |a| then ballot; 1 = integer max on the bit pattern of
// |a| (high word, then low word only when the high words tie) with the
// row maxima combined in SGPRs; 2 = the same with the rows combined in DPP.
template <int PV>
__device__ __forceinline__ unsigned pivot_umax(unsigned v) {
    if constexpr (PV == 2) return wave_umax_bcast(v);
    else return wave_umax(v);
}
template <int PV>
__device__ __forceinline__ unsigned long long pivot_ballot(bool pred) {
    if constexpr (PV == 2) return wave_ballot(pred);
    else return __ballot(pred);
}
template <int PV>
__device__ __forceinline__ int pivot_lane(double a, bool cand) {
    cand = cand && !__builtin_isnan(a);
    if constexpr (PV == 0) {
        const double c = cand ? fabs(a) : -1.0;
        const double mx = wave_max(c);
        const unsigned long long hit = __ballot(cand && c == mx);
        return hit ? __ffsll(hit) - 1 : -1;
    } else {
        const double fa = fabs(a);
        const unsigned hi = cand ? static_cast<unsigned>(__double2hiint(fa)) : 0u;
        const unsigned mh = pivot_umax<PV>(hi);
        unsigned long long hit = pivot_ballot<PV>(cand && hi == mh);
        if (__popcll(hit) > 1) {
            const bool top = cand && hi == mh;
            const unsigned lo = top ? static_cast<unsigned>(__double2loint(fa)) : 0u;
            const unsigned ml = pivot_umax<PV>(lo);
            hit = pivot_ballot<PV>(top && lo == ml);
        }
        return hit ? __ffsll(hit) - 1 : -1;
    }
}
template <int PV>
__device__ __forceinline__ int pivot_lane(float a, bool cand) {
    cand = cand && !__builtin_isnan(a);
    if constexpr (PV == 0) {
        const float c = cand ? fabsf(a) : -1.f;
        const float mx = wave_max(c);
        const unsigned long long hit = __ballot(cand && c == mx);
        return hit ? __ffsll(hit) - 1 : -1;
    } else {
        const unsigned key = cand ? __float_as_uint(fabsf(a)) : 0u;
        const unsigned mk = pivot_umax<PV>(key);
        const unsigned long long hit = pivot_ballot<PV>(cand && key == mk);
        return hit ? __ffsll(hit) - 1 : -1;
    }
}

// The same choice from a wave-uniform mask of the candidate rows. Every
// ballot is of a plain compare and the masks are combined in scalar
// registers (a ballot of a combined predicate costs a select and a compare
// on the vector unit).
template <int PV>
__device__ __forceinline__ int pivot_lane_mask(double a, unsigned long long act) {
    if constexpr (PV == 0) {
        return pivot_lane<0>(a, lane_in(act));
    } else {
        const unsigned long long cm = act & wave_ballot(a == a);  // not NaN
        const double fa = fabs(a);
        const unsigned hi = lane_in(cm) ? static_cast<unsigned>(__double2hiint(fa)) : 0u;
        const unsigned mh = pivot_umax<PV>(hi);
        unsigned long long hit = cm & wave_ballot(hi == mh);
        if (__popcll(hit) > 1) {
            const unsigned lo = lane_in(hit) ? static_cast<unsigned>(__double2loint(fa)) : 0u;
            const unsigned ml = pivot_umax<PV>(lo);
            hit &= wave_ballot(lo == ml);
        }
        return hit ? __ffsll(hit) - 1 : -1;
    }
}
template <int PV>
__device__ __forceinline__ int pivot_lane_mask(float a, unsigned long long act) {
    if constexpr (PV == 0) {
        return pivot_lane<0>(a, lane_in(act));
    } else {
        const unsigned long long cm = act & wave_ballot(a == a);
        const unsigned key = lane_in(cm) ? __float_as_uint(fabsf(a)) : 0u;
        const unsigned mk = pivot_umax<PV>(key);
        const unsigned long long hit = cm & wave_ballot(key == mk);
        return hit ? __ffsll(hit) - 1 : -1;
    }
}

// Orders this wave's LDS writes before its later LDS reads (LDS operations
// of one wave complete in order; this keeps the compiler from moving them).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Pivot-row columns routed through LDS (the rest go lane-to-SGPR):
// MODE 0 none, 1 all, 2 every other group of VW columns (VW = 16 bytes),
// 3 two groups of every three, 4 one group of every three.
template <int MODE, int VW>
__device__ constexpr bool via_lds(int j) {
    return MODE == 1 || (MODE == 2 && ((j / VW) & 1)) || (MODE == 3 && (j / VW) % 3 != 0) ||
           (MODE == 4 && (j / VW) % 3 == 2);
}

// Rank-1 update of columns J0..NC-1 of every lane's row with the pivot
// row held by lane p: row[j] = fma(-l, u_j, row[j]). Columns go in chunks
// of 8 (compile-time recursion, so every register index is a constant),
// with a scheduling barrier between chunks so only one chunk of fetched
// pivot-row values is live next to the row.
template <int MODE, int NC, int J0, int G, typename T>
__device__ __forceinline__ void publish_from(T* pb, const T (&row)[NC]) {
    constexpr int VW = 16 / static_cast<int>(sizeof(T));
    constexpr int j = G * VW;
    if constexpr (j < NC) {
        if constexpr (j + VW > J0 && via_lds<MODE, VW>(j)) {
            if constexpr (j + VW <= NC) {
                if constexpr (VW == 2) {
                    *reinterpret_cast<double2*>(pb + j) = make_double2(row[j], row[j + 1]);
                } else {
                    *reinterpret_cast<float4*>(pb + j) =
                        make_float4(row[j], row[j + 1], row[j + 2], row[j + 3]);
                }
            } else {
#pragma unroll
                for (int q = 0; q < VW; ++q)
                    if (j + q < NC) pb[j + q] = row[j + q];
            }
        }
        publish_from<MODE, NC, J0, G + 1>(pb, row);
    }
}

template <int MODE, int NC, int J0, int C0, typename T>
__device__ __forceinline__ void update_from(T (&row)[NC], T l, int p, const T* pb) {
    constexpr int VW = 16 / static_cast<int>(sizeof(T));
    constexpr int CH = 8;
    if constexpr (C0 < NC) {
        T u[CH];
#pragma unroll
        for (int i = 0; i < CH; i += VW) {
            const int j = C0 + i;
            const bool vec = MODE != 0 && j + VW <= NC && j + VW > J0 && via_lds<MODE, VW>(j);
            if (vec) {
                if constexpr (VW == 2) {
                    const double2 v = *reinterpret_cast<const double2*>(pb + j);
                    u[i] = v.x;
                    u[i + 1] = v.y;
                } else {
                    const float4 v = *reinterpret_cast<const float4*>(pb + j);
                    u[i] = v.x;
                    u[i + 1] = v.y;
                    u[i + 2] = v.z;
                    u[i + 3] = v.w;
                }
            } else {
#pragma unroll
                for (int q = 0; q < VW; ++q) {
                    if (j + q >= J0 && j + q < NC) {
                        if (MODE != 0 && via_lds<MODE, VW>(j + q))
                            u[i + q] = pb[j + q];
                        else
                            u[i + q] = bcast(row[j + q], p);
                    }
                }
            }
        }
#pragma unroll
        for (int i = 0; i < CH; ++i)
            if (C0 + i >= J0 && C0 + i < NC) row[C0 + i] = fma(-l, u[i], row[C0 + i]);
        __builtin_amdgcn_sched_barrier(0);
        update_from<MODE, NC, J0, C0 + CH>(row, l, p, pb);
    }
}

template <int J0, int MODE, int NC, typename T>
__device__ __forceinline__ void update_cols(T (&row)[NC], T l, int p, T* pb, bool me) {
    if constexpr (MODE != 0) {
        if (me) publish_from<MODE, NC, J0, 0>(pb, row);
        wave_lds_sync();
    }
    update_from<MODE, NC, J0, J0 / 8 * 8>(row, l, p, pb);
    if constexpr (MODE != 0) wave_lds_sync();  // reads done before the next publish
}

// State of the row-per-lane elimination.
template <int NT, typename T>
struct ElimState {
    bool done = false;  // this lane's row has been a pivot row (workgroup engines)
    unsigned long long done_mask = 0;  // lanes whose row has been a pivot row (wave engines)
    int mystep = 1 << 20;  // step at which it was (its unknown's index); never, for padding lanes
    T myrcp = T(0);     // 1 / its pivot
    int info = 0;       // first zero (or NaN-only) pivot, 1-based
    int perm[NT];       // lane of the pivot row of each step
};

// Elimination step K (compile time) of an NT x NT system with NC - NT
// right-hand sides appended to the rows. Rows never move: the pivot row's
// lane retires. STORE_L keeps the multiplier in row[K] (for later solves
// with L), otherwise row[K] of every row still active becomes 0.
template <int K, int NT, int NC, class Tune, bool STORE_L, typename T>
__device__ __forceinline__ void elim_from(T (&row)[NC], ElimState<NT, T>& st, T* pb, int lane) {
    if constexpr (K < NT) {
        // rows still active, the pivot lane and the rows eliminated as
        // wave-uniform masks: the per-lane selects read them directly
        constexpr unsigned long long rows = NT >= 64 ? ~0ull : ((1ull << NT) - 1);
        const unsigned long long act = rows & ~st.done_mask;
        const T a = row[K];
        int p = pivot_lane_mask<Tune::PV>(a, act);
        if (p < 0) {  // every candidate NaN
            p = __ffsll(act) - 1;
            if (!st.info) st.info = K + 1;
        }
        p = __builtin_amdgcn_readfirstlane(p);
        const T piv = bcast(a, p);
        if (piv == T(0) && !st.info) st.info = K + 1;
        // settle info at this step (deferred, it keeps every step's pivot
        // and candidate flag in scalar registers, which then spill)
        st.info = __builtin_amdgcn_readfirstlane(st.info);
        asm volatile("" : "+s"(st.info));
        T rcp;
        if constexpr (Tune::RX != 0)
            rcp = recip_uniform(piv);
        else
            rcp = T(1) / piv;
        const unsigned long long pm = 1ull << p;
        const bool is_piv = lane_in(pm);
        const bool elim = lane_in(act & ~pm);
        const T l = elim ? a * rcp : T(0);
        if constexpr (STORE_L) {
            row[K] = elim ? l : a;
        } else {
            row[K] = lane_in(act) ? T(0) : a;
        }
        st.mystep = is_piv ? K : st.mystep;
        st.myrcp = is_piv ? rcp : st.myrcp;
        st.done_mask |= pm;
        // Materialise this step's selects now: left alone, the compiler sinks
        // all NT of them to the end and keeps NT reciprocals and NT lane
        // masks live meanwhile (the masks spilled to VGPR lanes).
        if constexpr (Tune::PN != 0)
            asm volatile("" : "+v"(st.myrcp), "+v"(st.mystep), "+v"(row[K]));
        else
            asm volatile("" : "+v"(st.myrcp), "+v"(st.mystep));
        st.perm[K] = p;
        // The first steps update almost every column: their pivot rows go
        // lane-to-SGPR only, with no LDS-read temporaries at the register
        // peak (n = 35 at four waves per SIMD: 36 instead of 84 B/lane in
        // scratch, 13.72 instead of 13.94 ms).
        constexpr int BCK = K < 5 ? 0 : Tune::BC;
        update_cols<K + 1, BCK>(row, l, p, pb, is_piv);
        elim_from<K + 1, NT, NC, Tune, STORE_L>(row, st, pb, lane);
    }
}

// Knobs of the wave LU, chosen per stencil shape by the autotuner:
//   BC  pivot-row broadcast: 0 = v_readlane into SGPRs, 1 = LDS, 2 = half
//       of the columns each way (balances the VALU and LDS pipes), 3 = two
//       thirds through LDS, 4 = one third through LDS
//   PV  pivot search: see pivot_lane
//   FL  kernel-block fill: 0 = every lane evaluates its own row (the
//       other point through SGPRs), 1 = each pair once, through LDS,
//       2 = as 1 with y read from the node's lane (less LDS per wave)
//   BS  back substitution: 0 = t through SGPRs (v_readlane), 1 = LDS
//   PN  1 = the zeroed column of each step materialised at its step
//   RX  1 = pivot reciprocals by recip_uniform (four VALU instructions
//       fewer per step, a scalar branch per step; pays where registers
//       are not the limit)
template <int BC_, int PV_, int FL_ = 0, int BS_ = 0, int PN_ = 0, int RX_ = 0>
struct LuTune {
    static constexpr int BC = BC_, PV = PV_, FL = FL_, BS = BS_, PN = PN_, RX = RX_;
};
using DefaultTune = LuTune<0, 0>;

// Output table: reference kernels read a device table of NRHS pointers,
// the C ABI passes them by value.
template <typename T>
struct TableOut {
    T* const* A;
    __device__ T* operator()(int c) const { return A[c]; }
};
template <typename T, int R>
struct ValueOut {
    T* p[R];
    __device__ T* operator()(int c) const { return p[c]; }
};

// ---------------------------------------------------------------------------
// One stencil on one wave64: gather -> fill -> LU -> solve -> scatter (FP64)
// ---------------------------------------------------------------------------

template <int DIM, int N, int P, int Q, int R, bool INTERP, class Tune = DefaultTune>
struct WaveAssembly {
    static constexpr int M = npoly_dim(DIM, P);
    static constexpr int NT = N + M;
    static constexpr int NC = NT + R;
    static_assert(DIM == 2 || DIM == 3, "2-D or 3-D");
    static constexpr int nsten = N;  // stencil size
    static_assert(NT <= 64, "one system row per lane of a wave64");
    static_assert(INTERP || R == DIM + 1, "derivative set: Laplacian + one per axis");
    static_assert(Q >= 3 && Q % 2 == 1, "odd PHS exponent >= 3");

    using tune = Tune;
    // LDS per wave, in doubles. With the pairwise fill (Tune::FL = 1): a
    // region used three ways in turn -- monomials [M][N], interpolation
    // centres [DIM][R] and their monomials [R][M] (read while the rows are
    // built); then the kernel block phi [N][N], column-major (written pair
    // by pair, read one column per lane); then the pivot-row buffer [NC] of
    // the elimination (16-byte aligned) -- and after it the scaled points
    // [DIM][N]. Otherwise the pivot-row buffer, the points, the monomials
    // and the centres side by side. Rounded up to an even count so
    // consecutive waves' regions stay 16-byte aligned.
    static constexpr int lds_pivot = (NC + 1) / 2 * 2;
    static constexpr int lds_fill = M * N + (INTERP ? DIM * R + R * M : 0);
    static constexpr int lds_shared = cmax(cmax(lds_fill, N * N), lds_pivot);
    static constexpr bool pair_fill = Tune::FL != 0;
    // FL = 2: the pair fill reads y from the node's lane (cross-lane
    // shuffle) instead of LDS, which trims the wave's LDS below 10 KB
    static constexpr bool y_by_lane = Tune::FL == 2;
    static constexpr int off_pts = pair_fill ? (lds_shared + 1) / 2 * 2 : lds_pivot;
    static constexpr int off_poly = pair_fill ? 0 : lds_pivot + DIM * N;
    static constexpr int lds_doubles =
        pair_fill ? (off_pts + (y_by_lane ? DIM - 1 : DIM) * N + 1) / 2 * 2
                  : (off_poly + lds_fill + 1) / 2 * 2;
    // kernel-block entries above the diagonal, and how many per lane
    static constexpr int npairs = N * (N - 1) / 2;
    static constexpr int pair_slots = (npairs + 63) / 64;

    // Gather, shift, scale and build this lane's row of [M | B]:
    // lane k < N holds point row k, lane N + t polynomial row t.
    // Coordinates and centres of type C (double or float) are widened on load.
    template <typename T, typename C>
    __device__ __forceinline__ static void build_row(const int s, const C* __restrict__ x,
                                     const C* __restrict__ y, nodeduce<const C*> __restrict__ z,
                                     const int* __restrict__ ja_row, nodeduce<const C*> __restrict__ cx,
                                     nodeduce<const C*> __restrict__ cy,
                                     nodeduce<const C*> __restrict__ cz,
                                     double* lds, T (&row)[NC], double& rho2_out) {
#pragma clang fp contract(off)
        const int lane = lane_id();
        double* pts = lds + off_pts;     // [DIM][N]
        double* polyT = lds + off_poly;  // [M][N]
        double* cen = polyT + M * N;     // [DIM][R]
        double* cmono = cen + DIM * R;   // [R][M]
        double* phiT = lds;              // [N][N], entry (i, j) at j * N + i

        const int jn = lane < N ? ja_row[lane] : s;
        const double px = static_cast<double>(x[jn]) - static_cast<double>(x[s]);
        const double py = static_cast<double>(y[jn]) - static_cast<double>(y[s]);
        const double pz = DIM == 3 ? static_cast<double>(z[jn]) - static_cast<double>(z[s]) : 0.0;
        double d2 = px * px + py * py;
        if constexpr (DIM == 3) d2 = d2 + pz * pz;
        double rho2 = wave_uniform(wave_max(lane < N ? d2 : 0.0));
        if (!(rho2 > 0.0)) rho2 = 1.0;
        const double rho = sqrt(rho2);
        rho2_out = rho2;
        const double xs = px / rho, ys = py / rho, zs = pz / rho;

        constexpr PolyBasis<P, DIM> basis{};
        double mono[M];
        basis(xs, ys, zs, mono, 1);
        if (lane < N) {
            pts[lane] = xs;
            if constexpr (!y_by_lane) pts[N + lane] = ys;
            if constexpr (DIM == 3) pts[(y_by_lane ? 1 : 2) * N + lane] = zs;
#pragma unroll
            for (int t = 0; t < M; ++t) polyT[t * N + lane] = mono[t];
        }
        if constexpr (INTERP) {
            if (lane < R) {
                const double ex = static_cast<double>(cx[lane]) / rho;
                const double ey = static_cast<double>(cy[lane]) / rho;
                const double ez = DIM == 3 ? static_cast<double>(cz[lane]) / rho : 0.0;
                cen[lane] = ex;
                cen[R + lane] = ey;
                if constexpr (DIM == 3) cen[2 * R + lane] = ez;
                double cm[M];
                basis(ex, ey, ez, cm, 1);
#pragma unroll
                for (int t = 0; t < M; ++t) cmono[lane * M + t] = cm[t];
            }
        }
        wave_lds_sync();

        constexpr PHS<Q> phi{};
        if constexpr (!pair_fill) {
            // Every lane evaluates its own kernel row; stencil point col
            // comes from lane col through scalar registers (v_readlane).
            double bx[N], by[N], bz[DIM == 3 ? N : 1];
#pragma unroll
            for (int col = 0; col < N; ++col) {
                bx[col] = bcast(xs, col);
                by[col] = bcast(ys, col);
                if constexpr (DIM == 3) bz[col] = bcast(zs, col);
            }
            if (lane < N) {
#pragma unroll
                for (int col = 0; col < N; ++col) {
                    const double dx = bx[col] - xs;
                    const double dy = by[col] - ys;
                    double r2 = dx * dx + dy * dy;
                    if constexpr (DIM == 3) {
                        const double dz = bz[col] - zs;
                        r2 = r2 + dz * dz;
                    }
                    row[col] = static_cast<T>(phi(r2));
                    // keep the scheduler from interleaving all N evaluations
                    // (it would hold N sqrt temporaries live at once)
                    if ((col & 3) == 3) __builtin_amdgcn_sched_barrier(0);
                }
            }
        }
        // Polynomial columns, right-hand sides, and the polynomial rows
        // (which read the monomials of every point from LDS).
        if (lane < N) {
#pragma unroll
            for (int t = 0; t < M; ++t) row[N + t] = static_cast<T>(mono[t]);
            if constexpr (!INTERP) {
                double r2 = xs * xs + ys * ys;
                if constexpr (DIM == 3) r2 = r2 + zs * zs;
                const double tq = PHS<Q - 2>{}(r2);
                const double g = static_cast<double>(Q) * tq;
                row[NT + 0] = static_cast<T>(static_cast<double>(Q * (Q + DIM - 2)) * tq);
                row[NT + 1] = static_cast<T>(-g * xs);
                row[NT + 2] = static_cast<T>(-g * ys);
                if constexpr (DIM == 3) row[NT + 3] = static_cast<T>(-g * zs);
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
                    row[NT + c] = static_cast<T>(phi(r2));
                }
            }
        } else if (lane < NT) {
            const int t = lane - N;
#pragma unroll
            for (int col = 0; col < N; ++col) row[col] = static_cast<T>(polyT[t * N + col]);
#pragma unroll
            for (int c = N; c < NT; ++c) row[c] = T(0);
            if constexpr (!INTERP) {
                using Mk = PolyDerivMasks<DIM, P>;
                row[NT + 0] = ((Mk::lap >> t) & 1u) ? T(2) : T(0);
                row[NT + 1] = ((Mk::dx >> t) & 1u) ? T(1) : T(0);
                row[NT + 2] = ((Mk::dy >> t) & 1u) ? T(1) : T(0);
                if constexpr (DIM == 3) row[NT + 3] = ((Mk::dz >> t) & 1u) ? T(1) : T(0);
            } else {
#pragma unroll
                for (int c = 0; c < R; ++c) row[NT + c] = static_cast<T>(cmono[c * M + t]);
            }
        } else {
#pragma unroll
            for (int c = 0; c < NC; ++c) row[c] = T(0);
        }
        wave_lds_sync();  // monomials and centres read: the kernel block overwrites them
        if constexpr (pair_fill) fill_pairs(lane, ys, pts, phiT, row);
    }

    // Kernel block, each entry above the diagonal evaluated once: pair
    // q = (d - 1) N + i is nodes i and (i + d) mod N, d = 1 .. N/2 (the
    // pairs of d = N/2 counted once for even N), spread over all 64 lanes.
    // phi depends on the squared distance only, and (a - b)^2 == (b - a)^2
    // in floating point, so both entries get the value a direct fill
    // computes.
    template <typename T>
    __device__ __forceinline__ static void fill_pairs(int lane, double ys, const double* pts,
                                                      double* phiT, T (&row)[NC]) {
#pragma clang fp contract(off)
        constexpr PHS<Q> phi{};
        constexpr int oz = y_by_lane ? N : 2 * N;  // z in LDS (3-D)
#pragma unroll
        for (int sl = 0; sl < pair_slots; ++sl) {
            // lanes past the last pair take pair 0 and write nothing (every
            // lane joins the shuffles)
            const int q0 = sl * 64 + lane;
            const int q = q0 < npairs ? q0 : 0;
            const int d = q / N + 1;
            const int i = q - (d - 1) * N;
            const int j = i + d < N ? i + d : i + d - N;
            double yi, yj;
            if constexpr (y_by_lane) {
                yj = __shfl(ys, j, 64);
                yi = __shfl(ys, i, 64);
            } else {
                yj = pts[N + j];
                yi = pts[N + i];
            }
            if (q0 < npairs) {
                const double dx = pts[j] - pts[i];
                const double dy = yj - yi;
                double r2 = dx * dx + dy * dy;
                if constexpr (DIM == 3) {
                    const double dz = pts[oz + j] - pts[oz + i];
                    r2 = r2 + dz * dz;
                }
                const double v = phi(r2);
                phiT[j * N + i] = v;
                phiT[i * N + j] = v;
            }
        }
        if (lane < N) phiT[lane * N + lane] = phi(0.0);
        wave_lds_sync();
        if (lane < N) {
#pragma unroll
            for (int col = 0; col < N; ++col) row[col] = static_cast<T>(phiT[col * N + lane]);
        }
        wave_lds_sync();  // read before the elimination reuses the region
    }

    // Right-looking partial-pivot elimination of [M | B] held one row per
    // lane, then back substitution. On return lane i holds the solution
    // of unknown mystep in xout[c]. Returns info (0, or first zero pivot).
    __device__ __forceinline__ static int lu_solve(double (&row)[NC], int& mystep, double (&xout)[R],
                                   double* lds) {
#pragma clang fp contract(off)
        const int lane = lane_id();
        ElimState<NT, double> st;
        elim_from<0, NT, NC, Tune, false>(row, st, lds, lane);
        // back substitution: row[k] holds u_ik for rows pivotal before step
        // k and 0 for every other row, so all lanes update alike. The
        // pivot lane of step k (the lane whose mystep is k, found by a
        // ballot rather than kept in NT scalar registers through the
        // elimination) passes t_c = b_c / u_kk (as b_c * rcp) to every
        // lane, through SGPRs or (Tune::BS) through LDS in two slots taken
        // in turn.
        if constexpr (Tune::BS == 0) {
#pragma unroll
            for (int k = NT - 1; k >= 0; --k) {
                const int p = __ffsll(wave_ballot(st.mystep == k)) - 1;
#pragma unroll
                for (int c = 0; c < R; ++c) {
                    const double t = bcast(row[NT + c] * st.myrcp, p);
                    row[NT + c] = fma(-row[k], t, row[NT + c]);
                }
            }
        } else {
            constexpr int TS = (R + 1) / 2 * 2;  // slot stride, 16-byte aligned
            static_assert(2 * TS <= lds_pivot, "slots fit the pivot buffer");
#pragma unroll
            for (int k = NT - 1; k >= 0; --k) {
                double* tb = lds + (k & 1) * TS;
                if (st.mystep == k) {
#pragma unroll
                    for (int c = 0; c < R; ++c) tb[c] = row[NT + c] * st.myrcp;
                }
                wave_lds_sync();
#pragma unroll
                for (int c = 0; c < R; ++c) row[NT + c] = fma(-row[k], tb[c], row[NT + c]);
            }
            wave_lds_sync();  // slots read before the caller reuses LDS
        }
#pragma unroll
        for (int c = 0; c < R; ++c) xout[c] = row[NT + c] * st.myrcp;
        mystep = st.mystep;
        return st.info;
    }

    // Stencil s, whose CSR row starts at base (s*N for a fixed-width graph).
    // x, y, z, cx, cy, cz of type C (double or float); out(c) points at C.
    template <class Out, typename C>
    __device__ __forceinline__ static void run(const int s_arg, const size_t base_arg,
                               const C* __restrict__ x,
                               const C* __restrict__ y, nodeduce<const C*> __restrict__ z,
                               const int* __restrict__ ja, Out out,
                               nodeduce<const C*> __restrict__ cx,
                               nodeduce<const C*> __restrict__ cy,
                               nodeduce<const C*> __restrict__ cz, double* lds,
                               int* __restrict__ info) {
#pragma clang fp contract(off)
        const int s = __builtin_amdgcn_readfirstlane(s_arg);
        const size_t base = wave_uniform(base_arg);
        const int* ja_row = ja + base;
        double row[NC];
        double rho2;
        build_row<double, C>(s, x, y, z, ja_row, cx, cy, cz, lds, row, rho2);
        int mystep;
        double xo[R];
        const int inf = lu_solve(row, mystep, xo, lds);
        if (mystep < N) {
            const double rho = sqrt(rho2);
            const size_t at = base + mystep;
#pragma unroll
            for (int c = 0; c < R; ++c) {
                double v = xo[c];
                if constexpr (!INTERP) v = c == 0 ? v / rho2 : v / rho;
                out(c)[at] = static_cast<C>(v);
            }
        }
        if (info && lane_id() == 0) info[s] = inf;
    }
};

}  // namespace detail

}  // namespace rbf_operators

#endif  // RBF_WAVE_H
