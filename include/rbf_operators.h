// rbf_operators.h -- RBF-FD operator assembly for AMD gfx950 (MI355X).
//
// Drop-in for the reference's src/rbf_operators.h (ivan-pi/gpu-assembly):
// the same namespace, the same names and signatures --
//
//   PHS<Q>, npoly, PolyBasis<P>               rbf_operators.h:39-114
//   fill_matrix<N,P,Q>, fill_rhs<N,P,NRHS,Q>  rbf_operators.h:116-185
//   interp_config<N,NRHS,P,Q,ARCH,T>          rbf_operators.h:187-290
//   assemble_interp_weights<Config>           rbf_operators.h:292-376
//   laplace_config / assemble_laplace_weights the Laplacian entry point the
//                                             reference sketches (main.cpp:15-16,44-45)
//
// -- with the cuSolverDx block GESV replaced by a wavefront LU written for
// CDNA4: one wave64 per stencil, one matrix row per lane held in VGPRs
// together with its right-hand sides, pivot search by a wave max-reduction
// plus ballot, the pivot row broadcast lane-to-SGPR (v_readlane) straight
// into the FMA operands, and rows never moved (a lane that supplied a pivot
// is retired; the permutation lives in scalar registers). ARCH is a gfx
// code (950), the shared-memory table is gfx_arch.h.
//
// Differences from the reference, on purpose:
//   - the system is built in coordinates shifted to the stencil's node AND
//     scaled by the stencil radius rho = max_k |p_k - p_s|; interpolation
//     weights are scale-free, derivative weights are unscaled by rho^-1
//     (first derivatives) and rho^-2 (Laplacian). Without the scaling the
//     PHS systems at h = 1/1024 are numerically singular in FP64;
//   - fills round each product and sum on its own (no FMA contraction) and
//     the elimination uses explicit fma(), in the same order as the CPU
//     oracle, so GPU and CPU agree to a few ulps;
//   - one wave per stencil: block_dim is 64 and the launch contract
//     gridDim.x = nstencils is kept.
#ifndef RBF_OPERATORS_H
#define RBF_OPERATORS_H

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

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

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

template <int DIM, int N, int P, int Q, int R, bool INTERP>
struct WaveAssembly {
    static constexpr int M = npoly_dim(DIM, P);
    static constexpr int NT = N + M;
    static constexpr int NC = NT + R;
    static_assert(DIM == 2 || DIM == 3, "2-D or 3-D");
    static_assert(NT <= 64, "one system row per lane of a wave64");
    static_assert(INTERP || R == DIM + 1, "derivative set: Laplacian + one per axis");
    static_assert(Q >= 3 && Q % 2 == 1, "odd PHS exponent >= 3");

    // LDS per wave, in doubles: scaled points [DIM][N], monomials [M][N],
    // interpolation centres [DIM][R] and their monomials [R][M].
    static constexpr int lds_doubles = DIM * N + M * N + (INTERP ? DIM * R + R * M : 0);

    // Gather, shift, scale and build this lane's row of [M | B]:
    // lane k < N holds point row k, lane N + t polynomial row t.
    template <typename T>
    __device__ static void build_row(const int s, const double* __restrict__ x,
                                     const double* __restrict__ y, const double* __restrict__ z,
                                     const int* __restrict__ ja_row, const double* __restrict__ cx,
                                     const double* __restrict__ cy, const double* __restrict__ cz,
                                     double* lds, T (&row)[NC], double& rho2_out) {
#pragma clang fp contract(off)
        const int lane = lane_id();
        double* pts = lds;               // [DIM][N]
        double* polyT = lds + DIM * N;   // [M][N]
        double* cen = polyT + M * N;     // [DIM][R]
        double* cmono = cen + DIM * R;   // [R][M]

        const int jn = lane < N ? ja_row[lane] : s;
        const double px = x[jn] - x[s];
        const double py = y[jn] - y[s];
        const double pz = DIM == 3 ? z[jn] - z[s] : 0.0;
        double d2 = px * px + py * py;
        if constexpr (DIM == 3) d2 = d2 + pz * pz;
        double rho2 = wave_max(lane < N ? d2 : 0.0);
        if (!(rho2 > 0.0)) rho2 = 1.0;
        const double rho = sqrt(rho2);
        rho2_out = rho2;
        const double xs = px / rho, ys = py / rho, zs = pz / rho;

        constexpr PolyBasis<P, DIM> basis{};
        double mono[M];
        basis(xs, ys, zs, mono, 1);
        if (lane < N) {
            pts[lane] = xs;
            pts[N + lane] = ys;
            if constexpr (DIM == 3) pts[2 * N + lane] = zs;
#pragma unroll
            for (int t = 0; t < M; ++t) polyT[t * N + lane] = mono[t];
        }
        if constexpr (INTERP) {
            if (lane < R) {
                const double ex = cx[lane] / rho, ey = cy[lane] / rho;
                const double ez = DIM == 3 ? cz[lane] / rho : 0.0;
                cen[lane] = ex;
                cen[R + lane] = ey;
                if constexpr (DIM == 3) cen[2 * R + lane] = ez;
                double cm[M];
                basis(ex, ey, ez, cm, 1);
#pragma unroll
                for (int t = 0; t < M; ++t) cmono[lane * M + t] = cm[t];
            }
        }
        __syncthreads();

        constexpr PHS<Q> phi{};
        if (lane < N) {
#pragma unroll
            for (int col = 0; col < N; ++col) {
                const double dx = pts[col] - xs;
                const double dy = pts[N + col] - ys;
                double r2 = dx * dx + dy * dy;
                if constexpr (DIM == 3) {
                    const double dz = pts[2 * N + col] - zs;
                    r2 = r2 + dz * dz;
                }
                row[col] = static_cast<T>(phi(r2));
            }
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
    }

    // Right-looking partial-pivot elimination of [M | B] held one row per
    // lane, then back substitution. On return lane i holds the solution
    // of unknown mystep in xout[c]. Returns info (0, or first zero pivot).
    __device__ static int lu_solve(double (&row)[NC], int& mystep, double (&xout)[R]) {
#pragma clang fp contract(off)
        const int lane = lane_id();
        bool done = false;
        double myrcp = 0.0;
        int info = 0;
        int perm[NT];
        mystep = 64;
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            const bool active = lane < NT && !done;
            const double a = row[k];
            const double cand = active ? fabs(a) : -1.0;
            const double mx = wave_max(cand);
            const unsigned long long hit = __ballot(active && cand == mx);
            int p;
            if (hit) {
                p = __ffsll(hit) - 1;
            } else {  // every candidate NaN
                p = __ffsll(__ballot(active)) - 1;
                if (!info) info = k + 1;
            }
            p = __builtin_amdgcn_readfirstlane(p);
            const double piv = bcast(a, p);
            if (piv == 0.0 && !info) info = k + 1;
            const double rcp = 1.0 / piv;
            const bool is_piv = lane == p;
            const double l = (active && !is_piv) ? a * rcp : 0.0;
            row[k] = active ? 0.0 : row[k];
            if (is_piv) {
                done = true;
                mystep = k;
                myrcp = rcp;
            }
            perm[k] = p;
#pragma unroll
            for (int j = k + 1; j < NC; ++j) {
                const double u = bcast(row[j], p);
                row[j] = fma(-l, u, row[j]);
            }
        }
#pragma unroll
        for (int k = NT - 1; k >= 0; --k) {
            const int p = perm[k];
#pragma unroll
            for (int c = 0; c < R; ++c) {
                const double t = bcast(row[NT + c] * myrcp, p);
                row[NT + c] = fma(-row[k], t, row[NT + c]);
            }
        }
#pragma unroll
        for (int c = 0; c < R; ++c) xout[c] = row[NT + c] * myrcp;
        return info;
    }

    template <class Out>
    __device__ static void run(const int s, const double* __restrict__ x,
                               const double* __restrict__ y, const double* __restrict__ z,
                               const int* __restrict__ ja, Out out,
                               const double* __restrict__ cx, const double* __restrict__ cy,
                               const double* __restrict__ cz, double* lds,
                               int* __restrict__ info) {
#pragma clang fp contract(off)
        const int* ja_row = ja + static_cast<size_t>(s) * N;
        double row[NC];
        double rho2;
        build_row<double>(s, x, y, z, ja_row, cx, cy, cz, lds, row, rho2);
        int mystep;
        double xo[R];
        const int inf = lu_solve(row, mystep, xo);
        if (mystep < N) {
            const double rho = sqrt(rho2);
            const size_t base = static_cast<size_t>(s) * N + mystep;
#pragma unroll
            for (int c = 0; c < R; ++c) {
                double v = xo[c];
                if constexpr (!INTERP) v = c == 0 ? v / rho2 : v / rho;
                out(c)[base] = v;
            }
        }
        if (info && lane_id() == 0) info[s] = inf;
    }
};

}  // namespace detail

// ---------------------------------------------------------------------------
// Launch configurations (the reference's interp_config, rbf_operators.h:
// 187-290). Launch contract:
//   gridDim.x = nstencils, blockDim = block_dim (one wave64),
//   dynamic LDS = shared_memory_size() (no opt-in needed on CDNA).
// ---------------------------------------------------------------------------

template <int N_, int NRHS_, int P_, int Q_ = 3, unsigned int ARCH_ = 950, typename T_ = double>
struct interp_config {
    using value_type = T_;
    static constexpr int N = N_, NRHS = NRHS_, P = P_, Q = Q_;
    static constexpr unsigned int ARCH = ARCH_;
    static constexpr int npoly = PolyBasis<P>::np;
    static constexpr int nt = N + npoly;
    static constexpr int dim = 2;
    static_assert(std::is_same_v<value_type, double>, "the wave LU path computes in FP64");
    using engine = detail::WaveAssembly<2, N, P, Q, NRHS, true>;
    static constexpr auto& arch = gfx_arch::limits_v<ARCH>;
    static constexpr std::size_t shared_memory_size() {
        return sizeof(double) * static_cast<std::size_t>(engine::lds_doubles);
    }
    static_assert(arch.fits(shared_memory_size()), "LDS footprint exceeds the per-workgroup limit");
    static constexpr dim3 block_dim = dim3(64, 1, 1);
    static constexpr unsigned nthreads = 64;
    static constexpr bool needs_dynamic_smem_opt_in = arch.needs_dynamic_smem_opt_in(shared_memory_size());
};

// Laplacian, d/dx, d/dy (, d/dz) at the stencil's own node: NRHS = DIM + 1.
template <int N_, int P_, int Q_ = 3, unsigned int ARCH_ = 950, typename T_ = double, int DIM_ = 2>
struct laplace_config {
    using value_type = T_;
    static constexpr int N = N_, P = P_, Q = Q_, dim = DIM_;
    static constexpr int NRHS = DIM_ + 1;
    static constexpr unsigned int ARCH = ARCH_;
    static constexpr int npoly = npoly_dim(DIM_, P_);
    static constexpr int nt = N + npoly;
    static_assert(std::is_same_v<value_type, double>, "the wave LU path computes in FP64");
    using engine = detail::WaveAssembly<DIM_, N, P, Q, NRHS, false>;
    static constexpr auto& arch = gfx_arch::limits_v<ARCH>;
    static constexpr std::size_t shared_memory_size() {
        return sizeof(double) * static_cast<std::size_t>(engine::lds_doubles);
    }
    static_assert(arch.fits(shared_memory_size()), "LDS footprint exceeds the per-workgroup limit");
    static constexpr dim3 block_dim = dim3(64, 1, 1);
    static constexpr unsigned nthreads = 64;
    static constexpr bool needs_dynamic_smem_opt_in = false;
};

// Interpolation weights (rbf_operators.h:320-376). Same arguments: A is a
// device-readable table of NRHS value arrays laid out like ja; xc, yc are
// the NRHS centres in the stencil-local (shifted, unscaled) frame.
template <class Config, typename T = typename Config::value_type>
__global__ void __launch_bounds__(64) assemble_interp_weights(const int nstencils,
                                                              const T* __restrict__ x,
                                                              const T* __restrict__ y,
                                                              const int* __restrict__ ja,
                                                              T* const* __restrict__ A,
                                                              const T* __restrict__ xc,
                                                              const T* __restrict__ yc,
                                                              int* __restrict__ info) {
    extern __shared__ double rbf_lds_interp[];
    const int s = blockIdx.x;
    if (s >= nstencils) return;
    Config::engine::run(s, x, y, nullptr, ja, detail::TableOut<T>{A}, xc, yc, nullptr,
                        rbf_lds_interp, info);
}

// Laplacian / d/dx / d/dy weights at every node: A[0] Laplacian, A[1]
// d/dx, A[2] d/dy (A[3] d/dz in 3-D, with z non-null).
template <class Config, typename T = typename Config::value_type>
__global__ void __launch_bounds__(64) assemble_laplace_weights(const int nstencils,
                                                               const T* __restrict__ x,
                                                               const T* __restrict__ y,
                                                               const int* __restrict__ ja,
                                                               T* const* __restrict__ A,
                                                               int* __restrict__ info,
                                                               const T* __restrict__ z = nullptr) {
    extern __shared__ double rbf_lds_laplace[];
    const int s = blockIdx.x;
    if (s >= nstencils) return;
    Config::engine::run(s, x, y, z, ja, detail::TableOut<T>{A}, nullptr, nullptr, nullptr,
                        rbf_lds_laplace, info);
}

}  // namespace rbf_operators

#endif  // RBF_OPERATORS_H
