// rbf_operators.h -- RBF-FD operator assembly for AMD gfx950 (MI355X).
//
// Drop-in for the reference's src/rbf_operators.h (ivan-pi/gpu-assembly):
// the same namespace, the same names and signatures --
//
//   Solver, GESV, GELS, solve_system_AB       rbf_operators.h:16-37   (rbf_gesv.h)
//   PHS<Q>, npoly, PolyBasis<P>               rbf_operators.h:39-114  (rbf_wave.h)
//   fill_matrix<N,P,Q>, fill_rhs<N,P,NRHS,Q>  rbf_operators.h:116-185 (rbf_wave.h)
//   interp_config<N,NRHS,P,Q,ARCH,T>          rbf_operators.h:187-290
//     ::solver, ::slice(), ::shared_memory_size(), ::block_dim, ...
//   assemble_interp_weights<Config>           rbf_operators.h:292-376
//   laplace_config / assemble_laplace_weights the Laplacian entry point the
//                                             reference sketches (main.cpp:15-16,44-45)
//
// ARCH is a gfx code (950); gfx_arch.h holds the LDS table that the
// reference keeps in cuda_arch.h. T may be float or double: the engines
// compute in FP64 either way (float coordinates are widened on load, the
// weights rounded to float on store).
//
// Engines behind the kernels:
//   nt <= 64  one wave64 per stencil, one system row per lane
//             (detail::WaveAssembly, rbf_wave.h): block_dim 64;
//   nt >  64  (Laplacian configurations only) a four-wave workgroup per
//             stencil, blocked LU with FP64 MFMA trailing updates
//             (detail::TileAssembly, rbf_tile.h): block_dim 256.
// The launch contract is the reference's in both cases:
//   <<<nstencils, Config::block_dim, Config::shared_memory_size()>>>.
#ifndef RBF_OPERATORS_H
#define RBF_OPERATORS_H

#include <cstddef>
#include <type_traits>

#include "rbf_gesv.h"
#include "rbf_tile.h"
#include "rbf_wave.h"

namespace rbf_operators {

// ---------------------------------------------------------------------------
// Launch configurations. The kernels compile with the wave LU tuning that
// the run-time plans start from before autotuning: pivot row a third
// through LDS, DPP pivot search, pairwise fill with y read from the node's
// lane, LDS back substitution, registers sized for four waves per SIMD.
// ---------------------------------------------------------------------------

namespace detail {
using DropInTune = LuTune<4, 2, 2, 1, 1>;
constexpr int kDropInWavesPerSimd = 4;

constexpr std::size_t round_up(std::size_t n, std::size_t a) { return (n + a - 1) / a * a; }
constexpr std::size_t smax(std::size_t a, std::size_t b) { return a > b ? a : b; }

// The per-stencil arrays of a caller's own solve, in the reference's order
// (rbf_operators.h:206-207): xs[N], ys[N], As[nt*nt], Bs[nt*NRHS], ipiv[nt],
// each aligned to its element type.
template <typename T, int N, int NT, int NRHS>
struct StencilLayout {
    static constexpr std::size_t xs = 0;
    static constexpr std::size_t ys = round_up(xs + sizeof(T) * N, alignof(T));
    static constexpr std::size_t As = round_up(ys + sizeof(T) * N, alignof(T));
    static constexpr std::size_t Bs = round_up(As + sizeof(T) * NT * NT, alignof(T));
    static constexpr std::size_t ipiv = round_up(Bs + sizeof(T) * NT * NRHS, alignof(int));
    static constexpr std::size_t bytes = ipiv + sizeof(int) * NT;
};

template <typename T>
struct StencilSlices {
    T* xs;
    T* ys;
    T* As;
    T* Bs;
    int* ipiv;
};
}  // namespace detail

// Interpolation to NRHS centres (2-D). nt = N + npoly must be at most 64:
// one system row per lane of the wave that solves it.
template <int N_, int NRHS_, int P_, int Q_ = 3, unsigned int ARCH_ = 950, typename T_ = double>
struct interp_config {
    using value_type = T_;
    static constexpr int N = N_, NRHS = NRHS_, P = P_, Q = Q_;
    static constexpr unsigned int ARCH = ARCH_;
    static constexpr int npoly = PolyBasis<P>::np;
    static constexpr int nt = N + npoly;
    static constexpr int dim = 2;
    static_assert(std::is_same_v<value_type, float> || std::is_same_v<value_type, double>,
                  "real FP32 or FP64 data");
    static_assert(nt <= 64, "interpolation systems are solved one row per lane: nt <= 64");

    // --- Solver: the block GESV a caller's own kernel can run on slice()
    using solver = GESV<value_type, nt, NRHS, ARCH>;
    static_assert(solver::is_block_execution, "one cooperative solve per block");
    static_assert(solver::batches_per_block == 1, "one stencil per block");
    static_assert(solver::block_dim.y == 1 && solver::block_dim.z == 1, "1-D block");

    // --- Engine of assemble_interp_weights
    using tune = detail::DropInTune;
    static constexpr int waves_per_simd = detail::kDropInWavesPerSimd;
    using engine = detail::WaveAssembly<2, N, P, Q, NRHS, true, tune>;

    // --- Architecture and LDS
    static constexpr auto& arch = gfx_arch::limits_v<ARCH>;
    static constexpr std::size_t align_up(std::size_t n, std::size_t a) { return detail::round_up(n, a); }
    using layout = detail::StencilLayout<value_type, N, nt, NRHS>;
    // Room for the engine's working set and for the slice() arrays, so a
    // launch sized by it serves either kernel.
    static constexpr std::size_t shared_memory_size() {
        return detail::smax(sizeof(double) * static_cast<std::size_t>(engine::lds_doubles),
                            layout::bytes);
    }
    // xs, ys, As, Bs, ipiv in the block's dynamic LDS p (16-byte aligned);
    // binds as auto [xs, ys, As, Bs, ipiv] = Config::slice(p).
    __device__ static detail::StencilSlices<value_type> slice(void* p) {
        unsigned char* b = static_cast<unsigned char*>(p);
        return {reinterpret_cast<value_type*>(b + layout::xs),
                reinterpret_cast<value_type*>(b + layout::ys),
                reinterpret_cast<value_type*>(b + layout::As),
                reinterpret_cast<value_type*>(b + layout::Bs), reinterpret_cast<int*>(b + layout::ipiv)};
    }
    static_assert(arch.fits(shared_memory_size()), "LDS footprint exceeds the per-workgroup limit");
    static_assert(sizeof(value_type) * (nt * nt + nt * NRHS) >= solver::shared_memory_size,
                  "the A/B slices hold the solver's operands");

    // --- Launch-side parameters
    static constexpr dim3 block_dim = solver::block_dim;
    static constexpr unsigned nthreads = solver::max_threads_per_block;
    static constexpr bool needs_dynamic_smem_opt_in = arch.needs_dynamic_smem_opt_in(shared_memory_size());
};

// Laplacian, d/dx, d/dy (, d/dz) at the stencil's own node: NRHS = DIM + 1.
// Systems of more than 64 unknowns go to the four-wave tile engine.
template <int N_, int P_, int Q_ = 3, unsigned int ARCH_ = 950, typename T_ = double, int DIM_ = 2>
struct laplace_config {
    using value_type = T_;
    static constexpr int N = N_, P = P_, Q = Q_, dim = DIM_;
    static constexpr int NRHS = DIM_ + 1;
    static constexpr unsigned int ARCH = ARCH_;
    static constexpr int npoly = npoly_dim(DIM_, P_);
    static constexpr int nt = N + npoly;
    static_assert(std::is_same_v<value_type, float> || std::is_same_v<value_type, double>,
                  "real FP32 or FP64 data");
    static constexpr bool tiled = nt > 64;
    using tune = detail::DropInTune;
    using engine = std::conditional_t<tiled, detail::TileAssembly<DIM_, N, P, Q>,
                                      detail::WaveAssembly<DIM_, N, P, Q, NRHS, false, tune>>;
    // the block GESV of one system, where one wave solves it (nt <= 64)
    using solver = std::conditional_t<tiled, void, GESV<value_type, (tiled ? 64 : nt), NRHS, ARCH>>;
    static constexpr unsigned nthreads = tiled ? 256u : 64u;
    static constexpr int waves_per_simd = tiled ? 1 : detail::kDropInWavesPerSimd;
    static constexpr auto& arch = gfx_arch::limits_v<ARCH>;
    static constexpr std::size_t shared_memory_size() {
        return sizeof(double) * static_cast<std::size_t>(engine::lds_doubles);
    }
    static_assert(arch.fits(shared_memory_size()), "LDS footprint exceeds the per-workgroup limit");
    static constexpr dim3 block_dim = dim3(nthreads, 1, 1);
    static constexpr bool needs_dynamic_smem_opt_in = false;
};

// Interpolation weights (rbf_operators.h:320-376). Same arguments: A is a
// device-readable table of NRHS value arrays laid out like ja; xc, yc are
// the NRHS centres in the stencil-local (shifted, unscaled) frame.
template <class Config, typename T = typename Config::value_type>
__global__ void __launch_bounds__(Config::nthreads, Config::waves_per_simd)
assemble_interp_weights(const int nstencils,
                        const T* __restrict__ x,
                        const T* __restrict__ y,
                        const int* __restrict__ ja,
                        T* const* __restrict__ A,
                        const T* __restrict__ xc,
                        const T* __restrict__ yc,
                        int* __restrict__ info) {
    extern __shared__ __align__(16) double rbf_lds_interp[];
    const int s = blockIdx.x;
    if (s >= nstencils) return;
    Config::engine::run(s, static_cast<size_t>(s) * Config::N, x, y, nullptr, ja,
                        detail::TableOut<T>{A}, xc, yc, nullptr, rbf_lds_interp, info);
}

// Laplacian / d/dx / d/dy weights at every node: A[0] Laplacian, A[1]
// d/dx, A[2] d/dy (A[3] d/dz in 3-D, with z non-null).
template <class Config, typename T = typename Config::value_type>
__global__ void __launch_bounds__(Config::nthreads, Config::waves_per_simd)
assemble_laplace_weights(const int nstencils,
                         const T* __restrict__ x,
                         const T* __restrict__ y,
                         const int* __restrict__ ja,
                         T* const* __restrict__ A,
                         int* __restrict__ info,
                         const T* __restrict__ z = nullptr) {
    extern __shared__ __align__(16) double rbf_lds_laplace[];
    const int s = blockIdx.x;
    if (s >= nstencils) return;
    if constexpr (Config::tiled) {
        Config::engine::run(s, static_cast<size_t>(s) * Config::N, x, y, z, ja,
                            detail::TableOut<T>{A}, rbf_lds_laplace, info);
    } else {
        Config::engine::run(s, static_cast<size_t>(s) * Config::N, x, y, z, ja,
                            detail::TableOut<T>{A}, nullptr, nullptr, nullptr, rbf_lds_laplace,
                            info);
    }
}

}  // namespace rbf_operators

#endif  // RBF_OPERATORS_H
