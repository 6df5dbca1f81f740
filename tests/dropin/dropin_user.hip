// A user program of the drop-in header, written the way the reference's
// own driver uses its header (src/main.cpp:10-45): compile-time
// configurations, then one launch per batch with the configuration's block
// and LDS sizes,
//   kernel<<<nstencils, Config::block_dim, Config::shared_memory_size()>>>,
// and the same arguments (a device table of per-operator value arrays, the
// per-stencil info array). The tests call it through these extern "C"
// wrappers and compare its weights with the oracle and with the run-time
// plans of librbf_amd.so.
#include <hip/hip_runtime.h>

#include "rbf_operators.h"

namespace ro = rbf_operators;

namespace {

// main.cpp's configuration (N = 21, P = 2, Q = 3), the benchmarked n = 35
// stencil, a 3-D one (d/dz in A[3]), and two of more than 64 unknowns
// (71 and 100), which the header gives to its four-wave tile engine.
using Lap21 = ro::laplace_config<21, 2, 3, 950, double>;
using Lap35 = ro::laplace_config<35, 3, 5, 950, double>;
using Lap30x3 = ro::laplace_config<30, 2, 3, 950, double, 3>;
using Lap56 = ro::laplace_config<56, 4, 7, 950, double>;
using Lap80x3 = ro::laplace_config<80, 3, 3, 950, double, 3>;
// The CUDA-Fortran demo's D2Q9 interpolation (main.f90:16): 9 centres.
using Interp9 = ro::interp_config<9, 9, 2, 3, 950, double>;
using Interp21 = ro::interp_config<21, 9, 2, 3, 950, double>;
// T = float (the reference's interp_config accepts float or double)
using Lap35f = ro::laplace_config<35, 3, 5, 950, float>;
using Interp9f = ro::interp_config<9, 9, 2, 3, 950, float>;

static_assert(std::is_same_v<Interp9::solver, ro::GESV<double, 15, 9, 950>>);
static_assert(std::is_same_v<Interp9f::solver, ro::GESV<float, 15, 9, 950>>);
static_assert(Lap56::tiled && Lap56::block_dim.x == 256 && !Lap35::tiled);

// A kernel of the user's own built from the header's pieces, in the shape
// of the reference's assemble_interp_weights (rbf_operators.h:340-375): the
// block's LDS cut by Config::slice, the stencil gathered in shifted (not
// scaled) coordinates, fill_matrix / fill_rhs, the solve by
// solve_system_AB<Config::solver>, then the first N rows scattered.
template <class Config, typename T = typename Config::value_type>
__global__ void __launch_bounds__(64)
user_interp_weights(int nstencils, const T* __restrict__ x, const T* __restrict__ y,
                    const int* __restrict__ ja, T* const* __restrict__ A,
                    const T* __restrict__ xc, const T* __restrict__ yc, int* __restrict__ info) {
    extern __shared__ __align__(16) unsigned char user_lds[];
    constexpr int n = Config::N, nt = Config::nt, nrhs = Config::NRHS;
    constexpr int nth = static_cast<int>(Config::nthreads);
    const int s = blockIdx.x;
    if (s >= nstencils) return;
    auto [xs, ys, As, Bs, ipiv] = Config::slice(user_lds);
    const int tid = static_cast<int>(threadIdx.x);
    for (int k = tid; k < n; k += nth) {
        const int j = ja[static_cast<size_t>(s) * n + k];
        xs[k] = x[j] - x[s];
        ys[k] = y[j] - y[s];
    }
    __syncthreads();
    ro::fill_matrix<n, Config::P, Config::Q>(As, xs, ys, tid, nth);
    ro::fill_rhs<n, Config::P, nrhs, Config::Q>(Bs, xs, ys, xc, yc, tid, nth);
    __syncthreads();
    ro::solve_system_AB<typename Config::solver>(As, ipiv, Bs, &info[s]);
    __syncthreads();
    for (int c = 0; c < nrhs; ++c)
        for (int k = tid; k < n; k += nth) A[c][static_cast<size_t>(s) * n + k] = Bs[c * nt + k];
}

// Caller-filled systems through Solver::execute: system b (column-major
// A[nt*nt], B[nt*nrhs]) copied to the block's LDS, solved there, copied back.
template <class Solver, int NT, int NRHS, typename T>
__global__ void __launch_bounds__(64)
user_gesv(int nbatch, T* __restrict__ A, int* __restrict__ ipiv, T* __restrict__ B,
          int* __restrict__ info) {
    __shared__ T a[NT * NT], b[NT * NRHS];
    __shared__ int p[NT];
    const int s = blockIdx.x;
    if (s >= nbatch) return;
    const int tid = static_cast<int>(threadIdx.x);
    T* Ag = A + static_cast<size_t>(s) * NT * NT;
    T* Bg = B + static_cast<size_t>(s) * NT * NRHS;
    for (int i = tid; i < NT * NT; i += 64) a[i] = Ag[i];
    for (int i = tid; i < NT * NRHS; i += 64) b[i] = Bg[i];
    __syncthreads();
    ro::solve_system_AB<Solver>(a, p, b, &info[s]);
    __syncthreads();
    for (int i = tid; i < NT * NT; i += 64) Ag[i] = a[i];
    for (int i = tid; i < NT * NRHS; i += 64) Bg[i] = b[i];
    for (int i = tid; i < NT; i += 64) ipiv[static_cast<size_t>(s) * NT + i] = p[i];
}

template <class Config, typename T = double>
int launch_laplace(int n, const T* x, const T* y, ro::detail::nodeduce<const T*> z, const int* ja,
                   T* const* A, int* info, hipStream_t stream) {
    static_assert(!Config::needs_dynamic_smem_opt_in, "no LDS opt-in on CDNA");
    hipLaunchKernelGGL(ro::assemble_laplace_weights<Config>, dim3(n), Config::block_dim,
                       Config::shared_memory_size(), stream, n, x, y, ja, A, info, z);
    return static_cast<int>(hipGetLastError());
}

template <class Config, typename T = double>
int launch_interp(int n, const T* x, const T* y, const int* ja, T* const* A,
                  const T* xc, const T* yc, int* info, hipStream_t stream) {
    static_assert(!Config::needs_dynamic_smem_opt_in, "no LDS opt-in on CDNA");
    hipLaunchKernelGGL(ro::assemble_interp_weights<Config>, dim3(n), Config::block_dim,
                       Config::shared_memory_size(), stream, n, x, y, ja, A, xc, yc, info);
    return static_cast<int>(hipGetLastError());
}

template <class Config, typename T = double>
int launch_user_interp(int n, const T* x, const T* y, const int* ja, T* const* A,
                       const T* xc, const T* yc, int* info, hipStream_t stream) {
    hipLaunchKernelGGL(user_interp_weights<Config>, dim3(n), Config::block_dim,
                       Config::shared_memory_size(), stream, n, x, y, ja, A, xc, yc, info);
    return static_cast<int>(hipGetLastError());
}

}  // namespace

extern "C" {

// which: 0 = Lap21, 1 = Lap35, 2 = Lap30x3 (z non-null), 3 = Lap56,
// 4 = Lap80x3 (z non-null). A: device table
// of dim + 1 pointers. Returns the hipError_t of the launch; -1 for an
// unknown configuration or n < 1.
int dropin_laplace(int which, int n, const double* x, const double* y, const double* z,
                   const int* ja, double* const* A, int* info, void* stream) {
    if (n < 1) return -1;
    const auto st = static_cast<hipStream_t>(stream);
    switch (which) {
        case 0: return launch_laplace<Lap21>(n, x, y, nullptr, ja, A, info, st);
        case 1: return launch_laplace<Lap35>(n, x, y, nullptr, ja, A, info, st);
        case 2: return z ? launch_laplace<Lap30x3>(n, x, y, z, ja, A, info, st) : -1;
        case 3: return launch_laplace<Lap56>(n, x, y, nullptr, ja, A, info, st);
        case 4: return z ? launch_laplace<Lap80x3>(n, x, y, z, ja, A, info, st) : -1;
        default: return -1;
    }
}

// Lap35f on float data.
int dropin_laplace_f32(int n, const float* x, const float* y, const int* ja, float* const* A,
                       int* info, void* stream) {
    if (n < 1) return -1;
    return launch_laplace<Lap35f, float>(n, x, y, nullptr, ja, A, info,
                                         static_cast<hipStream_t>(stream));
}

// Interp9f on float data.
int dropin_interp_f32(int n, const float* x, const float* y, const int* ja, float* const* A,
                      const float* xc, const float* yc, int* info, void* stream) {
    if (n < 1) return -1;
    return launch_interp<Interp9f, float>(n, x, y, ja, A, xc, yc, info,
                                          static_cast<hipStream_t>(stream));
}

// The user's own kernel: which 0 = Interp9, 1 = Interp21 (double data).
int dropin_user_interp(int which, int n, const double* x, const double* y, const int* ja,
                       double* const* A, const double* xc, const double* yc, int* info,
                       void* stream) {
    if (n < 1) return -1;
    const auto st = static_cast<hipStream_t>(stream);
    switch (which) {
        case 0: return launch_user_interp<Interp9>(n, x, y, ja, A, xc, yc, info, st);
        case 1: return launch_user_interp<Interp21>(n, x, y, ja, A, xc, yc, info, st);
        default: return -1;
    }
}

// Batches of caller-filled 15 x 15 systems with 9 right-hand sides (the
// demo's shape, main.f90:22-57) through GESV<T, 15, 9, 950>::execute.
int dropin_gesv15_f64(int nbatch, double* A, int* ipiv, double* B, int* info, void* stream) {
    if (nbatch < 1) return -1;
    hipLaunchKernelGGL((user_gesv<ro::GESV<double, 15, 9, 950>, 15, 9, double>), dim3(nbatch),
                       dim3(64), 0, static_cast<hipStream_t>(stream), nbatch, A, ipiv, B, info);
    return static_cast<int>(hipGetLastError());
}
int dropin_gesv15_f32(int nbatch, float* A, int* ipiv, float* B, int* info, void* stream) {
    if (nbatch < 1) return -1;
    hipLaunchKernelGGL((user_gesv<ro::GESV<float, 15, 9, 950>, 15, 9, float>), dim3(nbatch),
                       dim3(64), 0, static_cast<hipStream_t>(stream), nbatch, A, ipiv, B, info);
    return static_cast<int>(hipGetLastError());
}

// which: 0 = Interp9, 1 = Interp21; A: device table of 9 pointers.
int dropin_interp(int which, int n, const double* x, const double* y, const int* ja,
                  double* const* A, const double* xc, const double* yc, int* info,
                  void* stream) {
    if (n < 1) return -1;
    const auto st = static_cast<hipStream_t>(stream);
    switch (which) {
        case 0: return launch_interp<Interp9>(n, x, y, ja, A, xc, yc, info, st);
        case 1: return launch_interp<Interp21>(n, x, y, ja, A, xc, yc, info, st);
        default: return -1;
    }
}

// The configurations' launch constants, for the tests to check against the
// documented contract: out = {block_dim.x, shared_memory_size(), nt, N,
// NRHS, waves_per_simd}.
void dropin_config(int interp, int which, long long* out) {
    auto put = [out](auto c) {
        using C = decltype(c);
        out[0] = C::block_dim.x;
        out[1] = static_cast<long long>(C::shared_memory_size());
        out[2] = C::nt;
        out[3] = C::N;
        out[4] = C::NRHS;
        out[5] = C::waves_per_simd;
    };
    if (interp) {
        if (which == 0) put(Interp9{});
        else put(Interp21{});
    } else {
        if (which == 0) put(Lap21{});
        else if (which == 1) put(Lap35{});
        else if (which == 2) put(Lap30x3{});
        else if (which == 3) put(Lap56{});
        else put(Lap80x3{});
    }
}

}  // extern "C"
