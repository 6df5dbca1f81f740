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
// stencil, and a 3-D one (d/dz in A[3]).
using Lap21 = ro::laplace_config<21, 2, 3, 950, double>;
using Lap35 = ro::laplace_config<35, 3, 5, 950, double>;
using Lap30x3 = ro::laplace_config<30, 2, 3, 950, double, 3>;
// The CUDA-Fortran demo's D2Q9 interpolation (main.f90:16): 9 centres.
using Interp9 = ro::interp_config<9, 9, 2, 3, 950, double>;
using Interp21 = ro::interp_config<21, 9, 2, 3, 950, double>;

template <class Config>
int launch_laplace(int n, const double* x, const double* y, const double* z, const int* ja,
                   double* const* A, int* info, hipStream_t stream) {
    static_assert(!Config::needs_dynamic_smem_opt_in, "no LDS opt-in on CDNA");
    hipLaunchKernelGGL(ro::assemble_laplace_weights<Config>, dim3(n), Config::block_dim,
                       Config::shared_memory_size(), stream, n, x, y, ja, A, info, z);
    return static_cast<int>(hipGetLastError());
}

template <class Config>
int launch_interp(int n, const double* x, const double* y, const int* ja, double* const* A,
                  const double* xc, const double* yc, int* info, hipStream_t stream) {
    static_assert(!Config::needs_dynamic_smem_opt_in, "no LDS opt-in on CDNA");
    hipLaunchKernelGGL(ro::assemble_interp_weights<Config>, dim3(n), Config::block_dim,
                       Config::shared_memory_size(), stream, n, x, y, ja, A, xc, yc, info);
    return static_cast<int>(hipGetLastError());
}

}  // namespace

extern "C" {

// which: 0 = Lap21, 1 = Lap35, 2 = Lap30x3 (z non-null). A: device table
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
        default: return -1;
    }
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
        else put(Lap30x3{});
    }
}

}  // extern "C"
