// assemble_group_mixed.hip -- librbf_amd kernels for the mixed-precision
// form of stencil systems larger than a wavefront (include/rbf_group.h,
// GroupMixed), and the FP64 pass that re-solves the stencils it could not
// converge.
#include <hip/hip_runtime.h>

#include "assemble_common.h"
#include "rbf_group.h"
#include "rbf_nullspace.h"
#include "rbf_tile.h"

namespace rbf_amd_internal {

using namespace rbf_operators;

__global__ void k_zero_count(int* c) {
    if (threadIdx.x == 0) *c = 0;
}

template <class E>
__global__ void __launch_bounds__(128)
k_assemble_group_mixed(int nstencils, const double* __restrict__ x, const double* __restrict__ y,
                       const double* __restrict__ z, const int* __restrict__ ja,
                       detail::ValueOut<double, E::nrhs> out, int* __restrict__ info,
                       uint8_t* __restrict__ status, int max_refine, int* __restrict__ ws) {
    __shared__ __align__(16) double lds[E::lds_doubles];
    const int s = xcd_block(nstencils);
    if (s >= nstencils) return;
    E::run(s, static_cast<size_t>(s) * E::nsten, x, y, z, ja, out, lds, info, status, max_refine,
           ws, ws + 1);
}

// The FP64 solve of the stencils the mixed kernel could not converge,
// persistent over the fallback list (the blocked-LU engine, four waves).
template <class G>
__global__ void __launch_bounds__(G::threads, G::occupancy)
k_group_fallback(const double* __restrict__ x, const double* __restrict__ y,
                 const double* __restrict__ z, const int* __restrict__ ja,
                 detail::ValueOut<double, G::nrhs> out, int* __restrict__ info,
                 const int* __restrict__ ws) {
    __shared__ __align__(16) double lds[G::lds_doubles];
    const int count = __builtin_amdgcn_readfirstlane(ws[0]);
    for (int i = static_cast<int>(blockIdx.x); i < count; i += static_cast<int>(gridDim.x)) {
        const int s = ws[1 + i];
        G::run(s, static_cast<size_t>(s) * G::nsten, x, y, z, ja, out, lds, info);
        __syncthreads();  // LDS reuse by the next stencil
    }
}

template <int DIM, int N, int P, int Q>
hipError_t launch_group_mixed(const LaunchArgs& a) {
    constexpr int NW = 2;
    using E = detail::GroupMixed<DIM, N, P, Q, NW>;
    using G = detail::TileAssembly<DIM, N, P, Q>;
    detail::ValueOut<double, E::nrhs> out;
    for (int c = 0; c < E::nrhs; ++c) out.p[c] = a.w[c];
    hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(64), 0, a.stream, a.ws);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_assemble_group_mixed<E>), dim3(xcd_grid(a.nstencils)), dim3(64 * NW), 0,
                       a.stream, a.nstencils, a.x, a.y, a.z, a.ja, out, a.info, a.status,
                       a.max_refine, a.ws);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_group_fallback<G>), dim3(1024), dim3(G::threads), 0, a.stream, a.x, a.y, a.z,
                       a.ja, out, a.info, a.ws);
    return hipGetLastError();
}

// Null-space mixed form of systems larger than a wavefront: one wave per
// stencil (n <= 128 nodes, two per lane above 64; n - m <= 64 free nodes),
// stencils that do not converge listed for the blocked-LU FP64 pass.
template <class E, int OCC>
__global__ void __launch_bounds__(64, OCC)
k_assemble_ns_list(int nstencils, const double* __restrict__ x, const double* __restrict__ y,
                   const double* __restrict__ z, const int* __restrict__ ja,
                   detail::ValueOut<double, E::R> out, int* __restrict__ info,
                   uint8_t* __restrict__ status, int max_refine, int* __restrict__ ws) {
    __shared__ __align__(16) double lds[E::lds_doubles];
    const int s = xcd_block(nstencils);
    if (s >= nstencils) return;
    E::run(s, static_cast<size_t>(s) * E::nsten, x, y, z, ja, out, lds, info, status, max_refine,
           ws, ws + 1);
}

template <int DIM, int N, int P, int Q, int SYM, int OCC>
hipError_t launch_ns_list(const LaunchArgs& a) {
    using E = detail::NullSpaceMixed<DIM, N, P, Q, detail::NsTune<SYM>>;
    using G = detail::TileAssembly<DIM, N, P, Q>;
    detail::ValueOut<double, E::R> out;
    for (int c = 0; c < E::R; ++c) out.p[c] = a.w[c];
    hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(64), 0, a.stream, a.ws);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_assemble_ns_list<E, OCC>), dim3(xcd_grid(a.nstencils)), dim3(64), 0,
                       a.stream, a.nstencils, a.x, a.y, a.z, a.ja, out, a.info, a.status,
                       a.max_refine, a.ws);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_group_fallback<G>), dim3(1024), dim3(G::threads), 0, a.stream, a.x, a.y,
                       a.z, a.ja, out, a.info, a.ws);
    return hipGetLastError();
}

template hipError_t launch_ns_list<2, 56, 4, 7, 16, 1>(const LaunchArgs&);
template hipError_t launch_ns_list<2, 56, 4, 7, 64, 1>(const LaunchArgs&);
template hipError_t launch_ns_list<2, 56, 4, 7, 0, 1>(const LaunchArgs&);
template hipError_t launch_ns_list<2, 56, 4, 7, 32, 1>(const LaunchArgs&);
template hipError_t launch_ns_list<3, 80, 3, 3, 16, 1>(const LaunchArgs&);
template hipError_t launch_ns_list<3, 80, 3, 3, 64, 1>(const LaunchArgs&);
template hipError_t launch_ns_list<3, 80, 3, 3, 32, 1>(const LaunchArgs&);
template hipError_t launch_ns_list<3, 80, 3, 3, 0, 1>(const LaunchArgs&);
template hipError_t launch_group_mixed<2, 56, 4, 7>(const LaunchArgs&);

}  // namespace rbf_amd_internal
