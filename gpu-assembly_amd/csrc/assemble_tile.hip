// assemble_tile.hip -- librbf_amd kernels for stencil systems of 64
// unknowns and more: blocked LU with FP64 MFMA trailing updates
// (include/rbf_tile.h), one workgroup per stencil, one wave per
// 16-column tile.
#include <hip/hip_runtime.h>

#include "assemble_common.h"
#include "rbf_tile.h"

namespace rbf_amd_internal {

using namespace rbf_operators;

template <class E>
__global__ void __launch_bounds__(E::threads, E::occupancy)
k_assemble_tile(int nstencils, const double* __restrict__ x, const double* __restrict__ y,
                const double* __restrict__ z, const int* __restrict__ ja,
                detail::ValueOut<double, E::nrhs> out, int* __restrict__ info,
                uint8_t* __restrict__ status) {
    __shared__ __align__(16) double lds[E::lds_doubles];
    const int s = xcd_block(nstencils);
    if (s >= nstencils) return;
    E::run(s, static_cast<size_t>(s) * E::nsten, x, y, z, ja, out, lds, info);
    if (status && threadIdx.x == 0) status[s] = 0;
}

template <int DIM, int N, int P, int Q>
hipError_t launch_tile_diff(const LaunchArgs& a) {
    using E = detail::TileAssembly<DIM, N, P, Q>;
    detail::ValueOut<double, E::nrhs> out;
    for (int c = 0; c < E::nrhs; ++c) out.p[c] = a.w[c];
    hipLaunchKernelGGL((k_assemble_tile<E>), dim3(xcd_grid(a.nstencils)), dim3(E::threads), 0,
                       a.stream, a.nstencils, a.x, a.y, a.z, a.ja, out, a.info, a.status);
    return hipGetLastError();
}

template hipError_t launch_tile_diff<2, 56, 4, 7>(const LaunchArgs&);
template hipError_t launch_tile_diff<3, 80, 3, 3>(const LaunchArgs&);

}  // namespace rbf_amd_internal
