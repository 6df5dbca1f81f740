// assemble_pack.hip -- kernels of the packed engine (include/rbf_wavepack.h):
// several stencils of at most 32 unknowns per wave64, one per segment of
// 16 or 32 lanes. Instantiated for the shipped shapes whose systems fit a
// segment; assemble.hip lists them among each shape's variants.

#include <hip/hip_runtime.h>

#include "assemble_common.h"
#include "rbf_wavepack.h"

namespace rbf_amd_internal {

namespace {

using namespace rbf_operators;

template <class E, int WPB, int OCC>
__global__ void __launch_bounds__(64 * WPB, OCC)
k_assemble_pack(int nstencils, const double* __restrict__ x, const double* __restrict__ y,
                const double* __restrict__ z, const int* __restrict__ ja,
                detail::ValueOut<double, E::NC - E::NT> out, const double* __restrict__ cx,
                const double* __restrict__ cy, const double* __restrict__ cz,
                int* __restrict__ info, uint8_t* __restrict__ status) {
    __shared__ __align__(16) double lds[WPB * E::lds_doubles];
    const int nwaves = (nstencils + E::S - 1) / E::S;
    const int nblocks = (nwaves + WPB - 1) / WPB;
    const int b = xcd_block(nblocks);
    const int wv = b * WPB + static_cast<int>(threadIdx.x >> 6);
    if (b >= nblocks || wv >= nwaves) return;
    E::run(wv * E::S, nstencils, x, y, z, ja, out, cx, cy, cz,
           lds + static_cast<int>(threadIdx.x >> 6) * E::lds_doubles, info, status);
}

}  // namespace

template <int DIM, int N, int P, int Q, int R, bool INTERP, int SL, int BCM, int OCC>
hipError_t launch_pack(const LaunchArgs& a) {
    constexpr int WPB = 1;
    using E = detail::WavePack<DIM, N, P, Q, R, INTERP, SL, 0, BCM>;
    detail::ValueOut<double, R> out;
    for (int c = 0; c < R; ++c) out.p[c] = a.w[c];
    const int nwaves = (a.nstencils + E::S - 1) / E::S;
    const int nblocks = (nwaves + WPB - 1) / WPB;
    hipLaunchKernelGGL((k_assemble_pack<E, WPB, OCC>), dim3(xcd_grid(nblocks)), dim3(64 * WPB), 0,
                       a.stream, a.nstencils, a.x, a.y, a.z, a.ja, out, a.cx, a.cy, a.cz, a.info,
                       a.status);
    return hipGetLastError();
}

// derivative sets (Laplacian, d/dx, d/dy) of the 2-D shapes up to 32
// unknowns, and interpolation to the D2Q9 lattice points (n = 9: 15 x 15;
// n = 21: 27 x 27); both broadcasts, registers unconstrained and sized for
// 4 (16-lane segments) or 3 (32-lane segments) waves per SIMD
#define RBFA_PACK(N, P, Q, R, INTERP, SL, OCC2)                                       \
    template hipError_t launch_pack<2, N, P, Q, R, INTERP, SL, 0, 1>(const LaunchArgs&);    \
    template hipError_t launch_pack<2, N, P, Q, R, INTERP, SL, 0, OCC2>(const LaunchArgs&); \
    template hipError_t launch_pack<2, N, P, Q, R, INTERP, SL, 1, 1>(const LaunchArgs&);    \
    template hipError_t launch_pack<2, N, P, Q, R, INTERP, SL, 1, OCC2>(const LaunchArgs&);
RBFA_PACK(13, 1, 3, 3, false, 16, 4)
RBFA_PACK(13, 2, 3, 3, false, 32, 3)
RBFA_PACK(21, 2, 3, 3, false, 32, 3)
RBFA_PACK(21, 2, 5, 3, false, 32, 3)
RBFA_PACK(9, 2, 3, 9, true, 16, 4)
RBFA_PACK(21, 2, 3, 9, true, 32, 3)
#undef RBFA_PACK

}  // namespace rbf_amd_internal
