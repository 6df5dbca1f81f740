// gesv.hip -- batched GESV of caller-filled systems (C ABI): the
// reference's extern "C" solve_dp (rbf_operators.cpp:7-11) and the
// execute(A, ipiv, B, info) it wraps (rbf_operators.h:29-37), for a batch
// of systems, one wave64 per system (include/rbf_gesv.h).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "assemble_common.h"
#include "rbf_assembly.h"
#include "rbf_gesv.h"

namespace {

using namespace rbf_operators;
using namespace rbf_amd_internal;

constexpr int kWaves = 4;  // systems per workgroup

template <int NTM, int NRM>
__global__ void __launch_bounds__(64 * kWaves)
k_gesv(int nt, int nrhs, int nbatch, double* __restrict__ A, int* __restrict__ ipiv,
       double* __restrict__ B, int* __restrict__ info) {
    const int nblocks = (nbatch + kWaves - 1) / kWaves;
    const int blk = xcd_block(nblocks);
    const int b = blk * kWaves + static_cast<int>(threadIdx.x >> 6);
    if (blk >= nblocks || b >= nbatch) return;
    const size_t sa = static_cast<size_t>(b) * nt * nt;
    const size_t sb = static_cast<size_t>(b) * nt * (nrhs > 0 ? nrhs : 1);
    const int r = detail::WaveGesv<NTM, NRM>::solve(nt, nrhs, A + sa, nt, ipiv + static_cast<size_t>(b) * nt,
                                                    nrhs > 0 ? B + sb : B, nt);
    if (info && detail::lane_id() == 0) info[b] = r;
}

template <int NTM, int NRM>
hipError_t launch_gesv(int nt, int nrhs, int nbatch, double* A, int* ipiv, double* B, int* info,
                       hipStream_t st) {
    const int nblocks = (nbatch + kWaves - 1) / kWaves;
    hipLaunchKernelGGL((k_gesv<NTM, NRM>), dim3(xcd_grid(nblocks)), dim3(64 * kWaves), 0, st, nt,
                       nrhs, nbatch, A, ipiv, B, info);
    return hipGetLastError();
}

}  // namespace

extern "C" {

RBFA_API int32_t rbfa_gesv_batched(int32_t nt, int32_t nrhs, int32_t nbatch, double* A,
                                   int32_t* ipiv, double* B, int32_t* info, void* stream) {
    if (nt < 1 || nt > 64 || nrhs < 0 || nrhs > 16 || nbatch < 0) return RBFA_ERR_ARG;
    if (nbatch == 0) return RBFA_OK;
    if (!A || !ipiv || (nrhs > 0 && !B)) return RBFA_ERR_ARG;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipError_t e;
    if (nt <= 16)
        e = nrhs <= 4 ? launch_gesv<16, 4>(nt, nrhs, nbatch, A, ipiv, B, info, st)
                      : launch_gesv<16, 16>(nt, nrhs, nbatch, A, ipiv, B, info, st);
    else if (nt <= 32)
        e = nrhs <= 4 ? launch_gesv<32, 4>(nt, nrhs, nbatch, A, ipiv, B, info, st)
                      : launch_gesv<32, 16>(nt, nrhs, nbatch, A, ipiv, B, info, st);
    else
        e = nrhs <= 4 ? launch_gesv<64, 4>(nt, nrhs, nbatch, A, ipiv, B, info, st)
                      : launch_gesv<64, 16>(nt, nrhs, nbatch, A, ipiv, B, info, st);
    return e == hipSuccess ? RBFA_OK : RBFA_ERR_HIP;
}

RBFA_API int32_t rbfa_solve_dp(double* A, int32_t* ipiv, double* B, int32_t* info, void* stream) {
    if (!info) return RBFA_ERR_ARG;
    return rbfa_gesv_batched(15, 9, 1, A, ipiv, B, info, stream);
}

}  // extern "C"
