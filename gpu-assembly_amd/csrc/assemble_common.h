// assemble_common.h -- internal to librbf_amd: launch arguments shared by
// the kernel translation units, and the XCD-aware block numbering.
#ifndef RBF_ASSEMBLE_COMMON_H
#define RBF_ASSEMBLE_COMMON_H

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rbf_amd_internal {

constexpr int kMaxOps = 16;

struct LaunchArgs {
    int32_t nstencils;
    const double *x, *y, *z;
    const int32_t* ja;
    double* w[kMaxOps];
    const double *cx, *cy, *cz;
    int32_t* info;
    uint8_t* status;
    int32_t max_refine;
    int* ws;  // plan workspace: [0] fallback count, [1..] fallback list
    hipStream_t stream;
};

// Sets the calling thread's rbfa_last_error text (host_io.cpp).
void set_last_error(const char* what);

// MI355X dispatches consecutive workgroups round-robin over its 8 XCDs,
// each with its own L2. Renumber so each XCD works on one contiguous range
// of (space-filling-curve ordered) stencils and its L2 sees neighbouring
// coordinates: block b runs block-range (b % 8) * chunk + b / 8.
__device__ __forceinline__ int xcd_block(int nblocks) {
    const int b = static_cast<int>(blockIdx.x);
    const int chunk = (nblocks + 7) / 8;
    return (b & 7) * chunk + (b >> 3);
}
inline unsigned xcd_grid(int nblocks) { return 8u * static_cast<unsigned>((nblocks + 7) / 8); }

// Systems larger than 64: one two-wave workgroup per stencil (defined in
// assemble_group.hip, instantiated there for the shipped shapes).
template <int DIM, int N, int P, int Q>
hipError_t launch_group_diff(const LaunchArgs& a);
template <int DIM, int N, int P, int Q>
hipError_t launch_group_mixed(const LaunchArgs& a);
// The null-space mixed form for systems larger than 64 unknowns whose
// stencils still fit a wave (assemble_group_mixed.hip): SYM symmetric
// steps of the reduced LU, OCC waves per SIMD.
template <int DIM, int N, int P, int Q, int SYM, int OCC>
hipError_t launch_ns_list(const LaunchArgs& a);
// Systems of 64 unknowns and more, blocked LU with FP64 MFMA trailing
// updates, one workgroup per stencil, one wave per 16-column tile
// (assemble_tile.hip).
template <int DIM, int N, int P, int Q>
hipError_t launch_tile_diff(const LaunchArgs& a);
// Systems of at most 32 unknowns, SL-lane segments, one stencil per
// segment, 64 / SL stencils per wave (assemble_pack.hip); BCM: pivot-row
// broadcast (0 LDS broadcast reads, 1 per-lane reads + DPP row_newbcast),
// OCC: waves per SIMD the registers are sized for.
template <int DIM, int N, int P, int Q, int R, bool INTERP, int SL, int BCM, int OCC>
hipError_t launch_pack(const LaunchArgs& a);

}  // namespace rbf_amd_internal

#endif  // RBF_ASSEMBLE_COMMON_H
