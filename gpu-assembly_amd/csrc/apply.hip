// apply.hip -- operator application: y_op = W_op u for the assembled CSR
// operators, then y_op[b] = u[b] on the listed boundary rows: the mv_op of
// the reference's pipeline sketch (main.cpp:59-66: A.spmv(x, y), then
// y[bnd] = x[bnd]).
//
// Two forms, both HBM-bound (per row: n weights per operator at 8 B, the
// column information once for all operators, the result per operator):
//
// rbfa_apply (CSR as it is): eight lanes share a row, lane q reads entries
// q, q+8, ..., and the eight partial sums meet in three shuffle steps. Every
// nonzero gathers u[ja[e]] from global memory; those 8-byte gathers, not the
// weight stream, bound this form.
//
// rbfa_apply_plan_* (blocked layout, built once per structure): a
// workgroup owns a block of consecutive rows, gathers each distinct column
// of the block once into LDS, streams weights and 16-bit local indices with
// coalesced non-temporal loads, forms every product into LDS and sums each
// row left to right -- the same products in the same order as a sequential
// CSR product, so the result is bit-identical to it.

#include <hip/hip_runtime.h>

#include <cstdint>

#include <cstring>
#include <new>

#include "apply_layout.h"
#include "assemble_common.h"
#include "rbf_assembly.h"

namespace {

using namespace rbf_amd_internal;

constexpr int kLanesPerRow = 8;

// Per-operator arrays passed by value (one pointer per operator).
template <int R>
struct OpPtrs {
    double* p[R];
};

template <int R>
__global__ void __launch_bounds__(256)
k_apply(int nrows, int k, const int* __restrict__ ia, const int* __restrict__ ja,
        OpPtrs<R> w, const double* __restrict__ u, OpPtrs<R> y) {
    const int gid = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    const int row = gid / kLanesPerRow;
    const int q = gid % kLanesPerRow;
    double acc[R];
#pragma unroll
    for (int c = 0; c < R; ++c) acc[c] = 0.0;
    if (row < nrows) {
        const long long a = ia ? ia[row] : static_cast<long long>(row) * k;
        const long long b = ia ? ia[row + 1] : a + k;
        for (long long e = a + q; e < b; e += kLanesPerRow) {
            const double ue = u[ja[e]];
#pragma unroll
            for (int c = 0; c < R; ++c) acc[c] = fma(w.p[c][e], ue, acc[c]);
        }
    }
#pragma unroll
    for (int off = kLanesPerRow / 2; off >= 1; off >>= 1) {
#pragma unroll
        for (int c = 0; c < R; ++c) acc[c] += __shfl_xor(acc[c], off, kLanesPerRow);
    }
    if (row < nrows && q == 0) {
#pragma unroll
        for (int c = 0; c < R; ++c) y.p[c][row] = acc[c];
    }
}

template <int R>
__global__ void k_identity_rows(int nb, const int* __restrict__ bnd, const double* __restrict__ u,
                                OpPtrs<R> y) {
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= nb) return;
    const int r = bnd[i];
#pragma unroll
    for (int c = 0; c < R; ++c) y.p[c][r] = u[r];
}

template <int R>
hipError_t launch_apply(int nrows, int k, const int* ia, const int* ja, double* const* w,
                        const double* u, double* const* y, int nb, const int* bnd,
                        hipStream_t st) {
    OpPtrs<R> wp, yp;
    for (int c = 0; c < R; ++c) {
        wp.p[c] = w[c];
        yp.p[c] = y[c];
    }
    const long long threads = static_cast<long long>(nrows) * kLanesPerRow;
    const unsigned grid = static_cast<unsigned>((threads + 255) / 256);
    hipLaunchKernelGGL(k_apply<R>, dim3(grid), dim3(256), 0, st, nrows, k, ia, ja, wp, u, yp);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || nb == 0) return e;
    hipLaunchKernelGGL(k_identity_rows<R>, dim3((nb + 255) / 256), dim3(256), 0, st, nb, bnd, u, yp);
    return hipGetLastError();
}

// Planned form: one workgroup per row block (XCD-contiguous block ranges).
// LDS: prod[R][kApplyEntries] then su[umax].
template <int R>
__global__ void __launch_bounds__(256)
k_apply_blocked(int nblk, const int* __restrict__ blk_row, const int* __restrict__ ia, int k,
                const int* __restrict__ uoff, const int* __restrict__ ucols,
                const uint16_t* __restrict__ lidx, OpPtrs<R> w, const double* __restrict__ u,
                OpPtrs<R> y, const uint8_t* __restrict__ ident) {
    extern __shared__ double lds[];
    constexpr int E = kApplyEntries;
    constexpr int NI = E / 256;
    double* prod = lds;
    double* su = lds + R * E;
    const int b = xcd_block(nblk);
    if (b >= nblk) return;
    const int t = static_cast<int>(threadIdx.x);
    const int r0 = blk_row[b], r1 = blk_row[b + 1];
    const long long e0 = ia ? ia[r0] : static_cast<long long>(r0) * k;
    const long long e1 = ia ? ia[r1] : static_cast<long long>(r1) * k;
    const int ne = static_cast<int>(e1 - e0);
    const int u0 = uoff[b], nu = uoff[b + 1] - u0;
    for (int i = t; i < nu; i += 256) su[i] = u[ucols[u0 + i]];
    // the block's weights and local indices are read once: non-temporal
    int li[NI];
    double wv[R][NI];
#pragma unroll
    for (int i = 0; i < NI; ++i) {
        const int e = t + 256 * i;
        li[i] = e < ne ? static_cast<int>(__builtin_nontemporal_load(lidx + e0 + e)) : 0;
#pragma unroll
        for (int c = 0; c < R; ++c)
            wv[c][i] = e < ne ? __builtin_nontemporal_load(w.p[c] + e0 + e) : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NI; ++i) {
        const int e = t + 256 * i;
        if (e < ne) {
            const double ue = su[li[i]];
#pragma unroll
            for (int c = 0; c < R; ++c) prod[c * E + e] = wv[c][i] * ue;
        }
    }
    __syncthreads();
    const int rows = r1 - r0;
    for (int p = t; p < R * rows; p += 256) {
        const int c = p / rows;
        const int row = r0 + (p - c * rows);
        const int a = static_cast<int>((ia ? ia[row] : static_cast<long long>(row) * k) - e0);
        const int z = static_cast<int>((ia ? ia[row + 1] : static_cast<long long>(row + 1) * k) - e0);
        const double* pr = prod + c * E;
        double s = 0.0;
#pragma unroll 8
        for (int j = a; j < z; ++j) s += pr[j];
        if (ident && ident[row]) s = u[row];
        y.p[c][row] = s;
    }
}

template <int R>
size_t blocked_lds_bytes(int umax) {
    return sizeof(double) * (static_cast<size_t>(R) * kApplyEntries + static_cast<size_t>(umax));
}

}  // namespace

struct rbfa_apply_plan {
    int32_t nrows = 0, ncols = 0, k = 0, nblk = 0, umax = 0;
    int64_t nnz = 0, nucols = 0;
    int32_t* blk_row = nullptr;
    int32_t* ia = nullptr;
    int32_t* uoff = nullptr;
    int32_t* ucols = nullptr;
    uint16_t* lidx = nullptr;
    uint8_t* ident = nullptr;
    int64_t device_bytes = 0;
};

namespace {

void free_apply_plan(rbfa_apply_plan* p) {
    if (!p) return;
    void* bufs[] = {p->blk_row, p->ia, p->uoff, p->ucols, p->lidx, p->ident};
    for (void* q : bufs)
        if (q) (void)hipFree(q);
    delete p;
}

template <typename T>
hipError_t upload(T** dst, const T* src, size_t count, int64_t* total) {
    *dst = nullptr;
    if (count == 0) return hipSuccess;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(dst), count * sizeof(T));
    if (e != hipSuccess) return e;
    *total += static_cast<int64_t>(count * sizeof(T));
    return hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice);
}

template <int R>
hipError_t launch_blocked(const rbfa_apply_plan* p, double* const* w, const double* u,
                          double* const* y, hipStream_t st) {
    OpPtrs<R> wp, yp;
    for (int c = 0; c < R; ++c) {
        wp.p[c] = w[c];
        yp.p[c] = y[c];
    }
    const size_t lds = blocked_lds_bytes<R>(p->umax);
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_apply_blocked<R>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       static_cast<int>(lds));
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_apply_blocked<R>, dim3(xcd_grid(p->nblk)), dim3(256), lds, st, p->nblk,
                       p->blk_row, p->ia, p->k, p->uoff, p->ucols, p->lidx, wp, u, yp, p->ident);
    return hipGetLastError();
}

}  // namespace

extern "C" {

RBFA_API int32_t rbfa_apply(int32_t nrows, int32_t k, const int32_t* ia, const int32_t* ja,
                            int32_t nops, double* const* w, const double* u, double* const* y,
                            int32_t nbnd, const int32_t* bnd, void* stream) {
    if (nrows < 0 || nops < 1 || nops > 4 || !w || !y || nbnd < 0 || (nbnd > 0 && !bnd))
        return RBFA_ERR_ARG;
    if (nrows == 0) return RBFA_OK;
    if (!ja || !u || (!ia && k < 1)) return RBFA_ERR_ARG;
    for (int c = 0; c < nops; ++c)
        if (!w[c] || !y[c]) return RBFA_ERR_ARG;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipError_t e = hipSuccess;
    switch (nops) {
        case 1: e = launch_apply<1>(nrows, k, ia, ja, w, u, y, nbnd, bnd, st); break;
        case 2: e = launch_apply<2>(nrows, k, ia, ja, w, u, y, nbnd, bnd, st); break;
        case 3: e = launch_apply<3>(nrows, k, ia, ja, w, u, y, nbnd, bnd, st); break;
        default: e = launch_apply<4>(nrows, k, ia, ja, w, u, y, nbnd, bnd, st); break;
    }
    return e == hipSuccess ? RBFA_OK : RBFA_ERR_HIP;
}

RBFA_API int32_t rbfa_apply_plan_create(int32_t nrows, int32_t ncols, int32_t k,
                                        const int32_t* ia, const int32_t* ja, int32_t nbnd,
                                        const int32_t* bnd, rbfa_apply_plan** plan) {
    if (!plan) return RBFA_ERR_ARG;
    *plan = nullptr;
    if (nrows < 1 || ncols < 1 || !ja || (!ia && k < 1) || nbnd < 0 || (nbnd > 0 && !bnd))
        return RBFA_ERR_ARG;
    for (int32_t i = 0; i < nbnd; ++i)
        if (bnd[i] < 0 || bnd[i] >= nrows || bnd[i] >= ncols) return RBFA_ERR_ARG;
    ApplyLayout L;
    try {
        if (build_apply_layout(nrows, ncols, ia ? 0 : k, ia, ja, L) != 0) return RBFA_ERR_ARG;
    } catch (const std::bad_alloc&) {
        return RBFA_ERR_NOMEM;
    }
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess ||
        std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return RBFA_ERR_NODEVICE;
    // the largest instantiation must fit the workgroup's LDS
    if (blocked_lds_bytes<4>(L.umax) > static_cast<size_t>(prop.sharedMemPerBlock))
        return RBFA_ERR_UNSUPPORTED;
    rbfa_apply_plan* p = new (std::nothrow) rbfa_apply_plan;
    if (!p) return RBFA_ERR_NOMEM;
    p->nrows = nrows;
    p->ncols = ncols;
    p->k = ia ? 0 : k;
    p->nblk = static_cast<int32_t>(L.blk_row.size()) - 1;
    p->umax = L.umax;
    p->nnz = static_cast<int64_t>(L.lidx.size());
    p->nucols = static_cast<int64_t>(L.ucols.size());
    std::vector<uint8_t> flag;
    if (nbnd > 0) {
        flag.assign(nrows, 0);
        for (int32_t i = 0; i < nbnd; ++i) flag[bnd[i]] = 1;
    }
    hipError_t e = hipSuccess;
    int64_t* tot = &p->device_bytes;
    if (e == hipSuccess) e = upload(&p->blk_row, L.blk_row.data(), L.blk_row.size(), tot);
    if (e == hipSuccess && ia) e = upload(&p->ia, ia, static_cast<size_t>(nrows) + 1, tot);
    if (e == hipSuccess) e = upload(&p->uoff, L.blk_uoff.data(), L.blk_uoff.size(), tot);
    if (e == hipSuccess) e = upload(&p->ucols, L.ucols.data(), L.ucols.size(), tot);
    if (e == hipSuccess) e = upload(&p->lidx, L.lidx.data(), L.lidx.size(), tot);
    if (e == hipSuccess && nbnd > 0) e = upload(&p->ident, flag.data(), flag.size(), tot);
    if (e != hipSuccess) {
        free_apply_plan(p);
        return e == hipErrorOutOfMemory ? RBFA_ERR_NOMEM : RBFA_ERR_HIP;
    }
    *plan = p;
    return RBFA_OK;
}

RBFA_API int32_t rbfa_apply_plan_destroy(rbfa_apply_plan* plan) {
    free_apply_plan(plan);
    return RBFA_OK;
}

RBFA_API int32_t rbfa_apply_plan_stats(const rbfa_apply_plan* plan, int64_t* out) {
    if (!plan || !out) return RBFA_ERR_ARG;
    out[0] = plan->nblk;
    out[1] = plan->umax;
    out[2] = plan->nucols;
    out[3] = plan->nnz;
    out[4] = plan->device_bytes;
    return RBFA_OK;
}

RBFA_API int32_t rbfa_apply_plan_run(const rbfa_apply_plan* plan, int32_t nops, double* const* w,
                                     const double* u, double* const* y, void* stream) {
    if (!plan || nops < 1 || nops > 4 || !w || !y || !u) return RBFA_ERR_ARG;
    for (int c = 0; c < nops; ++c)
        if (!w[c] || !y[c]) return RBFA_ERR_ARG;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipError_t e = hipSuccess;
    switch (nops) {
        case 1: e = launch_blocked<1>(plan, w, u, y, st); break;
        case 2: e = launch_blocked<2>(plan, w, u, y, st); break;
        case 3: e = launch_blocked<3>(plan, w, u, y, st); break;
        default: e = launch_blocked<4>(plan, w, u, y, st); break;
    }
    return e == hipSuccess ? RBFA_OK : RBFA_ERR_HIP;
}

}  // extern "C"
