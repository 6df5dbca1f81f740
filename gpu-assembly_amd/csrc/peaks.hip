// peaks.hip -- microbenchmarks that re-measure, on the GPU at hand, the
// three peaks the assembly kernels are priced against: FP64 vector FMA
// throughput, FP64 matrix (v_mfma_f64_16x16x4f64) throughput and HBM copy
// bandwidth. The benchmark reports them beside the datasheet figures.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "rbf_assembly.h"

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef double dbl2 __attribute__((ext_vector_type(2)));

constexpr int kThreads = 256;
constexpr int kChains = 8;  // independent FMA chains per lane

// Every lane runs kChains independent fma chains of `iters` steps. The sum
// is stored only when it equals a value it never takes, so the work stays
// and the stores do not cost bandwidth.
__global__ void __launch_bounds__(kThreads) k_peak_fma64(double* out, int iters, double b,
                                                         double c) {
    const int g = static_cast<int>(blockIdx.x * kThreads + threadIdx.x);
    double a[kChains];
#pragma unroll
    for (int j = 0; j < kChains; ++j) a[j] = 1.0 + 1e-3 * (g + j);
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int j = 0; j < kChains; ++j) a[j] = fma(a[j], b, c);
    }
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < kChains; ++j) s += a[j];
    if (s == -1.0) out[g] = s;
}

// Four independent accumulator chains of v_mfma_f64_16x16x4f64 per wave.
__global__ void __launch_bounds__(kThreads) k_peak_mfma64(double* out, int iters, double b) {
    const int g = static_cast<int>(blockIdx.x * kThreads + threadIdx.x);
    const double a = 1e-3 * (threadIdx.x & 63);
    dbl4 acc[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[j] = dbl4{0.0, 0.0, 0.0, 0.0};
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[j], 0, 0, 0);
    }
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j) s += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
    if (s == -1.0) out[g] = s;
}

// Grid-stride copy of n 16-byte vectors, non-temporal on both sides.
__global__ void __launch_bounds__(kThreads) k_peak_copy(const dbl2* __restrict__ src,
                                                        dbl2* __restrict__ dst, int64_t n) {
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kThreads;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x; i < n; i += stride)
        __builtin_nontemporal_store(__builtin_nontemporal_load(&src[i]), &dst[i]);
}

template <class Launch>
hipError_t time_ms(Launch&& launch, hipStream_t st, int reps, float& ms) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {
        launch();  // warm-up
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    for (int r = 0; r < reps && e == hipSuccess; ++r) {
        launch();
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess) ms /= static_cast<float>(reps);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return e;
}

}  // namespace

extern "C" {

RBFA_API int32_t rbfa_measure_peaks(double* out, void* stream) {
    if (!out) return RBFA_ERR_ARG;
    hipStream_t st = static_cast<hipStream_t>(stream);
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return RBFA_ERR_NODEVICE;
    const int cus = prop.multiProcessorCount;
    // compute kernels: 8 workgroups of 4 waves per CU (8 waves per SIMD)
    const int blocks = 8 * cus;
    const int64_t lanes = static_cast<int64_t>(blocks) * kThreads;
    // copy: 1 GiB each way
    const int64_t nvec = (int64_t(1) << 30) / 16;
    double* sink = nullptr;
    dbl2 *src = nullptr, *dst = nullptr;
    if (hipMalloc(&sink, sizeof(double) * lanes) != hipSuccess) return RBFA_ERR_NOMEM;
    if (hipMalloc(&src, 16 * nvec) != hipSuccess || hipMalloc(&dst, 16 * nvec) != hipSuccess) {
        (void)hipFree(sink);
        if (src) (void)hipFree(src);
        return RBFA_ERR_NOMEM;
    }
    const int it_fma = 4096, it_mfma = 256;
    float ms_fma = 0.f, ms_mfma = 0.f, ms_copy = 0.f;
    hipError_t e = time_ms(
        [&] {
            hipLaunchKernelGGL(k_peak_fma64, dim3(blocks), dim3(kThreads), 0, st, sink, it_fma,
                               1.0000001, 1e-9);
        },
        st, 3, ms_fma);
    if (e == hipSuccess)
        e = time_ms(
            [&] {
                hipLaunchKernelGGL(k_peak_mfma64, dim3(blocks), dim3(kThreads), 0, st, sink,
                                   it_mfma, 1.0);
            },
            st, 3, ms_mfma);
    if (e == hipSuccess) e = hipMemsetAsync(src, 0, 16 * nvec, st);
    if (e == hipSuccess)
        e = time_ms(
            [&] {
                hipLaunchKernelGGL(k_peak_copy, dim3(16 * cus), dim3(kThreads), 0, st, src, dst,
                                   nvec);
            },
            st, 3, ms_copy);
    (void)hipStreamSynchronize(st);
    (void)hipFree(sink);
    (void)hipFree(src);
    (void)hipFree(dst);
    if (e != hipSuccess) return RBFA_ERR_HIP;
    const double fma_flop = 2.0 * kChains * 4.0 * it_fma * static_cast<double>(lanes);
    const double waves = static_cast<double>(lanes) / 64.0;
    const double mfma_flop = 2.0 * 16 * 16 * 4 * 4.0 * 4.0 * it_mfma * waves;
    out[0] = fma_flop / (ms_fma * 1e-3) / 1e12;
    out[1] = mfma_flop / (ms_mfma * 1e-3) / 1e12;
    out[2] = 2.0 * 16.0 * static_cast<double>(nvec) / (ms_copy * 1e-3) / 1e9;
    return RBFA_OK;
}

}  // extern "C"
