// assemble.hip -- device side of librbf_amd: kernel instantiations for
// gfx950 and the C ABI (rbfa_plan_* / rbfa_assemble) that launches them.
//
// The kernels are the templates of include/rbf_operators.h; this file
// picks the shapes the library ships, maps blocks to stencils XCD-aware,
// and replaces the reference's compile-time interp_config choice
// (rbf_operators.h:187-290) with a run-time plan.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <new>

#include "rbf_assembly.h"
#include "rbf_operators.h"
#include "rbf_mixed.h"

namespace {

using namespace rbf_operators;

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
    hipStream_t stream;
};

// MI355X dispatches consecutive workgroups round-robin over its 8 XCDs,
// each with its own L2. Renumber so each XCD works on one contiguous range
// of (space-filling-curve ordered) stencils and its L2 sees neighbouring
// coordinates: block b runs stencil (b % 8) * chunk + b / 8.
__device__ __forceinline__ int xcd_stencil(int nstencils) {
    const int b = static_cast<int>(blockIdx.x);
    const int chunk = (nstencils + 7) / 8;
    return (b & 7) * chunk + (b >> 3);
}
inline unsigned xcd_grid(int nstencils) { return 8u * static_cast<unsigned>((nstencils + 7) / 8); }

template <class E>
__global__ void __launch_bounds__(64)
k_assemble_fp64(int nstencils, const double* __restrict__ x, const double* __restrict__ y,
                const double* __restrict__ z, const int* __restrict__ ja,
                detail::ValueOut<double, E::NC - E::NT> out, const double* __restrict__ cx,
                const double* __restrict__ cy, const double* __restrict__ cz,
                int* __restrict__ info, uint8_t* __restrict__ status) {
    __shared__ double lds[E::lds_doubles];
    const int s = xcd_stencil(nstencils);
    if (s >= nstencils) return;
    E::run(s, x, y, z, ja, out, cx, cy, cz, lds, info);
    if (status && detail::lane_id() == 0) status[s] = 0;
}

template <class E>
hipError_t launch_fp64(const LaunchArgs& a) {
    constexpr int R = E::NC - E::NT;
    detail::ValueOut<double, R> out;
    for (int c = 0; c < R; ++c) out.p[c] = a.w[c];
    hipLaunchKernelGGL(k_assemble_fp64<E>, dim3(xcd_grid(a.nstencils)), dim3(64), 0, a.stream,
                       a.nstencils, a.x, a.y, a.z, a.ja, out, a.cx, a.cy, a.cz, a.info, a.status);
    return hipGetLastError();
}

template <class E>
__global__ void __launch_bounds__(64)
k_assemble_mixed(int nstencils, const double* __restrict__ x, const double* __restrict__ y,
                 const double* __restrict__ z, const int* __restrict__ ja,
                 detail::ValueOut<double, E::R> out, int* __restrict__ info,
                 uint8_t* __restrict__ status, int max_refine) {
    __shared__ double lds[E::lds_doubles];
    const int s = xcd_stencil(nstencils);
    if (s >= nstencils) return;
    E::run(s, x, y, z, ja, out, lds, info, status, max_refine);
}

template <class E>
hipError_t launch_mixed(const LaunchArgs& a) {
    detail::ValueOut<double, E::R> out;
    for (int c = 0; c < E::R; ++c) out.p[c] = a.w[c];
    hipLaunchKernelGGL(k_assemble_mixed<E>, dim3(xcd_grid(a.nstencils)), dim3(64), 0, a.stream,
                       a.nstencils, a.x, a.y, a.z, a.ja, out, a.info, a.status, a.max_refine);
    return hipGetLastError();
}

using LaunchFn = hipError_t (*)(const LaunchArgs&);

struct Shape {
    int dim, n, P, Q, op_set, nrhs, precision;
    int nt, nops;
    LaunchFn fn;
};

template <int DIM, int N, int P, int Q>
constexpr Shape diff64() {
    using E = detail::WaveAssembly<DIM, N, P, Q, DIM + 1, false>;
    return {DIM, N, P, Q, RBFA_OPS_DIFF, 0, RBFA_PREC_FP64, E::NT, DIM + 1, &launch_fp64<E>};
}
template <int N, int P, int Q, int R>
constexpr Shape interp64() {
    using E = detail::WaveAssembly<2, N, P, Q, R, true>;
    return {2, N, P, Q, RBFA_OPS_INTERP, R, RBFA_PREC_FP64, E::NT, R, &launch_fp64<E>};
}
template <int DIM, int N, int P, int Q>
constexpr Shape diffmix() {
    using E = detail::WaveMixed<DIM, N, P, Q>;
    return {DIM, N, P, Q, RBFA_OPS_DIFF, 0, RBFA_PREC_MIXED, E::NT, DIM + 1, &launch_mixed<E>};
}

// The shapes this library ships: the stencil families of the workloads it
// is benchmarked on (n = 13/21/35 interior and boundary classes, r^3/r^5
// PHS, degree 2-3 polynomials) and the D2Q9 interpolation stencils of the
// reference's CUDA-Fortran demo (n = 9, NRHS = 9, main.f90:16).
const Shape kShapes[] = {
    diff64<2, 13, 2, 3>(),  diff64<2, 21, 2, 3>(), diff64<2, 35, 3, 5>(),
    diff64<2, 35, 3, 3>(),  diff64<2, 21, 2, 5>(), diff64<2, 13, 1, 3>(),
    interp64<9, 2, 3, 9>(), interp64<21, 2, 3, 9>(),
    diffmix<2, 21, 2, 3>(), diffmix<2, 35, 3, 5>(), diffmix<2, 13, 2, 3>(),
};

}  // namespace

struct rbfa_plan {
    rbfa_plan_desc desc;
    Shape shape;
};

extern "C" {

RBFA_API const char* rbfa_version(void) { return "rbf_amd 0.1 (gfx950)"; }

RBFA_API int32_t rbfa_plan_create(const rbfa_plan_desc* desc, rbfa_plan** plan) {
    if (!desc || !plan) return RBFA_ERR_ARG;
    *plan = nullptr;
    const rbfa_plan_desc& d = *desc;
    if ((d.dim != 2 && d.dim != 3) || d.n < 1 || d.degree < 0 || d.phs < 3 || d.phs % 2 == 0)
        return RBFA_ERR_ARG;
    if (d.op_set != RBFA_OPS_DIFF && d.op_set != RBFA_OPS_INTERP) return RBFA_ERR_ARG;
    if (d.op_set == RBFA_OPS_INTERP && (d.nrhs < 1 || d.nrhs > kMaxOps)) return RBFA_ERR_ARG;
    if (d.precision != RBFA_PREC_FP64 && d.precision != RBFA_PREC_MIXED) return RBFA_ERR_ARG;
    if (d.precision == RBFA_PREC_MIXED && (d.max_refine < 1 || d.max_refine > 3))
        return RBFA_ERR_ARG;
    for (const Shape& s : kShapes) {
        if (s.dim == d.dim && s.n == d.n && s.P == d.degree && s.Q == d.phs &&
            s.op_set == d.op_set && s.precision == d.precision &&
            (d.op_set == RBFA_OPS_DIFF || s.nrhs == d.nrhs)) {
            int dev = 0;
            if (hipGetDevice(&dev) != hipSuccess) return RBFA_ERR_NODEVICE;
            hipDeviceProp_t prop;
            if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return RBFA_ERR_NODEVICE;
            if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return RBFA_ERR_NODEVICE;
            auto* p = new (std::nothrow) rbfa_plan{d, s};
            if (!p) return RBFA_ERR_NOMEM;
            *plan = p;
            return RBFA_OK;
        }
    }
    return RBFA_ERR_UNSUPPORTED;
}

RBFA_API int32_t rbfa_plan_destroy(rbfa_plan* plan) {
    delete plan;
    return RBFA_OK;
}

RBFA_API int32_t rbfa_plan_nops(const rbfa_plan* plan) { return plan ? plan->shape.nops : -1; }

RBFA_API int32_t rbfa_plan_system_size(const rbfa_plan* plan) { return plan ? plan->shape.nt : -1; }

RBFA_API int32_t rbfa_assemble(const rbfa_plan* plan, int32_t nstencils, const double* x,
                               const double* y, const double* z, const int32_t* ja,
                               double* const* w, const double* xc, const double* yc,
                               const double* zc, int32_t* info, uint8_t* status, void* stream) {
    if (!plan || nstencils < 0) return RBFA_ERR_ARG;
    if (nstencils == 0) return RBFA_OK;
    const Shape& sh = plan->shape;
    if (!x || !y || !ja || !w || (sh.dim == 3 && !z)) return RBFA_ERR_ARG;
    if (static_cast<int64_t>(nstencils) * sh.n > INT32_MAX) return RBFA_ERR_ARG;
    if (sh.op_set == RBFA_OPS_INTERP && (!xc || !yc || (sh.dim == 3 && !zc))) return RBFA_ERR_ARG;
    LaunchArgs a{};
    a.nstencils = nstencils;
    a.x = x;
    a.y = y;
    a.z = z;
    a.ja = ja;
    for (int c = 0; c < sh.nops; ++c) {
        if (!w[c]) return RBFA_ERR_ARG;
        a.w[c] = w[c];
    }
    a.cx = xc;
    a.cy = yc;
    a.cz = zc;
    a.info = info;
    a.status = status;
    a.max_refine = plan->desc.max_refine;
    a.stream = static_cast<hipStream_t>(stream);
    return sh.fn(a) == hipSuccess ? RBFA_OK : RBFA_ERR_HIP;
}

}  // extern "C"
