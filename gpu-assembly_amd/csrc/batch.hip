// batch.hip -- uneven batches: stencils of several sizes (size classes) in
// one ragged CSR graph, assembled either
//   mode 0: one launch per size class over that class's stencil list, or
//   mode 1: one persistent launch whose workgroups draw stencils from
//           size-bucketed work queues (atomic tickets): the largest class
//           first, a whole four-wave workgroup per stencil (blocked LU with
//           FP64 MFMA trailing updates), then each wave on its own through
//           the smaller classes.
// Either pipeline is captured once into a hipGraph at bind time and
// replayed by rbfa_batch_launch; rbfa_batch_autotune picks each class's
// kernel variant and the faster pipeline.
//
// The reference kernel is fixed-width with an implicit row pointer
// (rbf_operators.h:307-310); ragged rows exist only in its CUDA-Fortran
// fill (rbf_cuda.f90:52-56) and graph reader (rbf_io.h:649-707).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <new>
#include <vector>

#include "assemble_common.h"
#include "rbf_assembly.h"
#include "rbf_group.h"
#include "rbf_wave.h"
#include "rbf_wavepack.h"
#include "rbf_tile.h"

namespace {

using namespace rbf_operators;
using namespace rbf_amd_internal;

constexpr int kMaxClasses = 4;
using Out3 = detail::ValueOut<double, 3>;

// ---- per-class list kernels ------------------------------------------------

template <class E, int WPB>
__global__ void __launch_bounds__(64 * WPB)
k_list_wave(int count, const int* __restrict__ order, const int* __restrict__ ia,
            const double* __restrict__ x, const double* __restrict__ y,
            const double* __restrict__ z, const int* __restrict__ ja, Out3 out,
            int* __restrict__ info, uint8_t* __restrict__ status) {
    __shared__ __align__(16) double lds[WPB * E::lds_doubles];
    const int nblocks = (count + WPB - 1) / WPB;
    const int b = xcd_block(nblocks);
    const int w = static_cast<int>(threadIdx.x >> 6);
    const int i = b * WPB + w;
    if (b >= nblocks || i >= count) return;
    const int s = order[i];
    E::run(s, static_cast<size_t>(ia[s]), x, y, z, ja, out, nullptr, nullptr, nullptr,
           lds + w * E::lds_doubles, info);
    if (status && detail::lane_id() == 0) status[s] = 0;
}

template <class G, int THREADS, int OCC>
__global__ void __launch_bounds__(THREADS, OCC)
k_list_group(int count, const int* __restrict__ order, const int* __restrict__ ia,
             const double* __restrict__ x, const double* __restrict__ y,
             const double* __restrict__ z, const int* __restrict__ ja, Out3 out,
             int* __restrict__ info, uint8_t* __restrict__ status) {
    __shared__ __align__(16) double lds[G::lds_doubles];
    const int i = xcd_block(count);
    if (i >= count) return;
    const int s = order[i];
    G::run(s, static_cast<size_t>(ia[s]), x, y, z, ja, out, lds, info);
    if (status && threadIdx.x == 0) status[s] = 0;
}

// Several stencils per wave (16- or 32-lane segments, rbf_wavepack.h) from
// the class list, one wave per workgroup.
template <class E, int OCC>
__global__ void __launch_bounds__(64, OCC)
k_list_pack(int count, const int* __restrict__ order, const int* __restrict__ ia,
            const double* __restrict__ x, const double* __restrict__ y,
            const double* __restrict__ z, const int* __restrict__ ja, Out3 out,
            int* __restrict__ info, uint8_t* __restrict__ status) {
    __shared__ __align__(16) double lds[E::lds_doubles];
    const int nwaves = (count + E::S - 1) / E::S;
    const int b = xcd_block(nwaves);
    if (b >= nwaves) return;
    E::run_list(b * E::S, count, order, ia, x, y, z, ja, out, lds, info, status);
}

// ---- the persistent kernel -------------------------------------------------

struct Queues {
    const int* order;          // stencil ids, grouped by class
    int start[kMaxClasses];    // first entry of each class in order
    int count[kMaxClasses];    // entries of each class
    int* tickets;              // next unclaimed entry of each class (zeroed before launch)
};

__global__ void k_zero(int* p, int n) {
    if (static_cast<int>(threadIdx.x) < n) p[threadIdx.x] = 0;
}

template <class G, class E2, class E1, class E0>
struct Persistent {
    static constexpr int waves = G::threads / 64;
    static constexpr int wave_lds =
        std::max(E2::lds_doubles, std::max(E1::lds_doubles, E0::lds_doubles));
    static constexpr int lds_doubles = std::max(G::lds_doubles, waves * wave_lds);

    template <class E>
    __device__ static void drain(const Queues& q, int c, const int* __restrict__ ia,
                                 const double* __restrict__ x, const double* __restrict__ y,
                                 const int* __restrict__ ja, Out3 out, int* __restrict__ info,
                                 uint8_t* __restrict__ status, double* lds) {
        const int lane = detail::lane_id();
        for (;;) {
            int t = 0;
            if (lane == 0) t = atomicAdd(&q.tickets[c], 1);
            t = __builtin_amdgcn_readfirstlane(__shfl(t, 0, 64));
            if (t >= q.count[c]) break;
            const int s = q.order[q.start[c] + t];
            E::run(s, static_cast<size_t>(ia[s]), x, y, nullptr, ja, out, nullptr, nullptr,
                   nullptr, lds, info);
            if (status && lane == 0) status[s] = 0;
        }
    }
};

// classes: 0 = n13 (P2, r^3), 1 = n21 (P2, r^3), 2 = n35 (P3, r^5), 3 = n56 (P4, r^7)
using EP13 = detail::WaveAssembly<2, 13, 2, 3, 3, false, detail::LuTune<2, 1>>;
using EP21 = detail::WaveAssembly<2, 21, 2, 3, 3, false, detail::LuTune<2, 1>>;
using EP35 = detail::WaveAssembly<2, 35, 3, 5, 3, false, detail::LuTune<2, 1>>;
using GP56 = detail::GroupAssembly<2, 56, 4, 7, 3, false, 2>;
using TP56 = detail::TileAssembly<2, 56, 4, 7>;
using PK = Persistent<TP56, EP35, EP21, EP13>;
constexpr int kPersistentThreads = TP56::threads;
// The persistent launch (mode 1) is switched off: a run of it faulted the
// GPU (memory aperture violation) after the single-wave engine's fill and
// back substitution changed, with no cause found in its code or in the
// compiled kernel. Mode 0 (one launch per class, the same engines, all in
// one captured graph) is unaffected. The kernel stays compiled.
constexpr bool kPersistentEnabled = false;

__global__ void __launch_bounds__(kPersistentThreads)
k_persistent(Queues q, const int* __restrict__ ia, const double* __restrict__ x,
             const double* __restrict__ y, const int* __restrict__ ja, Out3 out,
             int* __restrict__ info, uint8_t* __restrict__ status) {
    __shared__ __align__(16) double lds[PK::lds_doubles];
    __shared__ int tk;
    // largest class first: a whole workgroup per stencil
    for (;;) {
        if (threadIdx.x == 0) tk = atomicAdd(&q.tickets[3], 1);
        __syncthreads();
        const int t = tk;
        __syncthreads();
        if (t >= q.count[3]) break;
        const int s = q.order[q.start[3] + t];
        TP56::run(s, static_cast<size_t>(ia[s]), x, y, nullptr, ja, out, lds, info);
        if (status && threadIdx.x == 0) status[s] = 0;
        __syncthreads();  // LDS reuse by the next stencil
    }
    // then each wave on its own, larger classes first
    double* wl = lds + (threadIdx.x >> 6) * PK::wave_lds;
    PK::drain<EP35>(q, 2, ia, x, y, ja, out, info, status, wl);
    PK::drain<EP21>(q, 1, ia, x, y, ja, out, info, status, wl);
    PK::drain<EP13>(q, 0, ia, x, y, ja, out, info, status, wl);
}

// ---- class registry --------------------------------------------------------

struct ListArgs {
    int count;
    const int* order;
    const int* ia;
    const double *x, *y, *z;
    const int* ja;
    Out3 out;
    int* info;
    uint8_t* status;
    hipStream_t stream;
};
using ListFn = hipError_t (*)(const ListArgs&);

template <class E, int WPB>
hipError_t launch_list_wave(const ListArgs& a) {
    if (a.count == 0) return hipSuccess;
    const int nblocks = (a.count + WPB - 1) / WPB;
    hipLaunchKernelGGL((k_list_wave<E, WPB>), dim3(xcd_grid(nblocks)), dim3(64 * WPB), 0, a.stream,
                       a.count, a.order, a.ia, a.x, a.y, a.z, a.ja, a.out, a.info, a.status);
    return hipGetLastError();
}
template <class E, int OCC>
hipError_t launch_list_pack(const ListArgs& a) {
    if (a.count == 0) return hipSuccess;
    const int nwaves = (a.count + E::S - 1) / E::S;
    hipLaunchKernelGGL((k_list_pack<E, OCC>), dim3(xcd_grid(nwaves)), dim3(64), 0, a.stream,
                       a.count, a.order, a.ia, a.x, a.y, a.z, a.ja, a.out, a.info, a.status);
    return hipGetLastError();
}
template <class G, int THREADS, int OCC = 1>
hipError_t launch_list_group(const ListArgs& a) {
    if (a.count == 0) return hipSuccess;
    hipLaunchKernelGGL((k_list_group<G, THREADS, OCC>), dim3(xcd_grid(a.count)), dim3(THREADS), 0, a.stream,
                       a.count, a.order, a.ia, a.x, a.y, a.z, a.ja, a.out, a.info, a.status);
    return hipGetLastError();
}

struct ListVariant {
    const char* name;
    ListFn fn;
};
constexpr int kMaxClassVariants = 6;
struct BatchClass {
    int n, P, Q, nvariants, def;  // def: the variant used until autotuned
    ListVariant v[kMaxClassVariants];
};

// The shapes a class's list can run in: stencils per wave (packed 16- or
// 32-lane segments where the system fits 32 lanes), waves per workgroup
// (1, 2 or 4 for one stencil per wave), and the wave LU's tuning. The
// autotuner times each on the class's own stencils.
template <int N, int P, int Q>
constexpr BatchClass wave_class() {
    using T0 = detail::WaveAssembly<2, N, P, Q, 3, false, detail::LuTune<4, 2, 1, 0, 1>>;
    using T1 = detail::WaveAssembly<2, N, P, Q, 3, false, detail::LuTune<2, 1>>;
    using T2 = detail::WaveAssembly<2, N, P, Q, 3, false, detail::LuTune<0, 2, 1, 1, 1>>;
    constexpr int NT = N + npoly(P);
    if constexpr (NT <= 32) {
        constexpr int SL = NT <= 16 ? 16 : 32;
        constexpr int OCC2 = SL == 16 ? 4 : 3;
        using PK1 = detail::WavePack<2, N, P, Q, 3, false, SL, 0, 1>;
        using PK0 = detail::WavePack<2, N, P, Q, 3, false, SL, 0, 0>;
        return {N, P, Q, 6, 0,
                {{SL == 16 ? "packed: 4 stencils per wave bc=newbcast occ=4"
                           : "packed: 2 stencils per wave bc=newbcast occ=3",
                  &launch_list_pack<PK1, OCC2>},
                 {SL == 16 ? "packed: 4 stencils per wave bc=newbcast"
                           : "packed: 2 stencils per wave bc=newbcast",
                  &launch_list_pack<PK1, 1>},
                 {SL == 16 ? "packed: 4 stencils per wave" : "packed: 2 stencils per wave",
                  &launch_list_pack<PK0, 1>},
                 {"bc=third pv=dpp-bcast pin fill=pairs", &launch_list_wave<T0, 1>},
                 {"bc=half pv=dpp", &launch_list_wave<T1, 1>},
                 {"bc=readlane pv=dpp-bcast pin fill=pairs bs=lds", &launch_list_wave<T2, 1>}}};
    } else {
        return {N, P, Q, 5, 0,
                {{"bc=third pv=dpp-bcast pin fill=pairs", &launch_list_wave<T0, 1>},
                 {"bc=third pv=dpp-bcast pin fill=pairs, 2 waves per workgroup",
                  &launch_list_wave<T0, 2>},
                 {"bc=third pv=dpp-bcast pin fill=pairs, 4 waves per workgroup",
                  &launch_list_wave<T0, 4>},
                 {"bc=half pv=dpp", &launch_list_wave<T1, 1>},
                 {"bc=readlane pv=dpp-bcast pin fill=pairs bs=lds", &launch_list_wave<T2, 1>}}};
    }
}

const BatchClass kClasses[kMaxClasses] = {
    wave_class<13, 2, 3>(),
    wave_class<21, 2, 3>(),
    wave_class<35, 3, 5>(),
    {56, 4, 7, 2, 0,
     {{"blocked lu fp64 mfma", &launch_list_group<TP56, TP56::threads, TP56::occupancy>},
      {"workgroup nw=2", &launch_list_group<GP56, 128>}}},
};

}  // namespace

struct rbfa_batch {
    int nclasses = 0;
    // classes are kept sorted by size (slot u); user_slot maps the caller's
    // class number (its order at create) to its slot
    int cls[kMaxClasses];        // registry index of each slot's class
    int variant[kMaxClasses];    // chosen list-kernel variant per slot
    int user_slot[kMaxClasses];  // slot of the caller's class i
    int mode = 0;
    bool bound = false;
    int32_t nstencils = 0;
    int count[kMaxClasses] = {0, 0, 0, 0};
    int start[kMaxClasses] = {0, 0, 0, 0};
    int* d_order = nullptr;
    int* d_tickets = nullptr;
    const int32_t* ia = nullptr;
    const double *x = nullptr, *y = nullptr, *z = nullptr;
    const int32_t* ja = nullptr;
    double* w[3] = {nullptr, nullptr, nullptr};
    int32_t* info = nullptr;
    uint8_t* status = nullptr;
    int persistent_grid = 0;
    hipStream_t cap = nullptr;
    hipGraphExec_t exec[2] = {nullptr, nullptr};
};

namespace {

void drop_graphs(rbfa_batch* b) {
    for (auto& e : b->exec) {
        if (e) (void)hipGraphExecDestroy(e);
        e = nullptr;
    }
}

// Enqueue mode m's pipeline on stream st (used eagerly and under capture).
hipError_t enqueue(const rbfa_batch* b, int m, hipStream_t st) {
    Out3 out;
    for (int c = 0; c < 3; ++c) out.p[c] = b->w[c];
    if (m == 0) {
        // largest class first, so its long-running workgroups start early
        for (int u = b->nclasses - 1; u >= 0; --u) {
            const BatchClass& bc = kClasses[b->cls[u]];
            ListArgs a{b->count[u], b->d_order + b->start[u], b->ia, b->x, b->y, b->z, b->ja,
                       out, b->info, b->status, st};
            const hipError_t e = bc.v[b->variant[u]].fn(a);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    Queues q{};
    q.order = b->d_order;
    for (int r = 0; r < kMaxClasses; ++r) {
        q.start[r] = 0;
        q.count[r] = 0;
    }
    for (int u = 0; u < b->nclasses; ++u) {
        q.start[b->cls[u]] = b->start[u];
        q.count[b->cls[u]] = b->count[u];
    }
    q.tickets = b->d_tickets;
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, st, b->d_tickets, kMaxClasses);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_persistent, dim3(b->persistent_grid), dim3(kPersistentThreads), 0, st, q,
                       b->ia, b->x,
                       b->y, b->ja, out, b->info, b->status);
    return hipGetLastError();
}

int32_t capture(rbfa_batch* b, int m) {
    if (b->exec[m]) return RBFA_OK;
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return RBFA_ERR_HIP;
    const hipError_t e = enqueue(b, m, b->cap);
    const hipError_t e2 = hipStreamEndCapture(b->cap, &g);
    if (e != hipSuccess || e2 != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        return RBFA_ERR_HIP;
    }
    const hipError_t e3 = hipGraphInstantiate(&b->exec[m], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    return e3 == hipSuccess ? RBFA_OK : RBFA_ERR_HIP;
}

}  // namespace

extern "C" {

RBFA_API int32_t rbfa_batch_create(int32_t nclasses, const rbfa_plan_desc* classes,
                                   rbfa_batch** batch) {
    if (!batch || !classes || nclasses < 1 || nclasses > kMaxClasses) return RBFA_ERR_ARG;
    *batch = nullptr;
    int cls[kMaxClasses];
    for (int u = 0; u < nclasses; ++u) {
        const rbfa_plan_desc& d = classes[u];
        if (d.dim != 2 || d.op_set != RBFA_OPS_DIFF || d.precision != RBFA_PREC_FP64)
            return RBFA_ERR_UNSUPPORTED;
        cls[u] = -1;
        for (int r = 0; r < kMaxClasses; ++r)
            if (kClasses[r].n == d.n && kClasses[r].P == d.degree && kClasses[r].Q == d.phs)
                cls[u] = r;
        if (cls[u] < 0) return RBFA_ERR_UNSUPPORTED;
        for (int v = 0; v < u; ++v)
            if (classes[v].n == d.n) return RBFA_ERR_ARG;  // sizes must identify classes
    }
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return RBFA_ERR_NODEVICE;
    auto* b = new (std::nothrow) rbfa_batch;
    if (!b) return RBFA_ERR_NOMEM;
    // keep classes sorted by size so "largest first" is the last index
    std::vector<int> idx(nclasses);
    for (int u = 0; u < nclasses; ++u) idx[u] = u;
    std::sort(idx.begin(), idx.end(),
              [&](int a, int c) { return classes[a].n < classes[c].n; });
    b->nclasses = nclasses;
    for (int u = 0; u < nclasses; ++u) {
        b->cls[u] = cls[idx[u]];
        b->variant[u] = kClasses[b->cls[u]].def;
        b->user_slot[idx[u]] = u;
    }
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_persistent, kPersistentThreads,
                                                     0) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    b->persistent_grid = per_cu * prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&b->d_tickets, sizeof(int) * kMaxClasses) != hipSuccess) {
        if (b->cap) (void)hipStreamDestroy(b->cap);
        delete b;
        return RBFA_ERR_HIP;
    }
    *batch = b;
    return RBFA_OK;
}

RBFA_API int32_t rbfa_batch_destroy(rbfa_batch* b) {
    if (!b) return RBFA_OK;
    drop_graphs(b);
    if (b->d_order) (void)hipFree(b->d_order);
    if (b->d_tickets) (void)hipFree(b->d_tickets);
    if (b->cap) (void)hipStreamDestroy(b->cap);
    delete b;
    return RBFA_OK;
}

RBFA_API int32_t rbfa_batch_bind(rbfa_batch* b, int32_t nstencils, const int32_t* ia_host,
                                 const double* x, const double* y, const double* z,
                                 const int32_t* ia, const int32_t* ja, double* const* w,
                                 int32_t* info, uint8_t* status, void* stream) {
    if (!b) return RBFA_ERR_ARG;
    // unbound until every step below has succeeded: a failed bind leaves a
    // batch that launch and autotune refuse, never one that mixes the old
    // binding with parts of the new
    b->bound = false;
    if (nstencils < 0 || (nstencils > 0 && (!ia_host || !x || !y || !ia || !ja || !w)))
        return RBFA_ERR_ARG;
    for (int c = 0; c < 3; ++c)
        if (nstencils > 0 && !w[c]) return RBFA_ERR_ARG;
    // bucket the stencils by size (the row length names the class)
    std::vector<std::vector<int>> lists(b->nclasses);
    for (int32_t s = 0; s < nstencils; ++s) {
        const int len = ia_host[s + 1] - ia_host[s];
        int u = -1;
        for (int v = 0; v < b->nclasses; ++v)
            if (kClasses[b->cls[v]].n == len) u = v;
        if (u < 0) return RBFA_ERR_ARG;
        lists[u].push_back(s);
    }
    std::vector<int> order;
    order.reserve(static_cast<size_t>(nstencils));
    int start[kMaxClasses] = {}, count[kMaxClasses] = {};
    for (int u = 0; u < b->nclasses; ++u) {
        start[u] = static_cast<int>(order.size());
        count[u] = static_cast<int>(lists[u].size());
        order.insert(order.end(), lists[u].begin(), lists[u].end());
    }
    drop_graphs(b);
    if (b->d_order) (void)hipFree(b->d_order);
    b->d_order = nullptr;
    if (!order.empty()) {
        if (hipMalloc(&b->d_order, sizeof(int) * order.size()) != hipSuccess) return RBFA_ERR_NOMEM;
        if (hipMemcpyAsync(b->d_order, order.data(), sizeof(int) * order.size(),
                           hipMemcpyHostToDevice, static_cast<hipStream_t>(stream)) != hipSuccess ||
            hipStreamSynchronize(static_cast<hipStream_t>(stream)) != hipSuccess)
            return RBFA_ERR_HIP;
    }
    for (int u = 0; u < b->nclasses; ++u) {
        b->start[u] = start[u];
        b->count[u] = count[u];
    }
    b->nstencils = nstencils;
    b->ia = ia;
    b->x = x;
    b->y = y;
    b->z = z;
    b->ja = ja;
    for (int c = 0; c < 3; ++c) b->w[c] = w[c];
    b->info = info;
    b->status = status;
    b->bound = true;
    return RBFA_OK;
}

RBFA_API int32_t rbfa_batch_class_count(const rbfa_batch* b, int32_t cls) {
    if (!b || cls < 0 || cls >= b->nclasses) return -1;
    return b->count[b->user_slot[cls]];
}

RBFA_API int32_t rbfa_batch_set_mode(rbfa_batch* b, int32_t mode) {
    if (!b || (mode != 0 && mode != 1)) return RBFA_ERR_ARG;
    if (mode == 1 && !kPersistentEnabled) {
        rbf_amd_internal::set_last_error(
            "rbfa_batch_set_mode: the persistent launch (mode 1) is switched off in this build; "
            "mode 0 runs one launch per size class in the captured graph");
        return RBFA_ERR_UNSUPPORTED;
    }
    b->mode = mode;
    return RBFA_OK;
}

RBFA_API int32_t rbfa_batch_get_mode(const rbfa_batch* b) { return b ? b->mode : -1; }

RBFA_API int32_t rbfa_batch_launch(rbfa_batch* b, void* stream) {
    if (!b || !b->bound) return RBFA_ERR_ARG;
    if (b->nstencils == 0) return RBFA_OK;
    const int32_t rc = capture(b, b->mode);
    if (rc != RBFA_OK) return rc;
    return hipGraphLaunch(b->exec[b->mode], static_cast<hipStream_t>(stream)) == hipSuccess
               ? RBFA_OK
               : RBFA_ERR_HIP;
}

RBFA_API int32_t rbfa_batch_autotune(rbfa_batch* b, void* stream, int32_t reps, float* ms_out) {
    if (!b || !b->bound || reps < 1) return RBFA_ERR_ARG;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return RBFA_ERR_HIP;
    auto time_it = [&](auto&& fn, float& ms) -> int32_t {
        if (fn() != hipSuccess) return RBFA_ERR_HIP;  // warm-up
        if (hipEventRecord(e0, st) != hipSuccess) return RBFA_ERR_HIP;
        for (int r = 0; r < reps; ++r)
            if (fn() != hipSuccess) return RBFA_ERR_HIP;
        if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            return RBFA_ERR_HIP;
        ms /= reps;
        return RBFA_OK;
    };
    int32_t rc = RBFA_OK;
    Out3 out;
    for (int c = 0; c < 3; ++c) out.p[c] = b->w[c];
    // 1. each class's kernel variant, on that class's own stencils
    for (int u = 0; u < b->nclasses && rc == RBFA_OK; ++u) {
        const BatchClass& bc = kClasses[b->cls[u]];
        if (bc.nvariants < 2 || b->count[u] == 0) continue;
        float best = 0.f;
        for (int v = 0; v < bc.nvariants && rc == RBFA_OK; ++v) {
            ListArgs a{b->count[u], b->d_order + b->start[u], b->ia, b->x, b->y, b->z, b->ja,
                       out, b->info, b->status, st};
            float ms = 0.f;
            rc = time_it([&] { return bc.v[v].fn(a); }, ms);
            if (rc == RBFA_OK && (v == 0 || ms < best)) {
                best = ms;
                b->variant[u] = v;
            }
        }
    }
    // 2. the captured pipelines (the persistent one only while enabled;
    //    its time is reported as -1 otherwise)
    drop_graphs(b);
    float ms_mode[2] = {0.f, -1.f};
    for (int m = 0; m < (kPersistentEnabled ? 2 : 1) && rc == RBFA_OK; ++m) {
        rc = capture(b, m);
        if (rc == RBFA_OK) rc = time_it([&] { return hipGraphLaunch(b->exec[m], st); }, ms_mode[m]);
    }
    if (rc == RBFA_OK) {
        b->mode = kPersistentEnabled && ms_mode[1] < ms_mode[0] ? 1 : 0;
        if (ms_out) {
            ms_out[0] = ms_mode[0];
            ms_out[1] = ms_mode[1];
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

// Classes are numbered as the caller listed them at create; internally
// they are kept sorted by size.
static int batch_slot(const rbfa_batch* b, int32_t cls) {
    if (!b || cls < 0 || cls >= b->nclasses) return -1;
    return b->user_slot[cls];
}

RBFA_API const char* rbfa_batch_class_variant(const rbfa_batch* b, int32_t cls) {
    const int u = batch_slot(b, cls);
    return u < 0 ? nullptr : kClasses[b->cls[u]].v[b->variant[u]].name;
}

RBFA_API int32_t rbfa_batch_class_num_variants(const rbfa_batch* b, int32_t cls) {
    const int u = batch_slot(b, cls);
    return u < 0 ? -1 : kClasses[b->cls[u]].nvariants;
}

RBFA_API const char* rbfa_batch_class_variant_name(const rbfa_batch* b, int32_t cls, int32_t v) {
    const int u = batch_slot(b, cls);
    if (u < 0 || v < 0 || v >= kClasses[b->cls[u]].nvariants) return nullptr;
    return kClasses[b->cls[u]].v[v].name;
}

RBFA_API int32_t rbfa_batch_set_class_variant(rbfa_batch* b, int32_t cls, int32_t v) {
    const int u = batch_slot(b, cls);
    if (u < 0 || v < 0 || v >= kClasses[b->cls[u]].nvariants) {
        rbf_amd_internal::set_last_error("rbfa_batch_set_class_variant: no such class or variant "
                       "(rbfa_batch_class_num_variants lists them)");
        return RBFA_ERR_ARG;
    }
    if (b->variant[u] != v) {
        b->variant[u] = v;
        drop_graphs(b);  // the next launch captures the pipeline again
    }
    return RBFA_OK;
}

}  // extern "C"
