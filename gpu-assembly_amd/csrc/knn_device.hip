// knn_device.hip -- the node layer on the GPU: curve keys, stable key
// order, gathers, and the k-nearest-neighbour stencil search, each giving
// exactly what the host node layer (host_nodes.cpp) gives for the same
// points: the same keys (curve_keys.h), the same stable order, and the
// same stencils, found on the same bucket grid (knn_grid.h) with the same
// ring order and stopping rule, ordered by (squared distance, index).
// Replaces, on the device, NodeSet::stencils (rbf_nodeset.h:155-181) and
// hilbert_order / morton_order (rbf_reorder.h:198-229).
//
// Search: one wave per query point. The k best candidates so far stay
// sorted in registers (entry e in lane e % 64, two entries per lane for
// k > 64); each ring's cells are scanned row segment by row segment, 64
// candidates at a time, and every candidate better than the current k-th
// is inserted by a rank (ballot + popcount) and a one-lane shift
// (lane shuffles).

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <climits>
#include <cstdint>
#include <vector>

#include "assemble_common.h"
#include "curve_keys.h"
#include "knn_grid.h"
#include "rbf_assembly.h"

namespace {

using namespace rbf_amd_internal;

constexpr int kBboxBlocks = 512;

template <int D>
__global__ void __launch_bounds__(256)
k_bbox(int n, const double* __restrict__ X0, const double* __restrict__ X1,
       const double* __restrict__ X2, double* __restrict__ part) {
    __shared__ double red[2 * 3][256];
    const double* X[3] = {X0, X1, X2};
    double lo[3], hi[3];
    for (int d = 0; d < D; ++d) {
        lo[d] = X[d][0];
        hi[d] = X[d][0];
    }
    for (int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x); i < n;
         i += static_cast<int>(gridDim.x * blockDim.x))
        for (int d = 0; d < D; ++d) {
            lo[d] = fmin(lo[d], X[d][i]);
            hi[d] = fmax(hi[d], X[d][i]);
        }
    for (int d = 0; d < D; ++d) {
        red[2 * d][threadIdx.x] = lo[d];
        red[2 * d + 1][threadIdx.x] = hi[d];
    }
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w)
            for (int d = 0; d < D; ++d) {
                red[2 * d][threadIdx.x] = fmin(red[2 * d][threadIdx.x], red[2 * d][threadIdx.x + w]);
                red[2 * d + 1][threadIdx.x] =
                    fmax(red[2 * d + 1][threadIdx.x], red[2 * d + 1][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int d = 0; d < D; ++d) {
            part[blockIdx.x * 6 + 2 * d] = red[2 * d][0];
            part[blockIdx.x * 6 + 2 * d + 1] = red[2 * d + 1][0];
        }
}

// Bounding box of device points (lo[d], hi[d] on the host).
template <int D>
hipError_t point_bounds(int n, const double* const* X, double* lo, double* hi, hipStream_t st) {
    double* part = nullptr;
    hipError_t e = hipMalloc(&part, sizeof(double) * 6 * kBboxBlocks);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_bbox<D>, dim3(kBboxBlocks), dim3(256), 0, st, n, X[0], X[1],
                       D == 3 ? X[2] : X[0], part);
    std::vector<double> h(6 * kBboxBlocks);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(),
                                            hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(part);
    if (e != hipSuccess) return e;
    for (int d = 0; d < D; ++d) {
        lo[d] = h[2 * d];
        hi[d] = h[2 * d + 1];
        for (int b = 1; b < kBboxBlocks; ++b) {
            lo[d] = std::min(lo[d], h[b * 6 + 2 * d]);
            hi[d] = std::max(hi[d], h[b * 6 + 2 * d + 1]);
        }
    }
    return hipSuccess;
}

struct GridDev {
    double lo[3];
    double cell;
    int dims[3];
    const int* start;
    const int* order;
    const double* pts;
};

__device__ __forceinline__ int grid_coord(const GridDev& g, int d, double v) {
    const int c = static_cast<int>((v - g.lo[d]) / g.cell);
    return min(max(c, 0), g.dims[d] - 1);
}

template <int D>
__global__ void __launch_bounds__(256)
k_cell_count(int n, const double* __restrict__ X0, const double* __restrict__ X1,
             const double* __restrict__ X2, GridDev g, int* __restrict__ cell_of,
             int* __restrict__ count) {
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const double p[3] = {X0[i], X1[i], D == 3 ? X2[i] : 0.0};
    long long ci = 0;
    for (int d = D - 1; d >= 0; --d) ci = ci * g.dims[d] + grid_coord(g, d, p[d]);
    cell_of[i] = static_cast<int>(ci);
    atomicAdd(&count[ci], 1);
}

template <int D>
__global__ void __launch_bounds__(256)
k_cell_scatter(int n, const double* __restrict__ X0, const double* __restrict__ X1,
               const double* __restrict__ X2, const int* __restrict__ cell_of,
               const int* __restrict__ start, int* __restrict__ fill, int* __restrict__ order,
               double* __restrict__ pts) {
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const int c = cell_of[i];
    const int slot = start[c] + atomicAdd(&fill[c], 1);
    order[slot] = i;
    pts[static_cast<size_t>(slot) * D] = X0[i];
    pts[static_cast<size_t>(slot) * D + 1] = X1[i];
    if (D == 3) pts[static_cast<size_t>(slot) * D + 2] = X2[i];
}

__device__ __forceinline__ bool less_pair(double a, int ia, double b, int ib) {
    return a < b || (a == b && ia < ib);
}

// The k best (d2, index) pairs of one query: entry e in lane e & 63 of set
// e >> 6 (a0/i0: entries 0..63, a1/i1: entries 64..127).
struct TopK {
    double a0, a1;
    int i0, i1;
    int count;
    double wd;  // k-th entry once count == k
    int wi;

    __device__ __forceinline__ void init() {
        a0 = a1 = __builtin_inf();
        i0 = i1 = INT_MAX;
        count = 0;
        wd = 0.0;
        wi = 0;
    }
    __device__ __forceinline__ void insert(double cd, int ci, int k, int lane) {
        const bool v0 = lane < count, v1 = lane + 64 < count;
        const int pos = __popcll(__ballot(v0 && less_pair(a0, i0, cd, ci))) +
                        __popcll(__ballot(v1 && less_pair(a1, i1, cd, ci)));
        const double ua0 = __shfl_up(a0, 1, 64);
        const int ui0 = __shfl_up(i0, 1, 64);
        double ua1 = __shfl_up(a1, 1, 64);
        int ui1 = __shfl_up(i1, 1, 64);
        const double l63a = __shfl(a0, 63, 64);
        const int l63i = __shfl(i0, 63, 64);
        if (lane == 0) {
            ua1 = l63a;
            ui1 = l63i;
        }
        const int e0 = lane, e1 = lane + 64;
        if (e0 > pos) {
            a0 = ua0;
            i0 = ui0;
        } else if (e0 == pos) {
            a0 = cd;
            i0 = ci;
        }
        if (e1 > pos) {
            a1 = ua1;
            i1 = ui1;
        } else if (e1 == pos) {
            a1 = cd;
            i1 = ci;
        }
        count = count < k ? count + 1 : k;
        if (count == k) {
            const int last = k - 1;
            wd = __shfl(last < 64 ? a0 : a1, last & 63, 64);
            wi = __shfl(last < 64 ? i0 : i1, last & 63, 64);
        }
    }

    // Compare-exchange with lane ^ stride: keep the smaller pair (take_min)
    // or the larger one.
    __device__ __forceinline__ static void cmpx(double& d, int& i, int stride, bool take_min) {
        const double pd = __shfl_xor(d, stride, 64);
        const int pi = __shfl_xor(i, stride, 64);
        if (take_min == less_pair(pd, pi, d, i)) {
            d = pd;
            i = pi;
        }
    }

    // Merge a chunk of candidates (one per lane; lanes not kept hold
    // (inf, INT_MAX)) with `nkeep` kept into the list at once. The chunk
    // is sorted across the wave (bitonic) and reversed against the first
    // 64 entries: the pairwise minima are the 64 smallest of both, the
    // maxima the 64 largest, each a bitonic sequence sorted by six merge
    // stages. The minima are the new entries 0..63 (none exceeds an entry
    // 64.. of the list); for k > 64 the maxima are merged the same way
    // with entries 64..127, keeping the smaller 64. The first min(k, count
    // + nkeep) entries are then the k smallest pairs of list and chunk --
    // what inserting the candidates one by one leaves, since (d, index) is
    // a strict total order and a candidate above the k-th is never kept.
    __device__ __forceinline__ static void sort_merge(double& d, int& i, int lane) {
        for (int stride = 32; stride > 0; stride >>= 1) cmpx(d, i, stride, (lane & stride) == 0);
    }
    __device__ __forceinline__ void merge_chunk(double cd, int ci, int nkeep, int k, int lane) {
        for (int size = 2; size <= 64; size <<= 1)
            for (int stride = size >> 1; stride > 0; stride >>= 1)
                cmpx(cd, ci, stride, ((lane & stride) == 0) == ((lane & size) == 0));
        if (lane >= count) {
            a0 = __builtin_inf();
            i0 = INT_MAX;
        }
        double rd = __shfl(cd, 63 - lane, 64);
        int ri = __shfl(ci, 63 - lane, 64);
        const bool rl = less_pair(rd, ri, a0, i0);
        double hd = rl ? a0 : rd;  // the larger of the pair
        int hi = rl ? i0 : ri;
        if (rl) {
            a0 = rd;
            i0 = ri;
        }
        sort_merge(a0, i0, lane);
        if (k > 64) {
            sort_merge(hd, hi, lane);
            if (lane + 64 >= count) {
                a1 = __builtin_inf();
                i1 = INT_MAX;
            }
            rd = __shfl(a1, 63 - lane, 64);
            ri = __shfl(i1, 63 - lane, 64);
            if (less_pair(rd, ri, hd, hi)) {
                hd = rd;
                hi = ri;
            }
            sort_merge(hd, hi, lane);
            a1 = hd;
            i1 = hi;
        }
        count = count + nkeep < k ? count + nkeep : k;
        if (lane >= count) {
            a0 = __builtin_inf();
            i0 = INT_MAX;
        }
        if (lane + 64 >= count) {
            a1 = __builtin_inf();
            i1 = INT_MAX;
        }
        if (count == k) {
            const int last = k - 1;
            wd = __shfl(last < 64 ? a0 : a1, last & 63, 64);
            wi = __shfl(last < 64 ? i0 : i1, last & 63, 64);
        }
    }
};

// Chunks with at least this many candidates below the current k-th are
// merged at once rather than inserted one by one.
constexpr int kMergeMin = 8;

// Candidates in slots [sb, se), 64 at a time.
template <int D>
__device__ __forceinline__ void scan_slots(int sb, int se, const GridDev& g, const double (&q)[3],
                                           TopK& t, int k, int lane) {
#pragma clang fp contract(off)
    for (int base = sb; base < se; base += 64) {
        const int slot = base + lane;
        const bool valid = slot < se;
        double acc = 0.0;
        int idx = INT_MAX;
        if (valid) {
            const double* p = g.pts + static_cast<size_t>(slot) * D;
            for (int d = 0; d < D; ++d) {
                const double diff = q[d] - p[d];
                acc += diff * diff;
            }
            idx = g.order[slot];
        }
        const bool keep = valid && (t.count < k || less_pair(acc, idx, t.wd, t.wi));
        unsigned long long m = __ballot(keep);
        if (__popcll(m) >= kMergeMin) {
            t.merge_chunk(keep ? acc : __builtin_inf(), keep ? idx : INT_MAX, __popcll(m), k, lane);
            continue;
        }
        while (m) {
            const int j = __ffsll(m) - 1;
            m &= m - 1;
            const double cd = __shfl(acc, j, 64);
            const int ci = __shfl(idx, j, 64);
            if (t.count < k || less_pair(cd, ci, t.wd, t.wi)) t.insert(cd, ci, k, lane);
        }
    }
}

template <int D>
__global__ void __launch_bounds__(256)
k_knn(int n, int k, const double* __restrict__ X0, const double* __restrict__ X1,
      const double* __restrict__ X2, GridDev g, int* __restrict__ ja) {
#pragma clang fp contract(off)
    const int lane = static_cast<int>(threadIdx.x & 63);
    const int nb = (n + 3) / 4;
    const int b = xcd_block(nb);
    const int s = b * 4 + static_cast<int>(threadIdx.x >> 6);
    if (b >= nb || s >= n) return;
    const double q[3] = {X0[s], X1[s], D == 3 ? X2[s] : 0.0};
    int qc[3] = {0, 0, 0};
    for (int d = 0; d < D; ++d) qc[d] = grid_coord(g, d, q[d]);
    TopK t;
    t.init();
    for (int r = 0;; ++r) {
        int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        bool covers_all = true;
        for (int d = 0; d < D; ++d) {
            lo[d] = max(qc[d] - r, 0);
            hi[d] = min(qc[d] + r, g.dims[d] - 1);
            if (qc[d] - r > 0 || qc[d] + r < g.dims[d] - 1) covers_all = false;
        }
        // the cells at Chebyshev distance r from qc, as row segments along x
        const int z0 = D == 3 ? lo[2] : 0, z1 = D == 3 ? hi[2] : 0;
        for (int zz = z0; zz <= z1; ++zz) {
            const bool zshell = D == 3 && (zz == qc[2] - r || zz == qc[2] + r);
            for (int yy = lo[1]; yy <= hi[1]; ++yy) {
                const bool full = zshell || yy == qc[1] - r || yy == qc[1] + r;
                const int row = (D == 3 ? zz * g.dims[1] + yy : yy) * g.dims[0];
                if (full) {
                    scan_slots<D>(g.start[row + lo[0]], g.start[row + hi[0] + 1], g, q, t, k, lane);
                } else {
                    if (qc[0] - r >= 0)
                        scan_slots<D>(g.start[row + qc[0] - r], g.start[row + qc[0] - r + 1], g, q,
                                      t, k, lane);
                    if (r > 0 && qc[0] + r < g.dims[0])
                        scan_slots<D>(g.start[row + qc[0] + r], g.start[row + qc[0] + r + 1], g, q,
                                      t, k, lane);
                }
            }
        }
        if (covers_all) break;
        if (t.count == k) {
            // distance from q to the outside of the searched block
            double bd = __builtin_inf();
            for (int d = 0; d < D; ++d) {
                if (qc[d] - r > 0) bd = fmin(bd, q[d] - (g.lo[d] + (qc[d] - r) * g.cell));
                if (qc[d] + r < g.dims[d] - 1)
                    bd = fmin(bd, (g.lo[d] + (qc[d] + r + 1) * g.cell) - q[d]);
            }
            if (bd > 0 && bd * bd * (1.0 - 1e-9) > t.wd) break;
        }
    }
    int* out = ja + static_cast<size_t>(s) * k;
    if (lane < k) out[lane] = t.i0;
    if (lane + 64 < k) out[lane + 64] = t.i1;
}

template <int D>
int32_t knn_device(int32_t n, const double* const* X, int32_t k, int32_t* ja, hipStream_t st) {
    double lo[3], hi[3];
    if (point_bounds<D>(n, X, lo, hi, st) != hipSuccess) return RBFA_ERR_HIP;
    GridDev g{};
    for (int d = 0; d < D; ++d) g.lo[d] = lo[d];
    g.cell = grid_params<D>(lo, hi, n, k, g.dims);
    long long total = 1;
    for (int d = 0; d < D; ++d) total *= g.dims[d];
    int *count = nullptr, *start = nullptr, *fill = nullptr, *cell_of = nullptr, *order = nullptr;
    double* pts = nullptr;
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
    int32_t rc = RBFA_OK;
    auto ok = [&](hipError_t e) {
        if (e != hipSuccess && rc == RBFA_OK)
            rc = e == hipErrorOutOfMemory ? RBFA_ERR_NOMEM : RBFA_ERR_HIP;
        return rc == RBFA_OK;
    };
    const size_t cells = static_cast<size_t>(total) + 1;
    if (ok(hipMalloc(&count, sizeof(int) * cells)) && ok(hipMalloc(&start, sizeof(int) * cells)) &&
        ok(hipMalloc(&fill, sizeof(int) * cells)) && ok(hipMalloc(&cell_of, sizeof(int) * n)) &&
        ok(hipMalloc(&order, sizeof(int) * n)) && ok(hipMalloc(&pts, sizeof(double) * n * D)) &&
        ok(hipMemsetAsync(count, 0, sizeof(int) * cells, st)) &&
        ok(hipMemsetAsync(fill, 0, sizeof(int) * cells, st))) {
        const unsigned grid = static_cast<unsigned>((n + 255) / 256);
        hipLaunchKernelGGL(k_cell_count<D>, dim3(grid), dim3(256), 0, st, n, X[0], X[1],
                           D == 3 ? X[2] : X[0], g, cell_of, count);
        if (ok(hipGetLastError()) &&
            ok(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, count, start,
                                                static_cast<int>(cells), st)) &&
            ok(hipMalloc(&tmp, tmp_bytes)) &&
            ok(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, count, start,
                                                static_cast<int>(cells), st))) {
            hipLaunchKernelGGL(k_cell_scatter<D>, dim3(grid), dim3(256), 0, st, n, X[0], X[1],
                               D == 3 ? X[2] : X[0], cell_of, start, fill, order, pts);
            if (ok(hipGetLastError())) {
                g.start = start;
                g.order = order;
                g.pts = pts;
                const int nb = (n + 3) / 4;
                hipLaunchKernelGGL(k_knn<D>, dim3(xcd_grid(nb)), dim3(256), 0, st, n, k, X[0], X[1],
                                   D == 3 ? X[2] : X[0], g, ja);
                ok(hipGetLastError());
            }
        }
    }
    // the temporaries are freed once the stream has used them
    ok(hipStreamSynchronize(st));
    void* bufs[] = {count, start, fill, cell_of, order, pts, tmp};
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    return rc;
}

// ---- keys, order, gather ------------------------------------------------------

__global__ void __launch_bounds__(256)
k_keys2(int n, const double* __restrict__ x, const double* __restrict__ y, double x0, double y0,
        double x1, double y1, int ndiv, int hilbert, int64_t* __restrict__ keys) {
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    keys[i] = hilbert ? hilbert_key2(x[i], y[i], x0, y0, x1, y1, ndiv)
                      : morton_key2(x[i], y[i], x0, y0, x1, y1, ndiv);
}

__global__ void __launch_bounds__(256)
k_keys3(int n, const double* __restrict__ x, const double* __restrict__ y,
        const double* __restrict__ z, double b0, double b1, double b2, double b3, double b4,
        double b5, int ndiv, int64_t* __restrict__ keys) {
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const double bb[6] = {b0, b1, b2, b3, b4, b5};
    keys[i] = morton_key3(x[i], y[i], z[i], bb, ndiv);
}

__global__ void __launch_bounds__(256) k_iota(int n, int* __restrict__ v) {
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) v[i] = i;
}

__global__ void __launch_bounds__(256)
k_gather(int n, const int* __restrict__ perm, const double* __restrict__ src,
         double* __restrict__ dst) {
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) dst[i] = src[perm[i]];
}

}  // namespace

extern "C" {

RBFA_API int32_t rbfa_knn_stencils_device(int32_t n, int32_t dim, const double* x,
                                          const double* y, const double* z, int32_t k,
                                          int32_t* ja, void* stream) {
    if (n < 1 || (dim != 2 && dim != 3) || k < 1 || k > n || k > 128) return RBFA_ERR_ARG;
    if (!x || !y || !ja || (dim == 3 && !z)) return RBFA_ERR_ARG;
    if (static_cast<int64_t>(n) * k > INT32_MAX) return RBFA_ERR_ARG;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (dim == 2) {
        const double* X[2] = {x, y};
        return knn_device<2>(n, X, k, ja, st);
    }
    const double* X[3] = {x, y, z};
    return knn_device<3>(n, X, k, ja, st);
}

RBFA_API int32_t rbfa_curve_keys_device(int32_t n, int32_t dim, int32_t curve, const double* x,
                                        const double* y, const double* z, const double* bbox,
                                        int32_t ndiv, int64_t* keys, void* stream) {
    if (n < 1 || !x || !y || !keys || (dim != 2 && dim != 3)) return RBFA_ERR_ARG;
    if (dim == 2 && (curve != 0 && curve != 1)) return RBFA_ERR_ARG;
    if (dim == 3 && (curve != 0 || !z)) return RBFA_ERR_ARG;
    if (ndiv < 1 || ndiv > (dim == 2 ? 31 : 21)) return RBFA_ERR_ARG;
    hipStream_t st = static_cast<hipStream_t>(stream);
    double bb[6];
    if (bbox) {
        for (int i = 0; i < 2 * dim; ++i) bb[i] = bbox[i];
    } else {
        double lo[3], hi[3];
        const double* X[3] = {x, y, z};
        const hipError_t e = dim == 2 ? point_bounds<2>(n, X, lo, hi, st) : point_bounds<3>(n, X, lo, hi, st);
        if (e != hipSuccess) return RBFA_ERR_HIP;
        for (int d = 0; d < dim; ++d) {
            bb[d] = lo[d];
            bb[dim + d] = hi[d];
        }
    }
    const unsigned grid = static_cast<unsigned>((n + 255) / 256);
    if (dim == 2)
        hipLaunchKernelGGL(k_keys2, dim3(grid), dim3(256), 0, st, n, x, y, bb[0], bb[1], bb[2],
                           bb[3], ndiv, curve, keys);
    else
        hipLaunchKernelGGL(k_keys3, dim3(grid), dim3(256), 0, st, n, x, y, z, bb[0], bb[1], bb[2],
                           bb[3], bb[4], bb[5], ndiv, keys);
    return hipGetLastError() == hipSuccess ? RBFA_OK : RBFA_ERR_HIP;
}

RBFA_API int32_t rbfa_order_by_keys_device(int32_t n, const int64_t* keys, int32_t key_bits,
                                           int32_t* perm, void* stream) {
    if (n < 1 || !keys || !perm || key_bits < 1 || key_bits > 63) return RBFA_ERR_ARG;
    hipStream_t st = static_cast<hipStream_t>(stream);
    int* vals = nullptr;
    uint64_t* ksorted = nullptr;
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
    int32_t rc = RBFA_OK;
    auto ok = [&](hipError_t e) {
        if (e != hipSuccess && rc == RBFA_OK)
            rc = e == hipErrorOutOfMemory ? RBFA_ERR_NOMEM : RBFA_ERR_HIP;
        return rc == RBFA_OK;
    };
    const uint64_t* kin = reinterpret_cast<const uint64_t*>(keys);
    if (ok(hipMalloc(&vals, sizeof(int) * n)) && ok(hipMalloc(&ksorted, sizeof(uint64_t) * n))) {
        hipLaunchKernelGGL(k_iota, dim3((n + 255) / 256), dim3(256), 0, st, n, vals);
        if (ok(hipGetLastError()) &&
            ok(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, kin, ksorted, vals, perm, n,
                                                  0, key_bits, st)) &&
            ok(hipMalloc(&tmp, tmp_bytes)))
            ok(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, kin, ksorted, vals, perm, n, 0,
                                                  key_bits, st));
    }
    ok(hipStreamSynchronize(st));
    void* bufs[] = {vals, ksorted, tmp};
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    return rc;
}

RBFA_API int32_t rbfa_gather_f64_device(int32_t n, const int32_t* perm, const double* src,
                                        double* dst, void* stream) {
    if (n < 0 || (n > 0 && (!perm || !src || !dst)) || src == dst) return RBFA_ERR_ARG;
    if (n == 0) return RBFA_OK;
    hipLaunchKernelGGL(k_gather, dim3((n + 255) / 256), dim3(256), 0,
                       static_cast<hipStream_t>(stream), n, perm, src, dst);
    return hipGetLastError() == hipSuccess ? RBFA_OK : RBFA_ERR_HIP;
}

}  // extern "C"
