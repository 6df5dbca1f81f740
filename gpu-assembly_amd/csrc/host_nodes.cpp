// host_nodes.cpp -- host node layer of librbf_amd: curve keys, stable
// renumbering, k-nearest-neighbour stencils and seeded synthetic node sets.
//
// This is the producer of the CSR column indices ja that the device
// assembly consumes. The contract it keeps with the reference
// (ivan-pi/gpu-assembly) is bit-exact structure:
//
//   - curve keys follow hidx / zidx of rbf_ordering.F90:82-226 operation
//     for operation: the midpoint x0 + 0.5*(x1 - x0), the strict test
//     px > cx, the Hilbert state tables, int64 keys of 2 bits per level;
//   - the order is a stable sort on the keys (rbf_reorder.h:183-189);
//   - stencils are the k nearest by the squared distance nanoflann's
//     L2_Simple_Adaptor forms (query minus point, summed over the axes,
//     nanoflann.hpp:690-699), sorted ascending, so the node itself comes
//     first (rbf_nodeset.h:155-181). Equal distances are ordered by the
//     lower index (nanoflann keeps the first one its traversal finds,
//     nanoflann.hpp:253-283; on clouds without exact ties the two agree).
//
// Built with -ffp-contract=off: distance sums and midpoints must round the
// same way as the reference's, or near-ties could order differently.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>

#include <omp.h>

#include "curve_keys.h"
#include "knn_grid.h"
#include "rbf_assembly.h"

namespace {

// ---------------------------------------------------------------------------
// Curve keys
// ---------------------------------------------------------------------------

inline int64_t morton_key2(double px, double py, const rbfa_bbox2& bb, int ndiv) {
    return rbf_amd_internal::morton_key2(px, py, bb.xmin, bb.ymin, bb.xmax, bb.ymax, ndiv);
}
inline int64_t hilbert_key2(double px, double py, const rbfa_bbox2& bb, int ndiv) {
    return rbf_amd_internal::hilbert_key2(px, py, bb.xmin, bb.ymin, bb.xmax, bb.ymax, ndiv);
}
using rbf_amd_internal::morton_key3;

rbfa_bbox2 bbox_of(int32_t n, const double* x, const double* y) {
    rbfa_bbox2 b{x[0], y[0], x[0], y[0]};
    for (int32_t i = 1; i < n; ++i) {
        b.xmin = std::min(b.xmin, x[i]);
        b.xmax = std::max(b.xmax, x[i]);
        b.ymin = std::min(b.ymin, y[i]);
        b.ymax = std::max(b.ymax, y[i]);
    }
    return b;
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort of indices by int64 key (keys >= 0).
// ---------------------------------------------------------------------------

void stable_order(int32_t n, const int64_t* keys, int32_t* perm) {
    uint64_t kmax = 0;
    bool negative = false;
    for (int32_t i = 0; i < n; ++i) {
        if (keys[i] < 0) negative = true;
        kmax = std::max<uint64_t>(kmax, static_cast<uint64_t>(keys[i]));
    }
    std::iota(perm, perm + n, 0);
    if (negative) {  // general fallback: comparison sort, still stable
        std::stable_sort(perm, perm + n, [keys](int32_t a, int32_t b) { return keys[a] < keys[b]; });
        return;
    }
    int bits = 0;
    while (bits < 64 && (kmax >> bits) != 0) bits += 8;
    std::vector<int32_t> tmp(n);
    int32_t* src = perm;
    int32_t* dst = tmp.data();
    for (int shift = 0; shift < bits; shift += 8) {
        size_t count[257] = {0};
        for (int32_t i = 0; i < n; ++i) ++count[((keys[src[i]] >> shift) & 0xff) + 1];
        for (int b = 0; b < 256; ++b) count[b + 1] += count[b];
        for (int32_t i = 0; i < n; ++i) dst[count[(keys[src[i]] >> shift) & 0xff]++] = src[i];
        std::swap(src, dst);
    }
    if (src != perm) std::memcpy(perm, src, sizeof(int32_t) * static_cast<size_t>(n));
}

// ---------------------------------------------------------------------------
// k nearest neighbours on a uniform bucket grid
// ---------------------------------------------------------------------------

// Sorted candidate list of capacity k, ordered by (d2, index).
struct Best {
    int k = 0, count = 0;
    double* d2;
    int32_t* idx;
    inline bool worse(double a, int32_t ia, double b, int32_t ib) const {  // (a,ia) > (b,ib)
        return a > b || (a == b && ia > ib);
    }
    inline void add(double d, int32_t i) {
        if (count == k && !worse(d2[k - 1], idx[k - 1], d, i)) return;
        int pos = count < k ? count : k - 1;
        while (pos > 0 && worse(d2[pos - 1], idx[pos - 1], d, i)) {
            d2[pos] = d2[pos - 1];
            idx[pos] = idx[pos - 1];
            --pos;
        }
        d2[pos] = d;
        idx[pos] = i;
        if (count < k) ++count;
    }
};

template <int D>
struct Grid {
    double lo[D], cell;
    int dims[D];
    std::vector<int32_t> start;  // cell -> first slot
    std::vector<int32_t> order;  // slot -> point index
    std::vector<double> pts;     // slot -> coordinates (D per slot)

    int coord(int d, double v) const {
        int c = static_cast<int>((v - lo[d]) / cell);
        return std::min(std::max(c, 0), dims[d] - 1);
    }
};

template <int D>
Grid<D> build_grid(int32_t n, const double* const* X, int k) {
    Grid<D> g;
    double hi[D];
    for (int d = 0; d < D; ++d) {
        g.lo[d] = hi[d] = X[d][0];
        for (int32_t i = 1; i < n; ++i) {
            g.lo[d] = std::min(g.lo[d], X[d][i]);
            hi[d] = std::max(hi[d], X[d][i]);
        }
    }
    g.cell = rbf_amd_internal::grid_params<D>(g.lo, hi, n, k, g.dims);
    long long total = 1;
    for (int d = 0; d < D; ++d) total *= g.dims[d];
    std::vector<int32_t> cell_of(n);
    g.start.assign(static_cast<size_t>(total) + 1, 0);
    for (int32_t i = 0; i < n; ++i) {
        long long ci = 0;
        for (int d = D - 1; d >= 0; --d) ci = ci * g.dims[d] + g.coord(d, X[d][i]);
        cell_of[i] = static_cast<int32_t>(ci);
        ++g.start[ci + 1];
    }
    for (long long c2 = 0; c2 < total; ++c2) g.start[c2 + 1] += g.start[c2];
    g.order.resize(n);
    g.pts.resize(static_cast<size_t>(n) * D);
    std::vector<int32_t> fill(g.start.begin(), g.start.end() - 1);
    for (int32_t i = 0; i < n; ++i) {  // ascending i within each cell
        const int32_t slot = fill[cell_of[i]]++;
        g.order[slot] = i;
        for (int d = 0; d < D; ++d) g.pts[static_cast<size_t>(slot) * D + d] = X[d][i];
    }
    return g;
}

template <int D>
void knn_all(int32_t n, const double* const* X, int k, int32_t* ja) {
    const Grid<D> g = build_grid<D>(n, X, k);
#pragma omp parallel
    {
        std::vector<double> d2(k);
        std::vector<int32_t> id(k);
#pragma omp for schedule(dynamic, 256)
        for (int32_t s = 0; s < n; ++s) {
            Best best;
            best.k = k;
            best.d2 = d2.data();
            best.idx = id.data();
            double q[D];
            int qc[D];
            for (int d = 0; d < D; ++d) {
                q[d] = X[d][s];
                qc[d] = g.coord(d, q[d]);
            }
            for (int r = 0;; ++r) {
                // visit the cells at Chebyshev distance r from qc
                int lo[D], hi[D];
                bool covers_all = true;
                for (int d = 0; d < D; ++d) {
                    lo[d] = std::max(qc[d] - r, 0);
                    hi[d] = std::min(qc[d] + r, g.dims[d] - 1);
                    if (qc[d] - r > 0 || qc[d] + r < g.dims[d] - 1) covers_all = false;
                }
                int c[D];
                for (int d = 0; d < D; ++d) c[d] = lo[d];
                while (true) {
                    bool on_shell = false;
                    for (int d = 0; d < D; ++d)
                        if (c[d] == qc[d] - r || c[d] == qc[d] + r) on_shell = true;
                    if (on_shell) {
                        long long ci = 0;
                        for (int d = D - 1; d >= 0; --d) ci = ci * g.dims[d] + c[d];
                        for (int32_t slot = g.start[ci]; slot < g.start[ci + 1]; ++slot) {
                            const double* p = &g.pts[static_cast<size_t>(slot) * D];
                            double acc = 0.0;
                            for (int d = 0; d < D; ++d) {
                                const double diff = q[d] - p[d];
                                acc += diff * diff;
                            }
                            best.add(acc, g.order[slot]);
                        }
                    }
                    int d = 0;
                    while (d < D && c[d] == hi[d]) {
                        c[d] = lo[d];
                        ++d;
                    }
                    if (d == D) break;
                    ++c[d];
                }
                if (covers_all) break;
                if (best.count == k) {
                    // distance from q to the outside of the searched block
                    double b = std::numeric_limits<double>::infinity();
                    for (int d = 0; d < D; ++d) {
                        if (qc[d] - r > 0) b = std::min(b, q[d] - (g.lo[d] + (qc[d] - r) * g.cell));
                        if (qc[d] + r < g.dims[d] - 1)
                            b = std::min(b, (g.lo[d] + (qc[d] + r + 1) * g.cell) - q[d]);
                    }
                    if (b > 0 && b * b * (1.0 - 1e-9) > best.d2[k - 1]) break;
                }
            }
            std::memcpy(&ja[static_cast<size_t>(s) * k], best.idx, sizeof(int32_t) * k);
        }
    }
}

// xorshift64 (Marsaglia 2003, shift triple 13/7/17), 53-bit doubles in [0,1)
struct XorShift64 {
    uint64_t s;
    explicit XorShift64(uint64_t seed) : s(seed ? seed : 0x9E3779B97F4A7C15ull) {}
    double next() {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        return static_cast<double>(s >> 11) * 0x1.0p-53;
    }
};

bool is_permutation(int32_t n, const int32_t* perm) {
    std::vector<char> seen(static_cast<size_t>(n), 0);
    for (int32_t i = 0; i < n; ++i) {
        const int32_t v = perm[i];
        if (v < 0 || v >= n || seen[v]) return false;
        seen[v] = 1;
    }
    return true;
}

template <typename T>
int32_t permute_impl(int32_t n, const int32_t* perm, T* a) {
    if (n < 0) return RBFA_ERR_ARG;
    if (n == 0) return RBFA_OK;
    if (!perm || !a || !is_permutation(n, perm)) return RBFA_ERR_ARG;
    std::vector<T> tmp(a, a + n);
    for (int32_t i = 0; i < n; ++i) a[i] = tmp[perm[i]];
    return RBFA_OK;
}


// Graded unit-disk node set: concentric rings with spacing h(r) = hc (1 -
// 0.75 r) both along and across the rings (centre spacing hc, rim hc/4),
// a random phase per ring and a jitter of +-0.15 h in radius and angle, so
// neighbouring nodes keep a minimum distance (a Poisson-disk-like set).
// Returns the number of interior nodes a centre spacing hc gives.
static int64_t disk_rings(double hc, uint64_t seed, double* x, double* y, int64_t cap) {
    XorShift64 rng(seed);
    int64_t cnt = 0;
    auto emit = [&](double px, double py) {
        if (x && cnt < cap) {
            x[cnt] = px;
            y[cnt] = py;
        }
        ++cnt;
    };
    emit(0.0, 0.0);
    const double two_pi = 6.283185307179586;
    double r = 0.0;
    while (true) {
        r += hc * (1.0 - 0.75 * r);
        const double h = hc * (1.0 - 0.75 * r);
        if (r > 1.0 - 0.5 * h) break;
        const int64_t m = std::max<int64_t>(6, static_cast<int64_t>(std::llround(two_pi * r / h)));
        const double phase = rng.next();
        for (int64_t j = 0; j < m; ++j) {
            const double dr = 0.15 * h * (2.0 * rng.next() - 1.0);
            const double dt = 0.15 * (2.0 * rng.next() - 1.0);
            const double th = two_pi * (static_cast<double>(j) + phase + dt) / static_cast<double>(m);
            emit((r + dr) * std::cos(th), (r + dr) * std::sin(th));
        }
    }
    return cnt;
}

// Radially graded disk by dart throwing. Target spacing h(r) = hc (1 - a r),
// a = 3/4 (hc at the centre, hc/4 at the rim); candidates are drawn from the
// density 1 / h(r)^2 and a candidate p is kept when every node q already
// kept lies at least beta (h(p) + h(q)) / 2 away. The boundary nodes go in
// first, so interior nodes keep the same distance from them.
class GradedDarts {
  public:
    static constexpr double a = 0.75;
    // hard-core distance in units of the local spacing: a packing fraction
    // pi beta^2 / 4 = 0.40 at one node per h^2, well below the 0.547 at
    // which random sequential adsorption jams
    static constexpr double beta = 0.7136;

    GradedDarts(int64_t n_interior, uint64_t seed) : rng_(seed) {
        // nodes per unit area 1 / h(r)^2: n = 2 pi G(1) / hc^2
        hc_ = std::sqrt(6.283185307179586 * mass(1.0) / static_cast<double>(n_interior));
        cell_ = beta * spacing(1.0);
        side_ = static_cast<int64_t>(std::ceil(2.0 / cell_)) + 1;
        head_.assign(static_cast<size_t>(side_ * side_), -1);
    }

    double spacing(double r) const { return hc_ * (1.0 - a * r); }

    // Radius of a candidate: inverts the radial mass G(r) / G(1) by Newton
    // steps from a bisection bracket.
    double draw_radius() {
        const double target = rng_.next() * mass(1.0);
        double lo = 0.0, hi = 1.0, r = 0.5;
        for (int it = 0; it < 60; ++it) {
            const double f = mass(r) - target;
            if (f > 0) hi = r; else lo = r;
            const double step = f / (r / ((1.0 - a * r) * (1.0 - a * r)));
            double nr = r - step;
            if (!(nr > lo && nr < hi)) nr = 0.5 * (lo + hi);
            if (std::fabs(nr - r) < 1e-15) { r = nr; break; }
            r = nr;
        }
        return r;
    }
    double draw_unit() { return rng_.next(); }

    // Keeps (px, py) unless a kept node is too close; returns whether kept.
    bool offer(double px, double py) {
        const double r = std::min(1.0, std::hypot(px, py));
        const double hp = spacing(r);
        // the largest spacing a violating neighbour can have: the spacing
        // nearer the centre, one search radius in
        const double reach = beta * spacing(std::max(0.0, r - beta * hc_));
        const int64_t cx = cell_of(px), cy = cell_of(py);
        const int64_t span = static_cast<int64_t>(std::ceil(reach / cell_));
        for (int64_t j = std::max<int64_t>(0, cy - span); j <= std::min(side_ - 1, cy + span); ++j)
            for (int64_t i = std::max<int64_t>(0, cx - span); i <= std::min(side_ - 1, cx + span);
                 ++i)
                for (int64_t q = head_[static_cast<size_t>(j * side_ + i)]; q >= 0;
                     q = next_[static_cast<size_t>(q)]) {
                    const double dx = px_[static_cast<size_t>(q)] - px;
                    const double dy = py_[static_cast<size_t>(q)] - py;
                    const double need = 0.5 * beta * (hp + hq_[static_cast<size_t>(q)]);
                    if (dx * dx + dy * dy < need * need) return false;
                }
        keep(px, py, hp, cx, cy);
        return true;
    }
    void keep_unchecked(double px, double py) {
        keep(px, py, spacing(std::min(1.0, std::hypot(px, py))), cell_of(px), cell_of(py));
    }
    const std::vector<double>& xs() const { return px_; }
    const std::vector<double>& ys() const { return py_; }

  private:
    // G(r) = integral_0^r s / (1 - a s)^2 ds
    static double mass(double r) {
        const double u = 1.0 - a * r;
        return (1.0 / u - 1.0 + std::log(u)) / (a * a);
    }
    int64_t cell_of(double v) const {
        return std::min(side_ - 1, std::max<int64_t>(0, static_cast<int64_t>((v + 1.0) / cell_)));
    }
    void keep(double px, double py, double hp, int64_t cx, int64_t cy) {
        const size_t c = static_cast<size_t>(cy * side_ + cx);
        next_.push_back(head_[c]);
        head_[c] = static_cast<int64_t>(px_.size());
        px_.push_back(px);
        py_.push_back(py);
        hq_.push_back(hp);
    }

    XorShift64 rng_;
    double hc_, cell_;
    int64_t side_;
    std::vector<int64_t> head_, next_;
    std::vector<double> px_, py_, hq_;
};

}  // namespace

extern "C" {

RBFA_API const char* rbfa_status_string(int32_t status) {
    switch (status) {
        case RBFA_OK: return "ok";
        case RBFA_ERR_ARG: return "invalid argument";
        case RBFA_ERR_UNSUPPORTED: return "unsupported stencil shape";
        case RBFA_ERR_HIP: return "HIP runtime error";
        case RBFA_ERR_NOMEM: return "out of memory";
        case RBFA_ERR_NODEVICE: return "no gfx950 device";
        case RBFA_ERR_FORMAT: return "malformed input file";
        case RBFA_ERR_IO: return "file I/O error";
        default: return "unknown status";
    }
}

RBFA_API int32_t rbfa_morton_keys(int32_t n, const double* x, const double* y,
                                  const rbfa_bbox2* bbox, int32_t ndiv, int64_t* keys) {
    if (n < 0 || ndiv < 1 || ndiv > 31) return RBFA_ERR_ARG;
    if (n == 0) return RBFA_OK;
    if (!x || !y || !keys) return RBFA_ERR_ARG;
    const rbfa_bbox2 bb = bbox ? *bbox : bbox_of(n, x, y);
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; ++i) keys[i] = morton_key2(x[i], y[i], bb, ndiv);
    return RBFA_OK;
}

RBFA_API int32_t rbfa_hilbert_keys(int32_t n, const double* x, const double* y,
                                   const rbfa_bbox2* bbox, int32_t ndiv, int64_t* keys) {
    if (n < 0 || ndiv < 1 || ndiv > 31) return RBFA_ERR_ARG;
    if (n == 0) return RBFA_OK;
    if (!x || !y || !keys) return RBFA_ERR_ARG;
    const rbfa_bbox2 bb = bbox ? *bbox : bbox_of(n, x, y);
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; ++i) keys[i] = hilbert_key2(x[i], y[i], bb, ndiv);
    return RBFA_OK;
}

RBFA_API int32_t rbfa_morton_keys_3d(int32_t n, const double* x, const double* y,
                                     const double* z, const double* bbox6, int32_t ndiv,
                                     int64_t* keys) {
    if (n < 0 || ndiv < 1 || ndiv > 21) return RBFA_ERR_ARG;
    if (n == 0) return RBFA_OK;
    if (!x || !y || !z || !keys) return RBFA_ERR_ARG;
    double bb[6];
    if (bbox6) {
        std::memcpy(bb, bbox6, sizeof bb);
    } else {
        const double* X[3] = {x, y, z};
        for (int d = 0; d < 3; ++d) {
            bb[d] = bb[d + 3] = X[d][0];
            for (int32_t i = 1; i < n; ++i) {
                bb[d] = std::min(bb[d], X[d][i]);
                bb[d + 3] = std::max(bb[d + 3], X[d][i]);
            }
        }
    }
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; ++i) keys[i] = morton_key3(x[i], y[i], z[i], bb, ndiv);
    return RBFA_OK;
}

RBFA_API int32_t rbfa_order_by_keys(int32_t n, const int64_t* keys, int32_t* perm) {
    if (n < 0) return RBFA_ERR_ARG;
    if (n == 0) return RBFA_OK;
    if (!keys || !perm) return RBFA_ERR_ARG;
    stable_order(n, keys, perm);
    return RBFA_OK;
}

RBFA_API int32_t rbfa_permute_f64(int32_t n, const int32_t* perm, double* a) {
    return permute_impl(n, perm, a);
}

RBFA_API int32_t rbfa_permute_i32(int32_t n, const int32_t* perm, int32_t* a) {
    return permute_impl(n, perm, a);
}

RBFA_API int32_t rbfa_renumber_stencils(int32_t n, int32_t k, const int32_t* perm, int32_t* ja) {
    if (n < 0 || k < 1) return RBFA_ERR_ARG;
    if (n == 0) return RBFA_OK;
    if (!perm || !ja || !is_permutation(n, perm)) return RBFA_ERR_ARG;
    const size_t nnz = static_cast<size_t>(n) * k;
    for (size_t e = 0; e < nnz; ++e)
        if (ja[e] < 0 || ja[e] >= n) return RBFA_ERR_ARG;
    std::vector<int32_t> inv(n);
    for (int32_t i = 0; i < n; ++i) inv[perm[i]] = i;
    std::vector<int32_t> tmp(ja, ja + nnz);
#pragma omp parallel for schedule(static)
    for (int32_t s = 0; s < n; ++s)
        for (int32_t j = 0; j < k; ++j)
            ja[static_cast<size_t>(s) * k + j] = inv[tmp[static_cast<size_t>(perm[s]) * k + j]];
    return RBFA_OK;
}

RBFA_API int32_t rbfa_make_row_ptr(int32_t n, int32_t k, int32_t* ia) {
    if (n < 0 || k < 1 || !ia) return RBFA_ERR_ARG;
    if (static_cast<int64_t>(n) * k > std::numeric_limits<int32_t>::max()) return RBFA_ERR_ARG;
    for (int32_t s = 0; s <= n; ++s) ia[s] = s * k;
    return RBFA_OK;
}

RBFA_API int32_t rbfa_knn_stencils(int32_t n, int32_t dim, const double* x, const double* y,
                                   const double* z, int32_t k, int32_t* ja) {
    if (n < 0 || (dim != 2 && dim != 3) || k < 1 || k > n) return RBFA_ERR_ARG;
    if (!x || !y || !ja || (dim == 3 && !z)) return RBFA_ERR_ARG;
    if (static_cast<int64_t>(n) * k > std::numeric_limits<int32_t>::max()) return RBFA_ERR_ARG;
    if (dim == 2) {
        const double* X[2] = {x, y};
        knn_all<2>(n, X, k, ja);
    } else {
        const double* X[3] = {x, y, z};
        knn_all<3>(n, X, k, ja);
    }
    return RBFA_OK;
}

RBFA_API int32_t rbfa_gen_jittered_grid(int32_t dim, int32_t m, double jitter, uint64_t seed,
                                        double* x, double* y, double* z) {
    if ((dim != 2 && dim != 3) || m < 1 || !(jitter >= 0.0 && jitter < 0.5)) return RBFA_ERR_ARG;
    if (!x || !y || (dim == 3 && !z)) return RBFA_ERR_ARG;
    const double h = 1.0 / m;
    XorShift64 rng(seed);
    size_t p = 0;
    const int mz = dim == 3 ? m : 1;
    for (int kk = 0; kk < mz; ++kk)
        for (int j = 0; j < m; ++j)
            for (int i = 0; i < m; ++i, ++p) {
                x[p] = (i + 0.5) * h + jitter * h * (2.0 * rng.next() - 1.0);
                y[p] = (j + 0.5) * h + jitter * h * (2.0 * rng.next() - 1.0);
                if (dim == 3) z[p] = (kk + 0.5) * h + jitter * h * (2.0 * rng.next() - 1.0);
            }
    return RBFA_OK;
}


RBFA_API int32_t rbfa_gen_disk_darts(int32_t n_interior, int32_t n_boundary, uint64_t seed,
                                     double* x, double* y, int32_t* flag) {
    if (n_interior < 1 || n_boundary < 0 || !x || !y || !flag) return RBFA_ERR_ARG;
    GradedDarts d(n_interior, seed);
    const double two_pi = 6.283185307179586;
    for (int32_t j = 0; j < n_boundary; ++j) {
        const double th = two_pi * j / n_boundary;
        d.keep_unchecked(std::cos(th), std::sin(th));
    }
    // a dart count far past what the packing fraction needs means the
    // process has jammed (a boundary far denser than the rim spacing)
    const int64_t max_darts = 64 * static_cast<int64_t>(n_interior) + 1000;
    int64_t kept = 0;
    for (int64_t t = 0; t < max_darts && kept < n_interior; ++t) {
        const double r = d.draw_radius();
        const double th = two_pi * d.draw_unit();
        if (d.offer(r * std::cos(th), r * std::sin(th))) ++kept;
    }
    if (kept < n_interior) return RBFA_ERR_ARG;
    const auto& px = d.xs();
    const auto& py = d.ys();
    for (int64_t i = 0; i < n_interior; ++i) {
        x[i] = px[static_cast<size_t>(n_boundary + i)];
        y[i] = py[static_cast<size_t>(n_boundary + i)];
        flag[i] = 0;
    }
    for (int32_t j = 0; j < n_boundary; ++j) {
        x[n_interior + j] = px[static_cast<size_t>(j)];
        y[n_interior + j] = py[static_cast<size_t>(j)];
        flag[n_interior + j] = 1;
    }
    return RBFA_OK;
}

RBFA_API int32_t rbfa_gen_graded_disk(int32_t n_interior, int32_t n_boundary, uint64_t seed,
                                      double* x, double* y, int32_t* flag) {
    if (n_interior < 1 || n_boundary < 0 || !x || !y || !flag) return RBFA_ERR_ARG;
    // centre spacing for at least n_interior ring nodes (bisection on hc)
    double lo = 1e-6, hi = 0.5;
    if (disk_rings(hi, seed, nullptr, nullptr, 0) >= n_interior) lo = hi;
    for (int it = 0; it < 80 && hi - lo > 1e-15 * hi; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (disk_rings(mid, seed, nullptr, nullptr, 0) >= n_interior) lo = mid; else hi = mid;
    }
    const int64_t have = disk_rings(lo, seed, nullptr, nullptr, 0);
    std::vector<double> bx(static_cast<size_t>(have)), by(static_cast<size_t>(have));
    disk_rings(lo, seed, bx.data(), by.data(), have);
    // drop the surplus nodes at seeded random positions
    std::vector<char> keep(static_cast<size_t>(have), 1);
    XorShift64 rng(seed ^ 0xA5A5A5A5A5A5A5A5ull);
    for (int64_t extra = have - n_interior; extra > 0;) {
        const int64_t i = 1 + static_cast<int64_t>(rng.next() * static_cast<double>(have - 1));
        if (keep[i]) {
            keep[i] = 0;
            --extra;
        }
    }
    int64_t o = 0;
    for (int64_t i = 0; i < have; ++i)
        if (keep[i]) {
            x[o] = bx[i];
            y[o] = by[i];
            flag[o] = 0;
            ++o;
        }
    const double two_pi = 6.283185307179586;
    for (int32_t j = 0; j < n_boundary; ++j, ++o) {
        const double th = two_pi * j / n_boundary;
        x[o] = std::cos(th);
        y[o] = std::sin(th);
        flag[o] = 1;
    }
    return RBFA_OK;
}

}  // extern "C"
