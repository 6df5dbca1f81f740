// knn_grid.h -- internal to librbf_amd: the bucket-grid parameters of the
// k-nearest-neighbour search, shared by the host search (host_nodes.cpp)
// and the device search (knn_device.hip) so that both build the same grid,
// visit the same cells in the same ring order and stop at the same ring.
#ifndef RBF_KNN_GRID_H
#define RBF_KNN_GRID_H

#include <algorithm>
#include <cmath>

namespace rbf_amd_internal {

// From the bounding box lo/hi of n points: about max(2, k/8) points per
// cell; at most 8192 (2-D) / 512 (3-D) cells per axis; the cell edge is the
// largest per-axis extent over the cell count, widened by 1e-12 so the
// last cell holds the maximum. Returns the cell edge and writes dims.
template <int D>
inline double grid_params(const double* lo, const double* hi, long long n, int k, int* dims) {
    double vol = 1.0;
    int nonflat = 0;
    for (int d = 0; d < D; ++d)
        if (hi[d] > lo[d]) {
            vol *= hi[d] - lo[d];
            ++nonflat;
        }
    const double per_cell = std::max(2.0, k / 8.0);
    double cell = nonflat ? std::pow(vol * per_cell / static_cast<double>(n), 1.0 / nonflat) : 1.0;
    if (!(cell > 0) || !std::isfinite(cell)) cell = 1.0;
    for (int d = 0; d < D; ++d) {
        const double ext = hi[d] - lo[d];
        long long m = ext > 0 ? static_cast<long long>(ext / cell) + 1 : 1;
        m = std::min<long long>(m, D == 2 ? 8192 : 512);
        dims[d] = static_cast<int>(m);
    }
    double c = 0;
    for (int d = 0; d < D; ++d) {
        const double ext = hi[d] - lo[d];
        if (ext > 0) c = std::max(c, ext / dims[d]);
    }
    return c > 0 ? c * (1.0 + 1e-12) : 1.0;
}

}  // namespace rbf_amd_internal

#endif  // RBF_KNN_GRID_H
