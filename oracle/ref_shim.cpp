// ref_shim.cpp -- extern "C" entry points onto the reference's own host
// code, compiled from its sources where they lie (see oracle/Makefile).
// TEST INFRASTRUCTURE ONLY: the tests and the golden-vector script call
// these to pin the oracle and the product's node layer against the
// reference. Nothing here is copied from the reference; it only calls it:
//
//   ref_stencils_nanoflann  NodeSet::stencils        rbf_nodeset.h:162-181
//   ref_stencils_ckdtree    KdTree::knn_stencils     rbf_spatial.cpp:251-260
//   ref_hilbert_order       hilbert_order            rbf_reorder.h:222-229
//   ref_morton_order        morton_order             rbf_reorder.h:212-220
//   ref_renumber_stencils   renumber_stencils        rbf_reorder.h:278-288
//   ref_read_graph          read_graph_csr           rbf_io.h:649-707
//   ref_write_mm            write_matrix_market      rbf_io.h:782-812
//   rbf_morton_keys / rbf_hilbert_keys are the Fortran bind(c) entry
//   points themselves (rbf_ordering.F90:31-67), linked in unchanged.
#include <cstdint>
#include <cstring>
#include <span>
#include <vector>

#include "rbf_io.h"
#include "rbf_nodeset.h"
#include "rbf_reorder.h"
#include "rbf_spatial.h"

extern "C" {

int ref_stencils_nanoflann(int32_t n, const double* x, const double* y, int32_t k, int32_t* ja) {
    rbf::NodeSet<double, int32_t> ns(std::vector<double>(x, x + n), std::vector<double>(y, y + n),
                                     std::vector<int>(static_cast<size_t>(n), 0));
    const auto out = ns.stencils(k);
    std::memcpy(ja, out.data(), sizeof(int32_t) * out.size());
    return 0;
}

int ref_stencils_ckdtree(int32_t n, int32_t ndim, const double* pts_interleaved, int32_t k,
                         int32_t* ja) {
    rbf::spatial::KdTree tree(
        std::span<const double>(pts_interleaved, static_cast<size_t>(n) * ndim), ndim);
    const auto out = tree.knn_stencils(k);
    std::memcpy(ja, out.data(), sizeof(int32_t) * out.size());
    return 0;
}

int ref_hilbert_order(int32_t n, const double* x, const double* y, int32_t ndiv, int32_t* perm) {
    const auto p = rbf::hilbert_order<int32_t>(std::span<const double>(x, n),
                                               std::span<const double>(y, n), ndiv);
    std::memcpy(perm, p.map().data(), sizeof(int32_t) * n);
    return 0;
}

int ref_morton_order(int32_t n, const double* x, const double* y, int32_t ndiv, int32_t* perm) {
    const auto p = rbf::morton_order<int32_t>(std::span<const double>(x, n),
                                              std::span<const double>(y, n), ndiv);
    std::memcpy(perm, p.map().data(), sizeof(int32_t) * n);
    return 0;
}

int ref_renumber_stencils(int32_t n, int32_t k, const int32_t* perm, int32_t* ja) {
    const rbf::Permutation<int32_t> p(std::vector<int32_t>(perm, perm + n));
    rbf::renumber_stencils(std::span<int32_t>(ja, static_cast<size_t>(n) * k), k, p);
    return 0;
}


// Sizes (ia == NULL) or contents of a .graph file. Valid files only: the
// reference's reader ends the process (exit(1)) on a malformed one.
int ref_read_graph(const char* path, int32_t k, int32_t* n, int64_t* nnz, int32_t* ia,
                   int32_t* ja) {
    try {
        auto [ria, rja] = rbf::io::read_graph_csr<int32_t>(path, k);
        *n = static_cast<int32_t>(ria.size() - 1);
        *nnz = static_cast<int64_t>(rja.size());
        if (ia) {
            std::memcpy(ia, ria.data(), ria.size() * sizeof(int32_t));
            std::memcpy(ja, rja.data(), rja.size() * sizeof(int32_t));
        }
        return 0;
    } catch (...) {
        return -1;
    }
}

int ref_write_mm(const char* path, int64_t rows, int64_t cols, const int32_t* ia,
                 const int32_t* ja, const double* a) {
    try {
        if (a)
            rbf::io::write_matrix_market<double, int32_t>(path, rows, cols, ia, ja, a, 0);
        else
            rbf::io::write_matrix_market_pattern<int32_t>(path, rows, cols, ia, ja, 0);
        return 0;
    } catch (...) {
        return -1;
    }
}
}  // extern "C"
