// apply_layout.cpp -- host construction of the blocked apply layout (see
// apply_layout.h): greedy row blocking, then per block a sorted distinct
// column list and the local index of every nonzero (OpenMP over blocks).

#include "apply_layout.h"

#include <algorithm>

namespace rbf_amd_internal {

int build_apply_layout(int32_t nrows, int32_t ncols, int32_t k, const int32_t* ia,
                       const int32_t* ja, ApplyLayout& out) {
    auto row_start = [&](int64_t r) -> int64_t { return ia ? ia[r] : r * k; };
    const int64_t nnz = row_start(nrows);
    if (ia && ia[0] != 0) return 1;
    // 1. greedy blocking: grow a block while it stays within both limits
    out.blk_row.assign(1, 0);
    int64_t r = 0;
    while (r < nrows) {
        const int64_t e0 = row_start(r);
        int64_t r1 = r;
        while (r1 < nrows && r1 - r < kApplyRows) {
            const int64_t w = row_start(r1 + 1) - row_start(r1);
            if (w < 0 || w > kApplyEntries) return 1;
            if (row_start(r1 + 1) - e0 > kApplyEntries) break;
            ++r1;
        }
        out.blk_row.push_back(static_cast<int32_t>(r1));
        r = r1;
    }
    const int64_t nblk = static_cast<int64_t>(out.blk_row.size()) - 1;
    int bad = 0;
#pragma omp parallel for reduction(| : bad)
    for (int64_t e = 0; e < nnz; ++e) bad |= (ja[e] < 0 || ja[e] >= ncols) ? 1 : 0;
    if (bad) return 1;
    // 2. distinct columns per block (count, then fill at the prefix offsets)
    std::vector<int32_t> count(nblk);
#pragma omp parallel
    {
        std::vector<int32_t> buf;
#pragma omp for schedule(dynamic, 64)
        for (int64_t b = 0; b < nblk; ++b) {
            const int64_t e0 = row_start(out.blk_row[b]), e1 = row_start(out.blk_row[b + 1]);
            buf.assign(ja + e0, ja + e1);
            std::sort(buf.begin(), buf.end());
            count[b] = static_cast<int32_t>(std::unique(buf.begin(), buf.end()) - buf.begin());
        }
    }
    out.blk_uoff.assign(nblk + 1, 0);
    out.umax = 0;
    for (int64_t b = 0; b < nblk; ++b) {
        out.blk_uoff[b + 1] = out.blk_uoff[b] + count[b];
        out.umax = std::max(out.umax, count[b]);
    }
    out.ucols.resize(out.blk_uoff[nblk]);
    out.lidx.resize(nnz);
#pragma omp parallel
    {
        std::vector<int32_t> buf;
#pragma omp for schedule(dynamic, 64)
        for (int64_t b = 0; b < nblk; ++b) {
            const int64_t e0 = row_start(out.blk_row[b]), e1 = row_start(out.blk_row[b + 1]);
            buf.assign(ja + e0, ja + e1);
            std::sort(buf.begin(), buf.end());
            buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
            std::copy(buf.begin(), buf.end(), out.ucols.begin() + out.blk_uoff[b]);
            for (int64_t e = e0; e < e1; ++e)
                out.lidx[e] = static_cast<uint16_t>(
                    std::lower_bound(buf.begin(), buf.end(), ja[e]) - buf.begin());
        }
    }
    return 0;
}

}  // namespace rbf_amd_internal
