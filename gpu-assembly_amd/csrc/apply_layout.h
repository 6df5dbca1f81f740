// apply_layout.h -- internal to librbf_amd: the blocked layout of an
// operator's CSR structure used by the planned operator application.
//
// Rows are cut into consecutive blocks of at most kApplyRows rows and
// kApplyEntries nonzeros. Each block lists its distinct columns once,
// sorted (ucols[blk_uoff[b] .. blk_uoff[b+1])), and every nonzero carries
// its column's position in that list as a 16-bit local index, in the CSR's
// own entry order, so the assembled weight arrays are used as they are.
#ifndef RBF_APPLY_LAYOUT_H
#define RBF_APPLY_LAYOUT_H

#include <cstdint>
#include <vector>

namespace rbf_amd_internal {

constexpr int kApplyRows = 256;
constexpr int kApplyEntries = 2048;

struct ApplyLayout {
    std::vector<int32_t> blk_row;   // nblk + 1 row offsets
    std::vector<int32_t> blk_uoff;  // nblk + 1 offsets into ucols
    std::vector<int32_t> ucols;     // distinct columns per block, ascending
    std::vector<uint16_t> lidx;     // nnz local column indices
    int32_t umax = 0;               // most distinct columns in one block
};

// 0 on success; 1 if an argument is out of range (a column outside
// [0, ncols), ia not non-decreasing, a row wider than kApplyEntries).
int build_apply_layout(int32_t nrows, int32_t ncols, int32_t k, const int32_t* ia,
                       const int32_t* ja, ApplyLayout& out);

}  // namespace rbf_amd_internal

#endif  // RBF_APPLY_LAYOUT_H
