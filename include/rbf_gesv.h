// rbf_gesv.h -- the block GESV of the reference's drop-in interface on one
// gfx950 wavefront: LU with partial pivoting and solve of a caller-filled
// system, the factors, pivots and solution written back in place, as
// cuSolverDx's gesv_partial_pivot::execute(A, ipiv, B, info) leaves them
// (rbf_operators.h:16-37 of the reference; LAPACK dgetrf / dgetrs layout).
//
//   Solver<T, function::gesv_partial_pivot, NT, NRHS, ARCH>, GESV<...>
//     one system per 64-thread block: block_dim, is_block_execution,
//     batches_per_block, shared_memory_size, execute(A, ipiv, B, info)
//   GELS<...>   named for completeness; using it is a compile-time error
//   solve_system_AB<Solver>(A, ipiv, B, info)
//
// The elimination is the wave LU of rbf_operators.h (one matrix row per
// lane, rows never moved, pivot = first largest |a| in row order), so A,
// ipiv and X are bit-identical to the CPU oracle's textbook elimination:
//   A     column-major NT x NT (leading dimension NT): L (unit diagonal,
//         below) and U (on and above), the row that was pivot at step k
//         in row k;
//   ipiv  1-based, dgetrf's sequential interchanges (row k swapped with
//         row ipiv[k] - 1 before step k);
//   B     column-major NT x NRHS: the solution;
//   info  0, or k when the k-th pivot (1-based) was exactly zero (the
//         factors are still written; the solution is then not meaningful).
// Arithmetic is FP64 for T = double and T = float alike (float data is
// widened on load and rounded on store).
#ifndef RBF_GESV_H
#define RBF_GESV_H

#include "rbf_wave.h"

namespace rbf_operators {
namespace detail {

constexpr int gesv_pad(int n) { return n <= 16 ? 16 : (n <= 32 ? 32 : 64); }

// One system on one wave64. NTM / NRM: compile-time bounds on the order
// and the number of right-hand sides; nt <= NTM rows are real, the rows
// and columns up to NTM are the identity (never pivots of real columns).
template <int NTM, int NRM>
struct WaveGesv {
    static constexpr int NC = NTM + NRM;
    static_assert(NTM <= 64, "one row per lane");
    using tune = LuTune<0, 2>;

    template <typename T>
    __device__ __forceinline__ static int solve(int nt, int nrhs, T* A, int lda, int* ipiv, T* B,
                                                int ldb) {
#pragma clang fp contract(off)
        const int lane = lane_id();
        const bool real = lane < nt;
        double row[NC];
#pragma unroll
        for (int j = 0; j < NTM; ++j) {
            double v = j == lane ? 1.0 : 0.0;
            if (real) v = j < nt ? static_cast<double>(A[lane + j * lda]) : 0.0;
            row[j] = v;
        }
#pragma unroll
        for (int c = 0; c < NRM; ++c)
            row[NTM + c] = (real && c < nrhs) ? static_cast<double>(B[lane + c * ldb]) : 0.0;
        ElimState<NTM, double> st;
        elim_from<0, NTM, NC, tune, true>(row, st, static_cast<double*>(nullptr), lane);
        // dgetrf's interchanges: this row's position under the swaps of
        // the steps so far; step k swaps positions k and pos(p_k)
        int pos = lane, myipiv = 0;
#pragma unroll
        for (int k = 0; k < NTM; ++k) {
            if (k < nt) {
                const int p = st.perm[k];
                const int q = __builtin_amdgcn_readlane(pos, p);
                myipiv = lane == k ? q + 1 : myipiv;
                pos = lane == p ? k : (pos == k ? q : pos);
            }
        }
        // back substitution in the oracle's order (rows pivotal before step
        // k hold u_ik; the others take nothing)
#pragma unroll
        for (int k = NTM - 1; k >= 0; --k) {
            const int p = st.perm[k];
            const double uk = st.mystep < k ? row[k] : 0.0;
#pragma unroll
            for (int c = 0; c < NRM; ++c) {
                const double t = bcast(row[NTM + c] * st.myrcp, p);
                row[NTM + c] = fma(-uk, t, row[NTM + c]);
            }
        }
        if (real) {
            const int at = st.mystep;  // this row's place in the factors
#pragma unroll
            for (int j = 0; j < NTM; ++j)
                if (j < nt) A[at + j * lda] = static_cast<T>(row[j]);
#pragma unroll
            for (int c = 0; c < NRM; ++c)
                if (c < nrhs) B[at + c * ldb] = static_cast<T>(row[NTM + c] * st.myrcp);
            ipiv[lane] = myipiv;
        }
        return st.info;
    }
};

}  // namespace detail

// ---------------------------------------------------------------------------
// The reference's solver names (rbf_operators.h:16-37)
// ---------------------------------------------------------------------------

enum class function { gesv_partial_pivot, gels };

template <typename T, function F, int NT, int NRHS, unsigned int ARCH>
struct Solver {
    static_assert(F == function::gesv_partial_pivot,
                  "GELS (least squares by QR) is not provided on gfx950: use GESV");
};

template <typename T, int NT, int NRHS, unsigned int ARCH>
struct Solver<T, function::gesv_partial_pivot, NT, NRHS, ARCH> {
    static_assert(std::is_same_v<T, double> || std::is_same_v<T, float>, "real FP32 or FP64");
    static_assert(NT >= 1 && NT <= 64, "one system row per lane of a wave64");
    static_assert(NRHS >= 0 && NRHS <= 16, "at most 16 right-hand sides per solve");
    using a_data_type = T;
    using b_data_type = T;
    using status_type = int;
    static constexpr bool is_block_execution = true;
    static constexpr unsigned int batches_per_block = 1;
    static constexpr unsigned int sm = ARCH;
    static constexpr unsigned int max_threads_per_block = 64;
    static constexpr dim3 block_dim = dim3(64, 1, 1);
    // A and B, as the caller holds them (the factorisation itself works in
    // registers)
    static constexpr std::size_t shared_memory_size = sizeof(T) * (NT * NT + NT * NRHS);

    // Every lane of the block's wave calls it; A, ipiv, B in memory the
    // wave can read and write (LDS or global), info written by lane 0.
    __device__ void execute(a_data_type* A, int* ipiv, b_data_type* B, status_type* info) const {
        using E = detail::WaveGesv<detail::gesv_pad(NT), detail::gesv_pad(NRHS == 0 ? 1 : NRHS)>;
        const int r = E::solve(NT, NRHS, A, NT, ipiv, B, NT);
        if (detail::lane_id() == 0) *info = r;
    }
};

// Solve via LU factorization with partial pivoting
template <typename T, int NT, int NRHS, unsigned int ARCH>
using GESV = Solver<T, function::gesv_partial_pivot, NT, NRHS, ARCH>;

// General least squares (QR): named as in the reference, not provided
template <typename T, int NT, int NRHS, unsigned int ARCH>
using GELS = Solver<T, function::gels, NT, NRHS, ARCH>;

// One system solved in place by the block (rbf_operators.h:29-37).
template <class GESV_solver>
__device__ void solve_system_AB(typename GESV_solver::a_data_type* A, int* ipiv,
                                typename GESV_solver::b_data_type* B,
                                typename GESV_solver::status_type* info) {
    static_assert(GESV_solver::is_block_execution, "assumes block configuration");
    GESV_solver().execute(A, ipiv, B, info);
}

}  // namespace rbf_operators

#endif  // RBF_GESV_H
