// Prefactored Cholesky global solve for small bodies.
//
// Replaces the reference's CholeskySpLinearSolver (host AMD ordering +
// cuSOLVER csrchol factor, src/simulation/solver/linear/cholesky.cu:73-158,
// solve at :185-191) and the EigenCholesky path (pdSolver.cu:178-185).  Both
// produce x = A^-1 b for the prefactored SPD matrix; this solver does the
// same with its own factor:
//   * the scalar matrix A_s (the reference factors A_s (x) I_3) is reordered
//     by reverse Cuthill-McKee when that narrows its band;
//   * A_s = L L^T is factored once on the host in double (band Cholesky) and
//     L is uploaded in the solve precision, stored by diagonals
//     (Ld[d*N + i] = L[i][i-d]) so a wavefront reading one diagonal for 64
//     consecutive rows issues one coalesced load;
//   * the solve runs one workgroup per (independent system, right-hand side
//     component): rows are processed in 64-row blocks -- a band update
//     spread over 4 wavefronts, then the 64x64 triangular tile solved by one
//     wavefront with register broadcasts -- forward for L, backward for L^T;
//     the band update reads the previous (next) bw rows' values from a ring
//     of bw + 64 rows in LDS, so the limit is the bandwidth (up to the LDS a
//     workgroup may use), not the number of rows.
#pragma once

#include <vector>

#include "common.hpp"

namespace sbs {

template <class T> class BandCholesky {
public:
    // CSR of the full symmetric matrix (both triangles, sorted columns).
    void factor(int n, const std::vector<int>& rowptr, const std::vector<int>& col, const std::vector<T>& val,
                hipStream_t stream);
    // x = A^-1 b for 3 interleaved right-hand sides; kernels return early
    // when *skip != 0 (device flag, may be null).
    void solve(const T* b, T* x, hipStream_t stream, const int* skip) const;
    int n() const { return n_; }
    int bandwidth() const { return bw_; }
    int systems() const { return nsys_; }
    // largest half bandwidth the solve can hold (the LDS of one workgroup)
    static int max_bandwidth();

private:
    int n_ = 0, bw_ = 0, nsys_ = 0, ring_ = 0;
    DevBuf<T> ld_;          // (bw+1) diagonals of N
    DevBuf<int> perm_;      // position i holds original row perm[i]
    DevBuf<int> sys_;       // first row of each independent system (nsys+1)
    DevBuf<T> ybuf_;        // forward-sweep result, 3 per row
};

}  // namespace sbs
