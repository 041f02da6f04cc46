// General sparse linear solve from COO triplets on gfx950: the reference's
// LinearSolver<T>::Solve(N, d_b, d_x, A, nz, rowIdx, colIdx, d_guess)
// (src/simulation/solver/linear/linear.h:62) with the PCGJacobiSolver
// algorithm (pcgJacobi.cu:89-173), as IPCSolver calls it (ipc.cu:239):
//   * the triplets may be unsorted and repeat an entry: sort_coo
//     (linear.cu:7-40) sorts them by (row, column) and sums the repeats;
//   * x0 = d_guess (or 0), r = b - A x0;
//   * each iteration: stop when ||r||_2 < tol (absolute), z = D^-1 r,
//     rho = r.z, stop when |rho| < 1e-15, p = z (+ beta p), q = A p,
//     alpha = rho / p.q, x += alpha p, r -= alpha q; at most max_iter.
// Dot products accumulate in double whatever T is.  The scalars stay on the
// device; the host reads the control word every kCheck iterations.
#pragma once

#include <cstddef>
#include <cstdint>

#include "common.hpp"

namespace sbs {

// Size of the fixed device workspace of the allocation-free solve (one per
// GPU, for the life of the process): 32 MiB, about N = 100,000 unknowns with
// 400,000 triplets in FP64.
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;

// Device workspace of one solve, carved from one buffer: the sorted triplets,
// the CSR arrays, the five PCG vectors, reduction partials and the control
// block.  `bytes` is what solve() needs for (N, nz).
struct CooWorkspace {
    static size_t bytes(int n, long long nz, size_t elem, size_t sort_tmp);
};

template <class T> class CooPcgSolver {
public:
    CooPcgSolver(int max_iter, double tol) : max_iter_(max_iter), tol_(tol) {}
    // Solve with library-owned workspace (grown with hipMalloc as needed).
    int solve(int n, const T* b, T* x, const T* A, long long nz, const int* row, const int* col, const T* guess,
              hipStream_t stream);
    // Solve in caller-provided device memory of `ws_bytes` (no allocation);
    // throws when it is too small.
    int solve_in(int n, const T* b, T* x, const T* A, long long nz, const int* row, const int* col, const T* guess,
                 hipStream_t stream, void* ws, size_t ws_bytes);
    // Solve in the library's fixed per-GPU workspace (kNoAllocScratchBytes):
    // no hipMalloc or hipFree; throws when (n, nz) needs more.  Calls on one
    // GPU share that workspace, so calls that overlap in time on one GPU race.
    int solve_no_alloc(int n, const T* b, T* x, const T* A, long long nz, const int* row, const int* col,
                       const T* guess, hipStream_t stream);
    // hipCUB temporary bytes for sorting and reducing nz triplets.
    static size_t sort_tmp_bytes(long long nz, int n);
    long long last_nnz() const { return nnz_; }
    void set_limits(int max_iter, double tol) { max_iter_ = max_iter; tol_ = tol; }

private:
    int max_iter_;
    double tol_;
    long long nnz_ = 0;
    DevBuf<unsigned char> ws_;
};

}  // namespace sbs
