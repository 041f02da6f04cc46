// Dense small-body solver: per-body inverse applied with MFMA, corrected by
// FP64 iterative refinement on the GPU.
//
// Reference: CholeskyDnLinearSolver (src/simulation/solver/linear/cholesky.cu:
// 26-70, 161-166): dense potrf once, potrs per Solve.  Here each body's
// inverse A_b^-1 is formed once on the host in FP64 (Cholesky) and stored in
// FP32 or BF16 MFMA fragments; a Solve runs
//     x = 0;  r = b;  repeat:  x += A~^-1 r  (MFMA, FP32 accumulate, x in FP64)
//                              r  = b - A x   (FP64, the assembled sparse A)
// until max |D^-1 r| <= tol max |D^-1 b| per component (D = diag A; the pinned
// rows make an unscaled residual meaningless), or max_iter corrections.
// Bodies whose x, residual fragments and scratch fit 64 KB of LDS (up to 736
// padded rows with FP32 fragments, 1152 with BF16) run the whole loop in one
// workgroup of 16 wavefronts per body, each body with its own stop test;
// larger ones run each correction and residual as launches over all bodies
// with one stop test.
#pragma once

#include <cstdint>
#include <vector>

#include "common.hpp"

namespace sbs {

enum class DensePrecision : int { FP32 = 0, BF16 = 1 };

// Sparse A (sliced ELL off-diagonal + diagonal), device arrays owned by the caller.
template <class T> struct EllView {
    int n = 0;
    const int* slice_ptr = nullptr;
    const int* slice_w = nullptr;
    const uint32_t* ecol = nullptr;
    const T* eval = nullptr;
    const T* diag = nullptr;
};

template <class T> class DenseRefinedSolver {
public:
    // rowptr/col/val: the same matrix on the host (CSR, diagonal included),
    // whose connected components must be contiguous row ranges (bodies).
    void setup(const EllView<T>& A, const std::vector<int>& rowptr, const std::vector<int>& col,
               const std::vector<T>& val, DensePrecision prec, hipStream_t stream);
    // x = A^-1 b for 3 interleaved right-hand sides (device arrays, 3n each).
    // Skipped entirely when *skip != 0 (device flag; may be null).
    // x = A^-1 b to tol (max-norm residual relative to b); with `warm` the
    // one-workgroup-per-body form starts from the caller's x instead of 0
    void solve(const T* b, T* x, double tol, int max_iter, const int* skip, bool warm = false);
    // Corrections applied by the last solve (-1 - n if it stopped at max_iter)
    // and the scaled residual ratio reached; waits for the stream.
    void last_result(int* iterations, double* ratio);
    int bodies() const { return nb_; }
    int max_body() const { return max_n_; }
    size_t inverse_bytes() const { return tile_bytes_; }
    DensePrecision precision() const { return prec_; }
    // true when every body is solved by one workgroup in one launch per solve
    bool one_workgroup_per_body() const { return fused_lds_ != 0; }
    void set_stream(hipStream_t s) { stream_ = s; }

private:
    EllView<T> A_{};
    DensePrecision prec_ = DensePrecision::FP32;
    hipStream_t stream_ = nullptr;
    int n_ = 0, nb_ = 0, max_n_ = 0, nitems_ = 0;
    size_t tile_bytes_ = 0;
    DevBuf<int> body_row0_, body_n_, body_p_;
    DevBuf<long long> tile_off_, rhs_off_;
    DevBuf<int2> items_;                 // (body, 16-row block) per wavefront
    DevBuf<int> row_body_;               // body of each row
    DevBuf<unsigned char> tiles_;        // inverse fragments (FP32 or BF16)
    DevBuf<unsigned char> rhs_;          // residual as MFMA B fragments
    DevBuf<double> x64_;                 // 3n
    DevBuf<unsigned long long> norms_;   // per iteration: max |D^-1 r| xyz, then max |D^-1 b| xyz
    DevBuf<double> report_;              // [iterations, ratio]
    DevBuf<int> body_iters_;             // one-workgroup-per-body solve: corrections per body
    DevBuf<double> body_ratio_;          //   and the residual ratio each reached
    size_t fused_lds_ = 0;               // LDS of that solve; 0: bodies too large, per-row-block launches
    int fused_nseg_ = 1;                 // K segments of its products (partial sums in LDS)
    double* host_report_ = nullptr;
    int slots_ = 0;
    int last_max_iter_ = 0;

public:
    DenseRefinedSolver() = default;
    ~DenseRefinedSolver();
    DenseRefinedSolver(const DenseRefinedSolver&) = delete;
    DenseRefinedSolver& operator=(const DenseRefinedSolver&) = delete;
};

}  // namespace sbs
