// Host-side dense factorisation of small bodies: the connected components of
// a symmetric sparse matrix, and per component the FP64 Cholesky factor and
// the explicit inverse.  Plain C++17 (no HIP) so the CPU test suite compiles
// it alone to check it.
//
// Reference: CholeskyDnLinearSolver (src/simulation/solver/linear/cholesky.cu:
// 26-70, 161-166) factors the dense system once (potrf) and solves per call
// (potrs).  Here the per-body inverse is formed once so that the per-call
// solve is one batched matrix product (MFMA on the GPU), corrected to the
// requested accuracy by FP64 iterative refinement.
#pragma once

#include <cmath>
#include <cstdint>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

namespace sbs {

// Connected components of the pattern (rowptr/col over n rows).  Returns the
// component of each row; components are numbered in order of their smallest
// row, and every component is a contiguous row range when `contiguous` is
// reported true.
inline std::vector<int> row_components(int n, const std::vector<int>& rowptr, const std::vector<int>& col,
                                       int* ncomp, bool* contiguous) {
    std::vector<int> parent(n);
    std::iota(parent.begin(), parent.end(), 0);
    auto find = [&](int a) {
        while (parent[a] != a) a = parent[a] = parent[parent[a]];
        return a;
    };
    for (int i = 0; i < n; ++i)
        for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
            int a = find(i), b = find(col[k]);
            if (a != b) parent[a < b ? b : a] = a < b ? a : b;
        }
    std::vector<int> comp(n, -1), id(n, -1);
    int nc = 0;
    for (int i = 0; i < n; ++i) {
        const int r = find(i);
        if (id[r] < 0) id[r] = nc++;
        comp[i] = id[r];
    }
    bool contig = true;
    for (int i = 1; i < n; ++i)
        if (comp[i] != comp[i - 1] && comp[i] != comp[i - 1] + 1) contig = false;
    for (int i = 1; i < n && contig; ++i)
        if (comp[i] < comp[i - 1]) contig = false;
    *ncomp = nc;
    *contiguous = contig;
    return comp;
}

// In-place Cholesky A = L L^T of a dense SPD matrix (row-major, n x n, lower
// triangle used), then the inverse A^-1 = L^-T L^-1 written to `inv`
// (row-major, full).  Throws std::domain_error if A is not positive definite.
inline void dense_cholesky_inverse(int n, std::vector<double>& a, std::vector<double>& inv) {
    for (int j = 0; j < n; ++j) {
        double d = a[(size_t)j * n + j];
        for (int k = 0; k < j; ++k) d -= a[(size_t)j * n + k] * a[(size_t)j * n + k];
        if (!(d > 0.0)) throw std::domain_error("dense Cholesky: matrix not positive definite at row " + std::to_string(j));
        d = std::sqrt(d);
        a[(size_t)j * n + j] = d;
        for (int i = j + 1; i < n; ++i) {
            double s = a[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) s -= a[(size_t)i * n + k] * a[(size_t)j * n + k];
            a[(size_t)i * n + j] = s / d;
        }
    }
    // Linv = L^-1 (lower), column by column
    std::vector<double> li((size_t)n * n, 0.0);
    for (int c = 0; c < n; ++c) {
        li[(size_t)c * n + c] = 1.0 / a[(size_t)c * n + c];
        for (int i = c + 1; i < n; ++i) {
            double s = 0.0;
            for (int k = c; k < i; ++k) s += a[(size_t)i * n + k] * li[(size_t)k * n + c];
            li[(size_t)i * n + c] = -s / a[(size_t)i * n + i];
        }
    }
    // inv = Linv^T Linv (symmetric)
    inv.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) {
            double s = 0.0;
            for (int k = i; k < n; ++k) s += li[(size_t)k * n + i] * li[(size_t)k * n + j];
            inv[(size_t)i * n + j] = s;
            inv[(size_t)j * n + i] = s;
        }
}

// Dense block of rows [r0, r0 + n) of a CSR matrix whose columns of those rows
// stay inside the block.
template <class V>
std::vector<double> dense_block(int r0, int n, const std::vector<int>& rowptr, const std::vector<int>& col,
                                const std::vector<V>& val) {
    std::vector<double> a((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int k = rowptr[r0 + i]; k < rowptr[r0 + i + 1]; ++k) {
            const int j = col[k] - r0;
            if (j < 0 || j >= n) throw std::invalid_argument("dense_block: coupling outside the block");
            a[(size_t)i * n + j] = (double)val[k];
        }
    return a;
}

}  // namespace sbs
