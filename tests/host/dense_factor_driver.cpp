// Test driver: reads a CSR matrix from stdin; prints the component of each
// row and whether the components are contiguous, then per component its size
// followed by the dense inverse (row-major) from the library's host dense
// Cholesky (csrc/dense_factor.hpp).  Exit code 3: not positive definite.
#include <cstdio>
#include <iostream>
#include <vector>

#include "dense_factor.hpp"

int main() {
    int n = 0, nnz = 0;
    std::cin >> n >> nnz;
    std::vector<int> rowptr(n + 1), col(nnz);
    std::vector<double> val(nnz);
    for (auto& x : rowptr) std::cin >> x;
    for (auto& x : col) std::cin >> x;
    for (auto& x : val) std::cin >> x;
    int nc = 0;
    bool contiguous = false;
    const std::vector<int> comp = sbs::row_components(n, rowptr, col, &nc, &contiguous);
    std::printf("%d %d\n", nc, contiguous ? 1 : 0);
    for (int i = 0; i < n; ++i) std::printf("%d ", comp[i]);
    std::printf("\n");
    if (!contiguous) return 0;
    int r0 = 0;
    try {
        for (int c = 0; c < nc; ++c) {
            int m = 0;
            while (r0 + m < n && comp[r0 + m] == c) ++m;
            std::vector<double> a = sbs::dense_block(r0, m, rowptr, col, val), inv;
            sbs::dense_cholesky_inverse(m, a, inv);
            std::printf("%d\n", m);
            for (double v : inv) std::printf("%.17g ", v);
            std::printf("\n");
            r0 += m;
        }
    } catch (const std::domain_error&) {
        return 3;
    }
    return 0;
}
