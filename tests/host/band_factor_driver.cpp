// Test driver: reads a CSR matrix from stdin, factors it with the library's
// host band Cholesky (csrc/band_factor.hpp), prints bw, perm, L by rows, the
// first row of each independent system and the bandwidth estimate.
// Exit code 2: invalid input, 3: not positive definite.
#include <cstdio>
#include <iostream>
#include <vector>

#include "band_factor.hpp"

int main() {
    int n = 0, nnz = 0;
    std::cin >> n >> nnz;
    std::vector<int> rowptr(n + 1), col(nnz);
    std::vector<double> val(nnz);
    for (auto& x : rowptr) std::cin >> x;
    for (auto& x : col) std::cin >> x;
    for (auto& x : val) std::cin >> x;
    try {
        const sbs::BandFactor f = sbs::band_factor(n, rowptr, col, val);
        std::printf("%d %d\n", f.n, f.bw);
        for (int i = 0; i < n; ++i) std::printf("%d ", f.perm[i]);
        std::printf("\n");
        for (size_t i = 0; i < f.Lb.size(); ++i) std::printf("%.17g ", f.Lb[i]);
        std::printf("\n");
        for (int p : sbs::band_systems(n, rowptr, col, f.perm)) std::printf("%d ", p);
        std::printf("\n%d\n", sbs::band_width_estimate(n, rowptr, col));
    } catch (const std::invalid_argument&) {
        return 2;
    } catch (const std::domain_error&) {
        return 3;
    }
    return 0;
}
