// Test driver: reads a row pattern (CSR over vertex ids) and vertex positions
// from stdin, builds the halo tiles of csrc/halo_tiles.hpp with Morton keys
// of the positions and prints every array.  Exit code 3: the build refused
// the pattern (std::invalid_argument).
#include <cstdio>
#include <iostream>
#include <vector>

#include "halo_tiles.hpp"

template <class V> void put(const char* name, const V& v) {
    std::printf("%s", name);
    for (const auto& x : v) std::printf(" %lld", (long long)x);
    std::printf("\n");
}

int main() {
    int nv = 0, nnz = 0, rpg = 0, stride = 0;
    std::cin >> nv >> nnz >> rpg >> stride;
    std::vector<int> rowstart(nv + 1);
    std::vector<uint32_t> cols(nnz);
    std::vector<double> X(3 * (size_t)nv);
    for (auto& x : rowstart) std::cin >> x;
    for (auto& x : cols) std::cin >> x;
    for (auto& x : X) std::cin >> x;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int v = 0; v < nv; ++v)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], X[3 * (size_t)v + k]);
            hi[k] = std::max(hi[k], X[3 * (size_t)v + k]);
        }
    try {
        const sbs::HaloTiles h = sbs::build_halo_tiles(
            nv, rowstart, cols, [&](int v) { return sbs::morton_code(&X[3 * (size_t)v], lo, hi); },
            [](int k) { return (int64_t)k; }, rpg, stride);
        std::printf("max_halo %d\nmax_width %d\n", h.max_halo, h.max_width);
        std::vector<uint32_t> keys(nv);
        for (int v = 0; v < nv; ++v) keys[v] = sbs::morton_code(&X[3 * (size_t)v], lo, hi);
        put("keys", keys);
        put("row_perm", h.row_perm);
        put("row_of", h.row_of);
        put("g_row0", h.g_row0);
        put("g_rows", h.g_rows);
        put("g_halo_n", h.g_halo_n);
        put("g_width", h.g_width);
        put("g_eptr", h.g_eptr);
        put("g_own0", h.g_own0);
        put("halo", h.halo);
        put("halo_vertex", h.halo_vertex);
        put("col", h.col);
        put("src", h.src);
    } catch (const std::invalid_argument& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 3;
    }
    return 0;
}
