// Host-side band Cholesky factorisation with reverse Cuthill-McKee ordering.
// Plain C++17 (no HIP): the GPU direct solver uploads the factor this
// computes, and the CPU test suite compiles it alone to check it.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <queue>
#include <stdexcept>
#include <string>
#include <vector>

namespace sbs {

struct BandFactor {
    int n = 0;
    int bw = 0;                // half bandwidth of the reordered matrix
    std::vector<int> perm;     // position i holds original row perm[i]
    std::vector<double> Lb;    // Lb[i*(bw+1) + d] = L[i][i-d] (reordered)
};

// Reverse Cuthill-McKee ordering of a symmetric pattern (all components).
inline std::vector<int> rcm_order(int n, const std::vector<int>& rowptr, const std::vector<int>& col) {
    std::vector<int> deg(n);
    for (int i = 0; i < n; ++i) deg[i] = rowptr[i + 1] - rowptr[i];
    std::vector<int> order;
    order.reserve(n);
    std::vector<char> seen(n, 0), vis(n, 0);
    std::vector<int> level(n, -1), touched;
    auto bfs = [&](int s, std::vector<int>& out) {
        for (int u : touched) { vis[u] = 0; level[u] = -1; }
        touched.clear();
        out.clear();
        std::queue<int> qu;
        qu.push(s);
        vis[s] = 1;
        level[s] = 0;
        touched.push_back(s);
        std::vector<int> nb;
        while (!qu.empty()) {
            const int u = qu.front();
            qu.pop();
            out.push_back(u);
            nb.clear();
            for (int k = rowptr[u]; k < rowptr[u + 1]; ++k) {
                const int w = col[k];
                if (!vis[w] && !seen[w]) { vis[w] = 1; touched.push_back(w); nb.push_back(w); }
            }
            std::sort(nb.begin(), nb.end(), [&](int a, int b) { return deg[a] != deg[b] ? deg[a] < deg[b] : a < b; });
            for (int w : nb) { level[w] = level[u] + 1; qu.push(w); }
        }
    };
    std::vector<int> comp;
    for (int s0 = 0; s0 < n; ++s0) {
        if (seen[s0]) continue;
        int s = s0;  // pseudo-peripheral start: two BFS sweeps
        for (int rep = 0; rep < 2; ++rep) {
            bfs(s, comp);
            int best = s;
            for (int u : comp)
                if (level[u] > level[best] || (level[u] == level[best] && deg[u] < deg[best])) best = u;
            s = best;
        }
        bfs(s, comp);
        for (int u : comp) { seen[u] = 1; order.push_back(u); }
    }
    std::reverse(order.begin(), order.end());
    return order;
}

inline int bandwidth_of(int n, const std::vector<int>& rowptr, const std::vector<int>& col,
                        const std::vector<int>& pos) {
    int bw = 0;
    for (int i = 0; i < n; ++i)
        for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) bw = std::max(bw, std::abs(pos[i] - pos[col[k]]));
    return bw;
}

// Half bandwidth after the ordering band_factor would choose (the narrower of
// natural and RCM order), without factoring.
inline int band_width_estimate(int n, const std::vector<int>& rowptr, const std::vector<int>& col) {
    std::vector<int> natural(n);
    for (int i = 0; i < n; ++i) natural[i] = i;
    std::vector<int> rcm = rcm_order(n, rowptr, col), pos_rcm(n);
    for (int i = 0; i < n; ++i) pos_rcm[rcm[i]] = i;
    return std::min(bandwidth_of(n, rowptr, col, natural), bandwidth_of(n, rowptr, col, pos_rcm));
}

// Independent systems of the reordered matrix: position p starts a new one
// when no row before p couples to a row at or after p (the factor is then
// block diagonal there).  Returns the first position of each system and n.
inline std::vector<int> band_systems(int n, const std::vector<int>& rowptr, const std::vector<int>& col,
                                     const std::vector<int>& perm) {
    std::vector<int> pos(n);
    for (int i = 0; i < n; ++i) pos[perm[i]] = i;
    std::vector<int> sys = {0};
    int reach = -1;
    for (int p = 0; p < n; ++p) {
        if (p > 0 && reach < p) sys.push_back(p);
        const int v = perm[p];
        for (int k = rowptr[v]; k < rowptr[v + 1]; ++k) reach = std::max(reach, pos[col[k]]);
    }
    sys.push_back(n);
    return sys;
}

// A = L L^T for a symmetric positive definite CSR matrix (both triangles,
// any column order), in double, after the narrower of natural / RCM order.
// Throws std::invalid_argument for malformed input and std::domain_error
// when the matrix is not positive definite.
template <class V>
BandFactor band_factor(int n, const std::vector<int>& rowptr, const std::vector<int>& col, const std::vector<V>& val) {
    if (n <= 0) throw std::invalid_argument("band_factor: empty matrix");
    if (rowptr.size() != (size_t)n + 1 || rowptr[0] != 0 || col.size() != (size_t)rowptr[n] ||
        val.size() != col.size())
        throw std::invalid_argument("band_factor: inconsistent CSR input");
    for (int i = 0; i < n; ++i)
        if (rowptr[i + 1] < rowptr[i]) throw std::invalid_argument("band_factor: rowptr not monotone");
    for (int c : col)
        if (c < 0 || c >= n) throw std::invalid_argument("band_factor: column index out of range");
    BandFactor f;
    f.n = n;
    std::vector<int> natural(n), pos_nat(n);
    for (int i = 0; i < n; ++i) natural[i] = pos_nat[i] = i;
    std::vector<int> rcm = rcm_order(n, rowptr, col), pos_rcm(n);
    for (int i = 0; i < n; ++i) pos_rcm[rcm[i]] = i;
    const int bw_nat = bandwidth_of(n, rowptr, col, pos_nat), bw_rcm = bandwidth_of(n, rowptr, col, pos_rcm);
    const bool use_rcm = bw_rcm < bw_nat;
    f.perm = use_rcm ? rcm : natural;
    const std::vector<int>& pos = use_rcm ? pos_rcm : pos_nat;
    const int bw = use_rcm ? bw_rcm : bw_nat;
    f.bw = bw;
    f.Lb.assign((size_t)n * (bw + 1), 0.0);
    std::vector<double>& Lb = f.Lb;
    for (int i = 0; i < n; ++i)
        for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
            const int pi = pos[i], pj = pos[col[k]];
            if (pj <= pi) Lb[(size_t)pi * (bw + 1) + (pi - pj)] = (double)val[k];
        }
    for (int i = 0; i < n; ++i) {
        double* Li = &Lb[(size_t)i * (bw + 1)];
        for (int j = std::max(0, i - bw); j <= i; ++j) {
            const double* Lj = &Lb[(size_t)j * (bw + 1)];
            double s = Li[i - j];
            for (int k = std::max(0, i - bw); k < j; ++k) s -= Li[i - k] * Lj[j - k];
            if (j < i) {
                Li[i - j] = s / Lj[0];
            } else {
                if (!(s > 0.0))
                    throw std::domain_error("Cholesky factorisation failed: matrix not positive definite at row " +
                                            std::to_string(i));
                Li[0] = std::sqrt(s);
            }
        }
    }
    return f;
}

}  // namespace sbs
