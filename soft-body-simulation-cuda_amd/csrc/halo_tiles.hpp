// Host construction of the halo tiles of the PCG kernels: the rows of the
// scalar matrix grouped into workgroups along a Morton curve of the
// vertices' rest positions, each workgroup with the list of the rows its rows
// touch (themselves and their neighbours: the "halo", staged into LDS once
// per launch) and its rows' entries in ELL form with 16-bit halo-local column
// indices.  Rows are renumbered: internal row i is vertex row_perm[i], the
// rows of a workgroup are consecutive internal rows (per-row vectors kept in
// internal order are read and written contiguously), and inside a workgroup
// they ascend by vertex id.
//
// Plain C++17 (no HIP) so the CPU test suite compiles it alone to check it.
#pragma once

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <vector>

namespace sbs {

// 30-bit Morton code of a point in the box [lo, hi] (10 bits per axis).
inline uint32_t morton_code(const double p[3], const double lo[3], const double hi[3]) {
    auto spread = [](uint32_t x) {  // 10 bits -> every third bit
        x = (x * 0x00010001u) & 0xFF0000FFu;
        x = (x * 0x00000101u) & 0x0F00F00Fu;
        x = (x * 0x00000011u) & 0xC30C30C3u;
        x = (x * 0x00000005u) & 0x49249249u;
        return x;
    };
    uint32_t cell[3];
    for (int k = 0; k < 3; ++k) {
        const double ext = hi[k] - lo[k];
        double u = ext > 0 ? (p[k] - lo[k]) / ext : 0.0;
        u = std::min(std::max(u * 1024.0, 0.0), 1023.0);
        cell[k] = (uint32_t)u;
    }
    return (spread(cell[0]) << 2) | (spread(cell[1]) << 1) | spread(cell[2]);
}

struct HaloTiles {
    int rows_per_group = 256;
    int halo_stride = 0;                 // halo entries reserved per workgroup (the LDS capacity)
    int max_halo = 0;                    // largest halo built
    int max_width = 0;                   // longest row (entries, diagonal included)
    std::vector<uint32_t> row_perm;      // internal row -> vertex
    std::vector<uint32_t> row_of;        // vertex -> internal row
    // per workgroup
    std::vector<int> g_row0, g_rows, g_halo_n, g_width, g_eptr;
    std::vector<int> g_own0;             // halo-local index of the workgroup's first row (its rows follow it)
    // halo of workgroup g at [g * halo_stride, g * halo_stride + g_halo_n[g]),
    // ascending internal rows; unused positions repeat the workgroup's first row
    std::vector<uint32_t> halo;          // internal rows
    std::vector<uint32_t> halo_vertex;   // the same entries as vertex ids
    // entry e of the workgroup's thread t at g_eptr[g] + e * rows_per_group + t
    std::vector<uint16_t> col;           // halo-local column
    std::vector<int64_t> src;            // caller's value index of the entry; -1: padding (value 0)
};

// `rowstart`/`cols`: the pattern of each vertex's row (CSR over vertex ids,
// diagonal included or not, as the caller's product needs); `value_index(k)`
// the caller's value index of CSR position k.  `key(v)` orders the vertices
// (a Morton code of the rest position).  A workgroup takes up to
// rows_per_group consecutive rows in that order and stops early when its halo
// would exceed halo_stride entries; a single row whose halo exceeds it is an
// error (std::invalid_argument).  The workgroup's rows are then numbered in
// ascending vertex id.
template <class Key, class ValueIndex>
HaloTiles build_halo_tiles(int nv, const std::vector<int>& rowstart, const std::vector<uint32_t>& cols, Key key,
                           ValueIndex value_index, int rows_per_group, int halo_stride) {
    if (nv <= 0) throw std::invalid_argument("halo tiles: no rows");
    if (rows_per_group <= 0 || rows_per_group % 64 != 0)
        throw std::invalid_argument("halo tiles: rows per group must be a positive multiple of 64");
    if (halo_stride <= 0 || halo_stride > 65536) throw std::invalid_argument("halo tiles: halo indices are 16-bit");
    if ((int)rowstart.size() != nv + 1 || (size_t)rowstart[nv] != cols.size())
        throw std::invalid_argument("halo tiles: malformed pattern");
    for (uint32_t c : cols)
        if (c >= (uint32_t)nv) throw std::invalid_argument("halo tiles: column out of range");
    HaloTiles h;
    h.rows_per_group = rows_per_group;
    h.halo_stride = halo_stride;
    h.row_perm.resize(nv);
    h.row_of.resize(nv);
    {
        std::vector<std::pair<uint64_t, uint32_t>> order(nv);
        for (int v = 0; v < nv; ++v) order[v] = {(uint64_t)key(v), (uint32_t)v};
        std::sort(order.begin(), order.end());
        for (int i = 0; i < nv; ++i) {
            h.row_perm[i] = order[i].second;
            h.row_of[order[i].second] = (uint32_t)i;
        }
    }
    // pass 1: workgroup boundaries, growing each workgroup row by row (in
    // Morton order) while its halo fits
    std::vector<int> mark(nv, -1);  // workgroup that already holds the row (Morton position)
    int nset = 0;
    for (int r = 0, g = 0; r < nv; ++g) {
        nset = 0;
        int n = 0;
        while (n < rows_per_group && r + n < nv) {
            const uint32_t i = (uint32_t)(r + n), v = h.row_perm[i];
            int fresh = mark[i] != g;
            for (int k = rowstart[v]; k < rowstart[v + 1]; ++k) {
                const uint32_t c = h.row_of[cols[k]];
                if (mark[c] != g && c != i) {
                    mark[c] = -2 - g;  // counted once even if the row repeats it
                    ++fresh;
                }
            }
            for (int k = rowstart[v]; k < rowstart[v + 1]; ++k) {
                const uint32_t c = h.row_of[cols[k]];
                if (mark[c] == -2 - g) mark[c] = -1;
            }
            if (nset + fresh > halo_stride) {
                if (n == 0) throw std::invalid_argument("halo tiles: one row touches more rows than the halo holds");
                break;
            }
            if (mark[i] != g) { mark[i] = g; ++nset; }
            for (int k = rowstart[v]; k < rowstart[v + 1]; ++k) {
                const uint32_t c = h.row_of[cols[k]];
                if (mark[c] != g) { mark[c] = g; ++nset; }
            }
            ++n;
        }
        h.g_row0.push_back(r);
        h.g_rows.push_back(n);
        r += n;
    }
    // inside a workgroup, rows ascend by vertex id: the workgroup's reads of
    // arrays in the caller's vertex order become runs of consecutive vertices
    const int ng = (int)h.g_row0.size();
    for (int g = 0; g < ng; ++g) {
        std::sort(h.row_perm.begin() + h.g_row0[g], h.row_perm.begin() + h.g_row0[g] + h.g_rows[g]);
        for (int i = h.g_row0[g]; i < h.g_row0[g] + h.g_rows[g]; ++i) h.row_of[h.row_perm[i]] = (uint32_t)i;
    }
    // pass 2: halos (the same row sets, in the final numbering), columns, values
    std::fill(mark.begin(), mark.end(), -1);
    std::vector<uint32_t> set;
    for (int g = 0; g < ng; ++g) {
        const int r = h.g_row0[g], n = h.g_rows[g];
        set.clear();
        int width = 0;
        for (int i = r; i < r + n; ++i) {
            const uint32_t v = h.row_perm[i];
            if (mark[i] != g) { mark[i] = g; set.push_back((uint32_t)i); }
            for (int k = rowstart[v]; k < rowstart[v + 1]; ++k) {
                const uint32_t c = h.row_of[cols[k]];
                if (mark[c] != g) { mark[c] = g; set.push_back(c); }
            }
            width = std::max(width, rowstart[v + 1] - rowstart[v]);
        }
        std::sort(set.begin(), set.end());
        h.g_halo_n.push_back((int)set.size());
        h.g_width.push_back(width);
        h.g_eptr.push_back((int)h.col.size());
        h.max_halo = std::max(h.max_halo, (int)set.size());
        h.max_width = std::max(h.max_width, width);
        const size_t h0 = h.halo.size();
        h.halo.resize(h0 + halo_stride, (uint32_t)r);
        std::copy(set.begin(), set.end(), h.halo.begin() + h0);
        auto local = [&](uint32_t i) {
            return (uint16_t)(std::lower_bound(set.begin(), set.end(), i) - set.begin());
        };
        h.g_own0.push_back(local((uint32_t)r));  // rows r..r+n-1 are consecutive in the sorted halo
        const size_t base = h.col.size();
        if (base + (size_t)width * rows_per_group > (size_t)INT32_MAX)
            throw std::invalid_argument("halo tiles: too many entries");
        h.col.resize(base + (size_t)width * rows_per_group);
        h.src.resize(base + (size_t)width * rows_per_group, -1);
        for (int t = 0; t < rows_per_group; ++t) {
            const bool row = t < n;
            const uint32_t i = (uint32_t)(row ? r + t : r);
            const uint32_t v = h.row_perm[i];
            const int len = row ? rowstart[v + 1] - rowstart[v] : 0;
            for (int e = 0; e < width; ++e) {
                const size_t pos = base + (size_t)e * rows_per_group + t;
                if (e < len) {
                    h.col[pos] = local(h.row_of[cols[rowstart[v] + e]]);
                    h.src[pos] = (int64_t)value_index(rowstart[v] + e);
                } else {  // padding: the row's own column, value 0
                    h.col[pos] = local(i);
                    h.src[pos] = -1;
                }
            }
        }
    }
    h.halo_vertex.resize(h.halo.size());
    for (size_t k = 0; k < h.halo.size(); ++k) h.halo_vertex[k] = h.row_perm[h.halo[k]];
    return h;
}

}  // namespace sbs
