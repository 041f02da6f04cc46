// TetGen loading for one soft body (see scene_loader.hpp for the reference
// functions restated here).
#include "scene_loader.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace sbs {
namespace {

// Lines of a TetGen file with comments ('#' to the end of the line) and blank
// lines dropped; the reference reads raw lines, which is the same for the
// files TetGen writes (comments only after the data).
std::vector<std::string> data_lines(const std::string& path) {
    std::ifstream f(path);
    if (!f.is_open()) throw std::invalid_argument("unable to open file: " + path);
    std::vector<std::string> out;
    std::string line;
    while (std::getline(f, line)) {
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line.resize(hash);
        if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
        out.push_back(line);
    }
    return out;
}

int header_count(const std::vector<std::string>& lines, const std::string& path) {
    if (lines.empty()) throw std::invalid_argument("empty file: " + path);
    std::istringstream is(lines[0]);
    long long n = -1;
    if (!(is >> n) || n < 0 || n > (1LL << 31) - 1) throw std::invalid_argument("bad count line in " + path);
    if ((size_t)n + 1 > lines.size()) throw std::invalid_argument("fewer records than declared in " + path);
    return (int)n;
}

// `k` integers after the record's own index, minus start_index
void read_indices(const std::string& line, int k, int start, const std::string& path, uint32_t* out) {
    std::istringstream is(line);
    long long idx, v;
    if (!(is >> idx)) throw std::invalid_argument("malformed record in " + path + ": " + line);
    for (int j = 0; j < k; ++j) {
        if (!(is >> v)) throw std::invalid_argument("malformed record in " + path + ": " + line);
        v -= start;
        if (v < 0 || v > 0xFFFFFFFFLL) throw std::invalid_argument("vertex index out of range in " + path);
        out[j] = (uint32_t)v;
    }
}

// glm::rotate(m, radians(deg), axis) for a unit coordinate axis, float
void rotate_axis(float m[16], float deg, int axis) {
    const float a = deg * 0.01745329251994329576923690768489f;
    const float c = std::cos(a), s = std::sin(a);
    float ax[3] = {0.f, 0.f, 0.f};
    ax[axis] = 1.f;
    const float t[3] = {(1.f - c) * ax[0], (1.f - c) * ax[1], (1.f - c) * ax[2]};
    float R[3][3];  // R[col][row]
    R[0][0] = c + t[0] * ax[0];
    R[0][1] = 0.f + t[0] * ax[1] + s * ax[2];
    R[0][2] = 0.f + t[0] * ax[2] - s * ax[1];
    R[1][0] = 0.f + t[1] * ax[0] - s * ax[2];
    R[1][1] = c + t[1] * ax[1];
    R[1][2] = 0.f + t[1] * ax[2] + s * ax[0];
    R[2][0] = 0.f + t[2] * ax[0] + s * ax[1];
    R[2][1] = 0.f + t[2] * ax[1] - s * ax[0];
    R[2][2] = c + t[2] * ax[2];
    float out[12];
    for (int col = 0; col < 3; ++col)
        for (int r = 0; r < 4; ++r)
            out[4 * col + r] = m[r] * R[col][0] + m[4 + r] * R[col][1] + m[8 + r] * R[col][2];
    std::copy(out, out + 12, m);
}

template <class S> void transform(std::vector<double>& X, const float m[16]) {
    const size_t n = X.size() / 3;
    for (size_t i = 0; i < n; ++i) {
        const S v[3] = {(S)X[3 * i], (S)X[3 * i + 1], (S)X[3 * i + 2]};
        for (int r = 0; r < 3; ++r) {
            // (m0 x + m1 y) + (m2 z + m3 * 1), the column order of glm's mat4 * vec4
            const S a0 = (S)m[r] * v[0] + (S)m[4 + r] * v[1];
            const S a1 = (S)m[8 + r] * v[2] + (S)m[12 + r] * (S)1;
            X[3 * i + r] = (double)(a0 + a1);
        }
    }
}

template <class S> void centre(std::vector<double>& X) {
    const size_t n = X.size() / 3;
    S c[3] = {0, 0, 0};
    for (size_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) c[k] += (S)X[3 * i + k];
    for (int k = 0; k < 3; ++k) c[k] /= (S)(float)n;
    for (size_t i = 0; i < n; ++i) {
        S p[3];
        for (int k = 0; k < 3; ++k) p[k] = (S)X[3 * i + k] - c[k];
        const float tmp = (float)p[1];  // the swap goes through a float (dataLoader.cu:163-165)
        X[3 * i] = (double)p[0];
        X[3 * i + 1] = (double)p[2];
        X[3 * i + 2] = (double)(S)tmp;
    }
}

}  // namespace

void body_model_matrix(const float pos[3], const float scale[3], const float rot[3], float m[16]) {
    for (int i = 0; i < 16; ++i) m[i] = (i % 5 == 0) ? 1.f : 0.f;
    // translate: column 3 = m0 p0 + m1 p1 + m2 p2 + m3 (identity: (p, 1))
    for (int r = 0; r < 4; ++r) m[12 + r] = m[r] * pos[0] + m[4 + r] * pos[1] + m[8 + r] * pos[2] + m[12 + r];
    for (int col = 0; col < 3; ++col)
        for (int r = 0; r < 4; ++r) m[4 * col + r] *= scale[col];
    rotate_axis(m, rot[0], 0);
    rotate_axis(m, rot[1], 1);
    rotate_axis(m, rot[2], 2);
}

std::vector<uint32_t> boundary_faces(const std::vector<uint32_t>& tet) {
    struct Face {
        std::array<uint32_t, 3> key;   // sorted
        std::array<uint32_t, 3> wind;  // as the tet lists it
        size_t order;
    };
    const size_t nt = tet.size() / 4;
    std::vector<Face> f;
    f.reserve(4 * nt);
    static const int kFace[4][3] = {{0, 1, 2}, {0, 2, 3}, {0, 3, 1}, {1, 3, 2}};
    for (size_t t = 0; t < nt; ++t)
        for (int k = 0; k < 4; ++k) {
            Face x;
            for (int j = 0; j < 3; ++j) x.wind[j] = tet[4 * t + kFace[k][j]];
            x.key = x.wind;
            std::sort(x.key.begin(), x.key.end());
            x.order = 4 * t + k;
            f.push_back(x);
        }
    std::sort(f.begin(), f.end(), [](const Face& a, const Face& b) {
        return a.key != b.key ? a.key < b.key : a.order < b.order;
    });
    std::vector<uint32_t> out;
    for (size_t i = 0; i < f.size();) {
        size_t j = i;
        while (j < f.size() && f[j].key == f[i].key) ++j;
        if ((j - i) % 2 == 1) out.insert(out.end(), f[j - 1].wind.begin(), f[j - 1].wind.end());
        i = j;
    }
    return out;
}

std::vector<uint32_t> surface_edges(const std::vector<uint32_t>& tri) {
    std::vector<std::pair<uint32_t, uint32_t>> e;
    e.reserve(tri.size());
    for (size_t i = 0; i + 2 < tri.size(); i += 3)
        for (int k = 0; k < 3; ++k) {
            const uint32_t a = tri[i + k], b = tri[i + (k + 1) % 3];
            e.emplace_back(std::min(a, b), std::max(a, b));
        }
    std::sort(e.begin(), e.end());
    e.erase(std::unique(e.begin(), e.end()), e.end());
    std::vector<uint32_t> out;
    out.reserve(2 * e.size());
    for (auto& p : e) {
        out.push_back(p.first);
        out.push_back(p.second);
    }
    return out;
}

BodyMesh load_tetgen(const std::string& node_file, const std::string& ele_file, const std::string& face_file,
                     const TetgenOptions& opt) {
    if (opt.precision != 32 && opt.precision != 64) throw std::invalid_argument("precision must be 32 or 64");
    BodyMesh b;
    // .node: "<n> <dim> <attrs> <marker>", then "<index> x y z ...", read as float
    const auto nl = data_lines(node_file);
    const int nv = header_count(nl, node_file);
    if (nv == 0) throw std::invalid_argument("no vertices in " + node_file);
    b.X.resize(3 * (size_t)nv);
    for (int i = 0; i < nv; ++i) {
        std::istringstream is(nl[1 + i]);
        long long idx;
        float x, y, z;
        if (!(is >> idx >> x >> y >> z)) throw std::invalid_argument("malformed vertex in " + node_file);
        b.X[3 * (size_t)i] = x;
        b.X[3 * (size_t)i + 1] = y;
        b.X[3 * (size_t)i + 2] = z;
    }
    if (opt.centralize) {
        if (opt.precision == 64) centre<double>(b.X);
        else centre<float>(b.X);
    }
    // .ele: "<n> 4 <attrs>", then "<index> a b c d"
    const auto el = data_lines(ele_file);
    const int nt = header_count(el, ele_file);
    b.tet.resize(4 * (size_t)nt);
    for (int t = 0; t < nt; ++t) read_indices(el[1 + t], 4, opt.start_index, ele_file, &b.tet[4 * (size_t)t]);
    for (uint32_t v : b.tet)
        if (v >= (uint32_t)nv) throw std::invalid_argument("tet vertex index out of range in " + ele_file);
    // .face: "<n> <marker>", then "<index> a b c [marker]"; no file, one that
    // does not open, or zero faces: extract (dataLoader.cu:71-91)
    if (!face_file.empty() && std::ifstream(face_file).is_open()) {
        const auto fl = data_lines(face_file);
        const int nf = header_count(fl, face_file);
        b.tri.resize(3 * (size_t)nf);
        for (int t = 0; t < nf; ++t) read_indices(fl[1 + t], 3, opt.start_index, face_file, &b.tri[3 * (size_t)t]);
        for (uint32_t v : b.tri)
            if (v >= (uint32_t)nv) throw std::invalid_argument("face vertex index out of range in " + face_file);
        b.tri_from_file = nf > 0;
    }
    if (b.tri.empty()) b.tri = boundary_faces(b.tet);
    b.edges = surface_edges(b.tri);
    float m[16];
    body_model_matrix(opt.pos, opt.scale, opt.rot, m);
    if (opt.precision == 64) transform<double>(b.X, m);
    else transform<float>(b.X, m);
    return b;
}

}  // namespace sbs
