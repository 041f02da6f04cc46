// CPU oracle for the collision path of the PD step (broad phase, CCD narrow
// phase, hit resolution and the CCD response).
//
// TEST INFRASTRUCTURE ONLY.  Plain C++17, used by tests/ as the checker; never
// linked into, loaded by, or called from the product library.  Built with
// -ffp-contract=off so its float arithmetic is the IEEE sequence written here
// (the GPU kernels are compiled to the same sequence), which makes pair lists,
// times of impact and hit sets comparable bit for bit.
//
// What it restates (reference file:line):
//   * leaf boxes of a triangle's trajectory   src/collision/ccd.cu:53-65
//       (fminf/fmaxf: coordinates pass through float, then -/+ 0.01)
//   * box overlap test (touching counts)       src/collision/intersections.cu:225-231
//   * pair filters and the 12 queries per pair src/collision/broadphase.cu:256-319,331-398
//       (same triangle, shared vertex, same body when ignoreSelfCollision)
//   * canonical query form, sort, unique       src/collision/broadphase.cu:450-484
//   * CCD: cubic of the coplanarity condition  src/collision/intersections.cu:286-337
//       with solveCubic / solveQuadratic / newtonsMethod :44-113 and the signed
//       distances :114-165
//   * hit resolution (min toi per vertex / edge) src/collision/narrowphase.cu:9-57,129-131
//   * storeTi                                   src/collision/narrowphase.cu:72-112
//   * CCDKernel response                        src/simulation/collisionUtil.cu:50-70
//   * ComputeMinStepSize (min toi of the CCD candidates) src/collision/bvh.cu:183-190
//   * UpdateQueries' candidates: boxes of X grown by 2 dhat src/collision/broadphase.cu:557-577,
//       computeTriTrajBBox src/collision/ccd.cu:74-85
//
// The broad phase here enumerates all ordered triangle pairs (O(n^2)); the
// reference finds the same pairs through its LBVH, whose internal boxes are
// exact unions of the leaf boxes, so the pair set does not depend on the tree.
//
// Decisions where the reference is undefined:
//   * equal times of impact inside a group keep the query that comes first in
//     (v1, v2, v3) order (the reference's thrust::sort is unstable there);
//   * a vertex touched by several hits takes the normal of the hit that comes
//     last in the sorted list (the reference's storeTi races);
//   * the weights of a degenerate signed distance are 0 (uninitialised in the
//     reference; they only feed a non-hit).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>

namespace {

template <class T> struct V3 {
    T x, y, z;
};
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
template <class T> inline T length(V3<T> a) { return std::sqrt(dot(a, a)); }
template <class T> inline V3<T> normalize(V3<T> a) { return (T(1) / std::sqrt(dot(a, a))) * a; }
template <class T> inline T stp(V3<T> u, V3<T> v, V3<T> w) { return dot(u, cross(v, w)); }
template <class T> inline T sgn(T v) { return v > T(0) ? T(1) : (v < T(0) ? T(-1) : T(0)); }

template <class T> inline V3<T> ld(const T* p, uint32_t i) { return {p[3 * (size_t)i], p[3 * (size_t)i + 1], p[3 * (size_t)i + 2]}; }

// ---------------------------------------------------------------- CCD --
template <class T> int solve_quadratic(T a, T b, T c, T* x) {
    const T d = b * b - T(4) * a * c;
    if (d < T(0)) {
        x[0] = -b / (T(2) * a);
        return 0;
    }
    const T q = -(b + sgn(b) * std::sqrt(d)) / T(2);
    int i = 0;
    if ((double)std::fabs(a) > 1e-12 * (double)std::fabs(q)) x[i++] = q / a;
    if ((double)std::fabs(q) > 1e-12 * (double)std::fabs(c)) x[i++] = c / q;
    if (i == 2 && x[0] > x[1]) std::swap(x[0], x[1]);
    return i;
}

template <class T> T newton(T a, T b, T c, T d, T x0, int init_dir) {
    if (init_dir != 0) {
        const T y0 = d + x0 * (c + x0 * (b + x0 * a));
        const T ddy0 = T(2) * b + x0 * (T(6) * a);
        if (ddy0 != T(0)) x0 += T(init_dir) * std::sqrt(std::fabs(T(2) * y0 / ddy0));
    }
    for (int it = 0; it < 100; ++it) {
        const T y = d + x0 * (c + x0 * (b + x0 * a));
        const T dy = c + x0 * (T(2) * b + x0 * T(3) * a);
        if (dy == T(0)) return x0;
        const T x1 = x0 - y / dy;
        if ((double)std::fabs(x0 - x1) < 1e-6) return x0;
        x0 = x1;
    }
    return x0;
}

template <class T> int solve_cubic(T a, T b, T c, T d, T* x) {
    T xc[2] = {T(0), T(0)};
    const int ncrit = solve_quadratic(T(3) * a, T(2) * b, c, xc);
    if (ncrit == 0) {
        x[0] = newton(a, b, c, d, xc[0], 0);
        return 1;
    }
    if (ncrit == 1) return solve_quadratic(b, c, d, x);
    const T yc0 = d + xc[0] * (c + xc[0] * (b + xc[0] * a));
    const T yc1 = d + xc[1] * (c + xc[1] * (b + xc[1] * a));
    int i = 0;
    if (yc0 * a >= T(0)) x[i++] = newton(a, b, c, d, xc[0], -1);
    if (yc0 * yc1 <= T(0)) {
        const int closer = std::fabs(yc0) < std::fabs(yc1) ? 0 : 1;
        x[i++] = newton(a, b, c, d, xc[closer], closer == 0 ? 1 : -1);
    }
    if (yc1 * a <= T(0)) x[i++] = newton(a, b, c, d, xc[1], 1);
    return i;
}

template <class T> T vf_distance(V3<T> x, V3<T> y0, V3<T> y1, V3<T> y2, V3<T>& n, T* w) {
    n = cross(normalize(y1 - y0), normalize(y2 - y0));
    if ((double)dot(n, n) < 1e-6) return (T)FLT_MAX;
    n = normalize(n);
    const T h = dot(x - y0, n);
    const T b0 = stp(y1 - x, y2 - x, n), b1 = stp(y2 - x, y0 - x, n), b2 = stp(y0 - x, y1 - x, n);
    const T s = b0 + b1 + b2;
    w[0] = T(1);
    w[1] = -b0 / s;
    w[2] = -b1 / s;
    w[3] = -b2 / s;
    return h;
}

template <class T> T ee_distance(V3<T> x0, V3<T> x1, V3<T> y0, V3<T> y1, V3<T>& n, T* w) {
    n = cross(normalize(x1 - x0), normalize(y1 - y0));
    if ((double)dot(n, n) < 1e-6) return (T)FLT_MAX;
    n = normalize(n);
    const T h = dot(x0 - y0, n);
    const T a0 = stp(y1 - x1, y0 - x1, n), a1 = stp(y0 - x0, y1 - x0, n);
    const T b0 = stp(x0 - y1, x1 - y1, n), b1 = stp(x1 - y0, x0 - y0, n);
    w[0] = a0 / (a0 + a1);
    w[1] = a1 / (a0 + a1);
    w[2] = -b0 / (b0 + b1);
    w[3] = -b1 / (b0 + b1);
    return h;
}

// type 1 = VF (v0 against triangle v1 v2 v3), 2 = EE (edge v0 v1 against v2 v3)
template <class T>
T ccd_test(int type, const uint32_t* v, const T* X, const T* XT, V3<T>& n) {
    const V3<T> x0 = ld(X, v[0]), x1 = ld(X, v[1]), x2 = ld(X, v[2]), x3 = ld(X, v[3]);
    const V3<T> u0 = ld(XT, v[0]) - x0, u1 = ld(XT, v[1]) - x1, u2 = ld(XT, v[2]) - x2, u3 = ld(XT, v[3]) - x3;
    const V3<T> x01 = x1 - x0, x02 = x2 - x0, x03 = x3 - x0;
    const V3<T> u01 = u1 - u0, u02 = u2 - u0, u03 = u3 - u0;
    const T a0 = stp(x01, x02, x03);
    const T a1 = stp(u01, x02, x03) + stp(x01, u02, x03) + stp(x01, x02, u03);
    const T a2 = stp(x01, u02, u03) + stp(u01, x02, u03) + stp(u01, u02, x03);
    const T a3 = stp(u01, u02, u03);
    n = {T(0), T(0), T(0)};
    if ((double)std::fabs(a0) < 1e-12 * (double)length(x01) * (double)length(x02) * (double)length(x03)) return T(1);
    T t[3] = {T(0), T(0), T(0)};
    const int nsol = solve_cubic(a3, a2, a1, a0, t);
    for (int i = 0; i < nsol; ++i) {
        if ((double)t[i] < -1e-12 || t[i] > T(1)) continue;
        const V3<T> xt0 = x0 + t[i] * u0, xt1 = x1 + t[i] * u1, xt2 = x2 + t[i] * u2, xt3 = x3 + t[i] * u3;
        T w[4] = {T(0), T(0), T(0), T(0)};
        T d;
        bool inside;
        if (type == 1) {
            d = vf_distance(xt0, xt1, xt2, xt3, n, w);
            inside = (double)std::min(-w[1], std::min(-w[2], -w[3])) >= -1e-3;
        } else {
            d = ee_distance(xt0, xt1, xt2, xt3, n, w);
            inside = (double)std::min(w[0], std::min(w[1], std::min(-w[2], -w[3]))) >= -1e-3;
        }
        if (dot(n, w[1] * u1 + w[2] * u2 + w[3] * u3) > T(0)) n = -n;
        if ((double)std::fabs(d) < 1e-6 && inside) return t[i];
    }
    return T(1);
}

// --------------------------------------------------------- broad phase --
template <class T> struct Box {
    T lo[3], hi[3];
};

template <class T> Box<T> leaf_box(const uint32_t* tri, const T* X, const T* XT) {
    Box<T> b;
    for (int k = 0; k < 3; ++k) {
        float mn = (float)X[3 * (size_t)tri[0] + k], mx = mn;
        for (int j = 0; j < 3; ++j) {
            const float a = (float)X[3 * (size_t)tri[j] + k], c = (float)XT[3 * (size_t)tri[j] + k];
            mn = std::fmin(std::fmin(mn, a), c);
            mx = std::fmax(std::fmax(mx, a), c);
        }
        b.lo[k] = (T)mn - (T)0.01;
        b.hi[k] = (T)mx + (T)0.01;
    }
    return b;
}

// Leaf box of a triangle at rest, grown by `bound` (computeTriTrajBBox,
// ccd.cu:74-85; the coordinates pass through float there as well).
template <class T> Box<T> leaf_box_bound(const uint32_t* tri, const T* X, T bound) {
    Box<T> b;
    for (int k = 0; k < 3; ++k) {
        float mn = (float)X[3 * (size_t)tri[0] + k], mx = mn;
        for (int j = 1; j < 3; ++j) {
            const float a = (float)X[3 * (size_t)tri[j] + k];
            mn = std::fmin(mn, a);
            mx = std::fmax(mx, a);
        }
        b.lo[k] = (T)mn - bound;
        b.hi[k] = (T)mx + bound;
    }
    return b;
}

template <class T> bool overlap(const Box<T>& a, const Box<T>& b) {
    for (int k = 0; k < 3; ++k)
        if (a.hi[k] < b.lo[k] || a.lo[k] > b.hi[k]) return false;
    return true;
}

struct Q {
    uint32_t type, v[4];
};
inline bool q_less(const Q& a, const Q& b) {
    if (a.type != b.type) return a.type < b.type;
    for (int k = 0; k < 4; ++k)
        if (a.v[k] != b.v[k]) return a.v[k] < b.v[k];
    return false;
}
inline bool q_eq(const Q& a, const Q& b) { return !q_less(a, b) && !q_less(b, a); }

// The 12 queries of the ordered pair (A, B) in canonical form; type 0 = dropped.
void pair_queries(const uint32_t* a, const uint32_t* b, Q* q) {
    static const int ed[6] = {0, 1, 0, 2, 1, 2};
    for (int i = 0; i < 3; ++i) {
        q[i] = {1, {a[i], b[0], b[1], b[2]}};
    }
    for (int i = 0; i < 3; ++i) {
        const uint32_t e0 = a[ed[2 * i]], e1 = a[ed[2 * i + 1]];
        q[3 + 3 * i] = {2, {e0, e1, b[0], b[1]}};
        q[4 + 3 * i] = {2, {e0, e1, b[0], b[2]}};
        q[5 + 3 * i] = {2, {e0, e1, b[1], b[2]}};
    }
    for (int k = 0; k < 12; ++k) {
        uint32_t* v = q[k].v;
        if (q[k].type == 1) {
            if (v[0] == v[1] || v[0] == v[2] || v[0] == v[3]) { q[k].type = 0; continue; }
            if (v[1] > v[2]) std::swap(v[1], v[2]);
            if (v[2] > v[3]) std::swap(v[2], v[3]);
            if (v[1] > v[2]) std::swap(v[1], v[2]);
        } else {
            if (v[0] > v[1]) std::swap(v[0], v[1]);
            if (v[2] > v[3]) std::swap(v[2], v[3]);
            if (v[0] == v[2] && v[1] == v[3]) q[k].type = 0;
        }
    }
}

template <class T>
std::vector<Q> broad_phase_boxes(int ntri, const uint32_t* tri, const uint32_t* fathers, const std::vector<Box<T>>& box,
                                 int ignore_self) {
    std::vector<Q> out;
    Q q[12];
    for (int i = 0; i < ntri; ++i) {
        const uint32_t* a = tri + 3 * (size_t)i;
        for (int j = 0; j < ntri; ++j) {
            if (i == j || !overlap(box[i], box[j])) continue;
            const uint32_t* b = tri + 3 * (size_t)j;
            if (ignore_self && fathers[i] == fathers[j]) continue;
            bool adj = false;
            for (int s = 0; s < 3; ++s)
                for (int t = 0; t < 3; ++t) adj |= a[s] == b[t];
            if (adj) continue;
            pair_queries(a, b, q);
            for (int k = 0; k < 12; ++k)
                if (q[k].type) out.push_back(q[k]);
        }
    }
    std::sort(out.begin(), out.end(), q_less);
    out.erase(std::unique(out.begin(), out.end(), q_eq), out.end());
    return out;
}

// Candidates of DetectCollision / ComputeMinStepSize: trajectory boxes X -> XT.
template <class T>
std::vector<Q> broad_phase(int ntri, const uint32_t* tri, const uint32_t* fathers, const T* X, const T* XT,
                           int ignore_self) {
    std::vector<Box<T>> box(ntri);
    for (int i = 0; i < ntri; ++i) box[i] = leaf_box(tri + 3 * (size_t)i, X, XT);
    return broad_phase_boxes(ntri, tri, fathers, box, ignore_self);
}

// Candidates of UpdateQueries (broadphase.cu:574-577): boxes of X grown by `bound`.
template <class T>
std::vector<Q> broad_phase_rest(int ntri, const uint32_t* tri, const uint32_t* fathers, const T* X, T bound,
                                int ignore_self) {
    std::vector<Box<T>> box(ntri);
    for (int i = 0; i < ntri; ++i) box[i] = leaf_box_bound(tri + 3 * (size_t)i, X, bound);
    return broad_phase_boxes(ntri, tri, fathers, box, ignore_self);
}

// Hit resolution + storeTi on the sorted query list; returns the number of
// winning queries with toi < 1.
template <class T>
int resolve_hits(const std::vector<Q>& qs, const std::vector<T>& toi, const std::vector<V3<T>>& nrm, int nv,
                 T* tI, float* normals) {
    for (int v = 0; v < nv; ++v) tI[v] = T(1);
    int hits = 0;
    size_t g = 0;
    while (g < qs.size()) {
        // group: same type and v0 (VF) or v0, v1 (EE); contiguous in sorted order
        size_t e = g + 1;
        while (e < qs.size() && qs[e].type == qs[g].type && qs[e].v[0] == qs[g].v[0] &&
               (qs[g].type == 1 || qs[e].v[1] == qs[g].v[1]))
            ++e;
        size_t best = g;
        for (size_t k = g + 1; k < e; ++k)
            if (toi[k] < toi[best]) best = k;
        if (toi[best] < T(1)) {
            ++hits;
            const Q& q = qs[best];
            const V3<T> n = nrm[best];
            const float nx = (float)n.x, ny = (float)n.y, nz = (float)n.z;
            auto put = [&](uint32_t v, float s) {
                tI[v] = (T)0.5;
                normals[3 * (size_t)v] = s * nx;
                normals[3 * (size_t)v + 1] = s * ny;
                normals[3 * (size_t)v + 2] = s * nz;
            };
            if (q.type == 2) {
                put(q.v[0], 1.f);
                put(q.v[1], 1.f);
            } else {
                put(q.v[0], 1.f);
                put(q.v[1], -1.f);
                put(q.v[2], -1.f);
                put(q.v[3], -1.f);
            }
        }
        g = e;
    }
    return hits;
}

template <class T> struct Detector {
    std::vector<Q> qs;
    std::vector<T> toi;
    std::vector<V3<T>> nrm;
    int hits = 0;
};

template <class T>
void* detect(int nv, int ntri, const uint32_t* tri, const uint32_t* fathers, const T* X, const T* XT,
             int ignore_self, T* tI, float* normals) {
    auto* d = new Detector<T>;
    d->qs = broad_phase(ntri, tri, fathers, X, XT, ignore_self);
    d->toi.resize(d->qs.size());
    d->nrm.resize(d->qs.size());
    for (size_t k = 0; k < d->qs.size(); ++k)
        d->toi[k] = ccd_test<T>((int)d->qs[k].type, d->qs[k].v, X, XT, d->nrm[k]);
    if (tI) d->hits = resolve_hits(d->qs, d->toi, d->nrm, nv, tI, normals);
    return d;
}

template <class T>
void ccd_response(int nv, T* X, const T* XT, T* V, const T* tI, const float* normals) {
    for (int v = 0; v < nv; ++v) {
        const V3<T> x = ld(X, (uint32_t)v), xt = ld(XT, (uint32_t)v);
        if (tI[v] < T(1)) {
            const V3<T> n = {(T)normals[3 * (size_t)v], (T)normals[3 * (size_t)v + 1], (T)normals[3 * (size_t)v + 2]};
            const V3<T> vel = xt - x;
            const V3<T> vn = dot(vel, n) * n;
            V[3 * (size_t)v] = -vn.x;
            V[3 * (size_t)v + 1] = -vn.y;
            V[3 * (size_t)v + 2] = -vn.z;
        } else {
            X[3 * (size_t)v] = xt.x;
            X[3 * (size_t)v + 1] = xt.y;
            X[3 * (size_t)v + 2] = xt.z;
        }
    }
}

}  // namespace

extern "C" {

// precision 32 / 64; X, XT are 3*nv scalars; tI (nv) and normals (3*nv) may be
// null for a query-only run.  Returns a handle holding the sorted query list.
void* orc_col_detect(int precision, int nv, int ntri, const uint32_t* tri, const uint32_t* fathers, const void* X,
                     const void* XT, int ignore_self, void* tI, float* normals) {
    if (precision == 64)
        return detect<double>(nv, ntri, tri, fathers, (const double*)X, (const double*)XT, ignore_self,
                              (double*)tI, normals);
    return detect<float>(nv, ntri, tri, fathers, (const float*)X, (const float*)XT, ignore_self, (float*)tI,
                         normals);
}

// ComputeMinStepSize (bvh.cu:183-190): the smallest toi over the CCD
// candidates of X -> XT (1 when there are none).
double orc_col_min_step(int precision, int nv, int ntri, const uint32_t* tri, const uint32_t* fathers, const void* X,
                        const void* XT, int ignore_self) {
    auto run = [&](auto* x, auto* xt) {
        using T = std::remove_const_t<std::remove_pointer_t<decltype(x)>>;
        const std::vector<Q> qs = broad_phase<T>(ntri, tri, fathers, x, xt, ignore_self);
        T best = T(1);
        for (const Q& q : qs) {
            V3<T> n;
            best = std::min(best, ccd_test<T>((int)q.type, q.v, x, xt, n));
        }
        return (double)best;
    };
    (void)nv;
    if (precision == 64) return run((const double*)X, (const double*)XT);
    return run((const float*)X, (const float*)XT);
}

// Candidate queries of UpdateQueries: boxes of X grown by `bound`; returns
// the count and, when `q` is not null, 5 u32 per query (type, v0..v3).
long long orc_col_proximity(int ntri, const uint32_t* tri, const uint32_t* fathers, const double* X, double bound,
                            int ignore_self, uint32_t* q) {
    const std::vector<Q> qs = broad_phase_rest<double>(ntri, tri, fathers, X, bound, ignore_self);
    if (q)
        for (size_t k = 0; k < qs.size(); ++k) {
            q[5 * k] = qs[k].type;
            for (int j = 0; j < 4; ++j) q[5 * k + 1 + j] = qs[k].v[j];
        }
    return (long long)qs.size();
}

void orc_col_destroy(int precision, void* h) {
    if (precision == 64) delete (Detector<double>*)h;
    else delete (Detector<float>*)h;
}

long long orc_col_num_queries(int precision, void* h) {
    return precision == 64 ? (long long)((Detector<double>*)h)->qs.size() : (long long)((Detector<float>*)h)->qs.size();
}

int orc_col_num_hits(int precision, void* h) {
    return precision == 64 ? ((Detector<double>*)h)->hits : ((Detector<float>*)h)->hits;
}

// queries: 5 u32 each (type 1 VF / 2 EE, v0..v3); toi as double; normals 3 doubles
void orc_col_get(int precision, void* h, uint32_t* q, double* toi, double* nrm) {
    auto fill = [&](auto* d) {
        for (size_t k = 0; k < d->qs.size(); ++k) {
            q[5 * k] = d->qs[k].type;
            for (int j = 0; j < 4; ++j) q[5 * k + 1 + j] = d->qs[k].v[j];
            if (toi) toi[k] = (double)d->toi[k];
            if (nrm) {
                nrm[3 * k] = (double)d->nrm[k].x;
                nrm[3 * k + 1] = (double)d->nrm[k].y;
                nrm[3 * k + 2] = (double)d->nrm[k].z;
            }
        }
    };
    if (precision == 64) fill((Detector<double>*)h);
    else fill((Detector<float>*)h);
}

// One ccdCollisionTest on a query (type 1 VF / 2 EE) over the given arrays.
double orc_ccd_test(int precision, int type, const uint32_t* v, const void* X, const void* XT, double* n) {
    if (precision == 64) {
        V3<double> nn;
        const double t = ccd_test<double>(type, v, (const double*)X, (const double*)XT, nn);
        if (n) { n[0] = nn.x; n[1] = nn.y; n[2] = nn.z; }
        return t;
    }
    V3<float> nn;
    const float t = ccd_test<float>(type, v, (const float*)X, (const float*)XT, nn);
    if (n) { n[0] = nn.x; n[1] = nn.y; n[2] = nn.z; }
    return (double)t;
}

// Roots of a x^3 + b x^2 + c x + d (solveCubic), in double.
int orc_solve_cubic(double a, double b, double c, double d, double* x) { return solve_cubic(a, b, c, d, x); }

void orc_ccd_response(int precision, int nv, void* X, const void* XT, void* V, const void* tI, const float* normals) {
    if (precision == 64)
        ccd_response<double>(nv, (double*)X, (const double*)XT, (double*)V, (const double*)tI, normals);
    else
        ccd_response<float>(nv, (float*)X, (const float*)XT, (float*)V, (const float*)tI, normals);
}

}  // extern "C"
