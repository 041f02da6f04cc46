// LBVH broad phase, CCD narrow phase and hit resolution on gfx950.
//
// Everything in this file is compiled without floating-point contraction:
// leaf boxes, the cubic coefficients, the root finder and the signed
// distances are the written IEEE sequence, so pair lists, times of impact and
// hit sets are reproducible bit for bit (division and square root are
// correctly rounded in HIP by default).
#pragma clang fp contract(off)

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cfloat>
#include <cstring>

#include "collision.hpp"

namespace sbs {
namespace {

constexpr int kTpbC = 256;
constexpr int kStack = 64;  // traversal stack depth (a stack overflow is reported, not ignored)

template <class T> struct V3 {
    T x, y, z;
};
template <class T> __device__ __forceinline__ V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> __device__ __forceinline__ V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> __device__ __forceinline__ V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> __device__ __forceinline__ V3<T> neg(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> __device__ __forceinline__ T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> __device__ __forceinline__ V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
template <class T> __device__ __forceinline__ T length(V3<T> a) { return sqrt(dot(a, a)); }
template <class T> __device__ __forceinline__ V3<T> normalize(V3<T> a) { return ((T)1 / sqrt(dot(a, a))) * a; }
template <class T> __device__ __forceinline__ T stp(V3<T> u, V3<T> v, V3<T> w) { return dot(u, cross(v, w)); }
template <class T> __device__ __forceinline__ T sgn(T v) { return v > (T)0 ? (T)1 : (v < (T)0 ? (T)-1 : (T)0); }
template <class T> __device__ __forceinline__ V3<T> ldv(const T* p, uint32_t i) {
    return {p[3 * (size_t)i], p[3 * (size_t)i + 1], p[3 * (size_t)i + 2]};
}

// ------------------------------------------------------------------ CCD --
// solveQuadratic / newtonsMethod / solveCubic (intersections.cu:44-113).
template <class T> __device__ int solve_quadratic(T a, T b, T c, T* x) {
    const T d = b * b - (T)4 * a * c;
    if (d < (T)0) {
        x[0] = -b / ((T)2 * a);
        return 0;
    }
    const T q = -(b + sgn(b) * sqrt(d)) / (T)2;
    int i = 0;
    if ((double)fabs(a) > 1e-12 * (double)fabs(q)) x[i++] = q / a;
    if ((double)fabs(q) > 1e-12 * (double)fabs(c)) x[i++] = c / q;
    if (i == 2 && x[0] > x[1]) {
        const T tmp = x[0];
        x[0] = x[1];
        x[1] = tmp;
    }
    return i;
}

template <class T> __device__ T newton(T a, T b, T c, T d, T x0, int init_dir) {
    if (init_dir != 0) {
        const T y0 = d + x0 * (c + x0 * (b + x0 * a));
        const T ddy0 = (T)2 * b + x0 * ((T)6 * a);
        if (ddy0 != (T)0) x0 += (T)init_dir * sqrt(fabs((T)2 * y0 / ddy0));
    }
    for (int it = 0; it < 100; ++it) {
        const T y = d + x0 * (c + x0 * (b + x0 * a));
        const T dy = c + x0 * ((T)2 * b + x0 * (T)3 * a);
        if (dy == (T)0) return x0;
        const T x1 = x0 - y / dy;
        if ((double)fabs(x0 - x1) < 1e-6) return x0;
        x0 = x1;
    }
    return x0;
}

template <class T> __device__ int solve_cubic(T a, T b, T c, T d, T* x) {
    T xc[2] = {(T)0, (T)0};
    const int ncrit = solve_quadratic((T)3 * a, (T)2 * b, c, xc);
    if (ncrit == 0) {
        x[0] = newton(a, b, c, d, xc[0], 0);
        return 1;
    }
    if (ncrit == 1) return solve_quadratic(b, c, d, x);
    const T yc0 = d + xc[0] * (c + xc[0] * (b + xc[0] * a));
    const T yc1 = d + xc[1] * (c + xc[1] * (b + xc[1] * a));
    int i = 0;
    if (yc0 * a >= (T)0) x[i++] = newton(a, b, c, d, xc[0], -1);
    if (yc0 * yc1 <= (T)0) {
        const int closer = fabs(yc0) < fabs(yc1) ? 0 : 1;
        x[i++] = newton(a, b, c, d, xc[closer], closer == 0 ? 1 : -1);
    }
    if (yc1 * a <= (T)0) x[i++] = newton(a, b, c, d, xc[1], 1);
    return i;
}

// signed_vf_distance / signed_ee_distance (intersections.cu:114-165).
template <class T> __device__ T vf_distance(V3<T> x, V3<T> y0, V3<T> y1, V3<T> y2, V3<T>& n, T* w) {
    n = cross(normalize(y1 - y0), normalize(y2 - y0));
    if ((double)dot(n, n) < 1e-6) return (T)FLT_MAX;
    n = normalize(n);
    const T h = dot(x - y0, n);
    const T b0 = stp(y1 - x, y2 - x, n), b1 = stp(y2 - x, y0 - x, n), b2 = stp(y0 - x, y1 - x, n);
    const T s = b0 + b1 + b2;
    w[0] = (T)1;
    w[1] = -b0 / s;
    w[2] = -b1 / s;
    w[3] = -b2 / s;
    return h;
}

template <class T> __device__ T ee_distance(V3<T> x0, V3<T> x1, V3<T> y0, V3<T> y1, V3<T>& n, T* w) {
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

// ccdCollisionTest (intersections.cu:286-337): earliest root of the
// coplanarity cubic in [0, 1] at which the pair is within 1e-6 and inside.
template <class T> __device__ T ccd_test(const QueryKey& q, const T* X, const T* XT, V3<T>& n) {
    const V3<T> x0 = ldv(X, q.v0), x1 = ldv(X, q.v1), x2 = ldv(X, q.v2), x3 = ldv(X, q.v3);
    const V3<T> u0 = ldv(XT, q.v0) - x0, u1 = ldv(XT, q.v1) - x1, u2 = ldv(XT, q.v2) - x2, u3 = ldv(XT, q.v3) - x3;
    const V3<T> x01 = x1 - x0, x02 = x2 - x0, x03 = x3 - x0;
    const V3<T> u01 = u1 - u0, u02 = u2 - u0, u03 = u3 - u0;
    const T a0 = stp(x01, x02, x03);
    const T a1 = stp(u01, x02, x03) + stp(x01, u02, x03) + stp(x01, x02, u03);
    const T a2 = stp(x01, u02, u03) + stp(u01, x02, u03) + stp(u01, u02, x03);
    const T a3 = stp(u01, u02, u03);
    n = {(T)0, (T)0, (T)0};
    if ((double)fabs(a0) < 1e-12 * (double)length(x01) * (double)length(x02) * (double)length(x03)) return (T)1;
    T t[3] = {(T)0, (T)0, (T)0};
    const int nsol = solve_cubic(a3, a2, a1, a0, t);
    for (int i = 0; i < nsol; ++i) {
        if ((double)t[i] < -1e-12 || t[i] > (T)1) continue;
        const V3<T> xt0 = x0 + t[i] * u0, xt1 = x1 + t[i] * u1, xt2 = x2 + t[i] * u2, xt3 = x3 + t[i] * u3;
        T w[4] = {(T)0, (T)0, (T)0, (T)0};
        T d;
        bool inside;
        if (q.type == 1) {
            d = vf_distance(xt0, xt1, xt2, xt3, n, w);
            inside = (double)fmin(-w[1], fmin(-w[2], -w[3])) >= -1e-3;
        } else {
            d = ee_distance(xt0, xt1, xt2, xt3, n, w);
            inside = (double)fmin(w[0], fmin(w[1], fmin(-w[2], -w[3]))) >= -1e-3;
        }
        if (dot(n, w[1] * u1 + w[2] * u2 + w[3] * u3) > (T)0) n = neg(n);
        if ((double)fabs(d) < 1e-6 && inside) return t[i];
    }
    return (T)1;
}

// ----------------------------------------------------------- broad phase --
// Scene bounds over X and XTilde (Morton normalisation only: the pair set
// does not depend on the tree shape): per-block partials, then one block.
template <class T> __global__ void scene_partial_kernel(int nv, const T* X, const T* XT, T* part) {
    if (!XT) XT = X;
    __shared__ T red[6][kTpbC];
    const int v = blockIdx.x * kTpbC + threadIdx.x;
    T lo[3] = {(T)FLT_MAX, (T)FLT_MAX, (T)FLT_MAX}, hi[3] = {(T)-FLT_MAX, (T)-FLT_MAX, (T)-FLT_MAX};
    if (v < nv)
        for (int k = 0; k < 3; ++k) {
            const T a = X[3 * (size_t)v + k], b = XT[3 * (size_t)v + k];
            lo[k] = fmin(a, b);
            hi[k] = fmax(a, b);
        }
    for (int k = 0; k < 3; ++k) {
        red[k][threadIdx.x] = lo[k];
        red[3 + k][threadIdx.x] = hi[k];
    }
    __syncthreads();
    for (int s = kTpbC / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int k = 0; k < 3; ++k) {
                red[k][threadIdx.x] = fmin(red[k][threadIdx.x], red[k][threadIdx.x + s]);
                red[3 + k][threadIdx.x] = fmax(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[6 * (size_t)blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

template <class T> __global__ void scene_final_kernel(int nblocks, T* part, T* scene) {
    if (threadIdx.x >= 6) return;
    const int k = threadIdx.x;
    T r = part[k];
    for (int b = 1; b < nblocks; ++b) r = k < 3 ? fmin(r, part[6 * (size_t)b + k]) : fmax(r, part[6 * (size_t)b + k]);
    scene[k] = r;
}

__device__ __forceinline__ uint32_t spread_bits(uint32_t x) {  // 10 bits -> every third bit
    x = (x * 0x00010001u) & 0xFF0000FFu;
    x = (x * 0x00000101u) & 0x0F00F00Fu;
    x = (x * 0x00000011u) & 0xC30C30C3u;
    x = (x * 0x00000005u) & 0x49249249u;
    return x;
}

// Leaf box of a triangle's trajectory (ccd.cu:53-65): coordinates pass
// through float (fminf/fmaxf), then -/+ 0.01 in the solver precision.  With
// XT null, the box of the triangle at X grown by `bound` (computeTriTrajBBox,
// ccd.cu:74-85, the UpdateQueries candidates).  The key is the 30-bit Morton
// code of the box centre above the triangle index, so keys are unique.
template <class T>
__global__ void leaf_kernel(int nt, const uint32_t* tri, const T* X, const T* XT, T bound, const T* scene,
                            T* leaf_box, unsigned long long* keys, uint32_t* order) {
    const int i = blockIdx.x * kTpbC + threadIdx.x;
    if (i >= nt) return;
    const uint32_t t0 = tri[3 * (size_t)i], t1 = tri[3 * (size_t)i + 1], t2 = tri[3 * (size_t)i + 2];
    T lo[3], hi[3];
    uint32_t cell[3];
    for (int k = 0; k < 3; ++k) {
        const float p0 = (float)X[3 * (size_t)t0 + k], p1 = (float)X[3 * (size_t)t1 + k], p2 = (float)X[3 * (size_t)t2 + k];
        if (XT) {
            const float q0 = (float)XT[3 * (size_t)t0 + k], q1 = (float)XT[3 * (size_t)t1 + k],
                        q2 = (float)XT[3 * (size_t)t2 + k];
            const float mn = fminf(fminf(fminf(fminf(fminf(p0, p1), p2), q0), q1), q2);
            const float mx = fmaxf(fmaxf(fmaxf(fmaxf(fmaxf(p0, p1), p2), q0), q1), q2);
            lo[k] = (T)mn - (T)0.01;
            hi[k] = (T)mx + (T)0.01;
        } else {
            lo[k] = (T)fminf(fminf(p0, p1), p2) - bound;
            hi[k] = (T)fmaxf(fmaxf(p0, p1), p2) + bound;
        }
        const double ext = (double)scene[3 + k] - (double)scene[k];
        double u = ext > 0 ? (0.5 * ((double)lo[k] + (double)hi[k]) - (double)scene[k]) / ext : 0.0;
        u = fmin(fmax(u * 1024.0, 0.0), 1023.0);
        cell[k] = (uint32_t)u;
    }
    for (int k = 0; k < 3; ++k) {
        leaf_box[6 * (size_t)i + k] = lo[k];
        leaf_box[6 * (size_t)i + 3 + k] = hi[k];
    }
    const uint32_t code = (spread_bits(cell[0]) << 2) | (spread_bits(cell[1]) << 1) | spread_bits(cell[2]);
    keys[i] = ((unsigned long long)code << 32) | (uint32_t)i;
    order[i] = (uint32_t)i;
}

__device__ __forceinline__ int common_prefix(const unsigned long long* k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    return __clzll(k[i] ^ k[j]);
}

// Karras (2012) binary radix tree over the sorted unique keys: internal node
// i covers a key range starting or ending at i and splits it at the highest
// differing bit.  Children >= n-1 are leaves (leaf position + n - 1).
__global__ void karras_kernel(int n, const unsigned long long* keys, int* left, int* right, int* parent) {
    const int i = blockIdx.x * kTpbC + threadIdx.x;
    if (i == 0) parent[0] = -1;
    if (i >= n - 1) return;
    const int d = common_prefix(keys, n, i, i + 1) > common_prefix(keys, n, i, i - 1) ? 1 : -1;
    const int dmin = common_prefix(keys, n, i, i - d);
    int lmax = 2;
    while (common_prefix(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (common_prefix(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = common_prefix(keys, n, i, j);
    int s = 0;
    int t = l;
    do {
        t = (t + 1) >> 1;
        if (common_prefix(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + min(d, 0);
    const int lc = min(i, j) == gamma ? gamma + n - 1 : gamma;
    const int rc = max(i, j) == gamma + 1 ? gamma + 1 + n - 1 : gamma + 1;
    left[i] = lc;
    right[i] = rc;
    parent[lc] = i;
    parent[rc] = i;
}

template <class T>
__device__ __forceinline__ void union_box(const T* a, const T* b, T* out) {
    for (int k = 0; k < 3; ++k) {
        out[k] = fmin(a[k], b[k]);
        out[3 + k] = fmax(a[3 + k], b[3 + k]);
    }
}

// Bottom-up fit: each leaf writes its box and climbs; at each internal node
// the second arrival (agent-scope acquire/release counter) writes the union
// of the two children and continues, the first stops.  Unions of exact
// bounds are exact, so internal boxes contain their leaves' boxes exactly.
template <class T>
__global__ void refit_kernel(int n, const uint32_t* order, const T* leaf_box, const int* left, const int* right,
                             const int* parent, int* flags, T* node_box) {
    const int k = blockIdx.x * kTpbC + threadIdx.x;
    if (k >= n) return;
    int node = k + n - 1;
    const uint32_t tri = order[k];
    for (int c = 0; c < 6; ++c) node_box[6 * (size_t)node + c] = leaf_box[6 * (size_t)tri + c];
    for (int step = 0; step < n; ++step) {  // a path to the root has fewer than n edges
        const int p = parent[node];
        if (p < 0) break;
        const int prev = __hip_atomic_fetch_add(&flags[p], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (prev == 0) break;
        T a[6], b[6], u[6];
        for (int c = 0; c < 6; ++c) {
            a[c] = __hip_atomic_load(&node_box[6 * (size_t)left[p] + c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            b[c] = __hip_atomic_load(&node_box[6 * (size_t)right[p] + c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        union_box(a, b, u);
        for (int c = 0; c < 6; ++c)
            __hip_atomic_store(&node_box[6 * (size_t)p + c], u[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        node = p;
    }
}

// The Serial build type: leaves' boxes first, then one launch per level.  In
// launch k an internal node whose children were both fitted in earlier
// launches (ready < k; leaves count as fitted before launch 1) writes their
// union and records ready = k.  A child fitted in the same launch has
// ready = k or is not yet visible, so it is never read half written.
template <class T>
__global__ void leaf_boxes_kernel(int n, const uint32_t* order, const T* leaf_box, T* node_box) {
    const int k = blockIdx.x * kTpbC + threadIdx.x;
    if (k >= n) return;
    const uint32_t tri = order[k];
    for (int c = 0; c < 6; ++c) node_box[6 * ((size_t)k + n - 1) + c] = leaf_box[6 * (size_t)tri + c];
}

template <class T>
__global__ void refit_level_kernel(int n, int launch, const int* left, const int* right, int* ready, T* node_box) {
    const int i = blockIdx.x * kTpbC + threadIdx.x;
    if (i >= n - 1 || ready[i] < launch) return;
    const int l = left[i], r = right[i];
    const int rl = l >= n - 1 ? 0 : __hip_atomic_load(&ready[l], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int rr = r >= n - 1 ? 0 : __hip_atomic_load(&ready[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (rl >= launch || rr >= launch) return;
    T u[6];
    union_box(node_box + 6 * (size_t)l, node_box + 6 * (size_t)r, u);
    for (int c = 0; c < 6; ++c) node_box[6 * (size_t)i + c] = u[c];
    __hip_atomic_store(&ready[i], launch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void ready_init_kernel(int n, int* ready) {
    const int i = blockIdx.x * kTpbC + threadIdx.x;
    if (i < n) ready[i] = 0x7fffffff;
}

template <class T> __device__ __forceinline__ bool overlap(const T* a, const T* b) {
    for (int k = 0; k < 3; ++k)
        if (a[3 + k] < b[k] || a[k] > b[3 + k]) return false;
    return true;
}

// Every leaf walks the tree (its walk split over the subtrees `depth` levels
// below the root, one thread each) and records the ordered pairs
// (its triangle, hit triangle) that pass the filters of traverseTree
// (broadphase.cu:331-398): a different triangle, no shared vertex, and a
// different body when ignore_self.  Pairs past `cap` are counted, not stored.
template <class T>
__global__ void traverse_kernel(int n, int depth, const uint32_t* order, const uint32_t* tri, const uint32_t* fathers,
                                const int* left, const int* right, const T* node_box, int ignore_self, uint2* pairs,
                                unsigned long long cap, unsigned long long* counters) {
    // thread g = s * n + k: leaf k (Morton order), subtree s of the 2^depth
    // nodes `depth` levels below the root (consecutive threads take
    // neighbouring leaves, so a wavefront's walks stay coherent)
    const unsigned long long g = (unsigned long long)blockIdx.x * kTpbC + threadIdx.x;
    if (g >= ((unsigned long long)n << depth)) return;
    const int k = (int)(g % (unsigned long long)n);
    const unsigned s = (unsigned)(g / (unsigned long long)n);
    const uint32_t ti = order[k];
    T qb[6];
    for (int c = 0; c < 6; ++c) qb[c] = node_box[6 * (size_t)(k + n - 1) + c];
    const uint32_t a0 = tri[3 * (size_t)ti], a1 = tri[3 * (size_t)ti + 1], a2 = tri[3 * (size_t)ti + 2];
    const uint32_t fa = fathers[ti];
    auto leaf = [&](int m) {  // m overlaps the query box
        const uint32_t tj = order[m - (n - 1)];
        if (tj == ti) return;
        if (ignore_self && fathers[tj] == fa) return;
        const uint32_t b0 = tri[3 * (size_t)tj], b1 = tri[3 * (size_t)tj + 1], b2 = tri[3 * (size_t)tj + 2];
        const bool adj = a0 == b0 || a0 == b1 || a0 == b2 || a1 == b0 || a1 == b1 || a1 == b2 || a2 == b0 ||
                         a2 == b1 || a2 == b2;
        if (adj) return;
        const unsigned long long slot = atomicAdd(&counters[0], 1ull);
        if (slot < cap) pairs[slot] = make_uint2(ti, tj);
    };
    // descend along the bits of s; a subtree whose box misses the query box
    // holds no overlapping leaf, and a leaf met above `depth` is taken by the
    // one thread whose remaining bits are 0
    int root = 0;
    for (int l = 0; l < depth; ++l) {
        const int rest = depth - 1 - l;
        const int m = ((s >> rest) & 1u) ? right[root] : left[root];
        T cb[6];
        for (int e = 0; e < 6; ++e) cb[e] = node_box[6 * (size_t)m + e];
        if (!overlap(qb, cb)) return;
        if (m >= n - 1) {
            if ((s & ((1u << rest) - 1u)) == 0) leaf(m);
            return;
        }
        root = m;
    }
    int stack[kStack];
    int sp = 0;
    stack[sp++] = root;
    for (int visits = 0; sp > 0; ++visits) {
        if (visits >= n) {  // a tree has n - 1 internal nodes: more visits means a malformed tree
            atomicOr(&counters[1], 2ull);
            break;
        }
        const int node = stack[--sp];
        const int ch[2] = {left[node], right[node]};
        for (int c = 0; c < 2; ++c) {
            const int m = ch[c];
            T cb[6];
            for (int e = 0; e < 6; ++e) cb[e] = node_box[6 * (size_t)m + e];
            if (!overlap(qb, cb)) continue;
            if (m >= n - 1) {
                leaf(m);
            } else if (sp < kStack) {
                stack[sp++] = m;
            } else {
                atomicOr(&counters[1], 1ull);  // reported to the host as an error
            }
        }
    }
}

// The 12 queries of an ordered pair in canonical form (fillQuery and
// sortEachQuery, broadphase.cu:285-319,450-470); dropped ones get type ~0.
__global__ void expand_kernel(unsigned long long npairs, const uint2* pairs, const uint32_t* tri, QueryKey* q) {
    const unsigned long long p = (unsigned long long)blockIdx.x * kTpbC + threadIdx.x;
    if (p >= npairs) return;
    const uint2 pr = pairs[p];
    uint32_t a[3], b[3];
    for (int k = 0; k < 3; ++k) {
        a[k] = tri[3 * (size_t)pr.x + k];
        b[k] = tri[3 * (size_t)pr.y + k];
    }
    const int ed[6] = {0, 1, 0, 2, 1, 2};
    QueryKey out[12];
    for (int i = 0; i < 3; ++i) out[i] = {1u, a[i], b[0], b[1], b[2]};
    for (int i = 0; i < 3; ++i) {
        const uint32_t e0 = a[ed[2 * i]], e1 = a[ed[2 * i + 1]];
        out[3 + 3 * i] = {2u, e0, e1, b[0], b[1]};
        out[4 + 3 * i] = {2u, e0, e1, b[0], b[2]};
        out[5 + 3 * i] = {2u, e0, e1, b[1], b[2]};
    }
    for (int k = 0; k < 12; ++k) {
        QueryKey& x = out[k];
        uint32_t t;
        if (x.type == 1u) {
            if (x.v0 == x.v1 || x.v0 == x.v2 || x.v0 == x.v3) {
                x.type = ~0u;
            } else {
                if (x.v1 > x.v2) { t = x.v1; x.v1 = x.v2; x.v2 = t; }
                if (x.v2 > x.v3) { t = x.v2; x.v2 = x.v3; x.v3 = t; }
                if (x.v1 > x.v2) { t = x.v1; x.v1 = x.v2; x.v2 = t; }
            }
        } else {
            if (x.v0 > x.v1) { t = x.v0; x.v0 = x.v1; x.v1 = t; }
            if (x.v2 > x.v3) { t = x.v2; x.v2 = x.v3; x.v3 = t; }
            if (x.v0 == x.v2 && x.v1 == x.v3) x.type = ~0u;
        }
        q[12 * p + k] = x;
    }
}

struct QueryLess {
    __device__ bool operator()(const QueryKey& a, const QueryKey& b) const {
        if (a.type != b.type) return a.type < b.type;
        if (a.v0 != b.v0) return a.v0 < b.v0;
        if (a.v1 != b.v1) return a.v1 < b.v1;
        if (a.v2 != b.v2) return a.v2 < b.v2;
        return a.v3 < b.v3;
    }
};

__device__ __forceinline__ bool same_query(const QueryKey& a, const QueryKey& b) {
    return a.type == b.type && a.v0 == b.v0 && a.v1 == b.v1 && a.v2 == b.v2 && a.v3 == b.v3;
}

__global__ void unique_flag_kernel(unsigned long long n, const QueryKey* q, int* flag) {
    const unsigned long long i = (unsigned long long)blockIdx.x * kTpbC + threadIdx.x;
    if (i >= n) return;
    flag[i] = q[i].type != ~0u && (i == 0 || !same_query(q[i], q[i - 1])) ? 1 : 0;
}

__global__ void unique_scatter_kernel(unsigned long long n, const QueryKey* q, const int* flag, const int* scan,
                                      QueryKey* out) {
    const unsigned long long i = (unsigned long long)blockIdx.x * kTpbC + threadIdx.x;
    if (i >= n) return;
    if (flag[i]) out[scan[i]] = q[i];
}

template <class T>
__global__ void narrow_kernel(long long nq, const QueryKey* q, const T* X, const T* XT, T* toi, T* nrm) {
    const long long i = (long long)blockIdx.x * kTpbC + threadIdx.x;
    if (i >= nq) return;
    V3<T> n;
    toi[i] = ccd_test<T>(q[i], X, XT, n);
    nrm[3 * i] = n.x;
    nrm[3 * i + 1] = n.y;
    nrm[3 * i + 2] = n.z;
}

__device__ __forceinline__ bool same_group(const QueryKey& a, const QueryKey& b) {
    return a.type == b.type && a.v0 == b.v0 && (a.type == 1u || a.v1 == b.v1);
}

// First-minimum toi of the group led by query i (groups are contiguous in
// the sorted list: VF by v0, EE by (v0, v1)); -1 when i does not lead one.
template <class T> __device__ __forceinline__ long long group_winner(long long nq, const QueryKey* q, const T* toi, long long i) {
    if (i > 0 && same_group(q[i], q[i - 1])) return -1;
    long long best = i;
    for (long long k = i + 1; k < nq && same_group(q[k], q[i]); ++k)
        if (toi[k] < toi[best]) best = k;
    return best;
}

template <class T> __global__ void fill_kernel(int n, T* p, T v) {
    const int i = blockIdx.x * kTpbC + threadIdx.x;
    if (i < n) p[i] = v;
}

// storeTi (narrowphase.cu:72-112) for the group winners with toi < 1: their
// vertices get tI = 0.5; the last winner in list order that touches a vertex
// owns its normal.
template <class T>
__global__ void resolve_kernel(long long nq, const QueryKey* q, const T* toi, T* tI, int* owner,
                               unsigned long long* counters) {
    const long long i = (long long)blockIdx.x * kTpbC + threadIdx.x;
    if (i >= nq) return;
    const long long w = group_winner(nq, q, toi, i);
    if (w < 0 || !(toi[w] < (T)1)) return;
    atomicAdd(&counters[2], 1ull);
    const QueryKey& x = q[w];
    const uint32_t vs[4] = {x.v0, x.v1, x.v2, x.v3};
    const int nvq = x.type == 2u ? 2 : 4;
    for (int k = 0; k < nvq; ++k) {
        tI[vs[k]] = (T)0.5;
        atomicMax(&owner[vs[k]], (int)i);
    }
}

template <class T>
__global__ void normals_kernel(long long nq, const QueryKey* q, const T* toi, const T* nrm, const int* owner,
                               float* normals) {
    const long long i = (long long)blockIdx.x * kTpbC + threadIdx.x;
    if (i >= nq) return;
    const long long w = group_winner(nq, q, toi, i);
    if (w < 0 || !(toi[w] < (T)1)) return;
    const QueryKey& x = q[w];
    const uint32_t vs[4] = {x.v0, x.v1, x.v2, x.v3};
    const float nx = (float)nrm[3 * w], ny = (float)nrm[3 * w + 1], nz = (float)nrm[3 * w + 2];
    const int nvq = x.type == 2u ? 2 : 4;
    for (int k = 0; k < nvq; ++k) {
        if (owner[vs[k]] != (int)i) continue;
        const float s = (x.type == 1u && k > 0) ? -1.f : 1.f;
        normals[3 * (size_t)vs[k]] = s * nx;
        normals[3 * (size_t)vs[k] + 1] = s * ny;
        normals[3 * (size_t)vs[k] + 2] = s * nz;
    }
}

template <class T> __global__ void min_partial_kernel(long long n, const T* v, T* part) {
    __shared__ T red[kTpbC];
    const long long i = (long long)blockIdx.x * kTpbC + threadIdx.x;
    red[threadIdx.x] = i < n ? v[i] : (T)1;
    __syncthreads();
    for (int s = kTpbC / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

template <class T> __global__ void min_final_kernel(int nb, T* part) {
    if (threadIdx.x != 0) return;
    T r = (T)1;
    for (int b = 0; b < nb; ++b) r = fmin(r, part[b]);
    part[0] = r;
}

template <class T>
__global__ void ccd_response_kernel(int nv, T* X, const T* XT, T* V, const T* tI, const float* normals) {
    const int v = blockIdx.x * kTpbC + threadIdx.x;
    if (v >= nv) return;
    const V3<T> x = ldv(X, (uint32_t)v), xt = ldv(XT, (uint32_t)v);
    if (tI[v] < (T)1) {
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

inline unsigned blocks_for(unsigned long long n) { return (unsigned)((n + kTpbC - 1) / kTpbC); }

// --------------------------------------------------------- fixed bodies --
// The response shared by the three body types (fixedBodyData.cu:80-85,
// 101-106, 126-131): normal velocity reflected and scaled by muN, tangential
// velocity shortened by muT (1 + muN) |v_N| (Coulomb-like), clamped at 0.
template <class T> __device__ __forceinline__ V3<T> fixed_response(V3<T> vel, V3<T> n, T mu_t, T mu_n) {
    const V3<T> vn = dot(vel, n) * n;
    const V3<T> vt = vel - vn;
    const T mag = length(vt);
    const T t = (T)1 - mu_t * ((T)1 + mu_n) * length(vn) / mag;
    const T a = mag == (T)0 ? (T)0 : (t < (T)0 ? (T)0 : t);  // glm::max(t, 0)
    return (-mu_n) * vn + a * vt;
}

// HandleCollisions (fixedBodyData.cu:136-148): one thread per vertex visits
// the spheres, then the planes, then the cylinders, in the order of the
// reference's three kernels, each body seeing the position and velocity left
// by the previous one:
//   sphere (:91-109):   d = |x - c| < r  -> x += (r - d) n, n = (x - c)/d;
//   plane (:68-88):     d = (x - p).u < 0 and v.u < 0 -> x -= d u;
//   cylinder (:112-134): n = (I - a a^T)(x - c), d = |n| < r -> x += (r - d) n/|n|
// (an infinite cylinder: the reference has no height test), each followed by
// the response above.  Body parameters are float and widened to T.
template <class T>
__global__ void fixed_bodies_kernel(int nv, T* X, T* V, const float* __restrict__ body, int ns, int np, int nc,
                                    T mu_t, T mu_n) {
    const int v = blockIdx.x * kTpbC + threadIdx.x;
    if (v >= nv) return;
    V3<T> x = ldv(X, (uint32_t)v), vel = ldv(V, (uint32_t)v);
    auto p_of = [&](int j) {
        const float* b = body + kFixedFloats * (size_t)j;
        return V3<T>{(T)b[0], (T)b[1], (T)b[2]};
    };
    auto u_of = [&](int j) {
        const float* b = body + kFixedFloats * (size_t)j;
        return V3<T>{(T)b[3], (T)b[4], (T)b[5]};
    };
    for (int j = 0; j < ns; ++j) {
        const V3<T> c = p_of(j);
        const T r = (T)body[kFixedFloats * (size_t)j + 6];
        const V3<T> tc = x - c;
        const T d = length(tc);
        if (d < r) {
            const V3<T> n = normalize(tc);
            x = x + (r - d) * n;
            vel = fixed_response(vel, n, mu_t, mu_n);
        }
    }
    for (int j = ns; j < ns + np; ++j) {
        const V3<T> p = p_of(j), u = u_of(j);
        const T d = dot(x - p, u);
        if (d < (T)0 && dot(vel, u) < (T)0) {
            x = x - d * u;
            vel = fixed_response(vel, u, mu_t, mu_n);
        }
    }
    for (int j = ns + np; j < ns + np + nc; ++j) {
        const V3<T> c = p_of(j), a = u_of(j);
        const T r = (T)body[kFixedFloats * (size_t)j + 6];
        // nnT = I - outer(a, a), entries 1 - a_i a_j on the diagonal, 0 - a_i a_j off it
        const T m00 = (T)1 - a.x * a.x, m01 = (T)0 - a.x * a.y, m02 = (T)0 - a.x * a.z;
        const T m11 = (T)1 - a.y * a.y, m12 = (T)0 - a.y * a.z, m22 = (T)1 - a.z * a.z;
        const V3<T> w = x - c;
        const V3<T> n = {m00 * w.x + m01 * w.y + m02 * w.z, m01 * w.x + m11 * w.y + m12 * w.z,
                         m02 * w.x + m12 * w.y + m22 * w.z};
        const T d = length(n);
        if (d < r) {
            const V3<T> nn = normalize(n);
            x = x + (r - d) * nn;
            vel = fixed_response(vel, nn, mu_t, mu_n);
        }
    }
    X[3 * (size_t)v] = x.x;
    X[3 * (size_t)v + 1] = x.y;
    X[3 * (size_t)v + 2] = x.z;
    V[3 * (size_t)v] = vel.x;
    V[3 * (size_t)v + 1] = vel.y;
    V[3 * (size_t)v + 2] = vel.z;
}

}  // namespace

FixedBodies::FixedBodies(const std::vector<FixedBodySpec>& bodies) {
    std::vector<float> grouped;
    params_.assign(kFixedFloats * bodies.size(), 0.f);
    for (int type : {kFixedSphere, kFixedPlane, kFixedCylinder}) {
        for (size_t k = 0; k < bodies.size(); ++k) {
            const FixedBodySpec& b = bodies[k];
            SBS_REQUIRE(b.type >= kFixedSphere && b.type <= kFixedCylinder, "fixed body: unknown type");
            if (b.type != type) continue;
            const float* m = b.model;  // m[4 * col + row]
            float q[kFixedFloats] = {m[12], m[13], m[14], 0.f, 0.f, 0.f, b.radius, 0.f};
            if (type == kFixedPlane) {
                // row 1 of A^-1 = row 1 of adj(A) / det(A), A the model's 3x3 part
                auto A = [&](int r, int c) { return m[4 * c + r]; };
                const float det = A(0, 0) * (A(1, 1) * A(2, 2) - A(2, 1) * A(1, 2)) -
                                  A(0, 1) * (A(1, 0) * A(2, 2) - A(2, 0) * A(1, 2)) +
                                  A(0, 2) * (A(1, 0) * A(2, 1) - A(2, 0) * A(1, 1));
                SBS_REQUIRE(det != 0.f, "fixed body: singular plane model matrix");
                const float w[3] = {-(A(1, 0) * A(2, 2) - A(1, 2) * A(2, 0)) / det,
                                    (A(0, 0) * A(2, 2) - A(0, 2) * A(2, 0)) / det,
                                    -(A(0, 0) * A(1, 2) - A(0, 2) * A(1, 0)) / det};
                const float s = 1.f / std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
                q[3] = w[0] * s; q[4] = w[1] * s; q[5] = w[2] * s;
            } else if (type == kFixedCylinder) {
                const float s = 1.f / std::sqrt(m[4] * m[4] + m[5] * m[5] + m[6] * m[6] + m[7] * m[7]);
                q[3] = m[4] * s; q[4] = m[5] * s; q[5] = m[6] * s;
            }
            for (int i = 0; i < kFixedFloats; ++i) {
                SBS_REQUIRE(std::isfinite(q[i]), "fixed body: non-finite parameter");
                grouped.push_back(q[i]);
                params_[kFixedFloats * k + i] = q[i];
            }
            (type == kFixedSphere ? ns_ : type == kFixedPlane ? np_ : nc_)++;
        }
    }
    if (!grouped.empty()) {
        dev_.alloc(grouped.size());
        SBS_HIP(hipMemcpy(dev_.get(), grouped.data(), grouped.size() * sizeof(float), hipMemcpyHostToDevice));
    }
}

template <class T> void FixedBodies::handle(T* X, T* V, int nv, T mu_t, T mu_n, hipStream_t stream) const {
    SBS_REQUIRE(nv >= 0 && (nv == 0 || (X && V)), "HandleCollisions: null arrays");
    if (nv == 0 || ns_ + np_ + nc_ == 0) return;
    hipLaunchKernelGGL(fixed_bodies_kernel<T>, dim3(blocks_for((unsigned long long)nv)), dim3(kTpbC), 0, stream, nv,
                       X, V, (const float*)dev_.get(), ns_, np_, nc_, mu_t, mu_n);
    SBS_HIP(hipGetLastError());
}

template <class T>
CollisionDetector<T>::CollisionDetector(int num_verts, int num_tris, const uint32_t* tri, const uint32_t* fathers,
                                        hipStream_t stream)
    : nv_(num_verts), nt_(num_tris), tri_(tri), fathers_(fathers), stream_(stream) {
    SBS_REQUIRE(num_verts > 0 && num_tris >= 0, "CollisionDetection: bad sizes");
    SBS_REQUIRE(num_tris == 0 || (tri != nullptr && fathers != nullptr), "CollisionDetection: null triangle arrays");
    // triangle indices are validated on the host before any kernel reads them
    if (nt_ > 0) {
        std::vector<uint32_t> h(3 * (size_t)nt_);
        SBS_HIP(hipMemcpy(h.data(), tri, h.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
        for (uint32_t v : h) SBS_REQUIRE(v < (uint32_t)nv_, "CollisionDetection: triangle index out of range");
    }
    const int n = std::max(nt_, 1);
    leaf_box_.alloc(6 * (size_t)n);
    keys_.alloc(n);
    keys_sorted_.alloc(n);
    order_in_.alloc(n);
    order_.alloc(n);
    left_.alloc(std::max(n - 1, 1));
    right_.alloc(std::max(n - 1, 1));
    parent_.alloc(2 * (size_t)n);
    flags_.alloc(std::max(n - 1, 1));
    node_box_.alloc(6 * (2 * (size_t)n));
    scene_.alloc(6 * (size_t)blocks_for(nv_) + 6);
    counters_.alloc(4);
    owner_.alloc(nv_);
    SBS_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_tmp_bytes_, keys_.get(), keys_sorted_.get(),
                                               order_in_.get(), order_.get(), n, 0, 64, stream_));
    sort_tmp_.alloc(std::max<size_t>(sort_tmp_bytes_, 1));
    pair_cap_ = std::max<size_t>(1024, 8 * (size_t)n);
    pairs_.alloc(pair_cap_);
    SBS_HIP(hipHostMalloc((void**)&host_, 4 * sizeof(unsigned long long), hipHostMallocDefault));
}

template <class T> void CollisionDetector<T>::set_build_type(int t) {
    SBS_REQUIRE(t >= 0 && t <= 2, "CollisionDetection::SetBuildType: 0 Serial, 1 Atomic or 2 Cooperative");
    build_type_ = t;
}

template <class T> CollisionDetector<T>::~CollisionDetector() {
    if (host_) (void)hipHostFree(host_);
}

template <class T> void CollisionDetector<T>::broad_phase(const T* X, const T* XT, bool ignore_self, T bound) {
    nq_ = 0;
    npairs_ = 0;
    if (nt_ < 2) return;
    const int n = nt_;
    const unsigned vb = blocks_for(nv_);
    hipLaunchKernelGGL(scene_partial_kernel<T>, dim3(vb), dim3(kTpbC), 0, stream_, nv_, X, XT, scene_.get());
    hipLaunchKernelGGL(scene_final_kernel<T>, dim3(1), dim3(64), 0, stream_, (int)vb, scene_.get(),
                       scene_.get() + 6 * (size_t)vb);
    hipLaunchKernelGGL(leaf_kernel<T>, dim3(blocks_for(n)), dim3(kTpbC), 0, stream_, n, tri_, X, XT, bound,
                       (const T*)(scene_.get() + 6 * (size_t)vb), leaf_box_.get(), keys_.get(), order_in_.get());
    SBS_HIP(hipGetLastError());
    size_t bytes = sort_tmp_bytes_;
    SBS_HIP(hipcub::DeviceRadixSort::SortPairs(sort_tmp_.get(), bytes, keys_.get(), keys_sorted_.get(),
                                               order_in_.get(), order_.get(), n, 0, 64, stream_));
    hipLaunchKernelGGL(karras_kernel, dim3(blocks_for(n)), dim3(kTpbC), 0, stream_, n,
                       (const unsigned long long*)keys_sorted_.get(), left_.get(), right_.get(), parent_.get());
    if (build_type_ == 0) {
        hipLaunchKernelGGL(leaf_boxes_kernel<T>, dim3(blocks_for(n)), dim3(kTpbC), 0, stream_, n,
                           (const uint32_t*)order_.get(), (const T*)leaf_box_.get(), node_box_.get());
        hipLaunchKernelGGL(ready_init_kernel, dim3(blocks_for(n)), dim3(kTpbC), 0, stream_, n - 1, flags_.get());
        SBS_HIP(hipGetLastError());
        // a tree of n leaves has at most n - 1 levels above them
        int root_ready = 0x7fffffff;
        for (int launch = 1; launch <= n - 1 && root_ready == 0x7fffffff; ++launch) {
            hipLaunchKernelGGL(refit_level_kernel<T>, dim3(blocks_for(n)), dim3(kTpbC), 0, stream_, n, launch,
                               (const int*)left_.get(), (const int*)right_.get(), flags_.get(), node_box_.get());
            SBS_HIP(hipGetLastError());
            SBS_HIP(hipMemcpyAsync(&root_ready, flags_.get(), sizeof(int), hipMemcpyDeviceToHost, stream_));
            SBS_HIP(hipStreamSynchronize(stream_));
        }
        SBS_REQUIRE(root_ready != 0x7fffffff, "CollisionDetection: serial BVH fit did not reach the root");
    } else {
        SBS_HIP(hipMemsetAsync(flags_.get(), 0, flags_.bytes(), stream_));
        hipLaunchKernelGGL(refit_kernel<T>, dim3(blocks_for(n)), dim3(kTpbC), 0, stream_, n,
                           (const uint32_t*)order_.get(), (const T*)leaf_box_.get(), (const int*)left_.get(),
                           (const int*)right_.get(), (const int*)parent_.get(), flags_.get(), node_box_.get());
        SBS_HIP(hipGetLastError());
    }
    // each leaf's walk is split over the 2^depth subtrees below the root, so
    // that about 2^18 threads walk (one per leaf would leave most CUs idle)
    int depth = 0;
    while (depth < 6 && ((unsigned long long)n << depth) < (1ull << 18)) ++depth;
    for (int attempt = 0; attempt < 2; ++attempt) {
        SBS_HIP(hipMemsetAsync(counters_.get(), 0, counters_.bytes(), stream_));
        hipLaunchKernelGGL(traverse_kernel<T>, dim3(blocks_for((unsigned long long)n << depth)), dim3(kTpbC), 0,
                           stream_, n, depth, (const uint32_t*)order_.get(), tri_, fathers_, (const int*)left_.get(),
                           (const int*)right_.get(), (const T*)node_box_.get(), (int)ignore_self, pairs_.get(),
                           (unsigned long long)pair_cap_, counters_.get());
        SBS_HIP(hipGetLastError());
        SBS_HIP(hipMemcpyAsync(host_, counters_.get(), 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream_));
        SBS_HIP(hipStreamSynchronize(stream_));
        SBS_REQUIRE(host_[1] == 0, "CollisionDetection: BVH traversal stack overflow or malformed tree");
        if (host_[0] <= pair_cap_) break;
        pair_cap_ = (size_t)host_[0] + host_[0] / 4;  // grow and run the traversal again
        pairs_.alloc(pair_cap_);
    }
    npairs_ = (long long)host_[0];
    SBS_REQUIRE((size_t)npairs_ <= pair_cap_, "CollisionDetection: candidate pair count changed between traversals");
    if (npairs_ == 0) return;
    const unsigned long long nraw = 12ull * (unsigned long long)npairs_;
    if (nraw > qcap_) {
        qcap_ = (size_t)nraw;
        qraw_.alloc(qcap_);
        qsorted_.alloc(qcap_);
        qunique_.alloc(qcap_);
        qflag_.alloc(qcap_);
        qscan_.alloc(qcap_);
        toi_.alloc(qcap_);
        qnormal_.alloc(3 * qcap_);
        red_.alloc(blocks_for(qcap_) + 1);
        size_t b1 = 0, b2 = 0;
        SBS_HIP(hipcub::DeviceMergeSort::SortKeysCopy(nullptr, b1, qraw_.get(), qsorted_.get(), (int64_t)nraw,
                                                      QueryLess(), stream_));
        SBS_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, b2, qflag_.get(), qscan_.get(), (int)nraw, stream_));
        qtmp_bytes_ = std::max(b1, b2);
        qtmp_.alloc(std::max<size_t>(qtmp_bytes_, 1));
    }
    hipLaunchKernelGGL(expand_kernel, dim3(blocks_for(npairs_)), dim3(kTpbC), 0, stream_,
                       (unsigned long long)npairs_, (const uint2*)pairs_.get(), tri_, qraw_.get());
    size_t bytes1 = qtmp_bytes_;
    SBS_HIP(hipcub::DeviceMergeSort::SortKeysCopy(qtmp_.get(), bytes1, qraw_.get(), qsorted_.get(), (int64_t)nraw,
                                                  QueryLess(), stream_));
    hipLaunchKernelGGL(unique_flag_kernel, dim3(blocks_for(nraw)), dim3(kTpbC), 0, stream_, nraw,
                       (const QueryKey*)qsorted_.get(), qflag_.get());
    size_t bytes2 = qtmp_bytes_;
    SBS_HIP(hipcub::DeviceScan::ExclusiveSum(qtmp_.get(), bytes2, qflag_.get(), qscan_.get(), (int)nraw, stream_));
    hipLaunchKernelGGL(unique_scatter_kernel, dim3(blocks_for(nraw)), dim3(kTpbC), 0, stream_, nraw,
                       (const QueryKey*)qsorted_.get(), (const int*)qflag_.get(), (const int*)qscan_.get(),
                       qunique_.get());
    SBS_HIP(hipGetLastError());
    int last[2] = {0, 0};
    SBS_HIP(hipMemcpyAsync(&last[0], qscan_.get() + (nraw - 1), sizeof(int), hipMemcpyDeviceToHost, stream_));
    SBS_HIP(hipMemcpyAsync(&last[1], qflag_.get() + (nraw - 1), sizeof(int), hipMemcpyDeviceToHost, stream_));
    SBS_HIP(hipStreamSynchronize(stream_));
    nq_ = (long long)last[0] + last[1];
}

template <class T> void CollisionDetector<T>::narrow_phase(const T* X, const T* XT) {
    if (nq_ == 0) return;
    hipLaunchKernelGGL(narrow_kernel<T>, dim3(blocks_for(nq_)), dim3(kTpbC), 0, stream_, nq_,
                       (const QueryKey*)qunique_.get(), X, XT, toi_.get(), qnormal_.get());
    SBS_HIP(hipGetLastError());
}

template <class T>
void CollisionDetector<T>::detect(const T* X, const T* XT, bool ignore_self, T* tI, float* normals) {
    SBS_REQUIRE(X && XT && tI && normals, "DetectCollision: null array");
    hits_ = 0;
    hipLaunchKernelGGL(fill_kernel<T>, dim3(blocks_for(nv_)), dim3(kTpbC), 0, stream_, nv_, tI, (T)1);
    broad_phase(X, XT, ignore_self);
    narrow_phase(X, XT);
    if (nq_ > 0) {
        hipLaunchKernelGGL(fill_kernel<int>, dim3(blocks_for(nv_)), dim3(kTpbC), 0, stream_, nv_, owner_.get(), -1);
        SBS_HIP(hipMemsetAsync(counters_.get() + 2, 0, sizeof(unsigned long long), stream_));
        hipLaunchKernelGGL(resolve_kernel<T>, dim3(blocks_for(nq_)), dim3(kTpbC), 0, stream_, nq_,
                           (const QueryKey*)qunique_.get(), (const T*)toi_.get(), tI, owner_.get(), counters_.get());
        hipLaunchKernelGGL(normals_kernel<T>, dim3(blocks_for(nq_)), dim3(kTpbC), 0, stream_, nq_,
                           (const QueryKey*)qunique_.get(), (const T*)toi_.get(), (const T*)qnormal_.get(),
                           (const int*)owner_.get(), normals);
        SBS_HIP(hipGetLastError());
        SBS_HIP(hipMemcpyAsync(host_ + 2, counters_.get() + 2, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                               stream_));
        SBS_HIP(hipStreamSynchronize(stream_));
        hits_ = (int)host_[2];
    }
}

template <class T> double CollisionDetector<T>::min_step(const T* X, const T* XT, bool ignore_self) {
    SBS_REQUIRE(X && XT, "ComputeMinStepSize: null array");
    broad_phase(X, XT, ignore_self);
    if (nq_ == 0) return 1.0;
    narrow_phase(X, XT);
    const unsigned nb = blocks_for(nq_);
    hipLaunchKernelGGL(min_partial_kernel<T>, dim3(nb), dim3(kTpbC), 0, stream_, nq_, (const T*)toi_.get(),
                       red_.get());
    hipLaunchKernelGGL(min_final_kernel<T>, dim3(1), dim3(64), 0, stream_, (int)nb, red_.get());
    T r = 1;
    SBS_HIP(hipMemcpyAsync(&r, red_.get(), sizeof(T), hipMemcpyDeviceToHost, stream_));
    SBS_HIP(hipStreamSynchronize(stream_));
    return (double)r;
}

template <class T> long long CollisionDetector<T>::proximity(const T* X, T bound, bool ignore_self) {
    SBS_REQUIRE(X, "UpdateQueries: null array");
    broad_phase(X, nullptr, ignore_self, bound);
    return nq_;
}

template <class T> void CollisionDetector<T>::get_queries(uint32_t* q5, double* toi, double* n3) const {
    if (nq_ == 0) return;
    std::vector<QueryKey> q(nq_);
    std::vector<T> t(nq_), n(3 * nq_);
    SBS_HIP(hipStreamSynchronize(stream_));
    SBS_HIP(hipMemcpy(q.data(), qunique_.get(), nq_ * sizeof(QueryKey), hipMemcpyDeviceToHost));
    SBS_HIP(hipMemcpy(t.data(), toi_.get(), nq_ * sizeof(T), hipMemcpyDeviceToHost));
    SBS_HIP(hipMemcpy(n.data(), qnormal_.get(), 3 * nq_ * sizeof(T), hipMemcpyDeviceToHost));
    for (long long i = 0; i < nq_; ++i) {
        if (q5) {
            q5[5 * i] = q[i].type;
            q5[5 * i + 1] = q[i].v0;
            q5[5 * i + 2] = q[i].v1;
            q5[5 * i + 3] = q[i].v2;
            q5[5 * i + 4] = q[i].v3;
        }
        if (toi) toi[i] = (double)t[i];
        if (n3)
            for (int k = 0; k < 3; ++k) n3[3 * i + k] = (double)n[3 * i + k];
    }
}

template <class T>
void ccd_response(int nv, T* X, const T* XT, T* V, const T* tI, const float* normals, hipStream_t stream) {
    hipLaunchKernelGGL(ccd_response_kernel<T>, dim3(blocks_for(nv)), dim3(kTpbC), 0, stream, nv, X, XT, V, tI, normals);
    SBS_HIP(hipGetLastError());
}

template class CollisionDetector<float>;
template class CollisionDetector<double>;
template void ccd_response<float>(int, float*, const float*, float*, const float*, const float*, hipStream_t);
template void ccd_response<double>(int, double*, const double*, double*, const double*, const float*, hipStream_t);
template void FixedBodies::handle<float>(float*, float*, int, float, float, hipStream_t) const;
template void FixedBodies::handle<double>(double*, double*, int, double, double, hipStream_t) const;

}  // namespace sbs
