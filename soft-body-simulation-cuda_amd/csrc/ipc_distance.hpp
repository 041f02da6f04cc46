// Squared distances between the primitives of a contact query, their distance
// types, and closed-form gradients and Hessians, in FP64 on the device.
//
// Reference: the squared-distance functions and their per-type dispatch of
// src/simulation/distance/point_triangle.cuh, edge_edge.cuh (the ipc-toolkit
// formulas), and the type classification of distance_type.cuh:50-273.  The
// derivatives here are derived by hand from the four base functions
//   point-point  |a - b|^2
//   point-line   |(e0 - p) x (e1 - p)|^2 / |e1 - e0|^2
//   point-plane  ((p - t0) . n)^2 / |n|^2,          n = (t1 - t0) x (t2 - t0)
//   line-line    ((eb0 - ea0) . n)^2 / |n|^2,       n = (ea1 - ea0) x (eb1 - eb0)
// with the chain rule over 3x3 blocks (no generated expressions).
#pragma once

#include <hip/hip_runtime.h>

namespace sbs {
namespace ipc {

enum DType : int { P_T0, P_T1, P_T2, P_E0, P_E1, P_E2, P_T, EA0_EB0, EA0_EB1, EA1_EB0, EA1_EB1, EA_EB0, EA_EB1,
                   EA0_EB, EA1_EB, EA_EB };

struct D3 {
    double x, y, z;
};
__host__ __device__ inline D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline D3 operator*(double s, D3 a) { return {s * a.x, s * a.y, s * a.z}; }
__host__ __device__ inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline D3 cross(D3 a, D3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__host__ __device__ inline double at(D3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

struct M3 {
    double m[3][3];
};
// C(a)_ij = eps_ijk a_k (so that C(a) b = b x a)
__host__ __device__ inline M3 eps(D3 a) {
    return {{{0.0, a.z, -a.y}, {-a.z, 0.0, a.x}, {a.y, -a.x, 0.0}}};
}
__host__ __device__ inline M3 tmul(const M3& A, const M3& B) {  // A^T B
    M3 C;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) C.m[i][j] = A.m[0][i] * B.m[0][j] + A.m[1][i] * B.m[1][j] + A.m[2][i] * B.m[2][j];
    return C;
}
__host__ __device__ inline D3 tvec(const M3& A, D3 v) {  // A^T v
    return {A.m[0][0] * v.x + A.m[1][0] * v.y + A.m[2][0] * v.z, A.m[0][1] * v.x + A.m[1][1] * v.y + A.m[2][1] * v.z,
            A.m[0][2] * v.x + A.m[1][2] * v.y + A.m[2][2] * v.z};
}

// ------------------------------------------------------- classification --
// Coordinates of p - t0 in the basis (t1 - t0, (t1 - t0) x normal), least squares.
__host__ __device__ inline void projected(D3 p, D3 t0, D3 t1, D3 normal, double& a, double& b) {
    const D3 b0 = t1 - t0, b1 = cross(b0, normal), r = p - t0;
    const double m00 = dot(b0, b0), m01 = dot(b0, b1), m11 = dot(b1, b1);
    const double inv = 1.0 / (m00 * m11 - m01 * m01);
    const double r0 = dot(b0, r), r1 = dot(b1, r);
    a = (m11 * r0 - m01 * r1) * inv;
    b = (-m01 * r0 + m00 * r1) * inv;
}

__host__ __device__ inline int point_triangle_type(D3 p, D3 t0, D3 t1, D3 t2) {
    const D3 n = cross(t1 - t0, t2 - t0);
    double a0, b0, a1, b1, a2, b2;
    projected(p, t0, t1, n, a0, b0);
    if (a0 > 0.0 && a0 < 1.0 && b0 >= 0.0) return P_E0;
    projected(p, t1, t2, n, a1, b1);
    if (a1 > 0.0 && a1 < 1.0 && b1 >= 0.0) return P_E1;
    projected(p, t2, t0, n, a2, b2);
    if (a2 > 0.0 && a2 < 1.0 && b2 >= 0.0) return P_E2;
    if (a0 <= 0.0 && a2 >= 1.0) return P_T0;
    if (a1 <= 0.0 && a0 >= 1.0) return P_T1;
    if (a2 <= 0.0 && a1 >= 1.0) return P_T2;
    return P_T;
}

__host__ __device__ inline int edge_edge_parallel_type(D3 ea0, D3 ea1, D3 eb0, D3 eb1) {
    const D3 ea = ea1 - ea0;
    const double alpha = dot(eb0 - ea0, ea) / dot(ea, ea);
    const double beta = dot(eb1 - ea0, ea) / dot(ea, ea);
    int eac, ebc;  // 0: vertex 0, 1: vertex 1, 2: interior
    if (alpha < 0) {
        eac = (0 <= beta && beta <= 1) ? 2 : 0;
        ebc = (beta <= alpha) ? 0 : (beta <= 1 ? 1 : 2);
    } else if (alpha > 1) {
        eac = (0 <= beta && beta <= 1) ? 2 : 1;
        ebc = (beta >= alpha) ? 0 : (0 <= beta ? 1 : 2);
    } else {
        eac = 2;
        ebc = 0;
    }
    return (ebc < 2 ? (eac << 1 | ebc) : (6 + eac)) + 7;
}

__host__ __device__ inline int edge_edge_type(D3 ea0, D3 ea1, D3 eb0, D3 eb1) {
    const D3 u = ea1 - ea0, v = eb1 - eb0, w = ea0 - eb0;
    const double a = dot(u, u), b = dot(u, v), c = dot(v, v), d = dot(u, w), e = dot(v, w);
    const double D = a * c - b * b;
    if (a == 0.0 && c == 0.0) return EA0_EB0;
    if (a == 0.0) return EA0_EB;
    if (c == 0.0) return EA_EB0;
    const double tol = 1.0e-20 * fmax(1.0, a * c);
    const D3 uxv = cross(u, v);
    if (dot(uxv, uxv) < tol) return edge_edge_parallel_type(ea0, ea1, eb0, eb1);
    int def = EA_EB;
    const double sN = b * e - c * d;
    double tN, tD;
    if (sN <= 0.0) {
        tN = e;
        tD = c;
        def = EA0_EB;
    } else if (sN >= D) {
        tN = e + b;
        tD = c;
        def = EA1_EB;
    } else {
        tN = a * e - b * d;
        tD = D;
        if (tN > 0.0 && tN < tD && dot(uxv, uxv) < tol) {
            if (sN < D / 2) {
                tN = e;
                tD = c;
                def = EA0_EB;
            } else {
                tN = e + b;
                tD = c;
                def = EA1_EB;
            }
        }
    }
    if (tN <= 0.0) {
        if (-d <= 0.0) return EA0_EB0;
        if (-d >= a) return EA1_EB0;
        return EA_EB0;
    }
    if (tN >= tD) {
        if ((-d + b) <= 0.0) return EA0_EB1;
        if ((-d + b) >= a) return EA1_EB1;
        return EA_EB1;
    }
    return def;
}

// query type 1: point P[0] against triangle P[1..3]; type 2: edge P[0]P[1] against edge P[2]P[3]
__host__ __device__ inline int distance_type(int qtype, const D3 (&P)[4]) {
    return qtype == 1 ? point_triangle_type(P[0], P[1], P[2], P[3]) : edge_edge_type(P[0], P[1], P[2], P[3]);
}

// The base function of a type and the query points it takes, in order.
struct Base {
    int kind;  // 0 point-point, 1 point-line, 2 point-plane, 3 line-line
    int pt[4];
};
__host__ __device__ inline Base base_of(int qtype, int dt) {
    if (qtype == 1) {
        switch (dt) {
        case P_T0: return {0, {0, 1, 0, 0}};
        case P_T1: return {0, {0, 2, 0, 0}};
        case P_T2: return {0, {0, 3, 0, 0}};
        case P_E0: return {1, {0, 1, 2, 0}};
        case P_E1: return {1, {0, 2, 3, 0}};
        case P_E2: return {1, {0, 3, 1, 0}};
        default: return {2, {0, 1, 2, 3}};
        }
    }
    switch (dt) {
    case EA0_EB0: return {0, {0, 2, 0, 0}};
    case EA0_EB1: return {0, {0, 3, 0, 0}};
    case EA1_EB0: return {0, {1, 2, 0, 0}};
    case EA1_EB1: return {0, {1, 3, 0, 0}};
    case EA_EB0: return {1, {2, 0, 1, 0}};
    case EA_EB1: return {1, {3, 0, 1, 0}};
    case EA0_EB: return {1, {0, 2, 3, 0}};
    case EA1_EB: return {1, {1, 2, 3, 0}};
    default: return {3, {0, 1, 2, 3}};
    }
}

__host__ __device__ inline double distance(int qtype, int dt, const D3 (&P)[4]) {
    const Base b = base_of(qtype, dt);
    const D3 a0 = P[b.pt[0]], a1 = P[b.pt[1]], a2 = P[b.pt[2]], a3 = P[b.pt[3]];
    switch (b.kind) {
    case 0: {
        const D3 d = a0 - a1;
        return dot(d, d);
    }
    case 1: {
        const D3 c = cross(a1 - a0, a2 - a0), e = a2 - a1;
        return dot(c, c) / dot(e, e);
    }
    case 2: {
        const D3 n = cross(a2 - a1, a3 - a1);
        const double s = dot(a0 - a1, n);
        return s * s / dot(n, n);
    }
    default: {
        const D3 n = cross(a1 - a0, a3 - a2);
        const double s = dot(a2 - a0, n);
        return s * s / dot(n, n);
    }
    }
}

// ------------------------------------------------------------ derivatives --
// Gradient (12) and Hessian of the squared distance of a query.  The Hessian
// is delivered as 3x3 blocks through hb(P, Q, M3) (block of query points P, Q),
// each (P, Q) pair delivered at most once per call of a base function term,
// to be accumulated by the caller.
//
// A base function of variables y = (y_0 .. y_{k-1}) (3D each) is evaluated
// with its y-gradient and y-Hessian blocks; each y_i is a difference of two
// query points, y_i = P[plus_i] - P[minus_i], so
//   grad_P += sum_i (P == plus_i) g_i - (P == minus_i) g_i
//   H_PQ   += sum_ij sign_i(P) sign_j(Q) H_ij.
template <class HB>
__host__ __device__ inline void scatter(int k, const int* plus, const int* minus, const D3* gy, const M3 (&Hy)[3][3],
                                        double* g, HB&& hb) {
    for (int i = 0; i < k; ++i) {
        g[3 * plus[i]] += gy[i].x; g[3 * plus[i] + 1] += gy[i].y; g[3 * plus[i] + 2] += gy[i].z;
        g[3 * minus[i]] -= gy[i].x; g[3 * minus[i] + 1] -= gy[i].y; g[3 * minus[i] + 2] -= gy[i].z;
    }
    for (int i = 0; i < k; ++i)
        for (int j = 0; j < k; ++j) {
            hb(plus[i], plus[j], Hy[i][j], 1.0);
            hb(plus[i], minus[j], Hy[i][j], -1.0);
            hb(minus[i], plus[j], Hy[i][j], -1.0);
            hb(minus[i], minus[j], Hy[i][j], 1.0);
        }
}

__host__ __device__ inline M3 outer(D3 a, D3 b) {
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = at(a, i) * at(b, j);
    return r;
}
__host__ __device__ inline M3 transpose(const M3& a) {
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[j][i];
    return r;
}
__host__ __device__ inline M3 ident(double s) { return {{{s, 0, 0}, {0, s, 0}, {0, 0, s}}}; }
__host__ __device__ inline void axpy(M3& r, double s, const M3& a) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] += s * a.m[i][j];
}

// Squared distance of the query with gradient g (12, zeroed here) and Hessian
// blocks accumulated through hb(P, Q, block, sign).
template <class HB>
__host__ __device__ inline double distance_derivatives(int qtype, int dt, const D3 (&P)[4], double* g, HB&& hb) {
    for (int i = 0; i < 12; ++i) g[i] = 0.0;
    const Base b = base_of(qtype, dt);
    const int* pt = b.pt;
    M3 Hy[3][3];
    D3 gy[3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Hy[i][j] = ident(0.0);
    if (b.kind == 0) {  // f = |y|^2, y = a - b
        const D3 y = P[pt[0]] - P[pt[1]];
        gy[0] = 2.0 * y;
        Hy[0][0] = ident(2.0);
        const int plus[1] = {pt[0]}, minus[1] = {pt[1]};
        scatter(1, plus, minus, gy, Hy, g, hb);
        return dot(y, y);
    }
    if (b.kind == 1) {  // f = |a x c|^2 / |c - a|^2, a = e0 - p, c = e1 - p
        const D3 a = P[pt[1]] - P[pt[0]], c = P[pt[2]] - P[pt[0]];
        const D3 x = cross(a, c), e = c - a;
        const double num = dot(x, x), den = dot(e, e), f = num / den;
        const M3 Na = eps(c), Nc = transpose(eps(a));  // dx/da, dx/dc (eps(a)^T = -eps(a))
        const D3 gn[2] = {2.0 * tvec(Na, x), 2.0 * tvec(Nc, x)};
        const D3 gd[2] = {-2.0 * e, 2.0 * e};
        M3 Hn[2][2];
        Hn[0][0] = tmul(Na, Na); Hn[1][1] = tmul(Nc, Nc); Hn[0][1] = tmul(Na, Nc);
        axpy(Hn[0][1], 1.0, eps(x));
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j)
                if (!(i == 1 && j == 0))
                    for (int r = 0; r < 3; ++r)
                        for (int s = 0; s < 3; ++s) Hn[i][j].m[r][s] *= 2.0;
        Hn[1][0] = transpose(Hn[0][1]);
        const double sgnd[2][2] = {{2.0, -2.0}, {-2.0, 2.0}};
        const double i1 = 1.0 / den, i2 = i1 * i1, i3 = i2 * i1;
        for (int i = 0; i < 2; ++i) {
            gy[i] = i1 * gn[i] - (num * i2) * gd[i];
            for (int j = 0; j < 2; ++j) {
                M3 h = ident(0.0);
                axpy(h, i1, Hn[i][j]);
                axpy(h, -i2, outer(gn[i], gd[j]));
                axpy(h, -i2, outer(gd[i], gn[j]));
                axpy(h, 2.0 * num * i3, outer(gd[i], gd[j]));
                axpy(h, -num * i2, ident(sgnd[i][j]));
                Hy[i][j] = h;
            }
        }
        const int plus[2] = {pt[1], pt[2]}, minus[2] = {pt[0], pt[0]};
        scatter(2, plus, minus, gy, Hy, g, hb);
        return f;
    }
    // f = s^2 / m, s = w . n, n = u x v, m = n . n
    D3 w, u, v;
    int plus[3], minus[3];
    if (b.kind == 2) {  // point-plane (p, t0, t1, t2): w = p - t0, u = t1 - t0, v = t2 - t0
        w = P[pt[0]] - P[pt[1]]; u = P[pt[2]] - P[pt[1]]; v = P[pt[3]] - P[pt[1]];
        plus[0] = pt[0]; minus[0] = pt[1];
        plus[1] = pt[2]; minus[1] = pt[1];
        plus[2] = pt[3]; minus[2] = pt[1];
    } else {  // line-line (ea0, ea1, eb0, eb1): w = eb0 - ea0, u = ea1 - ea0, v = eb1 - eb0
        w = P[pt[2]] - P[pt[0]]; u = P[pt[1]] - P[pt[0]]; v = P[pt[3]] - P[pt[2]];
        plus[0] = pt[2]; minus[0] = pt[0];
        plus[1] = pt[1]; minus[1] = pt[0];
        plus[2] = pt[3]; minus[2] = pt[2];
    }
    const D3 n = cross(u, v);
    const double s = dot(w, n), m = dot(n, n), f = s * s / m;
    const M3 Nu = eps(v), Nv = transpose(eps(u));  // dn/du, dn/dv
    const D3 gs[3] = {n, cross(v, w), cross(w, u)};
    const D3 zero = {0.0, 0.0, 0.0};
    const D3 gm[3] = {zero, 2.0 * tvec(Nu, n), 2.0 * tvec(Nv, n)};
    M3 S[3][3], Mh[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            S[i][j] = ident(0.0);
            Mh[i][j] = ident(0.0);
        }
    S[0][1] = eps(v); S[0][2] = transpose(eps(u)); S[1][2] = eps(w);
    S[1][0] = transpose(S[0][1]); S[2][0] = transpose(S[0][2]); S[2][1] = transpose(S[1][2]);
    Mh[1][1] = tmul(Nu, Nu); Mh[2][2] = tmul(Nv, Nv); Mh[1][2] = tmul(Nu, Nv);
    axpy(Mh[1][2], 1.0, eps(n));
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            Mh[1][1].m[r][c] *= 2.0; Mh[2][2].m[r][c] *= 2.0; Mh[1][2].m[r][c] *= 2.0;
        }
    Mh[2][1] = transpose(Mh[1][2]);
    const double im = 1.0 / m, im2 = im * im, im3 = im2 * im;
    for (int i = 0; i < 3; ++i) {
        gy[i] = (2.0 * s * im) * gs[i] - (s * s * im2) * gm[i];
        for (int j = 0; j < 3; ++j) {
            M3 h = ident(0.0);
            axpy(h, 2.0 * im, outer(gs[i], gs[j]));
            axpy(h, 2.0 * s * im, S[i][j]);
            axpy(h, -2.0 * s * im2, outer(gs[i], gm[j]));
            axpy(h, -2.0 * s * im2, outer(gm[i], gs[j]));
            axpy(h, 2.0 * s * s * im3, outer(gm[i], gm[j]));
            axpy(h, -s * s * im2, Mh[i][j]);
            Hy[i][j] = h;
        }
    }
    scatter(3, plus, minus, gy, Hy, g, hb);
    return f;
}

}  // namespace ipc
}  // namespace sbs
