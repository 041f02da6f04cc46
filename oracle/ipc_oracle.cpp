// CPU oracle for the FP64 IPC time step (incremental potential contact):
// Newton on E(x) = 1/2 |x - x~|_M^2 + h^2 (gravity + corotated elasticity +
// squared-distance barrier), with a CCD-limited backtracking line search.
//
// TEST INFRASTRUCTURE ONLY.  Plain C++17, used by tests/ as the checker; never
// linked into, loaded by, or called from the product library.
//
// What it restates (reference file:line):
//   * IPCSolver::SolverStep, SearchDirection, EndCondition,
//     DOFElimination                          src/simulation/solver/IPC/ipc.cu:142-258
//   * IPEnergy Val / GradientHessian / UpdateKappa / InitStepSize
//                                             src/simulation/energy/ip.cu:49-128
//       (the elastic term is CorotatedEnergy, ip.cu:22, or, selected, its
//       one-line swap NeoHookean08Energy)
//   * InertiaEnergy                           src/simulation/energy/inertia.cu:55-112
//   * GravityEnergy (g = 9.8 along -y)        src/simulation/energy/gravity.cu:1-47, gravity.h:14
//   * CorotatedEnergy Val / GradHessianKern   src/simulation/energy/corotated.cu:100-209
//       (the Newton Hessian of the elastic term is not made positive definite there)
//   * NeoHookean08Energy Val / GradHessianKern src/simulation/energy/neohookean08.cu:99-206
//       (the Hessian it assembles omits the dC/dF term it forms, H3, and is
//       not made positive definite)
//   * BarrierEnergy Val / gradHessianKern     src/simulation/energy/barrier.cu:10-120,180-200
//   * makePD (cyclic Jacobi, clamp at 1e-6)   src/simulation/matrix.h:441-518
//   * distance types                          src/simulation/distance/distance_type.cuh:50-273
//   * squared distances point-point, point-line, point-plane, line-line and
//     their dispatch per type                 src/simulation/distance/point_triangle.cuh, edge_edge.cuh
//   * UpdateQueries (candidates, types, distances, d <= dhat kept)
//                                             src/collision/broadphase.cu:574-587
//
// Gradients and Hessians of the squared distances are formed here by
// second-order forward differentiation of the distance formulas, not by the
// closed forms the product uses, so that the two derivations check each other.
//
// Decisions where the reference is undefined or left out:
//   * the SVD is TR1690 in double precision (the reference's svd3 runs in float);
//   * the linear solve is Jacobi-PCG to a relative residual of `lin_tol`
//     (the reference's default is a sparse Cholesky; the solution is the same
//     up to that tolerance); a fixed degree of freedom gets the row and column
//     of the identity (the reference's per-entry elimination leaves a
//     diagonal equal to the number of contributing entries, which does not
//     change the solution since its right-hand side is 0);
//   * contact areas (computeInvDmV0, solverUtil.cuh:110-137, divided by the
//     degree, femSolver.cu:13-16) start from zero (the reference accumulates
//     into uninitialised memory).
//
// Fixed bodies (ImplicitBarrierEnergy, src/simulation/energy/implicitBarrier.cu):
// per vertex and body, planes then cylinders then spheres, the signed distance
// d (plane: (x - p).u; cylinder: |(I - a a^T)(x - c)| - r; sphere: |x - c| - r)
// and, when d < dhat, area * dhat * kappa/2 (s - 1) ln s with s = d/dhat and
// kappa = 100 (FixedBody::kappa, rigid.h:25; solverUtil.cuh:33-47 for the
// value, gradient and Hessian, the Hessian made positive definite per body);
// the first step is bounded by 0.9 of each vertex's distance to each body
// along -p (InitStepSize, implicitBarrier.cu:352-407).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

extern "C" {
void orc_svd3_f64(const double* A9, double* U9, double* s3, double* V9);
double orc_col_min_step(int precision, int nv, int ntri, const uint32_t* tri, const uint32_t* fathers, const void* X,
                        const void* XT, int ignore_self);
long long orc_col_proximity(int ntri, const uint32_t* tri, const uint32_t* fathers, const double* X, double bound,
                            int ignore_self, uint32_t* q);
}

namespace {

// ------------------------------------------------ second-order jets (12) --
constexpr int N = 12;
struct Jet {
    double v = 0;
    double g[N] = {};
    double h[N][N] = {};
};
Jet var(const double* x, int i) {
    Jet a;
    a.v = x[i];
    a.g[i] = 1.0;
    return a;
}
Jet operator+(const Jet& a, const Jet& b) {
    Jet c;
    c.v = a.v + b.v;
    for (int i = 0; i < N; ++i) {
        c.g[i] = a.g[i] + b.g[i];
        for (int j = 0; j < N; ++j) c.h[i][j] = a.h[i][j] + b.h[i][j];
    }
    return c;
}
Jet operator-(const Jet& a, const Jet& b) {
    Jet c;
    c.v = a.v - b.v;
    for (int i = 0; i < N; ++i) {
        c.g[i] = a.g[i] - b.g[i];
        for (int j = 0; j < N; ++j) c.h[i][j] = a.h[i][j] - b.h[i][j];
    }
    return c;
}
Jet operator*(const Jet& a, const Jet& b) {
    Jet c;
    c.v = a.v * b.v;
    for (int i = 0; i < N; ++i) {
        c.g[i] = a.v * b.g[i] + b.v * a.g[i];
        for (int j = 0; j < N; ++j)
            c.h[i][j] = a.v * b.h[i][j] + b.v * a.h[i][j] + a.g[i] * b.g[j] + b.g[i] * a.g[j];
    }
    return c;
}
Jet operator/(const Jet& a, const Jet& b) {
    Jet q;
    q.v = a.v / b.v;
    for (int i = 0; i < N; ++i) q.g[i] = (a.g[i] - q.v * b.g[i]) / b.v;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            q.h[i][j] = (a.h[i][j] - q.v * b.h[i][j] - q.g[i] * b.g[j] - b.g[i] * q.g[j]) / b.v;
    return q;
}
struct JV {
    Jet x, y, z;
};
JV jv(const double* x, int p) { return {var(x, 3 * p), var(x, 3 * p + 1), var(x, 3 * p + 2)}; }
JV operator-(const JV& a, const JV& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
Jet dot(const JV& a, const JV& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
JV cross(const JV& a, const JV& b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

// squared distances of the 12 coordinates of a query (points 0..3)
Jet d_pp(const JV& a, const JV& b) {
    const JV d = a - b;
    return dot(d, d);
}
Jet d_pl(const JV& p, const JV& e0, const JV& e1) {
    const JV c = cross(e0 - p, e1 - p);
    const JV e = e1 - e0;
    return dot(c, c) / dot(e, e);
}
Jet d_ppl(const JV& p, const JV& t0, const JV& t1, const JV& t2) {
    const JV n = cross(t1 - t0, t2 - t0);
    const Jet s = dot(p - t0, n);
    return (s * s) / dot(n, n);
}
Jet d_ll(const JV& ea0, const JV& ea1, const JV& eb0, const JV& eb1) {
    const JV n = cross(ea1 - ea0, eb1 - eb0);
    const Jet s = dot(eb0 - ea0, n);
    return (s * s) / dot(n, n);
}

enum DType { P_T0, P_T1, P_T2, P_E0, P_E1, P_E2, P_T, EA0_EB0, EA0_EB1, EA1_EB0, EA1_EB1, EA_EB0, EA_EB1, EA0_EB,
             EA1_EB, EA_EB };

// query type 1 (VF: point, triangle) or 2 (EE: edge a, edge b); x: 12 coordinates
Jet distance_jet(int qtype, int dt, const double* x) {
    const JV a = jv(x, 0), b = jv(x, 1), c = jv(x, 2), d = jv(x, 3);
    if (qtype == 1) {
        switch (dt) {
        case P_T0: return d_pp(a, b);
        case P_T1: return d_pp(a, c);
        case P_T2: return d_pp(a, d);
        case P_E0: return d_pl(a, b, c);
        case P_E1: return d_pl(a, c, d);
        case P_E2: return d_pl(a, d, b);
        default: return d_ppl(a, b, c, d);
        }
    }
    switch (dt) {
    case EA0_EB0: return d_pp(a, c);
    case EA0_EB1: return d_pp(a, d);
    case EA1_EB0: return d_pp(b, c);
    case EA1_EB1: return d_pp(b, d);
    case EA_EB0: return d_pl(c, a, b);
    case EA_EB1: return d_pl(d, a, b);
    case EA0_EB: return d_pl(a, c, d);
    case EA1_EB: return d_pl(b, c, d);
    default: return d_ll(a, b, c, d);
    }
}

// ------------------------------------------------------ distance types --
struct D3 {
    double x, y, z;
};
D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
D3 ld3(const double* x, int p) { return {x[3 * p], x[3 * p + 1], x[3 * p + 2]}; }

// Coordinates of p - t0 in the basis (t1 - t0, (t1 - t0) x normal), least squares.
void projected(D3 p, D3 t0, D3 t1, D3 normal, double& a, double& b) {
    const D3 b0 = t1 - t0, b1 = cross(b0, normal), r = p - t0;
    const double m00 = dot(b0, b0), m01 = dot(b0, b1), m11 = dot(b1, b1);
    const double inv = 1.0 / (m00 * m11 - m01 * m01);
    const double r0 = dot(b0, r), r1 = dot(b1, r);
    a = (m11 * r0 - m01 * r1) * inv;
    b = (-m01 * r0 + m00 * r1) * inv;
}

int point_triangle_type(D3 p, D3 t0, D3 t1, D3 t2) {
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

int edge_edge_parallel_type(D3 ea0, D3 ea1, D3 eb0, D3 eb1) {
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

int edge_edge_type(D3 ea0, D3 ea1, D3 eb0, D3 eb1) {
    const D3 u = ea1 - ea0, v = eb1 - eb0, w = ea0 - eb0;
    const double a = dot(u, u), b = dot(u, v), c = dot(v, v), d = dot(u, w), e = dot(v, w);
    const double D = a * c - b * b;
    if (a == 0.0 && c == 0.0) return EA0_EB0;
    if (a == 0.0) return EA0_EB;
    if (c == 0.0) return EA_EB0;
    const double tol = 1.0e-20 * std::max(1.0, a * c);
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

int distance_type(int qtype, const double* x) {
    const D3 a = ld3(x, 0), b = ld3(x, 1), c = ld3(x, 2), d = ld3(x, 3);
    return qtype == 1 ? point_triangle_type(a, b, c, d) : edge_edge_type(a, b, c, d);
}

// --------------------------------------------------------------- makePD --
// Cyclic Jacobi eigen-decomposition (pairs p < q in order, rotations skipped
// when |a_pq| < eps, up to maxSweeps sweeps), eigenvalues below 1e-6 raised to
// 1e-6, then V diag V^T.
void make_pd(int n, double* M, int max_sweeps = 20, double eps = 1e-9) {
    std::vector<double> V((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
    auto m = [&](int i, int j) -> double& { return M[(size_t)i * n + j]; };
    auto v = [&](int i, int j) -> double& { return V[(size_t)i * n + j]; };
    for (int sweep = 0; sweep < max_sweeps; ++sweep) {
        bool converged = true;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = m(p, q);
                if (std::fabs(apq) < eps) continue;
                converged = false;
                const double app = m(p, p), aqq = m(q, q);
                const double theta = 0.5 * (aqq - app) / apq;
                double t;
                if (std::fabs(theta) > 1e10) t = 0.5 / theta;
                else t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(1.0 + theta * theta));
                const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
                m(p, q) = 0.0;
                m(q, p) = 0.0;
                m(p, p) = c * c * app - 2.0 * s * c * apq + s * s * aqq;
                m(q, q) = s * s * app + 2.0 * s * c * apq + c * c * aqq;
                for (int k = 0; k < n; ++k) {
                    if (k == p || k == q) continue;
                    const double akp = m(k, p), akq = m(k, q);
                    m(k, p) = c * akp - s * akq;
                    m(p, k) = m(k, p);
                    m(k, q) = s * akp + c * akq;
                    m(q, k) = m(k, q);
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = v(k, p), vkq = v(k, q);
                    v(k, p) = c * vkp - s * vkq;
                    v(k, q) = s * vkp + c * vkq;
                }
            }
        if (converged) break;
    }
    std::vector<double> lam(n);
    for (int i = 0; i < n; ++i) lam[i] = std::max(m(i, i), 1e-6);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double s = 0.0;
            for (int k = 0; k < n; ++k) s += v(i, k) * lam[k] * v(j, k);
            m(i, j) = s;
        }
}

// -------------------------------------------------------------- barrier --
double barrier(double d, double dhat, double kappa) {
    const double s = d / (dhat * dhat);
    return 0.5 * dhat * kappa * 0.125 * (s - 1) * std::log(s);
}
double barrier_d1(double d, double dhat, double kappa) {
    const double dh2 = dhat * dhat, s = d / dh2;
    return 0.5 * dhat * (kappa / 8 * (std::log(s) / dh2 + (s - 1) / d));
}
double barrier_d2(double d, double dhat, double kappa) {
    const double dh2 = dhat * dhat;
    return 0.5 * dhat * (kappa / (8 * d * d * dh2) * (d + dh2));
}

// ------------------------------------------------------------ the solver --
// det F and the cofactor matrix C = dJ/dF: column c of C is the cross
// product of columns c+1 and c+2 of F (glm::cross(F[1], F[2]) and its
// rotations in neohookean08.cu:21-24).
double det3(const double F[3][3]) {
    return F[0][0] * (F[1][1] * F[2][2] - F[1][2] * F[2][1]) - F[0][1] * (F[1][0] * F[2][2] - F[1][2] * F[2][0]) +
           F[0][2] * (F[1][0] * F[2][1] - F[1][1] * F[2][0]);
}
void cofactor(const double F[3][3], double C[3][3]) {
    for (int c = 0; c < 3; ++c) {
        const int a = (c + 1) % 3, b = (c + 2) % 3;
        C[0][c] = F[1][a] * F[2][b] - F[2][a] * F[1][b];
        C[1][c] = F[2][a] * F[0][b] - F[0][a] * F[2][b];
        C[2][c] = F[0][a] * F[1][b] - F[1][a] * F[0][b];
    }
}

struct Query {
    uint32_t type, v[4];
    int dt;
    double d;
};

struct Ipc {
    int nv = 0, nt = 0, ntri = 0;
    std::vector<double> X, V, XT, mass, mu, lam, V0, Dm;  // Dm: Dm^-1 row-major per tet
    std::vector<uint32_t> tet, tri, fathers;
    std::vector<char> fixed;
    double dt = 0.01, damp = 0.999, dhat = 1e-3, tol = 5e-2, lin_tol = 1e-13, kappa = 100;
    int max_iter = 100, lin_max_iter = 20000;
    int ignore_self = 0;  // the detector's flag: UpdateQueries uses the last value set
    int elastic = 0;      // 0 CorotatedEnergy (ip.cu:22), 1 NeoHookean08Energy
    std::vector<Query> qs;
    // fixed bodies: type (0 sphere, 1 plane, 2 cylinder) and 8 floats each
    // (p[3], u[3], r, 0), and each vertex's contact area
    std::vector<int> fb_type;
    std::vector<float> fb_prm;
    std::vector<double> area;
    // statistics of the last step
    int newton_iters = 0, line_halvings = 0, lin_iters = 0, ok = 1;
    double last_alpha0 = 1.0;

    void gather(const std::vector<double>& x, const Query& q, double* c) const {
        for (int j = 0; j < 4; ++j)
            for (int k = 0; k < 3; ++k) c[3 * j + k] = x[3 * (size_t)q.v[j] + k];
    }

    void update_queries(const std::vector<double>& x) {
        const long long n = orc_col_proximity(ntri, tri.data(), fathers.data(), x.data(), 2 * dhat, ignore_self,
                                              nullptr);
        std::vector<uint32_t> raw(5 * (size_t)n);
        orc_col_proximity(ntri, tri.data(), fathers.data(), x.data(), 2 * dhat, ignore_self, raw.data());
        qs.clear();
        for (long long k = 0; k < n; ++k) {
            Query q;
            q.type = raw[5 * k];
            for (int j = 0; j < 4; ++j) q.v[j] = raw[5 * k + 1 + j];
            double c[12];
            gather(x, q, c);
            q.dt = distance_type((int)q.type, c);
            q.d = distance_jet((int)q.type, q.dt, c).v;
            if (q.d > dhat) continue;
            qs.push_back(q);
        }
    }

    void deformation(const std::vector<double>& x, int t, double F[3][3]) const {
        const uint32_t* tv = &tet[4 * (size_t)t];
        double Ds[3][3];
        for (int r = 0; r < 3; ++r)
            for (int j = 0; j < 3; ++j) Ds[r][j] = x[3 * (size_t)tv[j + 1] + r] - x[3 * (size_t)tv[0] + r];
        const double* m = &Dm[9 * (size_t)t];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) F[r][c] = Ds[r][0] * m[c] + Ds[r][1] * m[3 + c] + Ds[r][2] * m[6 + c];
    }

    static void svd(const double F[3][3], double U[3][3], double s[3], double Vm[3][3]) {
        orc_svd3_f64(&F[0][0], &U[0][0], s, &Vm[0][0]);
    }

    // Fixed-body barrier of vertex v at x: energy, gradient and the sum over
    // bodies of the made-PD 3x3 Hessians (any may be null), unscaled.
    void fixed_terms(const std::vector<double>& x, int v, double* e, double* g, double* H) const {
        const double kap = 1e2;
        const double* xv = &x[3 * (size_t)v];
        for (int pass : {1, 2, 0}) {  // planes, cylinders, spheres
            for (size_t j = 0; j < fb_type.size(); ++j) {
                if (fb_type[j] != pass) continue;
                const float* q = &fb_prm[8 * j];
                const double p[3] = {q[0], q[1], q[2]}, u[3] = {q[3], q[4], q[5]}, r = q[6];
                double n[3], d;
                if (pass == 1) {
                    d = 0.0;
                    for (int k = 0; k < 3; ++k) d += (xv[k] - p[k]) * u[k];
                    for (int k = 0; k < 3; ++k) n[k] = u[k];
                } else {
                    double w[3] = {xv[0] - p[0], xv[1] - p[1], xv[2] - p[2]};
                    if (pass == 2) {  // remove the axial part
                        const double au = w[0] * u[0] + w[1] * u[1] + w[2] * u[2];
                        for (int k = 0; k < 3; ++k) w[k] -= au * u[k];
                    }
                    const double len = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
                    d = len - r;
                    for (int k = 0; k < 3; ++k) n[k] = w[k] / len;
                }
                if (!(d < dhat)) continue;
                const double s_ = d / dhat, A = area[v];
                if (e) *e += A * dhat * kap * 0.5 * (s_ - 1) * std::log(s_);
                if (g) {
                    const double c = A * dhat * (kap / 2 * (std::log(s_) / dhat + (s_ - 1) / d));
                    for (int k = 0; k < 3; ++k) g[k] += c * n[k];
                }
                if (H) {
                    const double c = A * dhat * kap / (2 * d * d * dhat) * (d + dhat);
                    double M[9];
                    for (int a = 0; a < 3; ++a)
                        for (int b2 = 0; b2 < 3; ++b2) M[3 * a + b2] = c * (n[a] * n[b2]);
                    make_pd(3, M);
                    for (int k = 0; k < 9; ++k) H[k] += M[k];
                }
            }
        }
    }

    // InitStepSize of the fixed bodies: the largest alpha <= 1 that keeps
    // 10% of every vertex's distance to every body it approaches along -p.
    double fixed_step(const std::vector<double>& x, const std::vector<double>& pdir) const {
        double alpha = 1.0;
        for (int v = 0; v < nv; ++v) {
            const double* xv = &x[3 * (size_t)v];
            const double lp[3] = {-pdir[3 * (size_t)v], -pdir[3 * (size_t)v + 1], -pdir[3 * (size_t)v + 2]};
            for (size_t j = 0; j < fb_type.size(); ++j) {
                const float* q = &fb_prm[8 * j];
                const double p[3] = {q[0], q[1], q[2]}, u[3] = {q[3], q[4], q[5]}, r = q[6];
                double cand = NAN;
                if (fb_type[j] == 1) {
                    const double pn = lp[0] * u[0] + lp[1] * u[1] + lp[2] * u[2];
                    if (pn < 0) {
                        double dist = 0.0;
                        for (int k = 0; k < 3; ++k) dist += u[k] * (xv[k] - p[k]);
                        cand = 0.9 * dist / -pn;
                    }
                } else {
                    double w[3] = {xv[0] - p[0], xv[1] - p[1], xv[2] - p[2]}, pp[3] = {lp[0], lp[1], lp[2]};
                    if (fb_type[j] == 2) {
                        const double aw = w[0] * u[0] + w[1] * u[1] + w[2] * u[2];
                        const double ap = pp[0] * u[0] + pp[1] * u[1] + pp[2] * u[2];
                        for (int k = 0; k < 3; ++k) {
                            w[k] -= aw * u[k];
                            pp[k] -= ap * u[k];
                        }
                    }
                    // the reference tests the approach with the unprojected p for cylinders
                    const double pn = (fb_type[j] == 2 ? lp[0] * w[0] + lp[1] * w[1] + lp[2] * w[2]
                                                       : pp[0] * w[0] + pp[1] * w[1] + pp[2] * w[2]);
                    if (pn < 0) {
                        const double a2 = pp[0] * pp[0] + pp[1] * pp[1] + pp[2] * pp[2];
                        const double b = w[0] * pp[0] + w[1] * pp[1] + w[2] * pp[2];
                        const double c = w[0] * w[0] + w[1] * w[1] + w[2] * w[2] - r * r;
                        cand = 0.9 * (-b - std::sqrt(b * b - a2 * c)) / a2;
                    }
                }
                alpha = std::fmin(alpha, cand);  // a ray that misses (NaN) leaves alpha
            }
        }
        return alpha;
    }

    double energy(const std::vector<double>& x) const {
        const double h2 = dt * dt;
        double inertia = 0.0, grav = 0.0, elastic = 0.0, bar = 0.0;
        for (int v = 0; v < nv; ++v) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) {
                const double d = x[3 * (size_t)v + k] - XT[3 * (size_t)v + k];
                s += d * d;
            }
            inertia += mass[v] * s;
            grav += x[3 * (size_t)v + 1] * mass[v];
        }
        inertia *= 0.5;
        grav *= 9.8;
        for (int t = 0; t < nt; ++t) {
            double F[3][3], U[3][3], s[3], Vm[3][3];
            deformation(x, t, F);
            if (this->elastic == 1) {
                // NeoHookean08Energy::Val (neohookean08.cu:193-206)
                double f2 = 0.0;
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) f2 += F[r][c] * F[r][c];
                const double lj = std::log(det3(F));
                elastic += V0[t] * (mu[t] * 0.5 * (f2 - 3) - mu[t] * lj + 0.5 * lam[t] * lj * lj);
                continue;
            }
            svd(F, U, s, Vm);
            const double a = s[0] - 1, b = s[1] - 1, c = s[2] - 1, tr = a + b + c;
            elastic += V0[t] * (mu[t] * (a * a + b * b + c * c) + 0.5 * lam[t] * tr * tr);
        }
        for (const Query& q : qs)
            if (q.d < dhat * dhat) bar += barrier(q.d, dhat, kappa);
        double fixedb = 0.0;
        if (!fb_type.empty())
            for (int v = 0; v < nv; ++v) fixed_terms(x, v, &fixedb, nullptr, nullptr);
        return inertia + h2 * (grav + elastic + fixedb + bar);
    }

    // gradient of the tet's elastic energy (scaled by coef) and, optionally, its
    // 12x12 Newton Hessian (row-major, vertex-major coordinates)
    void elastic_local(const std::vector<double>& x, int t, double coef, double* g12, double* H) const {
        if (elastic == 1) {
            nh_local(x, t, coef, g12, H);
            return;
        }
        double F[3][3], U[3][3], s[3], Vm[3][3];
        deformation(x, t, F);
        svd(F, U, s, Vm);
        double R[3][3];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) R[r][c] = U[r][0] * Vm[c][0] + U[r][1] * Vm[c][1] + U[r][2] * Vm[c][2];
        double trRF = 0.0;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) trRF += R[r][c] * F[r][c];
        const double m_ = mu[t], l_ = lam[t], w = coef * V0[t];
        double P[3][3];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) P[r][c] = 2 * m_ * (F[r][c] - R[r][c]) + l_ * (trRF - 3.0) * R[r][c];
        // dE/dF : dF/dx ; F[r][c] depends on x_{j+1,r} with Dm[j][c] and on x_{0,r} with -sum_j Dm[j][c]
        const double* dm = &Dm[9 * (size_t)t];
        double G[9][12] = {};  // d vec(F) / d x, vec index 3c + r (column-major)
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                double s0 = 0.0;
                for (int j = 0; j < 3; ++j) {
                    G[3 * c + r][3 * (j + 1) + r] = dm[3 * j + c];
                    s0 += dm[3 * j + c];
                }
                G[3 * c + r][r] = -s0;
            }
        for (int i = 0; i < 12; ++i) {
            double acc = 0.0;
            for (int e = 0; e < 9; ++e) acc += P[e % 3][e / 3] * G[e][i];
            g12[i] = w * acc;
        }
        if (!H) return;
        // d2Psi/dF2 of CorotatedEnergy::GradHessianKern (vec index 3c + r)
        double g1[9], tw[3][9];
        for (int e = 0; e < 9; ++e) g1[e] = R[e % 3][e / 3];
        // twist matrices (1/sqrt 2) U T_k V^T with T_0 = e0 e1^T - e1 e0^T,
        // T_1 = e1 e2^T - e2 e1^T, T_2 = e2 e0^T - e0 e2^T (row, column)
        const int ta[3] = {0, 1, 2}, tb[3] = {1, 2, 0};
        for (int k = 0; k < 3; ++k)
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c)
                    tw[k][3 * c + r] = (U[r][ta[k]] * Vm[c][tb[k]] - U[r][tb[k]] * Vm[c][ta[k]]) / std::sqrt(2.0);
        double lk[3] = {2 / (s[0] + s[1]), 2 / (s[1] + s[2]), 2 / (s[0] + s[2])};
        if (s[0] + s[1] < 2) lk[0] = 1;
        if (s[1] + s[2] < 2) lk[1] = 1;
        if (s[0] + s[2] < 2) lk[2] = 1;
        const double cc = l_ * (s[0] + s[1] + s[2]) - 2 * m_ - 3 * l_;
        double D[9][9];
        for (int a = 0; a < 9; ++a)
            for (int b = 0; b < 9; ++b) {
                double v = (a == b ? 2 * m_ : 0.0) + l_ * g1[a] * g1[b];
                for (int k = 0; k < 3; ++k) v += cc * lk[k] * tw[k][a] * tw[k][b];
                D[a][b] = v * w;
            }
        for (int i = 0; i < 12; ++i)
            for (int j = 0; j < 12; ++j) {
                double acc = 0.0;
                for (int a = 0; a < 9; ++a) {
                    if (G[a][i] == 0.0) continue;
                    double row = 0.0;
                    for (int b = 0; b < 9; ++b) row += D[a][b] * G[b][j];
                    acc += G[a][i] * row;
                }
                H[12 * i + j] = acc;
            }
    }

    // NeoHookean08Energy::GradHessianKern (neohookean08.cu:99-183): with
    // J = det F and C = dJ/dF (cofactor matrix), P = mu (F - C/J) +
    // lambda ln J C/J, and the Hessian that kernel assembles,
    // mu I + (lambda (1 - ln J) + mu)/J^2 vec C vec C^T, mapped through
    // dvec(F)/dx (the H3 term it forms is never added; no projection).
    void nh_local(const std::vector<double>& x, int t, double coef, double* g12, double* H) const {
        double F[3][3];
        deformation(x, t, F);
        double C[3][3];
        cofactor(F, C);
        const double J = det3(F), lj = std::log(J);
        const double m_ = mu[t], l_ = lam[t], w = coef * V0[t];
        double P[3][3];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) P[r][c] = m_ * (F[r][c] - C[r][c] / J) + l_ * lj * C[r][c] / J;
        const double* dm = &Dm[9 * (size_t)t];
        double G[9][12] = {};  // d vec(F) / d x, vec index 3c + r
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                double s0 = 0.0;
                for (int j = 0; j < 3; ++j) {
                    G[3 * c + r][3 * (j + 1) + r] = dm[3 * j + c];
                    s0 += dm[3 * j + c];
                }
                G[3 * c + r][r] = -s0;
            }
        for (int i = 0; i < 12; ++i) {
            double acc = 0.0;
            for (int e = 0; e < 9; ++e) acc += P[e % 3][e / 3] * G[e][i];
            g12[i] = w * acc;
        }
        if (!H) return;
        const double alpha = (l_ * (1 - lj) + m_) / (J * J);
        double D[9][9];
        for (int a = 0; a < 9; ++a)
            for (int b = 0; b < 9; ++b)
                D[a][b] = w * ((a == b ? m_ : 0.0) + alpha * C[a % 3][a / 3] * C[b % 3][b / 3]);
        for (int i = 0; i < 12; ++i)
            for (int j = 0; j < 12; ++j) {
                double acc = 0.0;
                for (int a = 0; a < 9; ++a) {
                    if (G[a][i] == 0.0) continue;
                    double row = 0.0;
                    for (int b = 0; b < 9; ++b) row += D[a][b] * G[b][j];
                    acc += G[a][i] * row;
                }
                H[12 * i + j] = acc;
            }
    }

    void barrier_local(const std::vector<double>& x, const Query& q, double coef, double kap, double* g12,
                       double* H) const {
        double c[12];
        gather(x, q, c);
        const Jet d = distance_jet((int)q.type, q.dt, c);
        const double d1 = barrier_d1(q.d, dhat, kap), d2 = barrier_d2(q.d, dhat, kap);
        for (int i = 0; i < 12; ++i) g12[i] = coef * d1 * d.g[i];
        if (!H) return;
        for (int i = 0; i < 12; ++i)
            for (int j = 0; j < 12; ++j) H[12 * i + j] = coef * (d2 * d.g[i] * d.g[j] + d1 * d.h[i][j]);
        make_pd(12, H);
    }

    // UpdateKappa (ip.cu:77-115): kappa = max(100, max|grad E without barrier| / max|grad B(kappa=1)|)
    void update_kappa(const std::vector<double>& x) {
        const double h2 = dt * dt;
        std::vector<double> g(3 * (size_t)nv, 0.0);
        for (int v = 0; v < nv; ++v) {
            for (int k = 0; k < 3; ++k) g[3 * (size_t)v + k] += mass[v] * (x[3 * (size_t)v + k] - XT[3 * (size_t)v + k]);
            g[3 * (size_t)v + 1] += h2 * 9.8 * mass[v];
        }
        double g12[12];
        for (int t = 0; t < nt; ++t) {
            elastic_local(x, t, h2, g12, nullptr);
            for (int j = 0; j < 4; ++j)
                for (int k = 0; k < 3; ++k) g[3 * (size_t)tet[4 * (size_t)t + j] + k] += g12[3 * j + k];
        }
        if (!fb_type.empty())
            for (int v = 0; v < nv; ++v) {
                double gf[3] = {0, 0, 0};
                fixed_terms(x, v, nullptr, gf, nullptr);
                for (int k = 0; k < 3; ++k) g[3 * (size_t)v + k] += h2 * gf[k];
            }
        double ge = 0.0;
        for (double a : g) ge = std::max(ge, std::fabs(a));
        std::fill(g.begin(), g.end(), 0.0);
        for (const Query& q : qs) {
            if (q.d > dhat * dhat) continue;
            barrier_local(x, q, h2, 1.0, g12, nullptr);
            for (int j = 0; j < 4; ++j)
                for (int k = 0; k < 3; ++k) g[3 * (size_t)q.v[j] + k] += g12[3 * j + k];
        }
        double gb = 0.0;
        for (double a : g) gb = std::max(gb, std::fabs(a));
        kappa = gb > 1e-9 ? std::max(100.0, ge / gb) : 100.0;
    }

    // gradient and Newton Hessian at x (CSR over 3 nv), fixed DOFs eliminated
    void gradient_hessian(const std::vector<double>& x, std::vector<double>& g, std::vector<int>& rp,
                          std::vector<int>& ci, std::vector<double>& val) const {
        const double h2 = dt * dt;
        const int n = 3 * nv;
        g.assign(n, 0.0);
        std::vector<std::pair<long long, double>> coo;
        coo.reserve((size_t)n + 144 * ((size_t)nt + qs.size()));
        for (int v = 0; v < nv; ++v)
            for (int k = 0; k < 3; ++k) {
                const int r = 3 * v + k;
                g[r] += mass[v] * (x[r] - XT[r]);
                coo.push_back({(long long)r * n + r, mass[v]});
            }
        for (int v = 0; v < nv; ++v) g[3 * (size_t)v + 1] += h2 * 9.8 * mass[v];
        double g12[12], H[144];
        auto scatter = [&](const uint32_t* vv) {
            for (int a = 0; a < 4; ++a)
                for (int k = 0; k < 3; ++k) {
                    g[3 * (size_t)vv[a] + k] += g12[3 * a + k];
                    for (int b = 0; b < 4; ++b)
                        for (int l = 0; l < 3; ++l)
                            coo.push_back({(long long)(3 * vv[a] + k) * n + (3 * vv[b] + l), H[12 * (3 * a + k) + 3 * b + l]});
                }
        };
        for (int t = 0; t < nt; ++t) {
            elastic_local(x, t, h2, g12, H);
            scatter(&tet[4 * (size_t)t]);
        }
        if (!fb_type.empty())
            for (int v = 0; v < nv; ++v) {
                double gf[3] = {0, 0, 0}, Hf[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                fixed_terms(x, v, nullptr, gf, Hf);
                for (int k = 0; k < 3; ++k) {
                    g[3 * (size_t)v + k] += h2 * gf[k];
                    for (int l = 0; l < 3; ++l)
                        coo.push_back({(long long)(3 * v + k) * n + (3 * v + l), h2 * Hf[3 * k + l]});
                }
            }
        for (const Query& q : qs) {
            if (q.d > dhat * dhat) continue;
            barrier_local(x, q, h2, kappa, g12, H);
            scatter(q.v);
        }
        for (int v = 0; v < nv; ++v)
            if (fixed[v])
                for (int k = 0; k < 3; ++k) g[3 * (size_t)v + k] = 0.0;
        std::sort(coo.begin(), coo.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
        rp.assign(n + 1, 0);
        ci.clear();
        val.clear();
        for (size_t k = 0; k < coo.size();) {
            const long long key = coo[k].first;
            double s = 0.0;
            while (k < coo.size() && coo[k].first == key) s += coo[k++].second;
            const int r = (int)(key / n), c = (int)(key % n);
            if (fixed[r / 3] || fixed[c / 3]) s = (r == c) ? 1.0 : 0.0;
            ci.push_back(c);
            val.push_back(s);
            rp[r + 1]++;
        }
        for (int r = 0; r < n; ++r) rp[r + 1] += rp[r];
    }

    // Jacobi-PCG: H p = g to |r| <= lin_tol |g|
    void solve(const std::vector<int>& rp, const std::vector<int>& ci, const std::vector<double>& val,
               const std::vector<double>& b, std::vector<double>& x) {
        const int n = (int)b.size();
        x.assign(n, 0.0);
        std::vector<double> r = b, z(n), p(n), q(n), dinv(n, 1.0);
        for (int i = 0; i < n; ++i)
            for (int k = rp[i]; k < rp[i + 1]; ++k)
                if (ci[k] == i && val[k] != 0.0) dinv[i] = 1.0 / val[k];
        auto nrm = [](const std::vector<double>& a) {
            double s = 0.0;
            for (double v : a) s += v * v;
            return std::sqrt(s);
        };
        const double bn = nrm(b);
        lin_iters = 0;
        if (bn == 0.0) return;
        for (int i = 0; i < n; ++i) z[i] = dinv[i] * r[i];
        p = z;
        double rz = 0.0;
        for (int i = 0; i < n; ++i) rz += r[i] * z[i];
        for (int it = 0; it < lin_max_iter; ++it) {
            for (int i = 0; i < n; ++i) {
                double s = 0.0;
                for (int k = rp[i]; k < rp[i + 1]; ++k) s += val[k] * p[ci[k]];
                q[i] = s;
            }
            double pq = 0.0;
            for (int i = 0; i < n; ++i) pq += p[i] * q[i];
            if (!(pq > 0.0)) break;  // not positive definite along p: keep the iterate
            const double alpha = rz / pq;
            for (int i = 0; i < n; ++i) {
                x[i] += alpha * p[i];
                r[i] -= alpha * q[i];
            }
            ++lin_iters;
            if (nrm(r) <= lin_tol * bn) break;
            for (int i = 0; i < n; ++i) z[i] = dinv[i] * r[i];
            double rz1 = 0.0;
            for (int i = 0; i < n; ++i) rz1 += r[i] * z[i];
            const double beta = rz1 / rz;
            rz = rz1;
            for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        }
    }

    std::vector<double> search_direction() {
        std::vector<double> g, val, p;
        std::vector<int> rp, ci;
        gradient_hessian(X, g, rp, ci, val);
        solve(rp, ci, val, g, p);
        return p;
    }

    double min_step(const std::vector<double>& x, const std::vector<double>& xt) {
        ignore_self = 1;
        return orc_col_min_step(64, nv, ntri, tri.data(), fathers.data(), x.data(), xt.data(), 1);
    }

    bool step() {
        const double h = dt;
        newton_iters = 0;
        line_halvings = 0;
        const std::vector<double> xn = X;
        for (size_t i = 0; i < X.size(); ++i) XT[i] = X[i] + h * V[i] * damp;
        update_queries(X);
        update_kappa(X);
        double E_last = energy(X);
        if (!std::isfinite(E_last)) return ok = 0, false;
        std::vector<double> p = search_direction(), xt(X.size());
        auto pinf = [&]() {
            double m = 0.0;
            for (double v : p) m = std::max(m, std::fabs(v));
            return m / h;
        };
        while (!(pinf() < tol)) {
            if (++newton_iters > max_iter) return ok = 0, false;
            for (size_t i = 0; i < X.size(); ++i) xt[i] = X[i] - p[i];
            const double toi = min_step(X, xt);
            const double af = fb_type.empty() ? 1.0 : fixed_step(X, p);
            double alpha = 0.95 * std::min(af, 0.9 * toi);
            alpha = alpha < 1e-12 ? 0.0 : std::min(1.0, alpha);
            if (newton_iters == 1) last_alpha0 = alpha;
            while (true) {
                for (size_t i = 0; i < X.size(); ++i) xt[i] = X[i] - alpha * p[i];
                update_queries(xt);
                const double E = energy(xt);
                if (!std::isfinite(E)) return ok = 0, false;
                if (E > E_last) {
                    alpha /= 2;
                    ++line_halvings;
                } else {
                    break;
                }
                if (alpha < DBL_EPSILON) break;
            }
            X = xt;
            update_queries(X);
            E_last = energy(X);
            if (!std::isfinite(E_last)) return ok = 0, false;
            p = search_direction();
        }
        for (size_t i = 0; i < X.size(); ++i) V[i] = (X[i] - xn[i]) / h;
        return ok = 1, true;
    }
};

// Dm^-1 (row-major) and V0 = |det Dm| / 6 of each tet at the rest positions.
void precompute(Ipc& s) {
    s.Dm.assign(9 * (size_t)s.nt, 0.0);
    s.V0.assign(s.nt, 0.0);
    for (int t = 0; t < s.nt; ++t) {
        const uint32_t* tv = &s.tet[4 * (size_t)t];
        double D[3][3];
        for (int r = 0; r < 3; ++r)
            for (int j = 0; j < 3; ++j) D[r][j] = s.X[3 * (size_t)tv[j + 1] + r] - s.X[3 * (size_t)tv[0] + r];
        const double det = D[0][0] * (D[1][1] * D[2][2] - D[2][1] * D[1][2]) -
                           D[0][1] * (D[1][0] * D[2][2] - D[2][0] * D[1][2]) +
                           D[0][2] * (D[1][0] * D[2][1] - D[2][0] * D[1][1]);
        const double id = 1.0 / det;
        double* m = &s.Dm[9 * (size_t)t];
        m[0] = (D[1][1] * D[2][2] - D[1][2] * D[2][1]) * id;
        m[1] = (D[0][2] * D[2][1] - D[0][1] * D[2][2]) * id;
        m[2] = (D[0][1] * D[1][2] - D[0][2] * D[1][1]) * id;
        m[3] = (D[1][2] * D[2][0] - D[1][0] * D[2][2]) * id;
        m[4] = (D[0][0] * D[2][2] - D[0][2] * D[2][0]) * id;
        m[5] = (D[0][2] * D[1][0] - D[0][0] * D[1][2]) * id;
        m[6] = (D[1][0] * D[2][1] - D[1][1] * D[2][0]) * id;
        m[7] = (D[0][1] * D[2][0] - D[0][0] * D[2][1]) * id;
        m[8] = (D[0][0] * D[1][1] - D[0][1] * D[1][0]) * id;
        s.V0[t] = std::fabs(det) / 6.0;
    }
    // contact area per vertex: over its tets, a third of the areas of the
    // tet's three faces that contain it, averaged over the tets
    s.area.assign(s.nv, 0.0);
    std::vector<double> deg(s.nv, 0.0);
    for (int t = 0; t < s.nt; ++t) {
        const uint32_t* tv = &s.tet[4 * (size_t)t];
        double fa[4];  // faces (0 1 2), (1 2 3), (2 3 0), (3 0 1)
        for (int f = 0; f < 4; ++f) {
            const double* a = &s.X[3 * (size_t)tv[f]];
            const double* b = &s.X[3 * (size_t)tv[(f + 1) % 4]];
            const double* c = &s.X[3 * (size_t)tv[(f + 2) % 4]];
            const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
            const double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
            const double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
                         cz = e1[0] * e2[1] - e1[1] * e2[0];
            fa[f] = 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
        }
        for (int j = 0; j < 4; ++j) {  // vertex j lies on every face but face j+1
            double sum = 0.0;
            for (int f = 0; f < 4; ++f)
                if (f != (j + 1) % 4) sum += fa[f];
            s.area[tv[j]] += sum / 3.0;
            deg[tv[j]] += 1.0;
        }
    }
    for (int v = 0; v < s.nv; ++v)
        if (deg[v] > 0) s.area[v] /= deg[v];
}

}  // namespace

extern "C" {

// Distance type (query type 1 VF / 2 EE) of 12 coordinates (points 0..3).
int orc_ipc_distance_type(int qtype, const double* x) { return distance_type(qtype, x); }

// Squared distance of a given type with its gradient (12) and Hessian (144,
// row-major); grad / hess may be null.
double orc_ipc_distance(int qtype, int dtype, const double* x, double* grad, double* hess) {
    const Jet d = distance_jet(qtype, dtype, x);
    if (grad)
        for (int i = 0; i < 12; ++i) grad[i] = d.g[i];
    if (hess)
        for (int i = 0; i < 12; ++i)
            for (int j = 0; j < 12; ++j) hess[12 * i + j] = d.h[i][j];
    return d.v;
}

double orc_ipc_barrier(double d, double dhat, double kappa, double* d1, double* d2) {
    if (d1) *d1 = barrier_d1(d, dhat, kappa);
    if (d2) *d2 = barrier_d2(d, dhat, kappa);
    return barrier(d, dhat, kappa);
}

void orc_ipc_make_pd(int n, double* M) { make_pd(n, M); }

// An IPC solver over one scene (FP64).  X, V: 3 nv; tet: 4 nt; tri: 3 ntri;
// fixed: nv flags (DBC).  params: dt, damp, dhat, tol, lin_tol; ints: max_iter.
void* orc_ipc_create(int nv, int nt, int ntri, const double* X, const double* V, const uint32_t* tet,
                     const uint32_t* tri, const uint32_t* fathers, const double* mass, const double* mu,
                     const double* lam, const unsigned char* fixed, double dt, double damp, double dhat, double tol,
                     double lin_tol, int max_iter) {
    auto* s = new Ipc;
    s->nv = nv;
    s->nt = nt;
    s->ntri = ntri;
    s->X.assign(X, X + 3 * (size_t)nv);
    s->V.assign(V, V + 3 * (size_t)nv);
    s->XT.assign(3 * (size_t)nv, 0.0);
    s->tet.assign(tet, tet + 4 * (size_t)nt);
    s->tri.assign(tri, tri + 3 * (size_t)ntri);
    s->fathers.assign(fathers, fathers + ntri);
    s->mass.assign(mass, mass + nv);
    s->mu.assign(mu, mu + nt);
    s->lam.assign(lam, lam + nt);
    s->fixed.assign(fixed, fixed + nv);
    s->dt = dt;
    s->damp = damp;
    s->dhat = dhat;
    s->tol = tol;
    s->lin_tol = lin_tol;
    s->max_iter = max_iter;
    precompute(*s);
    return s;
}
void orc_ipc_destroy(void* h) { delete (Ipc*)h; }
// The elastic term: 0 CorotatedEnergy (ip.cu:22), 1 NeoHookean08Energy.
void orc_ipc_set_elastic(void* h, int model) { ((Ipc*)h)->elastic = model; }
// Neo-Hookean energy density, first Piola stress (9, row-major) and the
// 9x9 d2Psi/dF2 of GradHessianKern (vec index 3c + r) of one F (row-major).
double orc_nh_density(const double* F9, double mu, double lam, double* P9, double* D81) {
    double F[3][3], C[3][3];
    std::memcpy(F, F9, sizeof F);
    cofactor(F, C);
    const double J = det3(F), lj = std::log(J);
    double f2 = 0.0;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            f2 += F[r][c] * F[r][c];
            if (P9) P9[3 * r + c] = mu * (F[r][c] - C[r][c] / J) + lam * lj * C[r][c] / J;
        }
    if (D81) {
        const double alpha = (lam * (1 - lj) + mu) / (J * J);
        for (int a = 0; a < 9; ++a)
            for (int b = 0; b < 9; ++b)
                D81[9 * a + b] = (a == b ? mu : 0.0) + alpha * C[a % 3][a / 3] * C[b % 3][b / 3];
    }
    return mu * 0.5 * (f2 - 3) - mu * lj + 0.5 * lam * lj * lj;
}
// One SolverStep; returns 1 on success, 0 when it failed (the reference's `failed`).
int orc_ipc_step(void* h) { return ((Ipc*)h)->step() ? 1 : 0; }
void orc_ipc_get(void* h, double* X, double* V) {
    auto* s = (Ipc*)h;
    if (X) std::memcpy(X, s->X.data(), s->X.size() * sizeof(double));
    if (V) std::memcpy(V, s->V.data(), s->V.size() * sizeof(double));
}
void orc_ipc_set_x(void* h, const double* X) {
    auto* s = (Ipc*)h;
    std::memcpy(s->X.data(), X, s->X.size() * sizeof(double));
}
// Positions, velocities and the detector's ignore-self flag of a state taken
// from another run (the flag is set by the line search's ComputeMinStepSize,
// barrier.cu:227, so it is 1 after any step that searched a line).
void orc_ipc_set_xv(void* h, const double* X, const double* V, int ignore_self) {
    auto* s = (Ipc*)h;
    std::memcpy(s->X.data(), X, s->X.size() * sizeof(double));
    std::memcpy(s->V.data(), V, s->V.size() * sizeof(double));
    s->ignore_self = ignore_self;
}
// newton iterations, line-search halvings, last linear iterations, queries, kappa, first alpha0
void orc_ipc_stats(void* h, int* ints, double* reals) {
    auto* s = (Ipc*)h;
    ints[0] = s->newton_iters;
    ints[1] = s->line_halvings;
    ints[2] = s->lin_iters;
    ints[3] = (int)s->qs.size();
    reals[0] = s->kappa;
    reals[1] = s->last_alpha0;
}
// Energy terms at X with the current x~ (set by the last step), kappa and
// queries recomputed at X: total; and the gradient (3 nv, fixed DOFs not
// eliminated) when grad is not null.
double orc_ipc_energy(void* h, const double* X, double* grad) {
    auto* s = (Ipc*)h;
    std::vector<double> x(X, X + 3 * (size_t)s->nv);
    s->update_queries(x);
    const double e = s->energy(x);
    if (grad) {
        const double h2 = s->dt * s->dt;
        std::vector<double> g(3 * (size_t)s->nv, 0.0);
        for (int v = 0; v < s->nv; ++v) {
            for (int k = 0; k < 3; ++k) g[3 * (size_t)v + k] += s->mass[v] * (x[3 * (size_t)v + k] - s->XT[3 * (size_t)v + k]);
            g[3 * (size_t)v + 1] += h2 * 9.8 * s->mass[v];
        }
        double g12[12];
        for (int t = 0; t < s->nt; ++t) {
            s->elastic_local(x, t, h2, g12, nullptr);
            for (int j = 0; j < 4; ++j)
                for (int k = 0; k < 3; ++k) g[3 * (size_t)s->tet[4 * (size_t)t + j] + k] += g12[3 * j + k];
        }
        if (!s->fb_type.empty())
            for (int v = 0; v < s->nv; ++v) {
                double gf[3] = {0, 0, 0};
                s->fixed_terms(x, v, nullptr, gf, nullptr);
                for (int k = 0; k < 3; ++k) g[3 * (size_t)v + k] += h2 * gf[k];
            }
        for (const Query& q : s->qs) {
            if (q.d > s->dhat * s->dhat) continue;
            s->barrier_local(x, q, h2, s->kappa, g12, nullptr);
            for (int j = 0; j < 4; ++j)
                for (int k = 0; k < 3; ++k) g[3 * (size_t)q.v[j] + k] += g12[3 * j + k];
        }
        std::memcpy(grad, g.data(), g.size() * sizeof(double));
    }
    return e;
}
// Fixed bodies of the implicit barrier: n bodies, type (0 sphere, 1 plane,
// 2 cylinder) and 8 floats each (p[3], u[3], radius, 0).
void orc_ipc_set_fixed(void* h, int n, const int* types, const float* params) {
    auto* s = (Ipc*)h;
    s->fb_type.assign(types, types + n);
    s->fb_prm.assign(params, params + 8 * (size_t)n);
}
// Fixed-body barrier of one vertex at X (unscaled by h^2): energy, gradient
// (3) and the made-PD Hessian summed over bodies (9, row-major).
double orc_ipc_fixed_vertex(void* h, const double* X, int v, double* g, double* H) {
    auto* s = (Ipc*)h;
    std::vector<double> x(X, X + 3 * (size_t)s->nv);
    double e = 0.0, gl[3] = {0, 0, 0}, Hl[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    s->fixed_terms(x, v, &e, gl, Hl);
    if (g) std::memcpy(g, gl, sizeof gl);
    if (H) std::memcpy(H, Hl, sizeof Hl);
    return e;
}
// Contact area per vertex (nv).
void orc_ipc_contact_area(void* h, double* out) {
    auto* s = (Ipc*)h;
    std::memcpy(out, s->area.data(), s->area.size() * sizeof(double));
}
// Fixed-body step bound at X along -p.
double orc_ipc_fixed_step(void* h, const double* X, const double* p) {
    auto* s = (Ipc*)h;
    std::vector<double> x(X, X + 3 * (size_t)s->nv), pd(p, p + 3 * (size_t)s->nv);
    return s->fixed_step(x, pd);
}
// x~ and kappa for energy checks without a step.
void orc_ipc_set_state(void* h, const double* XT, double kappa) {
    auto* s = (Ipc*)h;
    std::memcpy(s->XT.data(), XT, s->XT.size() * sizeof(double));
    s->kappa = kappa;
}
// Queries of the last update: 5 u32 (type, v0..v3) + distance type + d.
long long orc_ipc_queries(void* h, uint32_t* q, int* dtype, double* d) {
    auto* s = (Ipc*)h;
    for (size_t k = 0; q && k < s->qs.size(); ++k) {
        q[5 * k] = s->qs[k].type;
        for (int j = 0; j < 4; ++j) q[5 * k + 1 + j] = s->qs[k].v[j];
        if (dtype) dtype[k] = s->qs[k].dt;
        if (d) d[k] = s->qs[k].d;
    }
    return (long long)s->qs.size();
}

}  // extern "C"
