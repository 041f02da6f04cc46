// FP64 IPC step on gfx950: energies, gradients, element Hessians with the
// positive-definite projection, matrix-free Jacobi-PCG and the Newton loop.
// Reference citations are per kernel; see ipc_engine.hpp for the layout.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <tuple>

#include <hipcub/hipcub.hpp>

#include "ipc_distance.hpp"
#include "ipc_engine.hpp"

#include <algorithm>
#include <tuple>
#include "svd3.hpp"

namespace sbs {

namespace {

constexpr int kT = 256;   // threads per block of the per-element kernels
constexpr int kQB = 16;   // queries per block of the barrier kernel (LDS: 2 x 144 doubles per query)
constexpr double kG = 9.8;  // GravityEnergy::g (gravity.h:14)

unsigned nblk(size_t n) { return (unsigned)std::max<size_t>(1, (n + kT - 1) / kT); }

using ipc::D3;

__device__ __forceinline__ D3 ldp(const double* X, uint32_t v) {
    return {X[3 * (size_t)v], X[3 * (size_t)v + 1], X[3 * (size_t)v + 2]};
}

// Block sums of NQ per-thread values, written to parts[NQ * block + k]:
// a butterfly sum in each wavefront (lane 0's total is used), then the
// wavefront totals in index order.  Every thread of the block calls it.
template <int NQ> __device__ void block_sums(const double (&v)[NQ], double* parts) {
    __shared__ double red[NQ][kT / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int k = 0; k < NQ; ++k) {
        const double t = wave_sum(v[k]);
        if (lane == 0) red[k][w] = t;
    }
    __syncthreads();
    if ((int)threadIdx.x < NQ) {
        double t = 0.0;
        for (int j = 0; j < kT / 64; ++j) t += red[threadIdx.x][j];
        parts[NQ * (size_t)blockIdx.x + threadIdx.x] = t;
    }
}

// Sums over blocks of parts[NQ * b + k] (one block of kT threads, every
// thread calls it): thread t adds blocks t, t + kT, ..., then the wavefront
// and block sums as above; every thread gets the totals, the same in every
// block that calls it.
template <int NQ> __device__ void sum_parts(int nb, const double* parts, double (&tot)[NQ]) {
    __shared__ double red[NQ][kT / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double acc[NQ];
    for (int k = 0; k < NQ; ++k) acc[k] = 0.0;
    for (int b = threadIdx.x; b < nb; b += kT)
        for (int k = 0; k < NQ; ++k) acc[k] += parts[NQ * (size_t)b + k];
    for (int k = 0; k < NQ; ++k) {
        const double t = wave_sum(acc[k]);
        if (lane == 0) red[k][w] = t;
    }
    __syncthreads();
    for (int k = 0; k < NQ; ++k) {
        double t = 0.0;
        for (int j = 0; j < kT / 64; ++j) t += red[k][j];
        tot[k] = t;
    }
}

// out[k] = sum over blocks of parts[NQ * b + k], in block order (one block).
template <int NQ> __global__ void sum_parts_kernel(int nb, const double* parts, double* out) {
    double tot[NQ];
    sum_parts<NQ>(nb, parts, tot);
    if ((int)threadIdx.x < NQ) out[threadIdx.x] = tot[threadIdx.x];
}

__global__ void max_abs_partial_kernel(size_t n, const double* v, double* parts) {
    __shared__ double red[kT];
    double m = 0.0;
    for (size_t i = (size_t)blockIdx.x * kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT) {
        const double a = fabs(v[i]);
        m = (a > m || a != a) ? a : m;  // NaN propagates
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = kT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            const double a = red[threadIdx.x + s];
            if (a > red[threadIdx.x] || a != a) red[threadIdx.x] = a;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) parts[blockIdx.x] = red[0];
}
__global__ void max_final_kernel(int nb, const double* parts, double* out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double m = 0.0;
    for (int b = 0; b < nb; ++b) {
        const double a = parts[b];
        if (a > m || a != a) m = a;
    }
    out[0] = m;
}

// FEMSolver ctor for the IPC data (femSolver.cu:7-17, solverUtil.cuh:109-137):
// Dm^-1 (row-major) and V0 = |det Dm| / 6 at the positions given at creation.
__global__ void ipc_precompute_kernel(int nt, const uint32_t* tet, const double* X, double* dm, double* v0) {
    const int t = blockIdx.x * kT + threadIdx.x;
    if (t >= nt) return;
    const uint32_t* tv = tet + 4 * (size_t)t;
    double D[3][3];
    for (int r = 0; r < 3; ++r)
        for (int j = 0; j < 3; ++j) D[r][j] = X[3 * (size_t)tv[j + 1] + r] - X[3 * (size_t)tv[0] + r];
    const double det = D[0][0] * (D[1][1] * D[2][2] - D[2][1] * D[1][2]) -
                       D[0][1] * (D[1][0] * D[2][2] - D[2][0] * D[1][2]) +
                       D[0][2] * (D[1][0] * D[2][1] - D[2][0] * D[1][1]);
    const double id = 1.0 / det;
    double* m = dm + 9 * (size_t)t;
    m[0] = (D[1][1] * D[2][2] - D[1][2] * D[2][1]) * id;
    m[1] = (D[0][2] * D[2][1] - D[0][1] * D[2][2]) * id;
    m[2] = (D[0][1] * D[1][2] - D[0][2] * D[1][1]) * id;
    m[3] = (D[1][2] * D[2][0] - D[1][0] * D[2][2]) * id;
    m[4] = (D[0][0] * D[2][2] - D[0][2] * D[2][0]) * id;
    m[5] = (D[0][2] * D[1][0] - D[0][0] * D[1][2]) * id;
    m[6] = (D[1][0] * D[2][1] - D[1][1] * D[2][0]) * id;
    m[7] = (D[0][1] * D[2][0] - D[0][0] * D[2][1]) * id;
    m[8] = (D[0][0] * D[1][1] - D[0][1] * D[1][0]) * id;
    v0[t] = fabs(det) / 6.0;
}

__global__ void ipc_fixed_kernel(int nv, const double* dbc, unsigned char* fixed) {
    const int v = blockIdx.x * kT + threadIdx.x;
    if (v < nv) fixed[v] = dbc[v] != 0.0;
}

// computeXTilde (ipc.cu:30-35) and x_n = X.
__global__ void ipc_begin_kernel(int n3, const double* X, const double* V, double h, double damp, double* XT,
                                 double* xn) {
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i >= n3) return;
    XT[i] = X[i] + h * V[i] * damp;
    xn[i] = X[i];
}

// computeXMinusAP (ipc.cu:37-44): out = X - alpha p.
__global__ void ipc_axpy_kernel(int n3, const double* X, const double* p, double alpha, double* out) {
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i < n3) out[i] = X[i] - alpha * p[i];
}

// updateVel (ipc.cu:46-53).
__global__ void ipc_velocity_kernel(int n3, const double* X, const double* xn, double h_inv, double* V) {
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i < n3) V[i] = (X[i] - xn[i]) * h_inv;
}

// GetDistanceType + ComputeDistance (distance_type.cuh:15-45) per candidate;
// count of the queries UpdateQueries keeps (d <= dhat, broadphase.cu:582).
__global__ void ipc_query_kernel(long long nq, const QueryKey* q, const double* X, double dhat, int* dtype, double* d,
                                 unsigned long long* kept) {
    const long long i = (long long)blockIdx.x * kT + threadIdx.x;
    if (i >= nq) return;
    const QueryKey k = q[i];
    const D3 P[4] = {ldp(X, k.v0), ldp(X, k.v1), ldp(X, k.v2), ldp(X, k.v3)};
    const int t = ipc::distance_type((int)k.type, P);
    const double dd = ipc::distance((int)k.type, t, P);
    dtype[i] = t;
    d[i] = dd;
    if (dd <= dhat) atomicAdd(kept, 1ull);
}

__device__ __forceinline__ void deformation(const double* X, const uint32_t* tv, const double* m, double F[3][3]) {
    double Ds[3][3];
    for (int r = 0; r < 3; ++r)
        for (int j = 0; j < 3; ++j) Ds[r][j] = X[3 * (size_t)tv[j + 1] + r] - X[3 * (size_t)tv[0] + r];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) F[r][c] = Ds[r][0] * m[c] + Ds[r][1] * m[3 + c] + Ds[r][2] * m[6 + c];
}

// b(d) = 1/2 dhat kappa / 8 (s - 1) ln s, s = d / dhat^2 (barrier.cu:11-15),
// and its first and second derivatives in d (:17-29).
__device__ __forceinline__ double barrier(double d, double dhat, double kappa) {
    const double s = d / (dhat * dhat);
    return 0.5 * dhat * kappa * 0.125 * (s - 1) * log(s);
}
__device__ __forceinline__ double barrier_d1(double d, double dhat, double kappa) {
    const double dh2 = dhat * dhat, s = d / dh2;
    return 0.5 * dhat * (kappa / 8 * (log(s) / dh2 + (s - 1) / d));
}
__device__ __forceinline__ double barrier_d2(double d, double dhat, double kappa) {
    const double dh2 = dhat * dhat;
    return 0.5 * dhat * (kappa / (8 * d * d * dh2) * (d + dh2));
}

// IPEnergy::Val (ip.cu:49-52) per element, summed per block:
//   inertia 1/2 m |x - x~|^2 (inertia.cu:71-85)
//   + h^2 (g m y (gravity.cu:23-34) + V0 (mu |S - I|^2 + lambda/2 tr(S - I)^2)
//          (corotated.cu:196-214) + b(d) for d < dhat^2 (barrier.cu:137-155)).
// ------------------------------------------------------ implicit barrier --
// Fixed-body records (collision.hpp: p[3], u[3], r, grouped spheres, planes,
// cylinders).  barrierFunc / Grad / Hess (solverUtil.cuh:33-47) with kappa =
// FixedBody::kappa = 100 (rigid.h:25), over planes, then cylinders, then
// spheres (implicitBarrier.cu:163-240), each body's Hessian made positive
// definite on its own (makePD, solverUtil.cuh:157-232).
struct Bodies {
    const float* rec;
    int ns, np, nc;
};
constexpr double kBodyKappa = 100.0;

// signed distance of x to body j and the unit normal; cyl/sphere: |n| - r
__device__ __forceinline__ double body_distance(const Bodies& b, int j, int kind, const double x[3], double n[3]) {
    const float* q = b.rec + kFixedFloats * (size_t)j;
    const double p[3] = {q[0], q[1], q[2]}, u[3] = {q[3], q[4], q[5]};
    if (kind == kFixedPlane) {
        for (int k = 0; k < 3; ++k) n[k] = u[k];
        return (x[0] - p[0]) * u[0] + (x[1] - p[1]) * u[1] + (x[2] - p[2]) * u[2];
    }
    const double w[3] = {x[0] - p[0], x[1] - p[1], x[2] - p[2]};
    double m[3];
    if (kind == kFixedCylinder) {  // (I - a a^T) w, entries 1 - a_i a_j / 0 - a_i a_j
        for (int r = 0; r < 3; ++r) {
            double acc = 0.0;
            for (int c = 0; c < 3; ++c) acc += ((r == c ? 1.0 : 0.0) - u[r] * u[c]) * w[c];
            m[r] = acc;
        }
    } else {
        for (int k = 0; k < 3; ++k) m[k] = w[k];
    }
    const double len = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
    const double inv = 1.0 / len;
    for (int k = 0; k < 3; ++k) n[k] = inv * m[k];
    return len - (double)q[6];
}

__device__ void make_pd3(double (&M)[3][3]) {
    double V[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};  // columns: eigenvectors
    for (int sweep = 0; sweep < 20; ++sweep) {
        bool any = false;
        for (int k = 0; k < 3; ++k) {
            const int p = k < 2 ? 0 : 1, q = k == 0 ? 1 : 2;
            const double apq = M[p][q];
            if (fabs(apq) < 1e-9) continue;
            any = true;
            const double app = M[p][p], aqq = M[q][q];
            const double theta = 0.5 * (aqq - app) / apq;
            double t;
            if (fabs(theta) > 1e10) t = 0.5 / theta;
            else t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(1.0 + theta * theta));
            const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
            M[p][q] = M[q][p] = 0.0;
            M[p][p] = c * c * app - 2.0 * s * c * apq + s * s * aqq;
            M[q][q] = s * s * app + 2.0 * s * c * apq + c * c * aqq;
            const int o = 3 - p - q;
            const double aop = M[o][p], aoq = M[o][q];
            M[o][p] = M[p][o] = c * aop - s * aoq;
            M[o][q] = M[q][o] = s * aop + c * aoq;
            for (int r = 0; r < 3; ++r) {
                const double vp = V[r][p], vq = V[r][q];
                V[r][p] = c * vp - s * vq;
                V[r][q] = s * vp + c * vq;
            }
        }
        if (!any) break;
    }
    double lam[3];
    for (int k = 0; k < 3; ++k) lam[k] = M[k][k] < 1e-6 ? 1e-6 : M[k][k];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double acc = 0.0;
            for (int k = 0; k < 3; ++k) acc += V[i][k] * lam[k] * V[j][k];
            M[i][j] = acc;
        }
}

// Energy, gradient and Hessian (any may be null) of vertex x over all bodies,
// unscaled.
__device__ void body_terms(const Bodies& b, const double x[3], double area, double dhat, double* e, double* g,
                           double (*H)[3]) {
    const int first[3] = {b.ns, b.ns + b.np, 0};           // planes, cylinders, spheres
    const int count[3] = {b.np, b.nc, b.ns};
    const int kind[3] = {kFixedPlane, kFixedCylinder, kFixedSphere};
    for (int pass = 0; pass < 3; ++pass)
        for (int j = first[pass]; j < first[pass] + count[pass]; ++j) {
            double n[3];
            const double d = body_distance(b, j, kind[pass], x, n);
            if (!(d < dhat)) continue;
            const double s = d / dhat, ls = log(s);
            if (e) *e += area * dhat * kBodyKappa * 0.5 * (s - 1) * ls;
            if (g) {
                const double c = area * dhat * (kBodyKappa / 2 * (ls / dhat + (s - 1) / d));
                for (int k = 0; k < 3; ++k) g[k] += c * n[k];
            }
            if (H) {
                const double c = area * dhat * kBodyKappa / (2 * d * d * dhat) * (d + dhat);
                double M[3][3];
                for (int r = 0; r < 3; ++r)
                    for (int q = 0; q < 3; ++q) M[r][q] = c * (n[r] * n[q]);
                make_pd3(M);
                for (int r = 0; r < 3; ++r)
                    for (int q = 0; q < 3; ++q) H[r][q] += M[r][q];
            }
        }
}

// g += coef * gradient; with hv: hv = coef * Hessian block, diag += its diagonal.
__global__ void ipc_body_grad_kernel(int nv, const double* X, Bodies b, const double* area, double dhat, double coef,
                                     double* g, double* hv, double* diag) {
    const int v = blockIdx.x * kT + threadIdx.x;
    if (v >= nv) return;
    const double x[3] = {X[3 * (size_t)v], X[3 * (size_t)v + 1], X[3 * (size_t)v + 2]};
    double gl[3] = {0, 0, 0}, H[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    body_terms(b, x, area[v], dhat, nullptr, gl, hv ? H : nullptr);
    for (int k = 0; k < 3; ++k) g[3 * (size_t)v + k] += coef * gl[k];
    if (hv) {
        for (int r = 0; r < 3; ++r)
            for (int q = 0; q < 3; ++q) hv[9 * (size_t)v + 3 * r + q] = coef * H[r][q];
        for (int k = 0; k < 3; ++k) diag[3 * (size_t)v + k] += coef * H[k][k];
    }
}

// InitStepSize of the bodies (implicitBarrier.cu:352-407): per block, the
// smallest 0.9 * (distance to the body along the motion -p) over its
// vertices, 1 when nothing is approached; a ray that misses a curved body
// (NaN) changes nothing.
__global__ void ipc_body_step_kernel(int nv, const double* X, const double* P, Bodies b, double* parts) {
    __shared__ double red[kT];
    double alpha = 1.0;
    const int v = blockIdx.x * kT + threadIdx.x;
    if (v < nv) {
        const double x[3] = {X[3 * (size_t)v], X[3 * (size_t)v + 1], X[3 * (size_t)v + 2]};
        const double lp[3] = {-P[3 * (size_t)v], -P[3 * (size_t)v + 1], -P[3 * (size_t)v + 2]};
        for (int j = 0; j < b.ns + b.np + b.nc; ++j) {
            const float* q = b.rec + kFixedFloats * (size_t)j;
            const double p[3] = {q[0], q[1], q[2]}, u[3] = {q[3], q[4], q[5]}, r = q[6];
            double cand = NAN;
            if (j >= b.ns && j < b.ns + b.np) {  // plane
                const double pn = lp[0] * u[0] + lp[1] * u[1] + lp[2] * u[2];
                if (pn < 0) cand = 0.9 * (u[0] * (x[0] - p[0]) + u[1] * (x[1] - p[1]) + u[2] * (x[2] - p[2])) / -pn;
            } else {
                const bool cyl = j >= b.ns + b.np;
                const double w0[3] = {x[0] - p[0], x[1] - p[1], x[2] - p[2]};
                double n[3], pp[3];
                for (int rr = 0; rr < 3; ++rr) {
                    double an = 0.0, ap = 0.0;
                    for (int c = 0; c < 3; ++c) {
                        const double m = cyl ? ((rr == c ? 1.0 : 0.0) - u[rr] * u[c]) : (rr == c ? 1.0 : 0.0);
                        an += m * w0[c];
                        ap += m * lp[c];
                    }
                    n[rr] = an;
                    pp[rr] = ap;
                }
                const double pn = lp[0] * n[0] + lp[1] * n[1] + lp[2] * n[2];
                if (pn < 0) {
                    const double pp2 = pp[0] * pp[0] + pp[1] * pp[1] + pp[2] * pp[2];
                    const double n2 = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
                    const double nd = n[0] * pp[0] + n[1] * pp[1] + n[2] * pp[2];
                    cand = 0.9 * (-nd - sqrt(nd * nd - pp2 * (n2 - r * r))) / pp2;
                }
            }
            alpha = fmin(alpha, cand);
        }
    }
    red[threadIdx.x] = alpha;
    __syncthreads();
    for (int s = kT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) parts[blockIdx.x] = red[0];
}
__global__ void min_final_kernel(int nb, const double* parts, double* out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double m = 1.0;
    for (int k = 0; k < nb; ++k) m = fmin(m, parts[k]);
    out[0] = m;
}

Bodies bodies_of(const FixedBodies* f) {
    if (!f || f->spheres() + f->planes() + f->cylinders() == 0) return Bodies{nullptr, 0, 0, 0};
    return Bodies{f->device(), f->spheres(), f->planes(), f->cylinders()};
}

__global__ void ipc_energy_kernel(int nv, int nt, long long nq, const double* X, const double* XT,
                                  const double* mass, const uint32_t* tet, const double* dm, const double* v0,
                                  const double* mu, const double* lam, const double* qd, double dhat, double kappa,
                                  double h2, Bodies bodies, const double* area, int neo_hookean, double* parts) {
    const long long i = (long long)blockIdx.x * kT + threadIdx.x;
    double e[1] = {0.0};
    if (i < nv) {
        double s = 0.0;
        for (int k = 0; k < 3; ++k) {
            const double d = X[3 * i + k] - XT[3 * i + k];
            s += d * d;
        }
        e[0] += 0.5 * mass[i] * s + h2 * kG * X[3 * i + 1] * mass[i];
        if (bodies.rec) {
            const double x[3] = {X[3 * i], X[3 * i + 1], X[3 * i + 2]};
            double eb = 0.0;
            body_terms(bodies, x, area[i], dhat, &eb, nullptr, nullptr);
            e[0] += h2 * eb;
        }
    }
    if (i < nt) {
        double F[3][3];
        deformation(X, tet + 4 * i, dm + 9 * i, F);
        if (neo_hookean) {
            // NeoHookean08Energy::Val (neohookean08.cu:193-206)
            double f2 = 0.0;
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) f2 += F[r][c] * F[r][c];
            const double lj = log(det3(F));
            e[0] += h2 * v0[i] * (mu[i] * 0.5 * (f2 - 3) - mu[i] * lj + 0.5 * lam[i] * lj * lj);
        } else {
            double U[3][3], sg[3], Vm[3][3];
            Svd3<double>::run(F, U, sg, Vm);
            const double a = sg[0] - 1, b = sg[1] - 1, c = sg[2] - 1, tr = a + b + c;
            e[0] += h2 * v0[i] * (mu[i] * (a * a + b * b + c * c) + 0.5 * lam[i] * tr * tr);
        }
    }
    if (i < nq) {
        const double d = qd[i];
        if (d < dhat * dhat) e[0] += h2 * barrier(d, dhat, kappa);
    }
    block_sums<1>(e, parts);
}

// Inertia (coef 1) and gravity (coef h^2) gradients, inertia.cu:88-105 and
// gravity.cu:6-13; diag = m (the inertia Hessian, inertia.cu:8-22).
__global__ void ipc_grad_vertex_kernel(int nv, const double* X, const double* XT, const double* mass, double h2,
                                       double* g, double* diag) {
    const int v = blockIdx.x * kT + threadIdx.x;
    if (v >= nv) return;
    const double m = mass[v];
    for (int k = 0; k < 3; ++k) {
        g[3 * (size_t)v + k] = m * (X[3 * (size_t)v + k] - XT[3 * (size_t)v + k]) + (k == 1 ? h2 * kG * m : 0.0);
        if (diag) diag[3 * (size_t)v + k] = m;
    }
}

// Corotated elasticity of one tet (CorotatedEnergy::GradHessianKern,
// corotated.cu:100-186), scaled by coef: the gradient V0 P Dm^-T with
// P = 2 mu (F - R) + lambda tr(R^T F - I) R into g (atomics), and with ht the
// 12x12 Hessian G^T d2Psi/dF2 G (G = dvec(F)/dx), where
//   d2Psi/dF2 = 2 mu I + lambda g1 g1^T + sum_k c lambda_k t_k t_k^T,
//   g1 = vec R, t_k = vec(U T_k V^T)/sqrt 2 (twists of the pairs 01, 12, 20),
//   c = lambda tr S - 2 mu - 3 lambda, lambda_k = 2/(s_a + s_b) (1 if s_a + s_b < 2);
// the elastic Newton Hessian is not projected to positive definite there.
__global__ void __launch_bounds__(kT) ipc_elastic_kernel(int nt, const uint32_t* tet, const double* X, const double* dm,
                                                         const double* v0, const double* mu, const double* lam,
                                                         double coef, const unsigned char* fixed, double* g, double* ht,
                                                         double* diag) {
    const int t = blockIdx.x * kT + threadIdx.x;
    if (t >= nt) return;
    const uint32_t* tv = tet + 4 * (size_t)t;
    const double* m = dm + 9 * (size_t)t;
    double F[3][3], U[3][3], sg[3], Vm[3][3];
    deformation(X, tv, m, F);
    Svd3<double>::run(F, U, sg, Vm);
    double R[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) R[r][c] = U[r][0] * Vm[c][0] + U[r][1] * Vm[c][1] + U[r][2] * Vm[c][2];
    double trRF = 0.0;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) trRF += R[r][c] * F[r][c];
    const double mu_ = mu[t], la = lam[t], w = coef * v0[t];
    // Gv[a][c]: d F[r][c] / d x_{a, r}
    double Gv[4][3];
    for (int c = 0; c < 3; ++c) {
        Gv[1][c] = m[c];
        Gv[2][c] = m[3 + c];
        Gv[3][c] = m[6 + c];
        Gv[0][c] = -(m[c] + m[3 + c] + m[6 + c]);
    }
    for (int a = 0; a < 4; ++a)
        for (int r = 0; r < 3; ++r) {
            double acc = 0.0;
            for (int c = 0; c < 3; ++c)
                acc += (2 * mu_ * (F[r][c] - R[r][c]) + la * (trRF - 3.0) * R[r][c]) * Gv[a][c];
            atomicAdd(&g[3 * (size_t)tv[a] + r], w * acc);
        }
    if (!ht) return;
    double tw[3][3][3];  // tw[k][r][c]
    const int ta[3] = {0, 1, 2}, tb[3] = {1, 2, 0};
    const double is2 = 1.0 / sqrt(2.0);
    for (int k = 0; k < 3; ++k)
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c)
                tw[k][r][c] = (U[r][ta[k]] * Vm[c][tb[k]] - U[r][tb[k]] * Vm[c][ta[k]]) * is2;
    double lk[3] = {2 / (sg[0] + sg[1]), 2 / (sg[1] + sg[2]), 2 / (sg[0] + sg[2])};
    if (sg[0] + sg[1] < 2) lk[0] = 1;
    if (sg[1] + sg[2] < 2) lk[1] = 1;
    if (sg[0] + sg[2] < 2) lk[2] = 1;
    const double cc = la * (sg[0] + sg[1] + sg[2]) - 2 * mu_ - 3 * la;
    double* H = ht + 144 * (size_t)t;
    for (int a = 0; a < 4; ++a)
        for (int k = 0; k < 3; ++k)
            for (int b = 0; b < 4; ++b)
                for (int l = 0; l < 3; ++l) {
                    // sum_{c,c'} Gv[a][c] Gv[b][c'] D((k,c), (l,c'))
                    double acc = 0.0;
                    for (int c = 0; c < 3; ++c)
                        for (int c2 = 0; c2 < 3; ++c2) {
                            double Dv = (k == l && c == c2 ? 2 * mu_ : 0.0) + la * R[k][c] * R[l][c2];
                            for (int q = 0; q < 3; ++q) Dv += cc * lk[q] * tw[q][k][c] * tw[q][l][c2];
                            acc += Gv[a][c] * Gv[b][c2] * Dv;
                        }
                    H[12 * (3 * a + k) + 3 * b + l] = w * acc;
                }
    if (diag)
        for (int a = 0; a < 4; ++a)
            if (!fixed[tv[a]])
                for (int k = 0; k < 3; ++k) atomicAdd(&diag[3 * (size_t)tv[a] + k], H[13 * (3 * a + k)]);
}

// Neo-Hookean elasticity of one tet (NeoHookean08Energy::GradHessianKern,
// neohookean08.cu:99-183), scaled by coef: with J = det F and C = dJ/dF (the
// cofactor matrix, columns F1 x F2, F2 x F0, F0 x F1), the gradient V0 P Dm^-T
// with P = mu (F - C/J) + lambda ln J C/J into g (atomics), and with ht the
// 12x12 Hessian G^T D G of the kernel's
//   D = mu I + (lambda (1 - ln J) + mu) / J^2 vec C vec C^T,
// which is what that kernel assembles: the term (lambda ln J - mu)/J dC/dF of
// the exact second derivative is formed there (H3) but never added, and the
// Hessian is not projected to positive definite.
__global__ void __launch_bounds__(kT) ipc_nh_kernel(int nt, const uint32_t* tet, const double* X, const double* dm,
                                                    const double* v0, const double* mu, const double* lam,
                                                    double coef, const unsigned char* fixed, double* g, double* ht,
                                                    double* diag) {
    const int t = blockIdx.x * kT + threadIdx.x;
    if (t >= nt) return;
    const uint32_t* tv = tet + 4 * (size_t)t;
    const double* m = dm + 9 * (size_t)t;
    double F[3][3];
    deformation(X, tv, m, F);
    double C[3][3];
    for (int c = 0; c < 3; ++c) {
        const int a = (c + 1) % 3, b = (c + 2) % 3;
        C[0][c] = F[1][a] * F[2][b] - F[2][a] * F[1][b];
        C[1][c] = F[2][a] * F[0][b] - F[0][a] * F[2][b];
        C[2][c] = F[0][a] * F[1][b] - F[1][a] * F[0][b];
    }
    const double J = det3(F), invJ = 1.0 / J, lj = log(J);
    const double mu_ = mu[t], la = lam[t], w = coef * v0[t];
    double Gv[4][3];
    for (int c = 0; c < 3; ++c) {
        Gv[1][c] = m[c];
        Gv[2][c] = m[3 + c];
        Gv[3][c] = m[6 + c];
        Gv[0][c] = -(m[c] + m[3 + c] + m[6 + c]);
    }
    for (int a = 0; a < 4; ++a)
        for (int r = 0; r < 3; ++r) {
            double acc = 0.0;
            for (int c = 0; c < 3; ++c)
                acc += (mu_ * (F[r][c] - invJ * C[r][c]) + la * lj * invJ * C[r][c]) * Gv[a][c];
            atomicAdd(&g[3 * (size_t)tv[a] + r], w * acc);
        }
    if (!ht) return;
    const double alpha = (la * (1.0 - lj) + mu_) * invJ * invJ;
    double q[4][3], gg[4][4];  // q[a][k] = sum_c Gv[a][c] C[k][c]; gg[a][b] = sum_c Gv[a][c] Gv[b][c]
    for (int a = 0; a < 4; ++a) {
        for (int k = 0; k < 3; ++k) q[a][k] = Gv[a][0] * C[k][0] + Gv[a][1] * C[k][1] + Gv[a][2] * C[k][2];
        for (int b = 0; b < 4; ++b) gg[a][b] = Gv[a][0] * Gv[b][0] + Gv[a][1] * Gv[b][1] + Gv[a][2] * Gv[b][2];
    }
    double* H = ht + 144 * (size_t)t;
    for (int a = 0; a < 4; ++a)
        for (int k = 0; k < 3; ++k)
            for (int b = 0; b < 4; ++b)
                for (int l = 0; l < 3; ++l)
                    H[12 * (3 * a + k) + 3 * b + l] = w * ((k == l ? mu_ * gg[a][b] : 0.0) + alpha * q[a][k] * q[b][l]);
    if (diag)
        for (int a = 0; a < 4; ++a)
            if (!fixed[tv[a]])
                for (int k = 0; k < 3; ++k) atomicAdd(&diag[3 * (size_t)tv[a] + k], H[13 * (3 * a + k)]);
}

// makePD (matrix.h:441-518) on a 12x12 matrix held in LDS (entry (i, j) of
// this thread at M[(12 i + j) * kQB + lane]): cyclic Jacobi over the pairs
// p < q, rotations skipped when |a_pq| < 1e-9, at most 20 sweeps; eigenvalues
// below 1e-6 raised to 1e-6; M <- V diag V^T.
__device__ void make_pd_lds(double* M, double* V, int lane) {
    auto m = [&](int i, int j) -> double& { return M[(12 * i + j) * kQB + lane]; };
    auto v = [&](int i, int j) -> double& { return V[(12 * i + j) * kQB + lane]; };
    for (int i = 0; i < 12; ++i)
        for (int j = 0; j < 12; ++j) v(i, j) = i == j ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 20; ++sweep) {
        bool converged = true;
        for (int p = 0; p < 12; ++p)
            for (int q = p + 1; q < 12; ++q) {
                const double apq = m(p, q);
                if (fabs(apq) < 1e-9) continue;
                converged = false;
                const double app = m(p, p), aqq = m(q, q);
                const double theta = 0.5 * (aqq - app) / apq;
                double t;
                if (fabs(theta) > 1e10) t = 0.5 / theta;
                else t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(1.0 + theta * theta));
                const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
                m(p, q) = 0.0;
                m(q, p) = 0.0;
                m(p, p) = c * c * app - 2.0 * s * c * apq + s * s * aqq;
                m(q, q) = s * s * app + 2.0 * s * c * apq + c * c * aqq;
                for (int k = 0; k < 12; ++k) {
                    if (k == p || k == q) continue;
                    const double akp = m(k, p), akq = m(k, q);
                    m(k, p) = c * akp - s * akq;
                    m(p, k) = m(k, p);
                    m(k, q) = s * akp + c * akq;
                    m(q, k) = m(k, q);
                }
                for (int k = 0; k < 12; ++k) {
                    const double vkp = v(k, p), vkq = v(k, q);
                    v(k, p) = c * vkp - s * vkq;
                    v(k, q) = s * vkp + c * vkq;
                }
            }
        if (converged) break;
    }
    double lamv[12];
    for (int i = 0; i < 12; ++i) lamv[i] = fmax(m(i, i), 1e-6);
    for (int i = 0; i < 12; ++i)
        for (int j = i; j < 12; ++j) {
            double s = 0.0;
            for (int k = 0; k < 12; ++k) s += v(i, k) * lamv[k] * v(j, k);
            m(i, j) = s;  // only the upper triangle is read back: rows below are still the input
        }
    for (int i = 1; i < 12; ++i)
        for (int j = 0; j < i; ++j) m(i, j) = m(j, i);
}

// Barrier of one query (Barrier::gradHessianKern, barrier.cu:99-173),
// scaled by coef: gradient coef b'(d) grad d into g (atomics), and with hq
// the Hessian coef (b''(d) grad d grad d^T + b'(d) hess d), made positive
// definite.  Queries with d > dhat^2 contribute nothing.
__global__ void __launch_bounds__(kQB) ipc_barrier_kernel(long long nq, const QueryKey* qk, const int* dtype,
                                                          const double* qd, const double* X, double dhat, double kappa,
                                                          double coef, const unsigned char* fixed, double* g,
                                                          double* hq, double* diag) {
    __shared__ double sM[144 * kQB];
    __shared__ double sV[144 * kQB];
    const long long i = (long long)blockIdx.x * kQB + threadIdx.x;
    if (i >= nq) return;  // no barrier below
    const double d = qd[i];
    if (d > dhat * dhat) return;
    const QueryKey k = qk[i];
    const D3 P[4] = {ldp(X, k.v0), ldp(X, k.v1), ldp(X, k.v2), ldp(X, k.v3)};
    const uint32_t vv[4] = {k.v0, k.v1, k.v2, k.v3};
    const int lane = threadIdx.x;
    double gd[12];
    double* M = sM;
    if (hq)
        for (int e = 0; e < 144; ++e) M[e * kQB + lane] = 0.0;
    ipc::distance_derivatives((int)k.type, dtype[i], P, gd, [&](int a, int b, const ipc::M3& blk, double sgn) {
        if (!hq) return;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) M[(12 * (3 * a + r) + 3 * b + c) * kQB + lane] += sgn * blk.m[r][c];
    });
    const double d1 = barrier_d1(d, dhat, kappa);
    for (int j = 0; j < 12; ++j) atomicAdd(&g[3 * (size_t)vv[j / 3] + j % 3], coef * d1 * gd[j]);
    if (!hq) return;
    const double d2 = barrier_d2(d, dhat, kappa);
    for (int a = 0; a < 12; ++a)
        for (int b = 0; b < 12; ++b) {
            double& e = M[(12 * a + b) * kQB + lane];
            e = coef * (d2 * gd[a] * gd[b] + d1 * e);
        }
    make_pd_lds(M, sV, lane);
    double* H = hq + 144 * (size_t)i;
    for (int e = 0; e < 144; ++e) H[e] = M[e * kQB + lane];
    if (diag)
        for (int a = 0; a < 4; ++a)
            if (!fixed[vv[a]])
                for (int r = 0; r < 3; ++r) atomicAdd(&diag[3 * (size_t)vv[a] + r], H[13 * (3 * a + r)]);
}

// DOFElimination (ipc.cu:55-80): zero gradient and identity row/column on fixed vertices.
__global__ void ipc_fix_kernel(int nv, const unsigned char* fixed, double* g, double* diag) {
    const int v = blockIdx.x * kT + threadIdx.x;
    if (v >= nv || !fixed[v]) return;
    for (int k = 0; k < 3; ++k) {
        g[3 * (size_t)v + k] = 0.0;
        if (diag) diag[3 * (size_t)v + k] = 1.0;
    }
}

// Start of a Jacobi-PCG solve (n = 3 N_v): x = 0, r = b, z = D^-1 b and the
// block partials of r.z and r.r.
__global__ void ipc_pcg_init_kernel(int n, const double* b, const double* diag, double* x, double* r, double* z,
                                    double* parts) {
    const int i = blockIdx.x * kT + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (i < n) {
        const double ri = b[i], zi = ri / diag[i];
        x[i] = 0.0;
        r[i] = ri;
        z[i] = zi;
        v[0] = ri * zi;
        v[1] = ri * ri;
    }
    block_sums<2>(v, parts);
}

// The tets' 12x12 Hessian blocks summed into 3x3 blocks per (vertex, vertex)
// pair, once per Newton iteration: entry e of block k is the sum, in
// ascending tet order, of the matching entries of the tets that hold both
// vertices (`contrib`: tet << 4 | a << 2 | b for row vertex a and column
// vertex b of the tet).  Blocks touching a fixed vertex are zero (the
// elimination the element product applies).
__global__ void ipc_assemble_rows_kernel(long long nnzb, const int* row_of, const int* col,
                                         const long long* contrib_ptr, const uint32_t* contrib, const double* ht,
                                         const unsigned char* fixed, double* val) {
    const long long g = (long long)blockIdx.x * kT + threadIdx.x;
    if (g >= 9 * nnzb) return;
    const long long k = g / 9;
    const int e = (int)(g - 9 * k), r = e / 3, c = e - 3 * r;
    double s = 0.0;
    if (!fixed[row_of[k]] && !fixed[col[k]]) {
        for (long long q = contrib_ptr[k]; q < contrib_ptr[k + 1]; ++q) {
            const uint32_t code = contrib[q];
            const size_t t = code >> 4;
            const int a = (code >> 2) & 3, b = code & 3;
            s += ht[144 * t + (size_t)(3 * a + r) * 12 + 3 * b + c];
        }
    }
    val[9 * k + e] = s;
}

// The active queries' 12x12 barrier Hessians assembled per Newton iteration
// into 3x3 blocks per (vertex, vertex) pair, like the elastic rows: every
// (query, a, b) with both vertices free gets the key row * nv + col (inactive
// ones the sentinel nv^2, which sorts last); a stable sort by key puts each
// pair's entries together in ascending query order; the heads of the runs
// number the blocks, each block sums its run's 3x3 pieces in that order, and
// the rows are the runs of equal row vertex.
__global__ void ipc_contact_expand_kernel(long long nq, const QueryKey* __restrict__ qk, const double* __restrict__ qd,
                                          double dhat2, const unsigned char* __restrict__ fixed,
                                          unsigned long long nvl, unsigned long long* __restrict__ key,
                                          uint32_t* __restrict__ code) {
    const long long t = (long long)blockIdx.x * kT + threadIdx.x;
    if (t >= 16 * nq) return;
    const long long i = t >> 4;
    const int a = (int)(t >> 2) & 3, b = (int)t & 3;
    unsigned long long k = nvl * nvl;
    if (qd[i] <= dhat2) {
        const QueryKey q = qk[i];
        const uint32_t vv[4] = {q.v0, q.v1, q.v2, q.v3};
        if (!fixed[vv[a]] && !fixed[vv[b]]) k = (unsigned long long)vv[a] * nvl + vv[b];
    }
    key[t] = k;
    code[t] = (uint32_t)t;  // 16 i + 4 a + b
}
__global__ void ipc_contact_heads_kernel(long long ne, const unsigned long long* __restrict__ key,
                                         unsigned long long sentinel, int* __restrict__ flag) {
    const long long e = (long long)blockIdx.x * kT + threadIdx.x;
    if (e >= ne) return;
    flag[e] = key[e] != sentinel && (e == 0 || key[e] != key[e - 1]);
}
// count = [blocks, entries before the first sentinel]
__global__ void ipc_contact_blocks_kernel(long long ne, const unsigned long long* __restrict__ key,
                                          unsigned long long nvl, const int* __restrict__ flag,
                                          const int* __restrict__ pos, int* __restrict__ head, int* __restrict__ brow,
                                          int* __restrict__ bcol, int* __restrict__ count) {
    const long long e = (long long)blockIdx.x * kT + threadIdx.x;
    if (e >= ne) return;
    const unsigned long long sentinel = nvl * nvl;
    const unsigned long long k = key[e];
    if (flag[e]) {
        const int b = pos[e];
        head[b] = (int)e;
        brow[b] = (int)(k / nvl);
        bcol[b] = (int)(k % nvl);
    }
    if (e == ne - 1) count[0] = pos[e] + flag[e];
    const bool valid = k != sentinel;
    if ((e == 0 && !valid) || (valid && (e == ne - 1 || key[e + 1] == sentinel))) count[1] = valid ? (int)e + 1 : 0;
}
__global__ void ipc_contact_values_kernel(long long nbmax, const int* __restrict__ count, const int* __restrict__ head,
                                          const uint32_t* __restrict__ code, const double* __restrict__ hq,
                                          double* __restrict__ val) {
    const long long g = (long long)blockIdx.x * kT + threadIdx.x;
    if (g >= 9 * nbmax) return;
    const int nb = count[0];
    const long long b = g / 9;
    if (b >= nb) return;
    const int rc = (int)(g - 9 * b), r = rc / 3, c = rc - 3 * r;
    const int e1 = b + 1 < nb ? head[b + 1] : count[1];
    double s = 0.0;
    for (int e = head[b]; e < e1; ++e) {
        const uint32_t t = code[e];
        const size_t i = t >> 4;
        const int a = (t >> 2) & 3, bb = t & 3;
        s += hq[144 * i + (size_t)(3 * a + r) * 12 + 3 * bb + c];
    }
    val[9 * b + rc] = s;
}
// crow[v] = the first block whose row is >= v (v = 0..nv)
__global__ void ipc_contact_rows_kernel(int nv, const int* __restrict__ count, const int* __restrict__ brow,
                                        int* __restrict__ crow) {
    const int v = blockIdx.x * kT + threadIdx.x;
    if (v > nv) return;
    int lo = 0, hi = count[0];
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (brow[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    crow[v] = lo;
}

// Jacobi-PCG on H p = g in two launches per iteration (n = 3 N_v).  The
// scalars stay on the device: out = [rz0, rr, pq, alpha, beta, |b|, rz (even
// iterations), rz (odd iterations)], ctl = [finished, completed iterations].
// Every block sums the previous launch's block partials itself in the same
// fixed order, so every block takes the same decision; block 0 records the
// scalars.  The search direction is double-buffered: the product launch of
// iteration `it` forms p_it = z + beta p_{it-1} (p_0 = z) wherever it needs
// it (its own rows and the columns of its rows' elastic and contact blocks)
// with one expression, so every copy is the same value.
// Both launches request their memory before the block-wide steps (the stop
// word, the partial sums), so those round trips overlap.
__device__ __forceinline__ double dir_value(bool first, double zj, double pj, double beta) {
    return first ? zj : fma(beta, pj, zj);
}

// The stop word of the solve, the same in every thread of the block (a
// block of a launch after the one that finished the solve returns whole).
__device__ __forceinline__ bool solve_stopped(const int* ctl) {
    __shared__ int stop;
    if (threadIdx.x == 0) stop = ctl[0];
    __syncthreads();
    return stop != 0;
}

constexpr int kRowLanes = 4;  // lanes per (vertex, component) row of the product
constexpr int kRowPer = 4;    // blocks per lane whose loads are in flight together

// Product launch of iteration it, kRowLanes lanes per row and component:
//   it > 0: the stop test of iteration it-1 (|r| <= tol |b| or it >= max_iter)
//   and beta = rz_it / rz_{it-1};
//   p_it and q = M p + (the implicit barrier's block) p + the assembled
//   elastic row times p + the assembled contact row times p (identity on
//   fixed rows), lane l taking the blocks l, l + kRowLanes, ... of each list;
//   block partials of p.q.
template <bool kFirst>
__global__ void __launch_bounds__(kT) ipc_pcg_prod_kernel(
    int nv, const long long* __restrict__ rowptr, const int* __restrict__ col, const double* __restrict__ val,
    const int* __restrict__ crow, const int* __restrict__ ccol, const double* __restrict__ cval,
    const unsigned char* __restrict__ fixed, const double* __restrict__ mass, const double* __restrict__ hv,
    const double* __restrict__ z, const double* __restrict__ pold, double* __restrict__ pnew, double* __restrict__ q,
    int nb_upd, const double* __restrict__ upd_parts, int it, int max_iter, double tol, double* __restrict__ parts,
    double* __restrict__ out, int* __restrict__ ctl) {
    const int t = blockIdx.x * kT + threadIdx.x;
    const int g = t / kRowLanes, l = t % kRowLanes;  // row g = 3 v + r
    const bool live = g < 3 * nv;
    const int v = live ? g / 3 : 0, r = live ? g - 3 * v : 0;
    bool work = false;
    long long k0 = 0, k1 = 0;
    int c0 = 0, c1 = 0;
    double own_z[3] = {0, 0, 0}, own_p[3] = {0, 0, 0}, m = 1.0, hb[3] = {0, 0, 0};
    if (live) {
        work = fixed[v] == 0;
        k0 = rowptr[v];
        k1 = rowptr[v + 1];
        if (crow) {
            c0 = crow[v];
            c1 = crow[v + 1];
        }
        for (int c = 0; c < 3; ++c) {
            own_z[c] = z[3 * (size_t)v + c];
            if (!kFirst) own_p[c] = pold[3 * (size_t)v + c];
        }
        if (work) {
            m = mass[v];
            if (hv)  // the implicit barrier's block on free vertices
                for (int c = 0; c < 3; ++c) hb[c] = hv[9 * (size_t)v + 3 * r + c];
        }
    }
    if (solve_stopped(ctl)) return;
    // the operands of kRowPer blocks of this lane, loaded before the scalars
    double bm[kRowPer][3], xz[kRowPer][3], xp[kRowPer][3];
    auto load_blocks = [&](const auto* __restrict__ cl, const double* __restrict__ vl, long long kb, long long ke) {
#pragma unroll
        for (int u = 0; u < kRowPer; ++u) {
            const long long kk = kb + l + (long long)kRowLanes * u;
            const long long kc = kk < ke ? kk : ke - 1;
            const size_t j = 3 * (size_t)cl[kc];
            for (int c = 0; c < 3; ++c) {
                bm[u][c] = vl[9 * kc + 3 * r + c];
                xz[u][c] = z[j + c];
                xp[u][c] = kFirst ? 0.0 : pold[j + c];
            }
        }
    };
    if (work && k0 < k1) load_blocks(col, val, k0, k1);
    double beta = 0.0;
    if (!kFirst) {
        double tot[2];
        sum_parts<2>(nb_upd, upd_parts, tot);
        const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
        if (sqrt(tot[1]) <= tol * out[5] || it >= max_iter) {
            if (lead) {
                out[1] = tot[1];
                ctl[1] = it;
                ctl[0] = 1;
            }
            return;  // the same decision in every block
        }
        beta = tot[0] / out[6 + ((it - 1) & 1)];
        if (lead) {
            out[1] = tot[1];
            out[4] = beta;
            out[6 + (it & 1)] = tot[0];
            ctl[1] = it;
        }
    }
    auto apply = [&](long long kb, long long ke, double& s) {
#pragma unroll
        for (int u = 0; u < kRowPer; ++u) {
            if (kb + l + (long long)kRowLanes * u >= ke) break;
            const double p0 = dir_value(kFirst, xz[u][0], xp[u][0], beta);
            const double p1 = dir_value(kFirst, xz[u][1], xp[u][1], beta);
            const double p2 = dir_value(kFirst, xz[u][2], xp[u][2], beta);
            s += bm[u][0] * p0 + bm[u][1] * p1 + bm[u][2] * p2;
        }
    };
    double s = 0.0, sc = 0.0;  // this lane's elastic and contact blocks, each in order
    if (work) {
        for (long long kb = k0; kb < k1; kb += kRowLanes * kRowPer) {
            if (kb != k0) load_blocks(col, val, kb, k1);
            apply(kb, k1, s);
        }
        for (long long kb = c0; kb < c1; kb += kRowLanes * kRowPer) {
            load_blocks(ccol, cval, kb, c1);
            apply(kb, c1, sc);
        }
    }
    // the row's lanes: (s0 + s1) + (s2 + s3)
    for (int o = 1; o < kRowLanes; o <<= 1) {
        s += __shfl_xor(s, o);
        sc += __shfl_xor(sc, o);
    }
    double pq[1] = {0.0};
    if (live && l == 0) {
        double pv[3];
        for (int c = 0; c < 3; ++c) pv[c] = dir_value(kFirst, own_z[c], own_p[c], beta);
        double y = m * pv[r];
        if (work) {
            y += hb[0] * pv[0] + hb[1] * pv[1] + hb[2] * pv[2];
            y += s;
            y += sc;
        }
        pnew[g] = pv[r];
        q[g] = y;
        pq[0] = pv[r] * y;
    }
    block_sums<1>(pq, parts);
}

// Update launch of iteration it: alpha = rz_it / p.q, x += alpha p,
// r -= alpha q, z = D^-1 r and the block partials of r.z and r.r.  H not positive along p
// finishes the solve with the iterate kept.
__global__ void __launch_bounds__(kT) ipc_pcg_update_kernel(int n, int nb_prod, const double* __restrict__ prod_parts,
                                                            int it, const double* __restrict__ p,
                                                            const double* __restrict__ q,
                                                            const double* __restrict__ diag, double* __restrict__ x,
                                                            double* __restrict__ r, double* __restrict__ z,
                                                            double* __restrict__ upd_parts, double* __restrict__ out,
                                                            int* __restrict__ ctl) {
    const int i = blockIdx.x * kT + threadIdx.x;
    double pi = 0.0, qi = 0.0, ri = 0.0, di = 1.0, xi = 0.0;
    if (i < n) {
        pi = p[i];
        qi = q[i];
        ri = r[i];
        di = diag[i];
        xi = x[i];
    }
    if (solve_stopped(ctl)) return;
    double tot[1];
    sum_parts<1>(nb_prod, prod_parts, tot);
    const double pq = tot[0];
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (!(pq > 0.0)) {
        if (lead) {
            out[2] = pq;
            ctl[1] = it;
            ctl[0] = 1;
        }
        return;
    }
    const double al = out[6 + (it & 1)] / pq;
    if (lead) {
        out[2] = pq;
        out[3] = al;
    }
    double v[2] = {0.0, 0.0};
    if (i < n) {
        x[i] = xi + al * pi;
        ri = ri - al * qi;
        const double zi = ri / di;
        r[i] = ri;
        z[i] = zi;
        v[0] = ri * zi;
        v[1] = ri * ri;
    }
    block_sums<2>(v, upd_parts);
}

}  // namespace

// ---------------------------------------------------------------------------
IpcEngine::IpcEngine(const IpcData& data) {
    SBS_REQUIRE(data.num_verts > 0 && data.num_tets > 0, "IPCSolver: empty mesh");
    SBS_REQUIRE(data.X && data.V && data.XTilde && data.Tet && data.mass && data.mu && data.lambda && data.DBC,
                "IPCSolver: a required SolverData pointer is null");
    SBS_REQUIRE(data.num_tris == 0 || (data.Tri && data.TriFathers), "IPCSolver: surface arrays are null");
    nv_ = data.num_verts;
    nt_ = data.num_tets;
    ntri_ = data.num_tris;
    d_ = data;
    {
        std::vector<uint32_t> tet(4 * (size_t)nt_);
        SBS_HIP(hipMemcpy(tet.data(), data.Tet, tet.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
        for (uint32_t v : tet) SBS_REQUIRE(v < (uint32_t)nv_, "IPCSolver: tet index out of range");
        build_rows(tet);
        // contact areas (computeInvDmV0, solverUtil.cuh:123-136, femSolver.cu:13-16):
        // over a vertex's tets, a third of the summed areas of the tet's three
        // faces through it, divided by the number of tets (accumulated from 0)
        std::vector<double> X(3 * (size_t)nv_), area(nv_, 0.0), deg(nv_, 0.0);
        SBS_HIP(hipMemcpy(X.data(), data.X, X.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int t = 0; t < nt_; ++t) {
            const uint32_t* tv = &tet[4 * (size_t)t];
            double fa[4];  // faces (0 1 2), (1 2 3), (2 3 0), (3 0 1)
            for (int f = 0; f < 4; ++f) {
                const double* a = &X[3 * (size_t)tv[f]];
                const double* b = &X[3 * (size_t)tv[(f + 1) % 4]];
                const double* c = &X[3 * (size_t)tv[(f + 2) % 4]];
                const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
                const double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
                const double cx = e1[1] * e2[2] - e2[1] * e1[2], cy = e1[2] * e2[0] - e2[2] * e1[0],
                             cz = e1[0] * e2[1] - e2[0] * e1[1];
                fa[f] = 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
            }
            const double per[4] = {(fa[0] + fa[2] + fa[3]) / 3.0, (fa[0] + fa[1] + fa[3]) / 3.0,
                                   (fa[0] + fa[1] + fa[2]) / 3.0, (fa[1] + fa[2] + fa[3]) / 3.0};
            for (int j = 0; j < 4; ++j) {
                area[tv[j]] += per[j];
                deg[tv[j]] += 1.0;
            }
        }
        for (int v = 0; v < nv_; ++v)
            if (deg[v] > 0) area[v] /= deg[v];
        area_.alloc(nv_);
        SBS_HIP(hipMemcpy(area_.get(), area.data(), area.size() * sizeof(double), hipMemcpyHostToDevice));
        if (ntri_ > 0) {
            std::vector<uint32_t> tri(3 * (size_t)ntri_);
            SBS_HIP(hipMemcpy(tri.data(), data.Tri, tri.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
            for (uint32_t v : tri) SBS_REQUIRE(v < (uint32_t)nv_, "IPCSolver: triangle index out of range");
        }
    }
    const size_t n3 = 3 * (size_t)nv_;
    dm_.alloc(9 * (size_t)nt_);
    v0_.alloc(nt_);
    fixed_.alloc(nv_);
    xn_.alloc(n3); xtmp_.alloc(n3); p_dir_.alloc(n3); g_.alloc(n3); diag_.alloc(n3);
    r_.alloc(n3); z_.alloc(n3); q_.alloc(n3); p2_.alloc(n3);
    crow_.alloc(nv_ + 1);
    ccount_.alloc(2);
    ht_.alloc(144 * (size_t)nt_);
    hv_.alloc(9 * (size_t)nv_);
    parts_.alloc(2 * (size_t)nblk(std::max<size_t>(n3, (size_t)nt_)) + 64);
    out_.alloc(8);
    pcg_ctl_.alloc(2);
    pcg_parts_.alloc(2 * (size_t)nblk(n3));
    prod_parts_.alloc(nblk((size_t)kRowLanes * n3));
    SBS_HIP(hipHostMalloc((void**)&host_, 8 * sizeof(double), hipHostMallocDefault));
    SBS_HIP(hipHostMalloc((void**)&host_ctl_, 2 * sizeof(int), hipHostMallocDefault));
    hipLaunchKernelGGL(ipc_precompute_kernel, dim3(nblk(nt_)), dim3(kT), 0, stream_, nt_, data.Tet,
                       (const double*)data.X, dm_.get(), v0_.get());
    hipLaunchKernelGGL(ipc_fixed_kernel, dim3(nblk(nv_)), dim3(kT), 0, stream_, nv_, data.DBC, fixed_.get());
    SBS_HIP(hipGetLastError());
    col_ = std::make_unique<CollisionDetector<double>>(nv_, ntri_, data.Tri, data.TriFathers, stream_);
    SBS_HIP(hipStreamSynchronize(stream_));
}

// The pattern of the assembled elastic Hessian: per vertex the sorted list of
// the vertices it shares a tet with (itself included), and per 3x3 block the
// (tet, a, b) entries that sum into it, in ascending tet order.
void IpcEngine::build_rows(const std::vector<uint32_t>& tet) {
    std::vector<std::vector<int>> nb(nv_);
    for (int t = 0; t < nt_; ++t)
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) nb[tet[4 * (size_t)t + a]].push_back((int)tet[4 * (size_t)t + b]);
    std::vector<long long> rowptr(nv_ + 1, 0);
    std::vector<int> col, row_of;
    for (int v = 0; v < nv_; ++v) {
        auto& l = nb[v];
        std::sort(l.begin(), l.end());
        l.erase(std::unique(l.begin(), l.end()), l.end());
        rowptr[v + 1] = rowptr[v] + (long long)l.size();
        for (int j : l) {
            col.push_back(j);
            row_of.push_back(v);
        }
    }
    nnzb_ = rowptr[nv_];
    std::vector<long long> cnt(nnzb_ + 1, 0);
    auto block = [&](int i, int j) {
        const auto& l = nb[i];
        return rowptr[i] + (long long)(std::lower_bound(l.begin(), l.end(), j) - l.begin());
    };
    for (int t = 0; t < nt_; ++t)
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) cnt[block((int)tet[4 * (size_t)t + a], (int)tet[4 * (size_t)t + b]) + 1]++;
    for (long long k = 0; k < nnzb_; ++k) cnt[k + 1] += cnt[k];
    std::vector<uint32_t> contrib(cnt[nnzb_]);
    std::vector<long long> fill(cnt.begin(), cnt.end() - 1);
    SBS_REQUIRE((size_t)nt_ < (1u << 28), "IPCSolver: too many tets for the assembled Hessian");
    for (int t = 0; t < nt_; ++t)  // ascending t, so each block's list is in tet order
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b)
                contrib[fill[block((int)tet[4 * (size_t)t + a], (int)tet[4 * (size_t)t + b])]++] =
                    ((uint32_t)t << 4) | ((uint32_t)a << 2) | (uint32_t)b;
    row_ptr_.alloc(nv_ + 1);
    row_col_.alloc(std::max<long long>(nnzb_, 1));
    row_of_.alloc(std::max<long long>(nnzb_, 1));
    row_val_.alloc(9 * std::max<long long>(nnzb_, 1));
    contrib_ptr_.alloc(nnzb_ + 1);
    contrib_.alloc(std::max<size_t>(contrib.size(), 1));
    SBS_HIP(hipMemcpy(row_ptr_.get(), rowptr.data(), rowptr.size() * sizeof(long long), hipMemcpyHostToDevice));
    SBS_HIP(hipMemcpy(row_col_.get(), col.data(), col.size() * sizeof(int), hipMemcpyHostToDevice));
    SBS_HIP(hipMemcpy(row_of_.get(), row_of.data(), row_of.size() * sizeof(int), hipMemcpyHostToDevice));
    SBS_HIP(hipMemcpy(contrib_ptr_.get(), cnt.data(), cnt.size() * sizeof(long long), hipMemcpyHostToDevice));
    SBS_HIP(hipMemcpy(contrib_.get(), contrib.data(), contrib.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
}

IpcEngine::~IpcEngine() {
    if (host_) (void)hipHostFree(host_);
    if (host_ctl_) (void)hipHostFree(host_ctl_);
}

void IpcEngine::set_stream(hipStream_t s) {
    stream_ = s;
    col_->set_stream(s);
}

double IpcEngine::reduce_sum(const double* parts, int n) {
    hipLaunchKernelGGL(sum_parts_kernel<1>, dim3(1), dim3(kT), 0, stream_, n, parts, out_.get());
    SBS_HIP(hipMemcpyAsync(host_, out_.get(), sizeof(double), hipMemcpyDeviceToHost, stream_));
    SBS_HIP(hipStreamSynchronize(stream_));
    return host_[0];
}

double IpcEngine::max_abs(const double* v, size_t n) {
    const unsigned nb = std::min<unsigned>(nblk(n), 1024);
    hipLaunchKernelGGL(max_abs_partial_kernel, dim3(nb), dim3(kT), 0, stream_, n, v, parts_.get());
    hipLaunchKernelGGL(max_final_kernel, dim3(1), dim3(64), 0, stream_, (int)nb, (const double*)parts_.get(),
                       out_.get());
    SBS_HIP(hipMemcpyAsync(host_, out_.get(), sizeof(double), hipMemcpyDeviceToHost, stream_));
    SBS_HIP(hipStreamSynchronize(stream_));
    return host_[0];
}

// UpdateQueries (broadphase.cu:574-587): candidates of the boxes at X grown by
// 2 dhat, then the distance type and squared distance of each.
void IpcEngine::update_queries(const double* X) {
    nq_ = ntri_ >= 2 ? col_->proximity(X, 2 * p_.dhat, ignore_self_) : 0;
    if ((size_t)nq_ > qcap_) {
        qcap_ = (size_t)nq_ + (size_t)nq_ / 4 + 16;
        qtype_.alloc(qcap_);
        qd_.alloc(qcap_);
        hq_.alloc(144 * qcap_);
        const size_t need = 2 * (size_t)nblk(std::max<size_t>({3 * (size_t)nv_, (size_t)nt_, qcap_})) + 64;
        if (need > parts_.n) parts_.alloc(need);
    }
    if (nq_ == 0) {
        stats_.queries = 0;
        return;
    }
    unsigned long long* kept = reinterpret_cast<unsigned long long*>(out_.get() + 7);
    SBS_HIP(hipMemsetAsync(kept, 0, sizeof(unsigned long long), stream_));
    hipLaunchKernelGGL(ipc_query_kernel, dim3(nblk(nq_)), dim3(kT), 0, stream_, nq_, col_->queries(),
                       (const double*)X, p_.dhat, qtype_.get(), qd_.get(), kept);
    SBS_HIP(hipGetLastError());
    unsigned long long nk = 0;
    SBS_HIP(hipMemcpyAsync(&nk, kept, sizeof(nk), hipMemcpyDeviceToHost, stream_));
    SBS_HIP(hipStreamSynchronize(stream_));
    stats_.queries = (long long)nk;
}

// UpdateQueries (broadphase.cu:574-587) as CollisionDetection exposes it: the
// candidates of the triangle pairs whose boxes at X, grown by 2 dhat, overlap
// (the detector's last ignore-self flag), their distance types and squared
// distances, and the queries with d <= dhat, sorted by (d, distance type) as
// the reference's final sort orders them, ties by the canonical vertex list.
long long update_queries_on(CollisionDetector<double>& cd, const double* X, double dhat, bool ignore_self,
                            std::vector<uint32_t>& q5, std::vector<int>& dtype, std::vector<double>& dist) {
    q5.clear();
    dtype.clear();
    dist.clear();
    const long long n = cd.proximity(X, 2 * dhat, ignore_self);
    if (n == 0) return 0;
    DevBuf<int> t(n);
    DevBuf<double> d(n);
    DevBuf<unsigned long long> kept(1);
    SBS_HIP(hipMemset(kept.get(), 0, sizeof(unsigned long long)));
    hipLaunchKernelGGL(ipc_query_kernel, dim3(nblk(n)), dim3(kT), 0, nullptr, n, cd.queries(), X, dhat, t.get(),
                       d.get(), kept.get());
    SBS_HIP(hipGetLastError());
    std::vector<QueryKey> qh(n);
    std::vector<int> th(n);
    std::vector<double> dh(n);
    SBS_HIP(hipMemcpy(qh.data(), cd.queries(), n * sizeof(QueryKey), hipMemcpyDeviceToHost));
    SBS_HIP(hipMemcpy(th.data(), t.get(), n * sizeof(int), hipMemcpyDeviceToHost));
    SBS_HIP(hipMemcpy(dh.data(), d.get(), n * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<long long> idx;
    for (long long i = 0; i < n; ++i)
        if (dh[i] <= dhat) idx.push_back(i);
    auto key = [&](long long i) {
        return std::make_tuple(dh[i], th[i], qh[i].type, qh[i].v0, qh[i].v1, qh[i].v2, qh[i].v3);
    };
    std::sort(idx.begin(), idx.end(), [&](long long a, long long b) { return key(a) < key(b); });
    for (long long i : idx) {
        q5.insert(q5.end(), {qh[i].type, qh[i].v0, qh[i].v1, qh[i].v2, qh[i].v3});
        dtype.push_back(th[i]);
        dist.push_back(dh[i]);
    }
    return (long long)idx.size();
}

double IpcEngine::energy(const double* X) {
    const size_t n = std::max<size_t>({(size_t)nv_, (size_t)nt_, (size_t)nq_});
    const unsigned nb = nblk(n);
    hipLaunchKernelGGL(ipc_energy_kernel, dim3(nb), dim3(kT), 0, stream_, nv_, nt_, nq_, X, (const double*)d_.XTilde,
                       d_.mass, d_.Tet, (const double*)dm_.get(), (const double*)v0_.get(), d_.mu, d_.lambda,
                       (const double*)qd_.get(), p_.dhat, kappa_, p_.dt * p_.dt, bodies_of(d_.fixed_bodies),
                       (const double*)area_.get(), (int)(elastic_ == kElasticNeoHookean), parts_.get());
    SBS_HIP(hipGetLastError());
    return reduce_sum(parts_.get(), (int)nb);
}

// IPEnergy::GradientHessian (ip.cu:54-75): inertia (coef 1), gravity,
// elasticity and barrier (coef h^2); with `hessian` the element Hessians,
// the Jacobi diagonal and the DOF elimination.
void IpcEngine::gradient(const double* X, bool hessian, double kap) {
    const double h2 = p_.dt * p_.dt;
    hipLaunchKernelGGL(ipc_grad_vertex_kernel, dim3(nblk(nv_)), dim3(kT), 0, stream_, nv_, X,
                       (const double*)d_.XTilde, d_.mass, h2, g_.get(), hessian ? diag_.get() : nullptr);
    elastic(X, h2, hessian);
    body_gradient(X, h2, hessian);
    if (nq_ > 0)
        hipLaunchKernelGGL(ipc_barrier_kernel, dim3((unsigned)((nq_ + kQB - 1) / kQB)), dim3(kQB), 0, stream_, nq_,
                           col_->queries(), (const int*)qtype_.get(), (const double*)qd_.get(), X, p_.dhat, kap, h2,
                           (const unsigned char*)fixed_.get(), g_.get(), hessian ? hq_.get() : nullptr,
                           hessian ? diag_.get() : nullptr);
    if (hessian)
        hipLaunchKernelGGL(ipc_fix_kernel, dim3(nblk(nv_)), dim3(kT), 0, stream_, nv_,
                           (const unsigned char*)fixed_.get(), g_.get(), diag_.get());
    SBS_HIP(hipGetLastError());
}

// The elastic term of IPEnergy (ip.cu:22 builds CorotatedEnergy; the
// Neo-Hookean energy is the one-line swap there): coef * gradient into g and,
// with `hessian`, the element Hessians and their Jacobi diagonal.
void IpcEngine::elastic(const double* X, double coef, bool hessian) {
    auto kern = elastic_ == kElasticNeoHookean ? ipc_nh_kernel : ipc_elastic_kernel;
    hipLaunchKernelGGL(kern, dim3(nblk(nt_)), dim3(kT), 0, stream_, nt_, d_.Tet, X, (const double*)dm_.get(),
                       (const double*)v0_.get(), d_.mu, d_.lambda, coef, (const unsigned char*)fixed_.get(), g_.get(),
                       hessian ? ht_.get() : nullptr, hessian ? diag_.get() : nullptr);
}

// ImplicitBarrierEnergy::GradientHessian (implicitBarrier.cu:150-240): adds
// coef * gradient to g and, with `hessian`, stores the per-vertex blocks and
// adds their diagonal to the Jacobi diagonal.
void IpcEngine::body_gradient(const double* X, double coef, bool hessian) {
    const Bodies b = bodies_of(d_.fixed_bodies);
    if (hessian) hv_live_ = b.rec != nullptr;
    if (!b.rec) return;
    hipLaunchKernelGGL(ipc_body_grad_kernel, dim3(nblk(nv_)), dim3(kT), 0, stream_, nv_, X, b,
                       (const double*)area_.get(), p_.dhat, coef, g_.get(), hessian ? hv_.get() : nullptr,
                       hessian ? diag_.get() : nullptr);
    SBS_HIP(hipGetLastError());
}

// ImplicitBarrierEnergy::InitStepSize (implicitBarrier.cu:352-407).
double IpcEngine::body_step(const double* X, const double* p) {
    const Bodies b = bodies_of(d_.fixed_bodies);
    if (!b.rec) return 1.0;
    const unsigned nb = nblk(nv_);
    hipLaunchKernelGGL(ipc_body_step_kernel, dim3(nb), dim3(kT), 0, stream_, nv_, X, p, b, parts_.get());
    hipLaunchKernelGGL(min_final_kernel, dim3(1), dim3(64), 0, stream_, (int)nb, (const double*)parts_.get(),
                       out_.get());
    SBS_HIP(hipGetLastError());
    SBS_HIP(hipMemcpyAsync(host_, out_.get(), sizeof(double), hipMemcpyDeviceToHost, stream_));
    SBS_HIP(hipStreamSynchronize(stream_));
    return host_[0];
}

// UpdateKappa (ip.cu:77-115): kappa = max(100, max|grad of the other terms| /
// max|grad of the barrier with kappa = 1|), 100 when the latter is below 1e-9.
void IpcEngine::update_kappa(const double* X) {
    const size_t n3 = 3 * (size_t)nv_;
    const double h2 = p_.dt * p_.dt;
    hipLaunchKernelGGL(ipc_grad_vertex_kernel, dim3(nblk(nv_)), dim3(kT), 0, stream_, nv_, X,
                       (const double*)d_.XTilde, d_.mass, h2, g_.get(), (double*)nullptr);
    elastic(X, h2, false);
    SBS_HIP(hipGetLastError());
    body_gradient(X, h2, false);
    const double ge = max_abs(g_.get(), n3);
    SBS_HIP(hipMemsetAsync(g_.get(), 0, g_.bytes(), stream_));
    if (nq_ > 0)
        hipLaunchKernelGGL(ipc_barrier_kernel, dim3((unsigned)((nq_ + kQB - 1) / kQB)), dim3(kQB), 0, stream_, nq_,
                           col_->queries(), (const int*)qtype_.get(), (const double*)qd_.get(), X, p_.dhat, 1.0, h2,
                           (const unsigned char*)fixed_.get(), g_.get(), (double*)nullptr, (double*)nullptr);
    SBS_HIP(hipGetLastError());
    const double gb = nq_ > 0 ? max_abs(g_.get(), n3) : 0.0;
    kappa_ = gb > 1e-9 ? std::max(100.0, ge / gb) : 100.0;
}

// Jacobi-PCG on H p = g; returns the iterations.  An iteration is two
// launches (the product with the previous iteration's stop test, then the
// update); the host reads the control word at points predicted from the
// previous solve (see the loop), and launches after the solve finished
// return at once.
int IpcEngine::solve() {
    constexpr int kPcgChunk = 16;
    const int n = 3 * nv_;
    const unsigned nb = nblk(n), nrb = nblk((size_t)kRowLanes * n);
    double* out = out_.get();  // [0] rz0, [1] rr, [2] pq, [3] alpha, [4] beta, [5] |b|, [6..7] rz by parity
    int* ctl = pcg_ctl_.get();
    double* upd_parts = pcg_parts_.get();
    double* prod_parts = prod_parts_.get();
    hipLaunchKernelGGL(ipc_pcg_init_kernel, dim3(nb), dim3(kT), 0, stream_, n, (const double*)g_.get(),
                       (const double*)diag_.get(), p_dir_.get(), r_.get(), z_.get(), upd_parts);
    hipLaunchKernelGGL(sum_parts_kernel<2>, dim3(1), dim3(kT), 0, stream_, (int)nb, (const double*)upd_parts, out);
    SBS_HIP(hipMemcpyAsync(host_, out, 2 * sizeof(double), hipMemcpyDeviceToHost, stream_));
    SBS_HIP(hipStreamSynchronize(stream_));
    const double bn = std::sqrt(host_[1]);
    if (bn == 0.0) return 0;
    host_[5] = bn;
    host_[6] = host_[0];
    host_ctl_[0] = 0;
    host_ctl_[1] = 0;
    SBS_HIP(hipMemcpyAsync(out + 5, host_ + 5, 2 * sizeof(double), hipMemcpyHostToDevice, stream_));
    SBS_HIP(hipMemcpyAsync(ctl, host_ctl_, 2 * sizeof(int), hipMemcpyHostToDevice, stream_));
    double* dir[2] = {xtmp_.get(), p2_.get()};  // p by iteration parity (x_tmp is free during the solve)
    const double* hv = hv_live_ ? (const double*)hv_.get() : nullptr;
    const bool contact = nq_ > 0;
    auto product = [&](int it) {
        auto kern = it == 0 ? ipc_pcg_prod_kernel<true> : ipc_pcg_prod_kernel<false>;
        hipLaunchKernelGGL(kern, dim3(nrb), dim3(kT), 0, stream_, nv_, (const long long*)row_ptr_.get(),
                           (const int*)row_col_.get(), (const double*)row_val_.get(),
                           contact ? (const int*)crow_.get() : nullptr, (const int*)cbcol_.get(),
                           (const double*)cval_.get(), (const unsigned char*)fixed_.get(), d_.mass, hv,
                           (const double*)z_.get(), (const double*)dir[it & 1], dir[(it + 1) & 1], q_.get(), (int)nb,
                           (const double*)upd_parts, it, p_.lin_max_iter, p_.lin_tol, prod_parts, out, ctl);
    };
    for (int launched = 0; launched < p_.lin_max_iter;) {
        // the first check after 9/10 of the previous solve's iterations, then
        // every kPcgChunk or launched/8 (a check costs the GPU ~30 us idle, a
        // launch after the solve finished ~2.5 us)
        const int want = launched == 0 ? std::max(kPcgChunk, pcg_last_ * 9 / 10) : std::max(kPcgChunk, launched / 8);
        const int chunk = std::min(want, p_.lin_max_iter - launched);
        for (int j = 0; j < chunk; ++j) {
            const int it = launched + j;
            product(it);
            hipLaunchKernelGGL(ipc_pcg_update_kernel, dim3(nb), dim3(kT), 0, stream_, n, (int)nrb,
                               (const double*)prod_parts, it, (const double*)dir[(it + 1) & 1], (const double*)q_.get(),
                               (const double*)diag_.get(), p_dir_.get(), r_.get(), z_.get(), upd_parts, out, ctl);
        }
        launched += chunk;
        if (launched == p_.lin_max_iter) product(launched);  // only the stop test of the last iteration
        SBS_HIP(hipGetLastError());
        SBS_HIP(hipMemcpyAsync(host_ctl_, ctl, 2 * sizeof(int), hipMemcpyDeviceToHost, stream_));
        SBS_HIP(hipStreamSynchronize(stream_));
        if (host_ctl_[0]) break;
    }
    pcg_last_ = host_ctl_[1];
    return host_ctl_[1];
}

// The contact rows of this Newton iteration from the active queries' Hessians.
void IpcEngine::assemble_contact() {
    const long long ne = 16 * nq_;
    const unsigned long long nvl = (unsigned long long)nv_;
    const unsigned long long sentinel = nvl * nvl;
    int bits = 1;
    while (bits < 64 && (sentinel >> bits) != 0) ++bits;
    SBS_REQUIRE(ne < INT32_MAX && sentinel / nvl == nvl, "IPCSolver: too many contact pairs");
    if ((size_t)ne > ecap_) {
        ecap_ = (size_t)ne + (size_t)ne / 4 + 64;
        ckey_.alloc(ecap_);
        ckey_s_.alloc(ecap_);
        ccode_.alloc(ecap_);
        ccode_s_.alloc(ecap_);
        cflag_.alloc(ecap_);
        cpos_.alloc(ecap_);
        chead_.alloc(ecap_);
        cbrow_.alloc(ecap_);
        cbcol_.alloc(ecap_);
        cval_.alloc(9 * ecap_);
    }
    size_t b1 = 0, b2 = 0;
    SBS_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, b1, ckey_.get(), ckey_s_.get(), ccode_.get(), ccode_s_.get(),
                                               (int)ne, 0, bits, stream_));
    SBS_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, b2, cflag_.get(), cpos_.get(), (int)ne, stream_));
    if (std::max(b1, b2) > ctmp_bytes_) {
        ctmp_bytes_ = std::max(b1, b2);
        ctmp_.alloc(ctmp_bytes_);
    }
    const unsigned eb = nblk((size_t)ne);
    hipLaunchKernelGGL(ipc_contact_expand_kernel, dim3(eb), dim3(kT), 0, stream_, nq_, col_->queries(),
                       (const double*)qd_.get(), p_.dhat * p_.dhat, (const unsigned char*)fixed_.get(), nvl,
                       ckey_.get(), ccode_.get());
    SBS_HIP(hipcub::DeviceRadixSort::SortPairs(ctmp_.get(), b1, ckey_.get(), ckey_s_.get(), ccode_.get(),
                                               ccode_s_.get(), (int)ne, 0, bits, stream_));
    hipLaunchKernelGGL(ipc_contact_heads_kernel, dim3(eb), dim3(kT), 0, stream_, ne,
                       (const unsigned long long*)ckey_s_.get(), sentinel, cflag_.get());
    SBS_HIP(hipcub::DeviceScan::ExclusiveSum(ctmp_.get(), b2, cflag_.get(), cpos_.get(), (int)ne, stream_));
    int* count = ccount_.get();
    hipLaunchKernelGGL(ipc_contact_blocks_kernel, dim3(eb), dim3(kT), 0, stream_, ne,
                       (const unsigned long long*)ckey_s_.get(), nvl, (const int*)cflag_.get(),
                       (const int*)cpos_.get(), chead_.get(), cbrow_.get(), cbcol_.get(), count);
    hipLaunchKernelGGL(ipc_contact_values_kernel, dim3(nblk(9 * (size_t)ne)), dim3(kT), 0, stream_, ne,
                       (const int*)count, (const int*)chead_.get(), (const uint32_t*)ccode_s_.get(),
                       (const double*)hq_.get(), cval_.get());
    hipLaunchKernelGGL(ipc_contact_rows_kernel, dim3(nblk((size_t)nv_ + 1)), dim3(kT), 0, stream_, nv_,
                       (const int*)count, (const int*)cbrow_.get(), crow_.get());
    SBS_HIP(hipGetLastError());
}

// SearchDirection (ipc.cu:230-239): gradient and Hessian at X, elimination,
// p = H^-1 g; false when the gradient holds NaN or Inf.
bool IpcEngine::search_direction(const double* X) {
    gradient(X, true, kappa_);
    hipLaunchKernelGGL(ipc_assemble_rows_kernel, dim3(nblk(9 * (size_t)nnzb_)), dim3(kT), 0, stream_, nnzb_,
                       (const int*)row_of_.get(), (const int*)row_col_.get(), (const long long*)contrib_ptr_.get(),
                       (const uint32_t*)contrib_.get(), (const double*)ht_.get(), (const unsigned char*)fixed_.get(),
                       row_val_.get());
    if (nq_ > 0) assemble_contact();
    SBS_HIP(hipGetLastError());
    const double gm = max_abs(g_.get(), 3 * (size_t)nv_);
    if (!std::isfinite(gm)) return false;
    stats_.linear_iterations += solve();
    return true;
}

bool IpcEngine::step(IpcData& data, const IpcParams& p) {
    d_ = data;
    p_ = p;
    const double h = p.dt;
    const int n3 = 3 * nv_;
    const unsigned vb = nblk(n3);
    stats_.newton_iterations = 0;
    stats_.line_search_halvings = 0;
    stats_.linear_iterations = 0;
    hipLaunchKernelGGL(ipc_begin_kernel, dim3(vb), dim3(kT), 0, stream_, n3, (const double*)data.X,
                       (const double*)data.V, h, p.damp, data.XTilde, xn_.get());
    SBS_HIP(hipGetLastError());
    update_queries(data.X);
    update_kappa(data.X);
    stats_.kappa = kappa_;
    double E_last = energy(data.X);
    if (!std::isfinite(E_last)) return false;
    if (!search_direction(data.X)) return false;
    double* X = data.X;
    double* xt = xtmp_.get();
    while (!(max_abs(p_dir_.get(), n3) / h < p.tol)) {  // EndCondition (ipc.cu:243-250)
        if (++stats_.newton_iterations > p.max_iterations) return false;
        hipLaunchKernelGGL(ipc_axpy_kernel, dim3(vb), dim3(kT), 0, stream_, n3, (const double*)X,
                           (const double*)p_dir_.get(), 1.0, xt);
        // InitStepSize (ip.cu:116-123, barrier.cu:225-228): CCD over X -> X - p,
        // and the fixed bodies' bound (1 without bodies)
        const double toi = ntri_ >= 2 ? col_->min_step(X, xt, true) : 1.0;
        ignore_self_ = true;
        const double af = body_step(X, p_dir_.get());
        double alpha = 0.95 * std::min(af, 0.9 * toi);
        alpha = alpha < 1e-12 ? 0.0 : std::min(1.0, alpha);
        if (stats_.newton_iterations == 1) stats_.alpha0 = alpha;
        double E = E_last;
        while (true) {
            hipLaunchKernelGGL(ipc_axpy_kernel, dim3(vb), dim3(kT), 0, stream_, n3, (const double*)X,
                               (const double*)p_dir_.get(), alpha, xt);
            update_queries(xt);
            E = energy(xt);
            if (!std::isfinite(E)) return false;
            if (E > E_last) {
                alpha /= 2;
                ++stats_.line_search_halvings;
            } else {
                break;
            }
            if (alpha < DBL_EPSILON) break;
        }
        // X takes the last trial point: its queries and energy are the ones
        // just computed there (the reference recomputes both at X, ipc.cu:209-214)
        SBS_HIP(hipMemcpyAsync(X, xt, n3 * sizeof(double), hipMemcpyDeviceToDevice, stream_));
        E_last = E;
        if (!search_direction(X)) return false;
    }
    stats_.energy = E_last;
    hipLaunchKernelGGL(ipc_velocity_kernel, dim3(vb), dim3(kT), 0, stream_, n3, (const double*)X,
                       (const double*)xn_.get(), 1.0 / h, data.V);
    SBS_HIP(hipGetLastError());
    SBS_HIP(hipStreamSynchronize(stream_));
    return true;
}

bool IpcEngine::update(IpcData& data, const IpcParams& params) {
    SBS_REQUIRE(params.dt > 0 && params.dhat > 0 && params.tol > 0, "IPCSolver: dt, dhat and tol must be positive");
    SBS_REQUIRE(params.max_iterations >= 1 && params.lin_max_iter >= 1, "IPCSolver: iteration limits must be >= 1");
    if (failed_) return false;
    const bool ok = step(data, params);
    if (!ok) failed_ = true;
    stats_.failed = failed_ ? 1 : 0;
    return ok;
}

void IpcEngine::set_state(const double* XTilde, double kappa) {
    SBS_HIP(hipMemcpy(d_.XTilde, XTilde, 3 * (size_t)nv_ * sizeof(double), hipMemcpyHostToDevice));
    kappa_ = kappa;
}

double IpcEngine::energy_at(const double* Xh, double* grad_host) {
    SBS_HIP(hipMemcpyAsync(xtmp_.get(), Xh, 3 * (size_t)nv_ * sizeof(double), hipMemcpyHostToDevice, stream_));
    update_queries(xtmp_.get());
    const double e = energy(xtmp_.get());
    if (grad_host) {
        gradient(xtmp_.get(), false, kappa_);
        SBS_HIP(hipMemcpyAsync(grad_host, g_.get(), 3 * (size_t)nv_ * sizeof(double), hipMemcpyDeviceToHost,
                               stream_));
        SBS_HIP(hipStreamSynchronize(stream_));
    }
    return e;
}

}  // namespace sbs
