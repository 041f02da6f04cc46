// gfx950 kernels and host orchestration of the projective-dynamics step.
// See pd_engine.hpp for the layout; reference citations are per kernel.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "direct_solver.hpp"
#include "pd_engine.hpp"
#include "svd3.hpp"

namespace sbs {

namespace {

constexpr int kTpb = 256;   // 4 wavefronts per block
constexpr int kSlice = 64;  // sliced-ELL slice height = one wavefront
constexpr int kPcgChunk = 8;   // iterations per host check in exact mode
constexpr int kPcgMargin = 2;  // adaptive cap = most iterations seen + margin

// Pin weight of computeDBCLocal: positional_weight * dt2Inv with
// dt2Inv = dtInv * dtInv in the solver precision (pdSolver.cu:145-147, :171;
// positional_weight = 1e6, pdSolver.h:31).
template <class T> T pin_weight(double dt) {
    const T dt_inv = (T)1 / (T)dt;
    const T dt2_inv = dt_inv * dt_inv;
    return (T)1e6 * dt2_inv;
}

template <class T> __device__ __forceinline__ void ld3(const T* __restrict__ p, unsigned i, T& x, T& y, T& z) {
    x = p[3 * (size_t)i];
    y = p[3 * (size_t)i + 1];
    z = p[3 * (size_t)i + 2];
}
template <class T> __device__ __forceinline__ void st3(T* __restrict__ p, unsigned i, T x, T y, T z) {
    p[3 * (size_t)i] = x;
    p[3 * (size_t)i + 1] = y;
    p[3 * (size_t)i + 2] = z;
}

// ---------------------------------------------------------------------------
// FEM precompute: computeInvDmV0 (src/simulation/solver/solverUtil.cuh:109-137).
// Dm = [x1-x0, x2-x0, x3-x0] (columns), Dm^-1 stored row-major in 9 planes,
// V0 = |det Dm| / 6, w = |V0| * mu (the coefficient of pdUtil.cu:19 and :126).
template <class T>
__global__ void __launch_bounds__(kTpb) fem_precompute_kernel(int nt, const T* __restrict__ X,
                                                              const uint4* __restrict__ tets,
                                                              const T* __restrict__ mu, T* __restrict__ dminv,
                                                              T* __restrict__ v0, T* __restrict__ w,
                                                              T* __restrict__ v0_out) {
    const unsigned t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= (unsigned)nt) return;
    const uint4 tv = tets[t];
    T p0[3], p1[3], p2[3], p3[3];
    ld3(X, tv.x, p0[0], p0[1], p0[2]);
    ld3(X, tv.y, p1[0], p1[1], p1[2]);
    ld3(X, tv.z, p2[0], p2[1], p2[2]);
    ld3(X, tv.w, p3[0], p3[1], p3[2]);
    T D[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        D[r][0] = p1[r] - p0[r];
        D[r][1] = p2[r] - p0[r];
        D[r][2] = p3[r] - p0[r];
    }
    const T d = det3(D);
    T inv[9];
    inv[0] = (D[1][1] * D[2][2] - D[1][2] * D[2][1]) / d;
    inv[1] = (D[0][2] * D[2][1] - D[0][1] * D[2][2]) / d;
    inv[2] = (D[0][1] * D[1][2] - D[0][2] * D[1][1]) / d;
    inv[3] = (D[1][2] * D[2][0] - D[1][0] * D[2][2]) / d;
    inv[4] = (D[0][0] * D[2][2] - D[0][2] * D[2][0]) / d;
    inv[5] = (D[0][2] * D[1][0] - D[0][0] * D[1][2]) / d;
    inv[6] = (D[1][0] * D[2][1] - D[1][1] * D[2][0]) / d;
    inv[7] = (D[0][1] * D[2][0] - D[0][0] * D[2][1]) / d;
    inv[8] = (D[0][0] * D[1][1] - D[0][1] * D[1][0]) / d;
#pragma unroll
    for (int k = 0; k < 9; ++k) dminv[(size_t)k * nt + t] = inv[k];
    const T vol = fabs(d) / (T)6;
    v0[t] = vol;
    w[t] = fabs(vol) * mu[t];
    if (v0_out) v0_out[t] = vol;
}

// Gradient operator rows of the linear tet: g0 = -(m0+m1+m2), g_{j+1} = m_j,
// where m_j is row j of Dm^-1 (pdUtil.cu:16, glm mat4x3 {-1,-1,-1,1,0,0,0,1,0,0,0,1}).
template <class T> __device__ __forceinline__ void load_dminv(const T* __restrict__ dminv, int nt, unsigned t, T (&m)[9]) {
#pragma unroll
    for (int k = 0; k < 9; ++k) m[k] = dminv[(size_t)k * nt + t];
}

// ---------------------------------------------------------------------------
// Assembly of L_s = sum_t w_t G Dm^-1 Dm^-T G^T (computeSiTSi, pdUtil.cu:9-40):
// diagonal into diagL (the reference's matrix_diag), off-diagonal entries
// into the sliced-ELL values found by a search of the (short) row.
template <class T>
__global__ void __launch_bounds__(kTpb) assemble_kernel(int nt, const uint4* __restrict__ tets,
                                                        const T* __restrict__ dminv, const T* __restrict__ w,
                                                        const int* __restrict__ slice_ptr,
                                                        const int* __restrict__ rowlen,
                                                        const uint32_t* __restrict__ ecol, T* __restrict__ eval,
                                                        T* __restrict__ diagL) {
    const unsigned t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= (unsigned)nt) return;
    const uint4 tv = tets[t];
    const unsigned v[4] = {tv.x, tv.y, tv.z, tv.w};
    T m[9];
    load_dminv(dminv, nt, t, m);
    T g[4][3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        g[0][k] = -(m[k] + m[3 + k] + m[6 + k]);
        g[1][k] = m[k];
        g[2][k] = m[3 + k];
        g[3][k] = m[6 + k];
    }
    const T ww = w[t];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const unsigned row = v[i];
        const size_t base = (size_t)slice_ptr[row / kSlice] + (row % kSlice);
        const int len = rowlen[row];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const T kij = ww * (g[i][0] * g[j][0] + g[i][1] * g[j][1] + g[i][2] * g[j][2]);
            if (i == j) {
                atomicAdd(&diagL[row], kij);
            } else {
                for (int e = 0; e < len; ++e) {
                    const size_t pos = base + (size_t)e * kSlice;
                    if (ecol[pos] == v[j]) {
                        atomicAdd(&eval[pos], kij);
                        break;
                    }
                }
            }
        }
    }
}

// c = (m + DBC*weight)/dt^2 (setMDt_2, pdUtil.cu:42-54), diag(A_s) = diagL + c.
template <class T>
__global__ void mass_diag_kernel(int nv, const T* __restrict__ mass, const T* __restrict__ dbc, T weight, T dt2,
                                 const T* __restrict__ diagL, T* __restrict__ c, T* __restrict__ diagA,
                                 T* __restrict__ dinv) {
    const unsigned v = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (v >= (unsigned)nv) return;
    const T cv = (mass[v] + dbc[v] * weight) / dt2;
    c[v] = cv;
    const T d = diagL[v] + cv;
    diagA[v] = d;
    dinv[v] = (T)1 / d;
}

// ---------------------------------------------------------------------------
// Step prologue: gravity force (pdSolver.cu:154-155), setMDt_2MoreDBC
// (pdUtil.cu:56-71), computeSn (pdUtil.cu:73-95), sn_old copy, prev init
// (pdSolver.cu:158-162) and the first addM_h2Sn (pdUtil.cu:168-178).
template <class T>
__global__ void begin_step_kernel(int nv, const T* __restrict__ X, const T* __restrict__ V,
                                  const T* __restrict__ mass, const T* __restrict__ dbc,
                                  const T* __restrict__ more_dbc, T dt, T gravity, T* __restrict__ c,
                                  T* __restrict__ sn, T* __restrict__ sn_old, T* __restrict__ prev,
                                  T* __restrict__ b, int prev_from_sn) {
    const unsigned v = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (v >= (unsigned)nv) return;
    const T m = mass[v];
    const T dt2 = dt * dt;
    T cv = c[v];
    if (dbc[v] == (T)0) {
        const T wi = more_dbc ? more_dbc[v] : (T)0;
        cv = wi > (T)0 ? (m + wi) / dt2 : m / dt2;
        c[v] = cv;
    }
    T x, y, z, vx, vy, vz;
    ld3(X, v, x, y, z);
    ld3(V, v, vx, vy, vz);
    const T inv = (T)1 / cv;
    const T fy = -gravity * m;
    const T sx = x + dt * vx, sy = y + dt * vy + inv * fy, sz = z + dt * vz;
    st3(sn, v, sx, sy, sz);
    st3(sn_old, v, sx, sy, sz);
    if (prev_from_sn) st3(prev, v, sx, sy, sz);
    else st3(prev, v, (T)0, (T)0, (T)0);
    st3(b, v, cv * sx, cv * sy, cv * sz);
}

// ---------------------------------------------------------------------------
// computeLocal (src/simulation/solver/projective/pdUtil.cu:97-145): one thread
// per tet, F = Ds Dm^-1, 3x3 SVD in registers, R (or R - F), force columns
// w R Dm^-T G^T scattered into b with 12 float atomics.
template <class T>
__global__ void __launch_bounds__(kTpb) local_step_kernel(int nt, const uint4* __restrict__ tets,
                                                          const T* __restrict__ dminv, const T* __restrict__ w,
                                                          const T* __restrict__ q, T* __restrict__ b,
                                                          int residual_form, const int* __restrict__ skip) {
    if (skip && *skip) return;
    const unsigned t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= (unsigned)nt) return;
    const uint4 tv = tets[t];
    T m[9];
    load_dminv(dminv, nt, t, m);
    const T ww = w[t];
    T p0[3], p1[3], p2[3], p3[3];
    ld3(q, tv.x, p0[0], p0[1], p0[2]);
    ld3(q, tv.y, p1[0], p1[1], p1[2]);
    ld3(q, tv.z, p2[0], p2[1], p2[2]);
    ld3(q, tv.w, p3[0], p3[1], p3[2]);
    T F[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const T d0 = p1[r] - p0[r], d1 = p2[r] - p0[r], d2 = p3[r] - p0[r];
#pragma unroll
        for (int c = 0; c < 3; ++c) F[r][c] = d0 * m[c] + d1 * m[3 + c] + d2 * m[6 + c];
    }
    T P[3][3];
    pd_projection(F, residual_form != 0, P);
    // Q[:, j] = w * P * (row j of Dm^-1)
    T Q[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int j = 0; j < 3; ++j) Q[r][j] = ww * (P[r][0] * m[3 * j] + P[r][1] * m[3 * j + 1] + P[r][2] * m[3 * j + 2]);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        atomicAdd(&b[3 * (size_t)tv.x + r], -(Q[r][0] + Q[r][1] + Q[r][2]));
        atomicAdd(&b[3 * (size_t)tv.y + r], Q[r][0]);
        atomicAdd(&b[3 * (size_t)tv.z + r], Q[r][1]);
        atomicAdd(&b[3 * (size_t)tv.w + r], Q[r][2]);
    }
}

// b entries of pinned vertices are overwritten with DBCX * 1e6/h^2
// (computeDBCLocal, pdUtil.cu:147-166); used by the kernels below as a
// substitution on read instead of a separate pass.
template <class T>
__device__ __forceinline__ void rhs_at(const T* __restrict__ b, const T* __restrict__ dbc,
                                       const T* __restrict__ more_dbc, const T* __restrict__ dbcx, T wdbc,
                                       unsigned v, T& x, T& y, T& z) {
    const T pin = dbc[v];
    const T more = more_dbc ? more_dbc[v] : (T)0;
    if (pin > (T)0) {
        ld3(dbcx, v, x, y, z);
        x *= wdbc; y *= wdbc; z *= wdbc;
    } else if (more > (T)0) {
        ld3(dbcx, v, x, y, z);
        x *= more; y *= more; z *= more;
    } else {
        ld3(b, v, x, y, z);
    }
}

// Jacobi + Chebyshev global step (getErrorKern pdUtil.cu:195-214 and
// chebyshevKern pdUtil.cu:216-226 fused per vertex), followed by the next
// iteration's addM_h2Sn (b = c*sn_old) so no separate pass is needed.
template <class T>
__global__ void jacobi_chebyshev_kernel(int nv, T* __restrict__ b, const T* __restrict__ c,
                                        const T* __restrict__ diagL, const T* __restrict__ dbc,
                                        const T* __restrict__ more_dbc, const T* __restrict__ dbcx, T wdbc,
                                        T* __restrict__ sn, T* __restrict__ prev, const T* __restrict__ sn_old,
                                        T omega) {
    const unsigned v = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (v >= (unsigned)nv) return;
    T bx, by, bz;
    rhs_at(b, dbc, more_dbc, dbcx, wdbc, v, bx, by, bz);
    T sx, sy, sz, px, py, pz;
    ld3(sn, v, sx, sy, sz);
    ld3(prev, v, px, py, pz);
    const T cv = c[v];
    T nx, ny, nz;
    if (more_dbc && more_dbc[v] > (T)0) {
        nx = sx; ny = sy; nz = sz;
    } else {
        const T den = cv + diagL[v];
        nx = (bx - cv * sx) / den + sx;
        ny = (by - cv * sy) / den + sy;
        nz = (bz - cv * sz) / den + sz;
    }
    nx = (T)0.9 * (nx - sx) + sx;
    ny = (T)0.9 * (ny - sy) + sy;
    nz = (T)0.9 * (nz - sz) + sz;
    nx = (nx - px) * omega + px;
    ny = (ny - py) * omega + py;
    nz = (nz - pz) * omega + pz;
    st3(prev, v, sx, sy, sz);
    st3(sn, v, nx, ny, nz);
    T ox, oy, oz;
    ld3(sn_old, v, ox, oy, oz);
    st3(b, v, cv * ox, cv * oy, cv * oz);
}

// Chebyshev recurrence applied to a PCG solution held in `next`.
template <class T>
__global__ void chebyshev_kernel(int nv, const T* __restrict__ next, T* __restrict__ sn, T* __restrict__ prev,
                                 T omega, const int* __restrict__ skip) {
    if (skip && *skip) return;
    const unsigned i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= 3u * (unsigned)nv) return;
    const T s = sn[i];
    T nx = (T)0.9 * (next[i] - s) + s;
    nx = (nx - prev[i]) * omega + prev[i];
    prev[i] = s;
    sn[i] = nx;
}

// updateVelPos (pdUtil.cu:180-193) and X <- XTilde (pdSolver.cu:227).
template <class T>
__global__ void end_step_kernel(int nv, const T* __restrict__ sn, T dt_inv, T* __restrict__ XTilde,
                                T* __restrict__ V, T* __restrict__ X, const T* __restrict__ more_dbc,
                                int copy_x) {
    const unsigned v = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (v >= (unsigned)nv) return;
    T nx, ny, nz, ox, oy, oz;
    ld3(sn, v, nx, ny, nz);
    ld3(XTilde, v, ox, oy, oz);
    if (more_dbc && more_dbc[v] > (T)0) st3(V, v, (T)0, (T)0, (T)0);
    else st3(V, v, (nx - ox) * dt_inv, (ny - oy) * dt_inv, (nz - oz) * dt_inv);
    st3(XTilde, v, nx, ny, nz);
    if (copy_x) st3(X, v, nx, ny, nz);
}

// ---------------------------------------------------------------------------
// Jacobi-PCG over the scalar matrix with 3 fused right-hand sides.
// Reductions: each block reduces its rows with wave shuffles + LDS and writes
// one partial per quantity; the last block to arrive (atomic ticket) sums the
// partials in double and updates the control block, so no kernel launch or
// host round trip is spent on a dot product.
struct CtlBlock {
    unsigned ticket;
    int done;        // PCG converged or out of iterations
    int outer_done;  // local/global loop early exit (direct / PCG modes)
    int iters;       // PCG iterations of the current solve
    int total_iters; // PCG iterations since the start of the step
    int max_iter;    // hard limit per solve
    int launch_cap;  // iterations the host launched for each solve
    int cap_hits;    // solves that ran out of launched iterations unconverged
    int max_used;    // most iterations any solve of this step needed
    int solves;      // PCG solves started this step
    int active[3];
    double rz[3], thr[3], alpha[3], beta[3];
    double tol2, floor2, outer_tol, err;
};

__global__ void ctl_init_kernel(CtlBlock* ctl, int max_iter, int launch_cap, double tol2, double floor2,
                                double outer_tol) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    ctl->ticket = 0;
    ctl->done = 1;
    ctl->outer_done = 0;
    ctl->iters = 0;
    ctl->total_iters = 0;
    ctl->max_iter = max_iter;
    ctl->launch_cap = launch_cap;
    ctl->cap_hits = 0;
    ctl->max_used = 0;
    ctl->solves = 0;
    for (int k = 0; k < 3; ++k) {
        ctl->active[k] = 0;
        ctl->rz[k] = ctl->thr[k] = ctl->alpha[k] = ctl->beta[k] = 0.0;
    }
    ctl->tol2 = tol2;
    ctl->floor2 = floor2;
    ctl->outer_tol = outer_tol;
    ctl->err = 1.0;
}

template <int NQ>
__device__ __forceinline__ bool block_partials(const double (&vals)[NQ], double* __restrict__ partial,
                                               CtlBlock* ctl) {
    __shared__ double red[kTpb / 64][NQ];
    __shared__ bool last;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        const double s = wave_sum(vals[k]);
        if (lane == 0) red[wid][k] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < NQ; ++k) {
            double s = 0;
            for (int w2 = 0; w2 < kTpb / 64; ++w2) s += red[w2][k];
            partial[(size_t)blockIdx.x * NQ + k] = s;
        }
        __threadfence();
        const unsigned tk = atomicAdd(&ctl->ticket, 1u);
        last = (tk == gridDim.x - 1);
    }
    __syncthreads();
    return last;
}

template <int NQ>
__device__ __forceinline__ void last_block_sum(const double* __restrict__ partial, double (&out)[NQ]) {
    __shared__ double red[kTpb / 64][NQ];
    __threadfence();
    double acc[NQ];
#pragma unroll
    for (int k = 0; k < NQ; ++k) acc[k] = 0;
    for (unsigned i = threadIdx.x; i < gridDim.x; i += blockDim.x)
#pragma unroll
        for (int k = 0; k < NQ; ++k) acc[k] += __builtin_nontemporal_load(&partial[(size_t)i * NQ + k]);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        const double s = wave_sum(acc[k]);
        if (lane == 0) red[wid][k] = s;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        double s = 0;
        for (int w2 = 0; w2 < kTpb / 64; ++w2) s += red[w2][k];
        out[k] = s;
    }
}

// y = A_s x for one row, 3 components (sliced ELL, row-per-lane).
template <class T>
__device__ __forceinline__ void spmv_row(unsigned v, const int* __restrict__ slice_ptr, const int* __restrict__ slice_w,
                                         const uint32_t* __restrict__ ecol, const T* __restrict__ eval,
                                         const T* __restrict__ diagA, const T* __restrict__ x, T& y0, T& y1, T& y2) {
    T x0, x1, x2;
    ld3(x, v, x0, x1, x2);
    const T d = diagA[v];
    y0 = d * x0; y1 = d * x1; y2 = d * x2;
    const unsigned s = v / kSlice, lane = v % kSlice;
    const size_t base = (size_t)slice_ptr[s] + lane;
    const int width = slice_w[s];
    for (int e = 0; e < width; ++e) {
        const size_t pos = base + (size_t)e * kSlice;
        const uint32_t u = ecol[pos];
        const T a = eval[pos];
        T u0, u1, u2;
        ld3(x, u, u0, u1, u2);
        y0 += a * u0; y1 += a * u1; y2 += a * u2;
    }
}

// r = b - A x, z = D^-1 r, p = z, and the initial norms.  Also re-initialises
// b <- c*sn_old for the next outer iteration when reset_b is set, and copies
// the warm start into x_out (x_out may equal x).
template <class T>
__global__ void __launch_bounds__(kTpb) pcg_init_kernel(
    int nv, T* __restrict__ b, const T* __restrict__ dbc, const T* __restrict__ more_dbc,
    const T* __restrict__ dbcx, T wdbc, int dbc_override, const T* __restrict__ x, T* __restrict__ x_out,
    const int* __restrict__ slice_ptr, const int* __restrict__ slice_w, const uint32_t* __restrict__ ecol,
    const T* __restrict__ eval, const T* __restrict__ diagA, const T* __restrict__ dinv, T* __restrict__ r,
    T* __restrict__ p, int reset_b, const T* __restrict__ c, const T* __restrict__ sn_old,
    double* __restrict__ partial, CtlBlock* ctl) {
    const unsigned v = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const bool skip_all = ctl->outer_done != 0;
    double vals[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // rz[3], zz[3], dd[3]
    if (v < (unsigned)nv && !skip_all) {
        T bx, by, bz;
        if (dbc_override) rhs_at(b, dbc, more_dbc, dbcx, wdbc, v, bx, by, bz);
        else ld3(b, v, bx, by, bz);
        T a0, a1, a2;
        spmv_row(v, slice_ptr, slice_w, ecol, eval, diagA, x, a0, a1, a2);
        const T di = dinv[v];
        const T r0 = bx - a0, r1 = by - a1, r2 = bz - a2;
        const T z0 = di * r0, z1 = di * r1, z2 = di * r2;
        st3(r, v, r0, r1, r2);
        st3(p, v, z0, z1, z2);
        if (x_out != x) {
            T x0, x1, x2;
            ld3(x, v, x0, x1, x2);
            st3(x_out, v, x0, x1, x2);
        }
        vals[0] = (double)r0 * z0; vals[1] = (double)r1 * z1; vals[2] = (double)r2 * z2;
        vals[3] = (double)z0 * z0; vals[4] = (double)z1 * z1; vals[5] = (double)z2 * z2;
        const double d0 = (double)di * bx, d1 = (double)di * by, d2 = (double)di * bz;
        vals[6] = d0 * d0; vals[7] = d1 * d1; vals[8] = d2 * d2;
        if (reset_b) {
            T ox, oy, oz;
            ld3(sn_old, v, ox, oy, oz);
            const T cv = c[v];
            st3(b, v, cv * ox, cv * oy, cv * oz);
        }
    }
    if (skip_all) return;
    if (block_partials<9>(vals, partial, ctl)) {
        double s[9];
        last_block_sum<9>(partial, s);
        if (threadIdx.x == 0) {
            int any = 0;
            for (int k = 0; k < 3; ++k) {
                ctl->rz[k] = s[k];
                const double thr = fmax(ctl->tol2 * s[3 + k], ctl->floor2 * s[6 + k]);
                ctl->thr[k] = thr;
                ctl->active[k] = s[3 + k] > thr;
                any |= ctl->active[k];
                ctl->beta[k] = 0;
            }
            ctl->iters = 0;
            ctl->done = !any;
            ctl->solves += 1;
            ctl->ticket = 0;
        }
    }
}

// q = A p and p.q; last block: alpha = rz / pq for active components.
template <class T>
__global__ void __launch_bounds__(kTpb) pcg_spmv_kernel(int nv, const int* __restrict__ slice_ptr,
                                                        const int* __restrict__ slice_w,
                                                        const uint32_t* __restrict__ ecol,
                                                        const T* __restrict__ eval, const T* __restrict__ diagA,
                                                        const T* __restrict__ p, T* __restrict__ q,
                                                        double* __restrict__ partial, CtlBlock* ctl) {
    if (ctl->done || ctl->outer_done) return;
    const unsigned v = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    double vals[3] = {0, 0, 0};
    if (v < (unsigned)nv) {
        T a0, a1, a2;
        spmv_row(v, slice_ptr, slice_w, ecol, eval, diagA, p, a0, a1, a2);
        st3(q, v, a0, a1, a2);
        T p0, p1, p2;
        ld3(p, v, p0, p1, p2);
        vals[0] = (double)p0 * a0; vals[1] = (double)p1 * a1; vals[2] = (double)p2 * a2;
    }
    if (block_partials<3>(vals, partial, ctl)) {
        double s[3];
        last_block_sum<3>(partial, s);
        if (threadIdx.x == 0) {
            for (int k = 0; k < 3; ++k) ctl->alpha[k] = (ctl->active[k] && s[k] != 0.0) ? ctl->rz[k] / s[k] : 0.0;
            ctl->ticket = 0;
        }
    }
}

// x += alpha p, r -= alpha q, z = D^-1 r, r.z, z.z; last block: beta, stop test.
template <class T>
__global__ void __launch_bounds__(kTpb) pcg_update_kernel(int nv, T* __restrict__ x, T* __restrict__ r,
                                                          const T* __restrict__ p, const T* __restrict__ q,
                                                          const T* __restrict__ dinv, double* __restrict__ partial,
                                                          CtlBlock* ctl) {
    if (ctl->done || ctl->outer_done) return;
    const unsigned v = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const T al0 = (T)ctl->alpha[0], al1 = (T)ctl->alpha[1], al2 = (T)ctl->alpha[2];
    double vals[6] = {0, 0, 0, 0, 0, 0};
    if (v < (unsigned)nv) {
        T x0, x1, x2, r0, r1, r2, p0, p1, p2, q0, q1, q2;
        ld3(x, v, x0, x1, x2);
        ld3(r, v, r0, r1, r2);
        ld3(p, v, p0, p1, p2);
        ld3(q, v, q0, q1, q2);
        x0 += al0 * p0; x1 += al1 * p1; x2 += al2 * p2;
        r0 -= al0 * q0; r1 -= al1 * q1; r2 -= al2 * q2;
        st3(x, v, x0, x1, x2);
        st3(r, v, r0, r1, r2);
        const T di = dinv[v];
        const T z0 = di * r0, z1 = di * r1, z2 = di * r2;
        vals[0] = (double)r0 * z0; vals[1] = (double)r1 * z1; vals[2] = (double)r2 * z2;
        vals[3] = (double)z0 * z0; vals[4] = (double)z1 * z1; vals[5] = (double)z2 * z2;
    }
    if (block_partials<6>(vals, partial, ctl)) {
        double s[6];
        last_block_sum<6>(partial, s);
        if (threadIdx.x == 0) {
            int any = 0;
            for (int k = 0; k < 3; ++k) {
                const bool act = ctl->active[k] != 0;
                ctl->beta[k] = (act && ctl->rz[k] != 0.0) ? s[k] / ctl->rz[k] : 0.0;
                ctl->rz[k] = s[k];
                ctl->active[k] = act && (s[3 + k] > ctl->thr[k]);
                any |= ctl->active[k];
            }
            ctl->iters += 1;
            ctl->total_iters += 1;
            ctl->done = (!any) || (ctl->iters >= ctl->max_iter);
            if (ctl->iters > ctl->max_used) ctl->max_used = ctl->iters;
            if (!ctl->done && ctl->iters >= ctl->launch_cap) ctl->cap_hits += 1;
            ctl->ticket = 0;
        }
    }
}

// p = D^-1 r + beta p.
template <class T>
__global__ void pcg_direction_kernel(int nv, const T* __restrict__ r, const T* __restrict__ dinv, T* __restrict__ p,
                                     CtlBlock* ctl) {
    if (ctl->done || ctl->outer_done) return;
    const unsigned v = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (v >= (unsigned)nv) return;
    const T b0 = (T)ctl->beta[0], b1 = (T)ctl->beta[1], b2 = (T)ctl->beta[2];
    T r0, r1, r2, p0, p1, p2;
    ld3(r, v, r0, r1, r2);
    ld3(p, v, p0, p1, p2);
    const T di = dinv[v];
    st3(p, v, di * r0 + b0 * p0, di * r1 + b1 * p1, di * r2 + b2 * p2);
}

// computeError (pdSolver.cu:243-253) + prev <- sn, and the early-exit flag of
// the local/global loop (pdSolver.cu:164): stop once sqrt(mean (prev-sn)^2) < tol.
template <class T>
__global__ void __launch_bounds__(kTpb) outer_error_kernel(int nv, const T* __restrict__ sn, T* __restrict__ prev,
                                                           double* __restrict__ partial, CtlBlock* ctl) {
    if (ctl->outer_done) return;
    const unsigned v = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    double vals[1] = {0};
    if (v < (unsigned)nv) {
        T s0, s1, s2, p0, p1, p2;
        ld3(sn, v, s0, s1, s2);
        ld3(prev, v, p0, p1, p2);
        const T d0 = p0 - s0, d1 = p1 - s1, d2 = p2 - s2;
        vals[0] = (double)(d0 * d0) + (double)(d1 * d1) + (double)(d2 * d2);
        st3(prev, v, s0, s1, s2);
    }
    if (block_partials<1>(vals, partial, ctl)) {
        double s[1];
        last_block_sum<1>(partial, s);
        if (threadIdx.x == 0) {
            const double err = s[0] / (3.0 * (double)nv);
            ctl->err = err;
            ctl->outer_done = !(sqrt(err) >= ctl->outer_tol);
            ctl->ticket = 0;
        }
    }
}

// b <- c*sn_old (addM_h2Sn) for the direct-solve modes.
template <class T>
__global__ void rhs_reset_kernel(int nv, T* __restrict__ b, const T* __restrict__ c, const T* __restrict__ sn_old,
                                 const int* __restrict__ skip) {
    if (skip && *skip) return;
    const unsigned v = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (v >= (unsigned)nv) return;
    T ox, oy, oz;
    ld3(sn_old, v, ox, oy, oz);
    const T cv = c[v];
    st3(b, v, cv * ox, cv * oy, cv * oz);
}

// b_eff = b with the pinned overwrite applied (for the direct solve).
template <class T>
__global__ void rhs_pinned_kernel(int nv, const T* __restrict__ b, const T* __restrict__ dbc,
                                  const T* __restrict__ more_dbc, const T* __restrict__ dbcx, T wdbc,
                                  T* __restrict__ out, const int* __restrict__ skip) {
    if (skip && *skip) return;
    const unsigned v = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (v >= (unsigned)nv) return;
    T x, y, z;
    rhs_at(b, dbc, more_dbc, dbcx, wdbc, v, x, y, z);
    st3(out, v, x, y, z);
}

template <class T> __global__ void copy_kernel(size_t n, const T* __restrict__ src, T* __restrict__ dst) {
    const size_t i = (size_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

}  // namespace

// ---------------------------------------------------------------------------
template <class T> PdEngine<T>::PdEngine(int threads_per_block, const PdData<T>& data) : tpb_(threads_per_block) {
    SBS_REQUIRE(data.num_verts > 0 && data.num_tets > 0, "PdEngine: empty mesh");
    SBS_REQUIRE(data.X && data.V && data.X0 && data.XTilde && data.DBCX && data.Tet && data.mass && data.mu && data.DBC,
                "PdEngine: a required SolverData pointer is null");
    SBS_REQUIRE(((uintptr_t)data.Tet & 15u) == 0, "PdEngine: Tet must be 16-byte aligned");
    nv_ = data.num_verts;
    nt_ = data.num_tets;
    data_ = data;
    perf_data_ = {{"local step", 0.0}, {"global step", 0.0}, {"collision handling(fixed)", 0.0},
                  {"collision handling(mesh)", 0.0}};
    // tet indices are validated on the host before any kernel dereferences them
    host_tets_.resize(4 * (size_t)nt_);
    SBS_HIP(hipMemcpy(host_tets_.data(), data.Tet, host_tets_.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    for (uint32_t v : host_tets_) SBS_REQUIRE(v < (uint32_t)nv_, "PdEngine: tet index out of range");
    dminv_.alloc(9 * (size_t)nt_);
    w_.alloc(nt_);
    v0_.alloc(nt_);
    // FEMSolver ctor: Dm^-1, V0 from the rest positions (femSolver.cu:7-17).
    hipLaunchKernelGGL(fem_precompute_kernel<T>, dim3(ceil_div(nt_, kTpb)), dim3(kTpb), 0, stream_, nt_, data.X0,
                       (const uint4*)data.Tet, data.mu, dminv_.get(), v0_.get(), w_.get(), data.V0);
    SBS_HIP(hipGetLastError());
    SBS_HIP(hipHostMalloc((void**)&host_counters_, 8 * sizeof(int), hipHostMallocDefault));
    std::memset(host_counters_, 0, 8 * sizeof(int));
    SBS_HIP(hipEventCreateWithFlags(&step_event_, hipEventDisableTiming));
}

template <class T> PdEngine<T>::~PdEngine() {
    if (step_event_) {
        (void)hipEventSynchronize(step_event_);
        (void)hipEventDestroy(step_event_);
    }
    if (host_counters_) (void)hipHostFree(host_counters_);
    for (auto& pool : tev_)
        for (auto& e : pool) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
}

template <class T> void PdEngine<T>::reset() {
    ready_ = false;
    for (auto& e : perf_data_) e.ms = 0.0;
}

// Sliced-ELL structure of the off-diagonal scalar pattern, built on the host
// from the tet list (one-time, like the reference's host-side prepare).
template <class T> void PdEngine<T>::build_structure() {
    SBS_REQUIRE(host_tets_.size() == 4 * (size_t)nt_, "PdEngine: tet list not loaded");
    // edge pairs (both directions), sorted and unique
    std::vector<uint64_t> keys;
    keys.reserve(12 * (size_t)nt_);
    for (size_t t = 0; t < (size_t)nt_; ++t) {
        const uint32_t* tv = &host_tets_[4 * t];
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                if (i != j && tv[i] != tv[j]) keys.push_back(((uint64_t)tv[i] << 32) | tv[j]);
    }
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    nnz_off_ = keys.size();
    std::vector<int> rowlen(nv_, 0);
    for (uint64_t k : keys) rowlen[k >> 32]++;
    std::vector<int> rowstart(nv_ + 1, 0);
    for (int v = 0; v < nv_; ++v) rowstart[v + 1] = rowstart[v] + rowlen[v];
    nslices_ = (nv_ + kSlice - 1) / kSlice;
    std::vector<int> sptr(nslices_ + 1, 0), sw(nslices_, 0);
    for (int s = 0; s < nslices_; ++s) {
        int wmax = 0;
        for (int v = s * kSlice; v < std::min(nv_, (s + 1) * kSlice); ++v) wmax = std::max(wmax, rowlen[v]);
        sw[s] = wmax;
        sptr[s + 1] = sptr[s] + wmax * kSlice;
    }
    std::vector<uint32_t> ecol((size_t)sptr[nslices_]);
    for (int s = 0; s < nslices_; ++s)
        for (int lane = 0; lane < kSlice; ++lane) {
            const int v = s * kSlice + lane;
            for (int e = 0; e < sw[s]; ++e) {
                const size_t pos = (size_t)sptr[s] + (size_t)e * kSlice + lane;
                if (v < nv_ && e < rowlen[v]) ecol[pos] = (uint32_t)(keys[rowstart[v] + e] & 0xffffffffu);
                else ecol[pos] = (uint32_t)std::min(v, nv_ - 1);  // padding: own row, value 0
            }
        }
    slice_ptr_.alloc(nslices_ + 1);
    slice_w_.alloc(nslices_);
    rowlen_.alloc(nv_);
    ecol_.alloc(std::max<size_t>(1, ecol.size()));
    eval_.alloc(std::max<size_t>(1, ecol.size()));
    SBS_HIP(hipMemcpyAsync(slice_ptr_.get(), sptr.data(), sptr.size() * sizeof(int), hipMemcpyHostToDevice, stream_));
    SBS_HIP(hipMemcpyAsync(slice_w_.get(), sw.data(), sw.size() * sizeof(int), hipMemcpyHostToDevice, stream_));
    SBS_HIP(hipMemcpyAsync(rowlen_.get(), rowlen.data(), rowlen.size() * sizeof(int), hipMemcpyHostToDevice, stream_));
    if (!ecol.empty())
        SBS_HIP(hipMemcpyAsync(ecol_.get(), ecol.data(), ecol.size() * sizeof(uint32_t), hipMemcpyHostToDevice, stream_));
    SBS_HIP(hipStreamSynchronize(stream_));
}

template <class T> void PdEngine<T>::assemble(const PdParams& params) {
    const T dt2 = (T)params.dt * (T)params.dt;
    SBS_HIP(hipMemsetAsync(eval_.get(), 0, eval_.bytes(), stream_));
    SBS_HIP(hipMemsetAsync(diagL_.get(), 0, diagL_.bytes(), stream_));
    hipLaunchKernelGGL(assemble_kernel<T>, dim3(ceil_div(nt_, kTpb)), dim3(kTpb), 0, stream_, nt_,
                       (const uint4*)data_.Tet, dminv_.get(), w_.get(), slice_ptr_.get(), rowlen_.get(), ecol_.get(),
                       eval_.get(), diagL_.get());
    SBS_HIP(hipGetLastError());
    hipLaunchKernelGGL(mass_diag_kernel<T>, dim3(ceil_div(nv_, kTpb)), dim3(kTpb), 0, stream_, nv_, data_.mass,
                       data_.DBC, (T)1e6, dt2, diagL_.get(), c_.get(), diagA_.get(), dinv_.get());
    SBS_HIP(hipGetLastError());
}

template <class T> void PdEngine<T>::prepare(PdData<T>& data, const PdParams& params) {
    data_ = data;
    SBS_REQUIRE(params.dt > 0, "PdSolver: dt must be positive");
    const size_t n3 = 3 * (size_t)nv_;
    c_.alloc(nv_); diagL_.alloc(nv_); diagA_.alloc(nv_); dinv_.alloc(nv_);
    sn_.alloc(n3); sn_old_.alloc(n3); prev_.alloc(n3); next_.alloc(n3); b_.alloc(n3);
    r_.alloc(n3); z_.alloc(n3); p_.alloc(n3); q_.alloc(n3);
    red_blocks_ = (int)ceil_div(nv_, kTpb);
    partial_.alloc((size_t)red_blocks_ * 9 + 16);
    ctl_.alloc(sizeof(CtlBlock));
    SBS_HIP(hipMemsetAsync(ctl_.get(), 0, sizeof(CtlBlock), stream_));
    build_structure();
    assemble(params);
    system_built_ = true;
    prepared_dt_ = params.dt;
    // DBCX <- X0 (pdSolver.cu:134)
    SBS_HIP(hipMemcpyAsync(data.DBCX, data.X0, n3 * sizeof(T), hipMemcpyDeviceToDevice, stream_));
    direct_.reset();
    if (solver_ == GlobalSolver::CuSolverCholesky || solver_ == GlobalSolver::EigenCholesky) direct_factor();
    SBS_HIP(hipStreamSynchronize(stream_));
    ready_ = true;
}

template <class T> void PdEngine<T>::direct_factor() {
    SBS_REQUIRE(system_built_, "direct factorisation requested before the system was assembled");
    Export ex = export_system();
    direct_ = std::make_unique<BandCholesky<T>>();
    direct_->factor(nv_, ex.rowptr, ex.col, ex.val, stream_);
}

template <class T> void PdEngine<T>::direct_solve(const T* b, T* x) {
    SBS_REQUIRE(direct_ != nullptr, "direct solve requested before factorisation");
    direct_->solve(b, x, stream_, (const int*)&reinterpret_cast<CtlBlock*>(ctl_.get())->outer_done);
}

template <class T>
void PdEngine<T>::pcg_iterations(T* x, int count) {
    CtlBlock* ctl = reinterpret_cast<CtlBlock*>(ctl_.get());
    const unsigned nb = (unsigned)red_blocks_;
    for (int it = 0; it < count; ++it) {
        tick(kKPcgSpmv, true);
        hipLaunchKernelGGL(pcg_spmv_kernel<T>, dim3(nb), dim3(kTpb), 0, stream_, nv_, slice_ptr_.get(),
                           slice_w_.get(), ecol_.get(), eval_.get(), diagA_.get(), p_.get(), q_.get(), partial_.get(),
                           ctl);
        tick(kKPcgSpmv, false);
        tick(kKPcgUpdate, true);
        hipLaunchKernelGGL(pcg_update_kernel<T>, dim3(nb), dim3(kTpb), 0, stream_, nv_, x, r_.get(), p_.get(),
                           q_.get(), dinv_.get(), partial_.get(), ctl);
        tick(kKPcgUpdate, false);
        tick(kKPcgDirection, true);
        hipLaunchKernelGGL(pcg_direction_kernel<T>, dim3(nb), dim3(kTpb), 0, stream_, nv_, r_.get(), dinv_.get(),
                           p_.get(), ctl);
        tick(kKPcgDirection, false);
    }
    SBS_HIP(hipGetLastError());
}

// Launch policy of the PCG loop.  Kernels of a converged solve return at
// once, but each launch still costs a dispatch, so:
//   exact    -- launch chunks of kPcgChunk iterations and read the device
//               `done` flag between chunks (one host sync per chunk);
//   adaptive -- launch cap_ iterations with no host sync.  The device counts
//               solves that ran out of launched iterations (cap_hits); the
//               host reads the counters of the last completed step and raises
//               cap_ when needed.  cap_ starts from an exact first step.
template <class T>
void PdEngine<T>::pcg_solve(const T* b_src, T* x, bool reset_b, const T* sn_old, bool use_guess, const PdParams& params,
                            bool dbc_override, bool exact, const T* guess_src) {
    CtlBlock* ctl = reinterpret_cast<CtlBlock*>(ctl_.get());
    const unsigned nb = (unsigned)red_blocks_;
    const T wdbc = pin_weight<T>(params.dt);
    if (!use_guess) SBS_HIP(hipMemsetAsync(x, 0, 3 * (size_t)nv_ * sizeof(T), stream_));
    tick(kKPcgInit, true);
    hipLaunchKernelGGL(pcg_init_kernel<T>, dim3(nb), dim3(kTpb), 0, stream_, nv_, const_cast<T*>(b_src), data_.DBC,
                       data_.moreDBC, data_.DBCX, wdbc, (int)dbc_override, guess_src ? guess_src : (const T*)x, x,
                       slice_ptr_.get(), slice_w_.get(),
                       ecol_.get(), eval_.get(), diagA_.get(), dinv_.get(), r_.get(), p_.get(), (int)reset_b, c_.get(),
                       sn_old, partial_.get(), ctl);
    tick(kKPcgInit, false);
    SBS_HIP(hipGetLastError());
    if (!exact) {
        pcg_iterations(x, std::min(cap_, params.pcg_max_iter));
        return;
    }
    int launched = 0;
    while (launched < params.pcg_max_iter) {
        const int n = std::min(kPcgChunk, params.pcg_max_iter - launched);
        pcg_iterations(x, n);
        launched += n;
        SBS_HIP(hipMemcpyAsync(host_counters_ + 4, &ctl->done, sizeof(int), hipMemcpyDeviceToHost, stream_));
        SBS_HIP(hipStreamSynchronize(stream_));
        if (host_counters_[4]) break;
    }
}

template <class T> void PdEngine<T>::review_counters() {
    if (!step_pending_) return;
    if (hipEventQuery(step_event_) != hipSuccess) return;  // previous step still running: keep the current cap
    step_pending_ = false;
    const int cap_hits = host_counters_[1], max_used = host_counters_[2];
    last_pcg_iters_ = host_counters_[0];
    cap_hits_total_ += cap_hits;
    if (max_used + kPcgMargin > cap_) cap_ = max_used + kPcgMargin;
    if (cap_hits > 0) cap_ += std::max(4, cap_ / 2);
}

template <class T> void PdEngine<T>::update(PdData<T>& data, const PdParams& params) {
    if (!ready_) prepare(data, params);
    data_ = data;
    SBS_REQUIRE(params.num_iterations >= 0, "PdSolver: numIterations must be >= 0");
    SBS_REQUIRE(params.pcg_max_iter >= 1, "PdSolver: pcg_max_iter must be >= 1");
    const bool direct = solver_ == GlobalSolver::CuSolverCholesky || solver_ == GlobalSolver::EigenCholesky;
    if (direct && !direct_) direct_factor();
    const T dt = (T)params.dt;
    const unsigned vb = ceil_div(nv_, kTpb), tb = ceil_div(nt_, kTpb);
    const T wdbc = pin_weight<T>(params.dt);
    CtlBlock* ctl = reinterpret_cast<CtlBlock*>(ctl_.get());
    const bool jacobi = solver_ == GlobalSolver::Jacobi;
    const bool pcg = solver_ == GlobalSolver::PCG;
    const bool cheb = jacobi || (pcg && params.pcg_chebyshev);
    review_counters();
    const bool exact = pcg && (!params.pcg_adaptive || cap_ <= 0);
    if (pcg && exact && step_pending_) {  // counters of an in-flight step are needed before an exact step
        SBS_HIP(hipEventSynchronize(step_event_));
        review_counters();
    }
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    if (perf_) for (auto& e : ev) SBS_HIP(hipEventCreate(&e));

    hipLaunchKernelGGL(ctl_init_kernel, dim3(1), dim3(64), 0, stream_, ctl, params.pcg_max_iter,
                       exact ? params.pcg_max_iter : std::min(cap_, params.pcg_max_iter),
                       params.pcg_tol * params.pcg_tol, params.pcg_floor * params.pcg_floor, params.tol);
    hipLaunchKernelGGL(begin_step_kernel<T>, dim3(vb), dim3(kTpb), 0, stream_, nv_, data.X, data.V, data.mass,
                       data.DBC, data.moreDBC, dt, (T)params.gravity, c_.get(), sn_.get(), sn_old_.get(), prev_.get(),
                       b_.get(), (int)cheb);
    SBS_HIP(hipGetLastError());
    T omega = 1;
    const T rho = (T)params.rho;
    const int* skip = cheb ? nullptr : &ctl->outer_done;
    for (int i = 0; i < params.num_iterations; ++i) {
        if (perf_) SBS_HIP(hipEventRecord(ev[0], stream_));
        tick(kKLocal, true);
        hipLaunchKernelGGL(local_step_kernel<T>, dim3(tb), dim3(kTpb), 0, stream_, nt_, (const uint4*)data.Tet,
                           dminv_.get(), w_.get(), sn_.get(), b_.get(), (int)jacobi, skip);
        tick(kKLocal, false);
        if (perf_) SBS_HIP(hipEventRecord(ev[1], stream_));
        if (cheb) {
            if (i <= 10) omega = 1;
            else if (i == 11) omega = (T)2 / ((T)2 - rho * rho);
            else omega = (T)4 / ((T)4 - rho * rho * omega);
        }
        if (jacobi) {
            tick(kKJacobi, true);
            hipLaunchKernelGGL(jacobi_chebyshev_kernel<T>, dim3(vb), dim3(kTpb), 0, stream_, nv_, b_.get(), c_.get(),
                               diagL_.get(), data.DBC, data.moreDBC, data.DBCX, wdbc, sn_.get(), prev_.get(),
                               sn_old_.get(), omega);
            tick(kKJacobi, false);
        } else if (pcg) {
            T* x = cheb ? next_.get() : sn_.get();
            pcg_solve(b_.get(), x, /*reset_b=*/true, sn_old_.get(), /*use_guess=*/true, params, /*dbc=*/true, exact,
                      cheb ? sn_.get() : nullptr);
            if (cheb) {
                tick(kKChebyshev, true);
                hipLaunchKernelGGL(chebyshev_kernel<T>, dim3(ceil_div(3 * (size_t)nv_, kTpb)), dim3(kTpb), 0, stream_,
                                   nv_, (const T*)next_.get(), sn_.get(), prev_.get(), omega, (const int*)nullptr);
                tick(kKChebyshev, false);
            } else
                hipLaunchKernelGGL(outer_error_kernel<T>, dim3(vb), dim3(kTpb), 0, stream_, nv_, (const T*)sn_.get(),
                                   prev_.get(), partial_.get(), ctl);
        } else {
            hipLaunchKernelGGL(rhs_pinned_kernel<T>, dim3(vb), dim3(kTpb), 0, stream_, nv_, (const T*)b_.get(),
                               data.DBC, data.moreDBC, data.DBCX, wdbc, next_.get(), skip);
            tick(kKDirect, true);
            direct_solve(next_.get(), sn_.get());
            tick(kKDirect, false);
            hipLaunchKernelGGL(outer_error_kernel<T>, dim3(vb), dim3(kTpb), 0, stream_, nv_, (const T*)sn_.get(),
                               prev_.get(), partial_.get(), ctl);
            hipLaunchKernelGGL(rhs_reset_kernel<T>, dim3(vb), dim3(kTpb), 0, stream_, nv_, b_.get(), c_.get(),
                               (const T*)sn_old_.get(), skip);
        }
        SBS_HIP(hipGetLastError());
        if (perf_) {
            SBS_HIP(hipEventRecord(ev[2], stream_));
            SBS_HIP(hipEventSynchronize(ev[2]));
            float a = 0, g = 0;
            SBS_HIP(hipEventElapsedTime(&a, ev[0], ev[1]));
            SBS_HIP(hipEventElapsedTime(&g, ev[1], ev[2]));
            perf_data_[0].ms += a;
            perf_data_[1].ms += g;
        }
    }
    hipLaunchKernelGGL(end_step_kernel<T>, dim3(vb), dim3(kTpb), 0, stream_, nv_, (const T*)sn_.get(), (T)1 / dt,
                       data.XTilde, data.V, data.X, data.moreDBC, 1);
    SBS_HIP(hipGetLastError());
    // counters: total PCG iterations, cap hits, max iterations of one solve, solves
    SBS_HIP(hipMemcpyAsync(host_counters_, &ctl->total_iters, sizeof(int), hipMemcpyDeviceToHost, stream_));
    SBS_HIP(hipMemcpyAsync(host_counters_ + 1, &ctl->cap_hits, 3 * sizeof(int), hipMemcpyDeviceToHost, stream_));
    SBS_HIP(hipEventRecord(step_event_, stream_));
    step_pending_ = true;
    if (perf_) {
        SBS_HIP(hipStreamSynchronize(stream_));
        for (auto& e : ev) (void)hipEventDestroy(e);
    }
    if (exact && pcg) {
        SBS_HIP(hipEventSynchronize(step_event_));
        review_counters();
    }
    last_iters_ = params.num_iterations;
}

template <class T> void PdEngine<T>::local_only(const T* q, const T* sn_old, bool residual_form, double dt, T* b_out) {
    SBS_REQUIRE(ready_, "local_only: engine not prepared");
    const unsigned vb = ceil_div(nv_, kTpb), tb = ceil_div(nt_, kTpb);
    hipLaunchKernelGGL(rhs_reset_kernel<T>, dim3(vb), dim3(kTpb), 0, stream_, nv_, b_.get(), c_.get(), sn_old,
                       (const int*)nullptr);
    hipLaunchKernelGGL(local_step_kernel<T>, dim3(tb), dim3(kTpb), 0, stream_, nt_, (const uint4*)data_.Tet,
                       dminv_.get(), w_.get(), q, b_.get(), (int)residual_form, (const int*)nullptr);
    const T wdbc = pin_weight<T>(dt);
    hipLaunchKernelGGL(rhs_pinned_kernel<T>, dim3(vb), dim3(kTpb), 0, stream_, nv_, (const T*)b_.get(), data_.DBC,
                       data_.moreDBC, data_.DBCX, wdbc, b_out, (const int*)nullptr);
    SBS_HIP(hipGetLastError());
}

template <class T> int PdEngine<T>::solve(const T* b, T* x, const PdParams& params, bool use_guess) {
    SBS_REQUIRE(ready_, "solve: engine not prepared");
    CtlBlock* ctl = reinterpret_cast<CtlBlock*>(ctl_.get());
    if (step_pending_) {
        SBS_HIP(hipEventSynchronize(step_event_));
        review_counters();
    }
    hipLaunchKernelGGL(ctl_init_kernel, dim3(1), dim3(64), 0, stream_, ctl, params.pcg_max_iter, params.pcg_max_iter,
                       params.pcg_tol * params.pcg_tol, params.pcg_floor * params.pcg_floor, 0.0);
    if (direct_) {
        direct_solve(b, x);
        SBS_HIP(hipStreamSynchronize(stream_));
        return 0;
    }
    pcg_solve(b, x, false, nullptr, use_guess, params, false, /*exact=*/true, nullptr);
    SBS_HIP(hipMemcpyAsync(host_counters_ + 4, &ctl->iters, sizeof(int), hipMemcpyDeviceToHost, stream_));
    SBS_HIP(hipStreamSynchronize(stream_));
    return host_counters_[4];
}

template <class T> typename PdEngine<T>::Export PdEngine<T>::export_system() const {
    Export ex;
    auto d2h = [&](std::vector<T>& dst, const DevBuf<T>& src, size_t n) {
        dst.resize(n);
        if (n) SBS_HIP(hipMemcpy(dst.data(), src.get(), n * sizeof(T), hipMemcpyDeviceToHost));
    };
    SBS_HIP(hipStreamSynchronize(stream_));
    d2h(ex.dminv, dminv_, 9 * (size_t)nt_);
    d2h(ex.V0, v0_, nt_);
    d2h(ex.w, w_, nt_);
    if (!system_built_) return ex;
    d2h(ex.c, c_, nv_);
    d2h(ex.diagA, diagA_, nv_);
    d2h(ex.diagL, diagL_, nv_);
    std::vector<int> sptr(nslices_ + 1), sw(nslices_), rl(nv_);
    SBS_HIP(hipMemcpy(sptr.data(), slice_ptr_.get(), sptr.size() * sizeof(int), hipMemcpyDeviceToHost));
    SBS_HIP(hipMemcpy(sw.data(), slice_w_.get(), sw.size() * sizeof(int), hipMemcpyDeviceToHost));
    SBS_HIP(hipMemcpy(rl.data(), rowlen_.get(), rl.size() * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<uint32_t> ec(ecol_.n);
    std::vector<T> ev(eval_.n);
    SBS_HIP(hipMemcpy(ec.data(), ecol_.get(), ec.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    SBS_HIP(hipMemcpy(ev.data(), eval_.get(), ev.size() * sizeof(T), hipMemcpyDeviceToHost));
    // full CSR of A_s with sorted columns (diagonal included)
    ex.rowptr.assign(nv_ + 1, 0);
    for (int v = 0; v < nv_; ++v) ex.rowptr[v + 1] = ex.rowptr[v] + rl[v] + 1;
    ex.col.resize(ex.rowptr[nv_]);
    ex.val.resize(ex.rowptr[nv_]);
    for (int v = 0; v < nv_; ++v) {
        std::vector<std::pair<int, T>> row;
        row.emplace_back(v, ex.diagA[v]);
        const int s = v / kSlice, lane = v % kSlice;
        for (int e = 0; e < rl[v]; ++e) {
            const size_t pos = (size_t)sptr[s] + (size_t)e * kSlice + lane;
            row.emplace_back((int)ec[pos], ev[pos]);
        }
        std::sort(row.begin(), row.end(), [](auto& a, auto& b) { return a.first < b.first; });
        for (size_t k = 0; k < row.size(); ++k) {
            ex.col[ex.rowptr[v] + k] = row[k].first;
            ex.val[ex.rowptr[v] + k] = row[k].second;
        }
    }
    return ex;
}

const char* kernel_name(int id) {
    static const char* names[kKCount] = {"local_step", "pcg_init", "pcg_spmv", "pcg_update",
                                         "pcg_direction", "jacobi_chebyshev", "chebyshev", "direct_solve"};
    return (id >= 0 && id < kKCount) ? names[id] : "?";
}

template <class T> void PdEngine<T>::set_kernel_timing(unsigned mask) {
    flush_timing();
    for (int k = 0; k < kKCount; ++k) ktime_[k] = KernelTime{};
    timing_mask_ = mask;
}

// Records the start (or stop) event of kernel `id` when it is being timed.
template <class T> void PdEngine<T>::tick(int id, bool start) {
    if (!(timing_mask_ & (1u << id))) return;
    auto& pool = tev_[id];
    size_t& used = tev_used_[id];
    if (start) {
        if (used == pool.size()) {
            hipEvent_t a, b;
            SBS_HIP(hipEventCreate(&a));
            SBS_HIP(hipEventCreate(&b));
            pool.emplace_back(a, b);
        }
        SBS_HIP(hipEventRecord(pool[used].first, stream_));
    } else {
        SBS_HIP(hipEventRecord(pool[used].second, stream_));
        ++used;
        if (used >= 4096) flush_timing();  // bound the pool
    }
}

template <class T> void PdEngine<T>::flush_timing() {
    bool any = false;
    for (int k = 0; k < kKCount; ++k) any |= tev_used_[k] > 0;
    if (!any) return;
    SBS_HIP(hipStreamSynchronize(stream_));
    for (int k = 0; k < kKCount; ++k) {
        for (size_t i = 0; i < tev_used_[k]; ++i) {
            float ms = 0.f;
            SBS_HIP(hipEventElapsedTime(&ms, tev_[k][i].first, tev_[k][i].second));
            ktime_[k].ms += ms;
            ktime_[k].launches += 1;
        }
        tev_used_[k] = 0;
    }
}

template <class T> KernelTime PdEngine<T>::kernel_time(int id) {
    SBS_REQUIRE(id >= 0 && id < kKCount, "kernel id out of range");
    flush_timing();
    return ktime_[id];
}

template class PdEngine<float>;
template class PdEngine<double>;

}  // namespace sbs
