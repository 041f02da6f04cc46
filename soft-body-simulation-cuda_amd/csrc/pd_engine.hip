// gfx950 kernels and host orchestration of the projective-dynamics step.
// See pd_engine.hpp for the layout; reference citations are per kernel.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <type_traits>

#include "collision.hpp"
#include "dense_refine.hpp"
#include "direct_solver.hpp"
#include "halo_tiles.hpp"
#include "pd_engine.hpp"
#include "batch_engine.hpp"
#include "svd3.hpp"

namespace sbs {

namespace {

#ifndef SBS_TPB
#define SBS_TPB 256
#endif
constexpr int kTpb = SBS_TPB;  // threads per block of the streaming kernels (and rows per halo tile)
constexpr int kSlice = 64;     // sliced-ELL slice height = one wavefront
constexpr int kStripes = 32;   // address stripes of each reduction slot
constexpr int kSlotQ = 16;     // doubles per stripe
constexpr int kSlotSize = kStripes * kSlotQ;
// reduction slots (each kSlotSize doubles)
constexpr int kSlotInit = 0;   // (D^-1 b)^2 per component
constexpr int kSlotIter = 1;   // 3 rotating slots: r.u, w.u, u.u per component
constexpr int kSlotErr = 4;    // sum (prev - x)^2 of computeError
constexpr int kNumSlots = 5;
constexpr int kCheckChunk = 8;  // iterations between host checks of an unlearned solve

// Pin weight of computeDBCLocal: positional_weight * dt2Inv with
// dt2Inv = dtInv * dtInv in the solver precision (pdSolver.cu:145-147, :171;
// positional_weight = 1e6, pdSolver.h:31).
template <class T> T pin_weight(double dt) {
    const T dt_inv = (T)1 / (T)dt;
    const T dt2_inv = dt_inv * dt_inv;
    return (T)1e6 * dt2_inv;
}

// The Chebyshev step of the global solve (chebyshevKern, pdUtil.cu:216-226):
// the damping 0.9 (x - s) + s uses the double literal 0.9, so it is evaluated
// in double and rounded once to the solver precision; then
// q = (damped - prev) omega + prev in the solver precision.  Every kernel that
// applies the recurrence uses this one sequence.
template <class T> __device__ __forceinline__ T cheb_update(T x, T s, T prev, T omega) {
    const T damped = (T)fma(0.9, (double)(x - s), (double)s);
    return fma(damped - prev, omega, prev);
}

template <class T> __device__ __forceinline__ void ld3(const T* __restrict__ p, size_t i, T& x, T& y, T& z) {
    x = p[3 * i];
    y = p[3 * i + 1];
    z = p[3 * i + 2];
}
template <class T> __device__ __forceinline__ void st3(T* __restrict__ p, size_t i, T x, T y, T z) {
    p[3 * i] = x;
    p[3 * i + 1] = y;
    p[3 * i + 2] = z;
}

// The PCG vectors a row's neighbours read (r, w, s and D^-1) live in one
// 12-scalar record per row: [r xyz, D^-1 | w xyz, 0 | s xyz, 0], so a
// neighbour costs three aligned 16-byte (float) loads from one or two cache
// lines instead of four gathers from four arrays, and r with D^-1 (all that
// u = D^-1 r needs) is the first 16 bytes.
constexpr int kRec = 12;
constexpr int kRecR = 0, kRecD = 3, kRecW = 4, kRecS = 8;
template <class T> struct Rec {
    T v[kRec];
};
template <class T> __device__ __forceinline__ Rec<T> ld_rec(const T* __restrict__ base, size_t i) {
    Rec<T> r;
    using V = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
    constexpr int per = 16 / sizeof(T);
    const V* p = reinterpret_cast<const V*>(base + kRec * i);
#pragma unroll
    for (int k = 0; k < kRec / per; ++k) {
        const V q = p[k];
        if constexpr (per == 4) {
            r.v[4 * k] = q.x; r.v[4 * k + 1] = q.y; r.v[4 * k + 2] = q.z; r.v[4 * k + 3] = q.w;
        } else {
            r.v[2 * k] = q.x; r.v[2 * k + 1] = q.y;
        }
    }
    return r;
}
// r and D^-1 of a record (its first 4 scalars).
template <class T>
__device__ __forceinline__ void ld_rec_rd(const T* __restrict__ base, size_t i, T& r0, T& r1, T& r2, T& d) {
    if constexpr (sizeof(T) == 4) {
        const float4 q = *reinterpret_cast<const float4*>(base + kRec * i);
        r0 = q.x; r1 = q.y; r2 = q.z; d = q.w;
    } else {
        const double2 q0 = *reinterpret_cast<const double2*>(base + kRec * i);
        const double2 q1 = *reinterpret_cast<const double2*>(base + kRec * i + 2);
        r0 = q0.x; r1 = q0.y; r2 = q1.x; d = q1.y;
    }
}
// w of a record (scalars 4..7, the pad written 0).
template <class T> __device__ __forceinline__ void st_rec_w(T* __restrict__ base, size_t i, T w0, T w1, T w2) {
    if constexpr (sizeof(T) == 4) {
        *reinterpret_cast<float4*>(base + kRec * i + kRecW) = make_float4(w0, w1, w2, 0.f);
    } else {
        *reinterpret_cast<double2*>(base + kRec * i + kRecW) = make_double2(w0, w1);
        *reinterpret_cast<double2*>(base + kRec * i + kRecW + 2) = make_double2(w2, 0.0);
    }
}
template <class T> __device__ __forceinline__ void st_rec(T* __restrict__ base, size_t i, const Rec<T>& r) {
    using V = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
    constexpr int per = 16 / sizeof(T);
    V* p = reinterpret_cast<V*>(base + kRec * i);
#pragma unroll
    for (int k = 0; k < kRec / per; ++k) {
        V q;
        if constexpr (per == 4) {
            q.x = r.v[4 * k]; q.y = r.v[4 * k + 1]; q.z = r.v[4 * k + 2]; q.w = r.v[4 * k + 3];
        } else {
            q.x = r.v[2 * k]; q.y = r.v[2 * k + 1];
        }
        p[k] = q;
    }
}

// Sums NQ per-thread values over the block (kTpb threads); thread k < NQ adds
// quantity k to the block's stripe of `slot` (stripe = block id mod kStripes),
// so no address sees more than gridDim/kStripes atomics.  Each row of 16
// lanes reduces with DPP moves and leaves its total in LDS; the 16 row totals
// are added in row order.  Every thread of the block calls it.
template <int NQ>
__device__ __forceinline__ void stripe_add(const double (&vals)[NQ], double* slot) {
    constexpr int kRows = kTpb / 16;
    __shared__ double red[NQ][kRows];
    const int lane = threadIdx.x & 63, rid = threadIdx.x >> 4;
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        const double s = row16_total(vals[k]);
        if ((lane & 15) == 0) red[k][rid] = s;
    }
    __syncthreads();
    if (threadIdx.x < NQ) {
        double s = 0;
#pragma unroll
        for (int r = 0; r < kRows; ++r) s += red[threadIdx.x][r];
        atomicAdd(&slot[(blockIdx.x % kStripes) * kSlotQ + threadIdx.x], s);
    }
}

// Totals of the first NQ quantities of a slot (written by an earlier kernel),
// computed by every wavefront on its own (no LDS, no barrier): lane 16 r + s
// adds stripes s and s + 16 of quantity 4 p + r, and one 16-lane reduction
// per row finishes four quantities at once.  Every lane must be active.
template <int NQ>
__device__ __forceinline__ void slot_totals(const double* slot, double (&out)[NQ]) {
    static_assert(kStripes == 32, "two stripes per lane, 16 lanes per quantity");
    constexpr int kPasses = (NQ + 3) / 4;
    const int lane = threadIdx.x & 63, r = lane >> 4, s = lane & 15;
    double v[kPasses];
#pragma unroll
    for (int p = 0; p < kPasses; ++p) {
        const int k = 4 * p + r;
        v[p] = k < NQ ? slot[s * kSlotQ + k] + slot[(s + 16) * kSlotQ + k] : 0.0;
    }
#pragma unroll
    for (int p = 0; p < kPasses; ++p) {
        const double t = row16_total(v[p]);
        if (4 * p + 0 < NQ) out[4 * p + 0] = readlane_f64<0>(t);
        if (4 * p + 1 < NQ) out[4 * p + 1] = readlane_f64<16>(t);
        if (4 * p + 2 < NQ) out[4 * p + 2] = readlane_f64<32>(t);
        if (4 * p + 3 < NQ) out[4 * p + 3] = readlane_f64<48>(t);
    }
}

// ---------------------------------------------------------------------------
// Device control block: per-step counters, the outer-loop early exit and
// the PCG state (double-buffered by iteration parity: launch i of a solve
// writes st[i&1] from st[(i-1)&1] and the reduction of launch i-1).
struct PcgState {
    double gamma[3];  // r.u of the current iterate
    double alpha[3];
    double beta[3];
    double thr[3];
    int active[3];
    int iters;       // completed iterations
    int done;        // converged (all components) or out of iterations
    int unconverged; // stopped by max_iter with a component still active
    int pad;
};

// The first kCounterInts ints are copied to the host at the end of a step.
struct CtlBlock {
    int pcg_iters;       // PCG iterations summed over the step
    int unconverged;     // PCG solves stopped before the tolerance (max_iter or launch cap)
    int max_used;        // most iterations of one solve this step
    int solves;          // PCG solves this step
    int last_iters;      // iterations of the last solve
    int pad0;
    int log[kMaxSolveLog];  // iterations of solve i of the step (-its-1 when stopped early)
    int outer_done;      // local/global loop early exit (direct / PCG modes)
    int pad1;
    double err;          // last computeError value
    PcgState st[2];
};
static_assert(offsetof(CtlBlock, log) + sizeof(int) * kMaxSolveLog == sizeof(int) * kCounterInts,
              "counter block layout");

__global__ void ctl_init_kernel(CtlBlock* ctl) {
    if (blockIdx.x != 0) return;
    for (int i = threadIdx.x; i < kCounterInts; i += blockDim.x) (&ctl->pcg_iters)[i] = 0;
    if (threadIdx.x == 0) {
        ctl->outer_done = 0;
        ctl->err = 1.0;
    }
}

__global__ void zero_kernel(double* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0.0;
}

// ---------------------------------------------------------------------------
// FEM precompute: computeInvDmV0 (src/simulation/solver/solverUtil.cuh:109-137).
// Dm = [x1-x0, x2-x0, x3-x0] (columns), Dm^-1 stored row-major in 9 planes
// at the tet's internal position, V0 = |det Dm| / 6, w = |V0| * mu (the
// coefficient of pdUtil.cu:19 and :126).
template <class T>
__global__ void __launch_bounds__(kTpb) fem_precompute_kernel(int nt, const T* __restrict__ X,
                                                              const uint4* __restrict__ tets,
                                                              const int* __restrict__ perm,
                                                              const T* __restrict__ mu, T* __restrict__ dminv,
                                                              T* __restrict__ v0, T* __restrict__ w,
                                                              T* __restrict__ v0_out) {
    const unsigned i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= (unsigned)nt) return;
    const int t = perm[i];
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
    for (int k = 0; k < 9; ++k) dminv[(size_t)k * nt + i] = inv[k];
    const T vol = fabs(d) / (T)6;
    v0[i] = vol;
    w[i] = fabs(vol) * mu[t];
    if (v0_out) v0_out[t] = vol;
}

template <class T> __device__ __forceinline__ void load_dminv(const T* __restrict__ dminv, int nt, unsigned t, T (&m)[9]) {
#pragma unroll
    for (int k = 0; k < 9; ++k) m[k] = dminv[(size_t)k * nt + t];
}

// ---------------------------------------------------------------------------
// Assembly of L_s = sum_t w_t G Dm^-1 Dm^-T G^T (computeSiTSi, pdUtil.cu:9-40),
// G rows (-1,-1,-1),(1,0,0),(0,1,0),(0,0,1) (pdUtil.cu:16): diagonal into
// diagL (the reference's matrix_diag), off-diagonal entries into the
// sliced-ELL values found by a search of the (short) row.  One-time prepare:
// one thread per row sums its tets' contributions in ascending internal tet
// order (inc: the row's (tet << 2 | corner) list), so the matrix is the same
// in every run (float atomics would make it depend on their order).
template <class T>
__global__ void __launch_bounds__(kTpb) assemble_rows_kernel(int nv, int nt, const int* __restrict__ inc_ptr,
                                                             const uint32_t* __restrict__ inc, const uint4* __restrict__ tets,
                                                             const int* __restrict__ perm, const T* __restrict__ dminv,
                                                             const T* __restrict__ w, const int* __restrict__ slice_ptr,
                                                             const int* __restrict__ rowlen,
                                                             const uint32_t* __restrict__ ecol, T* __restrict__ eval,
                                                             T* __restrict__ diagL) {
    const unsigned row = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (row >= (unsigned)nv) return;
    const size_t base = (size_t)slice_ptr[row / kSlice] + (row % kSlice);
    const int len = rowlen[row];
    T diag = 0;
    for (int k = inc_ptr[row]; k < inc_ptr[row + 1]; ++k) {
        const uint32_t code = inc[k];
        const unsigned i = code >> 2, a = code & 3u;
        const uint4 tv = tets[perm[i]];
        const unsigned v[4] = {tv.x, tv.y, tv.z, tv.w};
        T m[9];
        load_dminv(dminv, nt, i, m);
        T g[4][3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            g[0][c] = -(m[c] + m[3 + c] + m[6 + c]);
            g[1][c] = m[c];
            g[2][c] = m[3 + c];
            g[3][c] = m[6 + c];
        }
        const T ww = w[i];
        for (int bb = 0; bb < 4; ++bb) {
            const T kab = ww * (g[a][0] * g[bb][0] + g[a][1] * g[bb][1] + g[a][2] * g[bb][2]);
            if ((unsigned)bb == a) {
                diag += kab;
            } else {
                for (int e = 0; e < len; ++e) {
                    const size_t pos = base + (size_t)e * kSlice;
                    if (ecol[pos] == v[bb]) {
                        eval[pos] += kab;
                        break;
                    }
                }
            }
        }
    }
    diagL[row] = diag;
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
// (pdUtil.cu:56-71), computeSn (pdUtil.cu:73-95), the sn_old copy and the prev
// initialisation (pdSolver.cu:158-162; prev = 0 outside Jacobi mode, the
// reference memsets bytes there).  A vertex with a moreDBC weight (a dragged
// vertex) is predicted at its target, which the caller keeps in DBCX (computeSn
// writes the target into DBCX and sn).
template <class T>
__global__ void begin_step_kernel(int nv, const T* __restrict__ X, const T* __restrict__ V,
                                  const T* __restrict__ mass, const T* __restrict__ dbc,
                                  const T* __restrict__ more_dbc, const T* __restrict__ dbcx, T dt, T gravity,
                                  T* __restrict__ c,
                                  T* __restrict__ sn, T* __restrict__ sn_old, T* __restrict__ prev,
                                  int prev_from_sn, T* __restrict__ sn2, T* __restrict__ prev2) {
    const unsigned v = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (v >= (unsigned)nv) return;
    const T m = mass[v];
    const T dt2 = dt * dt;
    T cv = c[v];
    const T wi = more_dbc ? more_dbc[v] : (T)0;
    if (dbc[v] == (T)0) {
        cv = wi > (T)0 ? (m + wi) / dt2 : m / dt2;
        c[v] = cv;
    }
    T sx, sy, sz;
    if (wi > (T)0) {
        ld3(dbcx, v, sx, sy, sz);
    } else {
        T x, y, z, vx, vy, vz;
        ld3(X, v, x, y, z);
        ld3(V, v, vx, vy, vz);
        const T inv = (T)1 / cv;
        const T fy = -gravity * m;
        sx = x + dt * vx; sy = y + dt * vy + inv * fy; sz = z + dt * vz;
    }
    st3(sn, v, sx, sy, sz);
    st3(sn_old, v, sx, sy, sz);
    if (prev_from_sn) st3(prev, v, sx, sy, sz);
    else st3(prev, v, (T)0, (T)0, (T)0);
    // the second sn/prev pair of the folded Chebyshev finish: the fold writes
    // only vertices that some tet touches, and a vertex in no tet keeps
    // sn = prev = sn_old through every iteration (its row is c x = c sn_old)
    if (sn2) {
        st3(sn2, v, sx, sy, sz);
        st3(prev2, v, sx, sy, sz);
    }
}

// Right-hand side of one vertex: addM_h2Sn (pdUtil.cu:168-178) plus the
// local-step partials of every chunk touching it, with the pinned overwrite
// of computeDBCLocal (pdUtil.cu:147-166).
constexpr int kGatherBatch = 4;  // vertex->entry list entries loaded together

template <class T> struct RhsGather {
    const T* c;
    const T* sn_old;
    const int* vent_ptr;      // partials of vertex v: [vent_ptr[v], vent_ptr[v+1])
    const T* partial;
    const T* dbc;
    const T* more_dbc;
    const T* dbcx;
    T wdbc;

    // A vertex's partials are contiguous (the local step writes them in vertex
    // order); they are loaded kGatherBatch at a time, clamped in bounds.
    __device__ __forceinline__ void at(unsigned v, T& bx, T& by, T& bz) const {
        at_range(v, vent_ptr[v], vent_ptr[v + 1], bx, by, bz);
    }
    // The same with the vertex's partial range [e0, e1) given (a caller that
    // holds it per internal row need not wait for v to request it).
    __device__ __forceinline__ void at_range(unsigned v, int e0, int e1, T& bx, T& by, T& bz) const {
        const T pin = dbc[v];
        const T more = more_dbc ? more_dbc[v] : (T)0;
        T sx, sy, sz;
        ld3(sn_old, v, sx, sy, sz);
        const T cv = c[v];
        bx = cv * sx; by = cv * sy; bz = cv * sz;
        for (int e = e0; e < e1; e += kGatherBatch) {
#pragma unroll
            for (int j = 0; j < kGatherBatch; ++j) {
                T px, py, pz;
                ld3(partial, min(e + j, e1 - 1), px, py, pz);
                if (e + j < e1) { bx += px; by += py; bz += pz; }
            }
        }
        if (pin > (T)0 || more > (T)0) {
            const T wgt = pin > (T)0 ? wdbc : more;
            ld3(dbcx, v, bx, by, bz);
            bx *= wgt; by *= wgt; bz *= wgt;
        }
    }
};

// b_eff = gathered right-hand side, written out (direct solve input, tests).
template <class T>
__global__ void rhs_kernel(int nv, RhsGather<T> g, T* __restrict__ out, const int* __restrict__ skip) {
    if (skip && *skip) return;
    const unsigned v = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (v >= (unsigned)nv) return;
    T x, y, z;
    g.at(v, x, y, z);
    st3(out, v, x, y, z);
}

// Jacobi + Chebyshev global step (getErrorKern pdUtil.cu:195-214 and
// chebyshevKern pdUtil.cu:216-226 fused per vertex).
template <class T>
__global__ void jacobi_chebyshev_kernel(int nv, RhsGather<T> g, const T* __restrict__ diagL,
                                        T* __restrict__ sn, T* __restrict__ prev, T omega) {
    const unsigned v = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (v >= (unsigned)nv) return;
    T bx, by, bz;
    g.at(v, bx, by, bz);
    T sx, sy, sz, px, py, pz;
    ld3(sn, v, sx, sy, sz);
    ld3(prev, v, px, py, pz);
    const T cv = g.c[v];
    T nx, ny, nz;
    if (g.more_dbc && g.more_dbc[v] > (T)0) {
        nx = sx; ny = sy; nz = sz;
    } else {
        const T den = cv + diagL[v];
        nx = (bx - cv * sx) / den + sx;
        ny = (by - cv * sy) / den + sy;
        nz = (bz - cv * sz) / den + sz;
    }
    nx = cheb_update(nx, sx, px, omega);
    ny = cheb_update(ny, sy, py, omega);
    nz = cheb_update(nz, sz, pz, omega);
    st3(prev, v, sx, sy, sz);
    st3(sn, v, nx, ny, nz);
}

// updateVelPos (pdUtil.cu:180-193) and X <- XTilde (pdSolver.cu:227).
template <class T>
__global__ void end_step_kernel(int nv, const T* __restrict__ sn, T dt_inv, T* __restrict__ XTilde,
                                T* __restrict__ V, T* __restrict__ X, const T* __restrict__ more_dbc) {
    const unsigned v = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (v >= (unsigned)nv) return;
    T nx, ny, nz, ox, oy, oz;
    ld3(sn, v, nx, ny, nz);
    ld3(XTilde, v, ox, oy, oz);
    if (more_dbc && more_dbc[v] > (T)0) st3(V, v, (T)0, (T)0, (T)0);
    else st3(V, v, (nx - ox) * dt_inv, (ny - oy) * dt_inv, (nz - oz) * dt_inv);
    st3(XTilde, v, nx, ny, nz);
    if (X) st3(X, v, nx, ny, nz);  // X <- XTilde (pdSolver.cu:227) unless collisions decide it
}

// computeError (pdSolver.cu:243-253): sum (prev - src)^2 into the error slot,
// then prev <- src (and sn <- src when sn is given).  outer_decide_kernel
// turns the sum into the early exit of the local/global loop.
template <class T>
__global__ void __launch_bounds__(kTpb) outer_error_kernel(int nv, const T* __restrict__ src, T* __restrict__ sn,
                                                           T* __restrict__ prev, double* __restrict__ slots,
                                                           const int* __restrict__ skip) {
    if (skip && *skip) return;
    const unsigned v = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    double vals[1] = {0};
    if (v < (unsigned)nv) {
        T x0, x1, x2, p0, p1, p2;
        ld3(src, v, x0, x1, x2);
        ld3(prev, v, p0, p1, p2);
        const T d0 = p0 - x0, d1 = p1 - x1, d2 = p2 - x2;
        vals[0] = (double)(d0 * d0) + (double)(d1 * d1) + (double)(d2 * d2);
        if (sn) st3(sn, v, x0, x1, x2);
        st3(prev, v, x0, x1, x2);
    }
    stripe_add<1>(vals, slots + kSlotErr * kSlotSize);
}

// Off-diagonal entries of row v of the sliced-ELL matrix, issued kEllBatch at
// a time: column and value loads of a batch are independent, then the batch's
// gathers are independent, so a row costs width/kEllBatch dependent rounds of
// memory latency instead of width.  Positions past the row's width are
// clamped to its last entry (in bounds) and given the value 0, so every load
// is unconditional.  f(u, m) accumulates m * x[u] in entry order.
#ifndef SBS_ELL_BATCH
#define SBS_ELL_BATCH 8
#endif
constexpr int kEllBatch = SBS_ELL_BATCH;
template <int B = kEllBatch, class T, class F>
__device__ __forceinline__ void ell_row(const int* __restrict__ slice_ptr, const int* __restrict__ slice_w,
                                        const uint32_t* __restrict__ ecol, const T* __restrict__ eval, unsigned v,
                                        F&& f) {
    const unsigned s = v / kSlice, lane = v % kSlice;
    const size_t base = (size_t)slice_ptr[s] + lane;
    const int width = slice_w[s];
    for (int e0 = 0; e0 < width; e0 += B) {
        uint32_t u[B];
        T m[B];
#pragma unroll
        for (int j = 0; j < B; ++j) {
            const int e = min(e0 + j, width - 1);
            const size_t pos = base + (size_t)e * kSlice;
            u[j] = ecol[pos];
            const T mv = eval[pos];
            m[j] = e0 + j < width ? mv : (T)0;
        }
#pragma unroll
        for (int j = 0; j < B; ++j) f(u[j], m[j]);
    }
}

// ---------------------------------------------------------------------------
// Jacobi-PCG over the scalar matrix with 3 fused right-hand sides, in the
// Chronopoulos-Gear form: both inner products of an iteration are taken in
// one pass, so an iteration is ONE kernel launch (a kernel boundary is the
// cheapest grid-wide synchronisation on MI355X: a dependent launch costs
// ~2.5-3.7 us, less than any in-kernel grid barrier measured).
//
//   init1:  r0 = b - A x0, x = x0, (D^-1 b)^2
//   init2:  u0 = D^-1 r0, w0 = A u0, gamma0 = r0.u0, delta0 = w0.u0, u0.u0
//   K(i):   stop test on ||u_i|| and scalars (every block computes them from
//           the reduction stripes of the previous launch; block 0 records them)
//             beta_i  = gamma_i / gamma_{i-1}                     (beta_0 = 0)
//             alpha_i = gamma_i / (delta_i - beta_i gamma_i / alpha_{i-1})
//           then per row
//             p_i = u_i + beta_i p_{i-1},  s_i = w_i + beta_i s_{i-1}
//             x_{i+1} = x_i + alpha_i p_i, r_{i+1} = r_i - alpha_i s_i
//             u_{i+1} = D^-1 r_{i+1},      w_{i+1} = A u_{i+1}
//           and gamma_{i+1}, delta_{i+1}, u.u into the stripes.
// The neighbours' part of A u_{i+1} is formed from their r_i, w_i, s_{i-1}
// (double-buffered, read-only in K(i)) as three sums taken before alpha_i and
// beta_i are known (see pcg_iter_kernel).  This is the reference's
// Jacobi-PCG (pcgJacobi.cu:89-173) reorganised; in exact arithmetic the
// iterates are identical.  A component stops when
// ||u_k|| <= max(tol ||u_0||, floor ||D^-1 b||); launches after the stop
// return at once.
//
// Two kernel families run this iteration: the row-gather kernels below read
// each neighbour's record from memory per nonzero; the halo-staged kernels
// further down stage a workgroup's whole neighbourhood in LDS once per launch.

// Halo tiles: rows renumbered along a Morton curve of the rest positions
// (internal row i = vertex perm[i]) and grouped into workgroups of up to kTpb
// consecutive internal rows; a workgroup's halo is the ascending list of the
// internal rows its rows touch (their own and their neighbours').
#ifndef SBS_HALO_STRIDE
#define SBS_HALO_STRIDE 768
#endif
constexpr int kHaloStride = SBS_HALO_STRIDE;  // halo entries (LDS slots) per workgroup
constexpr int kHaloPer = kHaloStride / kTpb;  // halo entries staged per thread
constexpr int kGrpInts = 8;                   // per workgroup: row0, rows, halo size, width, first entry, own0
template <class T> struct HaloView {
    int groups = 0;
    const int* grp = nullptr;
    const uint32_t* halo = nullptr;    // groups x kHaloStride internal rows
    const uint32_t* halo_v = nullptr;  // the same entries as vertex ids
    const uint16_t* col = nullptr;     // entry e of thread t at first + e * kTpb + t: halo-local column
    const T* val = nullptr;            // its value (diagonal included, padding 0)
    const uint32_t* perm = nullptr;    // internal row -> vertex
    const uint2* vent = nullptr;       // internal row -> its vertex's partial range (null: none)
    T* x = nullptr;                    // the iterate in internal order
};

template <class T> struct PcgArgs {
    int nv;
    int gathered;             // 1: rhs from the local-step partials (+ pins), 0: plain b
    RhsGather<T> g;
    const T* b;
    const T* x0;              // warm start (null = 0)
    T* x;                     // solution
    const int* slice_ptr;
    const int* slice_w;
    const uint32_t* ecol;
    const T* eval;
    const T* diagA;
    const T* dinv;
    T* rec[2];                // records of iteration i in rec[i&1]: r_i, w_i, s_{i-1}, D^-1
    T* p;
    T* sn;
    T* prev;
    T omega;
    double* slots;            // kNumSlots x kSlotSize doubles
    CtlBlock* ctl;
    int max_iter;
    double tol2;
    double floor2;
    const int* skip;          // outer loop already converged
    HaloView<T> h;            // halo-staged kernels only: records and p are then in internal order
    T* hrec[2];               // halo kernels: R', W', S' planes (9 N_v scalars), double-buffered
};

__device__ __forceinline__ double* iter_slot(double* slots, int i) {
    return slots + (kSlotIter + ((i % 3) + 3) % 3) * kSlotSize;
}

template <class T>
__global__ void __launch_bounds__(kTpb) pcg_init1_kernel(PcgArgs<T> a) {
    if (a.skip && *a.skip) return;
    const unsigned v = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    double vals[3] = {0, 0, 0};
    if (v < (unsigned)a.nv) {
        T b0, b1, b2;
        if (a.gathered) a.g.at(v, b0, b1, b2);
        else ld3(a.b, v, b0, b1, b2);
        T ax0 = 0, ax1 = 0, ax2 = 0, x0 = 0, x1 = 0, x2 = 0;
        if (a.x0) {
            ld3(a.x0, v, x0, x1, x2);
            const T d = a.diagA[v];
            ax0 = d * x0; ax1 = d * x1; ax2 = d * x2;
            ell_row(a.slice_ptr, a.slice_w, a.ecol, a.eval, v, [&](uint32_t u, T m) {
                T u0, u1, u2;
                ld3(a.x0, u, u0, u1, u2);
                ax0 += m * u0; ax1 += m * u1; ax2 += m * u2;
            });
        }
        st3(a.x, v, x0, x1, x2);
        const T di = a.dinv[v];
        Rec<T> rc;
        for (int k = 0; k < kRec; ++k) rc.v[k] = (T)0;
        rc.v[kRecR] = b0 - ax0; rc.v[kRecR + 1] = b1 - ax1; rc.v[kRecR + 2] = b2 - ax2;
        rc.v[kRecD] = di;
        st_rec(a.rec[0], v, rc);
        const double d0 = (double)di * b0, d1 = (double)di * b1, d2 = (double)di * b2;
        vals[0] = d0 * d0; vals[1] = d1 * d1; vals[2] = d2 * d2;
    }
    stripe_add<3>(vals, a.slots + kSlotInit * kSlotSize);
}

template <class T>
__global__ void __launch_bounds__(kTpb) pcg_init2_kernel(PcgArgs<T> a) {
    if (a.skip && *a.skip) return;
    const unsigned v = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    double vals[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (v < (unsigned)a.nv) {
        T* rec = a.rec[0];
        T r0, r1, r2, dv;
        ld_rec_rd(rec, v, r0, r1, r2, dv);
        const T d = a.diagA[v];
        const T u0 = dv * r0, u1 = dv * r1, u2 = dv * r2;
        T w0 = d * u0, w1 = d * u1, w2 = d * u2;
        ell_row(a.slice_ptr, a.slice_w, a.ecol, a.eval, v, [&](uint32_t n, T m) {
            T q0, q1, q2, dn;
            ld_rec_rd(rec, n, q0, q1, q2, dn);
            w0 += m * (dn * q0); w1 += m * (dn * q1); w2 += m * (dn * q2);
        });
        // w0 into the record (neighbours read only r and D^-1 of it in this launch)
        st_rec_w(rec, v, w0, w1, w2);
        vals[0] = (double)r0 * u0; vals[1] = (double)r1 * u1; vals[2] = (double)r2 * u2;
        vals[3] = (double)w0 * u0; vals[4] = (double)w1 * u1; vals[5] = (double)w2 * u2;
        vals[6] = (double)u0 * u0; vals[7] = (double)u1 * u1; vals[8] = (double)u2 * u2;
    }
    stripe_add<9>(vals, iter_slot(a.slots, -1));
}

// Scalars of iteration i for a kernel body (uniform over the wavefront).
struct PcgStep {
    double alpha[3];
    double beta[3];
    int iters, done, unconverged;
};

// Row r < 3 of the wavefront gets the total of quantity K0 + r of a slot (a
// lane adds two stripes, then the 16-lane row sum); row 3 gets 0.  The lane
// value is loaded by every lane (row 3 at component 2's address and then
// zeroed) with no branch around it, so the loads of several quantities are
// in flight together instead of each waiting in its own masked block.
template <int K0> __device__ __forceinline__ double slot_pair3(const double* slot) {
    static_assert(kStripes == 32, "two stripes per lane, 16 lanes per quantity");
    const int lane = threadIdx.x & 63, r = lane >> 4, s = lane & 15;
    const int k = K0 + (r < 3 ? r : 2);
    const double v = slot[s * kSlotQ + k] + slot[(s + 16) * kSlotQ + k];
    return r < 3 ? v : 0.0;
}
template <int K0> __device__ __forceinline__ double slot_rows3(const double* slot) {
    return row16_total(slot_pair3<K0>(slot));
}

// Stop test and scalars of iteration i, identical in every wavefront; block 0
// records the state for the next launch.  Component c of the three
// right-hand sides is evaluated by row c of the wavefront (row 3 repeats
// component 2 and is ignored), so a lane does one component's divisions;
// alpha and beta are then read from lanes 0, 16 and 32.  Every lane of the
// wavefront must be active.
template <class T> __device__ __forceinline__ PcgStep pcg_state(const PcgArgs<T>& a, int i) {
    const int lane = threadIdx.x & 63, row = lane >> 4;
    const int c = row < 3 ? row : 2;
    PcgStep out;
    PcgState* rec = &a.ctl->st[i & 1];
    const bool writer = blockIdx.x == 0 && threadIdx.x < 64;
    bool act;
    double gamma, thr, alpha = 0.0, beta = 0.0;
    const PcgState* prev = i > 0 ? &a.ctl->st[(i - 1) & 1] : nullptr;
    bool carried = false;  // the solve stopped earlier: the state is carried forward
    int c_iters = 0, c_unconverged = 0;
    if (i == 0) {
        const double* sl = iter_slot(a.slots, -1);
        const double vg = slot_pair3<0>(sl), vd = slot_pair3<3>(sl), vu = slot_pair3<6>(sl);
        const double vb = slot_pair3<0>(a.slots + kSlotInit * kSlotSize);
        const double g = row16_total(vg), dl = row16_total(vd), uu = row16_total(vu), dd = row16_total(vb);
        thr = fmax(a.tol2 * uu, a.floor2 * dd);
        act = uu > thr;
        gamma = g;
        alpha = (act && dl != 0.0) ? g / dl : 0.0;
    } else {
        // The stop flag, the stripes of launch i-1 and this component's state
        // are requested together and consumed without a branch (the scalars are
        // formed in every case and selected), so they cost one memory round
        // trip instead of one each.
        const double* sl = iter_slot(a.slots, i - 1);
        const double vg = slot_pair3<0>(sl), vd = slot_pair3<3>(sl), vu = slot_pair3<6>(sl);
        const double pthr = prev->thr[c], pg = prev->gamma[c], pa = prev->alpha[c], pb = prev->beta[c];
        const int pact = prev->active[c];
        carried = prev->done != 0;
        c_iters = prev->iters;
        c_unconverged = prev->unconverged;
        const double g = row16_total(vg), dl = row16_total(vd), uu = row16_total(vu);
        thr = pthr;
        act = !carried && pact && uu > thr;
        const double bt = g / pg;
        const double den = dl - bt * g / pa;
        const double al = g / den;
        const bool ok = act && pg != 0.0;
        beta = ok ? bt : 0.0;
        alpha = ok && den != 0.0 ? al : 0.0;
        if (carried) {
            alpha = pa;
            beta = pb;
        }
        gamma = g;
    }
    const bool any = __ballot(act && row < 3) != 0ull;
    out.iters = carried ? c_iters : i;
    out.done = carried || !any || i >= a.max_iter;
    out.unconverged = carried ? c_unconverged : (any && i >= a.max_iter);
    out.alpha[0] = readlane_f64<0>(alpha);
    out.alpha[1] = readlane_f64<16>(alpha);
    out.alpha[2] = readlane_f64<32>(alpha);
    out.beta[0] = readlane_f64<0>(beta);
    out.beta[1] = readlane_f64<16>(beta);
    out.beta[2] = readlane_f64<32>(beta);
    if (writer) {
        if (carried) {
            if (threadIdx.x == 0) *rec = *prev;
        } else {
            if ((lane & 15) == 0 && row < 3) {
                rec->gamma[c] = gamma;
                rec->alpha[c] = alpha;
                rec->beta[c] = beta;
                rec->thr[c] = thr;
                rec->active[c] = act;
            }
            if (threadIdx.x == 0) {
                rec->iters = out.iters;
                rec->done = out.done;
                rec->unconverged = out.unconverged;
                rec->pad = 0;
            }
        }
    }
    return out;
}

#ifndef SBS_ITER_BATCH
#define SBS_ITER_BATCH 4
#endif
constexpr int kIterBatch = SBS_ITER_BATCH;  // neighbours in flight per row in the iteration kernel
// Unroll of the local step's per-vertex incidence sums: the LDS reads of
// several incidences are in flight together (the sum keeps its order).
#ifndef SBS_GATHER_UNROLL
#define SBS_GATHER_UNROLL 2
#endif
#define SBS_PRAGMA(x) _Pragma(#x)
#define SBS_UNROLL(n) SBS_PRAGMA(unroll n)

// One PCG iteration.  Everything that does not depend on alpha_i, beta_i is
// issued first: the row's own values and the three neighbour sums
//   Yr = sum m D^-1 r_i,  Yw = sum m D^-1 w_i,  Ys = sum m D^-1 s_{i-1}
// so that A u_{i+1} (off-diagonal part) = Yr - alpha (Yw + beta Ys) follows
// once the scalars are known; the reduction stripes of the previous launch
// are read while those loads are in flight.
// 5 wavefronts per SIMD keep every row of the 64^3 cube resident at once
// (4292 wavefronts on 1024 SIMDs).
template <class T, bool kFirst>
__global__ void __launch_bounds__(kTpb) __attribute__((amdgpu_waves_per_eu(5))) pcg_iter_kernel(PcgArgs<T> a, int i) {
    if (a.skip && *a.skip) return;
    if (!kFirst && a.ctl->st[(i - 1) & 1].done) {  // stopped earlier: carry the state forward
        if (blockIdx.x == 0 && threadIdx.x == 0) a.ctl->st[i & 1] = a.ctl->st[(i - 1) & 1];
        return;
    }
    const unsigned v = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const bool row = v < (unsigned)a.nv;
    const unsigned vr = row ? v : (unsigned)a.nv - 1;  // loads stay in bounds; results unused
    const int rd = i & 1, wr = rd ^ 1;
    const T* __restrict__ rec_in = a.rec[rd];
    const Rec<T> own = ld_rec(rec_in, vr);
    const T r0 = own.v[kRecR], r1 = own.v[kRecR + 1], r2 = own.v[kRecR + 2];
    const T w0 = own.v[kRecW], w1 = own.v[kRecW + 1], w2 = own.v[kRecW + 2];
    T s0 = own.v[kRecS], s1 = own.v[kRecS + 1], s2 = own.v[kRecS + 2];
    const T dv = own.v[kRecD];
    T x0, x1, x2, p0 = 0, p1 = 0, p2 = 0;
    if (!kFirst) ld3(a.p, vr, p0, p1, p2);
    ld3(a.x, vr, x0, x1, x2);
    const T d = a.diagA[vr];
    T yr0 = 0, yr1 = 0, yr2 = 0, yw0 = 0, yw1 = 0, yw2 = 0, ys0 = 0, ys1 = 0, ys2 = 0;
    ell_row<kIterBatch>(a.slice_ptr, a.slice_w, a.ecol, a.eval, vr, [&](uint32_t n, T m) {
        const Rec<T> nb = ld_rec(rec_in, n);
        const T md = m * nb.v[kRecD];
        yr0 += md * nb.v[kRecR]; yr1 += md * nb.v[kRecR + 1]; yr2 += md * nb.v[kRecR + 2];
        yw0 += md * nb.v[kRecW]; yw1 += md * nb.v[kRecW + 1]; yw2 += md * nb.v[kRecW + 2];
        if (!kFirst) { ys0 += md * nb.v[kRecS]; ys1 += md * nb.v[kRecS + 1]; ys2 += md * nb.v[kRecS + 2]; }
    });
    const PcgStep st = pcg_state(a, i);
    if (st.done) return;
    if (blockIdx.x == 0) {  // slot of launch i+1: last read by launch i-1, which has finished
        double* z = iter_slot(a.slots, i + 1);
        for (int k = threadIdx.x; k < kSlotSize; k += kTpb) z[k] = 0.0;
    }
    const T al0 = (T)st.alpha[0], al1 = (T)st.alpha[1], al2 = (T)st.alpha[2];
    const T be0 = (T)st.beta[0], be1 = (T)st.beta[1], be2 = (T)st.beta[2];
    double vals[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (row) {
        // p_i = u_i + beta p_{i-1}, s_i = w_i + beta s_{i-1}, x_{i+1}, r_{i+1}, u_{i+1}
        const T u0 = dv * r0, u1 = dv * r1, u2 = dv * r2;
        if (!kFirst) {
            p0 = fma(be0, p0, u0); p1 = fma(be1, p1, u1); p2 = fma(be2, p2, u2);
            s0 = fma(be0, s0, w0); s1 = fma(be1, s1, w1); s2 = fma(be2, s2, w2);
        } else {
            p0 = u0; p1 = u1; p2 = u2;
            s0 = w0; s1 = w1; s2 = w2;
        }
        x0 = fma(al0, p0, x0); x1 = fma(al1, p1, x1); x2 = fma(al2, p2, x2);
        const T rn0 = fma(-al0, s0, r0), rn1 = fma(-al1, s1, r1), rn2 = fma(-al2, s2, r2);
        const T un0 = dv * rn0, un1 = dv * rn1, un2 = dv * rn2;
        // w_{i+1} = diag u_{i+1} + Yr - alpha (Yw + beta Ys)
        const T y0 = d * un0 + (yr0 - al0 * (kFirst ? yw0 : fma(be0, ys0, yw0)));
        const T y1 = d * un1 + (yr1 - al1 * (kFirst ? yw1 : fma(be1, ys1, yw1)));
        const T y2 = d * un2 + (yr2 - al2 * (kFirst ? yw2 : fma(be2, ys2, yw2)));
        st3(a.p, v, p0, p1, p2);
        st3(a.x, v, x0, x1, x2);
        Rec<T> out;
        out.v[kRecR] = rn0; out.v[kRecR + 1] = rn1; out.v[kRecR + 2] = rn2; out.v[kRecD] = dv;
        out.v[kRecW] = y0; out.v[kRecW + 1] = y1; out.v[kRecW + 2] = y2; out.v[kRecW + 3] = (T)0;
        out.v[kRecS] = s0; out.v[kRecS + 1] = s1; out.v[kRecS + 2] = s2; out.v[kRecS + 3] = (T)0;
        st_rec(a.rec[wr], v, out);
        vals[0] = (double)rn0 * un0; vals[1] = (double)rn1 * un1; vals[2] = (double)rn2 * un2;
        vals[3] = (double)y0 * un0; vals[4] = (double)y1 * un1; vals[5] = (double)y2 * un2;
        vals[6] = (double)un0 * un0; vals[7] = (double)un1 * un1; vals[8] = (double)un2 * un2;
    }
    stripe_add<9>(vals, iter_slot(a.slots, i));
}

// Solve log and counters of the control block from the state after the L
// launched iterations.
__device__ __forceinline__ void pcg_log_solve(CtlBlock* c, int iters, int done, int unconverged, int L) {
    const int its = done ? iters : L;
    const bool cut = unconverged || !done;  // max_iter or the launch cap
    if (c->solves < kMaxSolveLog) c->log[c->solves] = cut ? -its - 1 : its;
    c->pcg_iters += its;
    c->solves += 1;
    c->last_iters = its;
    if (its > c->max_used) c->max_used = its;
    if (cut) c->unconverged += 1;
}

// State after the L launched iterations (the stop test of iteration L,
// evaluated by the first wavefront of block 0 only).
template <class T> __device__ __forceinline__ void pcg_close(const PcgArgs<T>& a, int L) {
    if (blockIdx.x != 0 || threadIdx.x >= 64) return;
    const PcgState* prev = L > 0 ? &a.ctl->st[(L - 1) & 1] : nullptr;
    if (prev && prev->done) {
        if (threadIdx.x == 0) pcg_log_solve(a.ctl, prev->iters, 1, prev->unconverged, L);
        return;
    }
    const PcgStep st = pcg_state(a, L);
    if (threadIdx.x == 0) pcg_log_solve(a.ctl, st.iters, st.done, st.unconverged, L);
}

// After K(L-1): the stop test of iteration L and bookkeeping (block 0), and with `cheb` the
// Chebyshev recurrence of the global step (pdSolver.cu:196-205) into sn/prev.
template <class T>
__global__ void __launch_bounds__(kTpb) pcg_finish_kernel(PcgArgs<T> a, int L, int cheb) {
    if (a.skip && *a.skip) return;
    pcg_close(a, L);
    if (!cheb) return;
    const unsigned v = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (v < (unsigned)a.nv) {
        for (int k = 0; k < 3; ++k) {
            const size_t i = 3 * (size_t)v + k;
            const T s = a.sn[i];
            const T nx = cheb_update(a.x[i], s, a.prev[i], a.omega);
            a.prev[i] = s;
            a.sn[i] = nx;
        }
    }
}

// ---------------------------------------------------------------------------
// computeLocal (src/simulation/solver/projective/pdUtil.cu:97-145), one
// workgroup per chunk of kChunkTets Morton-ordered tets:
//   1. the chunk's vertices (a sorted list) are staged from q into LDS, with
//      the chunk's incidence list (per vertex, the (tet, corner) pairs that
//      touch it, in tet order);
//   2. one thread per tet: F = Ds Dm^-1, 3x3 SVD in registers, R (or R - F),
//      the 4 corner forces of w R Dm^-T G^T written to LDS (plain stores);
//   3. one thread per chunk vertex sums its incidences in list order (the
//      three components together, each code decoded once) and writes the
//      chunk's partial force for that vertex at the vertex's slot of the
//      vertex-major partial array.
// No atomics anywhere, so the result is deterministic; float atomics on LDS
// serialise on the shared corners of neighbouring tets and were the
// bottleneck of the scatter form.
//
// With kFold the kernel also finishes the previous PCG solve (instead of a
// finish launch of its own): block 0 evaluates that solve's last stop test and
// bookkeeping (pcg_close) before it clears the reduction slots, and q is not
// read but formed at staging from that solve's iterate x (internal row order)
// by the Chebyshev recurrence (pdSolver.cu:196-205),
//   q = (0.9 (x - sn) + sn - prev) omega + prev,
// which every chunk holding the vertex evaluates identically (cheb_update);
// the vertex's first entry writes q to sn_out and sn to prev_out (double
// buffers: other chunks still read sn_in / prev_in).
template <class T> struct FoldIn {
    const T* x;           // the previous solve's iterate, internal row order
    const uint32_t* rows; // per chunk entry: internal row of its vertex, bit 31 on the vertex's first entry
    const T* sn_in;
    const T* prev_in;
    T* sn_out;
    T* prev_out;
    T omega;
};

// flags of the local step
constexpr int kLocalResidual = 1;  // P = R - F (Jacobi mode) instead of R
constexpr int kLocalPolar = 2;     // FP32: rotation by Newton's polar iteration (SVD fallback)

template <class T, bool kFold>
__global__ void __launch_bounds__(kChunkTets) local_chunk_kernel(
    int nt, const uint2* __restrict__ lidx, const int* __restrict__ chunk_vptr,
    const uint32_t* __restrict__ chunk_verts, const uint32_t* __restrict__ entry_pos,
    const uint16_t* __restrict__ inc_codes,
    const uint16_t* __restrict__ inc_end, const T* __restrict__ dminv, const T* __restrict__ w,
    const T* __restrict__ q, T* __restrict__ partial, int flags, const int* __restrict__ skip,
    double* __restrict__ zero_slots, int zero_n, FoldIn<T> fold, PcgArgs<T> close, int close_L) {
    if (skip && *skip) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    if (blockIdx.x == 0 && zero_slots) {  // the PCG solve that follows accumulates into these
        if (kFold) {
            pcg_close(close, close_L);     // reads the previous solve's last slot first
            __syncthreads();
        }
        for (int i = threadIdx.x; i < zero_n; i += kChunkTets) zero_slots[i] = 0.0;
    }
    const unsigned c = xcd_block(blockIdx.x, gridDim.x);
    const int e0 = chunk_vptr[c], nu = chunk_vptr[c + 1] - e0;
    T* fs = reinterpret_cast<T*>(smem_raw);                    // [12][kChunkTets] corner forces
    T* qs = fs + 12 * kChunkTets;                              // [nu][3] staged q
    uint16_t* codes = reinterpret_cast<uint16_t*>(qs + 3 * (size_t)nu);  // [4 kChunkTets] tet<<2 | corner
    uint16_t* ends = codes + 4 * kChunkTets;                   // [nu] end of each vertex's incidences
    for (int i = threadIdx.x; i < nu; i += kChunkTets) {
        const uint32_t g = chunk_verts[e0 + i];
        T x, y, z;
        if (kFold) {
            const uint32_t rw = fold.rows[e0 + i];
            T v[3], sv[3], pv[3];
            ld3(fold.x, rw & 0x7fffffffu, v[0], v[1], v[2]);
            ld3(fold.sn_in, g, sv[0], sv[1], sv[2]);
            ld3(fold.prev_in, g, pv[0], pv[1], pv[2]);
            for (int k = 0; k < 3; ++k) v[k] = cheb_update(v[k], sv[k], pv[k], fold.omega);
            if (rw >> 31) {
                st3(fold.sn_out, g, v[0], v[1], v[2]);
                st3(fold.prev_out, g, sv[0], sv[1], sv[2]);
            }
            x = v[0]; y = v[1]; z = v[2];
        } else {
            ld3(q, g, x, y, z);
        }
        qs[3 * i] = x; qs[3 * i + 1] = y; qs[3 * i + 2] = z;
        ends[i] = inc_end[e0 + i];
    }
    {
        const uint32_t* src = reinterpret_cast<const uint32_t*>(inc_codes) + (size_t)c * (2 * kChunkTets);
        uint32_t* dst = reinterpret_cast<uint32_t*>(codes);
        for (int k = threadIdx.x; k < 2 * kChunkTets; k += kChunkTets) dst[k] = src[k];
    }
    __syncthreads();
    const unsigned t = c * kChunkTets + threadIdx.x;
    if (t < (unsigned)nt) {
        const uint2 li = lidx[t];
        const unsigned l0 = li.x & 0xffffu, l1 = li.x >> 16, l2 = li.y & 0xffffu, l3 = li.y >> 16;
        T m[9];
        load_dminv(dminv, nt, t, m);
        const T ww = w[t];
        T F[3][3];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const T x0 = qs[3 * l0 + r];
            const T d0 = qs[3 * l1 + r] - x0, d1 = qs[3 * l2 + r] - x0, d2 = qs[3 * l3 + r] - x0;
#pragma unroll
            for (int k = 0; k < 3; ++k) F[r][k] = d0 * m[k] + d1 * m[3 + k] + d2 * m[6 + k];
        }
        T P[3][3];
        pd_projection(F, (flags & kLocalResidual) != 0, P, (flags & kLocalPolar) != 0);
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            // column j of w P Dm^-T is w P (row j of Dm^-1); corner 0 gets minus their sum
            const T f1 = ww * (P[r][0] * m[0] + P[r][1] * m[1] + P[r][2] * m[2]);
            const T f2 = ww * (P[r][0] * m[3] + P[r][1] * m[4] + P[r][2] * m[5]);
            const T f3 = ww * (P[r][0] * m[6] + P[r][1] * m[7] + P[r][2] * m[8]);
            fs[(0 * 3 + r) * kChunkTets + threadIdx.x] = -(f1 + f2 + f3);
            fs[(1 * 3 + r) * kChunkTets + threadIdx.x] = f1;
            fs[(2 * 3 + r) * kChunkTets + threadIdx.x] = f2;
            fs[(3 * 3 + r) * kChunkTets + threadIdx.x] = f3;
        }
    }
    __syncthreads();
    // one thread per chunk vertex sums all three components, so each
    // incidence code is read and decoded once
    for (int i = threadIdx.x; i < nu; i += kChunkTets) {
        const int k1 = ends[i];
        T a0 = 0, a1 = 0, a2 = 0;
        SBS_UNROLL(SBS_GATHER_UNROLL)
        for (int k = i ? ends[i - 1] : 0; k < k1; ++k) {
            const unsigned code = codes[k];
            const T* f = fs + (code & 3u) * 3 * kChunkTets + (code >> 2);
            a0 += f[0];
            a1 += f[kChunkTets];
            a2 += f[2 * kChunkTets];
        }
        st3(partial, entry_pos[e0 + i], a0, a1, a2);  // vertex-major: the consumers read contiguously
    }
}

// ---------------------------------------------------------------------------
// Halo-staged PCG kernels (same algorithm and stop test as above).  A launch
// stages what the product needs at every halo row into LDS once, then each
// thread forms its row's product from LDS through 16-bit halo-local columns,
// so a row's neighbours cost LDS reads instead of dependent memory round
// trips.  In the iteration the staged quantity is u_{i+1} = D^-1 r_{i+1} of
// the halo row, recomputed from its record (r_i, w_i, s_{i-1}, D^-1) with
// the owner's exact operation sequence, so every copy equals the owner's.
// Records, p and x are kept in internal order: a workgroup's own rows are
// contiguous and its halo is a few contiguous runs.

// kEllBatch entries of row t of the workgroup (columns and values), from
// entry e0 on; positions past the row's width repeat its last entry with value 0.
template <class T> struct RowBatch {
    unsigned c[kEllBatch];
    T m[kEllBatch];
};
// (Uniform base pointers and 32-bit byte offsets: the loads take a scalar base.)
template <class T>
__device__ __forceinline__ void row_batch(const HaloView<T>& h, int first, int width, int t, int e0, RowBatch<T>& b) {
    const char* __restrict__ c = reinterpret_cast<const char*>(h.col + first);
    const char* __restrict__ m = reinterpret_cast<const char*>(h.val + first);
#pragma unroll
    for (int j = 0; j < kEllBatch; ++j) {
        const unsigned o = (unsigned)min(e0 + j, width - 1) * kTpb + (unsigned)t;
        b.c[j] = *reinterpret_cast<const uint16_t*>(c + o * (unsigned)sizeof(uint16_t));
        const T mv = *reinterpret_cast<const T*>(m + o * (unsigned)sizeof(T));
        b.m[j] = e0 + j < width ? mv : (T)0;
    }
}
template <class T>
__device__ __forceinline__ void row_apply(const RowBatch<T>& b, T (*lu)[kHaloStride], T& y0, T& y1, T& y2) {
#pragma unroll
    for (int j = 0; j < kEllBatch; ++j) {
        y0 += b.m[j] * lu[0][b.c[j]];
        y1 += b.m[j] * lu[1][b.c[j]];
        y2 += b.m[j] * lu[2][b.c[j]];
    }
}
// Row t of the workgroup: y += sum over its entries of m * lu[.][col], the
// first batch already loaded (before the LDS stage, so its memory latency
// overlaps the stage's).
template <class T>
__device__ __forceinline__ void halo_row(const HaloView<T>& h, int first, int width, int t, const RowBatch<T>& b0,
                                         T (*lu)[kHaloStride], T& y0, T& y1, T& y2) {
    row_apply(b0, lu, y0, y1, y2);
    for (int e0 = kEllBatch; e0 < width; e0 += kEllBatch) {
        RowBatch<T> b;
        row_batch(h, first, width, t, e0, b);
        row_apply(b, lu, y0, y1, y2);
    }
}

struct HaloGroup {
    int row0, rows, hn, width, first, own0;  // own0: halo-local column of row row0 (row row0+t is own0+t)
};
template <class T> __device__ __forceinline__ HaloGroup halo_group(const HaloView<T>& h, unsigned g) {
    const int* gd = h.grp + kGrpInts * (size_t)g;
    return HaloGroup{gd[0], gd[1], gd[2], gd[3], gd[4], gd[5]};
}
// The descriptor as one vector load (lane k of each wavefront reads entry
// k mod kGrpInts) and its broadcast: a kernel that issues the load early and
// broadcasts later keeps the load in flight with its other first loads (the
// broadcasts are wavefront operations and stay where they are written).
template <class T> __device__ __forceinline__ int halo_group_lane(const HaloView<T>& h, unsigned g) {
    return h.grp[kGrpInts * (size_t)g + (threadIdx.x & (kGrpInts - 1))];
}
__device__ __forceinline__ HaloGroup halo_group_bcast(int v) {
    return HaloGroup{__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 1),
                     __builtin_amdgcn_readlane(v, 2), __builtin_amdgcn_readlane(v, 3),
                     __builtin_amdgcn_readlane(v, 4), __builtin_amdgcn_readlane(v, 5)};
}

// The halo kernels keep the PCG vectors that neighbours read scaled by D^-1,
// in three planes per buffer (xyz interleaved, internal row order):
//   R' = D^-1 r (= u), W' = D^-1 w, S' = D^-1 s,
// so a halo row costs three 12-byte loads (36 bytes, no D^-1 entry) and its
// u_{i+1} = R' - alpha (W' + beta S') needs no scaling.  The row's own
// diagonal (for the dot products r.u = sum d R'^2, w.u = sum w R') comes from
// its matrix entries.  The CG recurrences are linear, so scaling them row by
// row gives the same iterates in exact arithmetic.
template <class T> struct Planes {
    T* r;
    T* w;
    T* s;
};
template <class T> __device__ __forceinline__ Planes<T> planes(T* base, int nv) {
    return Planes<T>{base, base + 3 * (size_t)nv, base + 6 * (size_t)nv};
}

// halo_row, also summing the entries in the row's own column (its diagonal;
// the padding entries there carry 0).
template <class T>
__device__ __forceinline__ void halo_row_diag(const HaloView<T>& h, int first, int width, int t, const RowBatch<T>& b0,
                                              T (*lu)[kHaloStride], unsigned own, T& y0, T& y1, T& y2, T& d) {
    auto apply = [&](const RowBatch<T>& b) {
#pragma unroll
        for (int j = 0; j < kEllBatch; ++j) {
            y0 += b.m[j] * lu[0][b.c[j]];
            y1 += b.m[j] * lu[1][b.c[j]];
            y2 += b.m[j] * lu[2][b.c[j]];
            d += b.c[j] == own ? b.m[j] : (T)0;
        }
    };
    apply(b0);
    for (int e0 = kEllBatch; e0 < width; e0 += kEllBatch) {
        RowBatch<T> b;
        row_batch(h, first, width, t, e0, b);
        apply(b);
    }
}
// halo_row_diag with the first two batches already loaded.
template <class T>
__device__ __forceinline__ void halo_row_diag2(const HaloView<T>& h, int first, int width, int t,
                                               const RowBatch<T>& b0, const RowBatch<T>& b1, T (*lu)[kHaloStride],
                                               unsigned own, T& y0, T& y1, T& y2, T& d) {
    auto apply = [&](const RowBatch<T>& b) {
#pragma unroll
        for (int j = 0; j < kEllBatch; ++j) {
            y0 += b.m[j] * lu[0][b.c[j]];
            y1 += b.m[j] * lu[1][b.c[j]];
            y2 += b.m[j] * lu[2][b.c[j]];
            d += b.c[j] == own ? b.m[j] : (T)0;
        }
    };
    apply(b0);
    apply(b1);
    for (int e0 = 2 * kEllBatch; e0 < width; e0 += kEllBatch) {
        RowBatch<T> b;
        row_batch(h, first, width, t, e0, b);
        apply(b);
    }
}

// init1: r0 = b - A x0 (x0 staged at the halo), R'0 = D^-1 r0, x = x0, (D^-1 b)^2.
template <class T>
__global__ void __launch_bounds__(kTpb) pcg_init1_halo_kernel(PcgArgs<T> a) {
    if (a.skip && *a.skip) return;
    __shared__ T lx[3][kHaloStride];
    const unsigned g = xcd_block(blockIdx.x, gridDim.x);
    const HaloGroup G = halo_group(a.h, g);
    const int t = threadIdx.x;
    const bool row = t < G.rows;
    const size_t ir = (size_t)G.row0 + (row ? t : 0);
    const uint32_t v = a.h.perm[ir];
    T b0, b1, b2;
    if (a.gathered && a.h.vent) {
        const uint2 rv = a.h.vent[ir];  // requested with v, not after it
        a.g.at_range(v, (int)rv.x, (int)rv.y, b0, b1, b2);
    } else if (a.gathered) {
        a.g.at(v, b0, b1, b2);
    } else {
        ld3(a.b, v, b0, b1, b2);
    }
    T ax0 = 0, ax1 = 0, ax2 = 0, x0 = 0, x1 = 0, x2 = 0;
    if (a.x0) {
        RowBatch<T> bt;
        row_batch(a.h, G.first, G.width, t, 0, bt);
        const uint32_t* __restrict__ hv = a.h.halo_v + (size_t)g * kHaloStride;
        // every slot (the padded list holds valid vertices), so the loads go out together
#pragma unroll
        for (int k = 0; k < kHaloPer; ++k) {
            const int e = t + k * kTpb;
            T q0, q1, q2;
            ld3(a.x0, hv[e], q0, q1, q2);
            lx[0][e] = q0; lx[1][e] = q1; lx[2][e] = q2;
        }
        ld3(a.x0, v, x0, x1, x2);
        __syncthreads();
        halo_row(a.h, G.first, G.width, t, bt, lx, ax0, ax1, ax2);
    }
    double vals[3] = {0, 0, 0};
    if (row) {
        st3(a.h.x, ir, x0, x1, x2);
        const T di = a.dinv[v];
        st3(planes(a.hrec[0], a.nv).r, ir, di * (b0 - ax0), di * (b1 - ax1), di * (b2 - ax2));
        const double d0 = (double)di * b0, d1 = (double)di * b1, d2 = (double)di * b2;
        vals[0] = d0 * d0; vals[1] = d1 * d1; vals[2] = d2 * d2;
    }
    stripe_add<3>(vals, a.slots + kSlotInit * kSlotSize);
}

// init2: u0 = R'0 staged at the halo, w0 = A u0, W'0 = D^-1 w0;
// gamma0 = r0.u0 = sum d R'0^2, delta0 = w0.u0, u0.u0.
template <class T>
__global__ void __launch_bounds__(kTpb) pcg_init2_halo_kernel(PcgArgs<T> a) {
    if (a.skip && *a.skip) return;
    __shared__ T lu[3][kHaloStride];
    const unsigned g = xcd_block(blockIdx.x, gridDim.x);
    const HaloGroup G = halo_group(a.h, g);
    const int t = threadIdx.x;
    const bool row = t < G.rows;
    const size_t ir = (size_t)G.row0 + (row ? t : 0);
    const Planes<T> P = planes(a.hrec[0], a.nv);
    RowBatch<T> b0;
    row_batch(a.h, G.first, G.width, t, 0, b0);
    const uint32_t* __restrict__ hh = a.h.halo + (size_t)g * kHaloStride;
    // every slot (the padded list holds valid rows), so the loads go out together
#pragma unroll
    for (int k = 0; k < kHaloPer; ++k) {
        const int e = t + k * kTpb;
        T q0, q1, q2;
        ld3(P.r, hh[e], q0, q1, q2);
        lu[0][e] = q0; lu[1][e] = q1; lu[2][e] = q2;
    }
    T u0, u1, u2;
    ld3(P.r, ir, u0, u1, u2);
    __syncthreads();
    T w0 = 0, w1 = 0, w2 = 0, d = 0;
    halo_row_diag(a.h, G.first, G.width, t, b0, lu, (unsigned)(G.own0 + t), w0, w1, w2, d);
    double vals[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (row) {
        const T di = (T)1 / d;
        st3(P.w, ir, di * w0, di * w1, di * w2);
        vals[0] = (double)d * u0 * u0; vals[1] = (double)d * u1 * u1; vals[2] = (double)d * u2 * u2;
        vals[3] = (double)w0 * u0; vals[4] = (double)w1 * u1; vals[5] = (double)w2 * u2;
        vals[6] = (double)u0 * u0; vals[7] = (double)u1 * u1; vals[8] = (double)u2 * u2;
    }
    stripe_add<9>(vals, iter_slot(a.slots, -1));
}

// One PCG iteration K(i), halo-staged.  The group descriptor and the halo's
// row indices are requested with the reduction stripes and the state of the
// previous launch (a launch after the solve stopped carries that state
// forward inside pcg_state); once the scalars are formed the halo planes, the
// row's own planes, p, x and the row's first kEllBatch matrix entries are
// requested together, the next kEllBatch as soon as the halo values are
// staged, so the launch costs about two dependent memory round trips for rows
// of up to 2 kEllBatch entries.  The iteration that finds the solve
// converged returns after the scalars, before any of those loads (the
// decision is the same in every wavefront).
template <class T, bool kFirst>
__global__ void __launch_bounds__(kTpb) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 5 : 1)))
pcg_iter_halo_kernel(PcgArgs<T> a, int i) {
    if (a.skip && *a.skip) return;
    __shared__ T lu[3][kHaloStride];
    const unsigned g = xcd_block(blockIdx.x, gridDim.x);
    const int gdl = halo_group_lane(a.h, g);
    const int t = threadIdx.x;
    const int rd = i & 1, wr = rd ^ 1;
    const Planes<T> Pi = planes(a.hrec[rd], a.nv), Po = planes(a.hrec[wr], a.nv);
    const uint32_t* __restrict__ hh = a.h.halo + (size_t)g * kHaloStride;
    uint32_t hidx[kHaloPer];
#pragma unroll
    for (int k = 0; k < kHaloPer; ++k) hidx[k] = hh[t + k * kTpb];  // in bounds: kHaloStride per workgroup
    const PcgStep st = pcg_state(a, i);
    const HaloGroup G = halo_group_bcast(gdl);
    if (st.done) return;
    const bool row = t < G.rows;
    const size_t ir = (size_t)G.row0 + (row ? t : 0);
    // R', W', S' of the staged halo rows.  Every slot is loaded (the list is
    // padded with the workgroup's first row, so a slot past the halo reads a
    // valid row and stages a value no column refers to): loads under a
    // per-slot condition would each wait for the ones before them.
    T hq[kHaloPer][9];
#pragma unroll
    for (int k = 0; k < kHaloPer; ++k) {
        ld3(Pi.r, hidx[k], hq[k][0], hq[k][1], hq[k][2]);
        ld3(Pi.w, hidx[k], hq[k][3], hq[k][4], hq[k][5]);
        if (!kFirst) ld3(Pi.s, hidx[k], hq[k][6], hq[k][7], hq[k][8]);
    }
    T r0, r1, r2, w0, w1, w2, so0 = 0, so1 = 0, so2 = 0;
    ld3(Pi.r, ir, r0, r1, r2);
    ld3(Pi.w, ir, w0, w1, w2);
    if (!kFirst) ld3(Pi.s, ir, so0, so1, so2);
    T x0, x1, x2, p0 = 0, p1 = 0, p2 = 0;
    if (!kFirst) ld3(a.p, ir, p0, p1, p2);
    ld3(a.h.x, ir, x0, x1, x2);
    RowBatch<T> b0, b1;
    row_batch(a.h, G.first, G.width, t, 0, b0);
    const T al0 = (T)st.alpha[0], al1 = (T)st.alpha[1], al2 = (T)st.alpha[2];
    const T be0 = (T)st.beta[0], be1 = (T)st.beta[1], be2 = (T)st.beta[2];
    // u_{i+1} of the halo rows: S'_i = W'_i + beta S'_{i-1}, R'_{i+1} = R'_i - alpha S'_i
#pragma unroll
    for (int k = 0; k < kHaloPer; ++k) {
        const int e = t + k * kTpb;
        T sq0 = hq[k][3], sq1 = hq[k][4], sq2 = hq[k][5];
        if (!kFirst) {
            sq0 = fma(be0, hq[k][6], sq0); sq1 = fma(be1, hq[k][7], sq1); sq2 = fma(be2, hq[k][8], sq2);
        }
        lu[0][e] = fma(-al0, sq0, hq[k][0]);
        lu[1][e] = fma(-al1, sq1, hq[k][1]);
        lu[2][e] = fma(-al2, sq2, hq[k][2]);
    }
    // the second batch is requested once the staged halo values are out of
    // registers, so its latency overlaps the barrier and the first batch
    row_batch(a.h, G.first, G.width, t, kEllBatch, b1);
    if (blockIdx.x == 0) {  // slot of launch i+1: last read by launch i-1, which has finished
        double* z = iter_slot(a.slots, i + 1);
        for (int k = threadIdx.x; k < kSlotSize; k += kTpb) z[k] = 0.0;
    }
    __syncthreads();
    T y0 = 0, y1 = 0, y2 = 0, d = 0;  // w_{i+1} = A u_{i+1}
    halo_row_diag2(a.h, G.first, G.width, t, b0, b1, lu, (unsigned)(G.own0 + t), y0, y1, y2, d);
    double vals[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (row) {
        // p_i = u_i + beta p_{i-1} (u_i = R'_i), S'_i, x_{i+1}, R'_{i+1} = u_{i+1}
        T s0, s1, s2;
        if (!kFirst) {
            p0 = fma(be0, p0, r0); p1 = fma(be1, p1, r1); p2 = fma(be2, p2, r2);
            s0 = fma(be0, so0, w0); s1 = fma(be1, so1, w1); s2 = fma(be2, so2, w2);
        } else {
            p0 = r0; p1 = r1; p2 = r2;
            s0 = w0; s1 = w1; s2 = w2;
        }
        x0 = fma(al0, p0, x0); x1 = fma(al1, p1, x1); x2 = fma(al2, p2, x2);
        const T un0 = fma(-al0, s0, r0), un1 = fma(-al1, s1, r1), un2 = fma(-al2, s2, r2);
        const T di = (T)1 / d;
        st3(a.p, ir, p0, p1, p2);
        st3(a.h.x, ir, x0, x1, x2);
        st3(Po.r, ir, un0, un1, un2);
        st3(Po.w, ir, di * y0, di * y1, di * y2);
        st3(Po.s, ir, s0, s1, s2);
        vals[0] = (double)d * un0 * un0; vals[1] = (double)d * un1 * un1; vals[2] = (double)d * un2 * un2;
        vals[3] = (double)y0 * un0; vals[4] = (double)y1 * un1; vals[5] = (double)y2 * un2;
        vals[6] = (double)un0 * un0; vals[7] = (double)un1 * un1; vals[8] = (double)un2 * un2;
    }
    stripe_add<9>(vals, iter_slot(a.slots, i));
}

// After K(L-1): the stop test of iteration L and bookkeeping (block 0); the
// iterate goes back to vertex order, into x without `cheb`, and with `cheb`
// into the Chebyshev recurrence (pdSolver.cu:196-205) of sn/prev only.
template <class T>
__global__ void __launch_bounds__(kTpb) pcg_finish_halo_kernel(PcgArgs<T> a, int L, int cheb) {
    if (a.skip && *a.skip) return;
    pcg_close(a, L);
    const size_t ir = (size_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (ir >= (size_t)a.nv) return;
    const uint32_t v = a.h.perm[ir];
    T xv[3];
    ld3(a.h.x, ir, xv[0], xv[1], xv[2]);
    if (!cheb) {
        st3(a.x, v, xv[0], xv[1], xv[2]);
        return;
    }
    for (int k = 0; k < 3; ++k) {
        const size_t j = 3 * (size_t)v + k;
        const T s = a.sn[j];
        const T nx = cheb_update(xv[k], s, a.prev[j], a.omega);
        a.prev[j] = s;
        a.sn[j] = nx;
    }
}

// Entry values of the halo tiles from the assembled matrix: src >= 0 an
// off-diagonal sliced-ELL position, -1 padding, <= -2 the diagonal of row -2-src.
template <class T>
__global__ void halo_values_kernel(size_t n, const int* __restrict__ src, const T* __restrict__ eval,
                                   const T* __restrict__ diagA, T* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = src[i];
    out[i] = s >= 0 ? eval[s] : (s == -1 ? (T)0 : diagA[-2 - s]);
}

// Early exit of the local/global loop from the computeError sum (pdSolver.cu:164).
__global__ void outer_decide_kernel(const double* slots, int nv, double tol, CtlBlock* ctl, const int* skip) {
    if (skip && *skip) return;
    double t[1];
    slot_totals<1>(slots + kSlotErr * kSlotSize, t);
    if (threadIdx.x == 0) {
        const double err = t[0] / (3.0 * (double)nv);
        ctl->err = err;
        ctl->outer_done = !(sqrt(err) >= tol);
    }
}

}  // namespace

const char* kernel_name(int id) {
    static const char* names[kKCount] = {"local_step", "pcg_solve", "jacobi_chebyshev", "direct_solve",
                                         "rhs_gather", "pcg_iteration"};
    return (id >= 0 && id < kKCount) ? names[id] : "?";
}

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
    std::vector<T> X0t(3 * (size_t)nv_);
    SBS_HIP(hipMemcpy(X0t.data(), data.X0, X0t.size() * sizeof(T), hipMemcpyDeviceToHost));
    std::vector<double> X0h(X0t.begin(), X0t.end());
    build_chunks(X0h);
    {
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int v = 0; v < nv_; ++v)
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::min(lo[k], X0h[3 * (size_t)v + k]);
                hi[k] = std::max(hi[k], X0h[3 * (size_t)v + k]);
            }
        vertex_key_.resize(nv_);
        for (int v = 0; v < nv_; ++v) vertex_key_[v] = morton_code(&X0h[3 * (size_t)v], lo, hi);
    }
    dminv_.alloc(9 * (size_t)nt_);
    w_.alloc(nt_);
    v0_.alloc(nt_);
    // FEMSolver ctor: Dm^-1, V0 from the rest positions (femSolver.cu:7-17).
    hipLaunchKernelGGL(fem_precompute_kernel<T>, dim3(ceil_div(nt_, kTpb)), dim3(kTpb), 0, stream_, nt_, data.X0,
                       (const uint4*)data.Tet, tet_perm_d_.get(), data.mu, dminv_.get(), v0_.get(), w_.get(),
                       data.V0);
    SBS_HIP(hipGetLastError());
    SBS_HIP(hipHostMalloc((void**)&host_counters_, 2 * kCounterInts * sizeof(int), hipHostMallocDefault));
    std::memset(host_counters_, 0, 2 * kCounterInts * sizeof(int));
    SBS_HIP(hipHostMalloc(&host_ctl_, sizeof(CtlBlock), hipHostMallocDefault));
    std::memset(host_ctl_, 0, sizeof(CtlBlock));
    for (auto& e : step_event_) SBS_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    SBS_HIP(hipStreamSynchronize(stream_));
}

template <class T> PdEngine<T>::~PdEngine() {
    drop_graph();
    for (auto& e : step_event_)
        if (e) {
            (void)hipEventSynchronize(e);
            (void)hipEventDestroy(e);
        }
    if (host_counters_) (void)hipHostFree(host_counters_);
    if (host_ctl_) (void)hipHostFree(host_ctl_);
    for (auto& pool : tev_)
        for (auto& e : pool) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
}

template <class T> void PdEngine<T>::reset() {
    // the next update prepares again and reallocates the step's buffers, so a
    // captured step (which holds their addresses) must not be replayed
    drop_graph();
    ready_ = false;
    for (auto& e : perf_data_) e.ms = 0.0;
}

// Morton order of tets by rest centroid, chunks of kChunkTets, per-chunk
// sorted vertex lists, u16 local indices and the vertex -> entry lists.
template <class T> void PdEngine<T>::build_chunks(const std::vector<double>& X0h) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int v = 0; v < nv_; ++v)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], X0h[3 * (size_t)v + k]);
            hi[k] = std::max(hi[k], X0h[3 * (size_t)v + k]);
        }
    auto spread = [](uint32_t x) {  // 10 bits -> every third bit
        x = (x * 0x00010001u) & 0xFF0000FFu;
        x = (x * 0x00000101u) & 0x0F00F00Fu;
        x = (x * 0x00000011u) & 0xC30C30C3u;
        x = (x * 0x00000005u) & 0x49249249u;
        return x;
    };
    std::vector<uint64_t> key(nt_);
    for (int t = 0; t < nt_; ++t) {
        uint32_t cell[3];
        for (int k = 0; k < 3; ++k) {
            double cen = 0;
            for (int j = 0; j < 4; ++j) cen += X0h[3 * (size_t)host_tets_[4 * (size_t)t + j] + k];
            cen *= 0.25;
            const double ext = hi[k] - lo[k];
            double u = ext > 0 ? (cen - lo[k]) / ext : 0.0;
            u = std::min(std::max(u * 1024.0, 0.0), 1023.0);
            cell[k] = (uint32_t)u;
        }
        const uint32_t code = (spread(cell[0]) << 2) | (spread(cell[1]) << 1) | spread(cell[2]);
        key[t] = ((uint64_t)code << 32) | (uint32_t)t;
    }
    std::sort(key.begin(), key.end());
    tet_perm_.resize(nt_);
    for (int i = 0; i < nt_; ++i) tet_perm_[i] = (int)(key[i] & 0xffffffffu);
    nchunks_ = (nt_ + kChunkTets - 1) / kChunkTets;
    std::vector<int> cvptr(nchunks_ + 1, 0);
    std::vector<uint32_t> cverts;
    cverts.reserve((size_t)nt_ * 2);
    std::vector<uint2> lidx(nt_);
    std::vector<uint32_t> tmp;
    std::vector<int> cnt(4 * kChunkTets);
    std::vector<uint16_t> inc_end;
    inc_end.reserve((size_t)nt_ * 2);
    std::vector<uint16_t> inc_codes((size_t)((nt_ + kChunkTets - 1) / kChunkTets) * 4 * kChunkTets, 0);
    chunk_max_verts_ = 0;
    for (int c = 0; c < nchunks_; ++c) {
        const int t0 = c * kChunkTets, t1 = std::min(nt_, t0 + kChunkTets);
        tmp.clear();
        for (int i = t0; i < t1; ++i)
            for (int j = 0; j < 4; ++j) tmp.push_back(host_tets_[4 * (size_t)tet_perm_[i] + j]);
        std::sort(tmp.begin(), tmp.end());
        tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
        chunk_max_verts_ = std::max(chunk_max_verts_, (int)tmp.size());
        std::fill(cnt.begin(), cnt.end(), 0);
        for (int i = t0; i < t1; ++i) {
            uint32_t l[4];
            for (int j = 0; j < 4; ++j) {
                const uint32_t g = host_tets_[4 * (size_t)tet_perm_[i] + j];
                l[j] = (uint32_t)(std::lower_bound(tmp.begin(), tmp.end(), g) - tmp.begin());
                cnt[l[j]]++;
            }
            lidx[i] = make_uint2(l[0] | (l[1] << 16), l[2] | (l[3] << 16));
        }
        // incidence list of the chunk: per local vertex, (tet << 2 | corner) in tet order
        const size_t nloc = tmp.size();
        std::vector<int> pos(nloc + 1, 0);
        for (size_t k = 0; k < nloc; ++k) pos[k + 1] = pos[k] + cnt[k];
        for (size_t k = 0; k < nloc; ++k) inc_end.push_back((uint16_t)pos[k + 1]);
        uint16_t* codes = &inc_codes[(size_t)c * 4 * kChunkTets];
        for (int i = t0; i < t1; ++i)
            for (int j = 0; j < 4; ++j) {
                const uint32_t word = j < 2 ? lidx[i].x : lidx[i].y;
                const uint32_t lv = (j & 1) ? word >> 16 : word & 0xffffu;
                codes[pos[lv]++] = (uint16_t)(((i - t0) << 2) | j);
            }
        cverts.insert(cverts.end(), tmp.begin(), tmp.end());
        cvptr[c + 1] = (int)cverts.size();
    }
    nentries_ = cverts.size();
    // vertex -> entries
    std::vector<int> vptr(nv_ + 1, 0);
    for (uint32_t g : cverts) vptr[g + 1]++;
    for (int v = 0; v < nv_; ++v) vptr[v + 1] += vptr[v];
    std::vector<uint32_t> vent(nentries_);
    std::vector<int> fill(vptr.begin(), vptr.end() - 1);
    for (size_t e = 0; e < nentries_; ++e) vent[fill[cverts[e]]++] = (uint32_t)e;
    std::vector<uint32_t> epos(nentries_);  // entry -> its slot in vertex-major order
    for (size_t j = 0; j < nentries_; ++j) epos[vent[j]] = (uint32_t)j;
    host_cverts_ = cverts;
    host_vptr_ = vptr;
    host_first_.assign(nentries_, 0);
    for (int v = 0; v < nv_; ++v)
        if (vptr[v + 1] > vptr[v]) host_first_[vent[vptr[v]]] = 1;
    tet_perm_d_.alloc(nt_);
    tet_lidx_.alloc(nt_);
    chunk_vptr_.alloc(nchunks_ + 1);
    chunk_verts_.alloc(nentries_);
    vent_ptr_.alloc(nv_ + 1);
    entry_pos_.alloc(nentries_);
    partial_f_.alloc(3 * nentries_);
    inc_codes_.alloc(inc_codes.size());
    inc_end_.alloc(std::max<size_t>(1, inc_end.size()));
    SBS_HIP(hipMemcpy(inc_codes_.get(), inc_codes.data(), inc_codes.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    SBS_HIP(hipMemcpy(inc_end_.get(), inc_end.data(), inc_end.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    SBS_HIP(hipMemcpy(tet_perm_d_.get(), tet_perm_.data(), nt_ * sizeof(int), hipMemcpyHostToDevice));
    SBS_HIP(hipMemcpy(tet_lidx_.get(), lidx.data(), nt_ * sizeof(uint2), hipMemcpyHostToDevice));
    SBS_HIP(hipMemcpy(chunk_vptr_.get(), cvptr.data(), cvptr.size() * sizeof(int), hipMemcpyHostToDevice));
    SBS_HIP(hipMemcpy(chunk_verts_.get(), cverts.data(), nentries_ * sizeof(uint32_t), hipMemcpyHostToDevice));
    SBS_HIP(hipMemcpy(vent_ptr_.get(), vptr.data(), vptr.size() * sizeof(int), hipMemcpyHostToDevice));
    SBS_HIP(hipMemcpy(entry_pos_.get(), epos.data(), nentries_ * sizeof(uint32_t), hipMemcpyHostToDevice));
}

// Sliced-ELL structure of the off-diagonal scalar pattern, built on the host
// from the tet list (one-time, like the reference's host-side prepare).
template <class T> void PdEngine<T>::build_structure() {
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
    SBS_HIP(hipMemcpy(slice_ptr_.get(), sptr.data(), sptr.size() * sizeof(int), hipMemcpyHostToDevice));
    SBS_HIP(hipMemcpy(slice_w_.get(), sw.data(), sw.size() * sizeof(int), hipMemcpyHostToDevice));
    SBS_HIP(hipMemcpy(rowlen_.get(), rowlen.data(), rowlen.size() * sizeof(int), hipMemcpyHostToDevice));
    if (!ecol.empty())
        SBS_HIP(hipMemcpy(ecol_.get(), ecol.data(), ecol.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    build_halo(rowstart, keys, sptr);
}

// Halo tiles over the full pattern (diagonal first in each row).  A mesh
// whose rows cannot be tiled (one row touching more than kHaloStride rows, or
// more entries than 32-bit positions hold) keeps the row-gather kernels.
template <class T>
void PdEngine<T>::build_halo(const std::vector<int>& rowstart, const std::vector<uint64_t>& keys,
                             const std::vector<int>& sptr) {
    halo_groups_ = 0;
    std::vector<int> rs(nv_ + 1, 0);
    std::vector<uint32_t> cols;
    std::vector<int64_t> vidx;  // value code of each CSR position
    cols.reserve(keys.size() + nv_);
    vidx.reserve(keys.size() + nv_);
    for (int v = 0; v < nv_; ++v) {
        cols.push_back((uint32_t)v);
        vidx.push_back(-2 - (int64_t)v);
        const int s = v / kSlice, lane = v % kSlice;
        for (int e = 0; e < rowstart[v + 1] - rowstart[v]; ++e) {
            cols.push_back((uint32_t)(keys[rowstart[v] + e] & 0xffffffffu));
            vidx.push_back((int64_t)sptr[s] + (int64_t)e * kSlice + lane);
        }
        rs[v + 1] = (int)cols.size();
    }
    HaloTiles h;
    try {
        h = build_halo_tiles(nv_, rs, cols, [&](int v) { return vertex_key_[v]; },
                             [&](int k) { return vidx[k]; }, kTpb, kHaloStride);
    } catch (const std::invalid_argument&) {
        return;
    }
    for (int64_t s : h.src)
        if (s > INT32_MAX || s < INT32_MIN) return;
    const int ng = (int)h.g_row0.size();
    std::vector<int> grp((size_t)ng * kGrpInts, 0);
    for (int g = 0; g < ng; ++g) {
        int* d = &grp[(size_t)g * kGrpInts];
        d[0] = h.g_row0[g]; d[1] = h.g_rows[g]; d[2] = h.g_halo_n[g]; d[3] = h.g_width[g]; d[4] = h.g_eptr[g];
        d[5] = h.g_own0[g];
    }
    std::vector<int> src(h.src.begin(), h.src.end());
    h_grp_.alloc(grp.size());
    h_halo_.alloc(h.halo.size());
    h_halo_v_.alloc(h.halo_vertex.size());
    h_perm_.alloc(nv_);
    h_col_.alloc(std::max<size_t>(1, h.col.size()));
    h_src_.alloc(std::max<size_t>(1, src.size()));
    h_val_.alloc(std::max<size_t>(1, src.size()));
    cg_x_.alloc(3 * (size_t)nv_);
    for (int k = 0; k < 2; ++k) cg_plane_[k].alloc(9 * (size_t)nv_);
    {   // the folded finish reads x by internal row and writes sn/prev once per vertex
        std::vector<uint32_t> rows(nentries_);
        for (size_t e = 0; e < nentries_; ++e)
            rows[e] = h.row_of[host_cverts_[e]] | (host_first_[e] ? 0x80000000u : 0u);
        chunk_rows_.alloc(std::max<size_t>(1, nentries_));
        SBS_HIP(hipMemcpy(chunk_rows_.get(), rows.data(), nentries_ * sizeof(uint32_t), hipMemcpyHostToDevice));
        sn_alt_.alloc(3 * (size_t)nv_);
        prev_alt_.alloc(3 * (size_t)nv_);
    }
    SBS_HIP(hipMemcpy(h_grp_.get(), grp.data(), grp.size() * sizeof(int), hipMemcpyHostToDevice));
    SBS_HIP(hipMemcpy(h_halo_.get(), h.halo.data(), h.halo.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    SBS_HIP(hipMemcpy(h_halo_v_.get(), h.halo_vertex.data(), h.halo_vertex.size() * sizeof(uint32_t),
                      hipMemcpyHostToDevice));
    SBS_HIP(hipMemcpy(h_perm_.get(), h.row_perm.data(), nv_ * sizeof(uint32_t), hipMemcpyHostToDevice));
    if ((int)host_vptr_.size() == nv_ + 1) {
        std::vector<uint2> rv(nv_);
        for (int r = 0; r < nv_; ++r) {
            const uint32_t v = h.row_perm[r];
            rv[r] = make_uint2((unsigned)host_vptr_[v], (unsigned)host_vptr_[v + 1]);
        }
        h_vent_.alloc(nv_);
        SBS_HIP(hipMemcpy(h_vent_.get(), rv.data(), nv_ * sizeof(uint2), hipMemcpyHostToDevice));
    }
    if (!src.empty()) {
        SBS_HIP(hipMemcpy(h_col_.get(), h.col.data(), h.col.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
        SBS_HIP(hipMemcpy(h_src_.get(), src.data(), src.size() * sizeof(int), hipMemcpyHostToDevice));
    }
    halo_entries_ = src.size();
    halo_groups_ = ng;
}

template <class T> void PdEngine<T>::assemble(const PdParams& params) {
    const T dt2 = (T)params.dt * (T)params.dt;
    SBS_HIP(hipMemsetAsync(eval_.get(), 0, eval_.bytes(), stream_));
    if (!row_inc_ptr_.get()) {  // each row's (internal tet << 2 | corner) list, ascending tet
        std::vector<int> ptr(nv_ + 1, 0);
        for (int i = 0; i < nt_; ++i)
            for (int a = 0; a < 4; ++a) ptr[host_tets_[4 * (size_t)tet_perm_[i] + a] + 1]++;
        for (int v = 0; v < nv_; ++v) ptr[v + 1] += ptr[v];
        std::vector<uint32_t> inc(ptr[nv_]);
        std::vector<int> fill(ptr.begin(), ptr.end() - 1);
        for (int i = 0; i < nt_; ++i)
            for (int a = 0; a < 4; ++a) inc[fill[host_tets_[4 * (size_t)tet_perm_[i] + a]]++] = ((uint32_t)i << 2) | a;
        row_inc_ptr_.alloc(nv_ + 1);
        row_inc_.alloc(std::max<size_t>(1, inc.size()));
        SBS_HIP(hipMemcpy(row_inc_ptr_.get(), ptr.data(), ptr.size() * sizeof(int), hipMemcpyHostToDevice));
        SBS_HIP(hipMemcpy(row_inc_.get(), inc.data(), inc.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    hipLaunchKernelGGL(assemble_rows_kernel<T>, dim3(ceil_div(nv_, kTpb)), dim3(kTpb), 0, stream_, nv_, nt_,
                       (const int*)row_inc_ptr_.get(), (const uint32_t*)row_inc_.get(), (const uint4*)data_.Tet,
                       tet_perm_d_.get(), dminv_.get(), w_.get(), slice_ptr_.get(), rowlen_.get(), ecol_.get(),
                       eval_.get(), diagL_.get());
    SBS_HIP(hipGetLastError());
    hipLaunchKernelGGL(mass_diag_kernel<T>, dim3(ceil_div(nv_, kTpb)), dim3(kTpb), 0, stream_, nv_, data_.mass,
                       data_.DBC, (T)1e6, dt2, diagL_.get(), c_.get(), diagA_.get(), dinv_.get());
    SBS_HIP(hipGetLastError());
    if (halo_groups_ > 0) {
        hipLaunchKernelGGL(halo_values_kernel<T>, dim3(ceil_div(halo_entries_, kTpb)), dim3(kTpb), 0, stream_,
                           halo_entries_, (const int*)h_src_.get(), (const T*)eval_.get(), (const T*)diagA_.get(),
                           h_val_.get());
        SBS_HIP(hipGetLastError());
    }
}

template <class T> void PdEngine<T>::prepare(PdData<T>& data, const PdParams& params) {
    drop_graph();  // the buffers a captured step refers to are reallocated below
    data_ = data;
    SBS_REQUIRE(params.dt > 0, "PdSolver: dt must be positive");
    const size_t n3 = 3 * (size_t)nv_;
    c_.alloc(nv_); diagL_.alloc(nv_); diagA_.alloc(nv_); dinv_.alloc(nv_);
    sn_.alloc(n3); sn_old_.alloc(n3); prev_.alloc(n3); next_.alloc(n3); rhs_.alloc(n3);
    for (int k = 0; k < 2; ++k) cg_rec_[k].alloc(kRec * (size_t)nv_);
    cg_p_.alloc(n3);
    slots_.alloc((size_t)kNumSlots * kSlotSize);
    ctl_.alloc(sizeof(CtlBlock));
    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(256), 0, stream_, slots_.get(), kNumSlots * kSlotSize);
    hipLaunchKernelGGL(ctl_init_kernel, dim3(1), dim3(64), 0, stream_, reinterpret_cast<CtlBlock*>(ctl_.get()));
    SBS_HIP(hipGetLastError());
    build_structure();
    assemble(params);
    system_built_ = true;
    // DBCX <- X0 (pdSolver.cu:134)
    SBS_HIP(hipMemcpyAsync(data.DBCX, data.X0, n3 * sizeof(T), hipMemcpyDeviceToDevice, stream_));
    direct_.reset();
    dense_.reset();
    batch_.reset();
    if (is_direct()) direct_factor(params);
    SBS_HIP(hipStreamSynchronize(stream_));
    ready_ = true;
}

template <class T> bool PdEngine<T>::is_direct() const {
    return solver_ == GlobalSolver::CuSolverCholesky || solver_ == GlobalSolver::EigenCholesky ||
           solver_ == GlobalSolver::DenseCholesky;
}

// Band Cholesky for the sparse direct modes; per-body dense inverses with
// FP64 refinement for DenseCholesky.
template <class T> void PdEngine<T>::direct_factor(const PdParams& params) {
    SBS_REQUIRE(system_built_, "direct factorisation requested before the system was assembled");
    Export ex = export_system();
    if (solver_ == GlobalSolver::DenseCholesky) {
        EllView<T> A;
        A.n = nv_;
        A.slice_ptr = slice_ptr_.get();
        A.slice_w = slice_w_.get();
        A.ecol = ecol_.get();
        A.eval = eval_.get();
        A.diag = diagA_.get();
        SBS_REQUIRE(params.dense_precision == 0 || params.dense_precision == 1,
                    "PdSolver: dense_precision must be 0 (FP32) or 1 (BF16)");
        dense_ = std::make_unique<DenseRefinedSolver<T>>();
        dense_->setup(A, ex.rowptr, ex.col, ex.val, static_cast<DensePrecision>(params.dense_precision), stream_);
        dense_prec_ = params.dense_precision;
    } else {
        direct_ = std::make_unique<BandCholesky<T>>();
        direct_->factor(nv_, ex.rowptr, ex.col, ex.val, stream_);
    }
}

// BatchedPCG: the bodies (connected components), their systems from the
// assembled matrix, and the tuned workgroup size of each size class.
template <class T> void PdEngine<T>::build_batch(const PdParams& params) {
    SBS_REQUIRE(system_built_, "batched solver requested before the system was assembled");
    drop_graph();
    Export ex = export_system();
    BatchSystem<T> sys;
    sys.nv = nv_;
    sys.nt = nt_;
    sys.tets = host_tets_;
    sys.dminv = std::move(ex.dminv);
    sys.w = std::move(ex.w);
    sys.c0 = std::move(ex.c);
    sys.diagA = std::move(ex.diagA);
    sys.rowptr = std::move(ex.rowptr);
    sys.col = std::move(ex.col);
    sys.val = std::move(ex.val);
    sys.mass.resize(nv_);
    sys.dbc.resize(nv_);
    SBS_HIP(hipMemcpy(sys.mass.data(), data_.mass, nv_ * sizeof(T), hipMemcpyDeviceToHost));
    SBS_HIP(hipMemcpy(sys.dbc.data(), data_.DBC, nv_ * sizeof(T), hipMemcpyDeviceToHost));
    sys.x0.resize(3 * (size_t)nv_);
    SBS_HIP(hipMemcpy(sys.x0.data(), data_.X0, sys.x0.size() * sizeof(T), hipMemcpyDeviceToHost));
    batch_ = std::make_unique<BatchEngine<T>>(sys);
    if (params.batch_autotune) {
        BatchParams bp;
        bp.iterations = params.num_iterations;
        bp.dt = params.dt;
        bp.gravity = params.gravity;
        bp.rho = params.rho;
        bp.pcg_tol = params.pcg_tol;
        bp.pcg_floor = params.pcg_floor;
        bp.pcg_max_iter = params.pcg_max_iter;
        bp.polar = params.polar_rotation != 0;
        batch_->autotune(data_.X, data_.V, data_.DBCX, bp, stream_);
    }
}

template <class T> void PdEngine<T>::direct_solve(const T* b, T* x, const PdParams& params, bool warm) {
    const int* skip = &reinterpret_cast<CtlBlock*>(ctl_.get())->outer_done;
    if (dense_) {
        dense_->set_stream(stream_);
        dense_->solve(b, x, params.dense_tol, params.dense_max_iter, skip, warm);
        return;
    }
    SBS_REQUIRE(direct_ != nullptr, "direct solve requested before factorisation");
    direct_->solve(b, x, stream_, skip);
}

template <class T> void PdEngine<T>::dense_info(int* bodies, int* max_body, long long* bytes, int* iterations,
                                               double* ratio) {
    SBS_REQUIRE(dense_ != nullptr, "dense solver not in use (set the DenseCholesky global solver)");
    if (bodies) *bodies = dense_->bodies();
    if (max_body) *max_body = dense_->max_body();
    if (bytes) *bytes = (long long)dense_->inverse_bytes();
    if (iterations || ratio) dense_->last_result(iterations, ratio);
}

// LDS of one local-step workgroup: corner forces, staged q, incidence list.
template <class T> size_t PdEngine<T>::local_lds_bytes() const {
    const size_t b = (12 * (size_t)kChunkTets + 3 * (size_t)chunk_max_verts_) * sizeof(T) +
                     (4 * (size_t)kChunkTets + (size_t)chunk_max_verts_) * sizeof(uint16_t);
    return (b + 15) & ~(size_t)15;
}

template <class T> void PdEngine<T>::launch_local(const T* q, bool residual_form, const int* skip) {
    const size_t lds = local_lds_bytes();
    SBS_REQUIRE(lds <= 64 * 1024, "local step: chunk vertex list exceeds LDS");
    tick(kKLocal, true);
    // block 0 also clears the reduction slots of the global step that follows
    hipLaunchKernelGGL((local_chunk_kernel<T, false>), dim3(nchunks_), dim3(kChunkTets), lds, stream_, nt_,
                       tet_lidx_.get(), chunk_vptr_.get(), chunk_verts_.get(), entry_pos_.get(), inc_codes_.get(),
                       inc_end_.get(), dminv_.get(), w_.get(), q, partial_f_.get(),
                       (residual_form ? kLocalResidual : 0) | (polar_ ? kLocalPolar : 0), skip,
                       slots_.get(), kNumSlots * kSlotSize, FoldIn<T>{}, PcgArgs<T>{}, 0);
    tick(kKLocal, false);
    SBS_HIP(hipGetLastError());
}

template <class T>
void PdEngine<T>::launch_local_fold(const T* x_int, const T* sn_in, const T* prev_in, T* sn_out, T* prev_out,
                                    T omega, int close_L, const PdParams& params) {
    const size_t lds = local_lds_bytes();
    SBS_REQUIRE(lds <= 64 * 1024, "local step: chunk vertex list exceeds LDS");
    SBS_REQUIRE(close_L >= 1, "local step: folded finish without a solve");
    FoldIn<T> f{x_int, chunk_rows_.get(), sn_in, prev_in, sn_out, prev_out, omega};
    PcgArgs<T> c{};  // what the stop test of the previous solve reads
    c.nv = nv_;
    c.slots = slots_.get();
    c.ctl = reinterpret_cast<CtlBlock*>(ctl_.get());
    c.max_iter = params.pcg_max_iter;
    c.tol2 = params.pcg_tol * params.pcg_tol;
    c.floor2 = params.pcg_floor * params.pcg_floor;
    tick(kKLocal, true);
    hipLaunchKernelGGL((local_chunk_kernel<T, true>), dim3(nchunks_), dim3(kChunkTets), lds, stream_, nt_,
                       tet_lidx_.get(), chunk_vptr_.get(), chunk_verts_.get(), entry_pos_.get(), inc_codes_.get(),
                       inc_end_.get(), dminv_.get(), w_.get(), (const T*)nullptr, partial_f_.get(),
                       polar_ ? kLocalPolar : 0, (const int*)nullptr, slots_.get(), kNumSlots * kSlotSize, f, c,
                       close_L);
    tick(kKLocal, false);
    SBS_HIP(hipGetLastError());
}

// One global solve: init, L iterations (A, B), the probe A(L) and the finish
// kernel.  L is the learned launch cap of this solve index when `exact` is
// false and a cap is known; otherwise iterations are issued in chunks and the
// host reads the device state after each chunk until the solve has stopped.
// Returns L.
template <class T>
int PdEngine<T>::launch_pcg(int gathered, const T* b, const T* x0, T* x, bool cheb, T omega, const PdParams& params,
                            const int* skip, int solve_index, bool exact, T* sn, T* prev, bool defer_finish) {
    PcgArgs<T> a{};
    a.nv = nv_;
    a.gathered = gathered;
    a.g = RhsGather<T>{c_.get(), sn_old_.get(), vent_ptr_.get(), partial_f_.get(), data_.DBC,
                       data_.moreDBC, data_.DBCX, pin_weight<T>(params.dt)};
    a.b = b;
    a.x0 = x0;
    a.x = x;
    a.slice_ptr = slice_ptr_.get();
    a.slice_w = slice_w_.get();
    a.ecol = ecol_.get();
    a.eval = eval_.get();
    a.diagA = diagA_.get();
    a.dinv = dinv_.get();
    for (int k = 0; k < 2; ++k) {
        a.rec[k] = cg_rec_[k].get();
    }
    a.p = cg_p_.get();
    a.sn = sn ? sn : sn_.get();
    a.prev = prev ? prev : prev_.get();
    a.omega = omega;
    a.slots = slots_.get();
    a.ctl = reinterpret_cast<CtlBlock*>(ctl_.get());
    a.max_iter = params.pcg_max_iter;
    a.tol2 = params.pcg_tol * params.pcg_tol;
    a.floor2 = params.pcg_floor * params.pcg_floor;
    a.skip = skip;
    const bool halo = params.pcg_halo && halo_groups_ > 0;
    if (halo) {
        a.h.groups = halo_groups_;
        a.h.grp = h_grp_.get();
        a.h.halo = h_halo_.get();
        a.h.halo_v = h_halo_v_.get();
        a.h.col = h_col_.get();
        a.h.val = h_val_.get();
        a.h.perm = h_perm_.get();
        a.h.vent = h_vent_.get();
        a.h.x = cg_x_.get();
        for (int k = 0; k < 2; ++k) a.hrec[k] = cg_plane_[k].get();
    }
    const dim3 grid(ceil_div(nv_, kTpb)), block(kTpb), hgrid(halo ? halo_groups_ : 1);
    const int max_iter = params.pcg_max_iter;
    tick(kKPcg, true);
    if (halo) {
        hipLaunchKernelGGL(pcg_init1_halo_kernel<T>, hgrid, block, 0, stream_, a);
        hipLaunchKernelGGL(pcg_init2_halo_kernel<T>, hgrid, block, 0, stream_, a);
    } else {
        hipLaunchKernelGGL(pcg_init1_kernel<T>, grid, block, 0, stream_, a);
        hipLaunchKernelGGL(pcg_init2_kernel<T>, grid, block, 0, stream_, a);
    }
    // iterations are timed as runs of back-to-back launches (an event pair per
    // launch would add its own gap to every launch)
    auto iteration = [&](int i) {
        if (halo) {
            if (i == 0) hipLaunchKernelGGL((pcg_iter_halo_kernel<T, true>), hgrid, block, 0, stream_, a, i);
            else hipLaunchKernelGGL((pcg_iter_halo_kernel<T, false>), hgrid, block, 0, stream_, a, i);
        } else {
            if (i == 0) hipLaunchKernelGGL((pcg_iter_kernel<T, true>), grid, block, 0, stream_, a, i);
            else hipLaunchKernelGGL((pcg_iter_kernel<T, false>), grid, block, 0, stream_, a, i);
        }
    };
    const bool learned = !exact && solve_index >= 0 && solve_index < (int)cap_.size();
    int L = 0;
    if (learned && cap_[solve_index] > 0) {
        L = std::min(cap_[solve_index], max_iter);
        tick(kKPcgIter, true);
        for (int i = 0; i < L; ++i) iteration(i);
        tick(kKPcgIter, false, L);
    } else {
        const CtlBlock* hc = reinterpret_cast<const CtlBlock*>(host_ctl_);
        while (L < max_iter) {
            const int n = std::min(kCheckChunk, max_iter - L);
            tick(kKPcgIter, true);
            for (int j = 0; j < n; ++j) iteration(L++);
            tick(kKPcgIter, false, n);
            SBS_HIP(hipMemcpyAsync(host_ctl_, a.ctl, sizeof(CtlBlock), hipMemcpyDeviceToHost, stream_));
            SBS_HIP(hipStreamSynchronize(stream_));
            // st[(L-1)&1] is the state launch L-1 saw: done there means no later iteration ran
            if ((skip && hc->outer_done) || hc->st[(L - 1) & 1].done) break;
        }
        if (learned) cap_[solve_index] = std::min(max_iter, L + cap_margin(L, 0, 0));
    }
    if (defer_finish) SBS_REQUIRE(halo && cheb, "PCG: only the halo Chebyshev finish can be deferred");
    else if (halo) hipLaunchKernelGGL(pcg_finish_halo_kernel<T>, grid, block, 0, stream_, a, L, (int)cheb);
    else hipLaunchKernelGGL(pcg_finish_kernel<T>, grid, block, 0, stream_, a, L, (int)cheb);
    tick(kKPcg, false);
    SBS_HIP(hipGetLastError());
    return L;
}

// Iterations of headroom over the last observed count of a solve.
// Launches beyond the larger of the last two observed counts.  A count seen
// unchanged three times running gets one (the launch after the last
// iteration sees the solve converged) plus one per 16 iterations; otherwise
// 2 + used/8 plus the last change of the count (up to 3), since counts move
// during transients and the caps follow them with a lag of two steps.
template <class T> int PdEngine<T>::cap_margin(int used, int moved, int steady) {
    return steady >= 2 ? 1 + used / 16 : 2 + used / 8 + std::min(moved, 3);
}

template <class T> void PdEngine<T>::fold_slot(int slot) {
    if (!step_pending_[slot]) return;
    SBS_HIP(hipEventSynchronize(step_event_[slot]));
    step_pending_[slot] = false;
    const int* c = host_counters_ + kCounterInts * slot;
    last_pcg_iters_ = c[0];
    unconverged_total_ += c[1];
    pcg_iters_total_ += c[0];
    pcg_solves_total_ += c[3];
    // learned launch caps: follow the larger of the last two observed counts of
    // each solve index; a solve cut by its cap raises that cap by half
    if (step_caps_[slot] != cap_epoch_) return;  // counters of other settings
    const int* log = c + 6;
    const int nsol = std::min(std::min(c[3], kMaxSolveLog), (int)cap_.size());
    for (int i = 0; i < nsol; ++i) {
        const int v = log[i];
        if (v < 0) {
            const int its = -v - 1;
            if (its < cap_max_iter_) cap_[i] = std::min(cap_max_iter_, std::max(cap_[i] + cap_[i] / 2, its + 4));
            seen_[i] = std::max(seen_[i], its);
            moved_[i] = 3;
            steady_[i] = 0;
        } else {
            moved_[i] = seen_[i] > 0 ? std::abs(v - seen_[i]) : 0;
            steady_[i] = (seen_[i] > 0 && v == seen_[i]) ? steady_[i] + 1 : 0;
            const int used = std::max(v, seen_[i]);
            cap_[i] = std::min(cap_max_iter_, used + cap_margin(used, moved_[i], steady_[i]));
            seen_[i] = v;
        }
    }
}

template <class T> void PdEngine<T>::sync_counters() {
    fold_slot((int)(steps_ & 1));        // older step first
    fold_slot((int)((steps_ + 1) & 1));
}

template <class T> void PdEngine<T>::update(PdData<T>& data, const PdParams& params) {
    SBS_REQUIRE(params.polar_rotation == 0 || params.polar_rotation == 1, "PdSolver: polar_rotation must be 0 or 1");
    polar_ = params.polar_rotation != 0;
    if (!ready_) prepare(data, params);
    data_ = data;
    SBS_REQUIRE(params.num_iterations >= 0, "PdSolver: numIterations must be >= 0");
    SBS_REQUIRE(params.pcg_max_iter >= 1, "PdSolver: pcg_max_iter must be >= 1");
    const bool direct = is_direct();
    if (direct && ((solver_ == GlobalSolver::DenseCholesky) ? (!dense_ || dense_prec_ != params.dense_precision)
                                                             : !direct_)) {
        direct_.reset();
        dense_.reset();
        direct_factor(params);
    }
    if (solver_ == GlobalSolver::BatchedPCG) {
        SBS_REQUIRE(data.moreDBC == nullptr, "PdSolver: the batched solver does not take moreDBC weights");
        if (!batch_) build_batch(params);
    }
    // this step's pinned counter slot was last written two steps ago: fold it in
    const int slot = (int)(steps_ & 1);
    fold_slot(slot);
    const bool pcg = solver_ == GlobalSolver::PCG;
    if (pcg) reset_caps_if_changed(params);
    step_caps_[slot] = cap_epoch_;
    // A step whose launch sequence is known in advance (no host reads of the
    // device state, no timing events) is captured once into a HIP graph and
    // replayed while the sequence, its pointers and its parameters stay the
    // same; the learned PCG caps are part of that key.
    bool caps_known = true;
    if (pcg)
        for (int i = 0; i < params.num_iterations; ++i) caps_known = caps_known && cap_[i] > 0;
    const bool graph = params.use_graph && stream_ != nullptr && !perf_ && timing_mask_ == 0 && !params.handle_collision &&
                       (solver_ == GlobalSolver::Jacobi || solver_ == GlobalSolver::BatchedPCG ||
                        (pcg && params.pcg_adaptive && caps_known));
    if (graph) {
        const std::vector<uint64_t> key = step_key(data, params);
        if (!graph_exec_ || key != graph_key_ || graph_caps_stale(pcg ? params.num_iterations : 0)) {
            drop_graph();
            // the graph keeps its own caps, one launch above the learned ones,
            // so that counts moving by a few launches replay the same graph
            std::vector<int> learned;
            if (pcg) {
                learned = cap_;
                graph_caps_.assign(cap_.begin(), cap_.end());
                for (int& c : graph_caps_) c = std::min(cap_max_iter_, c + 1);
                cap_ = graph_caps_;
            }
            hipGraph_t gr = nullptr;
            SBS_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
            try {
                enqueue_step(data, params);
            } catch (...) {
                (void)hipStreamEndCapture(stream_, &gr);
                if (gr) (void)hipGraphDestroy(gr);
                if (pcg) cap_ = learned;
                throw;
            }
            if (pcg) cap_ = learned;
            SBS_HIP(hipStreamEndCapture(stream_, &gr));
            const hipError_t e = hipGraphInstantiate(&graph_exec_, gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            if (e != hipSuccess) {
                graph_exec_ = nullptr;
                throw Error(kHipError, std::string("PdSolver: graph instantiation failed: ") + hipGetErrorString(e));
            }
            graph_key_ = key;
            ++graph_captures_;
        }
        SBS_HIP(hipGraphLaunch(graph_exec_, stream_));
        ++graph_replays_;
    } else {
        enqueue_step(data, params);
    }
    // counters of the step (PCG iterations, unconverged solves, max per solve,
    // solves, last, per-solve log) land in this step's pinned slot
    CtlBlock* ctl = reinterpret_cast<CtlBlock*>(ctl_.get());
    SBS_HIP(hipMemcpyAsync(host_counters_ + kCounterInts * slot, &ctl->pcg_iters, kCounterInts * sizeof(int),
                           hipMemcpyDeviceToHost, stream_));
    SBS_HIP(hipEventRecord(step_event_[slot], stream_));
    step_pending_[slot] = true;
    ++steps_;
    if (perf_) SBS_HIP(hipStreamSynchronize(stream_));
    last_iters_ = params.num_iterations;
}

// Everything a captured step depends on: the caller's arrays, the parameters,
// the global solver and the learned launch caps.
template <class T> std::vector<uint64_t> PdEngine<T>::step_key(const PdData<T>& d, const PdParams& p) const {
    auto bits = [](double v) {
        uint64_t u;
        std::memcpy(&u, &v, sizeof u);
        return u;
    };
    std::vector<uint64_t> k = {(uint64_t)(uintptr_t)d.X, (uint64_t)(uintptr_t)d.V, (uint64_t)(uintptr_t)d.XTilde,
                               (uint64_t)(uintptr_t)d.DBCX, (uint64_t)(uintptr_t)d.mass, (uint64_t)(uintptr_t)d.DBC,
                               (uint64_t)(uintptr_t)d.moreDBC, (uint64_t)(uintptr_t)d.fixed,
                               (uint64_t)(uintptr_t)stream_, (uint64_t)solver_, (uint64_t)p.num_iterations,
                               bits(p.dt), bits(p.gravity), bits(p.tol), bits(p.rho), (uint64_t)p.pcg_max_iter,
                               bits(p.pcg_tol), bits(p.pcg_floor), (uint64_t)p.pcg_chebyshev, (uint64_t)p.pcg_halo,
                               bits(p.mu_n), bits(p.mu_t), (uint64_t)p.polar_rotation};
    k.push_back((uint64_t)cap_epoch_);
    return k;
}

// A captured PCG step is replayed while every solve's last observed count
// leaves it a spare launch under the graph's cap (count + 1 launches see the
// solve converge); it is re-captured when a count comes closer than that (or
// a solve was cut, which raises its count to the cap), or when the graph's
// caps have carried more than max(2, 1/16 of the learned total) launches
// above the learned caps for 16 steps running.
template <class T> bool PdEngine<T>::graph_caps_stale(int n) {
    if (n == 0) return false;
    if ((int)graph_caps_.size() != n) return true;
    int surplus = 0, learned = 0;
    for (int i = 0; i < n; ++i) {
        if (seen_[i] + 2 > graph_caps_[i] && graph_caps_[i] < cap_max_iter_) return true;
        surplus += std::max(graph_caps_[i] - cap_[i], 0);
        learned += cap_[i];
    }
    graph_slack_steps_ = surplus > std::max(2, learned / 16) ? graph_slack_steps_ + 1 : 0;
    return graph_slack_steps_ >= 16;
}

template <class T> void PdEngine<T>::drop_graph() {
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    graph_exec_ = nullptr;
    graph_key_.clear();
    graph_caps_.clear();
    graph_slack_steps_ = 0;
}

// The launches of one step (begin, the local/global iterations, end, the
// collision and fixed-body responses), on stream_.
template <class T> void PdEngine<T>::enqueue_step(PdData<T>& data, const PdParams& params) {
    const T dt = (T)params.dt;
    const unsigned vb = ceil_div(nv_, kTpb);
    CtlBlock* ctl = reinterpret_cast<CtlBlock*>(ctl_.get());
    const bool jacobi = solver_ == GlobalSolver::Jacobi;
    const bool pcg = solver_ == GlobalSolver::PCG;
    const bool cheb = jacobi || (pcg && params.pcg_chebyshev);
    const RhsGather<T> g{c_.get(), sn_old_.get(), vent_ptr_.get(), partial_f_.get(), data.DBC,
                         data.moreDBC, data.DBCX, pin_weight<T>(params.dt)};
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    if (perf_) for (auto& e : ev) SBS_HIP(hipEventCreate(&e));

    // PCG + Chebyshev on the halo kernels: the finish of solve i (its stop test
    // and the Chebyshev recurrence) runs inside the local step of iteration
    // i+1, which forms q from solve i's iterate; sn/prev then alternate
    // between two buffers.  The last solve keeps its finish launch.
    const bool fold = pcg && cheb && params.pcg_halo && halo_groups_ > 0 && chunk_rows_.get() != nullptr;
    hipLaunchKernelGGL(ctl_init_kernel, dim3(1), dim3(64), 0, stream_, ctl);
    if (solver_ == GlobalSolver::BatchedPCG) {
        // every body's whole step in one persistent launch per size class
        BatchParams bp;
        bp.iterations = params.num_iterations;
        bp.dt = params.dt;
        bp.gravity = params.gravity;
        bp.rho = params.rho;
        bp.pcg_tol = params.pcg_tol;
        bp.pcg_floor = params.pcg_floor;
        bp.pcg_max_iter = params.pcg_max_iter;
        bp.write_x = params.handle_collision == 0;
        bp.polar = params.polar_rotation != 0;
        if (perf_) SBS_HIP(hipEventRecord(ev[0], stream_));
        batch_->step(data.X, data.V, data.XTilde, data.DBCX, bp, stream_);
        if (perf_) {
            SBS_HIP(hipEventRecord(ev[1], stream_));
            SBS_HIP(hipEventSynchronize(ev[1]));
            float ms = 0;
            SBS_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
            perf_data_[1].ms += ms;  // local and global steps run fused, counted as the global step
        }
    } else {
        hipLaunchKernelGGL(begin_step_kernel<T>, dim3(vb), dim3(kTpb), 0, stream_, nv_, data.X, data.V, data.mass,
                           data.DBC, data.moreDBC, (const T*)data.DBCX, dt, (T)params.gravity, c_.get(), sn_.get(),
                           sn_old_.get(), prev_.get(),
                           (int)cheb, fold ? sn_alt_.get() : (T*)nullptr, fold ? prev_alt_.get() : (T*)nullptr);
        SBS_HIP(hipGetLastError());
        T omega = 1;
        const T rho = (T)params.rho;
        const int* skip = cheb ? nullptr : &ctl->outer_done;
        T* sn_cur = sn_.get();
        T* prev_cur = prev_.get();
        T* sn_nxt = sn_alt_.get();
        T* prev_nxt = prev_alt_.get();
        int open_L = 0;  // launches of the solve whose finish is pending (0: none)
        for (int i = 0; i < params.num_iterations; ++i) {
            if (perf_) SBS_HIP(hipEventRecord(ev[0], stream_));
            if (open_L > 0) {
                launch_local_fold(cg_x_.get(), sn_cur, prev_cur, sn_nxt, prev_nxt, omega, open_L, params);
                std::swap(sn_cur, sn_nxt);
                std::swap(prev_cur, prev_nxt);
                open_L = 0;
            } else {
                launch_local(sn_cur, jacobi, skip);
            }
            if (perf_) SBS_HIP(hipEventRecord(ev[1], stream_));
            if (cheb) {
                if (i <= 10) omega = 1;
                else if (i == 11) omega = (T)2 / ((T)2 - rho * rho);
                else omega = (T)4 / ((T)4 - rho * rho * omega);
            }
            if (jacobi) {
                tick(kKJacobi, true);
                hipLaunchKernelGGL(jacobi_chebyshev_kernel<T>, dim3(vb), dim3(kTpb), 0, stream_, nv_, g, diagL_.get(),
                                   sn_.get(), prev_.get(), omega);
                tick(kKJacobi, false);
            } else if (pcg) {
                // PCG warm-started from the current iterate, then the Chebyshev
                // recurrence, or computeError and the early exit
                const bool defer = fold && i + 1 < params.num_iterations;
                const int L = launch_pcg(1, nullptr, sn_cur, next_.get(), cheb, omega, params, skip, i,
                                         !params.pcg_adaptive, sn_cur, prev_cur, defer);
                if (defer) open_L = L;
                if (!cheb) {
                    hipLaunchKernelGGL(outer_error_kernel<T>, dim3(vb), dim3(kTpb), 0, stream_, nv_,
                                       (const T*)next_.get(), sn_.get(), prev_.get(), slots_.get(), skip);
                    hipLaunchKernelGGL(outer_decide_kernel, dim3(1), dim3(64), 0, stream_, (const double*)slots_.get(),
                                       nv_, params.tol, ctl, skip);
                }
            } else {
                tick(kKRhs, true);
                hipLaunchKernelGGL(rhs_kernel<T>, dim3(vb), dim3(kTpb), 0, stream_, nv_, g, rhs_.get(), skip);
                tick(kKRhs, false);
                tick(kKDirect, true);
                direct_solve(rhs_.get(), sn_.get(), params, true);  // sn holds the previous iterate
                tick(kKDirect, false);
                hipLaunchKernelGGL(outer_error_kernel<T>, dim3(vb), dim3(kTpb), 0, stream_, nv_, (const T*)sn_.get(),
                                   (T*)nullptr, prev_.get(), slots_.get(), skip);
                hipLaunchKernelGGL(outer_decide_kernel, dim3(1), dim3(64), 0, stream_, (const double*)slots_.get(), nv_,
                                   params.tol, ctl, skip);
            }
            SBS_HIP(hipGetLastError());
            if (perf_) {
                SBS_HIP(hipEventRecord(ev[2], stream_));
                SBS_HIP(hipEventSynchronize(ev[2]));
                float a = 0, gm = 0;
                SBS_HIP(hipEventElapsedTime(&a, ev[0], ev[1]));
                SBS_HIP(hipEventElapsedTime(&gm, ev[1], ev[2]));
                perf_data_[0].ms += a;
                perf_data_[1].ms += gm;
            }
        }
        hipLaunchKernelGGL(end_step_kernel<T>, dim3(vb), dim3(kTpb), 0, stream_, nv_, (const T*)sn_cur, (T)1 / dt,
                           data.XTilde, data.V, params.handle_collision ? (T*)nullptr : data.X, data.moreDBC);
        SBS_HIP(hipGetLastError());
    }
    const bool collide = params.handle_collision != 0;
    if (collide) {
        // DetectCollision between bodies on the step's motion X -> XTilde, then
        // CCDKernel (pdSolver.cu:218-224)
        SBS_REQUIRE(data.num_tris >= 0 && (data.num_tris == 0 || (data.Tri && data.TriFathers)),
                    "PdSolver: handleCollision needs Tri and TriFathers");
        SBS_REQUIRE(data.tIs && data.Normals, "PdSolver: handleCollision needs tIs and Normals");
        if (!collision_)
            collision_ = std::make_unique<CollisionDetector<T>>(nv_, data.num_tris, data.Tri, data.TriFathers, stream_);
        collision_->set_stream(stream_);
        if (perf_) SBS_HIP(hipEventRecord(ev[0], stream_));
        collision_->detect(data.X, data.XTilde, true, data.tIs, data.Normals);
        ccd_response<T>(nv_, data.X, data.XTilde, data.V, data.tIs, data.Normals, stream_);
        if (perf_) {
            SBS_HIP(hipEventRecord(ev[1], stream_));
            SBS_HIP(hipEventSynchronize(ev[1]));
            float ms = 0;
            SBS_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
            perf_data_[3].ms += ms;
        }
    }
    // FixedBodyData::HandleCollisions on XTilde and V, after X was taken
    // (pdSolver.cu:228-231)
    if (data.fixed) {
        if (perf_) SBS_HIP(hipEventRecord(ev[0], stream_));
        data.fixed->handle(data.XTilde, data.V, nv_, (T)params.mu_t, (T)params.mu_n, stream_);
        if (perf_) {
            SBS_HIP(hipEventRecord(ev[1], stream_));
            SBS_HIP(hipEventSynchronize(ev[1]));
            float ms = 0;
            SBS_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
            perf_data_[2].ms += ms;
        }
    }
    if (perf_) {
        SBS_HIP(hipStreamSynchronize(stream_));
        for (auto& e : ev) (void)hipEventDestroy(e);
    }
}

// Learned launch caps are per solve index and valid for one set of PCG settings.
template <class T> void PdEngine<T>::reset_caps_if_changed(const PdParams& p) {
    const bool same = (int)cap_.size() == p.num_iterations && cap_max_iter_ == p.pcg_max_iter &&
                      cap_tol_ == p.pcg_tol && cap_floor_ == p.pcg_floor && cap_cheb_ == p.pcg_chebyshev &&
                      cap_dt_ == p.dt;
    if (same) return;
    cap_.assign(p.num_iterations, 0);
    seen_.assign(p.num_iterations, 0);
    moved_.assign(p.num_iterations, 0);
    steady_.assign(p.num_iterations, 0);
    cap_max_iter_ = p.pcg_max_iter;
    cap_tol_ = p.pcg_tol;
    cap_floor_ = p.pcg_floor;
    cap_cheb_ = p.pcg_chebyshev;
    cap_dt_ = p.dt;
    ++cap_epoch_;
}

template <class T> int PdEngine<T>::pcg_cap() const {
    int m = 0;
    for (int c : cap_) m = std::max(m, c);
    return m;
}

template <class T> void PdEngine<T>::local_only(const T* q, const T* sn_old, bool residual_form, double dt, T* b_out) {
    SBS_REQUIRE(ready_, "local_only: engine not prepared");
    const unsigned vb = ceil_div(nv_, kTpb);
    launch_local(q, residual_form, nullptr);
    const RhsGather<T> g{c_.get(), sn_old, vent_ptr_.get(), partial_f_.get(), data_.DBC, data_.moreDBC,
                         data_.DBCX, pin_weight<T>(dt)};
    hipLaunchKernelGGL(rhs_kernel<T>, dim3(vb), dim3(kTpb), 0, stream_, nv_, g, b_out, (const int*)nullptr);
    SBS_HIP(hipGetLastError());
}

template <class T> int PdEngine<T>::solve(const T* b, T* x, const PdParams& params, bool use_guess) {
    SBS_REQUIRE(ready_, "solve: engine not prepared");
    SBS_REQUIRE(params.pcg_max_iter >= 1, "solve: pcg_max_iter must be >= 1");
    sync_counters();
    CtlBlock* ctl = reinterpret_cast<CtlBlock*>(ctl_.get());
    hipLaunchKernelGGL(ctl_init_kernel, dim3(1), dim3(64), 0, stream_, ctl);
    if (direct_ || dense_) {
        direct_solve(b, x, params);
        SBS_HIP(hipStreamSynchronize(stream_));
        int it = 0;
        if (dense_) dense_->last_result(&it, nullptr);
        return it;
    }
    // x0 = the caller's guess (copied: x is written while x0 is still read) or 0
    SBS_REQUIRE(!use_guess || x != nullptr, "solve: guess requested without x");
    if (use_guess) {
        SBS_HIP(hipMemcpyAsync(next_.get(), x, 3 * (size_t)nv_ * sizeof(T), hipMemcpyDeviceToDevice, stream_));
    }
    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(256), 0, stream_, slots_.get(), kNumSlots * kSlotSize);
    launch_pcg(0, b, use_guess ? next_.get() : nullptr, x, false, (T)1, params, nullptr, -1, true);
    SBS_HIP(hipMemcpyAsync(host_ctl_, ctl, sizeof(CtlBlock), hipMemcpyDeviceToHost, stream_));
    SBS_HIP(hipStreamSynchronize(stream_));
    const CtlBlock* hc = reinterpret_cast<const CtlBlock*>(host_ctl_);
    unconverged_total_ += hc->unconverged;
    pcg_iters_total_ += hc->pcg_iters;
    pcg_solves_total_ += hc->solves;
    return hc->last_iters;
}

template <class T> typename PdEngine<T>::Export PdEngine<T>::export_system() const {
    Export ex;
    auto d2h = [&](std::vector<T>& dst, const DevBuf<T>& src, size_t n) {
        dst.resize(n);
        if (n) SBS_HIP(hipMemcpy(dst.data(), src.get(), n * sizeof(T), hipMemcpyDeviceToHost));
    };
    SBS_HIP(hipStreamSynchronize(stream_));
    std::vector<T> dm, v0, w;
    d2h(dm, dminv_, 9 * (size_t)nt_);
    d2h(v0, v0_, nt_);
    d2h(w, w_, nt_);
    // internal (Morton) order -> caller order, Dm^-1 row-major per tet
    ex.V0.resize(nt_);
    ex.w.resize(nt_);
    ex.dminv.resize(9 * (size_t)nt_);
    for (int i = 0; i < nt_; ++i) {
        const int t = tet_perm_[i];
        ex.V0[t] = v0[i];
        ex.w[t] = w[i];
        for (int k = 0; k < 9; ++k) ex.dminv[9 * (size_t)t + k] = dm[(size_t)k * nt_ + i];
    }
    if (!system_built_) return ex;
    d2h(ex.c, c_, nv_);
    d2h(ex.diagA, diagA_, nv_);
    d2h(ex.diagL, diagL_, nv_);
    std::vector<int> sptr(nslices_ + 1), rl(nv_);
    SBS_HIP(hipMemcpy(sptr.data(), slice_ptr_.get(), sptr.size() * sizeof(int), hipMemcpyDeviceToHost));
    SBS_HIP(hipMemcpy(rl.data(), rowlen_.get(), rl.size() * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<uint32_t> ec(ecol_.n);
    std::vector<T> evv(eval_.n);
    SBS_HIP(hipMemcpy(ec.data(), ecol_.get(), ec.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    SBS_HIP(hipMemcpy(evv.data(), eval_.get(), evv.size() * sizeof(T), hipMemcpyDeviceToHost));
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
            row.emplace_back((int)ec[pos], evv[pos]);
        }
        std::sort(row.begin(), row.end(), [](auto& a, auto& b) { return a.first < b.first; });
        for (size_t k = 0; k < row.size(); ++k) {
            ex.col[ex.rowptr[v] + k] = row[k].first;
            ex.val[ex.rowptr[v] + k] = row[k].second;
        }
    }
    return ex;
}

template <class T> void PdEngine<T>::set_kernel_timing(unsigned mask) {
    flush_timing();
    for (int k = 0; k < kKCount; ++k) ktime_[k] = KernelTime{};
    timing_mask_ = mask;
}

// Records the start (or stop) event of kernel `id` when it is being timed; a
// stop event closes `launches` back-to-back launches of that kernel.
template <class T> void PdEngine<T>::tick(int id, bool start, int launches) {
    if (!(timing_mask_ & (1u << id))) return;
    auto& pool = tev_[id];
    size_t& used = tev_used_[id];
    if (start) {
        if (used == pool.size()) {
            hipEvent_t a, b;
            SBS_HIP(hipEventCreate(&a));
            SBS_HIP(hipEventCreate(&b));
            pool.emplace_back(a, b);
            tev_n_[id].push_back(0);
        }
        SBS_HIP(hipEventRecord(pool[used].first, stream_));
    } else {
        SBS_HIP(hipEventRecord(pool[used].second, stream_));
        tev_n_[id][used] = launches;
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
            ktime_[k].launches += tev_n_[k][i];
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
