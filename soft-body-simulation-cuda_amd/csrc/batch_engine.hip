// Batched PD step: one workgroup per body runs the whole step; see
// batch_engine.hpp.  Per body and outer iteration i (pdSolver.cu:164-205 with
// the Jacobi-PCG global solve, as the PCG global step of pd_engine):
//   computeLocal (pdUtil.cu:97-145) per tet, its 4 corner forces into `force`;
//   b = c sn_old + the forces of the vertex's corners, pinned rows
//   overwritten with DBCX 1e6/h^2 (pdUtil.cu:147-178);
//   Jacobi-PCG warm-started from sn, per component stopping when
//   ||D^-1 r_k|| <= max(tol ||D^-1 r_0||, floor ||D^-1 b||);
//   the Chebyshev recurrence (pdSolver.cu:196-198, pdUtil.cu:216-226).
// Reductions are block reductions in a fixed order (DPP inside a wavefront,
// then the wavefront totals in index order), so a body's result does not
// depend on the other bodies or on scheduling.
#include "batch_engine.hpp"

#include <algorithm>
#include <numeric>

#include "svd3.hpp"

namespace sbs {
namespace {

constexpr int kMaxWaves = 16;  // 1024-thread workgroups at most

template <class T> struct BatchArgs {
    const int* body_list;
    const int* vptr;
    const int* tptr;
    const int* eptr;
    const int* ew;
    const uint32_t* perm;
    const uint4* tets;
    const T* dminv;
    const T* w;
    const int* inc_ptr;
    const uint32_t* inc;
    const uint32_t* ecol;
    const T* eval;
    const T* diagA;
    const T* c0;
    const T* mass;
    const T* dbc;
    const T* origin;  // 3 per body: the body's local origin
    T* X;
    T* V;
    T* XT;
    const T* DBCX;
    T* sn;
    T* sn_old;
    T* prev;
    T* x;
    T* r;
    T* p;
    T* q;
    T* force;
    int* stats;
    int iters;
    int max_iter;
    int write_x;
    int polar;  // rotation by Newton's polar iteration (FP32)
    T dt;
    T gravity;
    T rho;
    T wdbc;
    T tol2;
    T floor2;
};

// The Chebyshev step's damping in double (the reference's 0.9 is a double
// literal), then (damped - prev) omega + prev.
template <class T> __device__ __forceinline__ T cheb(T x, T s, T prev, T omega) {
    const T damped = (T)fma(0.9, (double)(x - s), (double)s);
    return fma(damped - prev, omega, prev);
}

// Block sum of NQ doubles returned to every thread: DPP wavefront totals,
// then the wavefront totals added in index order.  `buf` alternates between
// two LDS buffers from one reduction to the next, so one barrier suffices.
template <int NQ> __device__ __forceinline__ void block_sum(double (&v)[NQ], double* buf) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int k = 0; k < NQ; ++k) v[k] = wave_total(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NQ; ++k) buf[wv * NQ + k] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        double s = 0.0;
        for (int j = 0; j < nw; ++j) s += buf[j * NQ + k];
        v[k] = s;
    }
}

// Entries j0 .. j0 + NB - 1 of row il of a body's ELL matrix
// (positions past the row's width repeat its last entry; the caller skips
// them), requested together so their latencies overlap.
template <class T> constexpr int kRowBatch = sizeof(T) == 4 ? 8 : 4;  // FP64: within the 128 registers
template <class T, int NB = kRowBatch<T>>
__device__ __forceinline__ void row_entries(const BatchArgs<T>& a, size_t ebase, int nrow, int il, int width, int j0,
                                            uint32_t (&uu)[NB], T (&vv)[NB]) {
#pragma unroll
    for (int jj = 0; jj < NB; ++jj) {
        const int j = j0 + jj < width ? j0 + jj : width - 1;
        const size_t e = ebase + (size_t)j * nrow + il;
        uu[jj] = a.ecol[e];
        vv[jj] = a.eval[e];
    }
}

// kLdsP: the PCG search direction of the body lives in LDS (3 values per
// vertex of the class's largest body), so the product's neighbour reads are
// LDS reads instead of L2 round trips; the values are the same.
template <class T, bool kLdsP>
__global__ void __launch_bounds__(1024) batch_step_kernel(BatchArgs<T> a, int first) {
    __shared__ double red[2][kMaxWaves * 9];
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    int rb = 0;
    const int b = a.body_list[first + blockIdx.x];
    const int v0 = a.vptr[b], v1 = a.vptr[b + 1], t0 = a.tptr[b], t1 = a.tptr[b + 1];
    T* const lp = reinterpret_cast<T*>(dyn_lds);
    const size_t pbase = 3 * (size_t)v0;
    // entry j (3 per vertex, global numbering) of the search direction
    auto P = [&](size_t j) -> T& { return kLdsP ? lp[j - pbase] : a.p[j]; };
    const int nrow = v1 - v0, width = a.ew[b];
    const size_t ebase = (size_t)a.eptr[b];
    const int B = blockDim.x, tid = threadIdx.x;
    const T dt = a.dt;
    // Every vector of the body is held relative to its origin o (the step is
    // translation invariant: L o = 0, so A (y + o) = b becomes A y = b - c o),
    // so FP32 rounds offsets of the body's size, not of its distance from the
    // world origin.
    const T o[3] = {a.origin[3 * b], a.origin[3 * b + 1], a.origin[3 * b + 2]};

    // prologue (pdSolver.cu:154-162): s_n = x + h v + f_ext / c, sn_old, prev
    for (int i = v0 + tid; i < v1; i += B) {
        const uint32_t g = a.perm[i];
        const T cv = a.c0[i];
        const T fy = -a.gravity * a.mass[i];
        const T inv = (T)1 / cv;
        const T s0 = (a.X[3 * (size_t)g] - o[0]) + dt * a.V[3 * (size_t)g];
        const T s1 = (a.X[3 * (size_t)g + 1] - o[1]) + dt * a.V[3 * (size_t)g + 1] + inv * fy;
        const T s2 = (a.X[3 * (size_t)g + 2] - o[2]) + dt * a.V[3 * (size_t)g + 2];
        T* sn = a.sn + 3 * (size_t)i;
        T* so = a.sn_old + 3 * (size_t)i;
        T* pv = a.prev + 3 * (size_t)i;
        sn[0] = s0; sn[1] = s1; sn[2] = s2;
        so[0] = s0; so[1] = s1; so[2] = s2;
        pv[0] = s0; pv[1] = s1; pv[2] = s2;
    }
    __syncthreads();
    int it_total = 0, solves = 0, max_solve = 0, cut = 0;
    T omega = 1;
    for (int i_outer = 0; i_outer < a.iters; ++i_outer) {
        // computeLocal (R form, as the direct and PCG global steps use it)
        for (int t = t0 + tid; t < t1; t += B) {
            const uint4 tv = a.tets[t];
            const T* m = a.dminv + 9 * (size_t)t;
            T mm[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) mm[k] = m[k];
            const T* q0 = a.sn + 3 * (size_t)tv.x;
            const T* q1 = a.sn + 3 * (size_t)tv.y;
            const T* q2 = a.sn + 3 * (size_t)tv.z;
            const T* q3 = a.sn + 3 * (size_t)tv.w;
            T F[3][3];
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                const T x0 = q0[r];
                const T d0 = q1[r] - x0, d1 = q2[r] - x0, d2 = q3[r] - x0;
#pragma unroll
                for (int k = 0; k < 3; ++k) F[r][k] = d0 * mm[k] + d1 * mm[3 + k] + d2 * mm[6 + k];
            }
            T P[3][3];
            pd_projection(F, false, P, a.polar != 0);
            const T ww = a.w[t];
            T* f = a.force + 12 * (size_t)t;
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                const T f1 = ww * (P[r][0] * mm[0] + P[r][1] * mm[1] + P[r][2] * mm[2]);
                const T f2 = ww * (P[r][0] * mm[3] + P[r][1] * mm[4] + P[r][2] * mm[5]);
                const T f3 = ww * (P[r][0] * mm[6] + P[r][1] * mm[7] + P[r][2] * mm[8]);
                f[r] = -(f1 + f2 + f3);
                f[3 + r] = f1;
                f[6 + r] = f2;
                f[9 + r] = f3;
            }
        }
        __syncthreads();
        // right-hand side and the PCG start: x = sn, r = b - A x, p = z = D^-1 r
        double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // r.z, z.z, (D^-1 b)^2 per component
        for (int i = v0 + tid; i < v1; i += B) {
            T bv[3];
            const T cv = a.c0[i];
#pragma unroll
            for (int k = 0; k < 3; ++k) bv[k] = cv * a.sn_old[3 * (size_t)i + k];
            // the vertex's corner forces in incidence order, kRowBatch requested together
            const int e0 = a.inc_ptr[i], e1 = a.inc_ptr[i + 1];
            for (int eb = e0; eb < e1; eb += kRowBatch<T>) {
                T fv[kRowBatch<T>][3];
#pragma unroll
                for (int jj = 0; jj < kRowBatch<T>; ++jj) {
                    const uint32_t code = a.inc[eb + jj < e1 ? eb + jj : e1 - 1];
                    const T* f = a.force + 12 * (size_t)(code >> 2) + 3 * (code & 3u);
#pragma unroll
                    for (int k = 0; k < 3; ++k) fv[jj][k] = f[k];
                }
#pragma unroll
                for (int jj = 0; jj < kRowBatch<T>; ++jj)
                    if (eb + jj < e1)
#pragma unroll
                        for (int k = 0; k < 3; ++k) bv[k] += fv[jj][k];
            }
            if (a.dbc[i] > (T)0) {
                // DBCX wdbc - c o (translation of b = DBCX wdbc by A o = c o):
                // (DBCX - o) wdbc - (c - wdbc) o, where both differences are
                // exact (nearby values), so nothing cancels
                const uint32_t g = a.perm[i];
                const T cw = cv - a.wdbc;
#pragma unroll
                for (int k = 0; k < 3; ++k) bv[k] = (a.DBCX[3 * (size_t)g + k] - o[k]) * a.wdbc - cw * o[k];
            }
            const T dA = a.diagA[i], di = (T)1 / dA;
            T Ax[3];
            const T* xs = a.sn + 3 * (size_t)i;
#pragma unroll
            for (int k = 0; k < 3; ++k) Ax[k] = dA * xs[k];
            const int il = i - v0;
            for (int j0 = 0; j0 < width; j0 += kRowBatch<T>) {
                uint32_t uu[kRowBatch<T>];
                T vv[kRowBatch<T>];
                row_entries(a, ebase, nrow, il, width, j0, uu, vv);
                T xu[kRowBatch<T>][3];
#pragma unroll
                for (int jj = 0; jj < kRowBatch<T>; ++jj)
#pragma unroll
                    for (int k = 0; k < 3; ++k) xu[jj][k] = a.sn[3 * (size_t)uu[jj] + k];
#pragma unroll
                for (int jj = 0; jj < kRowBatch<T>; ++jj)
                    if (j0 + jj < width)
#pragma unroll
                        for (int k = 0; k < 3; ++k) Ax[k] += vv[jj] * xu[jj][k];
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const T rk = bv[k] - Ax[k];
                const T zk = di * rk;
                const T dbk = di * bv[k];
                a.x[3 * (size_t)i + k] = xs[k];
                a.r[3 * (size_t)i + k] = rk;
                P(3 * (size_t)i + k) = zk;
                acc[k] += (double)(rk * zk);
                acc[3 + k] += (double)(zk * zk);
                acc[6 + k] += (double)(dbk * dbk);
            }
        }
        block_sum<9>(acc, red[rb]);
        rb ^= 1;
        T rz[3], thr[3];
        bool act[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            rz[k] = (T)acc[k];
            const T zz = (T)acc[3 + k];
            thr[k] = max(a.tol2 * zz, a.floor2 * (T)acc[6 + k]);
            act[k] = zz > thr[k];
        }
        int it = 0;
        for (; it < a.max_iter && (act[0] || act[1] || act[2]); ++it) {
            // q = A p and p.q (p of every row was completed before the last barrier)
            double pq[3] = {0, 0, 0};
            for (int i = v0 + tid; i < v1; i += B) {
                const T dA = a.diagA[i];
                const T pi[3] = {P(3 * (size_t)i), P(3 * (size_t)i + 1), P(3 * (size_t)i + 2)};
                T qv[3] = {dA * pi[0], dA * pi[1], dA * pi[2]};
                const int il = i - v0;
                for (int j0 = 0; j0 < width; j0 += kRowBatch<T>) {
                    uint32_t uu[kRowBatch<T>];
                    T vv[kRowBatch<T>];
                    row_entries(a, ebase, nrow, il, width, j0, uu, vv);
#pragma unroll
                    for (int jj = 0; jj < kRowBatch<T>; ++jj)
                        if (j0 + jj < width)
#pragma unroll
                            for (int k = 0; k < 3; ++k) qv[k] += vv[jj] * P(3 * (size_t)uu[jj] + k);
                }
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    a.q[3 * (size_t)i + k] = qv[k];
                    pq[k] += (double)(pi[k] * qv[k]);
                }
            }
            block_sum<3>(pq, red[rb]);
            rb ^= 1;
            T alpha[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) alpha[k] = (act[k] && pq[k] != 0.0) ? rz[k] / (T)pq[k] : (T)0;
            // x += alpha p, r -= alpha q, z = D^-1 r (kept in q: this row's q is used up)
            double nz[6] = {0, 0, 0, 0, 0, 0};
            for (int i = v0 + tid; i < v1; i += B) {
                const T di = (T)1 / a.diagA[i];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const size_t j = 3 * (size_t)i + k;
                    const T pk = P(j), qk = a.q[j];
                    a.x[j] += alpha[k] * pk;
                    const T rk = a.r[j] - alpha[k] * qk;
                    const T zk = di * rk;
                    a.r[j] = rk;
                    a.q[j] = zk;
                    nz[k] += (double)(rk * zk);
                    nz[3 + k] += (double)(zk * zk);
                }
            }
            block_sum<6>(nz, red[rb]);
            rb ^= 1;
            T beta[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                beta[k] = (act[k] && rz[k] != (T)0) ? (T)nz[k] / rz[k] : (T)0;
                rz[k] = (T)nz[k];
            }
            for (int i = v0 + tid; i < v1; i += B)
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const size_t j = 3 * (size_t)i + k;
                    P(j) = a.q[j] + beta[k] * P(j);
                }
#pragma unroll
            for (int k = 0; k < 3; ++k) act[k] = act[k] && (T)nz[3 + k] > thr[k];
            __syncthreads();
        }
        it_total += it;
        ++solves;
        max_solve = it > max_solve ? it : max_solve;
        cut += (act[0] || act[1] || act[2]) ? 1 : 0;
        // Chebyshev (omega = 1 for i <= 10, then the semi-iteration's schedule)
        if (i_outer <= 10) omega = 1;
        else if (i_outer == 11) omega = (T)2 / ((T)2 - a.rho * a.rho);
        else omega = (T)4 / ((T)4 - a.rho * a.rho * omega);
        for (int i = v0 + tid; i < v1; i += B)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const size_t j = 3 * (size_t)i + k;
                const T s = a.sn[j];
                const T n = cheb(a.x[j], s, a.prev[j], omega);
                a.prev[j] = s;
                a.sn[j] = n;
            }
        __syncthreads();
    }
    // updateVelPos (pdUtil.cu:180-193) and X <- XTilde (pdSolver.cu:227)
    const T dt_inv = (T)1 / dt;
    for (int i = v0 + tid; i < v1; i += B) {
        const uint32_t g = a.perm[i];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const T n = a.sn[3 * (size_t)i + k];
            const size_t j = 3 * (size_t)g + k;
            a.V[j] = (n - (a.XT[j] - o[k])) * dt_inv;
            const T xa = n + o[k];
            a.XT[j] = xa;
            if (a.write_x) a.X[j] = xa;
        }
    }
    if (tid == 0) {
        int* st = a.stats + 4 * (size_t)b;
        st[0] += it_total;
        st[1] += solves;
        st[2] = max(st[2], max_solve);
        st[3] += cut;
    }
}

__global__ void zero_int_kernel(int n, int* p) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0;
}

template <class T> void upload(DevBuf<T>& d, const std::vector<T>& h) {
    d.alloc(std::max<size_t>(h.size(), 1));
    if (!h.empty()) SBS_HIP(hipMemcpy(d.get(), h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
}

int round64(int n) { return std::max(64, std::min(1024, (n + 63) / 64 * 64)); }

}  // namespace

template <class T> BatchEngine<T>::BatchEngine(const BatchSystem<T>& s) {
    nv_ = s.nv;
    nt_ = s.nt;
    SBS_REQUIRE(nv_ > 0, "batched PD: empty scene");
    // bodies: connected components of the tet mesh (union-find over vertices)
    std::vector<int> parent(nv_);
    std::iota(parent.begin(), parent.end(), 0);
    auto find = [&](int v) {
        while (parent[v] != v) v = parent[v] = parent[parent[v]];
        return v;
    };
    for (int t = 0; t < nt_; ++t)
        for (int k = 1; k < 4; ++k) {
            const int a = find((int)s.tets[4 * (size_t)t]), b = find((int)s.tets[4 * (size_t)t + k]);
            if (a != b) parent[std::max(a, b)] = std::min(a, b);
        }
    std::vector<int> body_of_root(nv_, -1), body(nv_);
    nb_ = 0;
    for (int v = 0; v < nv_; ++v) {  // bodies numbered by their first vertex
        const int r = find(v);
        if (body_of_root[r] < 0) body_of_root[r] = nb_++;
        body[v] = body_of_root[r];
    }
    // internal vertex order: body by body, caller order inside a body
    vptr_h_.assign(nb_ + 1, 0);
    for (int v = 0; v < nv_; ++v) vptr_h_[body[v] + 1]++;
    for (int b = 0; b < nb_; ++b) vptr_h_[b + 1] += vptr_h_[b];
    std::vector<uint32_t> perm(nv_);
    std::vector<int> inv(nv_), fill(vptr_h_.begin(), vptr_h_.end() - 1);
    for (int v = 0; v < nv_; ++v) {
        const int i = fill[body[v]]++;
        perm[i] = (uint32_t)v;
        inv[v] = i;
    }
    // internal tet order: body by body, caller order inside a body
    tptr_h_.assign(nb_ + 1, 0);
    for (int t = 0; t < nt_; ++t) tptr_h_[body[s.tets[4 * (size_t)t]] + 1]++;
    for (int b = 0; b < nb_; ++b) tptr_h_[b + 1] += tptr_h_[b];
    std::vector<int> tfill(tptr_h_.begin(), tptr_h_.end() - 1);
    std::vector<uint4> tets(std::max(nt_, 1));
    std::vector<T> dminv(9 * (size_t)std::max(nt_, 1)), w(std::max(nt_, 1));
    for (int t = 0; t < nt_; ++t) {
        const int j = tfill[body[s.tets[4 * (size_t)t]]]++;
        const uint32_t* tv = &s.tets[4 * (size_t)t];
        tets[j] = make_uint4((unsigned)inv[tv[0]], (unsigned)inv[tv[1]], (unsigned)inv[tv[2]], (unsigned)inv[tv[3]]);
        for (int k = 0; k < 9; ++k) dminv[9 * (size_t)j + k] = s.dminv[9 * (size_t)t + k];
        w[j] = s.w[t];
    }
    // corner incidences per internal vertex, ascending internal tet
    std::vector<int> inc_ptr(nv_ + 1, 0);
    for (int j = 0; j < nt_; ++j) {
        inc_ptr[tets[j].x + 1]++; inc_ptr[tets[j].y + 1]++; inc_ptr[tets[j].z + 1]++; inc_ptr[tets[j].w + 1]++;
    }
    for (int i = 0; i < nv_; ++i) inc_ptr[i + 1] += inc_ptr[i];
    std::vector<uint32_t> inc(std::max(inc_ptr[nv_], 1));
    std::vector<int> ifill(inc_ptr.begin(), inc_ptr.end() - 1);
    for (int j = 0; j < nt_; ++j) {
        const unsigned c[4] = {tets[j].x, tets[j].y, tets[j].z, tets[j].w};
        for (int k = 0; k < 4; ++k) inc[ifill[c[k]]++] = ((uint32_t)j << 2) | (uint32_t)k;
    }
    // per-body ELL of the off-diagonal entries (column-major inside the body,
    // padded with the row's own column and value 0); per-vertex constants
    std::vector<int> eptr(nb_ + 1, 0), ew(nb_, 0);
    std::vector<T> diagA(nv_), c0(nv_), mass(nv_), dbc(nv_);
    for (int i = 0; i < nv_; ++i) {
        const int v = (int)perm[i];
        diagA[i] = s.diagA[v];
        c0[i] = s.c0[v];
        mass[i] = s.mass[v];
        dbc[i] = s.dbc[v];
        ew[body[v]] = std::max(ew[body[v]], s.rowptr[v + 1] - s.rowptr[v] - 1);
    }
    for (int b = 0; b < nb_; ++b) eptr[b + 1] = eptr[b] + ew[b] * (vptr_h_[b + 1] - vptr_h_[b]);
    std::vector<uint32_t> ecol(std::max(eptr[nb_], 1));
    std::vector<T> eval(std::max(eptr[nb_], 1), (T)0);
    for (int b = 0; b < nb_; ++b) {
        const int n = vptr_h_[b + 1] - vptr_h_[b];
        for (int il = 0; il < n; ++il) {
            const int i = vptr_h_[b] + il, v = (int)perm[i];
            int j = 0;
            for (int e = s.rowptr[v]; e < s.rowptr[v + 1]; ++e) {
                if (s.col[e] == v) continue;
                const size_t pos = (size_t)eptr[b] + (size_t)j * n + il;
                ecol[pos] = (uint32_t)inv[s.col[e]];
                eval[pos] = s.val[e];
                ++j;
            }
            for (; j < ew[b]; ++j) {
                const size_t pos = (size_t)eptr[b] + (size_t)j * n + il;
                ecol[pos] = (uint32_t)i;
                eval[pos] = 0;
            }
        }
    }
    SBS_REQUIRE(eptr[nb_] >= 0 && (size_t)nt_ * 4 < (1ull << 32), "batched PD: scene too large");
    // each body's origin: the rest position of its first vertex
    std::vector<T> origin(3 * (size_t)nb_);
    for (int b = 0; b < nb_; ++b)
        for (int k = 0; k < 3; ++k) origin[3 * (size_t)b + k] = s.x0[3 * (size_t)perm[vptr_h_[b]] + k];
    upload(origin_, origin);
    upload(vptr_, vptr_h_);
    upload(tptr_, tptr_h_);
    upload(eptr_, eptr);
    upload(ew_, ew);
    upload(perm_, perm);
    upload(tets_, tets);
    upload(dminv_, dminv);
    upload(w_, w);
    upload(inc_ptr_, inc_ptr);
    upload(inc_, inc);
    upload(ecol_, ecol);
    upload(eval_, eval);
    upload(diagA_, diagA);
    upload(c0_, c0);
    upload(mass_, mass);
    upload(dbc_, dbc);
    const size_t n3 = 3 * (size_t)nv_;
    sn_.alloc(n3); sn_old_.alloc(n3); prev_.alloc(n3); x_.alloc(n3); r_.alloc(n3); p_.alloc(n3); q_.alloc(n3);
    force_.alloc(12 * (size_t)std::max(nt_, 1));
    stats_.alloc(4 * (size_t)nb_);
    reset_stats();
    // size classes by tet count (factor 4 apart from 1,024 tets), largest bodies first
    const int bounds[] = {0, 1024, 4096, 16384, 1 << 30};
    std::vector<int> list;
    for (int c = 0; c < 4; ++c) {
        std::vector<int> mem;
        for (int b = 0; b < nb_; ++b) {
            const int nt = tptr_h_[b + 1] - tptr_h_[b];
            if (nt >= bounds[c] && nt < bounds[c + 1]) mem.push_back(b);
        }
        if (mem.empty()) continue;
        std::sort(mem.begin(), mem.end(), [&](int x, int y) {
            const int a = tptr_h_[x + 1] - tptr_h_[x], b2 = tptr_h_[y + 1] - tptr_h_[y];
            return a != b2 ? a > b2 : x < y;
        });
        BatchClass bc;
        bc.min_tets = tptr_h_[mem.back() + 1] - tptr_h_[mem.back()];
        bc.max_tets = tptr_h_[mem.front() + 1] - tptr_h_[mem.front()];
        bc.bodies = (int)mem.size();
        for (int b : mem) bc.max_verts = std::max(bc.max_verts, vptr_h_[b + 1] - vptr_h_[b]);
        bc.threads = round64(bc.max_tets / 16);  // until the tuner chooses
        cls_first_.push_back((int)list.size());
        list.insert(list.end(), mem.begin(), mem.end());
        cls_.push_back(bc);
    }
    upload(body_list_, list);
    for (size_t c = 0; c < cls_.size(); ++c) {
        hipStream_t st;
        SBS_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        streams_.push_back(st);
    }
    for (size_t k = 0; k < cls_.size() + 1; ++k) {
        hipEvent_t e;
        SBS_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        events_.push_back(e);
    }
    // LDS a workgroup may take for the search direction: the device's per-block
    // limit less the kernel's static 2.3 KB (64 KB unless it reports more)
    int dev = 0, lim = 0;
    SBS_HIP(hipGetDevice(&dev));
    if (hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || lim <= 0)
        lim = 64 * 1024;
    lds_p_limit_ = (size_t)lim - 4096;
    SBS_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&batch_step_kernel<T, true>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_p_limit_));
}

template <class T> BatchEngine<T>::~BatchEngine() {
    for (auto s : streams_) {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
    }
    for (auto e : events_) (void)hipEventDestroy(e);
}

template <class T> void BatchEngine<T>::reset_stats() {
    hipLaunchKernelGGL(zero_int_kernel, dim3(ceil_div(4 * (size_t)nb_, 256)), dim3(256), 0, nullptr, 4 * nb_,
                       stats_.get());
    SBS_HIP(hipGetLastError());
    SBS_HIP(hipDeviceSynchronize());
}

template <class T> std::vector<BatchBody> BatchEngine<T>::body_stats() {
    SBS_HIP(hipDeviceSynchronize());
    std::vector<int> h(4 * (size_t)nb_);
    SBS_HIP(hipMemcpy(h.data(), stats_.get(), h.size() * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<BatchBody> out(nb_);
    for (int b = 0; b < nb_; ++b) {
        out[b].verts = vptr_h_[b + 1] - vptr_h_[b];
        out[b].tets = tptr_h_[b + 1] - tptr_h_[b];
        out[b].iterations = h[4 * b];
        out[b].solves = h[4 * b + 1];
        out[b].max_per_solve = h[4 * b + 2];
        out[b].unconverged = h[4 * b + 3];
    }
    return out;
}

template <class T>
void BatchEngine<T>::launch_class(int c, T* X, T* V, T* XT, const T* DBCX, const BatchParams& p, hipStream_t s,
                                  int threads) {
    SBS_REQUIRE(threads >= 64 && threads <= 64 * kMaxWaves && threads % 64 == 0,
                "batched PD: workgroup size must be a multiple of 64 up to 1024");
    BatchArgs<T> a;
    a.body_list = body_list_.get();
    a.vptr = vptr_.get();
    a.tptr = tptr_.get();
    a.eptr = eptr_.get();
    a.ew = ew_.get();
    a.perm = perm_.get();
    a.tets = tets_.get();
    a.dminv = dminv_.get();
    a.w = w_.get();
    a.inc_ptr = inc_ptr_.get();
    a.inc = inc_.get();
    a.ecol = ecol_.get();
    a.eval = eval_.get();
    a.diagA = diagA_.get();
    a.c0 = c0_.get();
    a.mass = mass_.get();
    a.dbc = dbc_.get();
    a.origin = origin_.get();
    a.X = X;
    a.V = V;
    a.XT = XT;
    a.DBCX = DBCX;
    a.sn = sn_.get();
    a.sn_old = sn_old_.get();
    a.prev = prev_.get();
    a.x = x_.get();
    a.r = r_.get();
    a.p = p_.get();
    a.q = q_.get();
    a.force = force_.get();
    a.stats = stats_.get();
    a.iters = p.iterations;
    a.max_iter = p.pcg_max_iter;
    a.write_x = p.write_x ? 1 : 0;
    a.polar = p.polar ? 1 : 0;
    a.dt = (T)p.dt;
    a.gravity = (T)p.gravity;
    a.rho = (T)p.rho;
    const T dt_inv = (T)1 / (T)p.dt;
    a.wdbc = (T)1e6 * (dt_inv * dt_inv);
    a.tol2 = (T)(p.pcg_tol * p.pcg_tol);
    a.floor2 = (T)(p.pcg_floor * p.pcg_floor);
    const size_t lds = 3 * (size_t)cls_[c].max_verts * sizeof(T);
    if (lds <= lds_p_limit_)
        hipLaunchKernelGGL((batch_step_kernel<T, true>), dim3(cls_[c].bodies), dim3(threads), lds, s, a,
                           cls_first_[c]);
    else
        hipLaunchKernelGGL((batch_step_kernel<T, false>), dim3(cls_[c].bodies), dim3(threads), 0, s, a,
                           cls_first_[c]);
    SBS_HIP(hipGetLastError());
}

template <class T>
void BatchEngine<T>::step(T* X, T* V, T* XT, const T* DBCX, const BatchParams& p, hipStream_t stream) {
    SBS_REQUIRE(p.iterations >= 0 && p.pcg_max_iter >= 1 && p.dt > 0, "batched PD: invalid parameters");
    if (cls_.size() == 1) {
        launch_class(0, X, V, XT, DBCX, p, stream, cls_[0].threads);
        return;
    }
    // fork: every class on its own stream, joined back into `stream`
    SBS_HIP(hipEventRecord(events_[0], stream));
    for (size_t c = 0; c < cls_.size(); ++c) {
        SBS_HIP(hipStreamWaitEvent(streams_[c], events_[0], 0));
        launch_class((int)c, X, V, XT, DBCX, p, streams_[c], cls_[c].threads);
        SBS_HIP(hipEventRecord(events_[c + 1], streams_[c]));
    }
    for (size_t c = 0; c < cls_.size(); ++c) SBS_HIP(hipStreamWaitEvent(stream, events_[c + 1], 0));
}

template <class T>
void BatchEngine<T>::autotune(const T* X, const T* V, const T* DBCX, const BatchParams& p, hipStream_t stream) {
    const size_t n3 = 3 * (size_t)nv_;
    DevBuf<T> xc(n3), vc(n3), tc(n3);
    hipEvent_t e0, e1;
    SBS_HIP(hipEventCreate(&e0));
    SBS_HIP(hipEventCreate(&e1));
    const int cand[] = {64, 128, 256, 512, 1024};
    for (size_t c = 0; c < cls_.size(); ++c) {
        cls_[c].tried.clear();
        double best = 1e300;
        int best_t = cls_[c].threads;
        for (int th : cand) {
            if (th > 64 && th / 2 >= round64(cls_[c].max_tets)) continue;  // more threads than tets: no gain
            double ms_best = 1e300;
            for (int rep = 0; rep < 3; ++rep) {  // one warm-up, then the faster of two
                SBS_HIP(hipMemcpyAsync(xc.get(), X, n3 * sizeof(T), hipMemcpyDeviceToDevice, stream));
                SBS_HIP(hipMemcpyAsync(vc.get(), V, n3 * sizeof(T), hipMemcpyDeviceToDevice, stream));
                SBS_HIP(hipMemcpyAsync(tc.get(), X, n3 * sizeof(T), hipMemcpyDeviceToDevice, stream));
                SBS_HIP(hipEventRecord(e0, stream));
                launch_class((int)c, xc.get(), vc.get(), tc.get(), DBCX, p, stream, th);
                SBS_HIP(hipEventRecord(e1, stream));
                SBS_HIP(hipEventSynchronize(e1));
                float ms = 0;
                SBS_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (rep > 0) ms_best = std::min(ms_best, (double)ms);
            }
            cls_[c].tried.emplace_back(th, ms_best);
            if (ms_best < best) {
                best = ms_best;
                best_t = th;
            }
        }
        // the classes share the CUs when they run together: of the sizes within
        // 5% of the fastest, the smallest workgroup leaves the most room
        for (const auto& tm : cls_[c].tried)
            if (tm.second <= 1.05 * best && tm.first < best_t) best_t = tm.first;
        cls_[c].threads = best_t;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    reset_stats();
}

template class BatchEngine<float>;
template class BatchEngine<double>;

}  // namespace sbs
