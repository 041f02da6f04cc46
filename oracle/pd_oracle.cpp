// CPU oracle for the projective-dynamics (PD) step.
//
// TEST INFRASTRUCTURE ONLY.  This file is a plain C++17 restatement of the
// reference's PD path, used by tests/ as the checker and by bench.py as the
// CPU baseline.  It is never linked into, loaded by, or called from the
// product library.
//
// What it restates (reference file:line):
//   * FEM precompute  Dm^-1, V0          src/simulation/solver/solverUtil.cuh:109-137,
//                                        src/simulation/solver/femSolver.cu:7-17
//   * 3x3 SVD (McAdams et al. TR1690)    external/svd3_cuda/svd3_cuda.h:36-1043 (algorithm
//                                        restated from the technical report), used as in
//                                        src/simulation/solver/svd.cuh:8-14
//   * system assembly  S^T A^T A S, M/h^2 src/simulation/solver/projective/pdUtil.cu:9-54
//   * the PD step                         src/simulation/solver/projective/pdSolver.cu:141-208
//       computeSn                         pdUtil.cu:73-95
//       addM_h2Sn                         pdUtil.cu:168-178
//       computeLocal                      pdUtil.cu:97-145
//       computeDBCLocal                   pdUtil.cu:147-166
//       Jacobi + Chebyshev                pdUtil.cu:195-226, pdSolver.cu:195-199
//       direct solve + computeError       pdSolver.cu:176-190, 243-253
//       updateVelPos                      pdUtil.cu:180-193
//       X <- XTilde (no mesh contact)     pdSolver.cu:227
//
// The reference's system is (L_s + diag(c)) (x) I_3 with interleaved DOFs
// 3v+k: computeSiTSi only couples equal components (pdUtil.cu:35) and the
// mass term is per vertex (pdUtil.cu:48-53).  The oracle therefore solves the
// scalar N_v x N_v system for 3 right-hand sides; the answer is the same.
//
// Decisions recorded where the reference is undefined:
//   * prev_x starts at 0 in the direct-solve modes (the reference memsets
//     numVerts*3 BYTES, pdSolver.cu:162, leaving the rest uninitialised);
//   * the per-vertex scatter of the local step is summed in a fixed order
//     (the reference uses float atomics in arbitrary order, pdUtil.cu:129-143).
//
// Solver modes: 0 = Jacobi/Chebyshev (reference default), 1 = direct
// (CuSolverCholesky: exact solve, error-based early exit), 2 = EigenCholesky
// (same arithmetic on the host in the reference), 3 = Jacobi-PCG (the
// global step of the 64^3 workload: direct-solve semantics with the solve replaced by
// Jacobi-preconditioned CG to a relative residual, optional Chebyshev).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

template <class S> struct Tr1690 {
    // Approximate Givens half-angle for the 2x2 symmetric block [[a, b],[b, d]].
    static void jacobi_half(S a, S b, S d, S& ch, S& sh) {
        const S tiny = (S)1e-20, four_gamma_sq = (S)5.8284271247461900976;  // 3+2*sqrt(2)
        const S cpi8 = (S)0.92387953251128675613, spi8 = (S)0.38268343236508977173;
        sh = (S)0.5 * b;
        ch = a - d;
        if (sh * sh < tiny) { sh = 0; ch = 1; }
        S w = (S)1 / std::sqrt(ch * ch + sh * sh);
        bool keep = four_gamma_sq * sh * sh < ch * ch;
        if (keep) { ch *= w; sh *= w; } else { ch = cpi8; sh = spi8; }
    }
    // QR Givens half-angle that zeroes a2 against a1 (TR1690 Alg. 4):
    // sh = a2 unless a2^2 < 1e-12, m = max(|a1|, 1e-12), ch = m + sqrt(m^2 + sh^2),
    // swapped when a1 < 0, then normalised.
    static void qr_half(S a1, S a2, S& ch, S& sh) {
        const S eps = (S)1e-12;
        sh = a2 * a2 >= eps ? a2 : (S)0;
        const S m = std::max(std::fabs(a1), eps);
        ch = m + std::sqrt(m * m + sh * sh);
        if (!(a1 >= 0)) std::swap(ch, sh);
        S w = (S)1 / std::sqrt(ch * ch + sh * sh);
        ch *= w; sh *= w;
    }
    // A = U diag(sig) V^T with U, V proper rotations; sig[0] >= sig[1] >= |sig[2]|.
    static void svd(const S A[3][3], S U[3][3], S sig[3], S V[3][3]) {
        S M[3][3];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c)
                M[r][c] = A[0][r] * A[0][c] + A[1][r] * A[1][c] + A[2][r] * A[2][c];
        for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) V[r][c] = (r == c);
        static const int P[3] = {0, 1, 2}, Q[3] = {1, 2, 0};  // cyclic pairs (0,1), (1,2), (2,0)
        for (int sweep = 0; sweep < 4; ++sweep) {
            for (int e = 0; e < 3; ++e) {
                const int p = P[e], q = Q[e], k = 3 - p - q;
                S ch, sh;
                jacobi_half(M[p][p], M[p][q], M[q][q], ch, sh);
                const S c = ch * ch - sh * sh, s = (S)2 * ch * sh;
                const S a = M[p][p], b = M[p][q], d = M[q][q];
                const S app = c * c * a + (S)2 * c * s * b + s * s * d;
                const S aqq = s * s * a - (S)2 * c * s * b + c * c * d;
                const S apq = (c * c - s * s) * b + c * s * (d - a);
                const S apk = c * M[p][k] + s * M[q][k];
                const S aqk = -s * M[p][k] + c * M[q][k];
                M[p][p] = app; M[q][q] = aqq; M[p][q] = M[q][p] = apq;
                M[p][k] = M[k][p] = apk; M[q][k] = M[k][q] = aqk;
                for (int r = 0; r < 3; ++r) {
                    const S vp = V[r][p], vq = V[r][q];
                    V[r][p] = c * vp + s * vq;
                    V[r][q] = -s * vp + c * vq;
                }
            }
        }
        S B[3][3];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c)
                B[r][c] = A[r][0] * V[0][c] + A[r][1] * V[1][c] + A[r][2] * V[2][c];
        S rho[3];
        for (int c = 0; c < 3; ++c) rho[c] = B[0][c] * B[0][c] + B[1][c] * B[1][c] + B[2][c] * B[2][c];
        // swap columns i, j and negate column `neg` (one of them) so det V stays +1:
        // 1-2 negates the 2nd, 1-3 the 1st, 2-3 the 3rd column
        auto swap_neg = [&](int i, int j, int neg) {
            for (int r = 0; r < 3; ++r) {
                std::swap(B[r][i], B[r][j]);
                std::swap(V[r][i], V[r][j]);
                B[r][neg] = -B[r][neg];
                V[r][neg] = -V[r][neg];
            }
            std::swap(rho[i], rho[j]);
        };
        if (rho[0] < rho[1]) swap_neg(0, 1, 1);
        if (rho[0] < rho[2]) swap_neg(0, 2, 0);
        if (rho[1] < rho[2]) swap_neg(1, 2, 2);
        for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) U[r][c] = (r == c);
        static const int QP[3] = {0, 0, 1}, QQ[3] = {1, 2, 2};
        for (int e = 0; e < 3; ++e) {
            const int p = QP[e], q = QQ[e];
            S ch, sh;
            qr_half(B[p][p], B[q][p], ch, sh);
            const S c = ch * ch - sh * sh, s = (S)2 * ch * sh;
            for (int col = 0; col < 3; ++col) {  // B <- G^T B on rows p, q
                const S bp = B[p][col], bq = B[q][col];
                B[p][col] = c * bp + s * bq;
                B[q][col] = -s * bp + c * bq;
            }
            for (int r = 0; r < 3; ++r) {  // U <- U G
                const S up = U[r][p], uq = U[r][q];
                U[r][p] = c * up + s * uq;
                U[r][q] = -s * up + c * uq;
            }
        }
        sig[0] = B[0][0]; sig[1] = B[1][1]; sig[2] = B[2][2];
    }
};

template <class S> S det3(const S m[3][3]) {
    return m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2]) -
           m[0][1] * (m[1][0] * m[2][2] - m[2][0] * m[1][2]) +
           m[0][2] * (m[1][0] * m[2][1] - m[2][0] * m[1][1]);
}

// pdUtil.cu:112-125: F = Ds Dm^-1, R = U V^T, det<0 -> negate column 2,
// Jacobi mode returns R - F.
template <class S> void local_rotation(const S F[3][3], bool jacobi, S out[3][3]) {
    S U[3][3], sig[3], V[3][3];
    Tr1690<S>::svd(F, U, sig, V);
    S R[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            R[r][c] = U[r][0] * V[c][0] + U[r][1] * V[c][1] + U[r][2] * V[c][2];
    if (det3(R) < 0)
        for (int r = 0; r < 3; ++r) R[r][2] = -R[r][2];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) out[r][c] = jacobi ? R[r][c] - F[r][c] : R[r][c];
}

struct OrcParams {
    int solver;          // 0 Jacobi/Chebyshev, 1/2 direct, 3 PCG
    int num_iterations;  // local/global iterations per step
    double dt;
    double gravity;      // magnitude along -y
    double tol;          // outer early-exit tolerance (direct / PCG modes)
    double rho;          // Chebyshev spectral radius estimate
    double positional_weight;  // 1e6 (pdSolver.h:31)
    double pcg_tol;      // residual reduction ||z_k|| <= pcg_tol*||z_0|| per component (mode 3)
    double pcg_floor;    // ... or ||z_k|| <= pcg_floor*||D^-1 b|| (precision floor)
    int pcg_max_iter;
    int pcg_chebyshev;   // apply the Chebyshev recurrence after the PCG solve (mode 3)
};

template <class S> struct Pd {
    int nv = 0, nt = 0;
    std::vector<uint32_t> tet;
    std::vector<S> dminv, V0, w;      // 9*nt row-major, nt, nt
    std::vector<S> mass, dbc, c;      // nv
    std::vector<S> more;              // nv moreDBC weights (empty: all zero)
    std::vector<S> diagL;             // nv: diagonal of L_s (matrix_diag)
    std::vector<int> rowptr, col;     // CSR of A_s = L_s + diag(c)
    std::vector<S> val;
    std::vector<int> vinc_ptr, vinc;  // vertex -> (tet*4+slot) incidence, fixed order
    // banded/envelope Cholesky of A_s
    std::vector<int> env_first;       // first column of row i in the envelope
    std::vector<size_t> env_ptr;
    std::vector<double> env_L;        // factor always kept in double for the FP64 oracle;
                                      // the FP32 oracle uses a float copy below
    std::vector<S> env_Ls;
    std::vector<S> X, V, X0, XTilde, DBCX;  // AoS xyz

    void precompute(const double* X0in) {
        dminv.assign(9 * (size_t)nt, 0);
        V0.assign(nt, 0);
        for (int t = 0; t < nt; ++t) {
            const uint32_t* T = &tet[4 * (size_t)t];
            S Dm[3][3];
            for (int r = 0; r < 3; ++r)
                for (int k = 0; k < 3; ++k)
                    Dm[r][k] = (S)X0in[3 * T[k + 1] + r] - (S)X0in[3 * T[0] + r];
            S d = det3(Dm);
            S inv[3][3];
            inv[0][0] = (Dm[1][1] * Dm[2][2] - Dm[1][2] * Dm[2][1]) / d;
            inv[0][1] = (Dm[0][2] * Dm[2][1] - Dm[0][1] * Dm[2][2]) / d;
            inv[0][2] = (Dm[0][1] * Dm[1][2] - Dm[0][2] * Dm[1][1]) / d;
            inv[1][0] = (Dm[1][2] * Dm[2][0] - Dm[1][0] * Dm[2][2]) / d;
            inv[1][1] = (Dm[0][0] * Dm[2][2] - Dm[0][2] * Dm[2][0]) / d;
            inv[1][2] = (Dm[0][2] * Dm[1][0] - Dm[0][0] * Dm[1][2]) / d;
            inv[2][0] = (Dm[1][0] * Dm[2][1] - Dm[1][1] * Dm[2][0]) / d;
            inv[2][1] = (Dm[0][1] * Dm[2][0] - Dm[0][0] * Dm[2][1]) / d;
            inv[2][2] = (Dm[0][0] * Dm[1][1] - Dm[0][1] * Dm[1][0]) / d;
            for (int r = 0; r < 3; ++r)
                for (int k = 0; k < 3; ++k) dminv[9 * (size_t)t + 3 * r + k] = inv[r][k];
            V0[t] = std::fabs(d) / (S)6;
        }
    }

    // Local 4x4 matrix K = G Dm^-1 Dm^-T G^T, G rows: (-1,-1,-1),(1,0,0),(0,1,0),(0,0,1)
    // (pdUtil.cu:16-17; glm mat4x3 {-1,-1,-1,1,0,0,0,1,0,0,0,1}).
    void local_k(int t, S K[4][4]) const {
        const S* m = &dminv[9 * (size_t)t];
        S g[4][3];
        for (int k = 0; k < 3; ++k) {
            g[0][k] = -(m[0 * 3 + k] + m[1 * 3 + k] + m[2 * 3 + k]);
            g[1][k] = m[0 * 3 + k];
            g[2][k] = m[1 * 3 + k];
            g[3][k] = m[2 * 3 + k];
        }
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) K[i][j] = g[i][0] * g[j][0] + g[i][1] * g[j][1] + g[i][2] * g[j][2];
    }

    void assemble(const OrcParams& p) {
        const S dt2 = (S)p.dt * (S)p.dt;
        c.assign(nv, 0);
        diagL.assign(nv, 0);
        for (int v = 0; v < nv; ++v) c[v] = (mass[v] + dbc[v] * (S)p.positional_weight) / dt2;
        // adjacency
        std::vector<std::vector<int>> nb(nv);
        for (int t = 0; t < nt; ++t)
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) nb[tet[4 * (size_t)t + i]].push_back((int)tet[4 * (size_t)t + j]);
        rowptr.assign(nv + 1, 0);
        col.clear();
        for (int v = 0; v < nv; ++v) {
            nb[v].push_back(v);
            std::sort(nb[v].begin(), nb[v].end());
            nb[v].erase(std::unique(nb[v].begin(), nb[v].end()), nb[v].end());
            rowptr[v + 1] = rowptr[v] + (int)nb[v].size();
            col.insert(col.end(), nb[v].begin(), nb[v].end());
        }
        val.assign(col.size(), 0);
        auto slot = [&](int r, int cc) {
            auto b = col.begin() + rowptr[r], e = col.begin() + rowptr[r + 1];
            return (int)(std::lower_bound(b, e, cc) - col.begin());
        };
        for (int t = 0; t < nt; ++t) {
            S K[4][4];
            local_k(t, K);
            for (int i = 0; i < 4; ++i) {
                const int vi = (int)tet[4 * (size_t)t + i];
                diagL[vi] += K[i][i] * w[t];
                for (int j = 0; j < 4; ++j) val[slot(vi, (int)tet[4 * (size_t)t + j])] += K[i][j] * w[t];
            }
        }
        for (int v = 0; v < nv; ++v) val[slot(v, v)] += c[v];
        // incidence
        vinc_ptr.assign(nv + 1, 0);
        for (int t = 0; t < nt; ++t)
            for (int i = 0; i < 4; ++i) vinc_ptr[tet[4 * (size_t)t + i] + 1]++;
        for (int v = 0; v < nv; ++v) vinc_ptr[v + 1] += vinc_ptr[v];
        vinc.assign(4 * (size_t)nt, 0);
        std::vector<int> fill(vinc_ptr.begin(), vinc_ptr.end() - 1);
        for (int t = 0; t < nt; ++t)
            for (int i = 0; i < 4; ++i) vinc[fill[tet[4 * (size_t)t + i]]++] = 4 * t + i;
    }

    // Envelope Cholesky A_s = L L^T in double (the direct solve is exact up to
    // rounding; the reference's factor is Eigen/cuSOLVER, whose rounding is
    // not reproducible here).
    bool factor() {
        env_first.assign(nv, 0);
        env_ptr.assign(nv + 1, 0);
        for (int i = 0; i < nv; ++i) {
            env_first[i] = std::min(i, col[rowptr[i]]);
            env_ptr[i + 1] = env_ptr[i] + (size_t)(i - env_first[i] + 1);
        }
        env_L.assign(env_ptr[nv], 0.0);
        for (int i = 0; i < nv; ++i)
            for (int k = rowptr[i]; k < rowptr[i + 1]; ++k)
                if (col[k] <= i) env_L[env_ptr[i] + (col[k] - env_first[i])] = (double)val[k];
        for (int i = 0; i < nv; ++i) {
            double* Li = &env_L[env_ptr[i]];
            const int fi = env_first[i];
            for (int j = fi; j <= i; ++j) {
                double* Lj = &env_L[env_ptr[j]];
                const int fj = env_first[j];
                const int k0 = std::max(fi, fj);
                double s = Li[j - fi];
                for (int k = k0; k < j; ++k) s -= Li[k - fi] * Lj[k - fj];
                if (j < i) Li[j - fi] = s / Lj[j - fj];
                else {
                    if (!(s > 0)) return false;
                    Li[i - fi] = std::sqrt(s);
                }
            }
        }
        env_Ls.assign(env_L.begin(), env_L.end());
        return true;
    }

    // x = A_s^-1 b for the 3 interleaved components.
    void direct_solve(const S* b, S* x) const {
        std::vector<double> y(nv);
        for (int k = 0; k < 3; ++k) {
            for (int i = 0; i < nv; ++i) {
                const int fi = env_first[i];
                const double* Li = &env_L[env_ptr[i]];
                double s = (double)b[3 * i + k];
                for (int j = fi; j < i; ++j) s -= Li[j - fi] * y[j];
                y[i] = s / Li[i - fi];
            }
            for (int i = nv - 1; i >= 0; --i) {
                const int fi = env_first[i];
                const double* Li = &env_L[env_ptr[i]];
                y[i] /= Li[i - fi];
                for (int j = fi; j < i; ++j) y[j] -= Li[j - fi] * y[i];
            }
            for (int i = 0; i < nv; ++i) x[3 * i + k] = (S)y[i];
        }
    }

    void spmv3(const S* x, S* y) const {
#pragma omp parallel for schedule(static)
        for (int v = 0; v < nv; ++v) {
            S a0 = 0, a1 = 0, a2 = 0;
            for (int k = rowptr[v]; k < rowptr[v + 1]; ++k) {
                const S m = val[k];
                const int u = col[k];
                a0 += m * x[3 * u]; a1 += m * x[3 * u + 1]; a2 += m * x[3 * u + 2];
            }
            y[3 * v] = a0; y[3 * v + 1] = a1; y[3 * v + 2] = a2;
        }
    }

    // Sums of f(v) (3 components) over all vertices: fixed blocks of
    // vertices summed in double in parallel, then the block sums added in
    // block order, so the result does not depend on the thread count.
    template <class F> void vsum3(F&& f, double out[3]) const {
        constexpr int kBlocks = 256;
        double part[kBlocks][3];
#pragma omp parallel for schedule(static)
        for (int bk = 0; bk < kBlocks; ++bk) {
            const int v0 = (int)((long long)nv * bk / kBlocks), v1 = (int)((long long)nv * (bk + 1) / kBlocks);
            double a[3] = {0, 0, 0};
            for (int v = v0; v < v1; ++v) f(v, a);
            for (int k = 0; k < 3; ++k) part[bk][k] = a[k];
        }
        for (int k = 0; k < 3; ++k) out[k] = 0;
        for (int bk = 0; bk < kBlocks; ++bk)
            for (int k = 0; k < 3; ++k) out[k] += part[bk][k];
    }

    // Jacobi-preconditioned CG on the scalar system, 3 independent RHS,
    // warm start from x (pcgJacobi.cu:89-173, with cuBLAS's dot products as
    // double-accumulated block sums).  Residuals are measured Jacobi-scaled,
    // z = D^-1 r, which has units of displacement, so the pinned rows
    // (b = x0*1e6/h^2) do not swamp the norm the way they swamp ||b||.
    // Component k stops when ||z_k|| <= max(tol*||z_0||, floor*||D^-1 b||).
    int pcg(const S* b, S* x, double tol, double floor_tol, int max_iter) const {
        const size_t n3 = 3 * (size_t)nv;
        std::vector<S> r(n3), z(n3), p(n3), q(n3), dinv(nv);
#pragma omp parallel for schedule(static)
        for (int v = 0; v < nv; ++v)
            dinv[v] = (S)1 / val[std::lower_bound(col.begin() + rowptr[v], col.begin() + rowptr[v + 1], v) - col.begin()];
        spmv3(x, q.data());
        double rzd[3], bbd[3], rrd[3];
        vsum3([&](int v, double* a) {
            for (int k = 0; k < 3; ++k) {
                const size_t i = 3 * (size_t)v + k;
                r[i] = b[i] - q[i];
                z[i] = dinv[v] * r[i];
                p[i] = z[i];
                a[k] += (double)(r[i] * z[i]);
            }
        }, rzd);
        vsum3([&](int v, double* a) {
            for (int k = 0; k < 3; ++k) {
                const S db = dinv[v] * b[3 * (size_t)v + k];
                a[k] += (double)(db * db);
            }
        }, bbd);
        vsum3([&](int v, double* a) {
            for (int k = 0; k < 3; ++k) a[k] += (double)(z[3 * (size_t)v + k] * z[3 * (size_t)v + k]);
        }, rrd);
        S rz[3], rr[3];
        bool active[3];
        S thr[3];
        for (int k = 0; k < 3; ++k) {
            rz[k] = (S)rzd[k];
            rr[k] = (S)rrd[k];
            thr[k] = std::max((S)(tol * tol) * rr[k], (S)(floor_tol * floor_tol) * (S)bbd[k]);
            active[k] = rr[k] > thr[k];
        }
        int it = 0;
        for (; it < max_iter && (active[0] || active[1] || active[2]); ++it) {
            spmv3(p.data(), q.data());
            double pqd[3];
            vsum3([&](int v, double* a) {
                for (int k = 0; k < 3; ++k) a[k] += (double)(p[3 * (size_t)v + k] * q[3 * (size_t)v + k]);
            }, pqd);
            S alpha[3];
            for (int k = 0; k < 3; ++k) alpha[k] = (active[k] && pqd[k] != 0) ? rz[k] / (S)pqd[k] : (S)0;
            double rzn[3], rrn[3];
            vsum3([&](int v, double* a) {
                for (int k = 0; k < 3; ++k) {
                    const size_t i = 3 * (size_t)v + k;
                    x[i] += alpha[k] * p[i];
                    r[i] -= alpha[k] * q[i];
                    z[i] = dinv[v] * r[i];
                    a[k] += (double)(r[i] * z[i]);
                }
            }, rzn);
            vsum3([&](int v, double* a) {
                for (int k = 0; k < 3; ++k) a[k] += (double)(z[3 * (size_t)v + k] * z[3 * (size_t)v + k]);
            }, rrn);
            S beta[3];
            for (int k = 0; k < 3; ++k) {
                beta[k] = (active[k] && rz[k] != 0) ? (S)rzn[k] / rz[k] : (S)0;
                rz[k] = (S)rzn[k];
                rr[k] = (S)rrn[k];
            }
#pragma omp parallel for schedule(static)
            for (int v = 0; v < nv; ++v)
                for (int k = 0; k < 3; ++k) {
                    const size_t i = 3 * (size_t)v + k;
                    p[i] = z[i] + beta[k] * p[i];
                }
            for (int k = 0; k < 3; ++k) active[k] = active[k] && rr[k] > thr[k];
        }
        return it;
    }

    // computeLocal (pdUtil.cu:97-145) summed per vertex in a fixed order.
    void local_step(const S* q, bool jacobi, S* b, std::vector<S>& contrib) const {
#pragma omp parallel for schedule(static)
        for (int t = 0; t < nt; ++t) {
            const uint32_t* T = &tet[4 * (size_t)t];
            const S* m = &dminv[9 * (size_t)t];
            S Ds[3][3];
            for (int r = 0; r < 3; ++r)
                for (int k = 0; k < 3; ++k) Ds[r][k] = q[3 * T[k + 1] + r] - q[3 * T[0] + r];
            S F[3][3];
            for (int r = 0; r < 3; ++r)
                for (int k = 0; k < 3; ++k) F[r][k] = Ds[r][0] * m[0 * 3 + k] + Ds[r][1] * m[1 * 3 + k] + Ds[r][2] * m[2 * 3 + k];
            S R[3][3];
            local_rotation(F, jacobi, R);
            // Q = w R Dm^-T ; column j of Q = R (row j of Dm^-1)
            const S ww = w[t];
            S Qm[3][3];
            for (int r = 0; r < 3; ++r)
                for (int j = 0; j < 3; ++j) Qm[r][j] = ww * (R[r][0] * m[j * 3 + 0] + R[r][1] * m[j * 3 + 1] + R[r][2] * m[j * 3 + 2]);
            S* o = &contrib[12 * (size_t)t];
            for (int r = 0; r < 3; ++r) {
                o[0 * 3 + r] = -(Qm[r][0] + Qm[r][1] + Qm[r][2]);
                o[1 * 3 + r] = Qm[r][0];
                o[2 * 3 + r] = Qm[r][1];
                o[3 * 3 + r] = Qm[r][2];
            }
        }
#pragma omp parallel for schedule(static)
        for (int v = 0; v < nv; ++v) {
            S a0 = 0, a1 = 0, a2 = 0;
            for (int k = vinc_ptr[v]; k < vinc_ptr[v + 1]; ++k) {
                const S* o = &contrib[3 * (size_t)vinc[k]];
                a0 += o[0]; a1 += o[1]; a2 += o[2];
            }
            b[3 * v] += a0; b[3 * v + 1] += a1; b[3 * v + 2] += a2;
        }
    }

    // One PdSolver::Update without mesh contact (pdSolver.cu:141-208, 227).
    // moreDBC (the mouse-drag weights, empty: none) follows setMDt_2MoreDBC,
    // computeSn, computeDBCLocal, getErrorKern and updateVelPos: a vertex with
    // more > 0 (and not pinned) gets c = (m + more)/h^2, sn = its target
    // (DBCX, which computeSn sets to the target), b = DBCX * more, no Jacobi
    // move, and zero velocity.
    int step(const OrcParams& p, int* pcg_iters_total) {
        const S dt = (S)p.dt, dtInv = (S)1 / dt, dt2 = dt * dt, dt2Inv = dtInv * dtInv;
        const size_t n3 = 3 * (size_t)nv;
        std::vector<S> sn(n3), sn_old(n3), prev(n3), next(n3), b(n3), contrib(12 * (size_t)nt);
        const bool has_more = !more.empty();
        auto mw = [&](int v) { return has_more ? more[v] : (S)0; };
#pragma omp parallel for schedule(static)
        for (int v = 0; v < nv; ++v) {
            if (dbc[v] == 0) c[v] = mw(v) > 0 ? (mass[v] + mw(v)) / dt2 : mass[v] / dt2;  // setMDt_2MoreDBC
            if (mw(v) > 0) {
                for (int k = 0; k < 3; ++k) sn[3 * v + k] = DBCX[3 * v + k];
            } else {
                const S f[3] = {0, -(S)p.gravity * mass[v], 0};
                const S inv = (S)1 / c[v];
                for (int k = 0; k < 3; ++k) sn[3 * v + k] = X[3 * v + k] + dt * V[3 * v + k] + inv * f[k];
            }
        }
        sn_old = sn;
        if (p.solver == 0) prev = sn; else std::fill(prev.begin(), prev.end(), (S)0);
        const bool jacobi = p.solver == 0;
        double err = 1;
        S omega = 1;
        int iters = 0;
        const S wdbc = (S)p.positional_weight * dt2Inv;
        for (int i = 0; i < p.num_iterations && std::sqrt(err) >= p.tol; ++i) {
            ++iters;
#pragma omp parallel for schedule(static)
            for (int v = 0; v < nv; ++v)
                for (int k = 0; k < 3; ++k) b[3 * v + k] = c[v] * sn_old[3 * v + k];
            local_step(sn.data(), jacobi, b.data(), contrib);
#pragma omp parallel for schedule(static)
            for (int v = 0; v < nv; ++v) {  // computeDBCLocal
                if (dbc[v] > 0)
                    for (int k = 0; k < 3; ++k) b[3 * v + k] = DBCX[3 * v + k] * wdbc;
                else if (mw(v) > 0)
                    for (int k = 0; k < 3; ++k) b[3 * v + k] = DBCX[3 * v + k] * mw(v);
            }
            if (p.solver == 0 || (p.solver == 3 && p.pcg_chebyshev)) {
                if (p.solver == 0) {
#pragma omp parallel for schedule(static)
                    for (int v = 0; v < nv; ++v) {
                        const S cc = c[v], md = diagL[v];
                        for (int k = 0; k < 3; ++k)
                            next[3 * v + k] = mw(v) > 0 ? sn[3 * v + k]
                                                        : (b[3 * v + k] - cc * sn[3 * v + k]) / (cc + md) + sn[3 * v + k];
                    }
                } else {
                    next = sn;
                    int it = pcg(b.data(), next.data(), p.pcg_tol, p.pcg_floor, p.pcg_max_iter);
                    if (pcg_iters_total) *pcg_iters_total += it;
                }
                if (i <= 10) omega = 1;
                else if (i == 11) omega = (S)2 / ((S)2 - (S)p.rho * (S)p.rho);
                else omega = (S)4 / ((S)4 - (S)p.rho * (S)p.rho * omega);
                // chebyshevKern (pdUtil.cu:216-226): 0.9 is a double literal,
                // so the damping is evaluated in double and rounded once
#pragma omp parallel for schedule(static)
                for (size_t j = 0; j < n3; ++j) {
                    S nx = (S)(0.9 * (double)(next[j] - sn[j]) + (double)sn[j]);
                    nx = (nx - prev[j]) * omega + prev[j];
                    prev[j] = sn[j];
                    sn[j] = nx;
                }
            } else {
                if (p.solver == 3) {
                    int it = pcg(b.data(), sn.data(), p.pcg_tol, p.pcg_floor, p.pcg_max_iter);
                    if (pcg_iters_total) *pcg_iters_total += it;
                } else {
                    direct_solve(b.data(), sn.data());
                }
                double e3[3];
                vsum3([&](int v, double* a) {
                    for (int k = 0; k < 3; ++k) {
                        const S d = prev[3 * (size_t)v + k] - sn[3 * (size_t)v + k];
                        a[k] += (double)(d * d);
                    }
                }, e3);
                err = (e3[0] + e3[1] + e3[2]) / (double)n3;
                prev = sn;
            }
        }
#pragma omp parallel for schedule(static)
        for (int v = 0; v < nv; ++v)
            for (int k = 0; k < 3; ++k) {
                V[3 * v + k] = mw(v) > 0 ? (S)0 : (sn[3 * v + k] - XTilde[3 * v + k]) * dtInv;
                XTilde[3 * v + k] = sn[3 * v + k];
            }
        X = XTilde;
        return iters;
    }
};

template <class S> void* create(int nv, int nt, const double* X0, const double* V0in, const uint32_t* tet,
                                const double* mass, const double* mu, const double* dbc, const OrcParams* p) {
    auto* pd = new Pd<S>();
    pd->nv = nv; pd->nt = nt;
    pd->tet.assign(tet, tet + 4 * (size_t)nt);
    pd->precompute(X0);
    pd->w.resize(nt);
    for (int t = 0; t < nt; ++t) pd->w[t] = std::fabs(pd->V0[t]) * (S)mu[t];  // |V0| * mu (pdUtil.cu:126)
    pd->mass.assign(mass, mass + nv);
    pd->dbc.assign(dbc, dbc + nv);
    pd->X.resize(3 * (size_t)nv);
    pd->V.resize(3 * (size_t)nv);
    for (size_t i = 0; i < 3 * (size_t)nv; ++i) { pd->X[i] = (S)X0[i]; pd->V[i] = V0in ? (S)V0in[i] : (S)0; }
    pd->X0 = pd->X; pd->XTilde = pd->X; pd->DBCX = pd->X;
    pd->assemble(*p);
    if (p->solver == 1 || p->solver == 2) {
        if (!pd->factor()) { delete pd; return nullptr; }
    }
    return pd;
}

}  // namespace

extern "C" {

// 3x3 SVD of a row-major matrix (TR1690 algorithm, 4 Jacobi sweeps).
void orc_svd3_f32(const float* A9, float* U9, float* s3, float* V9) {
    float A[3][3], U[3][3], V[3][3];
    std::memcpy(A, A9, sizeof A);
    Tr1690<float>::svd(A, U, s3, V);
    std::memcpy(U9, U, sizeof U); std::memcpy(V9, V, sizeof V);
}
void orc_svd3_f64(const double* A9, double* U9, double* s3, double* V9) {
    double A[3][3], U[3][3], V[3][3];
    std::memcpy(A, A9, sizeof A);
    Tr1690<double>::svd(A, U, s3, V);
    std::memcpy(U9, U, sizeof U); std::memcpy(V9, V, sizeof V);
}
// Projection used by computeLocal: R (or R - F in Jacobi mode).
void orc_local_rotation_f64(const double* F9, int jacobi, double* R9) {
    double F[3][3], R[3][3];
    std::memcpy(F, F9, sizeof F);
    local_rotation(F, jacobi != 0, R);
    std::memcpy(R9, R, sizeof R);
}
void orc_local_rotation_f32(const float* F9, int jacobi, float* R9) {
    float F[3][3], R[3][3];
    std::memcpy(F, F9, sizeof F);
    local_rotation(F, jacobi != 0, R);
    std::memcpy(R9, R, sizeof R);
}

void* orc_pd_create(int fp64, int nv, int nt, const double* X0, const double* V0, const uint32_t* tet,
                    const double* mass, const double* mu, const double* dbc, const OrcParams* p) {
    return fp64 ? create<double>(nv, nt, X0, V0, tet, mass, mu, dbc, p)
                : create<float>(nv, nt, X0, V0, tet, mass, mu, dbc, p);
}

void orc_pd_destroy(int fp64, void* h) {
    if (fp64) delete static_cast<Pd<double>*>(h); else delete static_cast<Pd<float>*>(h);
}

// Runs nsteps PD steps; returns the total number of local/global iterations.
int orc_pd_step(int fp64, void* h, const OrcParams* p, int nsteps, int* pcg_iters) {
    int total = 0;
    for (int s = 0; s < nsteps; ++s)
        total += fp64 ? static_cast<Pd<double>*>(h)->step(*p, pcg_iters) : static_cast<Pd<float>*>(h)->step(*p, pcg_iters);
    return total;
}

// moreDBC weights (nv) and the targets of the vertices with a positive
// weight (3*nv, read only there): computeSn writes the target into DBCX.
void orc_pd_set_more_dbc(int fp64, void* h, const double* more, const double* target) {
    auto set = [&](auto* pd) {
        using S = std::remove_reference_t<decltype(pd->X[0])>;
        pd->more.assign(more, more + pd->nv);
        for (int v = 0; v < pd->nv; ++v)
            if (more[v] > 0)
                for (int k = 0; k < 3; ++k) pd->DBCX[3 * (size_t)v + k] = (S)target[3 * (size_t)v + k];
    };
    if (fp64) set(static_cast<Pd<double>*>(h)); else set(static_cast<Pd<float>*>(h));
}

// Copies out state: which 0 = X, 1 = V, 2 = XTilde (3*nv doubles).
void orc_pd_get(int fp64, void* h, int which, double* out) {
    auto get = [&](auto* pd) {
        const auto& src = which == 0 ? pd->X : which == 1 ? pd->V : pd->XTilde;
        for (size_t i = 0; i < src.size(); ++i) out[i] = (double)src[i];
    };
    if (fp64) get(static_cast<Pd<double>*>(h)); else get(static_cast<Pd<float>*>(h));
}

// Prepared quantities: dminv (9*nt row-major), V0 (nt), diagL (nv), c (nv),
// CSR of A_s (rowptr nv+1, col nnz, val nnz).  Pass nullptr to skip.
int orc_pd_nnz(int fp64, void* h) {
    return fp64 ? (int)static_cast<Pd<double>*>(h)->col.size() : (int)static_cast<Pd<float>*>(h)->col.size();
}
void orc_pd_export(int fp64, void* h, double* dminv, double* V0, double* diagL, double* c,
                   int* rowptr, int* col, double* val) {
    auto ex = [&](auto* pd) {
        if (dminv) for (size_t i = 0; i < pd->dminv.size(); ++i) dminv[i] = (double)pd->dminv[i];
        if (V0) for (size_t i = 0; i < pd->V0.size(); ++i) V0[i] = (double)pd->V0[i];
        if (diagL) for (size_t i = 0; i < pd->diagL.size(); ++i) diagL[i] = (double)pd->diagL[i];
        if (c) for (size_t i = 0; i < pd->c.size(); ++i) c[i] = (double)pd->c[i];
        if (rowptr) std::copy(pd->rowptr.begin(), pd->rowptr.end(), rowptr);
        if (col) std::copy(pd->col.begin(), pd->col.end(), col);
        if (val) for (size_t i = 0; i < pd->val.size(); ++i) val[i] = (double)pd->val[i];
    };
    if (fp64) ex(static_cast<Pd<double>*>(h)); else ex(static_cast<Pd<float>*>(h));
}

// One local step on a given iterate q (3*nv), b initialised to c*sn_old:
// exposes computeLocal + computeDBCLocal for kernel-level parity tests.
void orc_pd_local(int fp64, void* h, const double* q, const double* sn_old, int jacobi, double dt,
                  double positional_weight, double* b_out) {
    auto run = [&](auto* pd) {
        using S = std::remove_reference_t<decltype(pd->X[0])>;
        const size_t n3 = 3 * (size_t)pd->nv;
        std::vector<S> qs(n3), b(n3), contrib(12 * (size_t)pd->nt);
        for (size_t i = 0; i < n3; ++i) qs[i] = (S)q[i];
        for (int v = 0; v < pd->nv; ++v)
            for (int k = 0; k < 3; ++k) b[3 * v + k] = pd->c[v] * (S)sn_old[3 * v + k];
        pd->local_step(qs.data(), jacobi != 0, b.data(), contrib);
        const S dtInv = (S)1 / (S)dt, dt2Inv = dtInv * dtInv;
        const S wdbc = (S)positional_weight * dt2Inv;
        for (int v = 0; v < pd->nv; ++v)
            if (pd->dbc[v] > 0)
                for (int k = 0; k < 3; ++k) b[3 * v + k] = pd->DBCX[3 * v + k] * wdbc;
        for (size_t i = 0; i < n3; ++i) b_out[i] = (double)b[i];
    };
    if (fp64) run(static_cast<Pd<double>*>(h)); else run(static_cast<Pd<float>*>(h));
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
