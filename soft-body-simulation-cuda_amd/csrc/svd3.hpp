// 3x3 rotation-variant SVD held in registers, for the per-tet local step.
//
// Algorithm: McAdams, Selle, Tamstorf, Teran, Sifakis, "Computing the
// Singular Value Decomposition of 3x3 matrices with minimal branching and
// elementary floating point operations", UW-Madison TR1690 (2011) -- the
// method the reference vendors as external/svd3_cuda/svd3_cuda.h:36-1043 and
// calls through svdGLM (src/simulation/solver/svd.cuh:8-14):
//   1. S = A^T A;
//   2. 4 cyclic Jacobi sweeps over the pairs (0,1), (1,2), (2,0) with the
//      approximate half-angle Givens rotation (gamma = 3 + 2 sqrt 2 test,
//      pi/8 fallback), accumulating V;
//   3. B = A V, columns sorted by descending norm; each swap negates one
//      column (1-2: the 2nd, 1-3: the 1st, 2-3: the 3rd) so det V stays +1;
//   4. Givens QR of B (zeroing B10, B20, B21), U = G1 G2 G3, sigma = diag(B).
// U and V are proper rotations; only sigma[2] can be negative.
//
// Everything is unrolled with compile-time indices so the 3x3 matrices stay
// in VGPRs.  The sequence of operations matches the CPU oracle so FP64 runs
// agree to rounding.
#pragma once

#include <hip/hip_runtime.h>

namespace sbs {

template <class T> __device__ __forceinline__ T rsq(T x);
template <> __device__ __forceinline__ float rsq<float>(float x) { return rsqrtf(x); }
template <> __device__ __forceinline__ double rsq<double>(double x) { return rsqrt(x); }

template <class T> struct Svd3 {
    __device__ __forceinline__ static void jacobi_half(T a, T b, T d, T& ch, T& sh) {
        const T tiny = (T)1e-20, four_gamma_sq = (T)5.8284271247461900976;
        const T cpi8 = (T)0.92387953251128675613, spi8 = (T)0.38268343236508977173;
        sh = (T)0.5 * b;
        ch = a - d;
        const bool small = sh * sh < tiny;
        sh = small ? (T)0 : sh;
        ch = small ? (T)1 : ch;
        const T w = rsq<T>(ch * ch + sh * sh);
        const bool keep = four_gamma_sq * sh * sh < ch * ch;
        ch = keep ? w * ch : cpi8;
        sh = keep ? w * sh : spi8;
    }

    __device__ __forceinline__ static void qr_half(T a1, T a2, T& ch, T& sh) {
        const T eps = (T)1e-12;
        const T s = a2 * a2 >= eps ? a2 : (T)0;
        const T m = fmax(fabs(a1), eps);
        const T c = m + sqrt(m * m + s * s);
        const bool keep = a1 >= (T)0;
        ch = keep ? c : s;
        sh = keep ? s : c;
        const T w = rsq<T>(ch * ch + sh * sh);
        ch *= w;
        sh *= w;
    }

    // A row-major: A[r][c].  Outputs U, V row-major, sig.
    __device__ __forceinline__ static void run(const T (&A)[3][3], T (&U)[3][3], T (&sig)[3], T (&V)[3][3]) {
        T M[3][3];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) M[r][c] = A[0][r] * A[0][c] + A[1][r] * A[1][c] + A[2][r] * A[2][c];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) V[r][c] = (r == c) ? (T)1 : (T)0;
#pragma unroll
        for (int sweep = 0; sweep < 4; ++sweep) {
#pragma unroll
            for (int e = 0; e < 3; ++e) {
                const int p = e;             // pairs (0,1), (1,2), (2,0)
                const int q = (e + 1) % 3;
                const int k = 3 - p - q;
                T ch, sh;
                jacobi_half(M[p][p], M[p][q], M[q][q], ch, sh);
                const T c = ch * ch - sh * sh, s = (T)2 * ch * sh;
                const T a = M[p][p], b = M[p][q], d = M[q][q];
                const T app = c * c * a + (T)2 * c * s * b + s * s * d;
                const T aqq = s * s * a - (T)2 * c * s * b + c * c * d;
                const T apq = (c * c - s * s) * b + c * s * (d - a);
                const T apk = c * M[p][k] + s * M[q][k];
                const T aqk = -s * M[p][k] + c * M[q][k];
                M[p][p] = app;
                M[q][q] = aqq;
                M[p][q] = apq;
                M[q][p] = apq;
                M[p][k] = apk;
                M[k][p] = apk;
                M[q][k] = aqk;
                M[k][q] = aqk;
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    const T vp = V[r][p], vq = V[r][q];
                    V[r][p] = c * vp + s * vq;
                    V[r][q] = -s * vp + c * vq;
                }
            }
        }
        T B[3][3];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) B[r][c] = A[r][0] * V[0][c] + A[r][1] * V[1][c] + A[r][2] * V[2][c];
        T rho[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) rho[c] = B[0][c] * B[0][c] + B[1][c] * B[1][c] + B[2][c] * B[2][c];
        swap_neg<0, 1, 1>(B, V, rho);
        swap_neg<0, 2, 0>(B, V, rho);
        swap_neg<1, 2, 2>(B, V, rho);
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) U[r][c] = (r == c) ? (T)1 : (T)0;
        qr_step<0, 1>(B, U);
        qr_step<0, 2>(B, U);
        qr_step<1, 2>(B, U);
        sig[0] = B[0][0];
        sig[1] = B[1][1];
        sig[2] = B[2][2];
    }

    // When rho_i < rho_j: swap columns i, j and negate column N (branch-free).
    template <int I, int J, int N>
    __device__ __forceinline__ static void swap_neg(T (&B)[3][3], T (&V)[3][3], T (&rho)[3]) {
        const bool sw = rho[I] < rho[J];
        // selects of (possibly negated) values: the same bits as a product by -1/+1
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const T bi = B[r][I], bj = B[r][J], vi = V[r][I], vj = V[r][J];
            B[r][I] = sw ? (N == I ? -bj : bj) : bi;
            B[r][J] = sw ? (N == J ? -bi : bi) : bj;
            V[r][I] = sw ? (N == I ? -vj : vj) : vi;
            V[r][J] = sw ? (N == J ? -vi : vi) : vj;
        }
        const T ri = rho[I], rj = rho[J];
        rho[I] = sw ? rj : ri;
        rho[J] = sw ? ri : rj;
    }

    template <int P, int Q> __device__ __forceinline__ static void qr_step(T (&B)[3][3], T (&U)[3][3]) {
        T ch, sh;
        qr_half(B[P][P], B[Q][P], ch, sh);
        const T c = ch * ch - sh * sh, s = (T)2 * ch * sh;
#pragma unroll
        for (int col = 0; col < 3; ++col) {
            const T bp = B[P][col], bq = B[Q][col];
            B[P][col] = c * bp + s * bq;
            B[Q][col] = -s * bp + c * bq;
        }
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const T up = U[r][P], uq = U[r][Q];
            U[r][P] = c * up + s * uq;
            U[r][Q] = -s * up + c * uq;
        }
    }
};

template <class T> __device__ __forceinline__ T det3(const T (&m)[3][3]) {
    return m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2]) - m[0][1] * (m[1][0] * m[2][2] - m[2][0] * m[1][2]) +
           m[0][2] * (m[1][0] * m[2][1] - m[2][0] * m[1][1]);
}

// R = U V^T from the TR1690 SVD, column 2 negated if det R < 0.
template <class T> __device__ __forceinline__ void rotation_tr1690(const T (&F)[3][3], T (&P)[3][3]) {
    T U[3][3], sig[3], V[3][3];
    Svd3<T>::run(F, U, sig, V);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) P[r][c] = U[r][0] * V[c][0] + U[r][1] * V[c][1] + U[r][2] * V[c][2];
    const bool flip = det3(P) < (T)0;
#pragma unroll
    for (int r = 0; r < 3; ++r) P[r][2] = flip ? -P[r][2] : P[r][2];
}

// Rotation factor of the polar decomposition F = R S by Newton's iteration
// R <- (R + R^-T) / 2 from R = F (Higham 1986, unscaled): in the SVD basis
// each singular value goes s -> (s + 1/s) / 2, which converges
// quadratically to 1, so R -> U V^T, the same rotation TR1690 gives when
// det F > 0.  R^-T is the cofactor matrix over the determinant.  A step that
// moves no entry by more than 5e-4 leaves an error below ~1.3e-7 (the error
// after a step is about half the square of the step before it).  A
// wavefront stops after the 4th step when every lane has settled so (or
// will take the SVD); otherwise it runs a 5th, which settles singular values
// in about [0.27, 3.7].  Returns false (R unusable) when det F <= 0 or the
// last step moved an entry by more than 5e-4 or was not finite.
__device__ __forceinline__ bool rotation_newton(const float (&F)[3][3], float (&R)[3][3]) {
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) R[r][c] = F[r][c];
    bool det_ok = true, conv = false;
#pragma unroll
    for (int it = 0; it < 5; ++it) {
        float C[3][3];
        C[0][0] = R[1][1] * R[2][2] - R[1][2] * R[2][1];
        C[0][1] = R[1][2] * R[2][0] - R[1][0] * R[2][2];
        C[0][2] = R[1][0] * R[2][1] - R[1][1] * R[2][0];
        C[1][0] = R[0][2] * R[2][1] - R[0][1] * R[2][2];
        C[1][1] = R[0][0] * R[2][2] - R[0][2] * R[2][0];
        C[1][2] = R[0][1] * R[2][0] - R[0][0] * R[2][1];
        C[2][0] = R[0][1] * R[1][2] - R[0][2] * R[1][1];
        C[2][1] = R[0][2] * R[1][0] - R[0][0] * R[1][2];
        C[2][2] = R[0][0] * R[1][1] - R[0][1] * R[1][0];
        const float det = R[0][0] * C[0][0] + R[0][1] * C[0][1] + R[0][2] * C[0][2];
        if (it == 0) det_ok = det > 0.0f;
        const float h = 0.5f * __builtin_amdgcn_rcpf(det);
        conv = true;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float n = fmaf(0.5f, R[r][c], h * C[r][c]);
                if (it >= 3) conv = conv && fabsf(n - R[r][c]) <= 5e-4f;  // false for NaN
                R[r][c] = n;
            }
        if (it == 3 && __all(conv || !det_ok)) break;  // wavefront-uniform
    }
    return det_ok && conv;
}

// Projection of computeLocal (src/simulation/solver/projective/pdUtil.cu:112-125):
// R = U V^T, negate column 2 if det R < 0, and R - F in Jacobi (residual) mode.
// `polar` (FP32 only): take the rotation from Newton's polar iteration and
// fall back to the TR1690 SVD for the elements it does not settle (inverted,
// degenerate or strongly stretched).  Where TR1690's four sweeps have
// converged the two agree to rounding; on strongly sheared elements the
// sweeps can stop up to ~1e-3 short of the polar factor, which Newton's
// rotation is to FP32 rounding -- so `polar` is exact where the reference's
// SVD is not, and differs from it there.
template <class T>
__device__ __forceinline__ void pd_projection(const T (&F)[3][3], bool residual_form, T (&P)[3][3],
                                              bool polar = false) {
    if constexpr (sizeof(T) == 4) {
        if (!polar || !rotation_newton(F, P)) rotation_tr1690(F, P);
    } else {
        rotation_tr1690(F, P);
    }
    if (residual_form) {
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) P[r][c] -= F[r][c];
    }
}

}  // namespace sbs
