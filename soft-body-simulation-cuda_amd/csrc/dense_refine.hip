// Dense small-body solver on gfx950: per-body inverse fragments applied with
// MFMA (v_mfma_f32_16x16x4_f32 or v_mfma_f32_16x16x32_bf16, FP32
// accumulation), FP64 iterative refinement against the assembled sparse
// matrix.
//
// MFMA operand layout (one 16x16 output tile per wavefront, lane l):
//   FP32 16x16x4 : A[row l%16][k l/16],          B[k l/16][col l%16]
//   BF16 16x16x32: A[row l%16][k 8(l/16)+j],     B[k 8(l/16)+j][col l%16], j < 8
//   output       : C[row 4(l/16)+j][col l%16],   j < 4
// The inverse is stored pre-split into these fragments (one contiguous 64-lane
// fragment per (16-row block, k block)), so every operand load is one
// coalesced 256- or 1024-byte read.  The right-hand side occupies columns
// 0..2 of B (x, y, z); the other 13 columns are zero.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>

#include "dense_factor.hpp"
#include "dense_refine.hpp"

namespace sbs {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kWaves = 4;      // wavefronts per workgroup of the correction kernel
constexpr int kSliceD = 64;    // sliced-ELL slice height of the engine's matrix
constexpr int kBSlot = 0;      // norms: slot 0 holds max |D^-1 b|; slot k+1 holds max |D^-1 r_k|
constexpr int kTpbD = 256;
constexpr size_t kFusedLdsMax = 64 * 1024;  // one-workgroup-per-body solve when a body's LDS fits

inline uint16_t to_bf16(float f) {  // round to nearest even
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ double u2d(unsigned long long u) { return __longlong_as_double((long long)u); }

// Converged before correction k: max |D^-1 r_k| <= tol max |D^-1 b| for x, y, z.
__device__ __forceinline__ bool converged(const unsigned long long* norms, int k, double tol) {
    for (int c = 0; c < 3; ++c)
        if (!(u2d(norms[4 * (k + 1) + c]) <= tol * u2d(norms[4 * kBSlot + c]))) return false;
    return true;
}

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned lo = __shfl_xor((unsigned)v, off, 64), hi = __shfl_xor((unsigned)(v >> 32), off, 64);
        const unsigned long long o = ((unsigned long long)hi << 32) | lo;
        v = o > v ? o : v;
    }
    return v;
}

// Slot of row v, component c, of the packed right-hand side.
template <int MODE>
__device__ __forceinline__ long long rhs_pos(long long off, int kl, int c) {
    if (MODE == 0) return off + (long long)(kl / 4) * 64 + c + 16 * (kl % 4);
    return off + (long long)(kl / 32) * 512 + (long long)(c + 16 * ((kl % 32) / 8)) * 8 + (kl % 8);
}

template <int MODE> __device__ __forceinline__ void put_rhs(void* rhs, long long pos, double r) {
    if (MODE == 0) reinterpret_cast<float*>(rhs)[pos] = (float)r;
    else reinterpret_cast<__bf16*>(rhs)[pos] = (__bf16)(float)r;
}

// Per component: max over the block's rows of |D^-1 r|, one atomic per wavefront.
__device__ __forceinline__ void max3(const double (&s)[3], unsigned long long* slot) {
    for (int c = 0; c < 3; ++c) {
        const unsigned long long m = wave_max_u64((unsigned long long)__double_as_longlong(s[c]));
        if ((threadIdx.x & 63) == 0) atomicMax(&slot[c], m);
    }
}

__global__ void zero_u64_kernel(int n, unsigned long long* p) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0ull;
}

// x = 0, r_0 = b, norms of D^-1 b (slot 0 and slot 1).
template <class T, int MODE>
__global__ void refine_init_kernel(int n, const T* b, const T* diag, const int* row_body, const int* body_row0,
                                   const long long* rhs_off, void* rhs, double* x64, unsigned long long* norms,
                                   const int* skip) {
    if (skip && *skip) return;
    const int v = blockIdx.x * kTpbD + threadIdx.x;
    double s[3] = {0, 0, 0};
    if (v < n) {
        const int bd = row_body[v];
        const int kl = v - body_row0[bd];
        const double d = (double)diag[v];
        for (int c = 0; c < 3; ++c) {
            const double r = (double)b[3 * (size_t)v + c];
            x64[3 * (size_t)v + c] = 0.0;
            put_rhs<MODE>(rhs, rhs_pos<MODE>(rhs_off[bd], kl, c), r);
            s[c] = fabs(r / d);
        }
    }
    max3(s, norms + 4 * kBSlot);
    max3(s, norms + 4 * 1);
}

// Correction k: x += A~^-1 r_k, one 16-row block of one body per wavefront.
template <int MODE>
__global__ void __launch_bounds__(64 * kWaves)
    refine_correct_kernel(int k, int nitems, const int2* items, const int* body_row0, const int* body_n,
                          const int* body_p, const long long* tile_off, const long long* rhs_off, const void* tiles,
                          const void* rhs, double* x64, const unsigned long long* norms, double tol,
                          const int* skip) {
    if (skip && *skip) return;
    if (converged(norms, k, tol)) return;
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * kWaves + (threadIdx.x >> 6);
    if (wave >= nitems) return;  // whole wavefronts; no barrier below
    const int2 it = items[wave];
    const int bd = it.x, rb = it.y;
    const int P = body_p[bd];
    floatx4 acc0 = {0, 0, 0, 0}, acc1 = {0, 0, 0, 0};
    if (MODE == 0) {
        const int nk = P / 4;  // a multiple of 8
        const float* ta = reinterpret_cast<const float*>(tiles) + tile_off[bd] + (long long)rb * nk * 64;
        const float* tb = reinterpret_cast<const float*>(rhs) + rhs_off[bd];
        for (int q = 0; q < nk; q += 2) {
            acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(ta[q * 64 + lane], tb[q * 64 + lane], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(ta[(q + 1) * 64 + lane], tb[(q + 1) * 64 + lane], acc1, 0, 0,
                                                        0);
        }
    } else {
        const int nk = P / 32;
        const bf16x8* ta = reinterpret_cast<const bf16x8*>(tiles) + (tile_off[bd] + (long long)rb * nk * 512) / 8;
        const bf16x8* tb = reinterpret_cast<const bf16x8*>(rhs) + rhs_off[bd] / 8;
        int q = 0;
        for (; q + 1 < nk; q += 2) {
            acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ta[q * 64 + lane], tb[q * 64 + lane], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ta[(q + 1) * 64 + lane], tb[(q + 1) * 64 + lane], acc1, 0,
                                                           0, 0);
        }
        if (q < nk) acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ta[q * 64 + lane], tb[q * 64 + lane], acc0, 0, 0, 0);
    }
    const int col = lane & 15;
    if (col >= 3) return;
    const int nb = body_n[bd], r0 = body_row0[bd];
    for (int j = 0; j < 4; ++j) {
        const int row = rb * 16 + 4 * (lane >> 4) + j;
        if (row < nb) x64[3 * (size_t)(r0 + row) + col] += (double)(acc0[j] + acc1[j]);
    }
}

// r_{k+1} = b - A x in FP64 (matrix values of the solver precision), packed
// for the next correction, with max |D^-1 r_{k+1}| into slot k+2.
template <class T, int MODE>
__global__ void refine_residual_kernel(int k, EllView<T> A, const T* b, const int* row_body, const int* body_row0,
                                       const long long* rhs_off, void* rhs, const double* x64,
                                       unsigned long long* norms, double tol, const int* skip) {
    if (skip && *skip) return;
    if (converged(norms, k, tol)) return;
    const int v = blockIdx.x * kTpbD + threadIdx.x;
    double s[3] = {0, 0, 0};
    if (v < A.n) {
        const double d = (double)A.diag[v];
        double ax[3];
        for (int c = 0; c < 3; ++c) ax[c] = d * x64[3 * (size_t)v + c];
        const int sl = v / kSliceD, lane = v % kSliceD;
        const size_t base = (size_t)A.slice_ptr[sl] + lane;
        const int width = A.slice_w[sl];
        for (int e = 0; e < width; ++e) {
            const size_t pos = base + (size_t)e * kSliceD;
            const uint32_t u = A.ecol[pos];
            const double m = (double)A.eval[pos];
            for (int c = 0; c < 3; ++c) ax[c] += m * x64[3 * (size_t)u + c];
        }
        const int bd = row_body[v];
        const int kl = v - body_row0[bd];
        for (int c = 0; c < 3; ++c) {
            const double r = (double)b[3 * (size_t)v + c] - ax[c];
            put_rhs<MODE>(rhs, rhs_pos<MODE>(rhs_off[bd], kl, c), r);
            s[c] = fabs(r / d);
        }
    }
    max3(s, norms + 4 * (k + 2));
}

template <class T> __global__ void refine_store_kernel(int n3, const double* x64, T* x, const int* skip) {
    if (skip && *skip) return;
    const int i = blockIdx.x * kTpbD + threadIdx.x;
    if (i < n3) x[i] = (T)x64[i];
}

// First k with the stop test met: corrections applied, and the ratio reached.
__global__ void refine_report_kernel(int max_iter, const unsigned long long* norms, double tol, double* out,
                                     const int* skip) {
    if (threadIdx.x != 0) return;
    if (skip && *skip) return;
    int it = -1 - max_iter;
    int last = max_iter;
    for (int k = 0; k <= max_iter; ++k)
        if (converged(norms, k, tol)) {
            it = k;
            last = k;
            break;
        }
    double ratio = 0;
    for (int c = 0; c < 3; ++c) {
        const double bb = u2d(norms[4 * kBSlot + c]);
        const double rr = u2d(norms[4 * (last + 1) + c]);
        ratio = fmax(ratio, bb > 0 ? rr / bb : rr);
    }
    out[0] = (double)it;
    out[1] = ratio;
}

// ---------------------------------------------------------------------------
// Whole solve per body in one workgroup (bodies whose x, residual fragments
// and scratch fit the LDS): x (FP64) and the residual's MFMA B fragments stay
// in LDS across the corrections, each body stops on its own test, and there
// is no cross-workgroup reduction and one launch per solve.  A correction
// splits the body's (16-row block, K segment) products over the 16
// wavefronts -- inverse fragments streamed from HBM 8 (BF16: 4) per lane in
// flight, B fragments from LDS -- and sums the segments' partial products in
// a fixed order; the residual r = b - A x is formed row by row from the
// sliced-ELL matrix and x in LDS.

constexpr int kFusedThreads = 1024;
constexpr int kFusedWaves = kFusedThreads / 64;

template <int MODE> __host__ __device__ constexpr size_t fused_lds_bytes(int P, int nseg) {
    // x: 3 P doubles; 3 doubles per wavefront of reduction scratch; B
    // fragments: 16 P values (FP32 or BF16); partial products of the K
    // segments: 3 P floats each (none with one segment)
    return (3 * (size_t)P + 3 * kFusedWaves) * sizeof(double) + 16 * (size_t)P * (MODE == 0 ? 4 : 2) +
           (nseg > 1 ? (size_t)nseg * 3 * P * sizeof(float) : 0);
}

__device__ __forceinline__ double wave_max_f64(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// Maximum over the workgroup of three non-negative values (NaN counts as
// infinity, so a NaN residual never passes the stop test).
__device__ __forceinline__ void block_max3(const double (&s)[3], double* red, double (&out)[3]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double v = wave_max_f64(s[c] == s[c] ? s[c] : INFINITY);
        if (lane == 0) red[3 * wave + c] = v;
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double m = red[c];
        for (int w = 1; w < kFusedWaves; ++w) m = fmax(m, red[3 * w + c]);
        out[c] = m;
    }
    __syncthreads();
}

template <class T, int MODE>
__global__ void __launch_bounds__(kFusedThreads)
    dense_body_solve_kernel(EllView<T> A, const T* __restrict__ b, T* __restrict__ x, const int* __restrict__ body_row0,
                            const int* __restrict__ body_n, const int* __restrict__ body_p,
                            const long long* __restrict__ tile_off, const void* __restrict__ tiles, int nseg,
                            double tol, int max_iter, int* __restrict__ body_iters, double* __restrict__ body_ratio,
                            const int* skip, int warm) {
    if (skip && *skip) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int bd = blockIdx.x;
    const int r0 = body_row0[bd], nb = body_n[bd], P = body_p[bd];
    double* x64 = reinterpret_cast<double*>(smem_raw);  // [row][xyz]
    double* red = x64 + 3 * (size_t)P;
    void* rhs = red + 3 * kFusedWaves;
    float* part = reinterpret_cast<float*>(reinterpret_cast<unsigned char*>(rhs) + 16 * (size_t)P * (MODE == 0 ? 4 : 2));
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // x starts at 0, or with `warm` at the caller's x (the previous iterate)
    for (int i = tid; i < 3 * P; i += kFusedThreads)
        x64[i] = (warm && i < 3 * nb) ? (double)x[3 * (size_t)r0 + i] : 0.0;
    const int words = MODE == 0 ? 16 * P : 8 * P;  // B fragments as 32-bit words; padding stays 0
    for (int i = tid; i < words; i += kFusedThreads) reinterpret_cast<uint32_t*>(rhs)[i] = 0u;
    __syncthreads();
    // r = b - A x (FP64) packed as the next correction's B operand; s = max |r / d|
    auto residual = [&](double (&s)[3]) {
        s[0] = s[1] = s[2] = 0;
        for (int kl = tid; kl < nb; kl += kFusedThreads) {
            const int v = r0 + kl;
            const double d = (double)A.diag[v];
            double ax[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) ax[c] = d * x64[3 * kl + c];
            const int sl = v / kSliceD, ln = v % kSliceD;
            const size_t base = (size_t)A.slice_ptr[sl] + ln;
            const int width = A.slice_w[sl];
            // the row's entries 8 at a time, requested together, summed in order
            for (int e0 = 0; e0 < width; e0 += 8) {
                int uu[8];
                T mm[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const size_t pos = base + (size_t)(e0 + j < width ? e0 + j : width - 1) * kSliceD;
                    uu[j] = (int)A.ecol[pos] - r0;  // the body's own rows (padding: the row itself, value 0)
                    mm[j] = A.eval[pos];
                }
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    if (e0 + j < width && (unsigned)uu[j] < (unsigned)nb) {
                        const double m = (double)mm[j];
#pragma unroll
                        for (int c = 0; c < 3; ++c) ax[c] += m * x64[3 * uu[j] + c];
                    }
                }
            }
            const T bv[3] = {b[3 * (size_t)v], b[3 * (size_t)v + 1], b[3 * (size_t)v + 2]};
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double r = (double)bv[c] - ax[c];
                put_rhs<MODE>(rhs, rhs_pos<MODE>(0, kl, c), r);
                s[c] = fmax(s[c], fabs(r / d));
            }
        }
    };
    double s[3] = {0, 0, 0};
    for (int kl = tid; kl < nb; kl += kFusedThreads) {
        const int v = r0 + kl;
        const double d = (double)A.diag[v];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double r = (double)b[3 * (size_t)v + c];
            if (!warm) put_rhs<MODE>(rhs, rhs_pos<MODE>(0, kl, c), r);
            s[c] = fmax(s[c], fabs(r / d));
        }
    }
    double bn[3], rn[3];
    block_max3(s, red, bn);  // the stop test is relative to b, from either start
    if (warm) {
        residual(s);
        block_max3(s, red, rn);
    } else {
        rn[0] = bn[0]; rn[1] = bn[1]; rn[2] = bn[2];
    }
    const int nrb = P / 16, nk = MODE == 0 ? P / 4 : P / 32, seg = (nk + nseg - 1) / nseg;
    int k = 0;
    bool ok = false;
    for (;; ++k) {
        ok = rn[0] <= tol * bn[0] && rn[1] <= tol * bn[1] && rn[2] <= tol * bn[2];
        if (ok || k == max_iter) break;
        // x += A~^-1 r: (row block, K segment) products
        for (int item = wave; item < nrb * nseg; item += kFusedWaves) {
            const int rb = item / nseg, sg = item - rb * nseg;
            const int q0 = sg * seg, q1 = min(nk, q0 + seg);
            floatx4 acc0 = {0, 0, 0, 0}, acc1 = {0, 0, 0, 0};
            if (MODE == 0) {
                const float* ta = reinterpret_cast<const float*>(tiles) + tile_off[bd] + (long long)rb * nk * 64;
                const float* tb = reinterpret_cast<const float*>(rhs);
                int q = q0;
                for (; q + 8 <= q1; q += 8) {
                    float fa[8], fb[8];
#pragma unroll
                    for (int j = 0; j < 8; ++j) fa[j] = ta[(q + j) * 64 + lane];
#pragma unroll
                    for (int j = 0; j < 8; ++j) fb[j] = tb[(q + j) * 64 + lane];
#pragma unroll
                    for (int j = 0; j < 8; j += 2) {
                        acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[j], fb[j], acc0, 0, 0, 0);
                        acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[j + 1], fb[j + 1], acc1, 0, 0, 0);
                    }
                }
                for (; q < q1; ++q)
                    acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(ta[q * 64 + lane], tb[q * 64 + lane], acc0, 0, 0, 0);
            } else {
                const bf16x8* ta =
                    reinterpret_cast<const bf16x8*>(tiles) + (tile_off[bd] + (long long)rb * nk * 512) / 8;
                const bf16x8* tb = reinterpret_cast<const bf16x8*>(rhs);
                int q = q0;
                for (; q + 4 <= q1; q += 4) {
                    bf16x8 fa[4], fb[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) fa[j] = ta[(q + j) * 64 + lane];
#pragma unroll
                    for (int j = 0; j < 4; ++j) fb[j] = tb[(q + j) * 64 + lane];
#pragma unroll
                    for (int j = 0; j < 4; j += 2) {
                        acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[j], fb[j], acc0, 0, 0, 0);
                        acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[j + 1], fb[j + 1], acc1, 0, 0, 0);
                    }
                }
                for (; q < q1; ++q)
                    acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ta[q * 64 + lane], tb[q * 64 + lane], acc0, 0, 0, 0);
            }
            const int col = lane & 15;
            if (col < 3) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int row = rb * 16 + 4 * (lane >> 4) + j;
                    if (row < nb) {
                        if (nseg == 1) x64[3 * row + col] += (double)(acc0[j] + acc1[j]);
                        else part[(size_t)sg * 3 * P + 3 * row + col] = acc0[j] + acc1[j];
                    }
                }
            }
        }
        __syncthreads();
        if (nseg > 1) {  // the segments' partial products, summed in segment order
            for (int i = tid; i < 3 * nb; i += kFusedThreads) {
                double sum = 0;
                for (int sg = 0; sg < nseg; ++sg) sum += (double)part[(size_t)sg * 3 * P + i];
                x64[i] += sum;
            }
            __syncthreads();
        }
        residual(s);
        block_max3(s, red, rn);
    }
    for (int i = tid; i < 3 * nb; i += kFusedThreads) x[3 * (size_t)r0 + i] = (T)x64[i];
    if (tid == 0) {
        body_iters[bd] = ok ? k : -1 - max_iter;
        double ratio = 0;
        for (int c = 0; c < 3; ++c) ratio = fmax(ratio, bn[c] > 0 ? rn[c] / bn[c] : rn[c]);
        body_ratio[bd] = ratio;
    }
}

// The worst body: most corrections (-1 - max_iter if one stopped there) and
// the largest residual ratio.
__global__ void dense_body_report_kernel(int nb, int max_iter, const int* iters, const double* ratio, double* out,
                                         const int* skip) {
    if (skip && *skip) return;
    __shared__ int worst_it;
    __shared__ int failed;
    __shared__ double worst_ratio[256];
    if (threadIdx.x == 0) {
        worst_it = 0;
        failed = 0;
    }
    __syncthreads();
    double rmax = 0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        const int it = iters[i];
        if (it < 0) atomicOr(&failed, 1);
        else atomicMax(&worst_it, it);
        rmax = fmax(rmax, ratio[i]);
    }
    worst_ratio[threadIdx.x] = rmax;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0;
        for (int i = 0; i < (int)blockDim.x; ++i) r = fmax(r, worst_ratio[i]);
        out[0] = failed ? (double)(-1 - max_iter) : (double)worst_it;
        out[1] = r;
    }
}

inline unsigned blocks_of(long long n, int tpb) { return (unsigned)((n + tpb - 1) / tpb); }

}  // namespace

template <class T> DenseRefinedSolver<T>::~DenseRefinedSolver() {
    if (host_report_) (void)hipHostFree(host_report_);
}

template <class T>
void DenseRefinedSolver<T>::setup(const EllView<T>& A, const std::vector<int>& rowptr, const std::vector<int>& col,
                                  const std::vector<T>& val, DensePrecision prec, hipStream_t stream) {
    A_ = A;
    prec_ = prec;
    stream_ = stream;
    n_ = A.n;
    SBS_REQUIRE(n_ > 0 && rowptr.size() == (size_t)n_ + 1, "dense solver: inconsistent matrix");
    int nc = 0;
    bool contiguous = false;
    const std::vector<int> comp = row_components(n_, rowptr, col, &nc, &contiguous);
    SBS_REQUIRE(contiguous, "dense solver: bodies must be contiguous vertex ranges");
    std::vector<int> row0(nc, 0), nrows(nc, 0), pad(nc, 0);
    for (int i = 0; i < n_; ++i) {
        if (nrows[comp[i]] == 0) row0[comp[i]] = i;
        nrows[comp[i]]++;
    }
    nb_ = nc;
    max_n_ = *std::max_element(nrows.begin(), nrows.end());
    SBS_REQUIRE(max_n_ <= 4096, "dense solver: a body exceeds 4096 vertices (use a sparse global solver)");
    fused_lds_ = 0;  // chosen below once the padded sizes are known
    std::vector<long long> toff(nc + 1, 0), roff(nc + 1, 0);
    std::vector<int2> items;
    for (int b = 0; b < nc; ++b) {
        pad[b] = (nrows[b] + 31) / 32 * 32;
        toff[b + 1] = toff[b] + (long long)pad[b] * pad[b];
        roff[b + 1] = roff[b] + 16ll * pad[b];
        for (int rb = 0; rb < pad[b] / 16; ++rb) items.push_back(make_int2(b, rb));
    }
    nitems_ = (int)items.size();
    {
        // the most K segments (4, 2, 1) whose partial products still fit
        const int pmax = *std::max_element(pad.begin(), pad.end());
        fused_lds_ = 0;
        for (int ns : {4, 2, 1}) {
            const size_t lds = prec == DensePrecision::FP32 ? fused_lds_bytes<0>(pmax, ns) : fused_lds_bytes<1>(pmax, ns);
            if (lds <= kFusedLdsMax) {
                fused_lds_ = lds;
                fused_nseg_ = ns;
                break;
            }
        }
    }
    const size_t esz = prec == DensePrecision::FP32 ? 4 : 2;
    tile_bytes_ = (size_t)toff[nc] * esz;
    std::vector<unsigned char> tiles(tile_bytes_, 0);
    // per body: FP64 Cholesky and inverse, split into MFMA fragments (host
    // threads over bodies; each body writes only its own fragments)
    std::vector<std::string> errs(nc);
    std::atomic<int> next{0};
    auto work = [&]() {
        for (int b = next++; b < nc; b = next++) try {
            std::vector<double> a = dense_block(row0[b], nrows[b], rowptr, col, val), inv;
            dense_cholesky_inverse(nrows[b], a, inv);
            const int n = nrows[b], P = pad[b];
            if (prec == DensePrecision::FP32) {
                float* t = reinterpret_cast<float*>(tiles.data()) + toff[b];
                for (int rb = 0; rb < P / 16; ++rb)
                    for (int q = 0; q < P / 4; ++q)
                        for (int l = 0; l < 64; ++l) {
                            const int r = rb * 16 + l % 16, k = q * 4 + l / 16;
                            t[((long long)rb * (P / 4) + q) * 64 + l] = (r < n && k < n) ? (float)inv[(size_t)r * n + k] : 0.f;
                        }
            } else {
                uint16_t* t = reinterpret_cast<uint16_t*>(tiles.data()) + toff[b];
                for (int rb = 0; rb < P / 16; ++rb)
                    for (int q = 0; q < P / 32; ++q)
                        for (int l = 0; l < 64; ++l)
                            for (int j = 0; j < 8; ++j) {
                                const int r = rb * 16 + l % 16, k = q * 32 + 8 * (l / 16) + j;
                                t[((long long)rb * (P / 32) + q) * 512 + l * 8 + j] =
                                    (r < n && k < n) ? to_bf16((float)inv[(size_t)r * n + k]) : (uint16_t)0;
                            }
            }
        } catch (const std::exception& e) {
            errs[b] = e.what();
        }
    };
    const int nthreads = std::max(1, std::min<int>(nc, (int)std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (int t = 1; t < nthreads; ++t) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
    for (int b = 0; b < nc; ++b) SBS_REQUIRE(errs[b].empty(), "dense solver, body " + std::to_string(b) + ": " + errs[b]);
    std::vector<int> row_body(n_);
    for (int i = 0; i < n_; ++i) row_body[i] = comp[i];
    body_row0_.alloc(nc);
    body_n_.alloc(nc);
    body_p_.alloc(nc);
    tile_off_.alloc(nc);
    rhs_off_.alloc(nc);
    items_.alloc(std::max(nitems_, 1));
    row_body_.alloc(n_);
    tiles_.alloc(std::max<size_t>(tile_bytes_, 1));
    rhs_.alloc(std::max<size_t>((size_t)roff[nc] * esz, 1));
    x64_.alloc(3 * (size_t)n_);
    SBS_HIP(hipMemcpy(body_row0_.get(), row0.data(), nc * sizeof(int), hipMemcpyHostToDevice));
    SBS_HIP(hipMemcpy(body_n_.get(), nrows.data(), nc * sizeof(int), hipMemcpyHostToDevice));
    SBS_HIP(hipMemcpy(body_p_.get(), pad.data(), nc * sizeof(int), hipMemcpyHostToDevice));
    SBS_HIP(hipMemcpy(tile_off_.get(), toff.data(), nc * sizeof(long long), hipMemcpyHostToDevice));
    SBS_HIP(hipMemcpy(rhs_off_.get(), roff.data(), nc * sizeof(long long), hipMemcpyHostToDevice));
    SBS_HIP(hipMemcpy(items_.get(), items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice));
    SBS_HIP(hipMemcpy(row_body_.get(), row_body.data(), n_ * sizeof(int), hipMemcpyHostToDevice));
    SBS_HIP(hipMemcpy(tiles_.get(), tiles.data(), tile_bytes_, hipMemcpyHostToDevice));
    // padding rows and the 13 unused columns of the packed right-hand side stay zero
    SBS_HIP(hipMemset(rhs_.get(), 0, rhs_.bytes()));
    report_.alloc(2);
    body_iters_.alloc(nc);
    body_ratio_.alloc(nc);
    if (!host_report_) SBS_HIP(hipHostMalloc((void**)&host_report_, 2 * sizeof(double), hipHostMallocDefault));
}

template <class T>
void DenseRefinedSolver<T>::solve(const T* b, T* x, double tol, int max_iter, const int* skip, bool warm) {
    SBS_REQUIRE(nb_ > 0, "dense solver: not set up");
    SBS_REQUIRE(max_iter >= 1 && max_iter <= 64, "dense solver: max_iter must be in [1, 64]");
    const int need = 4 * (max_iter + 3);
    if (need > slots_) {
        norms_.alloc(need);
        slots_ = need;
    }
    last_max_iter_ = max_iter;
    if (fused_lds_) {
        const bool f32 = prec_ == DensePrecision::FP32;
        auto kern = f32 ? dense_body_solve_kernel<T, 0> : dense_body_solve_kernel<T, 1>;
        hipLaunchKernelGGL(kern, dim3(nb_), dim3(kFusedThreads), fused_lds_, stream_, A_, b, x,
                           (const int*)body_row0_.get(), (const int*)body_n_.get(), (const int*)body_p_.get(),
                           (const long long*)tile_off_.get(), (const void*)tiles_.get(), fused_nseg_, tol, max_iter,
                           body_iters_.get(), body_ratio_.get(), skip, (int)warm);
        hipLaunchKernelGGL(dense_body_report_kernel, dim3(1), dim3(256), 0, stream_, nb_, max_iter,
                           (const int*)body_iters_.get(), (const double*)body_ratio_.get(), report_.get(), skip);
        SBS_HIP(hipGetLastError());
        return;
    }
    hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(256), 0, stream_, slots_, norms_.get());
    const unsigned vb = blocks_of(n_, kTpbD);
    const bool f32 = prec_ == DensePrecision::FP32;
    if (f32)
        hipLaunchKernelGGL((refine_init_kernel<T, 0>), dim3(vb), dim3(kTpbD), 0, stream_, n_, b, A_.diag,
                           (const int*)row_body_.get(), (const int*)body_row0_.get(), (const long long*)rhs_off_.get(),
                           (void*)rhs_.get(), x64_.get(), norms_.get(), skip);
    else
        hipLaunchKernelGGL((refine_init_kernel<T, 1>), dim3(vb), dim3(kTpbD), 0, stream_, n_, b, A_.diag,
                           (const int*)row_body_.get(), (const int*)body_row0_.get(), (const long long*)rhs_off_.get(),
                           (void*)rhs_.get(), x64_.get(), norms_.get(), skip);
    const unsigned cb = blocks_of(nitems_, kWaves);
    for (int k = 0; k < max_iter; ++k) {
        if (f32) {
            hipLaunchKernelGGL(refine_correct_kernel<0>, dim3(cb), dim3(64 * kWaves), 0, stream_, k, nitems_,
                               (const int2*)items_.get(), (const int*)body_row0_.get(), (const int*)body_n_.get(),
                               (const int*)body_p_.get(), (const long long*)tile_off_.get(),
                               (const long long*)rhs_off_.get(), (const void*)tiles_.get(), (const void*)rhs_.get(),
                               x64_.get(), (const unsigned long long*)norms_.get(), tol, skip);
            hipLaunchKernelGGL((refine_residual_kernel<T, 0>), dim3(vb), dim3(kTpbD), 0, stream_, k, A_, b,
                               (const int*)row_body_.get(), (const int*)body_row0_.get(),
                               (const long long*)rhs_off_.get(), (void*)rhs_.get(), (const double*)x64_.get(),
                               norms_.get(), tol, skip);
        } else {
            hipLaunchKernelGGL(refine_correct_kernel<1>, dim3(cb), dim3(64 * kWaves), 0, stream_, k, nitems_,
                               (const int2*)items_.get(), (const int*)body_row0_.get(), (const int*)body_n_.get(),
                               (const int*)body_p_.get(), (const long long*)tile_off_.get(),
                               (const long long*)rhs_off_.get(), (const void*)tiles_.get(), (const void*)rhs_.get(),
                               x64_.get(), (const unsigned long long*)norms_.get(), tol, skip);
            hipLaunchKernelGGL((refine_residual_kernel<T, 1>), dim3(vb), dim3(kTpbD), 0, stream_, k, A_, b,
                               (const int*)row_body_.get(), (const int*)body_row0_.get(),
                               (const long long*)rhs_off_.get(), (void*)rhs_.get(), (const double*)x64_.get(),
                               norms_.get(), tol, skip);
        }
    }
    hipLaunchKernelGGL(refine_store_kernel<T>, dim3(blocks_of(3ll * n_, kTpbD)), dim3(kTpbD), 0, stream_, 3 * n_,
                       (const double*)x64_.get(), x, skip);
    hipLaunchKernelGGL(refine_report_kernel, dim3(1), dim3(64), 0, stream_, max_iter,
                       (const unsigned long long*)norms_.get(), tol, report_.get(), skip);
    SBS_HIP(hipGetLastError());
}

template <class T> void DenseRefinedSolver<T>::last_result(int* iterations, double* ratio) {
    SBS_HIP(hipMemcpyAsync(host_report_, report_.get(), 2 * sizeof(double), hipMemcpyDeviceToHost, stream_));
    SBS_HIP(hipStreamSynchronize(stream_));
    if (iterations) *iterations = (int)host_report_[0];
    if (ratio) *ratio = host_report_[1];
}

template class DenseRefinedSolver<float>;
template class DenseRefinedSolver<double>;

}  // namespace sbs
