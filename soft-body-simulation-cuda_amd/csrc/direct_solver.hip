// Band Cholesky factor (host, FP64) and blocked triangular solves (gfx950).
#include "band_factor.hpp"
#include "direct_solver.hpp"

namespace sbs {

namespace {

constexpr int kBlk = 64;       // rows per block = one wavefront
constexpr int kTileLd = 65;    // padded tile stride: conflict-free row and column reads
constexpr int kThreads = 256;  // 4 wavefronts
constexpr size_t kLdsBudget = 64 * 1024;

template <class T> size_t lds_bytes(int n) {
    return (3 * (size_t)n + (size_t)kBlk * kTileLd + 4 * (size_t)kBlk * 3) * sizeof(T);
}

template <class T>
__global__ void __launch_bounds__(kThreads) band_solve_kernel(int n, int bw, const T* __restrict__ Ld,
                                                              const int* __restrict__ perm, const T* __restrict__ b,
                                                              T* __restrict__ x, const int* __restrict__ skip) {
    if (skip && *skip) return;
    extern __shared__ unsigned char smem_raw[];
    T* ys = reinterpret_cast<T*>(smem_raw);
    T* tile = ys + 3 * (size_t)n;
    T* part = tile + kBlk * kTileLd;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    for (int i = tid; i < n; i += kThreads) {
        const int o = perm[i];
        ys[i] = b[3 * (size_t)o];
        ys[n + i] = b[3 * (size_t)o + 1];
        ys[2 * n + i] = b[3 * (size_t)o + 2];
    }
    __syncthreads();
    const int nblk = (n + kBlk - 1) / kBlk;
    // ---- forward: L y = b ----
    for (int J = 0; J < nblk; ++J) {
        const int i0 = J * kBlk, r = i0 + lane;
        T a0 = 0, a1 = 0, a2 = 0;
        if (r < n) {
            const int dlo = r - i0 + 1, dhi = min(bw, r);
            for (int d = dlo + wid; d <= dhi; d += 4) {
                const T l = Ld[(size_t)d * n + r];
                const int j = r - d;
                a0 += l * ys[j];
                a1 += l * ys[n + j];
                a2 += l * ys[2 * n + j];
            }
        }
        part[(wid * kBlk + lane) * 3 + 0] = a0;
        part[(wid * kBlk + lane) * 3 + 1] = a1;
        part[(wid * kBlk + lane) * 3 + 2] = a2;
        for (int e = tid; e < kBlk * kBlk; e += kThreads) {
            const int l = e / kBlk, t = e % kBlk, row = i0 + l, d = l - t;
            T v = 0;
            if (row < n && d >= 0 && d <= bw) v = Ld[(size_t)d * n + row];
            tile[l * kTileLd + t] = v;
        }
        __syncthreads();
        if (wid == 0) {
            T s0 = 0, s1 = 0, s2 = 0;
            if (r < n) {
                s0 = ys[r]; s1 = ys[n + r]; s2 = ys[2 * n + r];
                for (int w = 0; w < 4; ++w) {
                    s0 -= part[(w * kBlk + lane) * 3 + 0];
                    s1 -= part[(w * kBlk + lane) * 3 + 1];
                    s2 -= part[(w * kBlk + lane) * 3 + 2];
                }
            }
            const int tmax = min(kBlk, n - i0);
            for (int t = 0; t < tmax; ++t) {
                const T inv = (T)1 / tile[t * kTileLd + t];
                const T y0 = __shfl(s0, t, 64) * inv, y1 = __shfl(s1, t, 64) * inv, y2 = __shfl(s2, t, 64) * inv;
                if (lane > t) {
                    const T l = tile[lane * kTileLd + t];
                    s0 -= l * y0; s1 -= l * y1; s2 -= l * y2;
                } else if (lane == t) {
                    s0 = y0; s1 = y1; s2 = y2;
                }
            }
            if (r < n) { ys[r] = s0; ys[n + r] = s1; ys[2 * n + r] = s2; }
        }
        __syncthreads();
    }
    // ---- backward: L^T x = y ----
    for (int J = nblk - 1; J >= 0; --J) {
        const int i0 = J * kBlk, r = i0 + lane;
        T a0 = 0, a1 = 0, a2 = 0;
        if (r < n) {
            const int dlo = i0 + kBlk - r, dhi = min(bw, n - 1 - r);
            for (int d = dlo + wid; d <= dhi; d += 4) {
                const int j = r + d;
                const T l = Ld[(size_t)d * n + j];
                a0 += l * ys[j];
                a1 += l * ys[n + j];
                a2 += l * ys[2 * n + j];
            }
        }
        part[(wid * kBlk + lane) * 3 + 0] = a0;
        part[(wid * kBlk + lane) * 3 + 1] = a1;
        part[(wid * kBlk + lane) * 3 + 2] = a2;
        for (int e = tid; e < kBlk * kBlk; e += kThreads) {
            const int l = e / kBlk, t = e % kBlk, row = i0 + l, d = l - t;
            T v = 0;
            if (row < n && d >= 0 && d <= bw) v = Ld[(size_t)d * n + row];
            tile[l * kTileLd + t] = v;
        }
        __syncthreads();
        if (wid == 0) {
            T s0 = 0, s1 = 0, s2 = 0;
            if (r < n) {
                s0 = ys[r]; s1 = ys[n + r]; s2 = ys[2 * n + r];
                for (int w = 0; w < 4; ++w) {
                    s0 -= part[(w * kBlk + lane) * 3 + 0];
                    s1 -= part[(w * kBlk + lane) * 3 + 1];
                    s2 -= part[(w * kBlk + lane) * 3 + 2];
                }
            }
            const int tmax = min(kBlk, n - i0);
            for (int t = tmax - 1; t >= 0; --t) {
                const T inv = (T)1 / tile[t * kTileLd + t];
                const T x0 = __shfl(s0, t, 64) * inv, x1 = __shfl(s1, t, 64) * inv, x2 = __shfl(s2, t, 64) * inv;
                if (lane < t) {
                    const T l = tile[t * kTileLd + lane];
                    s0 -= l * x0; s1 -= l * x1; s2 -= l * x2;
                } else if (lane == t) {
                    s0 = x0; s1 = x1; s2 = x2;
                }
            }
            if (r < n) { ys[r] = s0; ys[n + r] = s1; ys[2 * n + r] = s2; }
        }
        __syncthreads();
    }
    for (int i = tid; i < n; i += kThreads) {
        const int o = perm[i];
        x[3 * (size_t)o] = ys[i];
        x[3 * (size_t)o + 1] = ys[n + i];
        x[3 * (size_t)o + 2] = ys[2 * n + i];
    }
}

}  // namespace

template <class T> int BandCholesky<T>::max_n() {
    const size_t fixed = ((size_t)kBlk * kTileLd + 4 * (size_t)kBlk * 3) * sizeof(T);
    return (int)((kLdsBudget - fixed) / (3 * sizeof(T)));
}

template <class T>
void BandCholesky<T>::factor(int n, const std::vector<int>& rowptr, const std::vector<int>& col,
                             const std::vector<T>& val, hipStream_t stream) {
    SBS_REQUIRE(n > 0, "BandCholesky: empty matrix");
    if (n > max_n())
        throw Error(kInvalidArgument, "direct (Cholesky) global solve supports at most " + std::to_string(max_n()) +
                                          " vertices per system in this precision; use the PCG global solver");
    BandFactor f;
    try {
        f = band_factor(n, rowptr, col, val);
    } catch (const std::invalid_argument& e) {
        throw Error(kInvalidArgument, e.what());
    } catch (const std::domain_error& e) {
        throw Error(kFactorFailed, e.what());
    }
    const int bw = f.bw;
    std::vector<T> ld((size_t)(bw + 1) * n);
    for (int i = 0; i < n; ++i)
        for (int d = 0; d <= bw; ++d) ld[(size_t)d * n + i] = (T)f.Lb[(size_t)i * (bw + 1) + d];
    n_ = n;
    bw_ = bw;
    ld_.alloc(ld.size());
    perm_.alloc(n);
    SBS_HIP(hipMemcpyAsync(ld_.get(), ld.data(), ld.size() * sizeof(T), hipMemcpyHostToDevice, stream));
    SBS_HIP(hipMemcpyAsync(perm_.get(), f.perm.data(), n * sizeof(int), hipMemcpyHostToDevice, stream));
    SBS_HIP(hipStreamSynchronize(stream));
}

template <class T> void BandCholesky<T>::solve(const T* b, T* x, hipStream_t stream, const int* skip) const {
    SBS_REQUIRE(n_ > 0, "BandCholesky: solve before factor");
    hipLaunchKernelGGL(band_solve_kernel<T>, dim3(1), dim3(kThreads), lds_bytes<T>(n_), stream, n_, bw_, ld_.get(),
                       perm_.get(), b, x, skip);
    SBS_HIP(hipGetLastError());
}

template class BandCholesky<float>;
template class BandCholesky<double>;

}  // namespace sbs
