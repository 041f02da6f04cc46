// Band Cholesky factor (host, FP64) and blocked triangular solves (gfx950).
#include "band_factor.hpp"
#include "direct_solver.hpp"

namespace sbs {

namespace {

constexpr int kBlk = 64;       // rows per block = one wavefront
constexpr int kTileLd = 65;    // padded tile stride: conflict-free row and column reads
constexpr int kWaves = 8;      // the band update of a block is split over 8 wavefronts
constexpr int kThreads = 64 * kWaves;
// host factor work N bw^2 / 2 beyond which the direct modes decline (the
// band factor is computed on the host; about a minute of one core)
constexpr double kMaxFactorWork = 6e10;

template <class T> size_t lds_bytes(int ring) {
    return ((size_t)ring + (size_t)kBlk * kTileLd + (size_t)kWaves * kBlk) * sizeof(T);
}

// smallest power of two holding the bw rows a block reads plus its own 64
int ring_rows(int bw) {
    int r = kBlk;
    while (r < bw + kBlk) r *= 2;
    return r;
}

// value of `v` in lane `l` of the wavefront (l uniform): v_readlane, no LDS trip
__device__ inline float lane_value(float v, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
__device__ inline double lane_value(double v, int l) {
    const long long u = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffLL), l);
    const int hi = __builtin_amdgcn_readlane((int)(u >> 32), l);
    return __builtin_bit_cast(double, (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// One workgroup per (independent system, right-hand side component): the
// rows of system s are [sys[s], sys[s+1]) of the reordered matrix, and no
// row of one system couples to another (the factor is block diagonal), so
// the systems and the x, y, z components are solved side by side.  Rows go
// in blocks of 64:
//   * band update: the block's rows times the y (forward) or x (backward)
//     values of the bw rows before (after) it, kept in a ring of R rows in
//     LDS (R a power of two >= bw + 64); kWaves wavefronts each take every
//     kWaves-th diagonal, 4 loads in flight per lane;
//   * the block's 64x64 lower triangle is read diagonal by diagonal (64
//     consecutive rows of one diagonal per load: coalesced) into registers
//     before the band update, so its loads overlap the band's, as do the
//     right-hand side and the next block's permutation (wavefront 0);
//   * the tile is solved by wavefront 0, the solved value of each row
//     broadcast by v_readlane and scaled by the reciprocal of the diagonal.
// y is also kept in `ybuf` for the backward sweep.  LDS grows with the
// bandwidth, not with the number of rows.
template <class T>
__global__ void __launch_bounds__(kThreads) band_solve_kernel(int n, int bw, int R, int nsys, const int* __restrict__ sys,
                                                              const T* __restrict__ Ld, const int* __restrict__ perm,
                                                              const T* __restrict__ b, T* __restrict__ x,
                                                              T* __restrict__ ybuf, const int* __restrict__ skip) {
    if (skip && *skip) return;
    constexpr int kDiagPerWave = kBlk / kWaves;  // tile diagonals each wavefront loads
    extern __shared__ unsigned char smem_raw[];
    T* ring = reinterpret_cast<T*>(smem_raw);
    T* tile = ring + R;
    T* part = tile + kBlk * kTileLd;
    const int mask = R - 1;
    // the three components of a system on one XCD (block b runs on XCD
    // b % 8), so their sweeps read the system's factor through one L2
    const int xcd = blockIdx.x % 8, jx = blockIdx.x / 8;
    const int comp = jx % 3, sy = 8 * (jx / 3) + xcd;
    if (sy >= nsys) return;
    const int s0 = sys[sy], s1 = sys[sy + 1];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const size_t step = (size_t)kWaves * n;
    const int dtile = min(bw, kBlk - 1);
    // diagonals wid, wid + kWaves, ... of the tile rows i0 + lane
    auto load_tile = [&](int i0, T (&tv)[kDiagPerWave]) {
        const int row = i0 + lane;
#pragma unroll
        for (int j = 0; j < kDiagPerWave; ++j) {
            const int d = wid + j * kWaves;
            tv[j] = (row < s1 && d <= dtile && d <= lane) ? Ld[(size_t)d * n + row] : T(0);
        }
    };
    auto store_tile = [&](const T (&tv)[kDiagPerWave]) {
#pragma unroll
        for (int j = 0; j < kDiagPerWave; ++j) {
            const int d = wid + j * kWaves;
            if (d <= lane) tile[lane * kTileLd + lane - d] = (d == 0) ? T(1) / tv[j] : tv[j];
        }
    };
    // ---- forward: L y = b ----
    int pr = (wid == 0 && s0 + lane < s1) ? perm[s0 + lane] : 0;
    for (int i0 = s0; i0 < s1; i0 += kBlk) {
        const int r = i0 + lane;
        T bv = 0;
        int pn = 0;
        if (wid == 0) {
            if (r < s1) bv = b[3 * (size_t)pr + comp];
            if (r + kBlk < s1) pn = perm[r + kBlk];
        }
        T tv[kDiagPerWave];
        load_tile(i0, tv);
        T a = 0;
        if (r < s1) {
            const int dlo = r - i0 + 1 + wid, dhi = min(bw, r - s0);
            const T* L = Ld + (size_t)dlo * n + r;
#pragma unroll 4
            for (int d = dlo; d <= dhi; d += kWaves, L += step) a += *L * ring[(r - d) & mask];
        }
        part[wid * kBlk + lane] = a;
        store_tile(tv);
        __syncthreads();
        if (wid == 0) {
            T sv = bv;
#pragma unroll
            for (int w = 0; w < kWaves; ++w) sv -= part[w * kBlk + lane];
            const int tmax = min(kBlk, s1 - i0);
            // this lane's tile row and the reciprocal diagonal into registers
            // first: the 64 dependent steps then wait on no LDS read
            T lrow[kBlk];
#pragma unroll
            for (int t = 0; t < kBlk; ++t) lrow[t] = tile[lane * kTileLd + t];
            const T dinv = tile[lane * kTileLd + lane];
#pragma unroll
            for (int t = 0; t < kBlk; ++t) {
                if (t < tmax) {
                    const T y = lane_value(sv, t) * lane_value(dinv, t);
                    if (lane > t) sv -= lrow[t] * y;
                    else if (lane == t) sv = y;
                }
            }
            if (r < s1) {
                ring[r & mask] = sv;
                ybuf[3 * (size_t)r + comp] = sv;
            }
            pr = pn;
        }
        __syncthreads();
    }
    // ---- backward: L^T x = y ----
    const int last = s0 + ((s1 - s0 - 1) / kBlk) * kBlk;
    pr = (wid == 0 && last + lane < s1) ? perm[last + lane] : 0;
    for (int i0 = last; i0 >= s0; i0 -= kBlk) {
        const int r = i0 + lane;
        T yv = 0;
        int pn = 0;
        if (wid == 0) {
            if (r < s1) yv = ybuf[3 * (size_t)r + comp];
            if (i0 > s0) pn = perm[r - kBlk];
        }
        T tv[kDiagPerWave];
        load_tile(i0, tv);
        T a = 0;
        if (r < s1) {
            const int dlo = i0 + kBlk - r + wid, dhi = min(bw, s1 - 1 - r);
            const T* L = Ld + (size_t)dlo * n + r + dlo;
#pragma unroll 4
            for (int d = dlo; d <= dhi; d += kWaves, L += step + kWaves) a += *L * ring[(r + d) & mask];
        }
        part[wid * kBlk + lane] = a;
        store_tile(tv);
        __syncthreads();
        if (wid == 0) {
            T sv = yv;
#pragma unroll
            for (int w = 0; w < kWaves; ++w) sv -= part[w * kBlk + lane];
            const int tmax = min(kBlk, s1 - i0);
            T lcol[kBlk];  // this lane's tile column
#pragma unroll
            for (int t = 0; t < kBlk; ++t) lcol[t] = tile[t * kTileLd + lane];
            const T dinv = tile[lane * kTileLd + lane];
#pragma unroll
            for (int t = kBlk - 1; t >= 0; --t) {
                if (t < tmax) {
                    const T xv = lane_value(sv, t) * lane_value(dinv, t);
                    if (lane < t) sv -= lcol[t] * xv;
                    else if (lane == t) sv = xv;
                }
            }
            if (r < s1) {
                ring[r & mask] = sv;
                x[3 * (size_t)pr + comp] = sv;
            }
            pr = pn;
        }
        __syncthreads();
    }
}

// LDS a workgroup may use on the current device (64 KB unless it reports more).
size_t lds_limit() {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 64 * 1024;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
        return 64 * 1024;
    return (size_t)v;
}

}  // namespace

template <class T> int BandCholesky<T>::max_bandwidth() {
    const size_t fixed = ((size_t)kBlk * kTileLd + (size_t)kWaves * kBlk) * sizeof(T);
    size_t ring = kBlk;  // largest power-of-two ring that fits
    while ((2 * ring) * sizeof(T) + fixed <= lds_limit()) ring *= 2;
    return (int)ring - kBlk;
}

template <class T>
void BandCholesky<T>::factor(int n, const std::vector<int>& rowptr, const std::vector<int>& col,
                             const std::vector<T>& val, hipStream_t stream) {
    SBS_REQUIRE(n > 0, "BandCholesky: empty matrix");
    const int bw_est = band_width_estimate(n, rowptr, col);
    if (bw_est > max_bandwidth())
        throw Error(kInvalidArgument, "direct (Cholesky) global solve: the reordered matrix's half bandwidth " +
                                          std::to_string(bw_est) + " exceeds the " + std::to_string(max_bandwidth()) +
                                          " rows the solve keeps in LDS in this precision; use the PCG global solver");
    if ((double)n * bw_est * bw_est / 2 > kMaxFactorWork)
        throw Error(kInvalidArgument, "direct (Cholesky) global solve: the band factor of " + std::to_string(n) +
                                          " rows with half bandwidth " + std::to_string(bw_est) +
                                          " is too much host work; use the PCG global solver");
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
    const std::vector<int> sys = band_systems(n, rowptr, col, f.perm);
    n_ = n;
    bw_ = bw;
    nsys_ = (int)sys.size() - 1;
    ring_ = ring_rows(bw);
    ld_.alloc(ld.size());
    perm_.alloc(n);
    sys_.alloc(sys.size());
    ybuf_.alloc(3 * (size_t)n);
    SBS_HIP(hipMemcpyAsync(ld_.get(), ld.data(), ld.size() * sizeof(T), hipMemcpyHostToDevice, stream));
    SBS_HIP(hipMemcpyAsync(perm_.get(), f.perm.data(), n * sizeof(int), hipMemcpyHostToDevice, stream));
    SBS_HIP(hipMemcpyAsync(sys_.get(), sys.data(), sys.size() * sizeof(int), hipMemcpyHostToDevice, stream));
    const size_t lds = lds_bytes<T>(ring_);
    if (lds > 64 * 1024)
        SBS_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&band_solve_kernel<T>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    SBS_HIP(hipStreamSynchronize(stream));
}

template <class T> void BandCholesky<T>::solve(const T* b, T* x, hipStream_t stream, const int* skip) const {
    SBS_REQUIRE(n_ > 0, "BandCholesky: solve before factor");
    hipLaunchKernelGGL(band_solve_kernel<T>, dim3(24 * ceil_div((size_t)nsys_, 8)), dim3(kThreads),
                       lds_bytes<T>(ring_), stream, n_, bw_, ring_, nsys_, (const int*)sys_.get(), ld_.get(), perm_.get(),
                       b, x, ybuf_.get(), skip);
    SBS_HIP(hipGetLastError());
}

template class BandCholesky<float>;
template class BandCholesky<double>;

}  // namespace sbs
