// General COO linear solve (LinearSolver<T>::Solve with PCGJacobiSolver's
// algorithm): sort_coo on the GPU with hipCUB, CSR rows, Jacobi-PCG with the
// scalars on the device.  See linear_solver.hpp for the reference lines.
#include "linear_solver.hpp"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

namespace sbs {
namespace {

constexpr int kTpbL = 256;
constexpr int kCheck = 16;  // iterations launched between host reads of the control word

struct LinCtl {
    double rho[2];   // r.z of iteration k, by parity of k
    int done;        // 1 once a stop test held
    int iters;       // iterations completed (the reference's k at its break)
    int bad_index;   // a triplet index outside [0, N)
    int pad;
    long long nnz;   // entries after summing repeats
};

unsigned nblk(size_t n) { return (unsigned)std::max<size_t>(1, (n + kTpbL - 1) / kTpbL); }

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Block sum of a double into parts[blockIdx.x] (fixed order inside the block).
__device__ __forceinline__ void block_sum_to(double v, double* parts) {
    __shared__ double red[kTpbL / 64];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < kTpbL / 64; ++w) s += red[w];
        parts[blockIdx.x] = s;
    }
}

// Sum of nb partials in index order: every block of a launch forms the same
// value (wave 0 strides the partials, then a fixed-order wave reduction).
__device__ __forceinline__ double total_of(const double* parts, int nb) {
    __shared__ double tot;
    if (threadIdx.x < 64) {
        double s = 0.0;
        for (int i = threadIdx.x; i < nb; i += 64) s += parts[i];
        s = wave_sum(s);
        if (threadIdx.x == 0) tot = s;
    }
    __syncthreads();
    const double t = tot;
    __syncthreads();
    return t;
}

__global__ void ctl_reset_kernel(LinCtl* ctl, long long* nruns) {
    if (threadIdx.x == 0) {
        ctl->rho[0] = ctl->rho[1] = 0.0;
        ctl->done = 0;
        ctl->iters = 0;
        ctl->bad_index = 0;
        ctl->nnz = 0;
        *nruns = 0;
    }
}

// The solve's done flag as one value for the whole block (thread 0 reads it).
__device__ __forceinline__ bool block_done(const LinCtl* ctl) {
    __shared__ int s_done;
    if (threadIdx.x == 0) s_done = ctl->done;
    __syncthreads();
    return s_done != 0;
}

// key = row * N + col; indices outside [0, N) flag the control block.
__global__ void coo_keys_kernel(long long nz, const int* row, const int* col, int n, unsigned long long* keys,
                                LinCtl* ctl) {
    const long long i = (long long)blockIdx.x * kTpbL + threadIdx.x;
    if (i >= nz) return;
    const int r = row[i], c = col[i];
    if (r < 0 || r >= n || c < 0 || c >= n) {
        ctl->bad_index = 1;
        keys[i] = 0;
        return;
    }
    keys[i] = (unsigned long long)r * (unsigned long long)n + (unsigned long long)c;
}

// CSR from the sorted unique keys: column of each entry, and rowptr[r] = the
// first entry of row r (rows without entries point at the next row's first).
__global__ void csr_kernel(long long nz_bound, const long long* nruns, const unsigned long long* ukeys, int n,
                           int* rowptr, int* colidx, LinCtl* ctl) {
    const long long i = (long long)blockIdx.x * kTpbL + threadIdx.x;
    const long long m = *nruns;
    if (i == 0) ctl->nnz = m;
    if (m == 0) {
        for (long long r = i; r <= n; r += (long long)gridDim.x * kTpbL) rowptr[r] = 0;
        return;
    }
    if (i >= m || i >= nz_bound) return;
    const unsigned long long k = ukeys[i];
    const long long r = (long long)(k / (unsigned long long)n);
    colidx[i] = (int)(k - (unsigned long long)r * (unsigned long long)n);
    const long long prev = i == 0 ? -1 : (long long)(ukeys[i - 1] / (unsigned long long)n);
    for (long long rr = prev + 1; rr <= r; ++rr) rowptr[rr] = (int)i;
    if (i == m - 1)
        for (long long rr = r + 1; rr <= n; ++rr) rowptr[rr] = (int)m;
}

// ExtractInverseDiagonalKernel (pcgJacobi.cu:6-24): 1 / the row's diagonal
// entry, 1 where it is missing or |d| < 1e-9.
template <class T>
__global__ void inv_diag_kernel(int n, const int* rowptr, const int* colidx, const T* val, T* dinv) {
    const int r = blockIdx.x * kTpbL + threadIdx.x;
    if (r >= n) return;
    T d = (T)1;
    for (int e = rowptr[r]; e < rowptr[r + 1]; ++e)
        if (colidx[e] == r) {
            d = val[e];
            break;
        }
    if (fabs((double)d) < 1e-9) d = (T)1;
    dinv[r] = (T)1 / d;
}

template <class T> __device__ __forceinline__ T csr_row(const int* rowptr, const int* colidx, const T* val, const T* v, int r) {
    T acc = 0;
    for (int e = rowptr[r]; e < rowptr[r + 1]; ++e) acc += val[e] * v[colidx[e]];
    return acc;
}

// x0 = guess (r = b - A x0) or 0 (r = b); z = D^-1 r; partials of r.r, r.z
// for iteration 0 (parity 0).
template <class T>
__global__ void pcg_start_kernel(int n, const T* b, const T* guess, T* x, T* r, T* z, const T* dinv, const int* rowptr,
                                 const int* colidx, const T* val, double* prr, double* prz) {
    const int i = blockIdx.x * kTpbL + threadIdx.x;
    double rr = 0.0, rz = 0.0;
    if (i < n) {
        T ri = b[i];
        if (guess) ri -= csr_row(rowptr, colidx, val, guess, i);
        const T zi = dinv[i] * ri;
        x[i] = guess ? guess[i] : (T)0;
        r[i] = ri;
        z[i] = zi;
        rr = (double)ri * (double)ri;
        rz = (double)ri * (double)zi;
    }
    block_sum_to(rr, prr);
    block_sum_to(rz, prz);
}

// Iteration k, part 1: the stop tests at the loop top (||r|| < tol, then
// |r.z| < 1e-15), beta = rho_k / rho_{k-1}, p = z + beta p (p = z at k = 0).
template <class T>
__global__ void pcg_dir_kernel(int n, int k, int nb, double tol, const double* prr, const double* prz, const T* z,
                               T* p, LinCtl* ctl) {
    if (block_done(ctl)) return;
    const double rr = total_of(prr, nb), rho = total_of(prz, nb);
    if (sqrt(rr) < tol || fabs(rho) < 1e-15) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            ctl->done = 1;
            ctl->iters = k;
        }
        return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->rho[k & 1] = rho;
    const T beta = k == 0 ? (T)0 : (T)(rho / ctl->rho[(k - 1) & 1]);
    const int i = blockIdx.x * kTpbL + threadIdx.x;
    if (i < n) p[i] = k == 0 ? z[i] : z[i] + beta * p[i];
}

// Iteration k, part 2: q = A p and the partials of p.q.
template <class T>
__global__ void pcg_spmv_kernel(int n, const int* rowptr, const int* colidx, const T* val, const T* p, T* q,
                                double* ppq, const LinCtl* ctl) {
    if (block_done(ctl)) return;
    const int i = blockIdx.x * kTpbL + threadIdx.x;
    double pq = 0.0;
    if (i < n) {
        const T qi = csr_row(rowptr, colidx, val, p, i);
        q[i] = qi;
        pq = (double)p[i] * (double)qi;
    }
    block_sum_to(pq, ppq);
}

// Iteration k, part 3: alpha = rho / p.q, x += alpha p, r -= alpha q, then
// z = D^-1 r and the partials of iteration k+1 (other parity).
template <class T>
__global__ void pcg_update_kernel(int n, int k, int nb, const double* ppq, const T* p, const T* q, T* x, T* r, T* z,
                                  const T* dinv, double* prr_next, double* prz_next, LinCtl* ctl) {
    if (block_done(ctl)) return;
    const double pq = total_of(ppq, nb);
    const T alpha = (T)(ctl->rho[k & 1] / pq);
    const int i = blockIdx.x * kTpbL + threadIdx.x;
    double rr = 0.0, rz = 0.0;
    if (i < n) {
        x[i] += alpha * p[i];
        const T ri = r[i] - alpha * q[i];
        const T zi = dinv[i] * ri;
        r[i] = ri;
        z[i] = zi;
        rr = (double)ri * (double)ri;
        rz = (double)ri * (double)zi;
    }
    block_sum_to(rr, prr_next);
    block_sum_to(rz, prz_next);
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->iters = k + 1;
}

// Workspace layout (each piece 256-byte aligned).
template <class T> struct Layout {
    unsigned long long *keys, *keys_sorted, *ukeys;
    T *vals_sorted, *uvals;
    long long* nruns;
    int *rowptr, *colidx;
    T *dinv, *r, *z, *p, *q;
    double *prr, *prz, *ppq;  // prr, prz: two parities of nb partials
    LinCtl* ctl;
    void* tmp;
    size_t tmp_bytes;
    size_t total;

    Layout(char* base, int n, long long nz, size_t tmp_b) {
        size_t off = 0;
        auto take = [&](size_t bytes) {
            char* ptr = base ? base + off : nullptr;
            off += align256(std::max<size_t>(bytes, 1));
            return ptr;
        };
        const size_t z1 = (size_t)std::max<long long>(nz, 1);
        const int nb = (int)nblk(n);
        keys = (unsigned long long*)take(z1 * 8);
        keys_sorted = (unsigned long long*)take(z1 * 8);
        ukeys = (unsigned long long*)take(z1 * 8);
        vals_sorted = (T*)take(z1 * sizeof(T));
        uvals = (T*)take(z1 * sizeof(T));
        nruns = (long long*)take(8);
        rowptr = (int*)take(((size_t)n + 1) * 4);
        colidx = (int*)take(z1 * 4);
        dinv = (T*)take((size_t)n * sizeof(T));
        r = (T*)take((size_t)n * sizeof(T));
        z = (T*)take((size_t)n * sizeof(T));
        p = (T*)take((size_t)n * sizeof(T));
        q = (T*)take((size_t)n * sizeof(T));
        prr = (double*)take(2 * (size_t)nb * 8);
        prz = (double*)take(2 * (size_t)nb * 8);
        ppq = (double*)take((size_t)nb * 8);
        ctl = (LinCtl*)take(sizeof(LinCtl));
        tmp = take(tmp_b);
        tmp_bytes = tmp_b;
        total = off;
    }
};

}  // namespace

// Workspace of the allocation-free solve (solve_no_alloc): one array per GPU,
// held until the process ends.  Host code never passes the array's name to a
// kernel or a copy (on the host it is only a stand-in for the symbol): it
// uses the device address hipGetSymbolAddress returns for the current GPU.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The current GPU's address of `scratch`, looked up once per device.
void* scratch_ptr() {
    int dev = 0;
    SBS_HIP(hipGetDevice(&dev));
    SBS_REQUIRE(dev >= 0 && dev < kMaxDevices, "allocation-free solve: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        SBS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        SBS_REQUIRE(p != nullptr, "allocation-free solve: no device address for the workspace");
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}
}  // namespace

template <class T>
int CooPcgSolver<T>::solve_no_alloc(int n, const T* b, T* x, const T* A, long long nz, const int* row,
                                    const int* col, const T* guess, hipStream_t stream) {
    SBS_REQUIRE(n >= 0 && nz >= 0, "LinearSolver::Solve: negative size");
    const size_t need = CooWorkspace::bytes(n, nz, sizeof(T), sort_tmp_bytes(nz, n));
    if (need > kNoAllocScratchBytes)
        throw Error(kInvalidArgument, "allocation-free solve: N = " + std::to_string(n) + ", nz = " +
                                          std::to_string(nz) + " needs " + std::to_string(need) +
                                          " bytes of workspace; its fixed workspace holds " +
                                          std::to_string(kNoAllocScratchBytes));
    return solve_in(n, b, x, A, nz, row, col, guess, stream, scratch_ptr(), kNoAllocScratchBytes);
}

template <class T> size_t CooPcgSolver<T>::sort_tmp_bytes(long long nz, int n) {
    const int items = (int)std::max<long long>(nz, 1);
    size_t a = 0, b = 0;
    unsigned long long* k = nullptr;
    T* v = nullptr;
    long long* nr = nullptr;
    SBS_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, a, k, k, v, v, items, 0, 64, nullptr));
    SBS_HIP(hipcub::DeviceReduce::ReduceByKey(nullptr, b, k, k, v, v, nr, hipcub::Sum(), items, nullptr));
    (void)n;
    return std::max(a, b);
}

size_t CooWorkspace::bytes(int n, long long nz, size_t elem, size_t sort_tmp) {
    if (elem == 4) return Layout<float>(nullptr, n, nz, sort_tmp).total;
    return Layout<double>(nullptr, n, nz, sort_tmp).total;
}

template <class T>
int CooPcgSolver<T>::solve(int n, const T* b, T* x, const T* A, long long nz, const int* row, const int* col,
                           const T* guess, hipStream_t stream) {
    SBS_REQUIRE(n >= 0 && nz >= 0, "LinearSolver::Solve: negative size");
    const size_t need = CooWorkspace::bytes(n, nz, sizeof(T), sort_tmp_bytes(nz, n));
    if (ws_.bytes() < need) ws_.alloc(need);
    return solve_in(n, b, x, A, nz, row, col, guess, stream, ws_.get(), ws_.bytes());
}

template <class T>
int CooPcgSolver<T>::solve_in(int n, const T* b, T* x, const T* A, long long nz, const int* row, const int* col,
                              const T* guess, hipStream_t stream, void* ws, size_t ws_bytes) {
    SBS_REQUIRE(n >= 0 && nz >= 0, "LinearSolver::Solve: negative size");
    SBS_REQUIRE(nz < (1ll << 31), "LinearSolver::Solve: more than 2^31 - 1 triplets");
    if (n == 0) return 0;
    SBS_REQUIRE(b && x, "LinearSolver::Solve: null b or x");
    SBS_REQUIRE(nz == 0 || (A && row && col), "LinearSolver::Solve: null triplet array");
    const size_t tmp_b = sort_tmp_bytes(nz, n);
    Layout<T> L(static_cast<char*>(ws), n, nz, tmp_b);
    SBS_REQUIRE(ws != nullptr && L.total <= ws_bytes, "LinearSolver::Solve: workspace too small for this system");
    LinCtl hc{};  // host copy of the control word (a plain copy: nothing is allocated)
    const int nb = (int)nblk(n);
    hipLaunchKernelGGL(ctl_reset_kernel, dim3(1), dim3(64), 0, stream, L.ctl, L.nruns);
    // sort_coo (linear.cu:7-40): sort by (row, column), sum repeated entries
    const int items = (int)nz;
    int end_bit = 1;
    while (end_bit < 64 && ((unsigned long long)n * (unsigned long long)n) > (1ull << end_bit)) ++end_bit;
    if (nz > 0) {
        hipLaunchKernelGGL(coo_keys_kernel, dim3(nblk(nz)), dim3(kTpbL), 0, stream, (long long)nz, row, col, n, L.keys,
                           L.ctl);
        size_t tb = L.tmp_bytes;
        SBS_HIP(hipcub::DeviceRadixSort::SortPairs(L.tmp, tb, L.keys, L.keys_sorted, A, L.vals_sorted, items, 0,
                                                   end_bit, stream));
        tb = L.tmp_bytes;
        SBS_HIP(hipcub::DeviceReduce::ReduceByKey(L.tmp, tb, L.keys_sorted, L.ukeys, L.vals_sorted, L.uvals, L.nruns,
                                                  hipcub::Sum(), items, stream));
    }
    hipLaunchKernelGGL(csr_kernel, dim3(nblk(std::max<long long>(nz, 1))), dim3(kTpbL), 0, stream,
                       (long long)std::max<long long>(nz, 1), (const long long*)L.nruns,
                       (const unsigned long long*)L.ukeys, n, L.rowptr, L.colidx, L.ctl);
    hipLaunchKernelGGL(inv_diag_kernel<T>, dim3(nb), dim3(kTpbL), 0, stream, n, (const int*)L.rowptr,
                       (const int*)L.colidx, (const T*)L.uvals, L.dinv);
    hipLaunchKernelGGL(pcg_start_kernel<T>, dim3(nb), dim3(kTpbL), 0, stream, n, b, guess, x, L.r, L.z,
                       (const T*)L.dinv, (const int*)L.rowptr, (const int*)L.colidx, (const T*)L.uvals, L.prr, L.prz);
    SBS_HIP(hipGetLastError());
    SBS_HIP(hipMemcpyAsync(&hc, L.ctl, sizeof(LinCtl), hipMemcpyDeviceToHost, stream));
    SBS_HIP(hipStreamSynchronize(stream));
    SBS_REQUIRE(!hc.bad_index, "LinearSolver::Solve: a row or column index is outside [0, N)");
    nnz_ = hc.nnz;
    int k = 0;
    while (k < max_iter_) {
        const int chunk = std::min(kCheck, max_iter_ - k);
        for (int j = 0; j < chunk; ++j, ++k) {
            const int par = k & 1;
            hipLaunchKernelGGL(pcg_dir_kernel<T>, dim3(nb), dim3(kTpbL), 0, stream, n, k, nb, tol_,
                               (const double*)(L.prr + par * nb), (const double*)(L.prz + par * nb), (const T*)L.z,
                               L.p, L.ctl);
            hipLaunchKernelGGL(pcg_spmv_kernel<T>, dim3(nb), dim3(kTpbL), 0, stream, n, (const int*)L.rowptr,
                               (const int*)L.colidx, (const T*)L.uvals, (const T*)L.p, L.q, L.ppq,
                               (const LinCtl*)L.ctl);
            hipLaunchKernelGGL(pcg_update_kernel<T>, dim3(nb), dim3(kTpbL), 0, stream, n, k, nb,
                               (const double*)L.ppq, (const T*)L.p, (const T*)L.q, x, L.r, L.z, (const T*)L.dinv,
                               L.prr + (1 - par) * nb, L.prz + (1 - par) * nb, L.ctl);
        }
        SBS_HIP(hipGetLastError());
        SBS_HIP(hipMemcpyAsync(&hc, L.ctl, sizeof(LinCtl), hipMemcpyDeviceToHost, stream));
        SBS_HIP(hipStreamSynchronize(stream));
        if (hc.done) break;
    }
    return hc.iters;
}

template class CooPcgSolver<float>;
template class CooPcgSolver<double>;

}  // namespace sbs
