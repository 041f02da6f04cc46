// Shared host/device helpers for the gfx950 PD library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace sbs {

// Status codes returned through the C ABI (0 = success).
enum Status : int {
    kOk = 0,
    kInvalidArgument = -1,
    kHipError = -2,
    kNotReady = -3,
    kFactorFailed = -4,
    kOutOfMemory = -5,
    kInternal = -6,
};

// Errors inside the library are C++ exceptions carrying a status; the C ABI
// catches them, stores the message in a thread-local slot and returns the code.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);
const char* last_error();

#define SBS_HIP(expr)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            throw ::sbs::Error(::sbs::kHipError, std::string(#expr) + ": " + hipGetErrorString(e_) + \
                                                     " (" __FILE__ ":" + std::to_string(__LINE__) + ")"); \
    } while (0)

#define SBS_REQUIRE(cond, msg)                                                   \
    do {                                                                         \
        if (!(cond)) throw ::sbs::Error(::sbs::kInvalidArgument, std::string(msg)); \
    } while (0)

// Device buffer owned by the library (hipMalloc / hipFree).
template <class T> struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        if (count == 0) return;
        hipError_t e = hipMalloc(&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            throw Error(kOutOfMemory, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed: " +
                                          hipGetErrorString(e));
        }
        n = count;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    T* get() const { return p; }
    size_t bytes() const { return n * sizeof(T); }
};

// Blocks of one launch are dispatched round-robin over the 8 XCDs of an
// MI355X (block b runs on XCD b % 8).  Remapping so each XCD owns one
// contiguous range of logical blocks keeps neighbouring rows/tets, and the
// vertex data they gather, in the same XCD's L2.
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nblocks) {
    constexpr unsigned kXcd = 8;
    const unsigned q = nblocks / kXcd, r = nblocks % kXcd;
    const unsigned x = b % kXcd, j = b / kXcd;
    return x * q + (x < r ? x : r) + j;
}

inline unsigned ceil_div(size_t a, unsigned b) { return (unsigned)((a + b - 1) / b); }

// Wave-level sum (64 lanes) with butterfly shuffles.
template <class T> __device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// One DPP move of a double (both 32-bit halves); lanes without a source, or
// outside ROW_MASK, read 0.  With every row enabled (and a control whose
// every lane has a source: quad_perm, mirrors) no old value is needed, so the
// destination is not cleared first.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ double dpp_f64(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    int lo, hi;
    if constexpr (ROW_MASK == 0xf) {
        lo = __builtin_amdgcn_mov_dpp((int)(unsigned)b, CTRL, 0xf, 0xf, true);
        hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), CTRL, 0xf, 0xf, true);
    } else {
        lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, ROW_MASK, 0xf, false);
        hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, ROW_MASK, 0xf, false);
    }
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Sum of a double over each row of 16 lanes with DPP moves (pairs and quads
// by quad_perm, 8 by row_half_mirror, 16 by row_mirror), returned to every
// lane of the row.  Every lane of the wavefront must be active.
__device__ __forceinline__ double row16_total(double v) {
    v += dpp_f64<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_f64<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_f64<0x141>(v);  // row_half_mirror
    v += dpp_f64<0x140>(v);  // row_mirror
    return v;
}

// The double held by lane L (uniform over the wavefront).
template <int L> __device__ __forceinline__ double readlane_f64(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, L);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), L);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Sum of a double over the 64 lanes of a wavefront with DPP moves (VALU only,
// no LDS round trips), returned to every lane.  The steps: pairs and quads
// (quad_perm), 8 (row_half_mirror), 16 (row_mirror), then row 1 and 3 take
// row 0 and 2's total (row_bcast:15) and row 3 takes rows 0-1 (row_bcast:31);
// lane 63 holds the total.  Every lane of the wavefront must be active.
__device__ __forceinline__ double wave_total(double v) {
    v += dpp_f64<0xB1>(v);        // quad_perm [1,0,3,2]
    v += dpp_f64<0x4E>(v);        // quad_perm [2,3,0,1]
    v += dpp_f64<0x141>(v);       // row_half_mirror
    v += dpp_f64<0x140>(v);       // row_mirror
    v += dpp_f64<0x142, 0xA>(v);  // row_bcast:15
    v += dpp_f64<0x143, 0xC>(v);  // row_bcast:31
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, 63);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), 63);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

}  // namespace sbs
