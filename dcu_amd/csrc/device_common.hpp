// Shared HIP helpers for gfx950 (CDNA4, 64-wide wavefronts).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

namespace spai {

// Error carried from the HIP runtime or an argument check up to the C-ABI,
// which turns it into a status code.
struct Error : std::runtime_error {
    int status;
    Error(int st, const std::string &msg) : std::runtime_error(msg), status(st) {}
};

enum Status {
    kOk = 0,
    kErrInvalid = -1,
    kErrHip = -2,
    kErrNoMem = -3,
    kErrUnsupported = -4,
    kErrBreakdown = -5,
    kErrNotBuilt = -6,
};

#define SPAI_HIP(call)                                                                          \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            throw ::spai::Error(e_ == hipErrorOutOfMemory ? ::spai::kErrNoMem : ::spai::kErrHip, \
                                std::string(#call) + ": " + hipGetErrorString(e_) + " at " +    \
                                    __FILE__ + ":" + std::to_string(__LINE__));                 \
    } while (0)

#define SPAI_CHECK_LAUNCH() SPAI_HIP(hipGetLastError())

inline void require(bool ok, const std::string &msg, int st = kErrInvalid) {
    if (!ok) throw Error(st, msg);
}
// literal messages: no string is built unless the check fails
inline void require(bool ok, const char *msg, int st = kErrInvalid) {
    if (!ok) throw Error(st, msg);
}

// Owning device allocation.
template <class T>
struct DeviceArray {
    T *ptr = nullptr;
    size_t count = 0;
    DeviceArray() = default;
    explicit DeviceArray(size_t n) { resize(n); }
    DeviceArray(const DeviceArray &) = delete;
    DeviceArray &operator=(const DeviceArray &) = delete;
    DeviceArray(DeviceArray &&o) noexcept : ptr(o.ptr), count(o.count) { o.ptr = nullptr; o.count = 0; }
    DeviceArray &operator=(DeviceArray &&o) noexcept {
        if (this != &o) { release(); ptr = o.ptr; count = o.count; o.ptr = nullptr; o.count = 0; }
        return *this;
    }
    ~DeviceArray() { release(); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        count = 0;
    }
    void resize(size_t n) {
        if (n == count && ptr) return;
        release();
        if (n) SPAI_HIP(hipMalloc((void **)&ptr, n * sizeof(T)));
        count = n;
    }
    // grow-only reservation for workspaces reused across calls
    void reserve(size_t n) {
        if (n > count) resize(n);
    }
    size_t bytes() const { return count * sizeof(T); }
    T *get() const { return ptr; }
};

// ------------------------------------------------------- device helpers ----
#ifdef __HIP_DEVICE_COMPILE__
#define SPAI_DEV __device__ __forceinline__
#else
#define SPAI_DEV __device__ inline
#endif

// Sum across a power-of-two group of W lanes (W <= 64) inside one wavefront;
// every lane of the group receives the total.
template <int W, class T>
SPAI_DEV T group_sum(T v) {
#pragma unroll
    for (int o = W / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, W);
    return v;
}

// A value from another lane of the same row of 16 lanes (a DPP move; CTRL
// is the DPP control: 0xB1 and 0x4E swap neighbours and pairs inside a quad,
// 0x141 mirrors each half-row of 8 lanes, 0x140 each row of 16).
template <int CTRL>
SPAI_DEV int dpp_move(int v) {
    return __builtin_amdgcn_update_dpp(v, v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
SPAI_DEV float dpp_move(float v) {
    return __int_as_float(dpp_move<CTRL>(__float_as_int(v)));
}
template <int CTRL>
SPAI_DEV double dpp_move(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_move<CTRL>((int)b), hi = dpp_move<CTRL>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Sum over aligned groups of W lanes, every lane receiving the total.  Inside
// a row of 16 lanes the partners come from DPP moves (quad swaps, then the
// half-row and row mirrors, each pairing the lanes symmetrically, so every
// lane adds the same two values); wider groups finish with cross-lane
// permutes.
template <int W, class T>
SPAI_DEV T group_allsum(T v) {
    static_assert(W == 1 || W == 2 || W == 4 || W == 8 || W == 16 || W == 32 || W == 64, "W: power of 2 up to 64");
    if constexpr (W >= 2) v += dpp_move<0xB1>(v);
    if constexpr (W >= 4) v += dpp_move<0x4E>(v);
    if constexpr (W >= 8) v += dpp_move<0x141>(v);
    if constexpr (W >= 16) v += dpp_move<0x140>(v);
    if constexpr (W >= 32) v += __shfl_xor(v, 16, W);
    if constexpr (W >= 64) v += __shfl_xor(v, 32, W);
    return v;
}

// A value every lane of the wavefront holds alike, marked uniform (scalar
// registers) for the code that uses it.
SPAI_DEV double wave_uniform(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)b), hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

template <int W>
SPAI_DEV int group_max(int v) {
#pragma unroll
    for (int o = W / 2; o > 0; o >>= 1) {
        int t = __shfl_xor(v, o, W);
        v = t > v ? t : v;
    }
    return v;
}

// Inclusive scan across W lanes (Hillis-Steele with shuffles).
template <int W>
SPAI_DEV int group_inclusive_scan(int v, int lane) {
#pragma unroll
    for (int o = 1; o < W; o <<= 1) {
        int t = __shfl_up(v, o, W);
        if (lane >= o) v += t;
    }
    return v;
}

// Orders LDS traffic between lanes of one wavefront: the compiler may not move
// LDS accesses across it and waits for outstanding LDS operations.
SPAI_DEV void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// 64-bit load that bypasses the non-coherent vector L1 (used for data written
// by other workgroups of the same launch).
template <class T>
SPAI_DEV T load_coherent(const T *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

inline int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

}  // namespace spai
