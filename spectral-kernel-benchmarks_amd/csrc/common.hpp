// Internal helpers shared by the HIP translation units of libskb_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "skb_hip.h"

namespace skb {

// ---- status plumbing ------------------------------------------------------
struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

void set_last_error(const std::string& message);

[[noreturn]] inline void fail(int status, const std::string& message) { throw Error(status, message); }

inline void check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) fail(SKB_RUNTIME_ERROR, std::string(what) + ": " + hipGetErrorString(e));
}

// Runs `body`, converting exceptions into status codes (no exception crosses
// the C ABI).
template <typename F>
int guarded(F&& body) noexcept {
    try {
        body();
        return SKB_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.status;
    } catch (const std::bad_alloc&) {
        set_last_error("allocation failed");
        return SKB_BAD_ALLOC;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return SKB_RUNTIME_ERROR;
    } catch (...) {
        set_last_error("unknown failure");
        return SKB_RUNTIME_ERROR;
    }
}

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Device allocation owned by a plan (allocated at creation only).
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() { reset(); }
    void allocate(size_t n) {
        reset();
        if (n == 0) return;
        if (hipMalloc(&ptr, n) != hipSuccess) {
            ptr = nullptr;
            (void)hipGetLastError();
            fail(SKB_BAD_ALLOC, "hipMalloc of " + std::to_string(n) + " bytes failed");
        }
        bytes = n;
    }
    void upload(const void* host, size_t n) {
        allocate(n);
        if (n) check_hip(hipMemcpy(ptr, host, n, hipMemcpyHostToDevice), "hipMemcpy");
    }
    void reset() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    template <typename T>
    T* as() const {
        return static_cast<T*>(ptr);
    }
};

inline bool is_pow2(uint64_t v) { return v && !(v & (v - 1)); }

// XCD-aware work mapping.  MI355X dispatches workgroup b to XCD b % 8, and
// each XCD has its own L2.  With one workgroup per item, item ranges are
// split into 8 contiguous chunks, one per XCD, so neighbouring items (which
// share the cache lines at their boundaries) run behind the same L2.
// Launch xcd_grid(items) workgroups; xcd_item() >= items means idle.
constexpr int kXcds = 8;
inline int xcd_grid(int items) { return ((items + kXcds - 1) / kXcds) * kXcds; }
__device__ __forceinline__ int xcd_item(int b, int items) {
    const int per = (items + kXcds - 1) / kXcds;
    return (b % kXcds) * per + b / kXcds;
}

// Sizes the hand-written FFT kernels are instantiated for.
constexpr uint64_t kMinFft = 4;
constexpr uint64_t kMaxFft = 1024;

// Rethrows a failed status of the library's own C entry points.
inline void check_status(int status) {
    if (status != SKB_OK) {
        char msg[1024];
        skb_last_error(msg, sizeof(msg));
        fail(status, msg);
    }
}

// Grid of a grid-stride elementwise kernel of 256-thread workgroups.
inline dim3 stream_grid(size_t total) {
    const size_t blocks = (total + 255) / 256;
    return dim3(static_cast<unsigned>(std::max<size_t>(1, std::min<size_t>(blocks, 8192))));
}

inline size_t complex_bytes(int precision) { return precision == SKB_F64 ? 16 : 8; }
inline size_t real_bytes(int precision) { return precision == SKB_F64 ? 8 : 4; }

inline void check_precision(int precision) {
    if (precision != SKB_F64 && precision != SKB_F32) fail(SKB_INVALID_ARGUMENT, "precision must be SKB_F64 or SKB_F32");
}

void validate_workload(const skb_workload* w);

// ---- device complex -------------------------------------------------------
template <typename R>
struct alignas(2 * sizeof(R)) cx {
    R x, y;
};

template <typename R>
__device__ __forceinline__ cx<R> operator+(cx<R> a, cx<R> b) {
    return {a.x + b.x, a.y + b.y};
}
template <typename R>
__device__ __forceinline__ cx<R> operator-(cx<R> a, cx<R> b) {
    return {a.x - b.x, a.y - b.y};
}
template <typename R>
__device__ __forceinline__ cx<R> cmul(cx<R> a, cx<R> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename R>
__device__ __forceinline__ cx<R> cmulc(cx<R> a, cx<R> b) {
    return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename R>
__device__ __forceinline__ cx<R> conjg(cx<R> a) {
    return {a.x, -a.y};
}
template <typename R>
__device__ __forceinline__ cx<R> scal(cx<R> a, R s) {
    return {a.x * s, a.y * s};
}

}  // namespace skb
