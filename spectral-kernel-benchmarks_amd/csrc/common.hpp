// Internal helpers shared by the HIP translation units of libskb_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>

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

// Every device allocation the library makes goes through DeviceBuffer and
// is counted here (skb_device_allocations), so a caller can check that
// repeated executions allocate nothing.
inline std::atomic<uint64_t> device_allocations{0};
inline std::atomic<uint64_t> device_allocated_bytes{0};

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
        device_allocations.fetch_add(1, std::memory_order_relaxed);
        device_allocated_bytes.fetch_add(n, std::memory_order_relaxed);
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

// Evaluations one batched launch takes at most (their pointers travel as
// kernel arguments).
constexpr int kBatchMax = 128;

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

// Workgroup barrier that orders LDS only.  __syncthreads() also makes the
// workgroup's global stores visible, which waits for every store in flight;
// kernels whose barriers guard LDS alone use this one so stores keep draining
// across it.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// ---- bounds-checked buffer access -----------------------------------------
// A raw buffer resource (stride 0) over `bytes` bytes: a load at an offset at
// or past `bytes` returns zero and such a store is dropped, so a kernel can
// issue a fixed number of memory operations whatever its edge cases.
// An offset past every resource this library builds (all are under 2 GiB);
// adding an element size to it cannot wrap to a valid offset.
constexpr uint32_t kBufferPast = 0x80000000u;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void* base, uint32_t bytes) {
    // raw buffer, stride 0: offsets at or past `bytes` read zero / are dropped
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), static_cast<short>(0), bytes, 0x00020000);
}

// The LLVM buffer intrinsics, bound directly: this compiler lowers
// __builtin_amdgcn_raw_buffer_load_b64 to a 32-bit load.
typedef float bf2 __attribute__((ext_vector_type(2)));
typedef float bf4 __attribute__((ext_vector_type(4)));
typedef double bd2 __attribute__((ext_vector_type(2)));
__device__ bf4 raw_buffer_load_f4(__amdgpu_buffer_rsrc_t r, int voffset, int soffset, int aux) __asm(
    "llvm.amdgcn.raw.ptr.buffer.load.v4f32");
__device__ float raw_buffer_load_f1(__amdgpu_buffer_rsrc_t r, int voffset, int soffset, int aux) __asm(
    "llvm.amdgcn.raw.ptr.buffer.load.f32");
__device__ bf2 raw_buffer_load_f2(__amdgpu_buffer_rsrc_t r, int voffset, int soffset, int aux) __asm(
    "llvm.amdgcn.raw.ptr.buffer.load.v2f32");
__device__ bd2 raw_buffer_load_d2(__amdgpu_buffer_rsrc_t r, int voffset, int soffset, int aux) __asm(
    "llvm.amdgcn.raw.ptr.buffer.load.v2f64");
__device__ void raw_buffer_store_f2(bf2 v, __amdgpu_buffer_rsrc_t r, int voffset, int soffset, int aux) __asm(
    "llvm.amdgcn.raw.ptr.buffer.store.v2f32");
__device__ void raw_buffer_store_f4(bf4 v, __amdgpu_buffer_rsrc_t r, int voffset, int soffset, int aux) __asm(
    "llvm.amdgcn.raw.ptr.buffer.store.v4f32");
__device__ void raw_buffer_store_d2(bd2 v, __amdgpu_buffer_rsrc_t r, int voffset, int soffset, int aux) __asm(
    "llvm.amdgcn.raw.ptr.buffer.store.v2f64");

template <typename R>
__device__ __forceinline__ cx<R> buffer_load_cx(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    if constexpr (sizeof(R) == 4) {
        const bf2 v = raw_buffer_load_f2(r, static_cast<int>(off), 0, 0);
        return cx<R>{v.x, v.y};
    } else {
        const bd2 v = raw_buffer_load_d2(r, static_cast<int>(off), 0, 0);
        return cx<R>{v.x, v.y};
    }
}

template <typename R>
__device__ __forceinline__ void buffer_store_cx(cx<R> x, __amdgpu_buffer_rsrc_t r, uint32_t off) {
    if constexpr (sizeof(R) == 4) raw_buffer_store_f2(bf2{x.x, x.y}, r, static_cast<int>(off), 0, 0);
    else raw_buffer_store_d2(bd2{x.x, x.y}, r, static_cast<int>(off), 0, 0);
}

__device__ void raw_buffer_store_f1(float v, __amdgpu_buffer_rsrc_t r, int voffset, int soffset, int aux) __asm(
    "llvm.amdgcn.raw.ptr.buffer.store.f32");
__device__ void raw_buffer_store_d1(double v, __amdgpu_buffer_rsrc_t r, int voffset, int soffset, int aux) __asm(
    "llvm.amdgcn.raw.ptr.buffer.store.f64");

template <typename R>
__device__ __forceinline__ void buffer_store_real(R x, __amdgpu_buffer_rsrc_t r, uint32_t off) {
    if constexpr (sizeof(R) == 4) raw_buffer_store_f1(x, r, static_cast<int>(off), 0, 0);
    else raw_buffer_store_d1(x, r, static_cast<int>(off), 0, 0);
}

}  // namespace skb
