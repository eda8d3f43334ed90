// Split <-> interleaved complex arrays: the conversions under the split-layout
// entry points of the providers (separate real and imaginary arrays, as the
// reference's split dgemm and split FFTW plans use, against the interleaved
// complex arrays the GEMM and FFT kernels read).  Memory-bound grid-stride
// kernels.
#include "common.hpp"
#include "internal.hpp"

namespace skb {
namespace {

template <typename R>
__global__ void interleave_kernel(const R* __restrict__ re, const R* __restrict__ im, cx<R>* __restrict__ out,
                                  size_t n) {
    for (size_t e = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; e < n;
         e += static_cast<size_t>(gridDim.x) * blockDim.x)
        out[e] = cx<R>{re[e], im[e]};
}

template <typename R>
__global__ void deinterleave_kernel(const cx<R>* __restrict__ in, R* __restrict__ re, R* __restrict__ im, R scale,
                                    size_t n) {
    for (size_t e = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; e < n;
         e += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const cx<R> v = in[e];
        if (re) re[e] = scale * v.x;  // a null array (a skipped component) is uniform over the grid
        if (im) im[e] = scale * v.y;
    }
}

template <typename R>
__global__ void scale_split_kernel(R* __restrict__ re, R* __restrict__ im, R scale, size_t n) {
    for (size_t e = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; e < n;
         e += static_cast<size_t>(gridDim.x) * blockDim.x) {
        re[e] *= scale;
        im[e] *= scale;
    }
}

template <typename R>
__global__ void copy_planes_kernel(const cx<R>* __restrict__ src, uint64_t src_stride, cx<R>* __restrict__ dst,
                                   uint64_t dst_stride, uint32_t planes, size_t n) {
    for (size_t e = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; e < n;
         e += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const size_t f = e / planes, p = e - f * planes;
        dst[p + dst_stride * f] = src[p + src_stride * f];
    }
}

}  // namespace

void copy_planes(int precision, uint64_t frequencies, int planes, const void* src, uint64_t src_stride, void* dst,
                 uint64_t dst_stride, hipStream_t s) {
    const size_t n = static_cast<size_t>(frequencies) * static_cast<size_t>(planes);
    if (n == 0) return;
    if (precision == SKB_F64)
        hipLaunchKernelGGL(copy_planes_kernel<double>, stream_grid(n), dim3(256), 0, s,
                           static_cast<const cx<double>*>(src), src_stride, static_cast<cx<double>*>(dst), dst_stride,
                           static_cast<uint32_t>(planes), n);
    else
        hipLaunchKernelGGL(copy_planes_kernel<float>, stream_grid(n), dim3(256), 0, s,
                           static_cast<const cx<float>*>(src), src_stride, static_cast<cx<float>*>(dst), dst_stride,
                           static_cast<uint32_t>(planes), n);
    check_hip(hipGetLastError(), "plane copy launch");
}

void interleave(int precision, uint64_t n, const void* re, const void* im, void* out, hipStream_t s) {
    if (n == 0) return;
    if (!re || !im || !out) fail(SKB_INVALID_ARGUMENT, "null device pointer");
    if (precision == SKB_F64)
        hipLaunchKernelGGL(interleave_kernel<double>, stream_grid(n), dim3(256), 0, s, static_cast<const double*>(re),
                           static_cast<const double*>(im), static_cast<cx<double>*>(out), static_cast<size_t>(n));
    else
        hipLaunchKernelGGL(interleave_kernel<float>, stream_grid(n), dim3(256), 0, s, static_cast<const float*>(re),
                           static_cast<const float*>(im), static_cast<cx<float>*>(out), static_cast<size_t>(n));
    check_hip(hipGetLastError(), "interleave launch");
}

void deinterleave(int precision, uint64_t n, const void* in, void* re, void* im, double scale, hipStream_t s) {
    if (n == 0) return;
    if (!in || (!re && !im)) fail(SKB_INVALID_ARGUMENT, "null device pointer");
    if (precision == SKB_F64)
        hipLaunchKernelGGL(deinterleave_kernel<double>, stream_grid(n), dim3(256), 0, s,
                           static_cast<const cx<double>*>(in), static_cast<double*>(re), static_cast<double*>(im),
                           scale, static_cast<size_t>(n));
    else
        hipLaunchKernelGGL(deinterleave_kernel<float>, stream_grid(n), dim3(256), 0, s,
                           static_cast<const cx<float>*>(in), static_cast<float*>(re), static_cast<float*>(im),
                           static_cast<float>(scale), static_cast<size_t>(n));
    check_hip(hipGetLastError(), "deinterleave launch");
}

}  // namespace skb

using namespace skb;

extern "C" {

int skb_interleave(int precision, uint64_t count, const void* real, const void* imaginary, void* out, void* stream) {
    return guarded([&] {
        check_precision(precision);
        interleave(precision, count, real, imaginary, out, as_stream(stream));
    });
}

int skb_deinterleave(int precision, uint64_t count, const void* in, void* real, void* imaginary, double scale,
                     void* stream) {
    return guarded([&] {
        check_precision(precision);
        deinterleave(precision, count, in, real, imaginary, scale, as_stream(stream));
    });
}

uint64_t skb_device_allocations(void) { return device_allocations.load(std::memory_order_relaxed); }

uint64_t skb_device_allocated_bytes(void) { return device_allocated_bytes.load(std::memory_order_relaxed); }

int skb_scale_split(int precision, uint64_t count, void* real, void* imaginary, double scale, void* stream) {
    return guarded([&] {
        check_precision(precision);
        if (count == 0) return;
        if (!real || !imaginary) fail(SKB_INVALID_ARGUMENT, "null device pointer");
        const hipStream_t s = as_stream(stream);
        if (precision == SKB_F64)
            hipLaunchKernelGGL(scale_split_kernel<double>, stream_grid(count), dim3(256), 0, s,
                               static_cast<double*>(real), static_cast<double*>(imaginary), scale,
                               static_cast<size_t>(count));
        else
            hipLaunchKernelGGL(scale_split_kernel<float>, stream_grid(count), dim3(256), 0, s,
                               static_cast<float*>(real), static_cast<float*>(imaginary), static_cast<float>(scale),
                               static_cast<size_t>(count));
        check_hip(hipGetLastError(), "scale launch");
    });
}

}  // extern "C"
