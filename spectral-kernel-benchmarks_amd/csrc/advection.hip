// Pointwise advection product target = -scale * (u*qx + v*qy + w*qz)
// (src/pointwise_advection.cpp:152-166).  Six input streams and one output
// stream; pure HBM-bound elementwise work, so the kernel moves 16 bytes per
// lane per stream with a grid-stride loop and keeps the sum in the order
// u*qx + v*qy + w*qz, as the reference's serial form does (FP32 streams sum
// in FP64 and round once).
#include "common.hpp"

namespace skb {

template <typename R>
__device__ __forceinline__ R advect(R u, R v, R w, R qx, R qy, R qz, R scale) {
    if constexpr (sizeof(R) == 4) {
        // FP32 operands: the products are exact in FP64, the sum and the
        // scaling run in FP64 and the target is rounded once to FP32
        const double s = static_cast<double>(u) * qx + static_cast<double>(v) * qy + static_cast<double>(w) * qz;
        return static_cast<R>(-static_cast<double>(scale) * s);
    } else {
        const R s = u * qx + v * qy + w * qz;
        return -scale * s;
    }
}

template <typename R>
__global__ __launch_bounds__(256) void advection_kernel(const R* __restrict__ shared, const R* __restrict__ deriv,
                                                       R* __restrict__ target, size_t volume, R scale) {
    constexpr int V = 16 / sizeof(R);  // elements per 16-byte vector
    const size_t vectors = volume / V;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    const size_t start = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    // Vector body needs every stream 16-byte aligned.
    const bool aligned = ((reinterpret_cast<uintptr_t>(shared) | reinterpret_cast<uintptr_t>(deriv) |
                           reinterpret_cast<uintptr_t>(target)) & 15) == 0 && (volume % V) == 0;
    if (aligned) {
        using Vec = typename std::conditional<sizeof(R) == 8, double2, float4>::type;
        const Vec* u = reinterpret_cast<const Vec*>(shared);
        const Vec* v = reinterpret_cast<const Vec*>(shared + volume);
        const Vec* w = reinterpret_cast<const Vec*>(shared + 2 * volume);
        const Vec* qx = reinterpret_cast<const Vec*>(deriv);
        const Vec* qy = reinterpret_cast<const Vec*>(deriv + volume);
        const Vec* qz = reinterpret_cast<const Vec*>(deriv + 2 * volume);
        Vec* t = reinterpret_cast<Vec*>(target);
        for (size_t i = start; i < vectors; i += stride) {
            const Vec a = u[i], b = v[i], c = w[i], d = qx[i], e = qy[i], f = qz[i];
            Vec r;
            const R* ap = reinterpret_cast<const R*>(&a);
            const R* bp = reinterpret_cast<const R*>(&b);
            const R* cp = reinterpret_cast<const R*>(&c);
            const R* dp = reinterpret_cast<const R*>(&d);
            const R* ep = reinterpret_cast<const R*>(&e);
            const R* fp = reinterpret_cast<const R*>(&f);
            R* rp = reinterpret_cast<R*>(&r);
#pragma unroll
            for (int k = 0; k < V; ++k) rp[k] = advect(ap[k], bp[k], cp[k], dp[k], ep[k], fp[k], scale);
            t[i] = r;
        }
        return;
    }
    for (size_t i = start; i < volume; i += stride)
        target[i] = advect(shared[i], shared[volume + i], shared[2 * volume + i], deriv[i], deriv[volume + i],
                           deriv[2 * volume + i], scale);
}

}  // namespace skb

using namespace skb;

extern "C" int skb_pointwise_advection(uint64_t volume, double scale, const void* shared, const void* derivative,
                                       void* target, int precision, void* stream) {
    return guarded([&] {
        check_precision(precision);
        if (volume == 0) fail(SKB_INVALID_ARGUMENT, "Pointwise advection volume must be positive.");
        if (!shared || !derivative || !target) fail(SKB_INVALID_ARGUMENT, "null device pointer");
        const size_t work = (volume + 3) / 4;
        const dim3 grid(static_cast<unsigned>(std::min<size_t>((work + 255) / 256, 2048 * 8)));
        if (precision == SKB_F64)
            hipLaunchKernelGGL(advection_kernel<double>, grid, dim3(256), 0, as_stream(stream),
                               static_cast<const double*>(shared), static_cast<const double*>(derivative),
                               static_cast<double*>(target), static_cast<size_t>(volume), scale);
        else
            hipLaunchKernelGGL(advection_kernel<float>, grid, dim3(256), 0, as_stream(stream),
                               static_cast<const float*>(shared), static_cast<const float*>(derivative),
                               static_cast<float*>(target), static_cast<size_t>(volume), static_cast<float>(scale));
        check_hip(hipGetLastError(), "advection launch");
    });
}
