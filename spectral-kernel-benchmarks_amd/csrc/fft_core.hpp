// In-register DFT building blocks shared by the FFT kernels of libskb_hip.so:
// radix-2/4/8 butterflies, exactly symmetric twiddle constants and the
// natural-order register DFTs of length 2..32.
#pragma once

#include "common.hpp"

namespace skb {

// x * (-i) for the forward sign, x * (+i) for the inverse.
template <bool INV, typename R>
__device__ __forceinline__ cx<R> rot_i(cx<R> a) {
    return INV ? cx<R>{-a.y, a.x} : cx<R>{a.y, -a.x};
}

template <bool INV, typename R>
__device__ __forceinline__ void dft4(cx<R>& a0, cx<R>& a1, cx<R>& a2, cx<R>& a3) {
    const cx<R> t0 = a0 + a2, t1 = a0 - a2, t2 = a1 + a3, t3 = rot_i<INV>(a1 - a3);
    a0 = t0 + t2;
    a2 = t0 - t2;
    a1 = t1 + t3;
    a3 = t1 - t3;
}

template <int RAD, bool INV, typename R>
__device__ __forceinline__ void dft(cx<R>* v) {
    if constexpr (RAD == 2) {
        const cx<R> a = v[0], b = v[1];
        v[0] = a + b;
        v[1] = a - b;
    } else if constexpr (RAD == 4) {
        dft4<INV>(v[0], v[1], v[2], v[3]);
    } else {
        static_assert(RAD == 8, "radix");
        cx<R> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
        cx<R> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
        dft4<INV>(e0, e1, e2, e3);
        dft4<INV>(o0, o1, o2, o3);
        const R h = R(0.707106781186547524400844362104849039);
        // o_k *= exp(-+ 2 pi i k / 8)
        o1 = INV ? cx<R>{h * (o1.x - o1.y), h * (o1.x + o1.y)} : cx<R>{h * (o1.x + o1.y), h * (o1.y - o1.x)};
        o2 = rot_i<INV>(o2);
        o3 = INV ? cx<R>{-h * (o3.x + o3.y), h * (o3.x - o3.y)} : cx<R>{h * (o3.y - o3.x), -h * (o3.x + o3.y)};
        v[0] = e0 + o0;
        v[4] = e0 - o0;
        v[1] = e1 + o1;
        v[5] = e1 - o1;
        v[2] = e2 + o2;
        v[6] = e2 - o2;
        v[3] = e3 + o3;
        v[7] = e3 - o3;
    }
}

// cos(2 pi m / 32), exact symmetric values from the first octant.
__host__ __device__ constexpr double cos32(int m) {
    constexpr double oct[9] = {1.0,
                               0.98078528040323044912618223613424,
                               0.92387953251128675612818318939679,
                               0.83146961230254523707878837761791,
                               0.70710678118654752440084436210485,
                               0.55557023301960222474283081394853,
                               0.38268343236508977172845998403040,
                               0.19509032201612826784828486847702,
                               0.0};
    m &= 31;
    if (m <= 8) return oct[m];
    if (m <= 16) return -oct[16 - m];
    if (m <= 24) return -oct[m - 16];
    return oct[32 - m];
}

// x * exp(-+ 2 pi i m / R) for a compile-time-known m (after unrolling).
template <int R, bool INV, typename T>
__device__ __forceinline__ cx<T> twc(int m, cx<T> x) {
    m = ((m % R) + R) % R;
    if (m == 0) return x;
    if (4 * m == R) return rot_i<INV>(x);
    if (2 * m == R) return cx<T>{-x.x, -x.y};
    if (4 * m == 3 * R) return rot_i<!INV>(x);
    const int m32 = m * (32 / R);
    const T c = static_cast<T>(cos32(m32));
    const T sn = static_cast<T>(cos32(m32 - 8));  // sin(2 pi m / R)
    const cx<T> w{c, INV ? sn : -sn};
    return cmul(x, w);
}

template <int R>
struct RSplit;
template <>
struct RSplit<16> {
    static constexpr int A = 4;
};
template <>
struct RSplit<32> {
    static constexpr int A = 8;
};

// In-register DFT of length R in natural order (R = 2, 4, 8, 16, 32).
template <int R, bool INV, typename T>
__device__ __forceinline__ void rdft(cx<T>* v) {
    if constexpr (R == 2 || R == 4 || R == 8) {
        dft<R, INV>(v);
    } else {
        constexpr int A = RSplit<R>::A, B = R / A;
        cx<T> y[R];
#pragma unroll
        for (int n2 = 0; n2 < B; ++n2) {
            cx<T> u[A];
#pragma unroll
            for (int n1 = 0; n1 < A; ++n1) u[n1] = v[B * n1 + n2];
            rdft<A, INV>(u);
#pragma unroll
            for (int k1 = 0; k1 < A; ++k1) y[k1 * B + n2] = twc<R, INV>(n2 * k1, u[k1]);
        }
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) {
            cx<T> w[B];
#pragma unroll
            for (int n2 = 0; n2 < B; ++n2) w[n2] = y[k1 * B + n2];
            rdft<B, INV>(w);
#pragma unroll
            for (int k2 = 0; k2 < B; ++k2) v[k1 + A * k2] = w[k2];
        }
    }
}

}  // namespace skb
