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

// The 16-point DFT (4 x 4: four 4-point DFTs over n1, then four over n2)
// with its inner twiddles in tangent form, after Linzer and Feig, "Modified
// FFTs for fused multiply-add architectures" (Math. Comp. 60, 1993):
// each twiddled input of a second-pass 4-point DFT is kappa * u, where u
// costs one FMA (or add) per component and kappa is a real scale.  In every
// second-pass DFT the two odd inputs share |kappa| (cos(pi/8) for k1 = 1
// and 3, sqrt(1/2) for k1 = 2), so both scales ride in the butterflies' FMAs:
// 80 instead of 96 operations for the second pass, where a twiddle product
// feeding two butterfly outputs cannot be contracted.  Natural order in and
// out, like rdft<16>.  The plane kernels use it; the two-pass kernels keep
// rdft<16>, on which it raises the register count of several instances.
template <typename T>
__device__ __forceinline__ T fmat(T a, T b, T c) {
    if constexpr (sizeof(T) == 4) return __builtin_fmaf(a, b, c);
    else return __builtin_fma(a, b, c);
}
template <typename T>
__device__ __forceinline__ cx<T> cfma(T k, cx<T> u, cx<T> t) {  // k * u + t
    return cx<T>{fmat(k, u.x, t.x), fmat(k, u.y, t.y)};
}
// a * (1 - s i r), s = +1 forward, -1 inverse
template <bool INV, typename T>
__device__ __forceinline__ cx<T> tan_form(cx<T> a, T r) {
    return INV ? cx<T>{fmat(-r, a.y, a.x), fmat(r, a.x, a.y)} : cx<T>{fmat(r, a.y, a.x), fmat(-r, a.x, a.y)};
}
// a * (r - s i)
template <bool INV, typename T>
__device__ __forceinline__ cx<T> cot_form(cx<T> a, T r) {
    return INV ? cx<T>{fmat(r, a.x, -a.y), fmat(r, a.y, a.x)} : cx<T>{fmat(r, a.x, a.y), fmat(r, a.y, -a.x)};
}
// a * (1 - s i)
template <bool INV, typename T>
__device__ __forceinline__ cx<T> f45(cx<T> a) {
    return INV ? cx<T>{a.x - a.y, a.y + a.x} : cx<T>{a.x + a.y, a.y - a.x};
}
// 4-point DFT of (a0, k2 u2 [or a2 exactly], k u1, e k u3), e = +-1
template <bool INV, typename T>
__device__ __forceinline__ void dft4_scaled(cx<T> t0, cx<T> t1, cx<T> u1, cx<T> u3, bool e, T k, cx<T>* out) {
    const cx<T> p = e ? u1 + u3 : u1 - u3, q = rot_i<INV>(e ? u1 - u3 : u1 + u3);
    out[0] = cfma(k, p, t0);
    out[8] = cfma(-k, p, t0);
    out[4] = cfma(k, q, t1);
    out[12] = cfma(-k, q, t1);
}
template <bool INV, typename T>
__device__ __forceinline__ void rdft16_fma(cx<T>* v) {
    const T C = static_cast<T>(0.92387953251128675612818318939679);  // cos(pi/8)
    const T Tn = static_cast<T>(0.41421356237309504880168872420970);  // tan(pi/8)
    const T H = static_cast<T>(0.70710678118654752440084436210485);  // sqrt(1/2)
    cx<T> y[4][4];  // y[n2][k1]: first pass
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        cx<T> a0 = v[n2], a1 = v[4 + n2], a2 = v[8 + n2], a3 = v[12 + n2];
        dft4<INV>(a0, a1, a2, a3);
        y[n2][0] = a0;
        y[n2][1] = a1;
        y[n2][2] = a2;
        y[n2][3] = a3;
    }
    // k1 = 0: no twiddles
    {
        cx<T> a0 = y[0][0], a1 = y[1][0], a2 = y[2][0], a3 = y[3][0];
        dft4<INV>(a0, a1, a2, a3);
        v[0] = a0;
        v[4] = a1;
        v[8] = a2;
        v[12] = a3;
    }
    // k1 = 1: W^1 = C (1 - s i Tn), W^2 = H (1 - s i), W^3 = C (Tn - s i)
    {
        const cx<T> u2 = f45<INV>(y[2][1]);
        dft4_scaled<INV>(cfma(H, u2, y[0][1]), cfma(-H, u2, y[0][1]), tan_form<INV>(y[1][1], Tn),
                         cot_form<INV>(y[3][1], Tn), true, C, v + 1);
    }
    // k1 = 2: W^2 = H (1 - s i), W^4 = -s i, W^6 = -H (1 + s i)
    {
        const cx<T> a2 = rot_i<INV>(y[2][2]);
        dft4_scaled<INV>(y[0][2] + a2, y[0][2] - a2, f45<INV>(y[1][2]), f45<!INV>(y[3][2]), false, H, v + 2);
    }
    // k1 = 3: W^3 = C (Tn - s i), W^6 = -H (1 + s i), W^9 = -C (1 - s i Tn)
    {
        const cx<T> u2 = f45<!INV>(y[2][3]);
        dft4_scaled<INV>(cfma(-H, u2, y[0][3]), cfma(H, u2, y[0][3]), cot_form<INV>(y[1][3], Tn),
                         tan_form<INV>(y[3][3], Tn), false, C, v + 3);
    }
}

// In-register DFT of length R in natural order (R = 2, 4, 8, 16, 32), the
// twiddles between its two passes applied as products.
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
