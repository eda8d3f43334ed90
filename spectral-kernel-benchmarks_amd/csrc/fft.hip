// Batched 2-D real FFTs in WVM order and the pruned retained-mode transforms
// for gfx950.
//
// Design (all sizes powers of two, 4..1024 per axis):
//   * Every pass works on a TILE OF CONSECUTIVE PLANES, so global traffic in
//     the plane-fastest WVM / retained layouts is contiguous runs of
//     tile*sizeof(complex) = 128 B.
//   * Row pass (x): a workgroup owns one row index y of `tile` planes.  Two
//     real rows are packed as one complex sequence (x_a + i x_b), transformed
//     with a mixed-radix Stockham FFT in LDS and separated by Hermitian
//     symmetry.  r2c keeps only kx < kc (kc = active k_x band for the pruned
//     transform, Nx/2+1 for the full one); c2r synthesises the zero band in
//     registers and never reads it.
//   * Column pass (y): a workgroup owns one kx column of `tile` planes.  Input
//     is either the strided intermediate / WVM spectrum or, for the pruned
//     inverse, a gather straight from the compact retained array (with the
//     k_x = 0 Hermitian partners); output is the strided array or a scatter
//     into the compact retained array.  Rows that hold no retained mode are
//     never read or written.
//   * Intermediate between the passes ("T"): z + Nz*field + P*(kx + kc*y).
//     For the full transforms T IS the WVM spectrum (kc = Nx/2+1) and the
//     column pass runs in place.
//   * Twiddles come from FP64 host tables rounded once to the working
//     precision and staged in LDS.
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "common.hpp"
#include "fft_core.hpp"
#include "internal.hpp"
#include "plane.hpp"

namespace skb {

// ---------------------------------------------------------------------------
// Stockham FFT in LDS
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int pick_radix(int n) {
    return (n % 8 == 0 && n != 16 && n != 32) ? 8 : (n % 4 == 0 ? 4 : 2);
}


// One Stockham stage of radix RAD over S sequences of length N stored at
// buf[s*SEQ ...].  NS = product of the radices already applied.  The stage
// reads every input before the first barrier and writes after it, so it runs
// in place.  Ends with a barrier.
template <typename R, int N, int SEQ, int RAD, int NS, int S, int NT, bool INV>
__device__ __forceinline__ void stockham_stage(cx<R>* buf, const cx<R>* tw) {
    constexpr int T = N / RAD;
    constexpr int TOTAL = S * T;
    constexpr int PER = (TOTAL + NT - 1) / NT;
    cx<R> v[PER][RAD];
    const int tid = threadIdx.x;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int b = tid + i * NT;
        if (TOTAL % NT == 0 || b < TOTAL) {
            const int s = b / T, j = b - s * T;
#pragma unroll
            for (int r = 0; r < RAD; ++r) v[i][r] = buf[s * SEQ + j + r * T];
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int b = tid + i * NT;
        if (TOTAL % NT == 0 || b < TOTAL) {
            const int s = b / T, j = b - s * T;
            const int t = j & (NS - 1);
            if constexpr (NS > 1) {
#pragma unroll
                for (int r = 1; r < RAD; ++r) {
                    const cx<R> w = tw[r * t * (N / (NS * RAD))];
                    v[i][r] = INV ? cmulc(v[i][r], w) : cmul(v[i][r], w);
                }
            }
            dft<RAD, INV>(v[i]);
            const int base = s * SEQ + (j / NS) * NS * RAD + t;
#pragma unroll
            for (int r = 0; r < RAD; ++r) buf[base + r * NS] = v[i][r];
        }
    }
    __syncthreads();
}

template <typename R, int N, int SEQ, int REM, int NS, int S, int NT, bool INV>
__device__ __forceinline__ void stockham(cx<R>* buf, const cx<R>* tw) {
    if constexpr (REM > 1) {
        constexpr int RAD = pick_radix(REM);
        stockham_stage<R, N, SEQ, RAD, NS, S, NT, INV>(buf, tw);
        stockham<R, N, SEQ, REM / RAD, NS * RAD, S, NT, INV>(buf, tw);
    }
}

constexpr int kThreads = 256;

template <typename R>
constexpr int row_tile() {
    return sizeof(R) == 8 ? 8 : 16;  // planes per row workgroup (even)
}
template <typename R, int N>
constexpr int col_tile() {
    // planes per column workgroup; keeps the LDS tile at <= ~64 KiB
    return sizeof(R) == 8 ? (N >= 1024 ? 4 : 8) : (N >= 1024 ? 8 : 16);
}

// Entry of a column's retained-mode list: the row (ky) it occupies, the
// compact mode index and whether the compact value is the conjugate of the
// stored half-spectrum value.
struct ColEntry {
    int row;
    int mode;
    int conj;
    int pad;
};

// ---------------------------------------------------------------------------
// Row pass: r2c along x.  real (x fastest, y, plane) -> T[p + P*(kx + kc*y)]
// ---------------------------------------------------------------------------
template <typename R, int NX>
__global__ __launch_bounds__(kThreads) void row_r2c_kernel(const R* __restrict__ real, cx<R>* __restrict__ out,
                                                          const cx<R>* __restrict__ tw, int ny, int planes, int kc) {
    constexpr int TP = row_tile<R>();
    constexpr int S = TP / 2;
    __shared__ cx<R> buf[S * NX];
    __shared__ cx<R> twl[NX];
    const int y = blockIdx.x;
    const int p0 = blockIdx.y * TP;
    for (int i = threadIdx.x; i < NX; i += kThreads) twl[i] = tw[i];
    for (int e = threadIdx.x; e < TP * NX; e += kThreads) {
        const int pit = e / NX, n = e - pit * NX;
        const int p = p0 + pit;
        const R v = p < planes ? real[(static_cast<size_t>(p) * ny + y) * NX + n] : R(0);
        reinterpret_cast<R*>(&buf[(pit >> 1) * NX + n])[pit & 1] = v;
    }
    __syncthreads();
    stockham<R, NX, NX, NX, 1, S, kThreads, false>(buf, twl);
    const size_t row_base = static_cast<size_t>(planes) * static_cast<size_t>(kc) * y;
    for (int e = threadIdx.x; e < kc * TP; e += kThreads) {
        const int pit = e % TP, kx = e / TP;
        const int p = p0 + pit;
        if (p >= planes) continue;
        const cx<R> zk = buf[(pit >> 1) * NX + kx];
        const cx<R> zn = buf[(pit >> 1) * NX + ((NX - kx) & (NX - 1))];
        // X_a = (Z_k + conj Z_{N-k})/2,  X_b = (Z_k - conj Z_{N-k})/(2i)
        const cx<R> x = (pit & 1) == 0 ? cx<R>{R(0.5) * (zk.x + zn.x), R(0.5) * (zk.y - zn.y)}
                                       : cx<R>{R(0.5) * (zk.y + zn.y), R(0.5) * (zn.x - zk.x)};
        out[row_base + static_cast<size_t>(planes) * kx + p] = x;
    }
}

// ---------------------------------------------------------------------------
// Row pass: c2r along x.  T[p + P*(kx + kc*y)] (kx < kc, zero above) -> real.
// The imaginary parts at kx = 0 and kx = Nx/2 are ignored, as FFTW's c2r does.
// ---------------------------------------------------------------------------
template <typename R, int NX>
__global__ __launch_bounds__(kThreads) void row_c2r_kernel(const cx<R>* __restrict__ in, R* __restrict__ real,
                                                          const cx<R>* __restrict__ tw, int ny, int planes, int kc) {
    constexpr int TP = row_tile<R>();
    constexpr int S = TP / 2;
    constexpr int H = NX / 2 + 1;
    __shared__ cx<R> buf[S * NX];
    __shared__ cx<R> twl[NX];
    const int y = blockIdx.x;
    const int p0 = blockIdx.y * TP;
    for (int i = threadIdx.x; i < NX; i += kThreads) twl[i] = tw[i];
    const size_t row_base = static_cast<size_t>(planes) * static_cast<size_t>(kc) * y;
    for (int e = threadIdx.x; e < H * S; e += kThreads) {
        const int s = e % S, kx = e / S;
        const int pa = p0 + 2 * s, pb = pa + 1;
        cx<R> xa{R(0), R(0)}, xb{R(0), R(0)};
        if (kx < kc) {
            const size_t base = row_base + static_cast<size_t>(planes) * kx;
            if (pa < planes) xa = in[base + pa];
            if (pb < planes) xb = in[base + pb];
        }
        const bool self = kx == 0 || kx == NX / 2;
        if (self) {
            xa.y = R(0);
            xb.y = R(0);
        }
        // Z_k = X_a + i X_b ;  Z_{N-k} = conj X_a + i conj X_b
        buf[s * NX + kx] = cx<R>{xa.x - xb.y, xa.y + xb.x};
        if (!self) buf[s * NX + NX - kx] = cx<R>{xa.x + xb.y, xb.x - xa.y};
    }
    __syncthreads();
    stockham<R, NX, NX, NX, 1, S, kThreads, true>(buf, twl);
    for (int e = threadIdx.x; e < TP * NX; e += kThreads) {
        const int pit = e / NX, n = e - pit * NX;
        const int p = p0 + pit;
        if (p >= planes) continue;
        const cx<R> z = buf[(pit >> 1) * NX + n];
        real[(static_cast<size_t>(p) * ny + y) * NX + n] = (pit & 1) ? z.y : z.x;
    }
}

// ---------------------------------------------------------------------------
// Column pass: c2c along y for column kx = blockIdx.x of `tile` planes.
// ---------------------------------------------------------------------------
enum ColIO { kStrided = 0, kRetained = 1, kSplit = 2 };

// Compact retained operand given as one view per field (kSplit).  Plane
// p = z + nz * f of mode m lives at re[f][zs*z + stride[f]*m] and
// im[f][...] (R elements).  Split arrays have zs = 1 and im apart from re;
// interleaved complex arrays have zs = 2 and im = re + 1.  The plain split
// layout (plane + P * mode) is one view with nz = P and stride = P.  Forward
// outputs are multiplied by `scale` (src/pruned_fftw_provider.cpp:1053-1083).
// Inputs may carry a spectral derivative factor per view, applied to the
// stored value at column kx, row ky: i*kx_scale*kx (kFactorDx) or
// i*ky_scale*l with l the signed row (kFactorDy) -- the stored-coordinate form
// of i*k*q at the logical (k, l), conjugation included.
constexpr int kMaxSplitViews = 16;
template <typename R>
struct SplitViews {
    R* re[kMaxSplitViews];
    R* im[kMaxSplitViews];
    long long stride[kMaxSplitViews];
    int factor[kMaxSplitViews];
    int nz, zs, ny;
    R scale, kx_scale, ky_scale;
};

// The view of plane p, with z folded into the pointers (lanes hold one plane).
template <typename R>
struct ViewLane {
    R* re;
    R* im;
    size_t stride;
    int factor;
};
template <typename R>
__device__ __forceinline__ ViewLane<R> view_lane(const SplitViews<R>& v, int p) {
    const int f = p / v.nz, z = p - f * v.nz;
    ViewLane<R> out{v.re[0], v.im[0], static_cast<size_t>(v.stride[0]), v.factor[0]};
    // unrolled select: a lane-varying index into the argument array would go
    // through private memory
#pragma unroll
    for (int i = 1; i < kMaxSplitViews; ++i)
        if (f == i) out = ViewLane<R>{v.re[i], v.im[i], static_cast<size_t>(v.stride[i]), v.factor[i]};
    out.re += static_cast<size_t>(v.zs) * z;
    out.im += static_cast<size_t>(v.zs) * z;
    return out;
}

// x * i * a for the view's derivative factor at stored (kx, row).
template <typename R>
__device__ __forceinline__ cx<R> apply_factor(const SplitViews<R>& v, int factor, int kx, int row, cx<R> x) {
    if (factor == kFactorNone) return x;
    const R a = factor == kFactorDx ? v.kx_scale * static_cast<R>(kx)
                                    : v.ky_scale * static_cast<R>(2 * row < v.ny ? row : row - v.ny);
    return cx<R>{-a * x.y, a * x.x};
}

template <typename R, int NY, bool INV, int IN, int OUT>
__global__ __launch_bounds__(kThreads) void col_kernel(const cx<R>* in, cx<R>* out, const cx<R>* __restrict__ tw,
                                                      int planes, int in_kc, int out_kc,
                                                      const int* __restrict__ col_start,
                                                      const ColEntry* __restrict__ entries, const SplitViews<R> views) {
    constexpr int TP = col_tile<R, NY>();
    constexpr int SEQ = NY + 1;  // one-element pad: plane-fastest LDS writes hit distinct banks
    __shared__ cx<R> buf[TP * SEQ];
    __shared__ cx<R> twl[NY];
    const int kx = blockIdx.x;
    const int p0 = blockIdx.y * TP;
    const size_t P = static_cast<size_t>(planes);
    for (int i = threadIdx.x; i < NY; i += kThreads) twl[i] = tw[i];
    if constexpr (IN == kStrided) {
        for (int e = threadIdx.x; e < NY * TP; e += kThreads) {
            const int pit = e % TP, yy = e / TP;
            const int p = p0 + pit;
            buf[pit * SEQ + yy] = p < planes ? in[P * (kx + static_cast<size_t>(in_kc) * yy) + p] : cx<R>{R(0), R(0)};
        }
    } else {
        for (int e = threadIdx.x; e < TP * SEQ; e += kThreads) buf[e] = cx<R>{R(0), R(0)};
        __syncthreads();
        const int beg = col_start[kx], cnt = col_start[kx + 1] - beg;
        for (int e = threadIdx.x; e < cnt * TP; e += kThreads) {
            const int pit = e % TP;
            const ColEntry ent = entries[beg + e / TP];
            const int p = p0 + pit;
            if (p >= planes) continue;
            cx<R> v;
            if constexpr (IN == kSplit) {
                const ViewLane<R> vl = view_lane(views, p);
                v = cx<R>{vl.re[vl.stride * ent.mode], vl.im[vl.stride * ent.mode]};
                if (ent.conj) v.y = -v.y;
                v = apply_factor(views, vl.factor, kx, ent.row, v);
            } else {
                v = in[P * ent.mode + p];
                if (ent.conj) v.y = -v.y;
            }
            buf[pit * SEQ + ent.row] = v;
        }
    }
    __syncthreads();
    stockham<R, NY, SEQ, NY, 1, TP, kThreads, INV>(buf, twl);
    if constexpr (OUT == kStrided) {
        for (int e = threadIdx.x; e < NY * TP; e += kThreads) {
            const int pit = e % TP, yy = e / TP;
            const int p = p0 + pit;
            if (p < planes) out[P * (kx + static_cast<size_t>(out_kc) * yy) + p] = buf[pit * SEQ + yy];
        }
    } else {
        const int beg = col_start[kx], cnt = col_start[kx + 1] - beg;
        for (int e = threadIdx.x; e < cnt * TP; e += kThreads) {
            const int pit = e % TP;
            const ColEntry ent = entries[beg + e / TP];
            const int p = p0 + pit;
            if (p >= planes) continue;
            cx<R> v = buf[pit * SEQ + ent.row];
            if (ent.conj) v.y = -v.y;
            if constexpr (OUT == kSplit) {
                const ViewLane<R> vl = view_lane(views, p);
                vl.re[vl.stride * ent.mode] = views.scale * v.x;
                vl.im[vl.stride * ent.mode] = views.scale * v.y;
            } else {
                out[P * ent.mode + p] = v;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Two-pass register FFT for N = N1 * N2 (N >= 64)
//
// X[k1 + N1 k2] = sum_n2 W_N^(n2 k1) W_N2^(n2 k2) sum_n1 x[N2 n1 + n2] W_N1^(n1 k1)
// Pass 1: a thread owns one n2 (of one sequence), loads x[N2 n1 + n2] for all
//         n1 straight from global memory, does a length-N1 DFT in registers
//         and applies W_N^(n2 k1).
// Pass 2: after ONE transpose through LDS a thread owns one k1, does a
//         length-N2 DFT in registers and stores X[k1 + N1 k2] straight to
//         global memory.
// One LDS round trip and one barrier per transform (plus the twiddle table).
// ---------------------------------------------------------------------------


template <int N>
struct Split2 {
    // N1 >= N2, both <= 32
    static constexpr int N1 = N == 64 ? 8 : (N == 128 ? 16 : (N == 256 ? 16 : 32));
    static constexpr int N2 = N / N1;
};

// Column kernels: planes per workgroup and the workgroup size; each global
// access run stays at >= 128 B.
template <typename R, int N>
struct Geo2 {
    static constexpr int N1 = Split2<N>::N1, N2 = Split2<N>::N2;
    static constexpr int LANES = (sizeof(R) == 4 ? 256 : 128);
    static constexpr int TP = LANES / N1;       // planes per workgroup
    static constexpr int THREADS = LANES;       // TP * N1 >= TP * N2
};

// Row kernels: plane pairs per workgroup and the workgroup size.
constexpr int kRowLanesF32 = 256;
template <typename R, int N>
struct RowGeo {
    static constexpr int N1 = Split2<N>::N1, N2 = Split2<N>::N2;
    static constexpr int THREADS = sizeof(R) == 4 ? kRowLanesF32 : 128;
    static constexpr int PAIRS = THREADS / N1;  // THREADS >= PAIRS * N2, as N2 <= N1
};

// First work item of a workgroup.  One workgroup per item (LOOP = false):
// the XCD-aware mapping; workgroups past the last item exit at once.
// Persistent workgroups (LOOP = true) take items blockIdx.x, blockIdx.x +
// gridDim.x, ... and prefetch the next item's inputs into registers.
template <bool LOOP>
__device__ __forceinline__ int first_item(int items) {
    return LOOP ? static_cast<int>(blockIdx.x) : xcd_item(static_cast<int>(blockIdx.x), items);
}

// Row passes.  A work item is one row index y of a tile of TP = 2*PAIRS
// consecutive planes; item = y * tiles + tile, so the workgroups running at
// the same time cover whole plane runs of the same rows (DRAM-page locality
// in the plane-fastest intermediate).  Persistent workgroups prefetch
// the next item's inputs into registers while transforming the current one.
template <typename R, int NX>
__device__ __forceinline__ void r2c_load(cx<R>* v, const R* real, int item, int ny, int planes) {
    using G = RowGeo<R, NX>;
    constexpr int N1 = G::N1, N2 = G::N2, S = G::PAIRS, TP = 2 * S;
    const int t = threadIdx.x, n2 = t % N2, s1 = t / N2;
    const int ntiles = (planes + TP - 1) / TP;
    const int y = item / ntiles, p0 = (item % ntiles) * TP;
    const int pa = p0 + 2 * s1, pb = pa + 1;
    const bool act = s1 < S;
    // Branch-free: a missing plane reads row 0 of plane 0 (in bounds) and
    // selects zero.
    const bool has_a = act && pa < planes, has_b = act && pb < planes;
    const R* ra = has_a ? real + (static_cast<size_t>(pa) * ny + y) * NX + n2 : real;
    const R* rb = has_b ? real + (static_cast<size_t>(pb) * ny + y) * NX + n2 : real;
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) {
        const R a = ra[N2 * n1], b = rb[N2 * n1];
        v[n1].x = has_a ? a : R(0);
        v[n1].y = has_b ? b : R(0);
    }
}

// Spectrum side: out[ps * (kx + kc * y) + p]; ps >= planes is the plane
// stride (planes for WVM order, padded to whole cache lines for the pruned
// intermediate).

// Row-pass transpose between the two register passes, through LDS in two
// rounds (real parts, then imaginary parts) over one array of R: half the LDS
// of a complex buffer, so the row kernels are limited by registers rather
// than LDS (7 rather than 4 workgroups per CU at Nx = 256, FP32).  Thread
// `put` writes z[k] at b[wo + k*YS]; every thread then reads w[m] from
// b[ro + m].  YS = N2 + 1 and a pair stride of N1*YS + 2 keep the writes
// and the reads of both row kernels' thread mappings within one or two
// cycles of the bank minimum.
template <typename R, int NW, int NR, int YS>
__device__ __forceinline__ void row_transpose(R* b, const cx<R>* z, bool put, int wo, int ro, cx<R>* w) {
    if (put) {
#pragma unroll
        for (int k = 0; k < NW; ++k) b[wo + k * YS] = z[k].x;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < NR; ++m) w[m].x = b[ro + m];
    __syncthreads();
    if (put) {
#pragma unroll
        for (int k = 0; k < NW; ++k) b[wo + k * YS] = z[k].y;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < NR; ++m) w[m].y = b[ro + m];
}

// The value held by lane l ^ S (S = 16 or 32: rows of S lanes swapped in
// pairs) through gfx950's v_permlane16/32_swap, a vector-ALU move: with both
// operands the same register, the swap's first result holds the even row in
// both rows of a pair and its second result the odd row, so a lane of an
// odd row (`odd`) takes the first.  Other S go through ds_bpermute.
template <int S>
__device__ __forceinline__ uint32_t lane_xor_u32(uint32_t v, bool odd) {
    if constexpr (S == 16) {
        const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        return odd ? r[0] : r[1];
    } else if constexpr (S == 32) {
        const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
        return odd ? r[0] : r[1];
    } else {
        return static_cast<uint32_t>(__shfl_xor(static_cast<int>(v), S));
    }
}
template <int S>
__device__ __forceinline__ float lane_xor(float v, bool odd) {
    return __uint_as_float(lane_xor_u32<S>(__float_as_uint(v), odd));
}
template <int S>
__device__ __forceinline__ double lane_xor(double v, bool odd) {
    const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(v));
    const uint64_t lo = lane_xor_u32<S>(static_cast<uint32_t>(b), odd);
    const uint64_t hi = lane_xor_u32<S>(static_cast<uint32_t>(b >> 32), odd);
    return __longlong_as_double(static_cast<long long>((hi << 32) | lo));
}

template <typename R, int NX, bool LOOP>
__global__ __launch_bounds__((RowGeo<R, NX>::THREADS)) void row2_r2c_kernel(const R* __restrict__ real,
                                                                         cx<R>* __restrict__ out,
                                                                         const cx<R>* __restrict__ tw, int ny,
                                                                         int planes, int ps, int kc, int items) {
    using G = RowGeo<R, NX>;
    constexpr int N1 = G::N1, N2 = G::N2, S = G::PAIRS, TP = 2 * S, NT = G::THREADS;
    constexpr int YS = N2 + 1;       // row of the transpose buffer
    constexpr int SS = N1 * YS + 2;  // one pair's transpose buffer
    __shared__ R buf[S * SS];
    __shared__ cx<R> twl[NX];
    const int t = threadIdx.x;
    int item = first_item<LOOP>(items);
    if (item >= items) return;  // uniform across the workgroup
    cx<R> v[N1];
    r2c_load<R, NX>(v, real, item, ny, planes);
    for (int i = t; i < NX; i += NT) twl[i] = tw[i];
    __syncthreads();
    const int n2 = t % N2, s1 = t / N2;
    // Pass 2: thread (pair s2 fastest, k1).  Rows of S lanes take k1 in the
    // order 0, N1/2, 1, N1-1, 2, N1-2, ..., so the threads of k1 and N1 - k1
    // are S lanes apart in one wavefront (64/S >= 2 rows per wavefront).
    const int s2 = t % S, kq = t / S;
    const int kj = kq >> 1;
    const int k1 = kj == 0 ? ((kq & 1) ? N1 / 2 : 0) : ((kq & 1) ? N1 - kj : kj);
    static_assert(64 % S == 0 && 64 / S >= 2, "partner rows must share a wavefront");
    const int ntiles = (planes + TP - 1) / TP;
    for (;;) {
        const int next = LOOP ? item + static_cast<int>(gridDim.x) : items;
        cx<R> vn[LOOP ? N1 : 1];
        if constexpr (LOOP) {
            if (next < items) r2c_load<R, NX>(vn, real, next, ny, planes);
        }
        const int y = item / ntiles, p0 = (item % ntiles) * TP;
        // pass 1: thread (n2 fastest, pair s)
        if (s1 < S) {
            rdft<N1, false>(v);
#pragma unroll
            for (int k = 0; k < N1; ++k) v[k] = cmul(v[k], twl[n2 * k]);
        }
        // pass 2: thread (k1 fastest, pair s)
        cx<R> w[N2];
        row_transpose<R, N1, N2, YS>(buf, v, s1 < S, s1 * SS + n2, s2 * SS + k1 * YS, w);
        rdft<N2, false>(w);
        // Thread (k1, s2) now holds Z[k1 + N1*k2] of pair s2.  The Hermitian
        // split X_a = (Z_k + conj Z_{N-k})/2, X_b = (Z_k - conj Z_{N-k})/(2i)
        // takes Z_{N-k} from the thread of k1' = (N1 - k1) mod N1 and the same
        // pair (its k2' = N2-1-k2), which sits S lanes away in the same
        // wavefront (see the pass-2 mapping above); k1 = 0 and k1 = N1/2 are
        // their own partners.  A wavefront holds S consecutive plane pairs of
        // 64/S values of k1, so each store instruction writes runs of S
        // plane pairs (256 bytes in FP32) per kx.  Each store is one 16-byte
        // plane pair (FP32, even ps, both planes present) or two elements,
        // through a buffer resource over the row: an element past kc or past
        // the last plane goes past the row and is dropped.  Planes past the
        // last one are never written (a strided WVM field view interleaves
        // other fields there).
        {
            constexpr uint32_t CB = sizeof(cx<R>);
            const int yu = __builtin_amdgcn_readfirstlane(y), p0u = __builtin_amdgcn_readfirstlane(p0);
            const __amdgpu_buffer_rsrc_t rs =
                buffer_rsrc(out + static_cast<size_t>(ps) * static_cast<size_t>(kc) * yu,
                            static_cast<uint32_t>(ps) * kc * CB);
            const int pa = p0u + 2 * s2;
            const bool has_a = pa < planes, has_b = pa + 1 < planes;
            const bool wide = sizeof(R) == 4 && (ps & 1) == 0;  // uniform
#pragma unroll
            for (int k2 = 0; k2 <= N2 / 2; ++k2) {
                if (N1 * k2 >= kc) break;  // uniform
                const int kx = k1 + N1 * k2;
                const cx<R> zk = w[k2];
                const cx<R> pz = w[N2 - 1 - k2];
                cx<R> zn{lane_xor<S>(pz.x, kq & 1), lane_xor<S>(pz.y, kq & 1)};
                if (2 * k1 == N1) zn = pz;
                if (k1 == 0) zn = w[(N2 - k2) & (N2 - 1)];
                const cx<R> xa{R(0.5) * (zk.x + zn.x), R(0.5) * (zk.y - zn.y)};
                const cx<R> xb{R(0.5) * (zk.y + zn.y), R(0.5) * (zn.x - zk.x)};
                const bool ok = kx < kc && has_a;
                const uint32_t off = ok ? static_cast<uint32_t>(ps * kx + pa) * CB : kBufferPast;
                if constexpr (sizeof(R) == 4) {
                    if (wide && has_b) {
                        raw_buffer_store_f4(bf4{xa.x, xa.y, xb.x, xb.y}, rs, static_cast<int>(off), 0, 0);
                        continue;
                    }
                }
                buffer_store_cx<R>(xa, rs, off);
                buffer_store_cx<R>(xb, rs, ok && has_b ? off + CB : kBufferPast);
            }
        }
        if (!LOOP || next >= items) break;
        __syncthreads();  // buf is rewritten by the next item
        if constexpr (LOOP) {
#pragma unroll
            for (int m = 0; m < N1; ++m) v[m] = vn[m];
        }
        item = next;
    }
}

// c2r pass-1 input: Z[k] = X_a[k] + i X_b[k] with the Hermitian extension;
// Im of the kx = 0 and kx = Nx/2 bins is ignored and kx >= kc is zero (never
// read).  Thread (pair s fastest, n2).
// Core: the pair (pa, pa + 1) of row y (row_base = ps*kc*y, wave-uniform);
// has_a / has_b say which of the two planes exist.  Thread (s1 = t % S,
// n2 = t / S).  Loads go through a buffer resource over the row's ps*kc
// elements: a lane with nothing to read passes an offset past the end and
// gets zero, so the 16 loads carry neither branches nor selects.  WIDE
// (FP32, even ps and pa): each pair is one 16-byte load; plane pa + 1 is then
// padding (ps > planes) when has_b is false, and is selected away.  Callers
// pick WIDE from wave-uniform values.
template <typename R, int NX, bool WIDE>
__device__ __forceinline__ void c2r_gather_t(cx<R>* v, const cx<R>* in, size_t row_base, int ps, int kc, int pa,
                                             bool has_a, bool has_b) {
    using G = RowGeo<R, NX>;
    constexpr int N1 = G::N1, N2 = G::N2, S = G::PAIRS;
    constexpr uint32_t CB = sizeof(cx<R>);
    const int n2 = static_cast<int>(threadIdx.x) / S;
    const bool act = n2 < N2 && has_a;
    const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(row_base));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(row_base >> 32));
    const __amdgpu_buffer_rsrc_t rs =
        buffer_rsrc(in + ((static_cast<size_t>(hi) << 32) | lo), static_cast<uint32_t>(ps) * kc * CB);
    constexpr bool wide = WIDE && sizeof(R) == 4;
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) {
        const int k = N2 * n1 + n2;
        // k = N2*n1 + n2 with n2 < N2: only n1 = N1/2 straddles Nx/2, and only
        // n1 = 0 (k = 0) and n1 = N1/2 (k = Nx/2) can hold a self-conjugate bin
        const bool upper = 2 * n1 > N1 || (2 * n1 == N1 && n2 > 0);
        const int kx = upper ? NX - k : k;
        const bool ok = act && kx < kc;
        const uint32_t off = ok ? static_cast<uint32_t>(ps * kx + pa) * CB : kBufferPast;
        cx<R> xa, xb;
        if constexpr (wide) {
            const bf4 q = raw_buffer_load_f4(rs, static_cast<int>(off), 0, 0);
            xa = cx<R>{q.x, q.y};
            xb = has_b ? cx<R>{q.z, q.w} : cx<R>{R(0), R(0)};
        } else {
            xa = buffer_load_cx<R>(rs, off);
            xb = buffer_load_cx<R>(rs, ok && has_b ? off + CB : kBufferPast);
        }
        if ((n1 == 0 || 2 * n1 == N1) && (kx == 0 || kx == NX / 2)) {
            xa.y = R(0);
            xb.y = R(0);
        }
        v[n1] = upper ? cx<R>{xa.x + xb.y, xb.x - xa.y} : cx<R>{xa.x - xb.y, xa.y + xb.x};
    }
}

// The same gather through flat global addresses (absent elements read row
// element 0 and are selected to zero), for the fused flux row kernel: at one
// wavefront per SIMD and a full register file there, this form measured
// faster than the buffer form above (512^2 x Nz129 flux 1.52 against 1.63 ms).
template <typename R, int NX, bool WIDE>
__device__ __forceinline__ void c2r_gather_flat_t(cx<R>* v, const cx<R>* in, size_t row_base, int ps, int kc,
                                                  int pa, bool has_a, bool has_b) {
    using G = RowGeo<R, NX>;
    constexpr int N1 = G::N1, N2 = G::N2, S = G::PAIRS;
    const int n2 = static_cast<int>(threadIdx.x) / S;
    const bool act = n2 < N2 && has_a;
    const cx<R>* row = in + row_base;
    constexpr bool wide = WIDE && sizeof(R) == 4;
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) {
        const int k = N2 * n1 + n2;
        const bool upper = k > NX / 2;
        const int kx = upper ? NX - k : k;
        const bool ok = act && kx < kc;
        const unsigned off = ok ? static_cast<unsigned>(ps * kx + pa) : 0u;
        cx<R> xa, xb;
        if constexpr (wide) {
            const float4 q = *reinterpret_cast<const float4*>(row + off);
            xa = ok ? cx<R>{q.x, q.y} : cx<R>{R(0), R(0)};
            xb = (ok && has_b) ? cx<R>{q.z, q.w} : cx<R>{R(0), R(0)};
        } else {
            const cx<R> a = row[off];
            const cx<R> b = row[(ok && has_b) ? off + 1 : 0u];
            xa = ok ? a : cx<R>{R(0), R(0)};
            xb = (ok && has_b) ? b : cx<R>{R(0), R(0)};
        }
        if (kx == 0 || kx == NX / 2) {
            xa.y = R(0);
            xb.y = R(0);
        }
        v[n1] = upper ? cx<R>{xa.x + xb.y, xb.x - xa.y} : cx<R>{xa.x - xb.y, xa.y + xb.x};
    }
}

// wide: FP32 with ps and the tile's first plane both even (wave-uniform).
template <typename R, int NX>
__device__ __forceinline__ void c2r_gather(cx<R>* v, const cx<R>* in, size_t row_base, int ps, int kc, int pa,
                                           bool has_a, bool has_b, bool wide) {
    if (sizeof(R) == 4 && wide) c2r_gather_t<R, NX, true>(v, in, row_base, ps, kc, pa, has_a, has_b);
    else c2r_gather_t<R, NX, false>(v, in, row_base, ps, kc, pa, has_a, has_b);
}

template <typename R, int NX>
__device__ __forceinline__ void c2r_load(cx<R>* v, const cx<R>* in, int item, int ny, int planes, int ps, int kc) {
    using G = RowGeo<R, NX>;
    constexpr int S = G::PAIRS, TP = 2 * S;
    const int s1 = static_cast<int>(threadIdx.x) % S;
    const int ntiles = (planes + TP - 1) / TP;
    const int y = item / ntiles, p0 = (item % ntiles) * TP;
    const int pa = p0 + 2 * s1;
    c2r_gather<R, NX>(v, in, static_cast<size_t>(ps) * static_cast<size_t>(kc) * y, ps, kc, pa, pa < planes,
                      pa + 1 < planes, (ps & 1) == 0);
}

template <typename R, int NX, bool LOOP>
__global__ __launch_bounds__((RowGeo<R, NX>::THREADS)) void row2_c2r_kernel(const cx<R>* __restrict__ in,
                                                                         R* __restrict__ real,
                                                                         const cx<R>* __restrict__ tw, int ny,
                                                                         int planes, int ps, int kc, int items) {
    using G = RowGeo<R, NX>;
    constexpr int N1 = G::N1, N2 = G::N2, S = G::PAIRS, TP = 2 * S, NT = G::THREADS;
    constexpr int YS = N2 + 1;
    constexpr int SS = N1 * YS + 2;
    __shared__ R buf[S * SS];
    __shared__ cx<R> twl[NX];
    const int t = threadIdx.x;
    int item = first_item<LOOP>(items);
    if (item >= items) return;  // uniform across the workgroup
    cx<R> v[N1];
    c2r_load<R, NX>(v, in, item, ny, planes, ps, kc);
    for (int i = t; i < NX; i += NT) twl[i] = tw[i];
    __syncthreads();
    const int s1 = t % S, n2 = t / S;
    const int k1 = t % N1, s2 = t / N1;
    const int ntiles = (planes + TP - 1) / TP;
    for (;;) {
        const int next = LOOP ? item + static_cast<int>(gridDim.x) : items;
        cx<R> vn[LOOP ? N1 : 1];
        if constexpr (LOOP) {
            if (next < items) c2r_load<R, NX>(vn, in, next, ny, planes, ps, kc);
        }
        // y and p0 are wave-uniform, but the division leaves them in vector
        // registers; without readfirstlane the buffer resource built from
        // them below is taken as divergent and every store is wrapped in a
        // waterfall loop.
        const int y = __builtin_amdgcn_readfirstlane(item / ntiles);
        const int p0 = __builtin_amdgcn_readfirstlane((item % ntiles) * TP);
        // pass 1: thread (pair s fastest, n2)
        if (n2 < N2) {
            rdft<N1, true>(v);
#pragma unroll
            for (int k = 0; k < N1; ++k) v[k] = cmulc(v[k], twl[n2 * k]);
        }
        // pass 2: thread (k1 fastest, pair s) -> real rows of planes 2s, 2s+1
        cx<R> w[N2];
        row_transpose<R, N1, N2, YS>(buf, v, n2 < N2, s1 * SS + n2, s2 * SS + k1 * YS, w);
        rdft<N2, true>(w);
        // Bounds-checked stores over the tile's rows (planes past the last
        // are dropped), so every thread issues 2*N2 stores: the wait for the
        // prefetched item then does not include them.
        {
            const int pa = p0 + 2 * s2;
            const R* base = real + (static_cast<size_t>(p0) * ny + y) * NX;
            const uint32_t plane_bytes = static_cast<uint32_t>(ny) * NX * sizeof(R);
            const int valid = planes - p0 < TP ? planes - p0 : TP;  // planes of this tile that exist (>= 1)
            const __amdgpu_buffer_rsrc_t r =
                buffer_rsrc(base, static_cast<uint32_t>(valid - 1) * plane_bytes + NX * sizeof(R));
            const uint32_t oa = static_cast<uint32_t>(pa - p0) * plane_bytes + k1 * sizeof(R);
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2) {
                buffer_store_real<R>(w[k2].x, r, oa + N1 * k2 * sizeof(R));
                buffer_store_real<R>(w[k2].y, r, oa + plane_bytes + N1 * k2 * sizeof(R));
            }
        }
        if (!LOOP || next >= items) break;
        __syncthreads();  // buf is rewritten by the next item
        if constexpr (LOOP) {
#pragma unroll
            for (int m = 0; m < N1; ++m) v[m] = vn[m];
        }
        item = next;
    }
}

// ---------------------------------------------------------------------------
// Fused physical-space stage of the flux graph (csrc/flux.hip).
// T_in holds 3 + 3*targets fields of nz planes (plane = z + nz*field, plane
// stride ps_in); T_out holds `targets` fields.  A work item is row y of a
// tile of TP = 2*S consecutive z.  For each target t
//     out_t = -scale * (u*q_t,x + v*q_t,y + w*q_t,z)
// with every factor the c2r row transform of its field; the product is
// formed in registers and transformed forward straight into T_out, so the
// physical fields never reach HBM.
// ---------------------------------------------------------------------------

// c2r of one field's tile: on return thread (k1 = t % N1, s2 = t / N1) holds
// out[k2] = (plane 2*s2, plane 2*s2 + 1) at x = k1 + N1*k2, k2 < N2.
template <typename R, int NX>
__device__ __forceinline__ void flux_c2r(cx<R>* out, const cx<R>* tin, cx<R>* buf, const cx<R>* twl, size_t row_base,
                                         int ps, int kc, int pbase, int z0, int nz) {
    using G = RowGeo<R, NX>;
    constexpr int N1 = G::N1, N2 = G::N2, S = G::PAIRS;
    constexpr int YS = N2 + 1, SS = N1 * YS + 1;
    const int t = threadIdx.x, s1 = t % S, n2 = t / S;
    const int k1 = t % N1, s2 = t / N1;
    const int za = z0 + 2 * s1;
    cx<R> v[N1];
    if (((ps | pbase) & 1) == 0)
        c2r_gather_flat_t<R, NX, true>(v, tin, row_base, ps, kc, pbase + 2 * s1, za < nz, za + 1 < nz);
    else
        c2r_gather_flat_t<R, NX, false>(v, tin, row_base, ps, kc, pbase + 2 * s1, za < nz, za + 1 < nz);
    __syncthreads();  // buf and twl ready / free
    if (n2 < N2) {
        rdft<N1, true>(v);
#pragma unroll
        for (int k = 0; k < N1; ++k) buf[s1 * SS + k * YS + n2] = cmulc(v[k], twl[n2 * k]);
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < N2; ++m) out[m] = buf[s2 * SS + k1 * YS + m];
    rdft<N2, true>(out);
}

// Forward transform of the register rows left by flux_c2r (x = k1 + N1*k2):
// N = A*B with a = k1 (A = N1), b = k2 (B = N2):
//     X[kk + N2*k2'] = sum_a W_N1^(a k2') W_N^(a kk) sum_b x[a + N1 b] W_N2^(b kk)
// then the Hermitian split of each pair and the kx < kc store into T_out.
template <typename R, int NX>
__device__ __forceinline__ void flux_r2c_store(cx<R>* acc, cx<R>* tout, cx<R>* buf, const cx<R>* twl,
                                               size_t row_base, int ps, int kc, int pbase, int z0, int nz) {
    using G = RowGeo<R, NX>;
    constexpr int N1 = G::N1, N2 = G::N2, S = G::PAIRS, TP = 2 * S, NT = G::THREADS;
    constexpr int YS2 = N1 + 1, SS2 = N2 * YS2 + 1, ZS = NX + 1;
    const int t = threadIdx.x;
    const int k1 = t % N1, s2 = t / N1;
    const int kk = t % N2, s = t / N2;
    rdft<N2, false>(acc);
#pragma unroll
    for (int k = 0; k < N2; ++k) acc[k] = cmul(acc[k], twl[k1 * k]);
    __syncthreads();  // buf free
#pragma unroll
    for (int k = 0; k < N2; ++k) buf[s2 * SS2 + k * YS2 + k1] = acc[k];
    __syncthreads();
    cx<R> y[N1];
    if (s < S) {
#pragma unroll
        for (int m = 0; m < N1; ++m) y[m] = buf[s * SS2 + kk * YS2 + m];
        rdft<N1, false>(y);
    }
    __syncthreads();
    if (s < S) {
#pragma unroll
        for (int k = 0; k < N1; ++k) buf[s * ZS + kk + N2 * k] = y[k];
    }
    __syncthreads();
    for (int e = t; e < kc * TP; e += NT) {
        const int pit = e % TP, kx = e / TP;
        const int z = z0 + pit;
        if (z >= nz) continue;
        const cx<R> zk = buf[(pit >> 1) * ZS + kx];
        const cx<R> zn = buf[(pit >> 1) * ZS + ((NX - kx) & (NX - 1))];
        const cx<R> x = (pit & 1) == 0 ? cx<R>{R(0.5) * (zk.x + zn.x), R(0.5) * (zk.y - zn.y)}
                                       : cx<R>{R(0.5) * (zk.y + zn.y), R(0.5) * (zn.x - zk.x)};
        tout[row_base + static_cast<size_t>(ps) * kx + pbase + pit] = x;
    }
}

template <typename R, int NX>
struct FluxGeo {
    using G = RowGeo<R, NX>;
    static constexpr int S = G::PAIRS;
    static constexpr int SS = G::N1 * (G::N2 + 1) + 1, SS2 = G::N2 * (G::N1 + 1) + 1, ZS = NX + 1;
    static constexpr int M1 = SS > SS2 ? SS : SS2;
    static constexpr int BUF = S * (M1 > ZS ? M1 : ZS);
};

// One work item of the fused flux rows: row y of a tile of TP planes.
template <typename R, int NX>
__device__ __forceinline__ void flux_rows_item(const cx<R>* __restrict__ tin, cx<R>* __restrict__ tout, cx<R>* buf,
                                               const cx<R>* twl, int nz, int ps_in, int ps_out, int kc, int targets,
                                               R scale, int item) {
    using G = RowGeo<R, NX>;
    constexpr int N2 = G::N2, TP = 2 * G::PAIRS;
    const int ztiles = (nz + TP - 1) / TP;
    const int y = item / ztiles, z0 = (item % ztiles) * TP;
    const size_t rin = static_cast<size_t>(ps_in) * static_cast<size_t>(kc) * y;
    const size_t rout = static_cast<size_t>(ps_out) * static_cast<size_t>(kc) * y;
    cx<R> sh[3][N2];
#pragma unroll
    for (int d = 0; d < 3; ++d) flux_c2r<R, NX>(sh[d], tin, buf, twl, rin, ps_in, kc, d * nz + z0, z0, nz);
    // The final summation of the flux, u q_x + v q_y + w q_z, runs in FP64
    // for FP32 plans of 256 and 512 columns: each product of two FP32 values
    // is exact in FP64, so the target is the FP64 sum rounded once to FP32.
    // Those kernels hold one wavefront per SIMD either way; at 64 and 128
    // columns the FP64 accumulators would halve the wavefronts per SIMD and
    // at 1024 they would spill, so there the sum is FP32 fused multiply-adds.
    constexpr bool kWideSum = sizeof(R) == 4 && (NX == 256 || NX == 512);
    using A = typename std::conditional<kWideSum, double, R>::type;
    for (int tg = 0; tg < targets; ++tg) {
        A ax[N2], ay[N2];
#pragma unroll
        for (int k = 0; k < N2; ++k) ax[k] = ay[k] = A(0);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const cx<R>* sd = sh[d];
            cx<R> q[N2];
            flux_c2r<R, NX>(q, tin, buf, twl, rin, ps_in, kc, (3 + 3 * tg + d) * nz + z0, z0, nz);
#pragma unroll
            for (int k = 0; k < N2; ++k) {
                ax[k] += static_cast<A>(sd[k].x) * static_cast<A>(q[k].x);
                ay[k] += static_cast<A>(sd[k].y) * static_cast<A>(q[k].y);
            }
        }
        cx<R> acc[N2];
#pragma unroll
        for (int k = 0; k < N2; ++k)
            acc[k] = cx<R>{static_cast<R>(-static_cast<A>(scale) * ax[k]), static_cast<R>(-static_cast<A>(scale) * ay[k])};
        flux_r2c_store<R, NX>(acc, tout, buf, twl, rout, ps_out, kc, tg * nz + z0, z0, nz);
    }
}

// The c2r rows of u, v, w stay in registers for all targets.  FP32 only: in
// FP64 the register set does not fit, and the separate-kernel path is faster.
// LOOP: fewer workgroups than items, each taking items blockIdx.x,
// blockIdx.x + gridDim.x, ... (a capped launch that leaves CUs to other
// streams); otherwise one workgroup per item in the XCD-aware order.
template <typename R, int NX, bool LOOP>
__global__ __launch_bounds__((RowGeo<R, NX>::THREADS)) void flux_rows_kernel(const cx<R>* __restrict__ tin,
                                                                          cx<R>* __restrict__ tout,
                                                                          const cx<R>* __restrict__ tw, int nz,
                                                                          int ps_in, int ps_out, int kc, int targets,
                                                                          R scale, int items) {
    using G = RowGeo<R, NX>;
    constexpr int NT = G::THREADS;
    __shared__ cx<R> buf[FluxGeo<R, NX>::BUF];
    __shared__ cx<R> twl[NX];
    int item = first_item<LOOP>(items);
    if (item >= items) return;  // uniform across the workgroup
    for (int i = threadIdx.x; i < NX; i += NT) twl[i] = tw[i];
    for (;;) {
        flux_rows_item<R, NX>(tin, tout, buf, twl, nz, ps_in, ps_out, kc, targets, scale, item);
        if (!LOOP) break;
        item += static_cast<int>(gridDim.x);
        if (item >= items) break;
        __syncthreads();  // buf free for the next item
    }
}

// Column pass along y.  A work item is one column kx and a tile of TP
// consecutive planes; item = kx * tiles + tile, so concurrently running
// workgroups read whole contiguous plane runs of the same (kx, y) rows --
// plane-tile-major order would leave each DRAM page 8% used per visit.  Workgroups are persistent:
// each loops over items blockIdx.x, blockIdx.x + gridDim.x, ... and issues the
// global loads of its NEXT item before transforming the current one, so HBM
// traffic overlaps the register/LDS work instead of alternating with it.
// IN  kStrided: in[ps*(kx + in_kc*y) + p]   kRetained: in[P*mode + p] via map
// OUT kStrided: out[ps*(kx + out_kc*y) + p] kRetained: out[P*mode + p] via map
// (P = planes; ps >= P is the plane stride of the strided operand)
// Maps: code = 2*mode + conj at [kx*NY + row], or -1 (no mode).
// lcode: the item's column of map codes staged in LDS, or null to read the
// codes from global memory.
template <typename R, int NY, int IN>
__device__ __forceinline__ void col2_load(cx<R>* v, const cx<R>* in, int item, int kx_count, int planes, int ps,
                                          int in_kc, const int* __restrict__ map, const SplitViews<R>& views,
                                          uint32_t in_bytes, const int* lcode = nullptr) {
    using G = Geo2<R, NY>;
    constexpr int N1 = G::N1, N2 = G::N2, TP = G::TP;
    const int t = threadIdx.x, pp = t % TP, n2 = t / TP;
    const int ntiles = (planes + TP - 1) / TP;
    const int kx = item / ntiles, p = (item % ntiles) * TP + pp;
    const size_t P = static_cast<size_t>(planes);
    // Branch-free loads: a lane with nothing to read loads element 0 (in
    // bounds) and selects zero.
    if constexpr (IN == kStrided) {
        const bool act = n2 < N2 && p < planes;
        constexpr uint32_t CB = sizeof(cx<R>);
        const uint64_t extent = static_cast<uint64_t>(ps) * static_cast<uint64_t>(in_kc) * NY * CB;
        if (extent < kBufferPast) {
            // The strided operand fits one buffer resource: 32-bit offsets,
            // and a lane with nothing to read gets zero from an offset past
            // the end (no branch, no select).
            const __amdgpu_buffer_rsrc_t rs = buffer_rsrc(in, static_cast<uint32_t>(extent));
#pragma unroll
            for (int n1 = 0; n1 < N1; ++n1) {
                const int yy = N2 * n1 + n2;
                const uint32_t off =
                    act ? static_cast<uint32_t>(ps * (kx + in_kc * yy) + p) * CB : kBufferPast;
                v[n1] = buffer_load_cx<R>(rs, off);
            }
        } else {
#pragma unroll
            for (int n1 = 0; n1 < N1; ++n1) {
                const int yy = N2 * n1 + n2;
                const size_t off =
                    act ? static_cast<size_t>(ps) * (kx + static_cast<size_t>(in_kc) * yy) + p : 0;
                const cx<R> x = in[off];
                v[n1] = act ? x : cx<R>{R(0), R(0)};
            }
        }
    } else {
        // all codes first, then all values: two batched round trips
        int code[N1];
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1)
            code[n1] = n2 < N2 ? (lcode ? lcode[N2 * n1 + n2] : map[kx * NY + N2 * n1 + n2]) : -1;
        ViewLane<R> vl{};
        if constexpr (IN == kSplit) {
            if (p < planes) vl = view_lane(views, p);
        }
        if constexpr (IN == kSplit) {
#pragma unroll
            for (int n1 = 0; n1 < N1; ++n1) {
                cx<R> x{R(0), R(0)};
                if (code[n1] >= 0 && p < planes) {
                    const size_t at = vl.stride * static_cast<size_t>(code[n1] >> 1);
                    x = cx<R>{vl.re[at], vl.im[at]};
                    if (code[n1] & 1) x.y = -x.y;
                    x = apply_factor(views, vl.factor, kx, N2 * n1 + n2, x);
                }
                v[n1] = x;
            }
        } else if (in_bytes) {
            // Compact input within 32-bit buffer offsets: an absent mode reads
            // zero from an offset past the end (no branch, no select).  The
            // uniform choice of form sits outside the loop, so all N1 loads
            // are in flight before the first is used.
            const __amdgpu_buffer_rsrc_t rs = buffer_rsrc(in, in_bytes);
#pragma unroll
            for (int n1 = 0; n1 < N1; ++n1) {
                const bool ok = code[n1] >= 0 && p < planes;
                const uint32_t off =
                    ok ? static_cast<uint32_t>((P * static_cast<size_t>(code[n1] >> 1) + p) * sizeof(cx<R>))
                       : kBufferPast;
                v[n1] = buffer_load_cx<R>(rs, off);
            }
#pragma unroll
            for (int n1 = 0; n1 < N1; ++n1)
                if (code[n1] & 1) v[n1].y = -v[n1].y;
        } else {
#pragma unroll
            for (int n1 = 0; n1 < N1; ++n1) {
                const bool ok = code[n1] >= 0 && p < planes;
                const cx<R> y = in[ok ? P * static_cast<size_t>(code[n1] >> 1) + p : 0];
                v[n1] = ok ? cx<R>{y.x, (code[n1] & 1) ? -y.y : y.y} : cx<R>{R(0), R(0)};
            }
        }
    }
}

// Column-pass transpose between the two register passes, through LDS in two
// rounds (real, then imaginary parts) over one array of R: half the LDS of a
// complex buffer, so the column kernels are limited by registers rather than
// LDS.  Entry (k, m) of the N1 x N2 exchange holds TP planes at
// ((N2 + 1)*k + m)*TP: the extra row per block of N2 puts the rows that one
// wavefront reads (consecutive k1 at the same m) on different banks.
// Buffer size: N1*(N2 + 1)*TP elements.
template <typename R, int N1, int N2, int TP>
__device__ __forceinline__ void col_transpose(R* b, const cx<R>* z, bool put, int n2, int k1, int pp, cx<R>* w) {
    auto at = [&](int k, int m) { return ((N2 + 1) * k + m) * TP + pp; };
    if (put) {
#pragma unroll
        for (int k = 0; k < N1; ++k) b[at(k, n2)] = z[k].x;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < N2; ++m) w[m].x = b[at(k1, m)];
    __syncthreads();
    if (put) {
#pragma unroll
        for (int k = 0; k < N1; ++k) b[at(k, n2)] = z[k].y;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < N2; ++m) w[m].y = b[at(k1, m)];
}

template <typename R, int NY, bool INV, int IN, int OUT, bool LOOP>
__global__ __launch_bounds__((Geo2<R, NY>::THREADS)) void col2_kernel(const cx<R>* in, cx<R>* out,
                                                                     const cx<R>* __restrict__ tw, int planes, int ps,
                                                                     int kx_count, int items, int in_kc, int out_kc,
                                                                     const int* __restrict__ map,
                                                                     const SplitViews<R> views, uint32_t in_bytes) {
    using G = Geo2<R, NY>;
    constexpr int N1 = G::N1, N2 = G::N2, TP = G::TP, NT = G::THREADS;
    __shared__ R buf[N1 * (N2 + 1) * TP];
    __shared__ cx<R> twl[NY];
    // One workgroup per item: the item's column of map codes is staged in
    // LDS with one coalesced load per thread, instead of N1 (or N2) loads
    // per thread that 16 lanes share.
    constexpr bool kStage = !LOOP && (IN != kStrided || OUT != kStrided);
    __shared__ int scode[kStage ? NY : 1];
    const int t = threadIdx.x;
    const size_t P = static_cast<size_t>(planes);
    const int ntiles = (planes + TP - 1) / TP;
    int item = first_item<LOOP>(items);
    if (item >= items) return;  // uniform across the workgroup
    cx<R> v[N1];
    if constexpr (kStage) {
        const int kx0 = item / ntiles;
        for (int i = t; i < NY; i += NT) scode[i] = map[kx0 * NY + i];
        for (int i = t; i < NY; i += NT) twl[i] = tw[i];
        __syncthreads();
        col2_load<R, NY, IN>(v, in, item, kx_count, planes, ps, in_kc, map, views, in_bytes, scode);
    } else {
        col2_load<R, NY, IN>(v, in, item, kx_count, planes, ps, in_kc, map, views, in_bytes);
        for (int i = t; i < NY; i += NT) twl[i] = tw[i];
        __syncthreads();
    }
    const int pp = t % TP, n2 = t / TP, k1 = t / TP;
    for (;;) {
        const int next = LOOP ? item + static_cast<int>(gridDim.x) : items;
        cx<R> vn[LOOP ? N1 : 1];
        if constexpr (LOOP) {
            if (next < items)
                col2_load<R, NY, IN>(vn, in, next, kx_count, planes, ps, in_kc, map, views, in_bytes);
        }
        const int kx = item / ntiles, p = (item % ntiles) * TP + pp;
        // Output codes are loaded with (not after) the item's inputs: on gfx9
        // one in-order counter tracks loads and stores, so a code loaded
        // between two stores would make the second store wait for the first.
        int ocode[N2];
        if constexpr (OUT != kStrided) {
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2) ocode[k2] = kStage ? scode[k1 + N1 * k2] : map[kx * NY + k1 + N1 * k2];
        }
        ViewLane<R> ovl{};
        if constexpr (OUT == kSplit) {
            if (p < planes) ovl = view_lane(views, p);
        }
        // pass 1: thread (plane fastest, n2); pass 2: thread (plane fastest, k1)
        cx<R> w[N2];
        if (n2 < N2) {
            rdft<N1, INV>(v);
#pragma unroll
            for (int k = 0; k < N1; ++k) v[k] = INV ? cmulc(v[k], twl[n2 * k]) : cmul(v[k], twl[n2 * k]);
        }
        col_transpose<R, N1, N2, TP>(buf, v, n2 < N2, n2, k1, pp, w);
        rdft<N2, INV>(w);
        if (p < planes) {
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2) {
                const int row = k1 + N1 * k2;
                if constexpr (OUT == kStrided) {
                    out[static_cast<size_t>(ps) * (kx + static_cast<size_t>(out_kc) * row) + p] = w[k2];
                } else {
                    const int code = ocode[k2];
                    if (code >= 0) {
                        cx<R> x = w[k2];
                        if (code & 1) x.y = -x.y;
                        if constexpr (OUT == kSplit) {
                            const size_t at = ovl.stride * static_cast<size_t>(code >> 1);
                            ovl.re[at] = views.scale * x.x;
                            ovl.im[at] = views.scale * x.y;
                        } else {
                            out[P * (code >> 1) + p] = x;
                        }
                    }
                }
            }
        }
        if (!LOOP || next >= items) break;
        __syncthreads();  // buf is rewritten by the next item
        if constexpr (LOOP) {
#pragma unroll
            for (int m = 0; m < N1; ++m) v[m] = vn[m];
        }
        item = next;
    }
}

// ---------------------------------------------------------------------------
// Retained gather / embed on a full WVM spectrum (P_h, E_h).
// ---------------------------------------------------------------------------
template <typename R>
__global__ void gather_kernel(const cx<R>* __restrict__ wvm, cx<R>* __restrict__ retained,
                              const ColEntry* __restrict__ modes /* row=ky, mode=kx */, int planes, int nxh,
                              size_t total) {
    for (size_t e = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; e < total;
         e += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const size_t m = e / planes, p = e - m * planes;
        const ColEntry md = modes[m];
        cx<R> v = wvm[static_cast<size_t>(planes) * (md.mode + static_cast<size_t>(nxh) * md.row) + p];
        if (md.conj) v.y = -v.y;
        retained[e] = v;
    }
}

template <typename R>
__global__ void zero_kernel(cx<R>* __restrict__ out, size_t total) {
    for (size_t e = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; e < total;
         e += static_cast<size_t>(gridDim.x) * blockDim.x)
        out[e] = cx<R>{R(0), R(0)};
}

// entries: (row=ky, mode=compact index, conj, pad=kx) incl. k_x = 0 partners
template <typename R>
__global__ void embed_kernel(const cx<R>* __restrict__ retained, cx<R>* __restrict__ wvm,
                             const ColEntry* __restrict__ entries, int planes, int nxh, size_t total) {
    for (size_t e = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; e < total;
         e += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const size_t i = e / planes, p = e - i * planes;
        const ColEntry en = entries[i];
        cx<R> v = retained[static_cast<size_t>(planes) * en.mode + p];
        if (en.conj) v.y = -v.y;
        wvm[static_cast<size_t>(planes) * (en.pad + static_cast<size_t>(nxh) * en.row) + p] = v;
    }
}

// ---------------------------------------------------------------------------
// Host-side dispatch
// ---------------------------------------------------------------------------
namespace {

template <typename R>
std::vector<cx<R>> twiddles(int n) {
    std::vector<cx<R>> t(n);
    for (int i = 0; i < n; ++i) {
        const double a = 2.0 * M_PI * static_cast<double>(i) / static_cast<double>(n);
        t[i] = cx<R>{static_cast<R>(std::cos(a)), static_cast<R>(-std::sin(a))};
    }
    return t;
}

void check_launch(const char* what) { check_hip(hipGetLastError(), what); }


// Workgroups of a launch: `requested` (> 0) or one per work item, capped at
// the item count.  Fewer workgroups than items makes them persistent.
int launch_grid(int items, int requested) {
    return requested > 0 && requested < items ? requested : xcd_grid(items);
}

template <typename R, int NX, bool LOOP>
void launch_row2(bool inverse, const void* in, void* out, const void* tw, int ny, int planes, int ps, int kc,
                 int items, int grid, hipStream_t stream) {
    using G = RowGeo<R, NX>;
    if (inverse)
        hipLaunchKernelGGL((row2_c2r_kernel<R, NX, LOOP>), dim3(grid), dim3(G::THREADS), 0, stream,
                           static_cast<const cx<R>*>(in), static_cast<R*>(out), static_cast<const cx<R>*>(tw), ny,
                           planes, ps, kc, items);
    else
        hipLaunchKernelGGL((row2_r2c_kernel<R, NX, LOOP>), dim3(grid), dim3(G::THREADS), 0, stream,
                           static_cast<const R*>(in), static_cast<cx<R>*>(out), static_cast<const cx<R>*>(tw), ny,
                           planes, ps, kc, items);
}

// ps: plane stride of the spectrum-side operand (>= planes); the Stockham
// kernels for Nx < 64 take ps == planes only.
template <typename R, int NX>
void launch_row(bool inverse, const void* in, void* out, const void* tw, int ny, int planes, int ps, int kc, int req,
                hipStream_t stream) {
    if constexpr (NX >= 64) {
        using G = RowGeo<R, NX>;
        constexpr int TP = 2 * G::PAIRS;
        const int items = ny * ((planes + TP - 1) / TP);
        const int grid = launch_grid(items, req);
        if (grid < items) launch_row2<R, NX, true>(inverse, in, out, tw, ny, planes, ps, kc, items, grid, stream);
        else launch_row2<R, NX, false>(inverse, in, out, tw, ny, planes, ps, kc, items, grid, stream);
    } else {
        if (ps != planes) fail(SKB_RUNTIME_ERROR, "padded plane stride needs Nx >= 64");
        constexpr int TP = row_tile<R>();
        const dim3 grid(ny, (planes + TP - 1) / TP);
        if (inverse)
            hipLaunchKernelGGL((row_c2r_kernel<R, NX>), grid, dim3(kThreads), 0, stream,
                               static_cast<const cx<R>*>(in), static_cast<R*>(out), static_cast<const cx<R>*>(tw), ny,
                               planes, kc);
        else
            hipLaunchKernelGGL((row_r2c_kernel<R, NX>), grid, dim3(kThreads), 0, stream, static_cast<const R*>(in),
                               static_cast<cx<R>*>(out), static_cast<const cx<R>*>(tw), ny, planes, kc);
    }
    check_launch(inverse ? "row c2r launch" : "row r2c launch");
}

enum class ColVariant { FwdStrided, FwdToRetained, FwdToSplit, InvStrided, InvFromRetained, InvFromSplit };

// Stage indices of the pruned plan (skb_pruned_plan_set_grid).
constexpr int kStageForwardRows = 0, kStageForwardColumns = 1, kStageInverseColumns = 2, kStageInverseRows = 3;

template <typename R, int NY, bool LOOP>
void launch_col2(ColVariant v, const cx<R>* i, cx<R>* o, const cx<R>* t, int planes, int ps, int kx_count, int items,
                 int in_kc, int out_kc, const int* map, const SplitViews<R>& views, uint32_t in_bytes, int grid_x,
                 hipStream_t stream) {
    using G = Geo2<R, NY>;
    const dim3 grid(grid_x), block(G::THREADS);
#define SKB_COL2(INV, IN, OUT)                                                                                    \
    hipLaunchKernelGGL((col2_kernel<R, NY, INV, IN, OUT, LOOP>), grid, block, 0, stream, i, o, t, planes, ps,     \
                       kx_count, items, in_kc, out_kc, map, views, in_bytes)
    switch (v) {
        case ColVariant::FwdStrided: SKB_COL2(false, kStrided, kStrided); break;
        case ColVariant::FwdToRetained: SKB_COL2(false, kStrided, kRetained); break;
        case ColVariant::FwdToSplit: SKB_COL2(false, kStrided, kSplit); break;
        case ColVariant::InvStrided: SKB_COL2(true, kStrided, kStrided); break;
        case ColVariant::InvFromRetained: SKB_COL2(true, kRetained, kStrided); break;
        case ColVariant::InvFromSplit: SKB_COL2(true, kSplit, kStrided); break;
    }
#undef SKB_COL2
}

template <typename R, int NY>
void launch_col(ColVariant v, const void* in, void* out, const void* tw, int planes, int ps, int kx_count, int in_kc,
                int out_kc, const int* col_start, const ColEntry* entries, const int* map,
                const SplitViews<R>& views, uint64_t in_bytes, int req, hipStream_t stream) {
    // buffer addressing of a compact input needs every offset below 2 GiB
    const uint32_t ib = in_bytes < kBufferPast ? static_cast<uint32_t>(in_bytes) : 0u;
    const auto* i = static_cast<const cx<R>*>(in);
    auto* o = static_cast<cx<R>*>(out);
    const auto* t = static_cast<const cx<R>*>(tw);
    if constexpr (NY >= 64) {
        using G = Geo2<R, NY>;
        const int items = kx_count * ((planes + G::TP - 1) / G::TP);
        const int grid = launch_grid(items, req);
        if (grid < items)
            launch_col2<R, NY, true>(v, i, o, t, planes, ps, kx_count, items, in_kc, out_kc, map, views, ib, grid,
                                     stream);
        else
            launch_col2<R, NY, false>(v, i, o, t, planes, ps, kx_count, items, in_kc, out_kc, map, views, ib, grid,
                                      stream);
    } else {
        if (ps != planes) fail(SKB_RUNTIME_ERROR, "padded plane stride needs Ny >= 64");
        constexpr int TP = col_tile<R, NY>();
        const dim3 grid(kx_count, (planes + TP - 1) / TP);
#define SKB_COL(INV, IN, OUT)                                                                                     \
    hipLaunchKernelGGL((col_kernel<R, NY, INV, IN, OUT>), grid, dim3(kThreads), 0, stream, i, o, t, planes, in_kc, \
                       out_kc, col_start, entries, views)
        switch (v) {
            case ColVariant::FwdStrided: SKB_COL(false, kStrided, kStrided); break;
            case ColVariant::FwdToRetained: SKB_COL(false, kStrided, kRetained); break;
            case ColVariant::FwdToSplit: SKB_COL(false, kStrided, kSplit); break;
            case ColVariant::InvStrided: SKB_COL(true, kStrided, kStrided); break;
            case ColVariant::InvFromRetained: SKB_COL(true, kRetained, kStrided); break;
            case ColVariant::InvFromSplit: SKB_COL(true, kSplit, kStrided); break;
        }
#undef SKB_COL
    }
    check_launch("column launch");
}

#define SKB_POW2_CASES(MACRO) \
    MACRO(4) MACRO(8) MACRO(16) MACRO(32) MACRO(64) MACRO(128) MACRO(256) MACRO(512) MACRO(1024)

template <typename R>
void row_pass(bool inverse, int nx, const void* in, void* out, const void* tw, int ny, int planes, int ps, int kc,
              int req, hipStream_t s) {
    switch (nx) {
#define SKB_ROW_CASE(N)                                                         \
    case N:                                                                     \
        launch_row<R, N>(inverse, in, out, tw, ny, planes, ps, kc, req, s);     \
        return;
        SKB_POW2_CASES(SKB_ROW_CASE)
#undef SKB_ROW_CASE
    }
    fail(SKB_INVALID_ARGUMENT, "unsupported Nx");
}

template <typename R>
void col_pass(ColVariant v, int ny, const void* in, void* out, const void* tw, int planes, int ps, int kx_count,
              int in_kc, int out_kc, const int* col_start, const ColEntry* entries, const int* map, int req,
              hipStream_t s, const SplitViews<R>& views = SplitViews<R>{}, uint64_t in_bytes = 0) {
    switch (ny) {
#define SKB_COL_CASE(N)                                                                                        \
    case N:                                                                                                    \
        launch_col<R, N>(v, in, out, tw, planes, ps, kx_count, in_kc, out_kc, col_start, entries, map, views,  \
                         in_bytes, req, s);                                                                    \
        return;
        SKB_POW2_CASES(SKB_COL_CASE)
#undef SKB_COL_CASE
    }
    fail(SKB_INVALID_ARGUMENT, "unsupported Ny");
}

void check_fft_sizes(const skb_workload& w) {
    validate_workload(&w);
    if (!is_pow2(w.nx) || !is_pow2(w.ny) || w.nx < kMinFft || w.ny < kMinFft || w.nx > kMaxFft || w.ny > kMaxFft)
        fail(SKB_UNSUPPORTED, "the hip provider supports power-of-two Nx, Ny in [4, 1024]");
    const uint64_t planes = w.nz * w.fields;
    // grid.y = ceil(planes / tile) with tiles of >= 4 planes must stay below 65536
    if (planes > (1u << 17)) fail(SKB_UNSUPPORTED, "plane count above 2^17");
    if (planes * (w.nx / 2 + 1) * w.ny > (1ull << 31) || planes * w.nx * w.ny > (1ull << 31))
        fail(SKB_UNSUPPORTED, "workload exceeds 2^31 elements per array");
}

// Twiddle tables for Nx and Ny in both precisions' layout.
struct Twiddles {
    DeviceBuffer x, y;
    void build(int precision, uint64_t nx, uint64_t ny) {
        if (precision == SKB_F64) {
            auto tx = twiddles<double>(static_cast<int>(nx));
            auto ty = twiddles<double>(static_cast<int>(ny));
            x.upload(tx.data(), tx.size() * sizeof(tx[0]));
            y.upload(ty.data(), ty.size() * sizeof(ty[0]));
        } else {
            auto tx = twiddles<float>(static_cast<int>(nx));
            auto ty = twiddles<float>(static_cast<int>(ny));
            x.upload(tx.data(), tx.size() * sizeof(tx[0]));
            y.upload(ty.data(), ty.size() * sizeof(ty[0]));
        }
    }
};

// Column lists of retained modes.  forward: one entry per mode at its stored
// (kx, ky).  inverse: additionally the k_x = 0 conjugate partners
// (src/model.cpp:332-335).
struct ColumnLists {
    uint64_t kc = 0;  // active k_x band: max stored_kx + 1
    std::vector<int> fwd_start, inv_start;
    std::vector<ColEntry> fwd, inv;
    // dense maps [kx * Ny + ky] -> 2*mode + conj, or -1
    std::vector<int> fmap, imap;
};

ColumnLists build_column_lists(const skb_workload& w, const skb_retained_mode* modes, uint64_t n) {
    if (!modes || n == 0) fail(SKB_INVALID_ARGUMENT, "The pruned provider requires retained modes.");
    if (n > (1u << 29)) fail(SKB_UNSUPPORTED, "too many retained modes");
    const uint64_t nxh = w.nx / 2 + 1;
    ColumnLists c;
    for (uint64_t m = 0; m < n; ++m) {
        if (modes[m].stored_kx >= nxh || modes[m].stored_ky >= w.ny)
            fail(SKB_INVALID_ARGUMENT, "A retained mode lies outside the half-spectrum.");
        c.kc = std::max<uint64_t>(c.kc, modes[m].stored_kx + 1);
    }
    std::vector<std::vector<ColEntry>> f(c.kc), i(c.kc);
    std::vector<uint8_t> used(c.kc * w.ny, 0);
    auto claim = [&](uint64_t kx, uint64_t row) {
        auto& u = used[kx * w.ny + row];
        if (u) fail(SKB_INVALID_ARGUMENT, "two retained modes (or a mode and a Hermitian partner) share a stored slot");
        u = 1;
    };
    for (uint64_t m = 0; m < n; ++m) {
        const auto& md = modes[m];
        const ColEntry e{static_cast<int>(md.stored_ky), static_cast<int>(m), md.conjugates_stored_value ? 1 : 0, 0};
        claim(md.stored_kx, md.stored_ky);
        f[md.stored_kx].push_back(e);
        i[md.stored_kx].push_back(e);
    }
    for (uint64_t m = 0; m < n; ++m) {
        const auto& md = modes[m];
        if (md.stored_kx == 0 && md.stored_ky != 0 && 2 * md.stored_ky != w.ny) {
            const uint64_t row = (w.ny - md.stored_ky) % w.ny;
            claim(0, row);
            i[0].push_back(ColEntry{static_cast<int>(row), static_cast<int>(m), md.conjugates_stored_value ? 0 : 1, 0});
        }
    }
    c.fmap.assign(c.kc * w.ny, -1);
    c.imap.assign(c.kc * w.ny, -1);
    for (uint64_t kx = 0; kx < c.kc; ++kx) {
        for (const auto& e : f[kx]) c.fmap[kx * w.ny + e.row] = 2 * e.mode + e.conj;
        for (const auto& e : i[kx]) c.imap[kx * w.ny + e.row] = 2 * e.mode + e.conj;
    }
    c.fwd_start.push_back(0);
    c.inv_start.push_back(0);
    for (uint64_t kx = 0; kx < c.kc; ++kx) {
        c.fwd.insert(c.fwd.end(), f[kx].begin(), f[kx].end());
        c.inv.insert(c.inv.end(), i[kx].begin(), i[kx].end());
        c.fwd_start.push_back(static_cast<int>(c.fwd.size()));
        c.inv_start.push_back(static_cast<int>(c.inv.size()));
    }
    return c;
}

}  // namespace

// ---------------------------------------------------------------------------
// Plans
// ---------------------------------------------------------------------------
}  // namespace skb

struct skb_fft_plan {
    skb_workload w{};
    int precision = SKB_F64;
    skb::Twiddles tw;
    // Contiguous spectrum that strided field views go through when the grid
    // is below the two-pass kernels' 64 x 64 (allocated at creation then).
    skb::DeviceBuffer view_stage;
    // Device copy of the mode table last handed to gather/embed.
    std::vector<skb_retained_mode> cached_modes;
    skb::DeviceBuffer gather_table, embed_table;
    uint64_t embed_entries = 0;
};

struct skb_pruned_plan {
    skb_workload w{};
    int grid[4] = {0, 0, 0, 0};  // per-stage workgroup count (0 = one per item)
    int precision = SKB_F64;
    uint64_t modes = 0;
    int planes = 0;  // P = Nz * fields
    int kc = 0;      // active k_x band
    // Plane stride of T: P rounded up to whole 128-byte lines (16 FP32 or 8
    // FP64 complex values) when both passes use the two-pass kernels, so every
    // plane tile of the column and row kernels starts on a cache line.
    int ps = 0;
    skb::Twiddles tw;
    skb::ColumnLists lists;
    skb::DeviceBuffer fwd_start, inv_start, fwd, inv, fmap, imap;
    skb::DeviceBuffer scratch;  // T: kc * Ny * ps complex, element (kx, y, p) at ps * (kx + kc * y) + p
    // Plane-resident path (FP32 256 x 256, k_x band 86): tables and the
    // chunked plane-major copy of the compact operand (16*P*(mode/16) +
    // 16*plane + mode%16).
    int path = SKB_PATH_AUTO;
    bool plane_built = false;
    skb::PlaneTables plane;
    skb::DeviceBuffer pm;
};

using namespace skb;

namespace {

// The gather / embed tables follow the mode list of the call; a new list is
// uploaded (allocating) on its first use, which must not happen while the
// stream is being captured into a graph.
void refresh_mode_tables(skb_fft_plan* plan, const skb_retained_mode* modes, uint64_t n, hipStream_t stream) {
    if (!modes || n == 0) fail(SKB_INVALID_ARGUMENT, "retained modes required");
    if (plan->cached_modes.size() == n &&
        std::memcmp(plan->cached_modes.data(), modes, n * sizeof(skb_retained_mode)) == 0)
        return;
    hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
    check_hip(hipStreamIsCapturing(stream, &capture), "hipStreamIsCapturing");
    if (capture != hipStreamCaptureStatusNone)
        fail(SKB_LOGIC_ERROR, "a new retained-mode list needs its tables built before graph capture: "
                              "call once with it outside the capture");
    const auto& w = plan->w;
    const uint64_t nxh = w.nx / 2 + 1;
    std::vector<ColEntry> g(n), e;
    e.reserve(n + n / 4);
    for (uint64_t m = 0; m < n; ++m) {
        const auto& md = modes[m];
        if (md.stored_kx >= nxh || md.stored_ky >= w.ny) fail(SKB_INVALID_ARGUMENT, "A retained mode lies outside the half-spectrum.");
        const int cj = md.conjugates_stored_value ? 1 : 0;
        g[m] = ColEntry{static_cast<int>(md.stored_ky), static_cast<int>(md.stored_kx), cj, 0};
        e.push_back(ColEntry{static_cast<int>(md.stored_ky), static_cast<int>(m), cj, static_cast<int>(md.stored_kx)});
    }
    for (uint64_t m = 0; m < n; ++m) {
        const auto& md = modes[m];
        if (md.stored_kx == 0 && md.stored_ky != 0 && 2 * md.stored_ky != w.ny)
            e.push_back(ColEntry{static_cast<int>((w.ny - md.stored_ky) % w.ny), static_cast<int>(m),
                                 md.conjugates_stored_value ? 0 : 1, 0});
    }
    plan->gather_table.upload(g.data(), g.size() * sizeof(ColEntry));
    plan->embed_table.upload(e.data(), e.size() * sizeof(ColEntry));
    plan->embed_entries = e.size();
    plan->cached_modes.assign(modes, modes + n);
}

// ps: frequency stride of the spectrum (the plane count, or a WVM field
// view's stride; ps > planes needs Nx, Ny >= 64).
template <typename R>
void fft_forward(skb_fft_plan* p, const void* real, void* wvm, hipStream_t s, int ps = 0) {
    const auto& w = p->w;
    const int planes = static_cast<int>(w.nz * w.fields);
    const int nxh = static_cast<int>(w.nx / 2 + 1);
    if (ps == 0) ps = planes;
    row_pass<R>(false, static_cast<int>(w.nx), real, wvm, p->tw.x.ptr, static_cast<int>(w.ny), planes, ps, nxh, 0, s);
    col_pass<R>(ColVariant::FwdStrided, static_cast<int>(w.ny), wvm, wvm, p->tw.y.ptr, planes, ps, nxh, nxh, nxh,
                nullptr, nullptr, nullptr, 0, s);
}

template <typename R>
void fft_inverse(skb_fft_plan* p, void* wvm, void* real, hipStream_t s, int ps = 0) {
    const auto& w = p->w;
    const int planes = static_cast<int>(w.nz * w.fields);
    const int nxh = static_cast<int>(w.nx / 2 + 1);
    if (ps == 0) ps = planes;
    col_pass<R>(ColVariant::InvStrided, static_cast<int>(w.ny), wvm, wvm, p->tw.y.ptr, planes, ps, nxh, nxh, nxh,
                nullptr, nullptr, nullptr, 0, s);
    row_pass<R>(true, static_cast<int>(w.nx), wvm, real, p->tw.x.ptr, static_cast<int>(w.ny), planes, ps, nxh, 0, s);
}

// A WVM field view of stride `stride`: native for Nx, Ny >= 64, otherwise
// through the plan's contiguous stage.
void check_view(const skb_fft_plan* p, uint64_t stride) {
    const uint64_t planes = p->w.nz * p->w.fields, freqs = (p->w.nx / 2 + 1) * p->w.ny;
    if (stride < planes) fail(SKB_INVALID_ARGUMENT, "WVM field-view frequency strides must contain every vertical plane.");
    if (stride * freqs > (1ull << 31)) fail(SKB_UNSUPPORTED, "WVM field view exceeds 2^31 elements");
}

template <typename R>
void fft_forward_view(skb_fft_plan* p, const void* real, void* view, uint64_t stride, hipStream_t s) {
    check_view(p, stride);
    const int planes = static_cast<int>(p->w.nz * p->w.fields);
    if (!p->view_stage.ptr) return fft_forward<R>(p, real, view, s, static_cast<int>(stride));
    fft_forward<R>(p, real, p->view_stage.ptr, s);
    copy_planes(p->precision, (p->w.nx / 2 + 1) * p->w.ny, planes, p->view_stage.ptr, planes, view, stride, s);
}

template <typename R>
void fft_inverse_view(skb_fft_plan* p, void* view, uint64_t stride, void* real, hipStream_t s) {
    check_view(p, stride);
    const int planes = static_cast<int>(p->w.nz * p->w.fields);
    if (!p->view_stage.ptr) return fft_inverse<R>(p, view, real, s, static_cast<int>(stride));
    copy_planes(p->precision, (p->w.nx / 2 + 1) * p->w.ny, planes, view, stride, p->view_stage.ptr, planes, s);
    fft_inverse<R>(p, p->view_stage.ptr, real, s);
}

// The four stages of the pruned transforms; T is the plan's kc x Ny x ps
// intermediate (plane stride ps, see skb_pruned_plan).
template <typename R>
void stage_forward_rows(skb_pruned_plan* p, const void* real, hipStream_t s) {
    const auto& w = p->w;
    row_pass<R>(false, static_cast<int>(w.nx), real, p->scratch.ptr, p->tw.x.ptr, static_cast<int>(w.ny), p->planes,
                p->ps, p->kc, p->grid[kStageForwardRows], s);
}

template <typename R>
void stage_forward_columns(skb_pruned_plan* p, void* retained, hipStream_t s) {
    // One workgroup per item: its retained-only stores beat the streaming
    // kernel's fixed store count here (about half the rows hold no mode).
    col_pass<R>(ColVariant::FwdToRetained, static_cast<int>(p->w.ny), p->scratch.ptr, retained, p->tw.y.ptr, p->planes,
                p->ps, p->kc, p->kc, p->kc, p->fwd_start.as<int>(), p->fwd.as<ColEntry>(), p->fmap.as<int>(),
                p->grid[kStageForwardColumns], s);
}

template <typename R>
void stage_inverse_columns(skb_pruned_plan* p, const void* retained, hipStream_t s) {
    col_pass<R>(ColVariant::InvFromRetained, static_cast<int>(p->w.ny), retained, p->scratch.ptr, p->tw.y.ptr,
                p->planes, p->ps, p->kc, p->kc, p->kc, p->inv_start.as<int>(), p->inv.as<ColEntry>(),
                p->imap.as<int>(), p->grid[kStageInverseColumns], s, SplitViews<R>{},
                p->modes * p->planes * sizeof(cx<R>));
}

template <typename R>
void stage_inverse_rows(skb_pruned_plan* p, void* real, hipStream_t s) {
    const auto& w = p->w;
    row_pass<R>(true, static_cast<int>(w.nx), p->scratch.ptr, real, p->tw.x.ptr, static_cast<int>(w.ny), p->planes,
                p->ps, p->kc, p->grid[kStageInverseRows], s);
}

// Forward columns into / inverse columns from a split compact operand.
template <typename R>
void stage_forward_columns_split(skb_pruned_plan* p, const SplitViews<R>& views, hipStream_t s) {
    col_pass<R>(ColVariant::FwdToSplit, static_cast<int>(p->w.ny), p->scratch.ptr, nullptr, p->tw.y.ptr, p->planes,
                p->ps, p->kc, p->kc, p->kc, p->fwd_start.as<int>(), p->fwd.as<ColEntry>(), p->fmap.as<int>(),
                p->grid[kStageForwardColumns], s, views);
}

template <typename R>
void stage_inverse_columns_split(skb_pruned_plan* p, const SplitViews<R>& views, hipStream_t s) {
    col_pass<R>(ColVariant::InvFromSplit, static_cast<int>(p->w.ny), nullptr, p->scratch.ptr, p->tw.y.ptr, p->planes,
                p->ps, p->kc, p->kc, p->kc, p->inv_start.as<int>(), p->inv.as<ColEntry>(), p->imap.as<int>(),
                p->grid[kStageInverseColumns], s, views);
}

// The plain split layout (plane + P * mode) as one view.
template <typename R>
SplitViews<R> plain_split(const void* re, const void* im, int planes, double scale) {
    if (!re || !im) fail(SKB_INVALID_ARGUMENT, "null split array");
    SplitViews<R> v{};
    v.re[0] = static_cast<R*>(const_cast<void*>(re));
    v.im[0] = static_cast<R*>(const_cast<void*>(im));
    v.stride[0] = planes;
    v.nz = planes;
    v.zs = 1;
    v.scale = static_cast<R>(scale);
    return v;
}

// One view per field; validation follows src/pruned_fftw_provider.cpp:929-958.
template <typename R>
SplitViews<R> field_split(const skb_pruned_plan* p, const skb_field_view* fields, uint64_t count, double scale,
                          bool output) {
    if (!fields || count != p->w.fields)
        fail(SKB_INVALID_ARGUMENT, output ? "Streaming pruned output field views must match the workload field count."
                                          : "Streaming pruned input field views must match the workload field count.");
    if (count > static_cast<uint64_t>(kMaxSplitViews))
        fail(SKB_UNSUPPORTED, "split field views support at most 16 fields");
    SplitViews<R> v{};
    for (uint64_t f = 0; f < count; ++f) {
        if (!fields[f].real || !fields[f].imaginary || fields[f].mode_stride < p->w.nz)
            fail(SKB_INVALID_ARGUMENT, output ? "Streaming pruned output field view is invalid."
                                              : "Streaming pruned input field view is invalid.");
        if (fields[f].mode_stride > (1ull << 40)) fail(SKB_UNSUPPORTED, "field view mode stride too large");
        v.re[f] = static_cast<R*>(fields[f].real);
        v.im[f] = static_cast<R*>(fields[f].imaginary);
        v.stride[f] = static_cast<long long>(fields[f].mode_stride);
    }
    v.nz = static_cast<int>(p->w.nz);
    v.zs = 1;
    v.scale = static_cast<R>(scale);
    return v;
}

// Runs stage F<float> or F<double> by the plan's precision.
#define SKB_BY_PRECISION(plan, F, ...)                \
    do {                                              \
        if ((plan)->precision == SKB_F64)             \
            F<double>(__VA_ARGS__);                   \
        else                                          \
            F<float>(__VA_ARGS__);                    \
    } while (0)

void require_ptrs(const void* a, const void* b) {
    if (!a || !b) fail(SKB_INVALID_ARGUMENT, "null device pointer");
}

bool uses_plane(const skb_pruned_plan* p) { return p->plane_built && p->path != SKB_PATH_TWO_PASS; }

}  // namespace

extern "C" {

int skb_fft_plan_create(const skb_workload* workload, int precision, skb_fft_plan** plan) {
    return guarded([&] {
        if (!plan) fail(SKB_INVALID_ARGUMENT, "null plan pointer");
        *plan = nullptr;
        check_precision(precision);
        validate_workload(workload);
        check_fft_sizes(*workload);
        auto p = std::make_unique<skb_fft_plan>();
        p->w = *workload;
        p->precision = precision;
        p->tw.build(precision, workload->nx, workload->ny);
        if (workload->nx < 64 || workload->ny < 64)
            p->view_stage.allocate((workload->nx / 2 + 1) * workload->ny * workload->nz * workload->fields *
                                   complex_bytes(precision));
        *plan = p.release();
    });
}

int skb_fft_plan_destroy(skb_fft_plan* plan) {
    return guarded([&] { delete plan; });
}

int skb_fft_forward(skb_fft_plan* plan, const void* real, void* wvm, void* stream) {
    return guarded([&] {
        if (!plan) fail(SKB_INVALID_ARGUMENT, "null plan");
        require_ptrs(real, wvm);
        if (plan->precision == SKB_F64) fft_forward<double>(plan, real, wvm, as_stream(stream));
        else fft_forward<float>(plan, real, wvm, as_stream(stream));
    });
}

int skb_fft_inverse(skb_fft_plan* plan, void* wvm, void* real, void* stream) {
    return guarded([&] {
        if (!plan) fail(SKB_INVALID_ARGUMENT, "null plan");
        require_ptrs(real, wvm);
        if (plan->precision == SKB_F64) fft_inverse<double>(plan, wvm, real, as_stream(stream));
        else fft_inverse<float>(plan, wvm, real, as_stream(stream));
    });
}

int skb_fft_forward_view(skb_fft_plan* plan, const void* real, void* spectrum, uint64_t frequency_stride,
                         void* stream) {
    return guarded([&] {
        if (!plan) fail(SKB_INVALID_ARGUMENT, "null plan");
        require_ptrs(real, spectrum);
        if (plan->precision == SKB_F64)
            fft_forward_view<double>(plan, real, spectrum, frequency_stride, as_stream(stream));
        else fft_forward_view<float>(plan, real, spectrum, frequency_stride, as_stream(stream));
    });
}

int skb_fft_inverse_view(skb_fft_plan* plan, void* spectrum, uint64_t frequency_stride, void* real, void* stream) {
    return guarded([&] {
        if (!plan) fail(SKB_INVALID_ARGUMENT, "null plan");
        require_ptrs(real, spectrum);
        if (plan->precision == SKB_F64)
            fft_inverse_view<double>(plan, spectrum, frequency_stride, real, as_stream(stream));
        else fft_inverse_view<float>(plan, spectrum, frequency_stride, real, as_stream(stream));
    });
}

int skb_fft_gather_retained(skb_fft_plan* plan, const skb_retained_mode* modes, uint64_t mode_count, const void* wvm,
                            void* retained, void* stream) {
    return guarded([&] {
        if (!plan) fail(SKB_INVALID_ARGUMENT, "null plan");
        require_ptrs(wvm, retained);
        refresh_mode_tables(plan, modes, mode_count, as_stream(stream));
        const int planes = static_cast<int>(plan->w.nz * plan->w.fields);
        const int nxh = static_cast<int>(plan->w.nx / 2 + 1);
        const size_t total = mode_count * static_cast<size_t>(planes);
        if (plan->precision == SKB_F64)
            hipLaunchKernelGGL(gather_kernel<double>, stream_grid(total), dim3(256), 0, as_stream(stream),
                               static_cast<const cx<double>*>(wvm), static_cast<cx<double>*>(retained),
                               plan->gather_table.as<ColEntry>(), planes, nxh, total);
        else
            hipLaunchKernelGGL(gather_kernel<float>, stream_grid(total), dim3(256), 0, as_stream(stream),
                               static_cast<const cx<float>*>(wvm), static_cast<cx<float>*>(retained),
                               plan->gather_table.as<ColEntry>(), planes, nxh, total);
        check_launch("gather launch");
    });
}

int skb_fft_gather_retained_modal(skb_fft_plan* plan, const skb_retained_mode* modes, uint64_t mode_count,
                                  uint64_t nj, const void* full_modal, void* modal, void* stream) {
    return guarded([&] {
        if (!plan || nj == 0) fail(SKB_INVALID_ARGUMENT, "null plan or zero vertical modes");
        require_ptrs(full_modal, modal);
        refresh_mode_tables(plan, modes, mode_count, as_stream(stream));
        const int block = static_cast<int>(nj * plan->w.fields);
        const int nxh = static_cast<int>(plan->w.nx / 2 + 1);
        const size_t total = mode_count * static_cast<size_t>(block);
        if (plan->precision == SKB_F64)
            hipLaunchKernelGGL(gather_kernel<double>, stream_grid(total), dim3(256), 0, as_stream(stream),
                               static_cast<const cx<double>*>(full_modal), static_cast<cx<double>*>(modal),
                               plan->gather_table.as<ColEntry>(), block, nxh, total);
        else
            hipLaunchKernelGGL(gather_kernel<float>, stream_grid(total), dim3(256), 0, as_stream(stream),
                               static_cast<const cx<float>*>(full_modal), static_cast<cx<float>*>(modal),
                               plan->gather_table.as<ColEntry>(), block, nxh, total);
        check_launch("modal gather launch");
    });
}

int skb_fft_embed_retained(skb_fft_plan* plan, const skb_retained_mode* modes, uint64_t mode_count,
                           const void* retained, void* wvm, void* stream) {
    return guarded([&] {
        if (!plan) fail(SKB_INVALID_ARGUMENT, "null plan");
        require_ptrs(wvm, retained);
        refresh_mode_tables(plan, modes, mode_count, as_stream(stream));
        const auto& w = plan->w;
        const int planes = static_cast<int>(w.nz * w.fields);
        const int nxh = static_cast<int>(w.nx / 2 + 1);
        const size_t full = static_cast<size_t>(nxh) * w.ny * planes;
        const size_t total = plan->embed_entries * static_cast<size_t>(planes);
        if (plan->precision == SKB_F64) {
            hipLaunchKernelGGL(zero_kernel<double>, stream_grid(full), dim3(256), 0, as_stream(stream),
                               static_cast<cx<double>*>(wvm), full);
            hipLaunchKernelGGL(embed_kernel<double>, stream_grid(total), dim3(256), 0, as_stream(stream),
                               static_cast<const cx<double>*>(retained), static_cast<cx<double>*>(wvm),
                               plan->embed_table.as<ColEntry>(), planes, nxh, total);
        } else {
            hipLaunchKernelGGL(zero_kernel<float>, stream_grid(full), dim3(256), 0, as_stream(stream),
                               static_cast<cx<float>*>(wvm), full);
            hipLaunchKernelGGL(embed_kernel<float>, stream_grid(total), dim3(256), 0, as_stream(stream),
                               static_cast<const cx<float>*>(retained), static_cast<cx<float>*>(wvm),
                               plan->embed_table.as<ColEntry>(), planes, nxh, total);
        }
        check_launch("embed launch");
    });
}

int skb_pruned_plan_create(const skb_workload* workload, const skb_retained_mode* modes, uint64_t mode_count,
                           int precision, skb_pruned_plan** plan) {
    return guarded([&] {
        if (!plan) fail(SKB_INVALID_ARGUMENT, "null plan pointer");
        *plan = nullptr;
        check_precision(precision);
        validate_workload(workload);
        check_fft_sizes(*workload);
        auto p = std::make_unique<skb_pruned_plan>();
        p->w = *workload;
        p->precision = precision;
        p->modes = mode_count;
        p->lists = build_column_lists(*workload, modes, mode_count);
        p->tw.build(precision, workload->nx, workload->ny);
        p->fwd_start.upload(p->lists.fwd_start.data(), p->lists.fwd_start.size() * sizeof(int));
        p->inv_start.upload(p->lists.inv_start.data(), p->lists.inv_start.size() * sizeof(int));
        p->fwd.upload(p->lists.fwd.data(), p->lists.fwd.size() * sizeof(ColEntry));
        p->inv.upload(p->lists.inv.data(), p->lists.inv.size() * sizeof(ColEntry));
        p->fmap.upload(p->lists.fmap.data(), p->lists.fmap.size() * sizeof(int));
        p->imap.upload(p->lists.imap.data(), p->lists.imap.size() * sizeof(int));
        p->planes = static_cast<int>(workload->nz * workload->fields);
        p->kc = static_cast<int>(p->lists.kc);
        const int line = precision == SKB_F64 ? 8 : 16;
        p->ps = workload->nx >= 64 && workload->ny >= 64 ? (p->planes + line - 1) / line * line : p->planes;
        p->scratch.allocate(p->lists.kc * workload->ny * static_cast<uint64_t>(p->ps) * complex_bytes(precision));
        if (plane_supported(*workload, p->lists.kc, mode_count, precision) &&
            plane_band_empty(p->lists.fmap, p->lists.imap, p->lists.kc)) {
            p->plane.build(*workload, p->lists.kc, mode_count, p->lists.fmap, p->lists.imap);
            p->pm.allocate(static_cast<uint64_t>(p->plane.np) * p->planes * sizeof(cx<float>));
            p->plane_built = true;
        }
        *plan = p.release();
    });
}

int skb_pruned_plan_destroy(skb_pruned_plan* plan) {
    return guarded([&] { delete plan; });
}

int skb_pruned_forward(skb_pruned_plan* plan, const void* real, void* retained, void* stream) {
    return guarded([&] {
        if (!plan) fail(SKB_INVALID_ARGUMENT, "null plan");
        require_ptrs(real, retained);
        const hipStream_t s = as_stream(stream);
        if (uses_plane(plan)) {
            plane_forward(plan->plane, real, plan->pm.ptr, plan->planes, s);
            pm_transpose(plan->plane, false, plan->pm.ptr, retained, plan->planes, s);
            return;
        }
        SKB_BY_PRECISION(plan, stage_forward_rows, plan, real, s);
        SKB_BY_PRECISION(plan, stage_forward_columns, plan, retained, s);
    });
}

int skb_pruned_inverse(skb_pruned_plan* plan, const void* retained, void* real, void* stream) {
    return guarded([&] {
        if (!plan) fail(SKB_INVALID_ARGUMENT, "null plan");
        require_ptrs(real, retained);
        const hipStream_t s = as_stream(stream);
        if (uses_plane(plan)) {
            pm_transpose(plan->plane, true, retained, plan->pm.ptr, plan->planes, s);
            plane_inverse(plan->plane, plan->pm.ptr, real, plan->planes, s);
            return;
        }
        SKB_BY_PRECISION(plan, stage_inverse_columns, plan, retained, s);
        SKB_BY_PRECISION(plan, stage_inverse_rows, plan, real, s);
    });
}

int skb_pruned_plan_set_path(skb_pruned_plan* plan, int path) {
    return guarded([&] {
        if (!plan || path < SKB_PATH_AUTO || path > SKB_PATH_PLANE) fail(SKB_INVALID_ARGUMENT, "bad transform path");
        if (path == SKB_PATH_PLANE && !plan->plane_built)
            fail(SKB_UNSUPPORTED, "the plane-resident path needs FP32 256 x 256 planes with the 2/3-rule k_x band");
        plan->path = path;
    });
}

int skb_pruned_plan_path(const skb_pruned_plan* plan) {
    return plan && uses_plane(plan) ? SKB_PATH_PLANE : SKB_PATH_TWO_PASS;
}

int skb_pruned_forward_split(skb_pruned_plan* plan, const void* real, void* retained_real, void* retained_imag,
                             double scale, void* stream) {
    return guarded([&] {
        if (!plan) fail(SKB_INVALID_ARGUMENT, "null plan");
        require_ptrs(real, retained_real);
        const hipStream_t s = as_stream(stream);
        if (plan->precision == SKB_F64) {
            const auto v = plain_split<double>(retained_real, retained_imag, plan->planes, scale);
            stage_forward_rows<double>(plan, real, s);
            stage_forward_columns_split<double>(plan, v, s);
        } else {
            const auto v = plain_split<float>(retained_real, retained_imag, plan->planes, scale);
            stage_forward_rows<float>(plan, real, s);
            stage_forward_columns_split<float>(plan, v, s);
        }
    });
}

int skb_pruned_inverse_split(skb_pruned_plan* plan, const void* retained_real, const void* retained_imag, void* real,
                             void* stream) {
    return guarded([&] {
        if (!plan) fail(SKB_INVALID_ARGUMENT, "null plan");
        require_ptrs(real, retained_real);
        const hipStream_t s = as_stream(stream);
        if (plan->precision == SKB_F64) {
            const auto v = plain_split<double>(retained_real, retained_imag, plan->planes, 1.0);
            stage_inverse_columns_split<double>(plan, v, s);
            stage_inverse_rows<double>(plan, real, s);
        } else {
            const auto v = plain_split<float>(retained_real, retained_imag, plan->planes, 1.0);
            stage_inverse_columns_split<float>(plan, v, s);
            stage_inverse_rows<float>(plan, real, s);
        }
    });
}

int skb_pruned_forward_split_fields(skb_pruned_plan* plan, const void* real, const skb_field_view* fields,
                                    uint64_t field_count, double scale, void* stream) {
    return guarded([&] {
        if (!plan || !real) fail(SKB_INVALID_ARGUMENT, "null argument");
        const hipStream_t s = as_stream(stream);
        if (plan->precision == SKB_F64) {
            const auto v = field_split<double>(plan, fields, field_count, scale, true);
            stage_forward_rows<double>(plan, real, s);
            stage_forward_columns_split<double>(plan, v, s);
        } else {
            const auto v = field_split<float>(plan, fields, field_count, scale, true);
            stage_forward_rows<float>(plan, real, s);
            stage_forward_columns_split<float>(plan, v, s);
        }
    });
}

int skb_pruned_inverse_split_fields(skb_pruned_plan* plan, const skb_field_view* fields, uint64_t field_count,
                                    void* real, void* stream) {
    return guarded([&] {
        if (!plan || !real) fail(SKB_INVALID_ARGUMENT, "null argument");
        const hipStream_t s = as_stream(stream);
        if (plan->precision == SKB_F64) {
            const auto v = field_split<double>(plan, fields, field_count, 1.0, false);
            stage_inverse_columns_split<double>(plan, v, s);
            stage_inverse_rows<double>(plan, real, s);
        } else {
            const auto v = field_split<float>(plan, fields, field_count, 1.0, false);
            stage_inverse_columns_split<float>(plan, v, s);
            stage_inverse_rows<float>(plan, real, s);
        }
    });
}

int skb_pruned_forward_rows(skb_pruned_plan* plan, const void* real, void* stream) {
    return guarded([&] {
        if (!plan || !real) fail(SKB_INVALID_ARGUMENT, "null argument");
        SKB_BY_PRECISION(plan, stage_forward_rows, plan, real, as_stream(stream));
    });
}

int skb_pruned_forward_columns(skb_pruned_plan* plan, void* retained, void* stream) {
    return guarded([&] {
        if (!plan || !retained) fail(SKB_INVALID_ARGUMENT, "null argument");
        SKB_BY_PRECISION(plan, stage_forward_columns, plan, retained, as_stream(stream));
    });
}

int skb_pruned_inverse_columns(skb_pruned_plan* plan, const void* retained, void* stream) {
    return guarded([&] {
        if (!plan || !retained) fail(SKB_INVALID_ARGUMENT, "null argument");
        SKB_BY_PRECISION(plan, stage_inverse_columns, plan, retained, as_stream(stream));
    });
}

int skb_pruned_inverse_rows(skb_pruned_plan* plan, void* real, void* stream) {
    return guarded([&] {
        if (!plan || !real) fail(SKB_INVALID_ARGUMENT, "null argument");
        SKB_BY_PRECISION(plan, stage_inverse_rows, plan, real, as_stream(stream));
    });
}

}  // extern "C"

namespace skb {

const PlaneTables* pruned_plane_tables(const skb_pruned_plan* p) {
    return p && p->plane_built ? &p->plane : nullptr;
}
void* pruned_plane_scratch(const skb_pruned_plan* p) { return p && p->plane_built ? p->pm.ptr : nullptr; }
int pruned_planes(const skb_pruned_plan* p) { return p ? p->planes : 0; }

bool flux_rows_supported(const skb_pruned_plan* pin, const skb_pruned_plan* pout) {
    if (!pin || !pout) return false;
    const auto &a = pin->w, &b = pout->w;
    return a.nx == b.nx && a.ny == b.ny && a.nz == b.nz && a.nx >= 64 && pin->kc == pout->kc &&
           pin->precision == SKB_F32 && pout->precision == SKB_F32 && a.fields == 3 + 3 * b.fields;
}

namespace {
template <typename R, int NX>
void launch_flux_rows(skb_pruned_plan* pin, skb_pruned_plan* pout, int targets, double scale, int workgroups,
                      hipStream_t s) {
    if constexpr (NX >= 64) {
        using G = RowGeo<R, NX>;
        constexpr int TP = 2 * G::PAIRS;
        const int nz = static_cast<int>(pin->w.nz);
        const int items = static_cast<int>(pin->w.ny) * ((nz + TP - 1) / TP);
        const int grid = launch_grid(items, workgroups);
        const auto* tin = static_cast<const cx<R>*>(pin->scratch.ptr);
        auto* tout = static_cast<cx<R>*>(pout->scratch.ptr);
        const auto* tw = pin->tw.x.as<cx<R>>();
        if (grid < items)
            hipLaunchKernelGGL((flux_rows_kernel<R, NX, true>), dim3(grid), dim3(G::THREADS), 0, s, tin, tout, tw, nz,
                               pin->ps, pout->ps, pin->kc, targets, static_cast<R>(scale), items);
        else
            hipLaunchKernelGGL((flux_rows_kernel<R, NX, false>), dim3(grid), dim3(G::THREADS), 0, s, tin, tout, tw,
                               nz, pin->ps, pout->ps, pin->kc, targets, static_cast<R>(scale), items);
        check_launch("flux rows launch");
    } else {
        fail(SKB_UNSUPPORTED, "fused flux rows need Nx >= 64");
    }
}

template <typename R>
void flux_rows_t(skb_pruned_plan* pin, skb_pruned_plan* pout, int targets, double scale, int workgroups,
                 hipStream_t s) {
    switch (pin->w.nx) {
#define SKB_FLUX_CASE(N)                                                     \
    case N:                                                                  \
        launch_flux_rows<R, N>(pin, pout, targets, scale, workgroups, s);    \
        return;
        SKB_POW2_CASES(SKB_FLUX_CASE)
#undef SKB_FLUX_CASE
    }
    fail(SKB_INVALID_ARGUMENT, "unsupported Nx");
}
}  // namespace

namespace {
template <typename R>
void flux_inverse_columns_t(skb_pruned_plan* pin, const FluxSource* src, int count, hipStream_t s) {
    SplitViews<R> v{};
    const int nz = static_cast<int>(pin->w.nz);
    for (int f = 0; f < count; ++f) {
        if (!src[f].base || src[f].field < 0 || src[f].field >= src[f].fields)
            fail(SKB_INVALID_ARGUMENT, "bad flux source");
        R* re = static_cast<R*>(const_cast<void*>(src[f].base)) + 2 * static_cast<size_t>(nz) * src[f].field;
        v.re[f] = re;
        v.im[f] = re + 1;
        v.stride[f] = 2LL * nz * src[f].fields;
        v.factor[f] = src[f].factor;
    }
    v.nz = nz;
    v.zs = 2;
    v.ny = static_cast<int>(pin->w.ny);
    v.scale = R(1);
    v.kx_scale = static_cast<R>(2.0 * M_PI / pin->w.lx);
    v.ky_scale = static_cast<R>(2.0 * M_PI / pin->w.ly);
    stage_inverse_columns_split<R>(pin, v, s);
}
}  // namespace

void flux_inverse_columns(skb_pruned_plan* pin, const FluxSource* sources, int count, void* stream) {
    if (!pin || !sources || count != static_cast<int>(pin->w.fields) || count > kMaxSplitViews)
        fail(SKB_INVALID_ARGUMENT, "flux sources must match the plan's fields (at most 16)");
    if (pin->precision == SKB_F64) flux_inverse_columns_t<double>(pin, sources, count, as_stream(stream));
    else flux_inverse_columns_t<float>(pin, sources, count, as_stream(stream));
}

void flux_rows(skb_pruned_plan* pin, skb_pruned_plan* pout, int targets, double scale, void* stream,
               int workgroups) {
    if (!flux_rows_supported(pin, pout)) fail(SKB_RUNTIME_ERROR, "fused flux rows: mismatched plans");
    flux_rows_t<float>(pin, pout, targets, scale, workgroups, as_stream(stream));
}

}  // namespace skb

extern "C" {

int skb_pruned_plan_set_grid(skb_pruned_plan* plan, int stage, int workgroups) {
    return guarded([&] {
        if (!plan || stage < 0 || stage > 3 || workgroups < 0) fail(SKB_INVALID_ARGUMENT, "bad stage or workgroup count");
        plan->grid[stage] = workgroups;
    });
}

uint64_t skb_pruned_plan_items(const skb_pruned_plan* plan, int stage) {
    if (!plan || stage < 0 || stage > 3) return 0;
    const auto& w = plan->w;
    const uint64_t planes = w.nz * w.fields;
    const bool f32 = plan->precision == SKB_F32;
    const bool rows = stage == kStageForwardRows || stage == kStageInverseRows;
    const uint64_t n = rows ? w.nx : w.ny;
    uint64_t tile;
    if (n >= 64) {
        const uint64_t n1 = n == 64 ? 8 : (n <= 256 ? 16 : 32);
        const uint64_t lanes = f32 ? 256 : 128;
        tile = rows ? 2 * (lanes / n1) : lanes / n1;
    } else {
        tile = rows ? (f32 ? 16 : 8) : (f32 ? 16 : 8);
    }
    const uint64_t tiles = (planes + tile - 1) / tile;
    return rows ? w.ny * tiles : plan->lists.kc * tiles;
}

uint64_t skb_pruned_active_kx(const skb_pruned_plan* plan) { return plan ? plan->lists.kc : 0; }

uint64_t skb_pruned_scratch_bytes(const skb_pruned_plan* plan) { return plan ? plan->scratch.bytes : 0; }

}  // extern "C"
