// K^2-grouped vertical GEMM on the matrix cores of gfx950.
//
// For group g (modes [first, first+count)) the columns are
// c = field + fields*mode, each a contiguous vector of K complex values
// (K = Nz forward, Nj inverse).  The product is
//     out[:, c] = A_g (M x K, real, row-major) * in[:, c]      (complex)
// i.e. one real matrix applied to the real and imaginary parts separately --
// the split-dgemm formulation of src/vertical_gemm.cpp:558-596, which the
// reference shows equals its complex zgemm (src/vertical_gemm.cpp:534-556).
//
// Two column layouts share one kernel:
//   compact    column c at in + K*c, out + M*c (VerticalGemmProvider,
//              src/vertical_gemm.cpp:395-422);
//   WVM direct column (mode, field) at in + K*(F*q + field), out + M*(F*q +
//              field) with q = stored_kx + (Nx/2+1)*stored_ky the mode's
//              frequency, and the k_x = 0 Hermitian partner written as the
//              conjugate (WvmDirectVerticalGemmProvider,
//              src/vertical_gemm.cpp:1081-1157).
//
// Work decomposition: one workgroup (4 wavefronts) per tile of 16 consecutive
// columns of one group.  The tile's B block (16 x K complex) is loaded with
// coalesced reads into LDS; each wavefront owns row blocks mb = wave,
// wave + 4, ... of 16 output rows and runs 16x16x4 MFMA
// (v_mfma_f32_16x16x4_f32 / v_mfma_f64_16x16x4_f64) over K in steps of 4,
// two accumulators (Re, Im) per row block, A fragments held in registers.
// When M = 16*MB + r with r <= 4 (M = 65: four blocks and one row), the r
// leftover rows are dot products on the vector ALU instead of a mostly-empty
// fifth MFMA block.  The 16 x M result is staged in the same LDS and written
// with coalesced stores.
//
// MFMA 16x16x4 operand layout (64 lanes, lane = l), measured on gfx950:
//   A[i][k]: i = l % 16, k = l / 16        B[k][j]: j = l % 16, k = l / 16
//   D[i][j]: j = l % 16 and, for accumulator element r,
//            f32: i = 4 * (l / 16) + r      f64: i = l / 16 + 4 * r
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "common.hpp"
#include "internal.hpp"

namespace skb {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

template <typename R>
struct Acc;
template <>
struct Acc<float> {
    using type = floatx4;
    static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
    static __device__ __forceinline__ type mfma(float a, float b, type c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};
template <>
struct Acc<double> {
    using type = doublex4;
    static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
    static __device__ __forceinline__ type mfma(double a, double b, type c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};

// A tile is up to 16 consecutive complex columns of one group.
struct GemmTile {
    int group;
    int col_begin;
    int ncols;  // 1..16
    int pad;
};

// Record i (wave-uniform) of the tile table.  The table is read-only while a
// kernel runs; reading it through the constant address space makes the fetch
// a scalar load, counted by lgkmcnt instead of the vector-memory counter, so
// waiting for a record does not also wait for the stores in flight.
__device__ __forceinline__ GemmTile tile_record(const GemmTile* tiles, int i) {
    using ConstTile = const __attribute__((address_space(4))) GemmTile;
    const ConstTile* t = (ConstTile*)tiles + __builtin_amdgcn_readfirstlane(i);
    return GemmTile{t->group, t->col_begin, t->ncols, t->pad};
}

// Row blocks per wavefront held at once (accumulator slots): 3 covers
// M <= 196; instantiations with KS >= 22 (K > 84, i.e. large Nz) hold 4,
// covering M <= 260.
template <int KS>
constexpr int gemm_slots() {
    return KS >= 22 ? 4 : 3;
}
constexpr bool gemm_ks_instantiated(int ks, bool fp64) {
    return (ks >= 1 && ks <= 17) || ks == 22 || ks == 33 || (!fp64 && (ks == 43 || ks == 65));
}

// LDS column stride of the B block, in complex elements.  With KP = 4*KS
// the MFMA B reads of one instruction (16 columns x 4 k) hit distinct banks:
// FP32 (8-byte reads, two half-waves) needs CS/2 odd, FP64 (16-byte reads,
// four quarter-waves) needs CS odd.
template <typename R, int KS>
constexpr int gemm_cs() {
    return sizeof(R) == 4 ? 4 * KS + 2 : 4 * KS + 1;
}

template <typename R>
struct GemmArgs {
    const cx<R>* in;
    cx<R>* out;
    const R* ap;    // [group][MB][KS][64]: A[16*mb + l%16][4*s + l/16] (zero outside M x K)
    const R* arow;  // [group][VR][4*KS]: rows 16*MB .. M-1 (zero for k >= K)
    const GemmTile* tiles;
    int ntiles, M, K, MB, VR, F;
    int in_ld, out_ld;   // compact: column c at in + in_ld*c and out + out_ld*c (complex elements)
    const int* qmap;     // WVM direct: frequency q of each mode
    const int* partner;  // WVM direct: frequency of the k_x = 0 conjugate partner, or -1
    // split-BF16 products (FP32 plans with K = 32*NK + 1): A = A0 + A1 + A2
    // in bf16, [group][MB][NK][3][64 lanes][8], lane (i, q) holding
    // A[16*mb + i][32*t + 8*q + e]; and the last column A[row][K - 1] in FP32,
    // [group][16*MB].
    const uint16_t* a3;
    const float* atail;
    // chunked plane-major input (split kernel, PM = true): element (mode, z,
    // field) at in[16*P*(mode / 16) + 16*(z + pm_nz * field) + mode % 16],
    // P = pm_planes, pm_np modes (a multiple of 16); a tile is up to 16
    // consecutive modes of one chunk, one group and one field
    // (GemmTile::pad = field), written to the modal layout
    // out + out_ld * (field + F * mode).  The input holds slot values: modes
    // with bit m % 16 of pm_cmask[m / 16] set are stored conjugated, and the
    // product, real-matrix times column, is conjugated back on output.
    int pm_np, pm_nz, pm_planes;
    const uint32_t* pm_cmask;
};

template <typename R, int KS, bool WVM>
__global__ __launch_bounds__(256) void grouped_gemm_kernel(const GemmArgs<R> g) {
    using A4 = typename Acc<R>::type;
    constexpr int KP = 4 * KS;
    constexpr int CS = gemm_cs<R, KS>();
    constexpr int kMaxSlots = gemm_slots<KS>();
    constexpr int NL = (16 * KP + 255) / 256;  // B elements loaded per thread
    extern __shared__ __align__(16) unsigned char gemm_lds[];
    cx<R>* sb = reinterpret_cast<cx<R>*>(gemm_lds);  // B block, then the output stage
    __shared__ int cbase[16], obase[16], pbase[16];
    __shared__ R sar[4 * KP];  // vector-row coefficients (VR <= 4)
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int j = lane & 15, q = lane >> 4;
    const bool one = gridDim.x >= static_cast<unsigned>(g.ntiles);
    const int step = one ? g.ntiles : static_cast<int>(gridDim.x);
    for (int tile = one ? xcd_item(blockIdx.x, g.ntiles) : static_cast<int>(blockIdx.x); tile < g.ntiles;
         tile += step) {
        const GemmTile t = tile_record(g.tiles, tile);
        // A fragments of this wavefront's first row block: issued before the
        // B block so both loads are in flight together.
        R a0[KS];
        if (wave < g.MB) {
            const R* a_ptr = g.ap + (static_cast<size_t>(t.group) * g.MB + wave) * KS * 64 + lane;
#pragma unroll
            for (int s = 0; s < KS; ++s) a0[s] = a_ptr[s * 64];
        }
        // vector-row coefficients of the group -> LDS with the B block
        {
            const R* ap = g.arow + static_cast<size_t>(t.group) * g.VR * KP;
            for (int e = tid; e < g.VR * KP; e += 256) sar[e] = ap[e];
        }
        if constexpr (WVM) {
            if (tid < 16) {
                int cb = 0, ob = 0, pb = -1;
                if (tid < t.ncols) {
                    const int col = t.col_begin + tid, mode = col / g.F, f = col - mode * g.F;
                    const int qf = g.qmap[mode], pf = g.partner[mode];
                    cb = g.K * (g.F * qf + f);
                    ob = g.M * (g.F * qf + f);
                    pb = pf >= 0 ? g.M * (g.F * pf + f) : -1;
                }
                cbase[tid] = cb;
                obase[tid] = ob;
                pbase[tid] = pb;
            }
            __syncthreads();
        }
        // B block -> LDS, zero outside the tile's columns and for k >= K
        cx<R> bv[NL];
#pragma unroll
        for (int n = 0; n < NL; ++n) {
            const int e = tid + 256 * n, c = e / KP, z = e - c * KP;
            bv[n] = cx<R>{R(0), R(0)};
            if (e < 16 * KP && c < t.ncols && z < g.K) {
                const size_t src =
                    WVM ? static_cast<size_t>(cbase[c]) + z
                        : static_cast<size_t>(g.in_ld) * (t.col_begin + c) + z;
                bv[n] = g.in[src];
            }
        }
#pragma unroll
        for (int n = 0; n < NL; ++n) {
            const int e = tid + 256 * n, c = e / KP, z = e - c * KP;
            if (e < 16 * KP) sb[c * CS + z] = bv[n];
        }
        __syncthreads();
        // MFMA row blocks mb = wave + 4 * slot
        A4 re[kMaxSlots], im[kMaxSlots];
#pragma unroll
        for (int sl = 0; sl < kMaxSlots; ++sl) {
            const int mb = wave + 4 * sl;
            if (mb < g.MB) {
                R a[KS];
                if (sl == 0) {
#pragma unroll
                    for (int s = 0; s < KS; ++s) a[s] = a0[s];
                } else {
                    const R* a_ptr = g.ap + (static_cast<size_t>(t.group) * g.MB + mb) * KS * 64 + lane;
#pragma unroll
                    for (int s = 0; s < KS; ++s) a[s] = a_ptr[s * 64];
                }
                A4 r = {0, 0, 0, 0}, i = {0, 0, 0, 0};
#pragma unroll
                for (int s = 0; s < KS; ++s) {
                    const cx<R> b = sb[j * CS + 4 * s + q];
                    r = Acc<R>::mfma(a[s], b.x, r);
                    i = Acc<R>::mfma(a[s], b.y, i);
                }
                re[sl] = r;
                im[sl] = i;
            }
        }
        // Leftover rows 16*MB + wave (wave < VR): lane (j, q) sums k in
        // [q*KS, q*KS + KS) of column j; the four partial sums are added
        // across lanes j, j+16, j+32, j+48.
        cx<R> vrow{R(0), R(0)};
        if (wave < g.VR) {
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                const R a = sar[wave * KP + q * KS + s];
                const cx<R> b = sb[j * CS + q * KS + s];
                vrow.x += a * b.x;
                vrow.y += a * b.y;
            }
            vrow.x += __shfl_xor(vrow.x, 16);
            vrow.y += __shfl_xor(vrow.y, 16);
            vrow.x += __shfl_xor(vrow.x, 32);
            vrow.y += __shfl_xor(vrow.y, 32);
        }
        __syncthreads();  // every wavefront is done reading B
        // stage the 16 x M result column-major (the output layout)
        cx<R>* st = sb;
#pragma unroll
        for (int sl = 0; sl < kMaxSlots; ++sl) {
            const int mb = wave + 4 * sl;
            if (mb < g.MB) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = 16 * mb + Acc<R>::row(lane, r);
                    if (row < g.M) st[j * g.M + row] = cx<R>{re[sl][r], im[sl][r]};
                }
            }
        }
        if (wave < g.VR && q == 0) st[j * g.M + 16 * g.MB + wave] = vrow;
        __syncthreads();
        const int total = g.M * t.ncols;
        for (int e = tid; e < total; e += 256) {
            const cx<R> v = st[e];
            if constexpr (WVM) {
                const int c = e / g.M, row = e - c * g.M;
                g.out[static_cast<size_t>(obase[c]) + row] = v;
                if (pbase[c] >= 0) g.out[static_cast<size_t>(pbase[c]) + row] = conjg(v);
            } else if (g.out_ld == g.M) {
                g.out[static_cast<size_t>(g.M) * t.col_begin + e] = v;
            } else {
                const int c = e / g.M, row = e - c * g.M;
                g.out[static_cast<size_t>(g.out_ld) * (t.col_begin + c) + row] = v;
            }
        }
        __syncthreads();  // the stage and the column bases are reused by the next tile
    }
}

// Streaming form for the compact layout with the output height M fixed at
// compile time.  Persistent workgroups each own a contiguous run of tiles
// (tiles are ordered by group, so a run mostly stays in one group).  The A
// fragments of every row block stay in registers and the vector-row
// coefficients in LDS until the group changes; the B block of the tile after
// next is loaded into registers while the current tile is multiplied.
// Loads and stores go through buffer resources sized to the tile, so a
// column past the tile reads zero and an element past it is not written:
// every thread issues the same, compile-time number of memory operations per
// tile.  On gfx9 one in-order counter tracks loads and stores, so a fixed
// store count lets the wait for the prefetched B block skip the current
// tile's stores.
constexpr int rows_mb(int M) { return (M >= 16 && M % 16 <= 4) ? M / 16 : (M + 15) / 16; }
constexpr int rows_vr(int M) { return (M >= 16 && M % 16 <= 4) ? M % 16 : 0; }

template <typename R, int KS, int M>
__global__ __launch_bounds__(256) void grouped_gemm_fixed_kernel(const GemmArgs<R> g) {
    using A4 = typename Acc<R>::type;
    constexpr int KP = 4 * KS;
    constexpr int CS = gemm_cs<R, KS>();
    constexpr int MB = rows_mb(M), VR = rows_vr(M), SL = (MB + 3) / 4;
    constexpr int NL = (16 * KP + 255) / 256;  // B elements per thread
    constexpr int NS = (16 * M + 255) / 256;   // output elements per thread
    constexpr uint32_t CB = sizeof(cx<R>);
    __shared__ __align__(16) cx<R> sb[16 * CS];  // B block
    __shared__ __align__(16) cx<R> st[16 * M];   // output stage, column-major
    __shared__ R sar[(VR > 0 ? VR : 1) * KP];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int j = lane & 15, q = lane >> 4;
    const int nwg = static_cast<int>(gridDim.x), wg = static_cast<int>(blockIdx.x);
    const int t_begin = static_cast<int>((static_cast<long long>(g.ntiles) * wg) / nwg);
    const int t_end = static_cast<int>((static_cast<long long>(g.ntiles) * (wg + 1)) / nwg);
    if (t_begin >= t_end) return;  // uniform across the workgroup
    const int in_step = g.in_ld;  // input column stride
    auto load_b = [&](cx<R>* bv, const GemmTile& t) {
        const cx<R>* base = g.in + static_cast<size_t>(g.in_ld) * t.col_begin;
        const uint32_t bytes = static_cast<uint32_t>(in_step * (t.ncols - 1) + g.K) * CB;
        const __amdgpu_buffer_rsrc_t r = buffer_rsrc(base, bytes);
#pragma unroll
        for (int n = 0; n < NL; ++n) {
            const int e = tid + 256 * n, c = e / KP, z = e - c * KP;
            bv[n] = buffer_load_cx<R>(r, static_cast<uint32_t>(in_step * c + z) * CB);
        }
    };
    // Padding rows z >= K are zeroed here rather than at load time, so the
    // prefetch is not waited for until its block is needed.
    auto put_b = [&](const cx<R>* bv) {
#pragma unroll
        for (int n = 0; n < NL; ++n) {
            const int e = tid + 256 * n, c = e / KP, z = e - c * KP;
            if (16 * KP % 256 == 0 || e < 16 * KP) sb[c * CS + z] = z < g.K ? bv[n] : cx<R>{R(0), R(0)};
        }
    };
    R a[SL][KS];
    // The vector-ALU row's coefficients pass through registers on their way
    // to LDS, so issuing them does not wait for the load.
    constexpr int NR = (VR * KP + 255) / 256;
    R rv[NR > 0 ? NR : 1];
    auto issue_group = [&](int group) {
#pragma unroll
        for (int sl = 0; sl < SL; ++sl) {
            const int mb = wave + 4 * sl;
            if (mb < MB) {
                const R* a_ptr = g.ap + (static_cast<size_t>(group) * MB + mb) * KS * 64 + lane;
#pragma unroll
                for (int s = 0; s < KS; ++s) a[sl][s] = a_ptr[s * 64];
            }
        }
        if constexpr (VR > 0) {
            const R* ap = g.arow + static_cast<size_t>(group) * VR * KP;
#pragma unroll
            for (int n = 0; n < NR; ++n) {
                const int e = tid + 256 * n;
                rv[n] = ap[e < VR * KP ? e : 0];
            }
        }
    };
    auto put_row = [&]() {
        if constexpr (VR > 0) {
#pragma unroll
            for (int n = 0; n < NR; ++n) {
                const int e = tid + 256 * n;
                if (e < VR * KP) sar[e] = rv[n];
            }
        }
    };
    auto load_group = [&](int group) {
        issue_group(group);
        put_row();
        // Group changes are rare: complete the fragment loads here, so the
        // multiply of the following tiles carries no wait for them.
        __builtin_amdgcn_s_waitcnt(0);
    };
    // Prologue: the group's fragments and the first two B blocks are all in
    // flight before the first wait.  A workgroup with one tile loads its
    // block twice rather than branch around the second load.
    GemmTile cur = tile_record(g.tiles, t_begin);
    int group = cur.group;
    issue_group(group);
    cx<R> bv[NL];
    load_b(bv, cur);
    GemmTile nxt = tile_record(g.tiles, t_begin + 1 < t_end ? t_begin + 1 : t_begin);
    cx<R> bn[NL];
    load_b(bn, nxt);
    put_b(bv);
    put_row();
#pragma unroll
    for (int n = 0; n < NL; ++n) bv[n] = bn[n];
    __syncthreads();
    for (int tile = t_begin; tile < t_end; ++tile) {
        A4 re[SL], im[SL];
#pragma unroll
        for (int sl = 0; sl < SL; ++sl) {
            const int mb = wave + 4 * sl;
            if (mb < MB) {
                A4 r = {0, 0, 0, 0}, i = {0, 0, 0, 0};
#pragma unroll
                for (int s = 0; s < KS; ++s) {
                    const cx<R> b = sb[j * CS + 4 * s + q];
                    r = Acc<R>::mfma(a[sl][s], b.x, r);
                    i = Acc<R>::mfma(a[sl][s], b.y, i);
                }
                re[sl] = r;
                im[sl] = i;
            }
        }
        cx<R> vrow{R(0), R(0)};
        if (VR > 0 && wave < VR) {
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                const R av = sar[wave * KP + q * KS + s];
                const cx<R> b = sb[j * CS + q * KS + s];
                vrow.x += av * b.x;
                vrow.y += av * b.y;
            }
            vrow.x += __shfl_xor(vrow.x, 16);
            vrow.y += __shfl_xor(vrow.y, 16);
            vrow.x += __shfl_xor(vrow.x, 32);
            vrow.y += __shfl_xor(vrow.y, 32);
        }
#pragma unroll
        for (int sl = 0; sl < SL; ++sl) {
            const int mb = wave + 4 * sl;
            if (mb < MB) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = 16 * mb + Acc<R>::row(lane, r);
                    if (row < M) st[j * M + row] = cx<R>{re[sl][r], im[sl][r]};
                }
            }
        }
        if (VR > 0 && wave < VR && q == 0) st[j * M + 16 * MB + wave] = vrow;
        __syncthreads();  // stage complete; B block and coefficients are free
        {
            // column c of the tile at out + out_ld*(col_begin + c); elements of
            // columns past the tile fall outside the resource and are dropped
            const int ld = g.out_ld;
            const __amdgpu_buffer_rsrc_t r =
                buffer_rsrc(g.out + static_cast<size_t>(ld) * cur.col_begin,
                            static_cast<uint32_t>(ld * (cur.ncols - 1) + M) * CB);
#pragma unroll
            for (int n = 0; n < NS; ++n) {
                const int e = tid + 256 * n, c = e / M, row = e - c * M;
                const cx<R> v = st[e < 16 * M ? e : 0];
                buffer_store_cx<R>(v, r, c < 16 ? static_cast<uint32_t>(ld * c + row) * CB : kBufferPast);
            }
        }
        if (tile + 1 < t_end) {
            put_b(bv);  // the B block prefetched for tile + 1
            cur = nxt;
            if (cur.group != group) {
                group = cur.group;
                load_group(group);
            }
            if (tile + 2 < t_end) {
                nxt = tile_record(g.tiles, tile + 2);
                load_b(bv, nxt);
            }
            __syncthreads();  // B block (and coefficients) of tile + 1 visible; the stage is drained
        }
    }
}

// Split-BF16 form of the fixed-height kernel (FP32 plans, K = 32*NK + 1):
// each product A*b is formed as the six leading terms of
// (A0 + A1 + A2)(b0 + b1 + b2) on the bf16 matrix cores
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation), smallest terms first; the
// three dropped terms are below 2^-24 of the product.  A is split once on
// the host (from the FP64 matrix); each B element is split once per tile, by
// the thread that moves it into LDS, into six bf16 planes (hi, mid, lo of Re
// and of Im).  The last k (K - 1 = 32*NK) and the vector-ALU rows use the
// FP32 copy of the B block.  Tile schedule, prefetch and stores are those of
// grouped_gemm_fixed_kernel.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ uint32_t bf16_pair(float lo, float hi) {
    // round to nearest even (v_cvt_pk_bf16_f32)
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    const bf16x2 v{static_cast<__bf16>(lo), static_cast<__bf16>(hi)};
    return __builtin_bit_cast(uint32_t, v);
}

// A batch of plane-major evaluations in one launch (PM): virtual tile v is
// tile v % ntiles of evaluation v / ntiles, whose input block sits
// in_stride complex elements further on and whose output is out[e].
struct GemmBatch {
    cx<float>* out[kBatchMax];
    int count;
    size_t in_stride;
};

template <int NK, int M, bool PM>
__global__ __launch_bounds__(256) void grouped_gemm_split_kernel(const GemmArgs<float> g, const GemmBatch bt) {
    using A4 = floatx4;
    constexpr int KS = (32 * NK + 1 + 3) / 4;  // FP32 copy: K rounded up to 4
    constexpr int KP = 4 * KS;
    constexpr int CS = gemm_cs<float, KS>();
    constexpr int MB = rows_mb(M), VR = rows_vr(M), SL = (MB + 3) / 4;
    constexpr int KB = 32 * NK;      // k handled on the bf16 matrix cores
    constexpr int BS = KB + 8;       // bf16 plane column stride (16-byte reads spread over banks)
    constexpr uint32_t CB = sizeof(cx<float>);
    static_assert(16 * MB <= M, "every MFMA row exists");
    static_assert(VR <= 1, "one vector-ALU row");
    // B blocks, double-buffered where LDS allows three workgroups per CU (the
    // next tile's block is staged while the current one is multiplied, one
    // barrier per tile); the largest height keeps one buffer and a barrier
    // before restaging it.
    constexpr int NB = NK >= 4 ? 1 : 2;
    __shared__ __align__(16) cx<float> sbb[NB][16 * CS];      // FP32 B block
    __shared__ __align__(16) uint16_t shb[NB][6][16 * BS];    // bf16 planes: Re hi, mid, lo, Im hi, mid, lo
    __shared__ float sar[(VR > 0 ? VR : 1) * KP];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int j = lane & 15, q = lane >> 4;
    const int nwg = static_cast<int>(gridDim.x), wg = static_cast<int>(blockIdx.x);
    const int total = PM ? g.ntiles * bt.count : g.ntiles;
    const int t_begin = static_cast<int>((static_cast<long long>(total) * wg) / nwg);
    const int t_end = static_cast<int>((static_cast<long long>(total) * (wg + 1)) / nwg);
    if (t_begin >= t_end) return;  // uniform across the workgroup
    const int in_step = g.in_ld;
    // tile record of virtual tile v; PM: the evaluation rides in pad's upper
    // bits (pad = field + 256 * evaluation)
    auto record = [&](int v) {
        if constexpr (PM) {
            const int e = v / g.ntiles;
            GemmTile r = tile_record(g.tiles, v - e * g.ntiles);
            r.pad += e << 8;
            return r;
        } else {
            return tile_record(g.tiles, v);
        }
    };
    // B staging: the first KB rows of each column as pairs of complex
    // elements (one 16-byte load, split into bf16 pairs written as 32-bit
    // words), NK pairs per thread; the last row k = KB as one 8-byte load by
    // the first 16 threads.  Columns past the tile read zero.
    struct BRegs {
        bf4 pair[NK];
        cx<float> tail;
    };
    auto load_b = [&](BRegs& bv, const GemmTile& t) {
        if constexpr (PM) {
            // column j = e % 16 of the tile, rows z, z + 1 with z = 2 (e / 16):
            // 16 lanes read one 128-byte run of consecutive modes per row (two
            // pieces when the tile's modes straddle two chunks)
            const uint32_t P = static_cast<uint32_t>(g.pm_planes);
            const __amdgpu_buffer_rsrc_t r =
                buffer_rsrc(g.in + bt.in_stride * (t.pad >> 8), static_cast<uint32_t>(g.pm_np) * P * CB);
            const uint32_t plane0 = static_cast<uint32_t>(g.pm_nz * (t.pad & 255));
            auto at = [&](int c, int z) {
                const uint32_t m = static_cast<uint32_t>(t.col_begin + c);
                return (16u * P * (m >> 4) + 16u * (plane0 + static_cast<uint32_t>(z)) + (m & 15u)) * CB;
            };
#pragma unroll
            for (int n = 0; n < NK; ++n) {
                const int e = tid + 256 * n, c = e & 15, z = 2 * (e >> 4);
                const bool ok = c < t.ncols;
                const cx<float> lo = buffer_load_cx<float>(r, ok ? at(c, z) : kBufferPast);
                const cx<float> hi = buffer_load_cx<float>(r, ok ? at(c, z + 1) : kBufferPast);
                bv.pair[n] = bf4{lo.x, lo.y, hi.x, hi.y};
            }
            bv.tail = buffer_load_cx<float>(r, tid < t.ncols ? at(tid, KB) : kBufferPast);
        } else {
            const cx<float>* base = g.in + static_cast<size_t>(g.in_ld) * t.col_begin;
            const uint32_t bytes = static_cast<uint32_t>(in_step * (t.ncols - 1) + g.K) * CB;
            const __amdgpu_buffer_rsrc_t r = buffer_rsrc(base, bytes);
#pragma unroll
            for (int n = 0; n < NK; ++n) {
                const int e = tid + 256 * n, c = e / (KB / 2), z = 2 * (e - c * (KB / 2));
                bv.pair[n] =
                    raw_buffer_load_f4(r, static_cast<int>(static_cast<uint32_t>(in_step * c + z) * CB), 0, 0);
            }
            bv.tail = buffer_load_cx<float>(r, tid < 16 ? static_cast<uint32_t>(in_step * tid + KB) * CB : kBufferPast);
        }
    };
    auto put_b = [&](const BRegs& bv, int buf) {
        cx<float>* sb = sbb[buf];
        uint16_t(*sh)[16 * BS] = shb[buf];
#pragma unroll
        for (int n = 0; n < NK; ++n) {
            const int e = tid + 256 * n;
            const int c = PM ? (e & 15) : e / (KB / 2);
            const int z = PM ? 2 * (e >> 4) : 2 * (e - c * (KB / 2));
            const bf4 x = bv.pair[n];  // Re z, Im z, Re z+1, Im z+1
            *reinterpret_cast<bf4*>(&sb[c * CS + z]) = x;
            const uint32_t hr = bf16_pair(x.x, x.z), hi = bf16_pair(x.y, x.w);
            const float r1[4] = {x.x - __uint_as_float(hr << 16), x.y - __uint_as_float(hi << 16),
                                 x.z - __uint_as_float(hr & 0xFFFF0000u), x.w - __uint_as_float(hi & 0xFFFF0000u)};
            const uint32_t mr = bf16_pair(r1[0], r1[2]), mi = bf16_pair(r1[1], r1[3]);
            const uint32_t lr = bf16_pair(r1[0] - __uint_as_float(mr << 16), r1[2] - __uint_as_float(mr & 0xFFFF0000u));
            const uint32_t li = bf16_pair(r1[1] - __uint_as_float(mi << 16), r1[3] - __uint_as_float(mi & 0xFFFF0000u));
            const int at = (c * BS + z) / 2;  // 32-bit words: z is even
            reinterpret_cast<uint32_t*>(sh[0])[at] = hr;
            reinterpret_cast<uint32_t*>(sh[1])[at] = mr;
            reinterpret_cast<uint32_t*>(sh[2])[at] = lr;
            reinterpret_cast<uint32_t*>(sh[3])[at] = hi;
            reinterpret_cast<uint32_t*>(sh[4])[at] = mi;
            reinterpret_cast<uint32_t*>(sh[5])[at] = li;
        }
        if (tid < 16) sb[tid * CS + KB] = bv.tail;
    };
    bf16x8 a[SL][NK][3];
    floatx4 at4[SL];
    constexpr int NR = (VR * KP + 255) / 256;
    float rv[NR > 0 ? NR : 1];
    auto issue_group = [&](int group) {
#pragma unroll
        for (int sl = 0; sl < SL; ++sl) {
            const int mb = wave + 4 * sl;
            if (mb < MB) {
#pragma unroll
                for (int t = 0; t < NK; ++t)
#pragma unroll
                    for (int p = 0; p < 3; ++p)
                        a[sl][t][p] = *reinterpret_cast<const bf16x8*>(
                            g.a3 + ((((static_cast<size_t>(group) * MB + mb) * NK + t) * 3 + p) * 64 + lane) * 8);
                at4[sl] = *reinterpret_cast<const floatx4*>(g.atail + static_cast<size_t>(group) * 16 * MB + 16 * mb +
                                                            4 * q);
            }
        }
        if constexpr (VR > 0) {
            const float* ap = g.arow + static_cast<size_t>(group) * VR * KP;
#pragma unroll
            for (int n = 0; n < NR; ++n) {
                const int e = tid + 256 * n;
                rv[n] = ap[e < VR * KP ? e : 0];
            }
        }
    };
    auto put_row = [&]() {
        if constexpr (VR > 0) {
#pragma unroll
            for (int n = 0; n < NR; ++n) {
                const int e = tid + 256 * n;
                if (e < VR * KP) sar[e] = rv[n];
            }
        }
    };
    // Barriers guard LDS only (lds_barrier): the output stores and the
    // prefetched B loads stay in flight across them.
    // The B blocks of the first DEPTH tiles of the range are requested with
    // the group's operand; each later block is requested as the block
    // DEPTH - 1 tiles ahead of it is staged (a deeper prefetch measured
    // slower: the workgroups' first blocks then queue behind their later ones).
    constexpr int DEPTH = 2;
    GemmTile rec[DEPTH];
#pragma unroll
    for (int i = 0; i < DEPTH; ++i) rec[i] = record(t_begin + i < t_end ? t_begin + i : t_begin);
    int group = rec[0].group;
    issue_group(group);
    BRegs pre[DEPTH];
#pragma unroll
    for (int i = 0; i < DEPTH; ++i)
        if (t_begin + i < t_end) load_b(pre[i], rec[i]);
    // rows K .. KP-1 of the FP32 copies stay zero (the vector-ALU rows read them)
    for (int e = tid; e < NB * 16 * (KP - 32 * NK - 1); e += 256) {
        const int b = e / (16 * (KP - 32 * NK - 1)), f = e - b * 16 * (KP - 32 * NK - 1);
        const int c = f / (KP - 32 * NK - 1);
        sbb[b][c * CS + 32 * NK + 1 + (f - c * (KP - 32 * NK - 1))] = cx<float>{0.f, 0.f};
    }
    put_b(pre[0], 0);
    put_row();
    lds_barrier();
    int cb = 0;  // buffer holding the current tile's block
    for (int base = t_begin; base < t_end; base += DEPTH) {
#pragma unroll
        for (int slot = 0; slot < DEPTH; ++slot) {
            const int tile = base + slot;
            if (tile >= t_end) break;  // uniform across the workgroup
            const GemmTile cur = rec[slot];
            const cx<float>* sb = sbb[cb];
            const uint16_t(*sh)[16 * BS] = shb[cb];
            A4 re[SL], im[SL];
#pragma unroll
            for (int sl = 0; sl < SL; ++sl) {
                const int mb = wave + 4 * sl;
                if (mb < MB) {
                    A4 r = {0, 0, 0, 0}, i = {0, 0, 0, 0};
#pragma unroll
                    for (int t = 0; t < NK; ++t) {
                        const int off = j * BS + 32 * t + 8 * q;
                        const bf16x8 r0 = *reinterpret_cast<const bf16x8*>(&sh[0][off]);
                        const bf16x8 r1 = *reinterpret_cast<const bf16x8*>(&sh[1][off]);
                        const bf16x8 r2 = *reinterpret_cast<const bf16x8*>(&sh[2][off]);
                        const bf16x8 i0 = *reinterpret_cast<const bf16x8*>(&sh[3][off]);
                        const bf16x8 i1 = *reinterpret_cast<const bf16x8*>(&sh[4][off]);
                        const bf16x8 i2 = *reinterpret_cast<const bf16x8*>(&sh[5][off]);
                        const bf16x8* A = a[sl][t];
                        r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[2], r0, r, 0, 0, 0);
                        i = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[2], i0, i, 0, 0, 0);
                        r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[1], r1, r, 0, 0, 0);
                        i = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[1], i1, i, 0, 0, 0);
                        r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[0], r2, r, 0, 0, 0);
                        i = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[0], i2, i, 0, 0, 0);
                        r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[1], r0, r, 0, 0, 0);
                        i = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[1], i0, i, 0, 0, 0);
                        r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[0], r1, r, 0, 0, 0);
                        i = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[0], i1, i, 0, 0, 0);
                        r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[0], r0, r, 0, 0, 0);
                        i = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[0], i0, i, 0, 0, 0);
                    }
                    // k = K - 1 in FP32: rows 4q + e of the block, column j
                    const cx<float> bt = sb[j * CS + KB];
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        r[e] += at4[sl][e] * bt.x;
                        i[e] += at4[sl][e] * bt.y;
                    }
                    re[sl] = r;
                    im[sl] = i;
                }
            }
            cx<float> vrow{0.f, 0.f};
            if (VR > 0 && wave < VR) {
                // row 16 MB (VR = 1 for these shapes): lane q sums k = 32 t +
                // 8 q + e (and k = KB on q = 0), then the sum over q -- the
                // order of gemm_pm_range_kernel, which gives the same bits
#pragma unroll
                for (int t = 0; t < NK; ++t)
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const int k = 32 * t + 8 * q + e;
                        const float av = sar[k];
                        const cx<float> b = sb[j * CS + k];
                        vrow.x += av * b.x;
                        vrow.y += av * b.y;
                    }
                if (q == 0) {
                    const float av = sar[KB];
                    const cx<float> b = sb[j * CS + KB];
                    vrow.x += av * b.x;
                    vrow.y += av * b.y;
                }
                vrow.x += __shfl_xor(vrow.x, 16);
                vrow.y += __shfl_xor(vrow.y, 16);
                vrow.x += __shfl_xor(vrow.x, 32);
                vrow.y += __shfl_xor(vrow.y, 32);
            }
            if constexpr (PM) {
                // slot values back to the caller's values (uniform word, lane bit)
                using ConstWord = const __attribute__((address_space(4))) uint32_t;
                const uint32_t word = *((ConstWord*)g.pm_cmask + (cur.col_begin >> 4));
                if ((word >> ((cur.col_begin & 15) + j)) & 1u) {
#pragma unroll
                    for (int sl = 0; sl < SL; ++sl) im[sl] = -im[sl];
                    vrow.y = -vrow.y;
                }
            }
            {
                // results straight from the accumulators: lane (j, q) holds rows
                // 4q .. 4q+3 of column j in each block, two 16-byte row pairs
                const int ld = PM ? g.out_ld * g.F : g.out_ld;
                cx<float>* obase = PM ? bt.out[cur.pad >> 8] + static_cast<size_t>(g.out_ld) *
                                                                   ((cur.pad & 255) + static_cast<size_t>(g.F) * cur.col_begin)
                                      : g.out + static_cast<size_t>(ld) * cur.col_begin;
                const __amdgpu_buffer_rsrc_t r =
                    buffer_rsrc(obase, static_cast<uint32_t>(ld * (cur.ncols - 1) + M) * CB);
                const uint32_t col = static_cast<uint32_t>(ld * j) * CB;
#pragma unroll
                for (int sl = 0; sl < SL; ++sl) {
                    const int mb = wave + 4 * sl;
                    if (mb < MB) {
                        const uint32_t at = col + static_cast<uint32_t>(16 * mb + 4 * q) * CB;
                        raw_buffer_store_f4(bf4{re[sl][0], im[sl][0], re[sl][1], im[sl][1]}, r, static_cast<int>(at),
                                            0, 0);
                        raw_buffer_store_f4(bf4{re[sl][2], im[sl][2], re[sl][3], im[sl][3]}, r,
                                            static_cast<int>(at + 2 * CB), 0, 0);
                    }
                }
                if (VR > 0 && wave < VR)
                    buffer_store_cx<float>(vrow, r,
                                           q == 0 ? col + static_cast<uint32_t>(16 * MB + wave) * CB : kBufferPast);
            }
            if (tile + 1 < t_end) {
                const int next = slot + 1 < DEPTH ? slot + 1 : 0;  // constant once the slot loop is unrolled
                if constexpr (NB == 1) lds_barrier();  // every wavefront is done with the block
                put_b(pre[next], NB == 1 ? 0 : cb ^ 1);  // the B block of tile + 1
                if (rec[next].group != group) {
                    group = rec[next].group;
                    issue_group(group);
                    lds_barrier();  // every vector-ALU row of this tile has read the old coefficients
                    put_row();
                }
                if (tile + DEPTH < t_end) {  // refill the slot this tile's block came from
                    rec[slot] = record(tile + DEPTH);
                    load_b(pre[slot], rec[slot]);
                }
                lds_barrier();  // block of tile + 1 complete; this tile's block is no longer read
                if constexpr (NB == 2) cb ^= 1;
            }
        }
    }
}

// Batched plane-major split-BF16 product (FP32 square families, K = M =
// 32*NK + 1): the tiles of many evaluations in one launch, one wavefront per
// chunk range.  A chunk range is up to 16 modes of one 16-mode chunk and one
// group; for one evaluation its F fields are one contiguous run of the modal
// layout (j + M*(field + F*mode)).  Workgroup (b, y) takes block b (up to
// kPmWaves ranges of ONE group, PmBlock) for the (evaluation, field) tiles
// [y*tpw, (y+1)*tpw) of each range, in the order e*F + f: it stages the
// group's split operand in LDS once, then every wavefront works through
// those tiles of its range with no further barrier, the B block of the next tile in flight while the
// current one is multiplied.  A B block is loaded straight into
// MFMA operand order (lane (j, q) holds z = 32t + 8q + e of mode j; each load
// instruction reads four whole 128-byte lines), split into bf16 hi/mid/lo in
// registers, and gives all 16*MB MFMA rows, the FP32 last column and the
// vector-ALU row M - 1, stored straight from the registers; the next tile's
// block is in flight meanwhile.  Rows 0 .. 16*MB - 1 use the products, order
// and rounding of grouped_gemm_split_kernel (equal bit for bit); row M - 1
// sums in another order (equal to rounding).  Slot values of modes stored
// conjugated (pm_cmask) are conjugated back on output.
struct PmBlock {
    int group;
    int range_begin;  // index into the range list
    int ranges;       // 1 .. kPmWaves
    int pad;
};

__device__ __forceinline__ PmBlock pm_block_record(const PmBlock* blocks, int i) {
    using ConstBlock = const __attribute__((address_space(4))) PmBlock;
    const ConstBlock* b = (ConstBlock*)blocks + __builtin_amdgcn_readfirstlane(i);
    return PmBlock{b->group, b->range_begin, b->ranges, b->pad};
}

// hi = RNE_bf16(x), mid = RNE_bf16(x - hi), lo = RNE_bf16(x - hi - mid) of
// two values, packed as bf16 pairs (a in the low half).
__device__ __forceinline__ void split3_pair(float a, float b, uint32_t& h, uint32_t& m, uint32_t& l) {
    h = bf16_pair(a, b);
    const float ra = a - __uint_as_float(h << 16), rb = b - __uint_as_float(h & 0xFFFF0000u);
    m = bf16_pair(ra, rb);
    l = bf16_pair(ra - __uint_as_float(m << 16), rb - __uint_as_float(m & 0xFFFF0000u));
}

constexpr int kPmWaves = 4;  // chunk ranges (wavefronts) per workgroup

template <int NK, int M>
__global__ __launch_bounds__(64 * kPmWaves, 2) void gemm_pm_range_kernel(const GemmArgs<float> g, const PmBlock* blocks,
                                                                      const GemmBatch bt, int tpw) {
    constexpr int W = kPmWaves;
    constexpr int MB = rows_mb(M), VR = rows_vr(M);
    constexpr int KB = 32 * NK;  // k on the matrix cores; k = KB is the FP32 last column
    constexpr int KS = (KB + 1 + 3) / 4, KP = 4 * KS;
    constexpr int AF = MB * NK * 3 * 64;  // 16-byte A fragments of one group
    constexpr uint32_t CB = sizeof(cx<float>);
    static_assert(M == KB + 1 && 16 * MB + VR == M && VR <= 1, "square split family");
    __shared__ __align__(16) bf16x8 sa[AF];
    __shared__ float stail[16 * MB];
    __shared__ float srow[VR > 0 ? KP : 1];
    const int tid = static_cast<int>(threadIdx.x), lane = tid & 63, wave = tid >> 6;
    const int j = lane & 15, q = lane >> 4;
    const PmBlock blk = pm_block_record(blocks, static_cast<int>(blockIdx.x));
    const int F = g.F;
    const int i0 = static_cast<int>(blockIdx.y) * tpw;  // first tile e*F + f of this workgroup
    const int all = bt.count * F;
    const bool active = wave < blk.ranges;  // uniform per wavefront
    const int n = active ? (all - i0 < tpw ? all - i0 : tpw) : 0;  // tiles of this wavefront

    // 1. the group's operand (requested first: the barrier below waits for
    //    it, and the B blocks requested after it stay in flight)
    constexpr int PER = (AF + 64 * W - 1) / (64 * W);
    bf16x8 av[PER];
    float tv = 0.f, rv0 = 0.f, rv1 = 0.f;
    {
        const bf16x8* src = reinterpret_cast<const bf16x8*>(g.a3) + static_cast<size_t>(blk.group) * AF;
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int at = tid + 64 * W * i;
            if (at < AF) av[i] = src[at];
        }
        if (tid < 16 * MB) tv = g.atail[static_cast<size_t>(blk.group) * 16 * MB + tid];
        if constexpr (VR > 0) {
            const float* ar = g.arow + static_cast<size_t>(blk.group) * KP;
            if (tid < KP) rv0 = ar[tid];
            if (tid + 64 * W < KP) rv1 = ar[tid + 64 * W];
        }
    }

    // 2. the B blocks of this wavefront's first two tiles
    const uint32_t P = static_cast<uint32_t>(g.pm_planes);
    const GemmTile rng = tile_record(g.tiles, blk.range_begin + (active ? wave : 0));
    const uint32_t m = static_cast<uint32_t>(rng.col_begin + j);
    const bool jok = j < rng.ncols;
    struct BBlock {
        cx<float> v[NK][8];
        cx<float> t;  // k = KB
    };
    auto load_tile = [&](int i, BBlock& b) {
        const bool on = jok && i < n;
        // the resource must be wave-uniform: evaluation from the uniform i only
        const int iu = i < n ? i : 0;
        const int e = (i0 + iu) / F, f = i0 + iu - e * F;
        const __amdgpu_buffer_rsrc_t rin = buffer_rsrc(g.in + bt.in_stride * e, static_cast<uint32_t>(g.pm_np) * P * CB);
        const uint32_t col = 16u * P * (m >> 4) + 16u * static_cast<uint32_t>(g.pm_nz * f) + (m & 15u);
#pragma unroll
        for (int t = 0; t < NK; ++t)
#pragma unroll
            for (int e8 = 0; e8 < 8; ++e8) {
                const uint32_t z = static_cast<uint32_t>(32 * t + 8 * q + e8);
                b.v[t][e8] = buffer_load_cx<float>(rin, on ? (col + 16u * z) * CB : kBufferPast);
            }
        b.t = buffer_load_cx<float>(rin, on ? (col + 16u * KB) * CB : kBufferPast);
    };
    BBlock cur, nxt;
    load_tile(0, cur);

    // 3. stage the operand
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int at = tid + 64 * W * i;
        if (at < AF) sa[at] = av[i];
    }
    if (tid < 16 * MB) stail[tid] = tv;
    if constexpr (VR > 0) {
        if (tid < KP) srow[tid] = rv0;
        if (tid + 64 * W < KP) srow[tid + 64 * W] = rv1;
    }
    lds_barrier();
    if (!active) return;

    // conjugation back to the caller's values: bit j of this range's modes
    using ConstWord = const __attribute__((address_space(4))) uint32_t;
    const uint32_t cword = *((ConstWord*)g.pm_cmask + (rng.col_begin >> 4));
    const bool cj = (cword >> ((rng.col_begin & 15) + j)) & 1u;

    // tile i of this wavefront from block b
    auto process = [&](int i, const BBlock& b) {
        const int e = (i0 + i) / F, f = i0 + i - e * F;
        // the range's run of evaluation e: mode j, field f, row r at (F j + f) M + r
        cx<float>* run = bt.out[e] + static_cast<size_t>(g.out_ld) * F * rng.col_begin;
        const __amdgpu_buffer_rsrc_t rout = buffer_rsrc(run, static_cast<uint32_t>(rng.ncols * F * M) * CB);
        const uint32_t ocol = static_cast<uint32_t>((F * j + f) * M) * CB;
        // row M - 1 on the vector ALU: this lane's k = 32 t + 8 q + e, then
        // the sum over q
        if constexpr (VR > 0) {
            cx<float> v{0.f, 0.f};
#pragma unroll
            for (int t = 0; t < NK; ++t)
#pragma unroll
                for (int e8 = 0; e8 < 8; ++e8) {
                    const float a = srow[32 * t + 8 * q + e8];
                    v.x += a * b.v[t][e8].x;
                    v.y += a * b.v[t][e8].y;
                }
            if (q == 0) {
                v.x += srow[KB] * b.t.x;
                v.y += srow[KB] * b.t.y;
            }
            v.x += __shfl_xor(v.x, 16);
            v.y += __shfl_xor(v.y, 16);
            v.x += __shfl_xor(v.x, 32);
            v.y += __shfl_xor(v.y, 32);
            if (cj) v.y = -v.y;
            buffer_store_cx<float>(v, rout, q == 0 && jok ? ocol + static_cast<uint32_t>(16 * MB) * CB : kBufferPast);
        }
        // the product transposed on the matrix cores, B^T A^T: the data
        // fragment is the A operand and the matrix fragment the B operand, in
        // the registers they occupy as B and A, so lane (n, q) ends with rows
        // 16 mb + n of modes 4 q + e (e = 0..3).  For each k block t, B split
        // into bf16 planes [re/im][hi, mid, lo], then the products of every
        // row block.
        floatx4 r[MB], im[MB];
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) r[mb] = im[mb] = floatx4{0, 0, 0, 0};
#pragma unroll
        for (int t = 0; t < NK; ++t) {
            bf16x8 bb[2][3];
            {
                uint32_t w[2][3][4];
#pragma unroll
                for (int e4 = 0; e4 < 4; ++e4) {
                    split3_pair(b.v[t][2 * e4].x, b.v[t][2 * e4 + 1].x, w[0][0][e4], w[0][1][e4], w[0][2][e4]);
                    split3_pair(b.v[t][2 * e4].y, b.v[t][2 * e4 + 1].y, w[1][0][e4], w[1][1][e4], w[1][2][e4]);
                }
#pragma unroll
                for (int c = 0; c < 2; ++c)
#pragma unroll
                    for (int p = 0; p < 3; ++p) {
                        typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
                        const u32x4 u{w[c][p][0], w[c][p][1], w[c][p][2], w[c][p][3]};
                        bb[c][p] = __builtin_bit_cast(bf16x8, u);
                    }
            }
#pragma unroll
            for (int mb = 0; mb < MB; ++mb) {
                const bf16x8* A = &sa[((mb * NK + t) * 3) * 64 + lane];
                const bf16x8 a0 = A[0], a1 = A[64], a2 = A[128];
                r[mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bb[0][0], a2, r[mb], 0, 0, 0);
                im[mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bb[1][0], a2, im[mb], 0, 0, 0);
                r[mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bb[0][1], a1, r[mb], 0, 0, 0);
                im[mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bb[1][1], a1, im[mb], 0, 0, 0);
                r[mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bb[0][2], a0, r[mb], 0, 0, 0);
                im[mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bb[1][2], a0, im[mb], 0, 0, 0);
                r[mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bb[0][0], a1, r[mb], 0, 0, 0);
                im[mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bb[1][0], a1, im[mb], 0, 0, 0);
                r[mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bb[0][1], a0, r[mb], 0, 0, 0);
                im[mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bb[1][1], a0, im[mb], 0, 0, 0);
                r[mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bb[0][0], a0, r[mb], 0, 0, 0);
                im[mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bb[1][0], a0, im[mb], 0, 0, 0);
            }
        }
        // k = KB of modes 4 q + e, from the lanes that loaded them (mode = lane % 16)
        cx<float> bt[4];
#pragma unroll
        for (int e4 = 0; e4 < 4; ++e4) {
            bt[e4].x = __shfl(b.t.x, 4 * q + e4);
            bt[e4].y = __shfl(b.t.y, 4 * q + e4);
        }
        const uint32_t cbits = cword >> (rng.col_begin & 15);
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) {
            const float at = stail[16 * mb + j];
#pragma unroll
            for (int e4 = 0; e4 < 4; ++e4) {
                r[mb][e4] += at * bt[e4].x;
                im[mb][e4] += at * bt[e4].y;
                const int md = 4 * q + e4;  // mode of this element; row 16 mb + j
                const float y = (cbits >> md) & 1u ? -im[mb][e4] : im[mb][e4];
                const uint32_t o = static_cast<uint32_t>((F * md + f) * M + 16 * mb + j) * CB;
                raw_buffer_store_f2(bf2{r[mb][e4], y}, rout, static_cast<int>(md < rng.ncols ? o : kBufferPast), 0, 0);
            }
        }
    };
    for (int i = 0; i < n; ++i) {  // uniform per wavefront
        if (i + 1 < n) load_tile(i + 1, nxt);  // in flight while tile i is multiplied
        process(i, cur);
        cur = nxt;
    }
}

// Modal round trip of a K^2 group's columns in one kernel:
//     out[:, c] = I_g * (w[:, c] .* (F_g * in[:, c]))
// with F_g (Nj x Nz) the forward and I_g (Nz x Nj) the inverse matrix of the
// column's group and w the real diagonal modal weights (layout j + Nj*c, the
// modal layout).  The tile's 16 x Nj modal block never leaves the
// workgroup: the forward product is scaled in registers and written to LDS
// as the B block of the inverse product, so HBM sees the physical columns
// once in and once out.  `out` may equal `in`: a tile reads all of its
// columns into LDS before it writes any of them, and tiles own disjoint
// columns.
template <typename R>
struct RoundTripArgs {
    const cx<R>* in;
    cx<R>* out;
    const R* w;
    const R *ap1, *arow1;  // forward operand (M = Nj, K = Nz)
    const R *ap2, *arow2;  // inverse operand (M = Nz, K = Nj)
    const GemmTile* tiles;
    int ntiles, nz, nj, mb1, vr1, mb2, vr2;
};

template <typename R, int KS1, int KS2>
__global__ __launch_bounds__(256) void modal_round_trip_kernel(const RoundTripArgs<R> g) {
    using A4 = typename Acc<R>::type;
    constexpr int KP1 = 4 * KS1, KP2 = 4 * KS2;
    constexpr int CS1 = gemm_cs<R, KS1>(), CS2 = gemm_cs<R, KS2>();
    constexpr int S1 = gemm_slots<KS1>(), S2 = gemm_slots<KS2>();
    constexpr int NL = (16 * KP1 + 255) / 256;
    __shared__ __align__(16) cx<R> sb1[16 * CS1];  // forward B block, then the 16 x Nz output stage
    __shared__ __align__(16) cx<R> sb2[16 * CS2];  // weighted modal tile = inverse B block
    __shared__ R sar1[4 * KP1], sar2[4 * KP2];     // vector-row coefficients (VR <= 4)
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int j = lane & 15, q = lane >> 4;
    const bool one = gridDim.x >= static_cast<unsigned>(g.ntiles);
    const int step = one ? g.ntiles : static_cast<int>(gridDim.x);
    for (int tile = one ? xcd_item(blockIdx.x, g.ntiles) : static_cast<int>(blockIdx.x); tile < g.ntiles;
         tile += step) {
        const GemmTile t = tile_record(g.tiles, tile);
        const size_t grp = static_cast<size_t>(t.group);
        R a0[KS1];
        if (wave < g.mb1) {
            const R* a_ptr = g.ap1 + (grp * g.mb1 + wave) * KS1 * 64 + lane;
#pragma unroll
            for (int s = 0; s < KS1; ++s) a0[s] = a_ptr[s * 64];
        }
        // weights of this lane's forward rows; zero outside Nj and the tile
        const bool col_ok = j < t.ncols;
        const R* w_col = g.w + static_cast<size_t>(g.nj) * (t.col_begin + (col_ok ? j : 0));
        R wt[S1][4];
#pragma unroll
        for (int sl = 0; sl < S1; ++sl) {
            const int mb = wave + 4 * sl;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * mb + Acc<R>::row(lane, r);
                const bool ok = col_ok && mb < g.mb1 && row < g.nj;
                const R v = w_col[ok ? row : 0];
                wt[sl][r] = ok ? v : R(0);
            }
        }
        R wv = R(0);
        if (wave < g.vr1 && col_ok) wv = w_col[16 * g.mb1 + wave];
        for (int e = tid; e < g.vr1 * KP1; e += 256) sar1[e] = g.arow1[grp * g.vr1 * KP1 + e];
        for (int e = tid; e < g.vr2 * KP2; e += 256) sar2[e] = g.arow2[grp * g.vr2 * KP2 + e];
        // forward B block -> LDS, zero outside the tile's columns and for k >= Nz
        cx<R> bv[NL];
#pragma unroll
        for (int n = 0; n < NL; ++n) {
            const int e = tid + 256 * n, c = e / KP1, z = e - c * KP1;
            bv[n] = cx<R>{R(0), R(0)};
            if (e < 16 * KP1 && c < t.ncols && z < g.nz)
                bv[n] = g.in[static_cast<size_t>(g.nz) * (t.col_begin + c) + z];
        }
        // inverse A fragments of this wavefront's first row block
        R b0[KS2];
        if (wave < g.mb2) {
            const R* a_ptr = g.ap2 + (grp * g.mb2 + wave) * KS2 * 64 + lane;
#pragma unroll
            for (int s = 0; s < KS2; ++s) b0[s] = a_ptr[s * 64];
        }
#pragma unroll
        for (int n = 0; n < NL; ++n) {
            const int e = tid + 256 * n, c = e / KP1, z = e - c * KP1;
            if (e < 16 * KP1) sb1[c * CS1 + z] = bv[n];
        }
        __syncthreads();
        // forward product, scaled by the weights into the inverse B block
#pragma unroll
        for (int sl = 0; sl < S1; ++sl) {
            const int mb = wave + 4 * sl;
            if (mb < g.mb1) {
                R a[KS1];
                if (sl == 0) {
#pragma unroll
                    for (int s = 0; s < KS1; ++s) a[s] = a0[s];
                } else {
                    const R* a_ptr = g.ap1 + (grp * g.mb1 + mb) * KS1 * 64 + lane;
#pragma unroll
                    for (int s = 0; s < KS1; ++s) a[s] = a_ptr[s * 64];
                }
                A4 re = {0, 0, 0, 0}, im = {0, 0, 0, 0};
#pragma unroll
                for (int s = 0; s < KS1; ++s) {
                    const cx<R> b = sb1[j * CS1 + 4 * s + q];
                    re = Acc<R>::mfma(a[s], b.x, re);
                    im = Acc<R>::mfma(a[s], b.y, im);
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = 16 * mb + Acc<R>::row(lane, r);
                    if (row < KP2) sb2[j * CS2 + row] = cx<R>{re[r] * wt[sl][r], im[r] * wt[sl][r]};
                }
            }
        }
        if (wave < g.vr1) {
            cx<R> v{R(0), R(0)};
#pragma unroll
            for (int s = 0; s < KS1; ++s) {
                const R a = sar1[wave * KP1 + q * KS1 + s];
                const cx<R> b = sb1[j * CS1 + q * KS1 + s];
                v.x += a * b.x;
                v.y += a * b.y;
            }
            v.x += __shfl_xor(v.x, 16);
            v.y += __shfl_xor(v.y, 16);
            v.x += __shfl_xor(v.x, 32);
            v.y += __shfl_xor(v.y, 32);
            if (q == 0) sb2[j * CS2 + 16 * g.mb1 + wave] = cx<R>{v.x * wv, v.y * wv};
        }
        {
            // rows the forward product does not cover, up to the inverse K padding
            const int covered = g.vr1 > 0 ? g.nj : 16 * g.mb1;
            const int gap = KP2 > covered ? KP2 - covered : 0;
            for (int e = tid; e < 16 * gap; e += 256) sb2[(e / gap) * CS2 + covered + e % gap] = cx<R>{R(0), R(0)};
        }
        __syncthreads();
        // inverse product; the result is staged in the forward B block's LDS
        cx<R>* st = sb1;
#pragma unroll
        for (int sl = 0; sl < S2; ++sl) {
            const int mb = wave + 4 * sl;
            if (mb < g.mb2) {
                R a[KS2];
                if (sl == 0) {
#pragma unroll
                    for (int s = 0; s < KS2; ++s) a[s] = b0[s];
                } else {
                    const R* a_ptr = g.ap2 + (grp * g.mb2 + mb) * KS2 * 64 + lane;
#pragma unroll
                    for (int s = 0; s < KS2; ++s) a[s] = a_ptr[s * 64];
                }
                A4 re = {0, 0, 0, 0}, im = {0, 0, 0, 0};
#pragma unroll
                for (int s = 0; s < KS2; ++s) {
                    const cx<R> b = sb2[j * CS2 + 4 * s + q];
                    re = Acc<R>::mfma(a[s], b.x, re);
                    im = Acc<R>::mfma(a[s], b.y, im);
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = 16 * mb + Acc<R>::row(lane, r);
                    if (row < g.nz) st[j * g.nz + row] = cx<R>{re[r], im[r]};
                }
            }
        }
        if (wave < g.vr2) {
            cx<R> v{R(0), R(0)};
#pragma unroll
            for (int s = 0; s < KS2; ++s) {
                const R a = sar2[wave * KP2 + q * KS2 + s];
                const cx<R> b = sb2[j * CS2 + q * KS2 + s];
                v.x += a * b.x;
                v.y += a * b.y;
            }
            v.x += __shfl_xor(v.x, 16);
            v.y += __shfl_xor(v.y, 16);
            v.x += __shfl_xor(v.x, 32);
            v.y += __shfl_xor(v.y, 32);
            if (q == 0) st[j * g.nz + 16 * g.mb2 + wave] = v;
        }
        __syncthreads();
        const int total = g.nz * t.ncols;
        for (int e = tid; e < total; e += 256) g.out[static_cast<size_t>(g.nz) * t.col_begin + e] = st[e];
        __syncthreads();  // LDS is reused by the next tile
    }
}

// applySyntheticModalWork{Interleaved,Split,Wvm} (src/model.cpp:269-308):
// out = w * in, one real weight per complex modal element.
template <typename R>
__global__ void modal_work_kernel(const R* __restrict__ w, const cx<R>* __restrict__ in, cx<R>* __restrict__ out,
                                  size_t n) {
    for (size_t e = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; e < n;
         e += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const R s = w[e];
        const cx<R> v = in[e];
        out[e] = cx<R>{s * v.x, s * v.y};
    }
}

template <typename R>
__global__ void modal_work_split_kernel(const R* __restrict__ w, const R* __restrict__ in_re,
                                        const R* __restrict__ in_im, R* __restrict__ out_re, R* __restrict__ out_im,
                                        size_t n) {
    for (size_t e = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; e < n;
         e += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const R s = w[e];
        out_re[e] = s * in_re[e];
        out_im[e] = s * in_im[e];
    }
}

// Full WVM modal arrays: element e = r + block*m (r = j + Nj*field, block =
// Nj*F) of the logical modal order is read and written at frequency q[m];
// the k_x = 0 partner frequency receives the conjugate.
template <typename R>
__global__ void modal_work_wvm_kernel(const R* __restrict__ w, const cx<R>* __restrict__ in, cx<R>* __restrict__ out,
                                      const int* __restrict__ qmap, const int* __restrict__ partner, int block,
                                      size_t n) {
    for (size_t e = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; e < n;
         e += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const size_t m = e / block, r = e - m * block;
        const size_t at = static_cast<size_t>(block) * qmap[m] + r;
        const R s = w[e];
        const cx<R> v = in[at];
        const cx<R> o{s * v.x, s * v.y};
        out[at] = o;
        const int p = partner[m];
        if (p >= 0) out[static_cast<size_t>(block) * p + r] = conjg(o);
    }
}

template <typename R>
__global__ void fill_zero_kernel(cx<R>* __restrict__ out, size_t total) {
    for (size_t e = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; e < total;
         e += static_cast<size_t>(gridDim.x) * blockDim.x)
        out[e] = cx<R>{R(0), R(0)};
}

// Split of M output rows into MB MFMA row blocks and VR <= 4 vector rows.
struct RowSplit {
    int mb, vr;
};
inline RowSplit row_split(int M) {
    if (M >= 16 && M % 16 <= 4) return RowSplit{M / 16, M % 16};
    return RowSplit{(M + 15) / 16, 0};
}

// Host: matrices (row-major M x K per group) -> fragment order and vector rows.
template <typename R>
void operand_order(const double* mats, uint64_t groups, int M, int K, std::vector<R>& ap, std::vector<R>& arow) {
    const int KS = (K + 3) / 4, KP = 4 * KS;
    const RowSplit sp = row_split(M);
    ap.assign(static_cast<size_t>(groups) * sp.mb * KS * 64, R(0));
    arow.assign(static_cast<size_t>(groups) * std::max(sp.vr, 1) * KP, R(0));
    for (uint64_t g = 0; g < groups; ++g) {
        for (int mb = 0; mb < sp.mb; ++mb)
            for (int s = 0; s < KS; ++s)
                for (int l = 0; l < 64; ++l) {
                    const int i = 16 * mb + (l & 15), k = 4 * s + (l >> 4);
                    if (i < M && k < K)
                        ap[((g * sp.mb + mb) * KS + s) * 64 + l] =
                            static_cast<R>(mats[(g * M + i) * static_cast<uint64_t>(K) + k]);
                }
        for (int r = 0; r < sp.vr; ++r)
            for (int k = 0; k < K; ++k)
                arow[(g * sp.vr + r) * KP + k] =
                    static_cast<R>(mats[(g * M + 16 * sp.mb + r) * static_cast<uint64_t>(K) + k]);
    }
}

namespace {

// Workgroups of the fixed-height streaming kernel when the plan does not set
// a count: three per CU of the 256-CU part.
constexpr int kStreamGrid = 768;
// Evaluations one wavefront of the batched plane-major product takes (its
// chunk range's tiles of that many evaluations, two always in flight).
constexpr int kPmEvalsPerWave = 1;

template <typename R, int KS, int M>
void launch_gemm_fixed(const GemmArgs<R>& a, int req, hipStream_t s) {
    const int grid = std::min(a.ntiles, req > 0 ? req : kStreamGrid);
    hipLaunchKernelGGL((grouped_gemm_fixed_kernel<R, KS, M>), dim3(grid), dim3(256), 0, s, a);
    check_hip(hipGetLastError(), "vertical GEMM launch");
}

// Output heights of the fixed-height kernel: the (K, M) pairs of the vertical
// families at Nz = 33, 65, 129 (M = K = Nz for square families, Nj x Nz
// forward and Nz x Nj inverse for the truncated ones); FP64 up to Nz = 65.
template <int NK, int M, bool PM = false>
void launch_gemm_split(const GemmArgs<float>& a, int req, hipStream_t s, const GemmBatch* batch = nullptr) {
    GemmBatch b{};
    b.count = 1;
    b.out[0] = a.out;
    if (batch) b = *batch;
    const int grid = std::min(a.ntiles * b.count, req > 0 ? req : kStreamGrid);
    hipLaunchKernelGGL((grouped_gemm_split_kernel<NK, M, PM>), dim3(grid), dim3(256), 0, s, a, b);
    check_hip(hipGetLastError(), "vertical GEMM launch");
}

template <typename R>
bool launch_gemm_fixed_any(const GemmArgs<R>& a, int req, hipStream_t s) {
    const int ks = (a.K + 3) / 4;
    if constexpr (sizeof(R) == 4) {
        if (a.a3) {
            // split-BF16 products (the plan built the split operand)
            if (a.K == 65 && a.M == 65) return launch_gemm_split<2, 65>(a, req, s), true;
            if (a.K == 33 && a.M == 33) return launch_gemm_split<1, 33>(a, req, s), true;
            if (a.K == 129 && a.M == 129) return launch_gemm_split<4, 129>(a, req, s), true;
        }
    }
#define SKB_FX(KSV, MV)                                  \
    if (ks == KSV && a.M == MV) {                        \
        launch_gemm_fixed<R, KSV, MV>(a, req, s);        \
        return true;                                     \
    }
    SKB_FX(9, 33) SKB_FX(9, 21) SKB_FX(6, 33) SKB_FX(17, 65) SKB_FX(17, 42) SKB_FX(11, 65)
    if constexpr (sizeof(R) == 4) {
        SKB_FX(33, 129) SKB_FX(33, 85) SKB_FX(22, 129)
    }
#undef SKB_FX
    return false;
}

template <typename R, int KS, bool WVM>
void launch_gemm_ks(const GemmArgs<R>& a, int req, hipStream_t s) {
    constexpr int CS = gemm_cs<R, KS>();
    if (!WVM && launch_gemm_fixed_any<R>(a, req, s)) return;
    const size_t lds = static_cast<size_t>(std::max(16 * CS, 16 * a.M)) * sizeof(cx<R>);
    const int grid = req > 0 && req < a.ntiles ? req : xcd_grid(a.ntiles);
    hipLaunchKernelGGL((grouped_gemm_kernel<R, KS, WVM>), dim3(grid), dim3(256), lds, s, a);
    check_hip(hipGetLastError(), "vertical GEMM launch");
}

template <typename R, bool WVM>
void launch_gemm(const GemmArgs<R>& a, int req, hipStream_t s) {
    if (a.ntiles == 0) return;
    switch ((a.K + 3) / 4) {
#define SKB_KS(N) \
    case N: launch_gemm_ks<R, N, WVM>(a, req, s); return;
        SKB_KS(1) SKB_KS(2) SKB_KS(3) SKB_KS(4) SKB_KS(5) SKB_KS(6) SKB_KS(7) SKB_KS(8) SKB_KS(9) SKB_KS(10)
        SKB_KS(11) SKB_KS(12) SKB_KS(13) SKB_KS(14) SKB_KS(15) SKB_KS(16) SKB_KS(17) SKB_KS(22) SKB_KS(33)
#undef SKB_KS
    }
    if constexpr (sizeof(R) == 4) {
        switch ((a.K + 3) / 4) {
            case 43: launch_gemm_ks<R, 43, WVM>(a, req, s); return;
            case 65: launch_gemm_ks<R, 65, WVM>(a, req, s); return;
        }
    }
    fail(SKB_UNSUPPORTED, "vertical GEMM inner dimension not instantiated");
}

// Round-trip kernel instantiations, (KS1, KS2) = (ceil(Nz/4), ceil(Nj/4)) for
// Nj = 2(Nz-1)/3: Nz 5-8, 9-12, 17-20, 33-36, 65-68, 129-132 and (FP32)
// 257-260.  Other sizes run the three-launch form (forward GEMM, weights,
// inverse GEMM).
template <typename R, int KS1, int KS2>
void launch_round_trip_ks(const RoundTripArgs<R>& a, int req, hipStream_t s) {
    const int grid = req > 0 && req < a.ntiles ? req : xcd_grid(a.ntiles);
    hipLaunchKernelGGL((modal_round_trip_kernel<R, KS1, KS2>), dim3(grid), dim3(256), 0, s, a);
    check_hip(hipGetLastError(), "modal round-trip launch");
}

template <typename R>
bool launch_round_trip(const RoundTripArgs<R>& a, int req, hipStream_t s, bool dry_run) {
    const int ks1 = (a.nz + 3) / 4, ks2 = (a.nj + 3) / 4;
#define SKB_RT(A, B)                                                  \
    if (ks1 == A && ks2 == B) {                                       \
        if (!dry_run && a.ntiles > 0) launch_round_trip_ks<R, A, B>(a, req, s); \
        return true;                                                  \
    }
    SKB_RT(2, 1) SKB_RT(3, 2) SKB_RT(5, 3) SKB_RT(9, 6) SKB_RT(17, 11) SKB_RT(33, 22)
    if constexpr (sizeof(R) == 4) {
        SKB_RT(65, 43)
    }
#undef SKB_RT
    return false;
}

}  // namespace

// Matrices of one direction on the device.
// bf16 with round-to-nearest-even, and back.
inline uint16_t bf16_bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) return static_cast<uint16_t>((u >> 16) | 0x40u);  // NaN stays NaN
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}
inline double bf16_value(uint16_t h) {
    const uint32_t u = static_cast<uint32_t>(h) << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// Split-BF16 operand: A = A0 + A1 + A2 (each bf16; the residuals are taken
// in FP64, so A0 + A1 + A2 agrees with A to about 2^-27 relative) in the
// v_mfma_f32_16x16x32_bf16 fragment order, and the last column A[:, K-1] in
// FP32.  Only for K = 32*NK + 1 and M = 16*MB + r with r <= 4.
bool split_bf16_shape(int M, int K) { return K % 32 == 1 && K > 1 && M >= 16 && M % 16 <= 4; }
void split_order(const double* mats, uint64_t groups, int M, int K, std::vector<uint16_t>& a3,
                 std::vector<float>& atail) {
    const int MB = M / 16, NK = K / 32;
    a3.assign(static_cast<size_t>(groups) * MB * NK * 3 * 64 * 8, 0);
    atail.assign(static_cast<size_t>(groups) * 16 * MB, 0.f);
    for (uint64_t g = 0; g < groups; ++g)
        for (int mb = 0; mb < MB; ++mb) {
            for (int t = 0; t < NK; ++t)
                for (int l = 0; l < 64; ++l)
                    for (int e = 0; e < 8; ++e) {
                        const int i = 16 * mb + (l & 15), k = 32 * t + 8 * (l >> 4) + e;
                        const double v = mats[(g * M + i) * static_cast<uint64_t>(K) + k];
                        const uint16_t h = bf16_bits(static_cast<float>(v));
                        const double r1 = v - bf16_value(h);
                        const uint16_t m = bf16_bits(static_cast<float>(r1));
                        const uint16_t lo = bf16_bits(static_cast<float>(r1 - bf16_value(m)));
                        const size_t at = ((((g * MB + mb) * NK + t) * 3) * 64 + l) * 8 + e;
                        a3[at] = h;
                        a3[at + 64 * 8] = m;
                        a3[at + 2 * 64 * 8] = lo;
                    }
            for (int r = 0; r < 16; ++r)
                atail[g * 16 * MB + 16 * mb + r] =
                    static_cast<float>(mats[(g * M + 16 * mb + r) * static_cast<uint64_t>(K) + K - 1]);
        }
}

struct GemmOperand {
    DeviceBuffer ap, arow;
    DeviceBuffer a3, atail;  // split-BF16 form (FP32 plans of a split shape)
    int M = 0, K = 0;
    RowSplit split{0, 0};
    void build(const double* mats, uint64_t groups, int M_, int K_, int precision) {
        M = M_;
        K = K_;
        split = row_split(M);
        if (precision == SKB_F32 && split_bf16_shape(M, K)) {
            std::vector<uint16_t> h;
            std::vector<float> t;
            split_order(mats, groups, M, K, h, t);
            a3.upload(h.data(), h.size() * sizeof(uint16_t));
            atail.upload(t.data(), t.size() * sizeof(float));
        }
        if (precision == SKB_F64) {
            std::vector<double> a, r;
            operand_order<double>(mats, groups, M, K, a, r);
            ap.upload(a.data(), a.size() * sizeof(double));
            arow.upload(r.data(), r.size() * sizeof(double));
        } else {
            std::vector<float> a, r;
            operand_order<float>(mats, groups, M, K, a, r);
            ap.upload(a.data(), a.size() * sizeof(float));
            arow.upload(r.data(), r.size() * sizeof(float));
        }
    }
};

}  // namespace skb

struct skb_vertical_plan {
    skb_workload w{};
    uint64_t nj = 0;
    int precision = SKB_F64;
    uint64_t groups = 0;
    int ntiles = 0;
    int grid = 0;  // workgroups per launch (0 = one per tile)
    int products = SKB_PRODUCTS_SPLIT_BF16;
    uint64_t matrix_bytes = 0;  // logical matrix bytes per direction
    skb::GemmOperand fwd, inv;
    skb::DeviceBuffer tiles;
    // WVM-direct plans: per-mode frequency and k_x = 0 partner frequency
    bool wvm = false;
    uint64_t modes = 0;
    skb::DeviceBuffer qmap, partner;
    // Tiles over a plane-major input: up to 16 consecutive modes of one group
    // and one field (GemmTile::pad = field), group by group.
    int pm_ntiles = 0;
    skb::DeviceBuffer pm_tiles;
    // batched plane-major product (Nz 33 and 65): the chunk ranges (group,
    // first mode, modes) and the workgroup blocks over them
    int pm_nranges = 0, pm_nblocks = 0;
    skb::DeviceBuffer pm_ranges, pm_blocks;
    // interleaved staging of the split-layout entry points (reserved by
    // skb_vertical_plan_reserve_split): physical-size and modal-size arrays
    skb::DeviceBuffer split_physical, split_modal;
};

using namespace skb;

namespace {

void validate_groups(const skb_mode_group* groups, uint64_t group_count, uint64_t* modes_out) {
    if (!groups || group_count == 0) fail(SKB_INVALID_ARGUMENT, "vertical GEMM requires at least one group");
    // Groups must tile [0, modes) contiguously (src/model.cpp:191-206).
    uint64_t next = 0;
    for (uint64_t g = 0; g < group_count; ++g) {
        if (groups[g].first_mode != next || groups[g].mode_count == 0)
            fail(SKB_INVALID_ARGUMENT, "vertical groups must be contiguous, non-empty and start at mode 0");
        next += groups[g].mode_count;
    }
    *modes_out = next;
}

std::unique_ptr<skb_vertical_plan> make_plan(const skb_workload* workload, uint64_t nj, const skb_mode_group* groups,
                                             uint64_t group_count, const double* forward, const double* inverse,
                                             int precision) {
    check_precision(precision);
    validate_workload(workload);
    if (!forward || !inverse) fail(SKB_INVALID_ARGUMENT, "vertical GEMM requires forward and inverse matrices");
    const uint64_t nz = workload->nz;
    if (nj == 0 || nz == 0) fail(SKB_INVALID_ARGUMENT, "Vertical operator dimensions must be positive.");
    const bool fp64 = precision == SKB_F64;
    const uint64_t limit = fp64 ? 132 : 260;
    if (nz > limit || nj > limit)
        fail(SKB_UNSUPPORTED, fp64 ? "FP64 vertical dimensions (Nz, Nj) above 132 are not instantiated"
                                   : "FP32 vertical dimensions (Nz, Nj) above 260 are not instantiated");
    // both directions: K in the instantiated set, row blocks within the slots
    for (const auto& mk : {std::pair<uint64_t, uint64_t>{nj, nz}, std::pair<uint64_t, uint64_t>{nz, nj}}) {
        const int M = static_cast<int>(mk.first), K = static_cast<int>(mk.second), ks = (K + 3) / 4;
        if (!gemm_ks_instantiated(ks, fp64))
            fail(SKB_UNSUPPORTED, "vertical GEMM inner dimension " + std::to_string(K) + " is not instantiated");
        const int slots = ks >= 22 ? 4 : 3;
        if ((row_split(M).mb + 3) / 4 > slots)
            fail(SKB_UNSUPPORTED, "vertical GEMM output dimension " + std::to_string(M) + " with inner dimension " +
                                      std::to_string(K) + " is not instantiated");
    }
    uint64_t modes = 0;
    validate_groups(groups, group_count, &modes);
    const uint64_t fields = workload->fields;
    if (modes * fields * std::max(nz, nj) > (1ull << 31)) fail(SKB_UNSUPPORTED, "too many vertical columns");
    std::vector<GemmTile> tiles;
    for (uint64_t g = 0; g < group_count; ++g) {
        const uint64_t c0 = groups[g].first_mode * fields;
        const uint64_t nc = groups[g].mode_count * fields;
        for (uint64_t c = 0; c < nc; c += 16)
            tiles.push_back(GemmTile{static_cast<int>(g), static_cast<int>(c0 + c),
                                     static_cast<int>(std::min<uint64_t>(16, nc - c)), 0});
    }
    auto p = std::make_unique<skb_vertical_plan>();
    p->w = *workload;
    p->nj = nj;
    p->precision = precision;
    p->groups = group_count;
    p->modes = modes;
    p->ntiles = static_cast<int>(tiles.size());
    const int inj = static_cast<int>(nj), inz = static_cast<int>(nz);
    p->fwd.build(forward, group_count, inj, inz, precision);
    p->inv.build(inverse, group_count, inz, inj, precision);
    p->tiles.upload(tiles.data(), tiles.size() * sizeof(GemmTile));
    // Plane-major tiles stay inside one 16-mode chunk (a group's first and
    // last tiles may be partial): a tile's 16 modes at one z are then one
    // 128-byte line that no other tile reads.
    std::vector<GemmTile> pm;
    for (uint64_t g = 0; g < group_count; ++g)
        for (uint64_t f = 0; f < fields; ++f) {
            const uint64_t first = groups[g].first_mode, end = first + groups[g].mode_count;
            for (uint64_t m = first; m < end; m = (m / 16 + 1) * 16)
                pm.push_back(GemmTile{static_cast<int>(g), static_cast<int>(m),
                                      static_cast<int>(std::min<uint64_t>((m / 16 + 1) * 16, end) - m),
                                      static_cast<int>(f)});
        }
    p->pm_ntiles = static_cast<int>(pm.size());
    p->pm_tiles.upload(pm.data(), pm.size() * sizeof(GemmTile));
    if (precision == SKB_F32 && nj == nz && (nz == 33 || nz == 65)) {
        std::vector<GemmTile> ranges;
        for (uint64_t g = 0; g < group_count; ++g) {
            const uint64_t first = groups[g].first_mode, end = first + groups[g].mode_count;
            for (uint64_t m = first; m < end; m = (m / 16 + 1) * 16)
                ranges.push_back(GemmTile{static_cast<int>(g), static_cast<int>(m),
                                          static_cast<int>(std::min<uint64_t>((m / 16 + 1) * 16, end) - m), 0});
        }
        std::vector<PmBlock> blocks;
        for (int r = 0; r < static_cast<int>(ranges.size());) {
            int k = 1;
            while (k < kPmWaves && r + k < static_cast<int>(ranges.size()) && ranges[r + k].group == ranges[r].group) ++k;
            blocks.push_back(PmBlock{ranges[r].group, r, k, 0});
            r += k;
        }
        p->pm_nranges = static_cast<int>(ranges.size());
        p->pm_ranges.upload(ranges.data(), ranges.size() * sizeof(GemmTile));
        p->pm_nblocks = static_cast<int>(blocks.size());
        p->pm_blocks.upload(blocks.data(), blocks.size() * sizeof(PmBlock));
    }
    p->matrix_bytes = nj * nz * group_count * (precision == SKB_F64 ? 8 : 4);
    return p;
}

template <typename R>
GemmArgs<R> args(const skb_vertical_plan* p, const GemmOperand& op, const void* in, void* out) {
    GemmArgs<R> a{};
    a.in = static_cast<const cx<R>*>(in);
    a.out = static_cast<cx<R>*>(out);
    a.ap = op.ap.as<R>();
    a.arow = op.arow.as<R>();
    a.tiles = p->tiles.as<GemmTile>();
    a.ntiles = p->ntiles;
    a.M = op.M;
    a.K = op.K;
    a.MB = op.split.mb;
    a.VR = op.split.vr;
    a.F = static_cast<int>(p->w.fields);
    a.in_ld = op.K;
    a.out_ld = op.M;
    a.qmap = p->qmap.as<int>();
    a.partner = p->partner.as<int>();
    const bool split = p->products == SKB_PRODUCTS_SPLIT_BF16 && op.a3.ptr;
    a.a3 = split ? op.a3.as<uint16_t>() : nullptr;
    a.atail = split ? op.atail.as<float>() : nullptr;
    return a;
}

void run(const skb_vertical_plan* p, bool forward, const void* in, void* out, void* stream, bool wvm) {
    if (!p || !in || !out) fail(SKB_INVALID_ARGUMENT, "null argument");
    if (p->wvm != wvm)
        fail(SKB_LOGIC_ERROR, wvm ? "WVM-direct entry point called on a compact vertical plan"
                                  : "compact entry point called on a WVM-direct vertical plan");
    const GemmOperand& op = forward ? p->fwd : p->inv;
    const hipStream_t s = as_stream(stream);
    if (p->precision == SKB_F64) {
        const auto a = args<double>(p, op, in, out);
        if (wvm) launch_gemm<double, true>(a, p->grid, s);
        else launch_gemm<double, false>(a, p->grid, s);
    } else {
        const auto a = args<float>(p, op, in, out);
        if (wvm) launch_gemm<float, true>(a, p->grid, s);
        else launch_gemm<float, false>(a, p->grid, s);
    }
}

}  // namespace

namespace skb {

// Forward product over a column subset of a wider compact array: input
// column c (c = field + fields*mode of the plan) is read at
// in + K*(in_cs*c + in_co); the output is the plan's dense layout.
// (GemmArgs carries the column stride in_ld = K*in_cs and an input pointer
// advanced by K*in_co.)
void vertical_forward_columns(const skb_vertical_plan* p, const void* in, int in_cs, int in_co, void* out,
                              void* stream) {
    if (!p || !in || !out) fail(SKB_INVALID_ARGUMENT, "null argument");
    if (p->wvm || in_cs < 1 || in_co < 0 || in_co >= in_cs) fail(SKB_INVALID_ARGUMENT, "bad column selection");
    const size_t shift = static_cast<size_t>(p->fwd.K) * in_co;
    const hipStream_t s = as_stream(stream);
    if (p->precision == SKB_F64) {
        auto a = args<double>(p, p->fwd, static_cast<const cx<double>*>(in) + shift, out);
        a.in_ld = p->fwd.K * in_cs;
        launch_gemm<double, false>(a, p->grid, s);
    } else {
        auto a = args<float>(p, p->fwd, static_cast<const cx<float>*>(in) + shift, out);
        a.in_ld = p->fwd.K * in_cs;
        launch_gemm<float, false>(a, p->grid, s);
    }
}

// Forward product from the plane-major layout (element (mode, z, field) at
// pm[mode + np * (z + Nz * field)]) into the plan's modal layout: the
// split-BF16 kernel of square FP32 plans (Nz = 33, 65, 129).
bool vertical_pm_supported(const skb_vertical_plan* p) {
    if (!p || p->wvm || p->precision != SKB_F32 || p->products != SKB_PRODUCTS_SPLIT_BF16 || !p->fwd.a3.ptr)
        return false;
    const int M = p->fwd.M, K = p->fwd.K;
    return M == K && (K == 33 || K == 65 || K == 129);
}

void vertical_forward_pm(const skb_vertical_plan* p, const void* pm, int np, const uint32_t* cmask, void* out,
                         void* stream) {
    // one evaluation: the persistent split kernel (its row 16 MB sums in the
    // batched kernel's order, so a batched evaluation equals this bit for bit)
    if (!vertical_pm_supported(p)) fail(SKB_UNSUPPORTED, "plane-major vertical product: unsupported plan");
    if (!pm || !out || !cmask || np < static_cast<int>(p->modes) || np % 16 != 0)
        fail(SKB_INVALID_ARGUMENT, "bad plane-major operand");
    auto a = args<float>(p, p->fwd, pm, out);
    a.tiles = p->pm_tiles.as<GemmTile>();
    a.ntiles = p->pm_ntiles;
    a.pm_np = np;
    a.pm_nz = static_cast<int>(p->w.nz);
    a.pm_planes = static_cast<int>(p->w.nz * p->w.fields);
    a.pm_cmask = cmask;
    const hipStream_t s = as_stream(stream);
    switch (a.K) {
        case 65: launch_gemm_split<2, 65, true>(a, p->grid, s); break;
        case 33: launch_gemm_split<1, 33, true>(a, p->grid, s); break;
        default: launch_gemm_split<4, 129, true>(a, p->grid, s); break;
    }
}

void vertical_forward_pm_batch(const skb_vertical_plan* p, const void* pm, size_t pm_stride, int np,
                               const uint32_t* cmask, const BatchOutputs& out, int count, void* stream) {
    if (!vertical_pm_supported(p)) fail(SKB_UNSUPPORTED, "plane-major vertical product: unsupported plan");
    if (count < 1 || count > kBatchMax) fail(SKB_INVALID_ARGUMENT, "batch of evaluations out of range");
    if (!pm || !cmask || np < static_cast<int>(p->modes) || np % 16 != 0)
        fail(SKB_INVALID_ARGUMENT, "bad plane-major operand");
    if (static_cast<long long>(p->pm_ntiles) * count >= (1ll << 31) || p->w.fields > 255)
        fail(SKB_UNSUPPORTED, "plane-major vertical product: batch too large");
    GemmBatch b{};
    b.count = count;
    b.in_stride = pm_stride;
    for (int e = 0; e < count; ++e) {
        if (!out.out[e]) fail(SKB_INVALID_ARGUMENT, "null modal output");
        b.out[e] = static_cast<cx<float>*>(out.out[e]);
    }
    auto a = args<float>(p, p->fwd, pm, b.out[0]);
    a.tiles = p->pm_tiles.as<GemmTile>();
    a.ntiles = p->pm_ntiles;
    a.pm_np = np;
    a.pm_nz = static_cast<int>(p->w.nz);
    a.pm_planes = static_cast<int>(p->w.nz * p->w.fields);
    a.pm_cmask = cmask;
    const hipStream_t s = as_stream(stream);
    if (p->pm_nblocks > 0 && p->grid == 0) {
        // one wavefront per chunk range, kPmEvalsPerWave evaluations each
        a.tiles = p->pm_ranges.as<GemmTile>();
        // tiles per wavefront: the fields of kPmEvalsPerWave evaluations
        const int F = static_cast<int>(p->w.fields), tpw = kPmEvalsPerWave * F;
        const dim3 grid(p->pm_nblocks, (count * F + tpw - 1) / tpw), block(64 * kPmWaves);
        if (a.K == 65)
            hipLaunchKernelGGL((gemm_pm_range_kernel<2, 65>), grid, block, 0, s, a, p->pm_blocks.as<PmBlock>(), b, tpw);
        else
            hipLaunchKernelGGL((gemm_pm_range_kernel<1, 33>), grid, block, 0, s, a, p->pm_blocks.as<PmBlock>(), b, tpw);
        check_hip(hipGetLastError(), "vertical GEMM launch");
        return;
    }
    switch (a.K) {
        case 65: launch_gemm_split<2, 65, true>(a, p->grid, s, &b); break;
        case 33: launch_gemm_split<1, 33, true>(a, p->grid, s, &b); break;
        default: launch_gemm_split<4, 129, true>(a, p->grid, s, &b); break;
    }
}

}  // namespace skb

namespace {

template <typename R>
RoundTripArgs<R> round_trip_args(const skb_vertical_plan* p, const void* weights, const void* in, void* out) {
    RoundTripArgs<R> a{};
    a.in = static_cast<const cx<R>*>(in);
    a.out = static_cast<cx<R>*>(out);
    a.w = static_cast<const R*>(weights);
    a.ap1 = p->fwd.ap.as<R>();
    a.arow1 = p->fwd.arow.as<R>();
    a.ap2 = p->inv.ap.as<R>();
    a.arow2 = p->inv.arow.as<R>();
    a.tiles = p->tiles.as<GemmTile>();
    a.ntiles = p->ntiles;
    a.nz = static_cast<int>(p->w.nz);
    a.nj = static_cast<int>(p->nj);
    a.mb1 = p->fwd.split.mb;
    a.vr1 = p->fwd.split.vr;
    a.mb2 = p->inv.split.mb;
    a.vr2 = p->inv.split.vr;
    return a;
}

bool round_trip(const skb_vertical_plan* p, const void* weights, const void* in, void* out, hipStream_t s,
                bool dry_run) {
    // row blocks must fit the accumulator slots of each phase
    const int ks1 = static_cast<int>((p->w.nz + 3) / 4), ks2 = static_cast<int>((p->nj + 3) / 4);
    if ((p->fwd.split.mb + 3) / 4 > (ks1 >= 22 ? 4 : 3) || (p->inv.split.mb + 3) / 4 > (ks2 >= 22 ? 4 : 3))
        return false;
    if (p->precision == SKB_F64)
        return launch_round_trip(round_trip_args<double>(p, weights, in, out), p->grid, s, dry_run);
    return launch_round_trip(round_trip_args<float>(p, weights, in, out), p->grid, s, dry_run);
}

template <typename R>
void launch_modal_work(uint64_t n, const void* w, const void* in, void* out, hipStream_t s) {
    hipLaunchKernelGGL(modal_work_kernel<R>, stream_grid(n), dim3(256), 0, s, static_cast<const R*>(w),
                       static_cast<const cx<R>*>(in), static_cast<cx<R>*>(out), static_cast<size_t>(n));
    check_hip(hipGetLastError(), "modal work launch");
}

}  // namespace

namespace skb {

bool vertical_round_trip_supported(const skb_vertical_plan* p) {
    return p && !p->wvm && round_trip(p, nullptr, nullptr, nullptr, nullptr, true);
}

void vertical_round_trip(const skb_vertical_plan* p, const void* weights, const void* in, void* out, void* stream) {
    if (!p || !weights || !in || !out) fail(SKB_INVALID_ARGUMENT, "null argument");
    if (p->wvm) fail(SKB_LOGIC_ERROR, "the modal round trip runs on compact vertical plans");
    if (!round_trip(p, weights, in, out, as_stream(stream), false))
        fail(SKB_UNSUPPORTED, "modal round trip not instantiated for this vertical size");
}

void modal_work(int precision, uint64_t n, const void* weights, const void* in, void* out, void* stream) {
    if (n == 0) return;
    if (!weights || !in || !out) fail(SKB_INVALID_ARGUMENT, "null argument");
    if (precision == SKB_F64) launch_modal_work<double>(n, weights, in, out, as_stream(stream));
    else launch_modal_work<float>(n, weights, in, out, as_stream(stream));
}

}  // namespace skb

namespace {

void fill_zero(const skb_vertical_plan* p, void* out, uint64_t elements, void* stream) {
    if (!p || !out) fail(SKB_INVALID_ARGUMENT, "null argument");
    if (!p->wvm) fail(SKB_LOGIC_ERROR, "output initialisation belongs to WVM-direct vertical plans");
    const dim3 grid(static_cast<unsigned>(std::max<uint64_t>(1, std::min<uint64_t>((elements + 255) / 256, 8192))));
    if (p->precision == SKB_F64)
        hipLaunchKernelGGL(fill_zero_kernel<double>, grid, dim3(256), 0, as_stream(stream),
                           static_cast<cx<double>*>(out), static_cast<size_t>(elements));
    else
        hipLaunchKernelGGL(fill_zero_kernel<float>, grid, dim3(256), 0, as_stream(stream),
                           static_cast<cx<float>*>(out), static_cast<size_t>(elements));
    check_hip(hipGetLastError(), "fill launch");
}

}  // namespace

// ---------------------------------------------------------------------------
// Type-I vertical transforms of the constant-stratification operator
// (src/constant_stratification.cpp:95-147 plans, 196-235 normalisation,
// 271-292 schedule).  Each transform is an Nz x Nz real matrix applied to
// the complex z-columns of one channel of a [rows][channels][Nz] array, in
// place, on the same MFMA GEMM as the K^2 family (one group, column stride
// Nz*channels).  The FFTW kinds map to matrices:
//   DCT          REDFT00 of length Nz:  Y_k = x_0 + (-1)^k x_{Nz-1}
//                                           + 2 sum_{j=1}^{Nz-2} x_j cos(pi j k/(Nz-1))
//   DST          RODFT00 of the Nz-2 interior points (endpoint rows identity,
//                as an in-place plan on the interior leaves them untouched):
//                Y_k = 2 sum_{j=1}^{Nz-2} x_j sin(pi j k/(Nz-1))
//   DCT_FORWARD  DCT, then x 1/(Nz-1) and the last row x 0.5
//                (normalizeForwardDct)
//   DST_FORWARD  DST, then x 1/(Nz-1) and zero endpoint rows
//                (normalizeForwardDst)
// A dense matrix on the matrix cores rather than a length-2(Nz-1) FFT: at
// Nz <= 260 the product is far below the matrix-core roofline and the stage
// stays bound by HBM, with no bit-reversal or twiddle passes.
// ---------------------------------------------------------------------------
namespace {

// cos / sin of pi*m/n with m reduced modulo 2n in integers first.
double cos_pi(long long m, long long n) {
    m %= 2 * n;
    return std::cos(3.14159265358979323846264338327950288 * static_cast<double>(m) / static_cast<double>(n));
}
double sin_pi(long long m, long long n) {
    m %= 2 * n;
    return std::sin(3.14159265358979323846264338327950288 * static_cast<double>(m) / static_cast<double>(n));
}

void type1_matrix(int kind, uint64_t nz_, std::vector<double>& a) {
    const long long nz = static_cast<long long>(nz_), n1 = nz - 1;
    a.assign(static_cast<size_t>(nz * nz), 0.0);
    const bool forward = kind == SKB_TYPE1_DCT_FORWARD || kind == SKB_TYPE1_DST_FORWARD;
    const double scale = forward ? 1.0 / static_cast<double>(n1) : 1.0;
    if (kind == SKB_TYPE1_DCT || kind == SKB_TYPE1_DCT_FORWARD) {
        for (long long k = 0; k < nz; ++k) {
            const double rs = scale * (forward && k == n1 ? 0.5 : 1.0);
            for (long long j = 0; j < nz; ++j) {
                double v;
                if (j == 0) v = 1.0;
                else if (j == n1) v = (k % 2 == 0) ? 1.0 : -1.0;
                else v = 2.0 * cos_pi(j * k, n1);
                a[static_cast<size_t>(k * nz + j)] = rs * v;
            }
        }
    } else {
        for (long long k = 1; k < n1; ++k)
            for (long long j = 1; j < n1; ++j) a[static_cast<size_t>(k * nz + j)] = scale * 2.0 * sin_pi(j * k, n1);
        if (!forward) {
            a[0] = 1.0;
            a[static_cast<size_t>(n1 * nz + n1)] = 1.0;
        }
    }
}

void check_kind(int kind) {
    if (kind < SKB_TYPE1_DCT || kind > SKB_TYPE1_DST_FORWARD) fail(SKB_INVALID_ARGUMENT, "unknown type-I transform kind");
}

}  // namespace

struct skb_type1_plan {
    uint64_t rows = 0, nz = 0;
    int precision = SKB_F64;
    int ntiles = 0;
    skb::GemmOperand op[4];  // indexed by kind
    skb::DeviceBuffer tiles;
};

namespace {

template <typename R>
void type1_run(const skb_type1_plan* p, int kind, uint64_t channels, uint64_t channel, void* data, hipStream_t s) {
    const GemmOperand& op = p->op[kind];
    GemmArgs<R> a{};
    cx<R>* col0 = static_cast<cx<R>*>(data) + p->nz * channel;
    a.in = col0;
    a.out = col0;
    a.ap = op.ap.as<R>();
    a.arow = op.arow.as<R>();
    a.tiles = p->tiles.as<GemmTile>();
    a.ntiles = p->ntiles;
    a.M = op.M;
    a.K = op.K;
    a.MB = op.split.mb;
    a.VR = op.split.vr;
    a.F = 1;
    a.in_ld = static_cast<int>(p->nz * channels);
    a.out_ld = a.in_ld;
    launch_gemm<R, false>(a, 0, s);
}

void type1_execute(const skb_type1_plan* p, int kind, uint64_t channels, uint64_t first, uint64_t count, void* data,
                   void* stream) {
    if (!p || !data) fail(SKB_INVALID_ARGUMENT, "null argument");
    check_kind(kind);
    if (channels == 0 || count == 0 || first + count > channels)
        fail(SKB_INVALID_ARGUMENT, "Invalid constant-stratification R2R plan shape.");
    if (p->rows * channels * p->nz > (1ull << 31)) fail(SKB_UNSUPPORTED, "type-I arrays above 2^31 elements");
    const hipStream_t s = as_stream(stream);
    for (uint64_t c = first; c < first + count; ++c) {
        if (p->precision == SKB_F64) type1_run<double>(p, kind, channels, c, data, s);
        else type1_run<float>(p, kind, channels, c, data, s);
    }
}

}  // namespace

extern "C" {

int skb_vertical_plan_create(const skb_workload* workload, uint64_t nj, const skb_mode_group* groups,
                             uint64_t group_count, const double* forward, const double* inverse, int precision,
                             skb_vertical_plan** plan) {
    return guarded([&] {
        if (!plan) fail(SKB_INVALID_ARGUMENT, "null plan pointer");
        *plan = nullptr;
        *plan = make_plan(workload, nj, groups, group_count, forward, inverse, precision).release();
    });
}

int skb_vertical_wvm_plan_create(const skb_workload* workload, const skb_retained_mode* modes, uint64_t mode_count,
                                 uint64_t nj, const skb_mode_group* groups, uint64_t group_count,
                                 const double* forward, const double* inverse, int precision,
                                 skb_vertical_plan** plan) {
    return guarded([&] {
        if (!plan) fail(SKB_INVALID_ARGUMENT, "null plan pointer");
        *plan = nullptr;
        validate_workload(workload);
        // src/vertical_gemm.cpp:965-1034
        if (nj != (workload->nz < 2 ? 0 : 2 * (workload->nz - 1) / 3))
            fail(SKB_INVALID_ARGUMENT, "Direct WVM vertical GEMM operator dimensions do not match the workload.");
        if (!modes || mode_count == 0 || !groups || group_count == 0)
            fail(SKB_INVALID_ARGUMENT, "Direct WVM vertical GEMM requires modes, groups, and fields.");
        auto p = make_plan(workload, nj, groups, group_count, forward, inverse, precision);
        if (p->modes != mode_count)
            fail(SKB_INVALID_ARGUMENT, "Direct WVM vertical GEMM groups do not cover every retained mode.");
        const uint64_t nxh = workload->nx / 2 + 1, ny = workload->ny;
        const uint64_t frequencies = nxh * ny;
        const uint64_t P = workload->nz * workload->fields;
        if (frequencies * P > (1ull << 31) || frequencies * nj * workload->fields > (1ull << 31))
            fail(SKB_UNSUPPORTED, "WVM arrays above 2^31 elements");
        std::vector<int> q(mode_count), partner(mode_count, -1);
        std::vector<uint8_t> used(frequencies, 0);
        for (uint64_t m = 0; m < mode_count; ++m) {
            const auto& md = modes[m];
            if (md.stored_kx >= nxh || md.stored_ky >= ny)
                fail(SKB_INVALID_ARGUMENT, "A retained mode lies outside the half-spectrum.");
            const uint64_t f = md.stored_kx + nxh * md.stored_ky;
            if (used[f])
                fail(SKB_INVALID_ARGUMENT,
                     "Direct WVM vertical GEMM requires a unique stored frequency for every retained mode.");
            used[f] = 1;
            q[m] = static_cast<int>(f);
            if (md.stored_kx == 0 && md.stored_ky != 0 && 2 * md.stored_ky != ny)
                partner[m] = static_cast<int>(nxh * ((ny - md.stored_ky) % ny));
        }
        p->wvm = true;
        p->qmap.upload(q.data(), q.size() * sizeof(int));
        p->partner.upload(partner.data(), partner.size() * sizeof(int));
        *plan = p.release();
    });
}

int skb_vertical_plan_destroy(skb_vertical_plan* plan) {
    return guarded([&] { delete plan; });
}

int skb_vertical_forward(skb_vertical_plan* plan, const void* physical, void* modal, void* stream) {
    return guarded([&] { run(plan, true, physical, modal, stream, false); });
}

int skb_vertical_inverse(skb_vertical_plan* plan, const void* modal, void* physical, void* stream) {
    return guarded([&] { run(plan, false, modal, physical, stream, false); });
}

int skb_vertical_wvm_forward(skb_vertical_plan* plan, const void* full_spectrum, void* full_modal, void* stream) {
    return guarded([&] { run(plan, true, full_spectrum, full_modal, stream, true); });
}

int skb_vertical_wvm_inverse(skb_vertical_plan* plan, const void* full_modal, void* full_spectrum, void* stream) {
    return guarded([&] { run(plan, false, full_modal, full_spectrum, stream, true); });
}

uint64_t skb_vertical_wvm_modal_elements(const skb_vertical_plan* plan) {
    if (!plan) return 0;
    return (plan->w.nx / 2 + 1) * plan->w.ny * plan->nj * plan->w.fields;
}

int skb_vertical_wvm_initialize_modal(skb_vertical_plan* plan, void* full_modal, void* stream) {
    return guarded([&] { fill_zero(plan, full_modal, skb_vertical_wvm_modal_elements(plan), stream); });
}

int skb_vertical_wvm_initialize_spectrum(skb_vertical_plan* plan, void* full_spectrum, void* stream) {
    return guarded([&] {
        if (!plan) fail(SKB_INVALID_ARGUMENT, "null plan");
        const auto& w = plan->w;
        fill_zero(plan, full_spectrum, (w.nx / 2 + 1) * w.ny * w.nz * w.fields, stream);
    });
}

int skb_modal_work(int precision, uint64_t count, const void* weights, const void* input, void* output,
                   void* stream) {
    return guarded([&] {
        check_precision(precision);
        modal_work(precision, count, weights, input, output, stream);
    });
}

int skb_modal_work_split(int precision, uint64_t count, const void* weights, const void* input_real,
                         const void* input_imag, void* output_real, void* output_imag, void* stream) {
    return guarded([&] {
        check_precision(precision);
        if (count == 0) return;
        if (!weights || !input_real || !input_imag || !output_real || !output_imag)
            fail(SKB_INVALID_ARGUMENT, "null argument");
        const hipStream_t s = as_stream(stream);
        if (precision == SKB_F64)
            hipLaunchKernelGGL(modal_work_split_kernel<double>, stream_grid(count), dim3(256), 0, s,
                               static_cast<const double*>(weights), static_cast<const double*>(input_real),
                               static_cast<const double*>(input_imag), static_cast<double*>(output_real),
                               static_cast<double*>(output_imag), static_cast<size_t>(count));
        else
            hipLaunchKernelGGL(modal_work_split_kernel<float>, stream_grid(count), dim3(256), 0, s,
                               static_cast<const float*>(weights), static_cast<const float*>(input_real),
                               static_cast<const float*>(input_imag), static_cast<float*>(output_real),
                               static_cast<float*>(output_imag), static_cast<size_t>(count));
        check_hip(hipGetLastError(), "modal work launch");
    });
}

int skb_vertical_wvm_modal_work(skb_vertical_plan* plan, const void* weights, const void* full_modal_input,
                                void* full_modal_output, void* stream) {
    return guarded([&] {
        if (!plan || !weights || !full_modal_input || !full_modal_output) fail(SKB_INVALID_ARGUMENT, "null argument");
        if (!plan->wvm) fail(SKB_LOGIC_ERROR, "WVM modal work needs a WVM-direct vertical plan");
        if (full_modal_input == full_modal_output)
            fail(SKB_INVALID_ARGUMENT, "WVM modal work is out of place (partner entries are rewritten)");
        const int block = static_cast<int>(plan->nj * plan->w.fields);
        const uint64_t n = plan->modes * block;
        const hipStream_t s = as_stream(stream);
        if (plan->precision == SKB_F64)
            hipLaunchKernelGGL(modal_work_wvm_kernel<double>, stream_grid(n), dim3(256), 0, s,
                               static_cast<const double*>(weights), static_cast<const cx<double>*>(full_modal_input),
                               static_cast<cx<double>*>(full_modal_output), plan->qmap.as<int>(),
                               plan->partner.as<int>(), block, static_cast<size_t>(n));
        else
            hipLaunchKernelGGL(modal_work_wvm_kernel<float>, stream_grid(n), dim3(256), 0, s,
                               static_cast<const float*>(weights), static_cast<const cx<float>*>(full_modal_input),
                               static_cast<cx<float>*>(full_modal_output), plan->qmap.as<int>(),
                               plan->partner.as<int>(), block, static_cast<size_t>(n));
        check_hip(hipGetLastError(), "WVM modal work launch");
    });
}

uint64_t skb_vertical_matrix_bytes(const skb_vertical_plan* plan) { return plan ? plan->matrix_bytes : 0; }

}  // extern "C"

namespace {

// Split layout over the interleaved kernels: interleave the input pair into
// interleaved staging, one grouped launch, deinterleave into the output pair.
// The matrices are real, so the real (imaginary) output part depends on the
// real (imaginary) input part alone: a component whose input and output are
// both null is left out (the other input fills both staged parts and only
// the requested output part is written).  Every staged entry is written
// before it is read, so the staging never needs clearing.
void run_split_staged(skb_vertical_plan* plan, bool forward, const void* in_re, const void* in_im, void* out_re,
                      void* out_im, void* stage_physical, void* stage_modal, void* stream) {
    if ((in_re == nullptr) != (out_re == nullptr) || (in_im == nullptr) != (out_im == nullptr) ||
        (!in_re && !in_im))
        fail(SKB_INVALID_ARGUMENT, "a split component needs both its input and its output array");
    const hipStream_t s = as_stream(stream);
    const uint64_t np = plan->modes * plan->w.nz * plan->w.fields, nm = plan->modes * plan->nj * plan->w.fields;
    void* stage_in = forward ? stage_physical : stage_modal;
    void* stage_out = forward ? stage_modal : stage_physical;
    interleave(plan->precision, forward ? np : nm, in_re ? in_re : in_im, in_im ? in_im : in_re, stage_in, s);
    run(plan, forward, stage_in, stage_out, stream, false);
    deinterleave(plan->precision, forward ? nm : np, stage_out, out_re, out_im, 1.0, s);
}

// Staging of skb_vertical_forward_split_noalloc: a file-scope device array,
// one copy per GPU at its own address, held until the process ends.  Host
// code never passes the array's name to a kernel or copy (on the host it is
// only a stand-in symbol): it uses the per-device address that
// hipGetSymbolAddress returns, looked up once per device.
constexpr uint64_t kNoAllocScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex scratch_mutex;
constexpr int kMaxDevices = 64;
void* scratch_ptr[kMaxDevices] = {};

void* scratch_device_address() {
    int device = 0;
    check_hip(hipGetDevice(&device), "hipGetDevice");
    if (device < 0 || device >= kMaxDevices) fail(SKB_UNSUPPORTED, "device ordinal above 63");
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_ptr[device]) {
        void* ptr = nullptr;
        check_hip(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!ptr) fail(SKB_RUNTIME_ERROR, "hipGetSymbolAddress(scratch) returned a null address");
        scratch_ptr[device] = ptr;
    }
    return scratch_ptr[device];
}

void run_split(skb_vertical_plan* plan, bool forward, const void* in_re, const void* in_im, void* out_re,
               void* out_im, void* stream) {
    if (!plan) fail(SKB_INVALID_ARGUMENT, "null plan");
    if (!plan->split_physical.ptr)
        fail(SKB_LOGIC_ERROR, "Split GEMM component requested from the complex GEMM provider.");
    run_split_staged(plan, forward, in_re, in_im, out_re, out_im, plan->split_physical.ptr, plan->split_modal.ptr,
                     stream);
}

}  // namespace

extern "C" {

int skb_vertical_plan_reserve_split(skb_vertical_plan* plan) {
    return guarded([&] {
        if (!plan) fail(SKB_INVALID_ARGUMENT, "null plan");
        if (plan->wvm) fail(SKB_LOGIC_ERROR, "the WVM-direct GEMM has no split layout");
        if (plan->split_physical.ptr) return;
        const uint64_t cb = complex_bytes(plan->precision);
        plan->split_physical.allocate(plan->modes * plan->w.nz * plan->w.fields * cb);
        plan->split_modal.allocate(plan->modes * plan->nj * plan->w.fields * cb);
    });
}

int skb_vertical_forward_split(skb_vertical_plan* plan, const void* physical_real, const void* physical_imag,
                               void* modal_real, void* modal_imag, void* stream) {
    return guarded([&] { run_split(plan, true, physical_real, physical_imag, modal_real, modal_imag, stream); });
}

int skb_vertical_inverse_split(skb_vertical_plan* plan, const void* modal_real, const void* modal_imag,
                               void* physical_real, void* physical_imag, void* stream) {
    return guarded([&] { run_split(plan, false, modal_real, modal_imag, physical_real, physical_imag, stream); });
}

int skb_vertical_forward_split_noalloc(skb_vertical_plan* plan, const void* physical_real, const void* physical_imag,
                                       void* modal_real, void* modal_imag, void* stream) {
    return guarded([&] {
        if (!plan) fail(SKB_INVALID_ARGUMENT, "null plan");
        if (plan->wvm) fail(SKB_LOGIC_ERROR, "the WVM-direct GEMM has no split layout");
        const uint64_t cb = complex_bytes(plan->precision);
        const uint64_t physical_bytes = (plan->modes * plan->w.nz * plan->w.fields * cb + 255) & ~uint64_t(255);
        const uint64_t modal_bytes = plan->modes * plan->nj * plan->w.fields * cb;
        if (physical_bytes + modal_bytes > kNoAllocScratchBytes)
            fail(SKB_INVALID_ARGUMENT, "split staging of " + std::to_string(physical_bytes + modal_bytes) +
                                           " bytes exceeds the " + std::to_string(kNoAllocScratchBytes) +
                                           "-byte no-allocation scratch");
        char* base = static_cast<char*>(scratch_device_address());
        run_split_staged(plan, true, physical_real, physical_imag, modal_real, modal_imag, base,
                         base + physical_bytes, stream);
    });
}

int skb_vertical_plan_set_grid(skb_vertical_plan* plan, int workgroups) {
    return guarded([&] {
        if (!plan || workgroups < 0) fail(SKB_INVALID_ARGUMENT, "bad workgroup count");
        plan->grid = workgroups;
    });
}

uint64_t skb_vertical_plan_items(const skb_vertical_plan* plan) { return plan ? static_cast<uint64_t>(plan->ntiles) : 0; }

int skb_vertical_plan_set_products(skb_vertical_plan* plan, int products) {
    return guarded([&] {
        if (!plan) fail(SKB_INVALID_ARGUMENT, "null plan");
        if (products != SKB_PRODUCTS_FP32 && products != SKB_PRODUCTS_SPLIT_BF16)
            fail(SKB_INVALID_ARGUMENT, "products must be SKB_PRODUCTS_FP32 or SKB_PRODUCTS_SPLIT_BF16");
        plan->products = products;
    });
}

int skb_vertical_plan_products(const skb_vertical_plan* plan, int direction) {
    if (!plan || (direction != 0 && direction != 1)) return -1;
    const GemmOperand& op = direction == 0 ? plan->fwd : plan->inv;
    return plan->products == SKB_PRODUCTS_SPLIT_BF16 && op.a3.ptr && !plan->wvm ? SKB_PRODUCTS_SPLIT_BF16
                                                                                 : SKB_PRODUCTS_FP32;
}

int skb_type1_matrix(uint64_t nz, int kind, double* matrix) {
    return guarded([&] {
        if (!matrix || nz < 3) fail(SKB_INVALID_ARGUMENT, "type-I transforms need Nz >= 3 and an output matrix");
        check_kind(kind);
        std::vector<double> a;
        type1_matrix(kind, nz, a);
        std::copy(a.begin(), a.end(), matrix);
    });
}

int skb_type1_plan_create(uint64_t rows, uint64_t nz, int precision, skb_type1_plan** plan) {
    return guarded([&] {
        if (!plan) fail(SKB_INVALID_ARGUMENT, "null plan pointer");
        *plan = nullptr;
        check_precision(precision);
        if (rows == 0 || nz < 3) fail(SKB_INVALID_ARGUMENT, "Invalid constant-stratification R2R plan shape.");
        const bool fp64 = precision == SKB_F64;
        const int ks = static_cast<int>((nz + 3) / 4);
        if (nz > (fp64 ? 132u : 260u) || !gemm_ks_instantiated(ks, fp64) ||
            (row_split(static_cast<int>(nz)).mb + 3) / 4 > (ks >= 22 ? 4 : 3))
            fail(SKB_UNSUPPORTED, "type-I transform length " + std::to_string(nz) + " is not instantiated");
        if (rows > (1ull << 26)) fail(SKB_UNSUPPORTED, "too many type-I rows");
        auto p = std::make_unique<skb_type1_plan>();
        p->rows = rows;
        p->nz = nz;
        p->precision = precision;
        std::vector<double> a;
        for (int kind = 0; kind < 4; ++kind) {
            type1_matrix(kind, nz, a);
            p->op[kind].build(a.data(), 1, static_cast<int>(nz), static_cast<int>(nz), precision);
        }
        std::vector<GemmTile> tiles;
        for (uint64_t c = 0; c < rows; c += 16)
            tiles.push_back(GemmTile{0, static_cast<int>(c), static_cast<int>(std::min<uint64_t>(16, rows - c)), 0});
        p->ntiles = static_cast<int>(tiles.size());
        p->tiles.upload(tiles.data(), tiles.size() * sizeof(GemmTile));
        *plan = p.release();
    });
}

int skb_type1_plan_destroy(skb_type1_plan* plan) {
    return guarded([&] { delete plan; });
}

int skb_type1_execute(skb_type1_plan* plan, int kind, uint64_t channels, uint64_t first_channel, uint64_t count,
                      void* data, void* stream) {
    return guarded([&] { type1_execute(plan, kind, channels, first_channel, count, data, stream); });
}

int skb_type1_vertical_schedule(skb_type1_plan* plan, void* arena3, void* arena1, void* stream) {
    return guarded([&] {
        if (!plan || !arena3 || !arena1) fail(SKB_INVALID_ARGUMENT, "null argument");
        // executeCompleteVerticalGraph (src/constant_stratification.cpp:271-292): one
        // velocity reconstruction and two cosine- and two sine-family derivative
        // reconstructions on the three-channel arena, then four projected flux
        // targets on the one-channel arena.
        for (int g = 0; g < 3; ++g) {
            type1_execute(plan, SKB_TYPE1_DCT, 3, 0, 2, arena3, stream);
            type1_execute(plan, SKB_TYPE1_DST, 3, 2, 1, arena3, stream);
        }
        for (int g = 0; g < 2; ++g) {
            type1_execute(plan, SKB_TYPE1_DST, 3, 0, 2, arena3, stream);
            type1_execute(plan, SKB_TYPE1_DCT, 3, 2, 1, arena3, stream);
        }
        for (int g = 0; g < 2; ++g) type1_execute(plan, SKB_TYPE1_DCT_FORWARD, 1, 0, 1, arena1, stream);
        for (int g = 0; g < 2; ++g) type1_execute(plan, SKB_TYPE1_DST_FORWARD, 1, 0, 1, arena1, stream);
    });
}

}  // extern "C"
