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
// Work decomposition: one wavefront per 16 consecutive complex columns of a
// group computes all ceil(M/16) row blocks with 16x16x4 MFMA
// (v_mfma_f64_16x16x4_f64 for FP64, v_mfma_f32_16x16x4_f32 for FP32) into
// two accumulators (Re, Im).  See grouped_gemm_kernel for the operand order.
//
// MFMA 16x16x4 operand layout (64 lanes, lane = l), measured on gfx950:
//   A[i][k]: i = l % 16, k = l / 16        B[k][j]: j = l % 16, k = l / 16
//   D[i][j]: j = l % 16 and, for accumulator element r,
//            f32: i = 4 * (l / 16) + r      f64: i = l / 16 + 4 * r
#include <algorithm>
#include <memory>
#include <vector>

#include "common.hpp"

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

// A tile is 16 consecutive complex columns of one group (one wavefront).
struct GemmTile {
    int group;
    int col_begin;
    int ncols;  // 1..16
    int pad;
};

constexpr int kMaxK = 260;

// The reduction index is permuted so that every lane's operands are
// contiguous: MFMA k-step s, lane group q = lane / 16 covers
//     z = KS * q + s,   s = 0 .. KS-1,   KS = ceil(K / 4).
// B (the input columns): lane (j = lane % 16, q) needs z = KS*q .. KS*q+KS-1
// of column j -- one contiguous run of KS complex values.  A is stored by the
// plan in fragment order
//     Ap[g][mb][s][lane] = A_g[16*mb + lane % 16][KS * (lane / 16) + s]
// (zero outside M x K), so each A fragment is one coalesced 64-element read.
//
// Wavefronts are persistent over tiles (tile = wave, wave + waves, ...).
// Memory ordering matters on gfx9, where one in-order counter covers loads
// and stores: every load a wavefront will wait on is issued BEFORE the
// stores that precede its use -- the next tile's B and the next row block's A
// are prefetched, and all stores of a tile are issued after its last MFMA.
template <typename R, int KS>
__device__ __forceinline__ void gemm_load_b(cx<R>* b, const cx<R>* in, const GemmTile& t, int K, int lane) {
    const int j = lane & 15, q = lane >> 4;
    const bool col_ok = j < t.ncols;
    const int z0 = KS * q;
    const cx<R>* src = in + static_cast<size_t>(K) * (t.col_begin + (col_ok ? j : 0)) + z0;
    if (col_ok && z0 + KS <= K) {
#pragma unroll
        for (int s = 0; s < KS; ++s) b[s] = src[s];
    } else {
#pragma unroll
        for (int s = 0; s < KS; ++s) b[s] = (col_ok && z0 + s < K) ? src[s] : cx<R>{R(0), R(0)};
    }
}

template <typename R, int KS, int RB>
__global__ __launch_bounds__(256) void grouped_gemm_kernel(const cx<R>* __restrict__ in, cx<R>* __restrict__ out,
                                                          const R* __restrict__ ap,
                                                          const GemmTile* __restrict__ tiles, int ntiles, int M,
                                                          int K) {
    using A4 = typename Acc<R>::type;
    // per-wave output stage: 16 columns x (16*RB) rows, complex
    __shared__ cx<R> stage[4][16 * RB * 16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int waves = gridDim.x * 4;
    // wave-uniform tile index -> scalar descriptor loads (counted apart from
    // the vector loads and stores)
    // One wavefront per tile: XCD-aware (each XCD takes a contiguous range of
    // workgroups' tiles, whose input and output blocks share boundary lines).
    const int wgs_needed = (ntiles + 3) / 4;
    const int wg = gridDim.x >= static_cast<unsigned>(wgs_needed) ? xcd_item(blockIdx.x, wgs_needed) : blockIdx.x;
    int tile = __builtin_amdgcn_readfirstlane(wg * 4 + wave);
    if (wg >= wgs_needed || tile >= ntiles) return;
    const int j = lane & 15;
    cx<R>* st = stage[wave];
    while (true) {
        const GemmTile t = tiles[tile];
        cx<R> b[KS];
        gemm_load_b<R, KS>(b, in, t, K, lane);
        const R* a_ptr = ap + static_cast<size_t>(t.group) * RB * KS * 64 + lane;
        A4 re[RB], im[RB];
        R a[KS], an[KS];
#pragma unroll
        for (int s = 0; s < KS; ++s) a[s] = a_ptr[s * 64];
#pragma unroll
        for (int mb = 0; mb < RB; ++mb) {
            if (mb + 1 < RB) {
#pragma unroll
                for (int s = 0; s < KS; ++s) an[s] = a_ptr[((mb + 1) * KS + s) * 64];
            }
            A4 r = {0, 0, 0, 0}, i = {0, 0, 0, 0};
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                r = Acc<R>::mfma(a[s], b[s].x, r);
                i = Acc<R>::mfma(a[s], b[s].y, i);
            }
            re[mb] = r;
            im[mb] = i;
            if (mb + 1 < RB) {
#pragma unroll
                for (int s = 0; s < KS; ++s) a[s] = an[s];
            }
        }
        // Stage the 16-column block column-major with row stride M (the
        // output layout), then store it as one contiguous run.
#pragma unroll
        for (int mb = 0; mb < RB; ++mb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = mb * 16 + Acc<R>::row(lane, r);
                if (row < M) st[j * M + row] = cx<R>{re[mb][r], im[mb][r]};
            }
        // the stage is private to this wavefront: order its LDS writes
        // before the reads of other lanes
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        cx<R>* dst = out + static_cast<size_t>(M) * t.col_begin;
        const int total = M * t.ncols;
        for (int e = lane; e < total; e += 64) dst[e] = st[e];
        tile += waves;
        if (tile >= ntiles) break;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// Host: matrices (row-major M x K per group) -> fragment order.
template <typename R>
std::vector<R> fragment_order(const double* mats, uint64_t groups, int M, int K) {
    const int KS = (K + 3) / 4, rblocks = (M + 15) / 16;
    std::vector<R> ap(static_cast<size_t>(groups) * rblocks * KS * 64, R(0));
    for (uint64_t g = 0; g < groups; ++g)
        for (int mb = 0; mb < rblocks; ++mb)
            for (int s = 0; s < KS; ++s)
                for (int l = 0; l < 64; ++l) {
                    const int i = 16 * mb + (l & 15), k = KS * (l >> 4) + s;
                    if (i < M && k < K)
                        ap[((g * rblocks + mb) * KS + s) * 64 + l] =
                            static_cast<R>(mats[(g * M + i) * static_cast<uint64_t>(K) + k]);
                }
    return ap;
}

namespace {

template <typename R, int KS, int RB>
void launch_gemm_rb(const void* in, void* out, const void* ap, const GemmTile* tiles, int ntiles, int M, int K,
                    int req, hipStream_t s) {
    const int wgs = (ntiles + 3) / 4;
    const int grid = req > 0 && (req + 3) / 4 < wgs ? (req + 3) / 4 : xcd_grid(wgs);
    hipLaunchKernelGGL((grouped_gemm_kernel<R, KS, RB>), dim3(grid), dim3(256), 0, s,
                       static_cast<const cx<R>*>(in), static_cast<cx<R>*>(out), static_cast<const R*>(ap), tiles,
                       ntiles, M, K);
    check_hip(hipGetLastError(), "vertical GEMM launch");
}

template <typename R, int KS>
void launch_gemm_ks(const void* in, void* out, const void* ap, const GemmTile* tiles, int ntiles, int M, int K,
                    int req, hipStream_t s) {
    switch ((M + 15) / 16) {
#define SKB_RB(N) \
    case N: launch_gemm_rb<R, KS, N>(in, out, ap, tiles, ntiles, M, K, req, s); return;
        SKB_RB(1) SKB_RB(2) SKB_RB(3) SKB_RB(4) SKB_RB(5) SKB_RB(6) SKB_RB(7) SKB_RB(8) SKB_RB(9)
#undef SKB_RB
    }
    fail(SKB_UNSUPPORTED, "vertical GEMM output dimension above 144");
}

template <typename R>
void launch_gemm(const void* in, void* out, const void* ap, const GemmTile* tiles, int ntiles, int M, int K, int req,
                 hipStream_t s) {
    switch ((K + 3) / 4) {
#define SKB_KS(N) \
    case N: launch_gemm_ks<R, N>(in, out, ap, tiles, ntiles, M, K, req, s); return;
        SKB_KS(1) SKB_KS(2) SKB_KS(3) SKB_KS(4) SKB_KS(5) SKB_KS(6) SKB_KS(7) SKB_KS(8) SKB_KS(9) SKB_KS(10)
        SKB_KS(11) SKB_KS(12) SKB_KS(13) SKB_KS(14) SKB_KS(15) SKB_KS(16) SKB_KS(17) SKB_KS(22) SKB_KS(33)
#undef SKB_KS
    }
    fail(SKB_UNSUPPORTED, "vertical GEMM inner dimension: ceil(K/4) must be <= 17, 22 or 33");
}

}  // namespace
}  // namespace skb

struct skb_vertical_plan {
    skb_workload w{};
    uint64_t nj = 0;
    int precision = SKB_F64;
    uint64_t groups = 0;
    int ntiles = 0;
    int waves = 0;  // wavefronts per launch (0 = one per tile)
    uint64_t matrix_bytes = 0;  // logical matrix bytes per direction
    skb::DeviceBuffer fwd, inv, tiles;  // matrices in fragment order
};

using namespace skb;

extern "C" {

int skb_vertical_plan_create(const skb_workload* workload, uint64_t nj, const skb_mode_group* groups,
                             uint64_t group_count, const double* forward, const double* inverse, int precision,
                             skb_vertical_plan** plan) {
    return guarded([&] {
        if (!plan) fail(SKB_INVALID_ARGUMENT, "null plan pointer");
        *plan = nullptr;
        check_precision(precision);
        validate_workload(workload);
        if (!groups || group_count == 0) fail(SKB_INVALID_ARGUMENT, "vertical GEMM requires at least one group");
        if (!forward || !inverse) fail(SKB_INVALID_ARGUMENT, "vertical GEMM requires forward and inverse matrices");
        const uint64_t nz = workload->nz;
        if (nj == 0 || nz == 0) fail(SKB_INVALID_ARGUMENT, "Vertical operator dimensions must be positive.");
        if (nz > 132 || nj > 132)
            fail(SKB_UNSUPPORTED, "vertical dimensions above 132 (Nz, Nj) are not instantiated");
        // Groups must tile [0, modes) contiguously (src/model.cpp:191-206).
        uint64_t next = 0;
        for (uint64_t g = 0; g < group_count; ++g) {
            if (groups[g].first_mode != next || groups[g].mode_count == 0)
                fail(SKB_INVALID_ARGUMENT, "vertical groups must be contiguous, non-empty and start at mode 0");
            next += groups[g].mode_count;
        }
        const uint64_t fields = workload->fields;
        if (next * fields > (1ull << 31)) fail(SKB_UNSUPPORTED, "too many vertical columns");
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
        p->ntiles = static_cast<int>(tiles.size());
        const int inj = static_cast<int>(nj), inz = static_cast<int>(nz);
        if (precision == SKB_F64) {
            const auto f = fragment_order<double>(forward, group_count, inj, inz);
            const auto i = fragment_order<double>(inverse, group_count, inz, inj);
            p->fwd.upload(f.data(), f.size() * sizeof(double));
            p->inv.upload(i.data(), i.size() * sizeof(double));
        } else {
            const auto f = fragment_order<float>(forward, group_count, inj, inz);
            const auto i = fragment_order<float>(inverse, group_count, inz, inj);
            p->fwd.upload(f.data(), f.size() * sizeof(float));
            p->inv.upload(i.data(), i.size() * sizeof(float));
        }
        p->tiles.upload(tiles.data(), tiles.size() * sizeof(GemmTile));
        p->matrix_bytes = nj * nz * group_count * (precision == SKB_F64 ? 8 : 4);
        *plan = p.release();
    });
}

int skb_vertical_plan_destroy(skb_vertical_plan* plan) {
    return guarded([&] { delete plan; });
}

int skb_vertical_forward(skb_vertical_plan* plan, const void* physical, void* modal, void* stream) {
    return guarded([&] {
        if (!plan || !physical || !modal) fail(SKB_INVALID_ARGUMENT, "null argument");
        const int M = static_cast<int>(plan->nj), K = static_cast<int>(plan->w.nz);
        if (plan->precision == SKB_F64)
            launch_gemm<double>(physical, modal, plan->fwd.ptr, plan->tiles.as<GemmTile>(), plan->ntiles, M, K, plan->waves,
                                as_stream(stream));
        else
            launch_gemm<float>(physical, modal, plan->fwd.ptr, plan->tiles.as<GemmTile>(), plan->ntiles, M, K, plan->waves,
                               as_stream(stream));
    });
}

int skb_vertical_inverse(skb_vertical_plan* plan, const void* modal, void* physical, void* stream) {
    return guarded([&] {
        if (!plan || !physical || !modal) fail(SKB_INVALID_ARGUMENT, "null argument");
        const int M = static_cast<int>(plan->w.nz), K = static_cast<int>(plan->nj);
        if (plan->precision == SKB_F64)
            launch_gemm<double>(modal, physical, plan->inv.ptr, plan->tiles.as<GemmTile>(), plan->ntiles, M, K, plan->waves,
                                as_stream(stream));
        else
            launch_gemm<float>(modal, physical, plan->inv.ptr, plan->tiles.as<GemmTile>(), plan->ntiles, M, K, plan->waves,
                               as_stream(stream));
    });
}

uint64_t skb_vertical_matrix_bytes(const skb_vertical_plan* plan) { return plan ? plan->matrix_bytes : 0; }

int skb_vertical_plan_set_grid(skb_vertical_plan* plan, int wavefronts) {
    return guarded([&] {
        if (!plan || wavefronts < 0) fail(SKB_INVALID_ARGUMENT, "bad wavefront count");
        plan->waves = wavefronts;
    });
}

uint64_t skb_vertical_plan_items(const skb_vertical_plan* plan) { return plan ? static_cast<uint64_t>(plan->ntiles) : 0; }

}  // extern "C"
