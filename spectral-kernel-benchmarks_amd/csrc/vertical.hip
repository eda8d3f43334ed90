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
// Work decomposition: a TILE is 16 consecutive complex columns of one group;
// one wavefront owns one tile and computes all ceil(M/16) row blocks with
// 16x16x4 MFMA (v_mfma_f64_16x16x4_f64 for FP64, v_mfma_f32_16x16x4_f32 for
// FP32), two accumulators (Re, Im) per row block.  The B fragments (the
// tile's input columns) are loaded once into registers; A fragments stream
// from the group's matrix (L2-resident: 22 KiB at 42x65 FP64).
//
// MFMA 16x16x4 operand layout (64 lanes, lane = l):
//   A[i][k]: i = l % 16, k = l / 16        B[k][j]: j = l % 16, k = l / 16
//   D[i][j]: j = l % 16, i = 4 * (l / 16) + r for accumulator element r
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
    static __device__ __forceinline__ type mfma(float a, float b, type c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};
template <>
struct Acc<double> {
    using type = doublex4;
    static __device__ __forceinline__ type mfma(double a, double b, type c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};

struct GemmTile {
    int group;      // matrix index
    int col_begin;  // first complex column
    int ncols;      // 1..16
    int pad;
};

constexpr int kMaxKSteps = 65;  // K <= 260

template <typename R, int KSTEPS>
__global__ __launch_bounds__(256) void grouped_gemm_kernel(const cx<R>* __restrict__ in, cx<R>* __restrict__ out,
                                                          const R* __restrict__ mats,
                                                          const GemmTile* __restrict__ tiles, int ntiles, int M,
                                                          int K) {
    using A4 = typename Acc<R>::type;
    const int lane = threadIdx.x & 63;
    const int tile_id = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (tile_id >= ntiles) return;
    const GemmTile t = tiles[tile_id];
    const int j = lane & 15;
    const int kq = lane >> 4;
    const bool col_ok = j < t.ncols;
    const cx<R>* src = in + static_cast<size_t>(K) * (t.col_begin + (col_ok ? j : 0));
    // B fragments for every k-step.
    cx<R> b[KSTEPS];
#pragma unroll
    for (int s = 0; s < KSTEPS; ++s) {
        const int k = 4 * s + kq;
        b[s] = (col_ok && k < K) ? src[k] : cx<R>{R(0), R(0)};
    }
    const R* A = mats + static_cast<size_t>(t.group) * M * K;
    const int row_blocks = (M + 15) >> 4;
    for (int mb = 0; mb < row_blocks; ++mb) {
        const int ai = mb * 16 + j;  // A row for this lane
        const bool row_ok = ai < M;
        A4 re = {0, 0, 0, 0}, im = {0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < KSTEPS; ++s) {
            const int k = 4 * s + kq;
            const R a = (row_ok && k < K) ? A[static_cast<size_t>(ai) * K + k] : R(0);
            re = Acc<R>::mfma(a, b[s].x, re);
            im = Acc<R>::mfma(a, b[s].y, im);
        }
        if (col_ok) {
            cx<R>* dst = out + static_cast<size_t>(M) * (t.col_begin + j);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = mb * 16 + 4 * kq + r;
                if (i < M) dst[i] = cx<R>{re[r], im[r]};
            }
        }
    }
}

namespace {

template <typename R, int KSTEPS>
void launch_gemm_k(const void* in, void* out, const void* mats, const GemmTile* tiles, int ntiles, int M, int K,
                   hipStream_t s) {
    const dim3 grid((ntiles + 3) / 4);
    hipLaunchKernelGGL((grouped_gemm_kernel<R, KSTEPS>), grid, dim3(256), 0, s, static_cast<const cx<R>*>(in),
                       static_cast<cx<R>*>(out), static_cast<const R*>(mats), tiles, ntiles, M, K);
    check_hip(hipGetLastError(), "vertical GEMM launch");
}

// K is rounded up to a multiple of 4; instantiate a handful of k-step counts.
template <typename R>
void launch_gemm(const void* in, void* out, const void* mats, const GemmTile* tiles, int ntiles, int M, int K,
                 hipStream_t s) {
    const int steps = (K + 3) / 4;
    if (steps <= 2) launch_gemm_k<R, 2>(in, out, mats, tiles, ntiles, M, K, s);
    else if (steps <= 4) launch_gemm_k<R, 4>(in, out, mats, tiles, ntiles, M, K, s);
    else if (steps <= 9) launch_gemm_k<R, 9>(in, out, mats, tiles, ntiles, M, K, s);
    else if (steps <= 11) launch_gemm_k<R, 11>(in, out, mats, tiles, ntiles, M, K, s);
    else if (steps <= 17) launch_gemm_k<R, 17>(in, out, mats, tiles, ntiles, M, K, s);
    else if (steps <= 22) launch_gemm_k<R, 22>(in, out, mats, tiles, ntiles, M, K, s);
    else if (steps <= 33) launch_gemm_k<R, 33>(in, out, mats, tiles, ntiles, M, K, s);
    else if (steps <= 43) launch_gemm_k<R, 43>(in, out, mats, tiles, ntiles, M, K, s);
    else if (steps <= kMaxKSteps) launch_gemm_k<R, kMaxKSteps>(in, out, mats, tiles, ntiles, M, K, s);
    else fail(SKB_UNSUPPORTED, "vertical GEMM inner dimension above 260");
}

}  // namespace
}  // namespace skb

struct skb_vertical_plan {
    skb_workload w{};
    uint64_t nj = 0;
    int precision = SKB_F64;
    uint64_t groups = 0;
    int ntiles = 0;
    skb::DeviceBuffer fwd, inv, tiles;
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
        if (nz > 4 * kMaxKSteps || nj > 4 * kMaxKSteps) fail(SKB_UNSUPPORTED, "vertical dimension above 260");
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
        const size_t me = nj * nz * group_count;
        if (precision == SKB_F64) {
            p->fwd.upload(forward, me * sizeof(double));
            p->inv.upload(inverse, me * sizeof(double));
        } else {
            std::vector<float> f(forward, forward + me), i(inverse, inverse + me);
            p->fwd.upload(f.data(), me * sizeof(float));
            p->inv.upload(i.data(), me * sizeof(float));
        }
        p->tiles.upload(tiles.data(), tiles.size() * sizeof(GemmTile));
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
            launch_gemm<double>(physical, modal, plan->fwd.ptr, plan->tiles.as<GemmTile>(), plan->ntiles, M, K,
                                as_stream(stream));
        else
            launch_gemm<float>(physical, modal, plan->fwd.ptr, plan->tiles.as<GemmTile>(), plan->ntiles, M, K,
                               as_stream(stream));
    });
}

int skb_vertical_inverse(skb_vertical_plan* plan, const void* modal, void* physical, void* stream) {
    return guarded([&] {
        if (!plan || !physical || !modal) fail(SKB_INVALID_ARGUMENT, "null argument");
        const int M = static_cast<int>(plan->w.nz), K = static_cast<int>(plan->nj);
        if (plan->precision == SKB_F64)
            launch_gemm<double>(modal, physical, plan->inv.ptr, plan->tiles.as<GemmTile>(), plan->ntiles, M, K,
                                as_stream(stream));
        else
            launch_gemm<float>(modal, physical, plan->inv.ptr, plan->tiles.as<GemmTile>(), plan->ntiles, M, K,
                               as_stream(stream));
    });
}

uint64_t skb_vertical_matrix_bytes(const skb_vertical_plan* plan) { return plan ? plan->fwd.bytes : 0; }

}  // extern "C"
