// The retained-mode operator as one C-ABI plan: pruned inverse c2r of the
// compact retained spectrum into real planes, forward r2c truncated back to
// the retained modes, K^2-grouped vertical forward product into the modal
// layout.  Every kernel of an evaluation is enqueued from here, so a caller
// pays one C-ABI call per evaluation.
//
// Plane path (FP32, 256 x 256 planes, 2/3-rule k_x band): the input is
// copied into the chunked plane-major layout (16*P*(mode/16) + 16*plane +
// mode%16), each plane's inverse and forward 2-D transform runs in one
// workgroup with the band on chip, and the vertical product reads the
// forward's plane-major output directly (square split-BF16 families) or
// through a copy back to the compact layout.
// Two-pass path: the pruned plan's row and column kernels and the compact
// vertical product.
// Batched execute, two forms.  Stage-batched (reserve_batch): up to
// kBatchMax evaluations per group, each stage one launch over the whole
// group (layout copies, plane transforms, vertical products), with per-
// evaluation intermediates; the plane workgroups of successive evaluations
// fill every CU and the vertical product runs in steady state.  Lanes
// (reserve_lanes): evaluations alternate over the plan's lanes (own streams
// and intermediates), forked from and joined back to the caller's stream.
#include <algorithm>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "internal.hpp"
#include "plane.hpp"

namespace skb {
// One lane of a batched execute: a stream of the plan's own and the
// intermediates one evaluation in flight needs (lane 0 uses the plan's).
struct OperatorLane {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    DeviceBuffer pm, real, compact;
    ~OperatorLane() {
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
        if (done) (void)hipEventDestroy(done);
    }
};
// The intermediates one evaluation writes and reads.
struct OperatorBuffers {
    void* pm;
    void* real;
    void* compact;
};
}  // namespace skb

struct skb_retained_operator {
    skb_workload w{};
    int precision = SKB_F32;
    uint64_t modes = 0, nj = 0;
    int planes = 0;
    int path = SKB_PATH_AUTO;
    bool fused = true;  // plane path: both transforms of a plane in one workgroup
    skb_pruned_plan* pruned = nullptr;
    skb_vertical_plan* vertical = nullptr;
    skb::DeviceBuffer real;     // Nx * Ny * P real values
    skb::DeviceBuffer compact;  // the forward's compact output on the two-pass path (and as the vertical input
                                // when the vertical plan does not read the plane-major layout)
    // batched execute (skb_retained_operator_reserve_lanes)
    std::vector<std::unique_ptr<skb::OperatorLane>> lanes;
    hipEvent_t fork = nullptr;
    bool lanes_apart = true;  // every lane on a hardware queue of its own
    skb::DeviceBuffer probe;  // target of the queue check's tiny kernel
    // stage-batched execute (skb_retained_operator_reserve_batch): the
    // plane-major blocks and real planes of batch_cap evaluations
    int batch_cap = 0;
    skb::DeviceBuffer batch_pm, batch_real;
    void drop_lanes() {
        lanes.clear();
        if (fork) (void)hipEventDestroy(fork);
        fork = nullptr;
        lanes_apart = true;
    }
    ~skb_retained_operator() {
        drop_lanes();
        if (vertical) skb_vertical_plan_destroy(vertical);
        if (pruned) skb_pruned_plan_destroy(pruned);
    }
};

using namespace skb;

namespace {

bool plane_path(const skb_retained_operator* op) {
    return op->path != SKB_PATH_TWO_PASS && pruned_plane_tables(op->pruned) != nullptr;
}

OperatorBuffers own_buffers(skb_retained_operator* op) {
    return {pruned_plane_scratch(op->pruned), op->real.ptr, op->compact.ptr};
}

void run_stage(skb_retained_operator* op, int stage, const void* retained, void* modal, hipStream_t s,
               const OperatorBuffers& b) {
    const bool plane = plane_path(op);
    const PlaneTables* t = pruned_plane_tables(op->pruned);
    void* pm = b.pm;
    switch (stage) {
        case SKB_OP_STAGE_LAYOUT:
            if (!retained) fail(SKB_INVALID_ARGUMENT, "null retained input");
            if (plane) pm_transpose(*t, true, retained, pm, op->planes, s);
            return;
        case SKB_OP_STAGE_INVERSE:
            if (plane) {
                plane_inverse(*t, pm, b.real, op->planes, s);
            } else {
                if (!retained) fail(SKB_INVALID_ARGUMENT, "null retained input");
                check_status(skb_pruned_inverse(op->pruned, retained, b.real, s));
            }
            return;
        case SKB_OP_STAGE_FORWARD:
            if (plane) {
                plane_forward(*t, b.real, pm, op->planes, s);
                if (!vertical_pm_supported(op->vertical))
                    pm_transpose(*t, false, pm, b.compact, op->planes, s);
            } else {
                check_status(skb_pruned_forward(op->pruned, b.real, b.compact, s));
            }
            return;
        case SKB_OP_STAGE_TRANSFORMS:
            // both transforms as execute runs them
            if (plane && op->fused) {
                plane_operator(*t, pm, pm, b.real, op->planes, s);
                if (!vertical_pm_supported(op->vertical))
                    pm_transpose(*t, false, pm, b.compact, op->planes, s);
            } else {
                run_stage(op, SKB_OP_STAGE_INVERSE, retained, modal, s, b);
                run_stage(op, SKB_OP_STAGE_FORWARD, retained, modal, s, b);
            }
            return;
        case SKB_OP_STAGE_VERTICAL:
            if (!modal) fail(SKB_INVALID_ARGUMENT, "null modal output");
            if (plane && vertical_pm_supported(op->vertical)) vertical_forward_pm(op->vertical, pm, t->np, t->cmask.as<uint32_t>(), modal, s);
            else check_status(skb_vertical_forward(op->vertical, b.compact, modal, s));
            return;
    }
    fail(SKB_INVALID_ARGUMENT, "bad operator stage");
}

void run_eval(skb_retained_operator* op, const void* retained, void* modal, hipStream_t s, const OperatorBuffers& b) {
    run_stage(op, SKB_OP_STAGE_LAYOUT, retained, modal, s, b);
    run_stage(op, SKB_OP_STAGE_TRANSFORMS, retained, modal, s, b);
    run_stage(op, SKB_OP_STAGE_VERTICAL, retained, modal, s, b);
}

// Queue check of the lanes: a bounded spin (`ticks` of the device's
// constant-rate wall clock) and a 64-word store.
__global__ void lane_spin_kernel(uint64_t ticks) {
    const uint64_t t0 = wall_clock64();
    while (wall_clock64() - t0 < ticks) __builtin_amdgcn_s_sleep(8);
}
__global__ void lane_touch_kernel(int* p) { p[threadIdx.x] = 1; }  // 64 ints

// True when work on stream b waits for work on stream a: the two streams
// share one hardware queue (in-order), so b's store completes only after
// a's spin has ended.
bool shares_queue(hipStream_t a, hipStream_t b, int* probe, uint64_t ticks) {
    hipEvent_t ea = nullptr, eb = nullptr;
    check_hip(hipEventCreateWithFlags(&ea, hipEventDisableTiming), "hipEventCreate");
    check_hip(hipEventCreateWithFlags(&eb, hipEventDisableTiming), "hipEventCreate");
    hipLaunchKernelGGL(lane_spin_kernel, dim3(1), dim3(64), 0, a, ticks);
    check_hip(hipGetLastError(), "lane spin launch");
    check_hip(hipEventRecord(ea, a), "hipEventRecord");
    hipLaunchKernelGGL(lane_touch_kernel, dim3(1), dim3(64), 0, b, probe);
    check_hip(hipGetLastError(), "lane touch launch");
    check_hip(hipEventRecord(eb, b), "hipEventRecord");
    check_hip(hipEventSynchronize(eb), "hipEventSynchronize");
    const bool shared = hipEventQuery(ea) == hipSuccess;
    check_hip(hipStreamSynchronize(a), "hipStreamSynchronize");
    (void)hipEventDestroy(ea);
    (void)hipEventDestroy(eb);
    return shared;
}

}  // namespace

extern "C" {

int skb_retained_operator_create(const skb_workload* workload, const skb_retained_mode* modes, uint64_t mode_count,
                                 uint64_t nj, const skb_mode_group* groups, uint64_t group_count,
                                 const double* forward, const double* inverse, int precision,
                                 skb_retained_operator** op) {
    return guarded([&] {
        if (!op) fail(SKB_INVALID_ARGUMENT, "null plan pointer");
        *op = nullptr;
        check_precision(precision);
        validate_workload(workload);
        auto p = std::make_unique<skb_retained_operator>();
        p->w = *workload;
        p->precision = precision;
        p->modes = mode_count;
        p->nj = nj;
        check_status(skb_pruned_plan_create(workload, modes, mode_count, precision, &p->pruned));
        check_status(skb_vertical_plan_create(workload, nj, groups, group_count, forward, inverse, precision,
                                              &p->vertical));
        p->planes = pruned_planes(p->pruned);
        const uint64_t cb = complex_bytes(precision);
        p->real.allocate(workload->nx * workload->ny * static_cast<uint64_t>(p->planes) * real_bytes(precision));
        p->compact.allocate(mode_count * static_cast<uint64_t>(p->planes) * cb);
        *op = p.release();
    });
}

int skb_retained_operator_destroy(skb_retained_operator* op) {
    return guarded([&] { delete op; });
}

int skb_retained_operator_execute(skb_retained_operator* op, const void* retained, void* modal, void* stream) {
    return guarded([&] {
        if (!op || !retained || !modal) fail(SKB_INVALID_ARGUMENT, "null argument");
        run_eval(op, retained, modal, as_stream(stream), own_buffers(op));
    });
}

int skb_retained_operator_reserve_lanes(skb_retained_operator* op, int lanes) {
    return guarded([&] {
        if (!op) fail(SKB_INVALID_ARGUMENT, "null plan");
        if (lanes < 1 || lanes > 8) fail(SKB_INVALID_ARGUMENT, "lanes must lie in [1, 8]");
        op->drop_lanes();
        if (lanes == 1 || !plane_path(op)) return;
        const PlaneTables* t = pruned_plane_tables(op->pruned);
        const uint64_t pm_bytes = static_cast<uint64_t>(t->np) * op->planes * complex_bytes(op->precision);
        if (!op->probe.ptr) op->probe.allocate(256);
        int device = 0, rate_khz = 0;
        check_hip(hipGetDevice(&device), "hipGetDevice");
        check_hip(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, device), "wall clock rate");
        const uint64_t ticks = static_cast<uint64_t>(rate_khz > 0 ? rate_khz : 100000) * 2;  // 2 ms
        check_hip(hipEventCreateWithFlags(&op->fork, hipEventDisableTiming), "hipEventCreate");
        // Streams are drawn until one sits on a hardware queue no other lane
        // uses (at most 8 draws per lane; the last one is kept otherwise).
        // Rejected streams stay alive until the lanes are chosen: the runtime
        // gives a new stream its least-used queue, which a destroyed stream
        // would hand straight back.
        std::vector<hipStream_t> rejected;
        struct Release {
            std::vector<hipStream_t>& v;
            ~Release() {
                for (hipStream_t st : v) (void)hipStreamDestroy(st);
            }
        } release{rejected};
        for (int k = 0; k < lanes; ++k) {
            auto lane = std::make_unique<OperatorLane>();
            for (int attempt = 0; attempt < 8; ++attempt) {
                hipStream_t st = nullptr;
                check_hip(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
                bool shared = false;
                for (const auto& o : op->lanes)
                    if (shares_queue(o->stream, st, static_cast<int*>(op->probe.ptr), ticks)) shared = true;
                if (!shared || attempt == 7) {
                    if (shared) op->lanes_apart = false;
                    lane->stream = st;
                    break;
                }
                rejected.push_back(st);
            }
            check_hip(hipEventCreateWithFlags(&lane->done, hipEventDisableTiming), "hipEventCreate");
            if (k > 0) {
                lane->pm.allocate(pm_bytes);
                lane->real.allocate(op->real.bytes);
                lane->compact.allocate(op->compact.bytes);
            }
            op->lanes.push_back(std::move(lane));
        }
    });
}

int skb_retained_operator_lanes(const skb_retained_operator* op, int* apart) {
    if (apart) *apart = op && op->lanes_apart ? 1 : 0;
    return op ? std::max<int>(1, static_cast<int>(op->lanes.size())) : 0;
}

int skb_retained_operator_reserve_batch(skb_retained_operator* op, int evaluations) {
    return guarded([&] {
        if (!op) fail(SKB_INVALID_ARGUMENT, "null plan");
        if (evaluations < 0 || evaluations > kBatchMax)
            fail(SKB_INVALID_ARGUMENT, "batched evaluations must lie in [0, " + std::to_string(kBatchMax) + "]");
        op->batch_cap = 0;
        op->batch_pm.reset();
        op->batch_real.reset();
        if (evaluations < 2) return;
        if (!plane_path(op) || !vertical_pm_supported(op->vertical))
            fail(SKB_UNSUPPORTED, "stage-batched execute: plane path with a plane-major vertical product only");
        const PlaneTables* t = pruned_plane_tables(op->pruned);
        op->batch_pm.allocate(static_cast<uint64_t>(evaluations) * t->np * op->planes * complex_bytes(op->precision));
        op->batch_real.allocate(static_cast<uint64_t>(evaluations) * op->real.bytes);
        op->batch_cap = evaluations;
    });
}

int skb_retained_operator_batch(const skb_retained_operator* op) { return op ? op->batch_cap : 0; }

namespace {

bool stage_batched(const skb_retained_operator* op) {
    return op->batch_cap >= 2 && op->fused && plane_path(op) && vertical_pm_supported(op->vertical);
}

// Stage-batched groups of up to batch_cap evaluations: the stages in `stages`
// (bit 0 layout, 1 transforms, 2 vertical), one launch each per group.
void run_batch_groups(skb_retained_operator* op, int count, const void* const* retained, void* const* modal,
                      hipStream_t s, int stages) {
    const PlaneTables* t = pruned_plane_tables(op->pruned);
    const size_t pm_stride = static_cast<size_t>(t->np) * op->planes;
    const size_t real_stride = op->real.bytes / real_bytes(op->precision);
    for (int i0 = 0; i0 < count; i0 += op->batch_cap) {
        const int n = std::min(op->batch_cap, count - i0);
        BatchInputs in{};
        BatchOutputs out{};
        for (int e = 0; e < n; ++e) {
            in.in[e] = retained[i0 + e];
            out.out[e] = modal[i0 + e];
        }
        if (stages & 1) pm_transpose_batch(*t, in, n, op->batch_pm.ptr, pm_stride, op->planes, s);
        if (stages & 2)
            plane_operator(*t, op->batch_pm.ptr, op->batch_pm.ptr, op->batch_real.ptr, op->planes, s, n, pm_stride,
                           real_stride);
        if (stages & 4)
            vertical_forward_pm_batch(op->vertical, op->batch_pm.ptr, pm_stride, t->np, t->cmask.as<uint32_t>(), out,
                                      n, s);
    }
}

void check_batch_args(const skb_retained_operator* op, int count, const void* const* retained, void* const* modal) {
    if (!op || count < 0 || (count > 0 && (!retained || !modal))) fail(SKB_INVALID_ARGUMENT, "null argument");
    for (int i = 0; i < count; ++i)
        if (!retained[i] || !modal[i]) fail(SKB_INVALID_ARGUMENT, "null retained input or modal output");
}

}  // namespace

int skb_retained_operator_execute_batch_stage(skb_retained_operator* op, int stage, int count,
                                              const void* const* retained, void* const* modal, void* stream) {
    return guarded([&] {
        check_batch_args(op, count, retained, modal);
        if (!stage_batched(op)) fail(SKB_LOGIC_ERROR, "no stage-batched execute reserved (reserve_batch)");
        int mask = 0;
        switch (stage) {
            case SKB_OP_STAGE_LAYOUT: mask = 1; break;
            case SKB_OP_STAGE_TRANSFORMS: mask = 2; break;
            case SKB_OP_STAGE_VERTICAL: mask = 4; break;
            default: fail(SKB_INVALID_ARGUMENT, "batched stages are LAYOUT, TRANSFORMS and VERTICAL");
        }
        run_batch_groups(op, count, retained, modal, as_stream(stream), mask);
    });
}

int skb_retained_operator_execute_batch(skb_retained_operator* op, int count, const void* const* retained,
                                        void* const* modal, void* stream) {
    return guarded([&] {
        check_batch_args(op, count, retained, modal);
        const hipStream_t s = as_stream(stream);
        if (stage_batched(op)) {
            run_batch_groups(op, count, retained, modal, s, 7);
            return;
        }
        const int nl = static_cast<int>(op->lanes.size());
        if (nl <= 1 || count <= 1 || !plane_path(op)) {
            for (int i = 0; i < count; ++i) run_eval(op, retained[i], modal[i], s, own_buffers(op));
            return;
        }
        const int used = std::min(nl, count);
        check_hip(hipEventRecord(op->fork, s), "hipEventRecord");
        for (int k = 0; k < used; ++k) check_hip(hipStreamWaitEvent(op->lanes[k]->stream, op->fork, 0), "wait");
        for (int i = 0; i < count; ++i) {
            const int k = i % used;
            const auto& lane = *op->lanes[k];
            const OperatorBuffers b = k == 0 ? own_buffers(op)
                                             : OperatorBuffers{lane.pm.ptr, lane.real.ptr, lane.compact.ptr};
            run_eval(op, retained[i], modal[i], lane.stream, b);
        }
        for (int k = 0; k < used; ++k) {
            check_hip(hipEventRecord(op->lanes[k]->done, op->lanes[k]->stream), "hipEventRecord");
            check_hip(hipStreamWaitEvent(s, op->lanes[k]->done, 0), "wait");
        }
    });
}

int skb_retained_operator_set_fused(skb_retained_operator* op, int fused) {
    return guarded([&] {
        if (!op) fail(SKB_INVALID_ARGUMENT, "null plan");
        op->fused = fused != 0;
    });
}

int skb_retained_operator_fused(const skb_retained_operator* op) { return op && op->fused ? 1 : 0; }

int skb_retained_operator_stage(skb_retained_operator* op, int stage, const void* retained, void* modal,
                                void* stream) {
    return guarded([&] {
        if (!op) fail(SKB_INVALID_ARGUMENT, "null plan");
        run_stage(op, stage, retained, modal, as_stream(stream), own_buffers(op));
    });
}

int skb_retained_operator_set_path(skb_retained_operator* op, int path) {
    return guarded([&] {
        if (!op) fail(SKB_INVALID_ARGUMENT, "null plan");
        check_status(skb_pruned_plan_set_path(op->pruned, path));
        op->path = path;
        if (!plane_path(op)) op->drop_lanes();  // lanes hold plane-path intermediates
    });
}

int skb_retained_operator_path(const skb_retained_operator* op) {
    return op && plane_path(op) ? SKB_PATH_PLANE : SKB_PATH_TWO_PASS;
}

void* skb_retained_operator_real(const skb_retained_operator* op) { return op ? op->real.ptr : nullptr; }

uint64_t skb_retained_operator_device_bytes(const skb_retained_operator* op) {
    if (!op) return 0;
    uint64_t lanes = 0;
    for (const auto& l : op->lanes) lanes += l->pm.bytes + l->real.bytes + l->compact.bytes;
    lanes += op->batch_pm.bytes + op->batch_real.bytes;
    return op->real.bytes + op->compact.bytes + skb_pruned_scratch_bytes(op->pruned) +
           (pruned_plane_scratch(op->pruned) ? plane_np(op->modes) * op->planes * 8 : 0) +
           2 * skb_vertical_matrix_bytes(op->vertical) + lanes;
}

}  // extern "C"
