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
#include <memory>

#include "common.hpp"
#include "internal.hpp"
#include "plane.hpp"

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
    ~skb_retained_operator() {
        if (vertical) skb_vertical_plan_destroy(vertical);
        if (pruned) skb_pruned_plan_destroy(pruned);
    }
};

using namespace skb;

namespace {

bool plane_path(const skb_retained_operator* op) {
    return op->path != SKB_PATH_TWO_PASS && pruned_plane_tables(op->pruned) != nullptr;
}

void run_stage(skb_retained_operator* op, int stage, const void* retained, void* modal, hipStream_t s) {
    const bool plane = plane_path(op);
    const PlaneTables* t = pruned_plane_tables(op->pruned);
    void* pm = pruned_plane_scratch(op->pruned);
    switch (stage) {
        case SKB_OP_STAGE_LAYOUT:
            if (!retained) fail(SKB_INVALID_ARGUMENT, "null retained input");
            if (plane) pm_transpose(true, retained, pm, t->modes, op->planes, t->np, s);
            return;
        case SKB_OP_STAGE_INVERSE:
            if (plane) {
                plane_inverse(*t, pm, op->real.ptr, op->planes, s);
            } else {
                if (!retained) fail(SKB_INVALID_ARGUMENT, "null retained input");
                check_status(skb_pruned_inverse(op->pruned, retained, op->real.ptr, s));
            }
            return;
        case SKB_OP_STAGE_FORWARD:
            if (plane) {
                plane_forward(*t, op->real.ptr, pm, op->planes, s);
                if (!vertical_pm_supported(op->vertical))
                    pm_transpose(false, pm, op->compact.ptr, t->modes, op->planes, t->np, s);
            } else {
                check_status(skb_pruned_forward(op->pruned, op->real.ptr, op->compact.ptr, s));
            }
            return;
        case SKB_OP_STAGE_TRANSFORMS:
            // both transforms as execute runs them
            if (plane && op->fused) {
                plane_operator(*t, pm, pm, op->real.ptr, op->planes, s);
                if (!vertical_pm_supported(op->vertical))
                    pm_transpose(false, pm, op->compact.ptr, t->modes, op->planes, t->np, s);
            } else {
                run_stage(op, SKB_OP_STAGE_INVERSE, retained, modal, s);
                run_stage(op, SKB_OP_STAGE_FORWARD, retained, modal, s);
            }
            return;
        case SKB_OP_STAGE_VERTICAL:
            if (!modal) fail(SKB_INVALID_ARGUMENT, "null modal output");
            if (plane && vertical_pm_supported(op->vertical)) vertical_forward_pm(op->vertical, pm, t->np, modal, s);
            else check_status(skb_vertical_forward(op->vertical, op->compact.ptr, modal, s));
            return;
    }
    fail(SKB_INVALID_ARGUMENT, "bad operator stage");
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
        const hipStream_t s = as_stream(stream);
        run_stage(op, SKB_OP_STAGE_LAYOUT, retained, modal, s);
        run_stage(op, SKB_OP_STAGE_TRANSFORMS, retained, modal, s);
        run_stage(op, SKB_OP_STAGE_VERTICAL, retained, modal, s);
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
        run_stage(op, stage, retained, modal, as_stream(stream));
    });
}

int skb_retained_operator_set_path(skb_retained_operator* op, int path) {
    return guarded([&] {
        if (!op) fail(SKB_INVALID_ARGUMENT, "null plan");
        check_status(skb_pruned_plan_set_path(op->pruned, path));
        op->path = path;
    });
}

int skb_retained_operator_path(const skb_retained_operator* op) {
    return op && plane_path(op) ? SKB_PATH_PLANE : SKB_PATH_TWO_PASS;
}

void* skb_retained_operator_real(const skb_retained_operator* op) { return op ? op->real.ptr : nullptr; }

uint64_t skb_retained_operator_device_bytes(const skb_retained_operator* op) {
    if (!op) return 0;
    return op->real.bytes + op->compact.bytes + skb_pruned_scratch_bytes(op->pruned) +
           (pruned_plane_scratch(op->pruned) ? plane_np(op->modes) * op->planes * 8 : 0) +
           2 * skb_vertical_matrix_bytes(op->vertical);
}

}  // extern "C"
