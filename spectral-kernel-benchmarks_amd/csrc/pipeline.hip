// Synthetic antialiased spectral pipeline (the spectral-pipeline kernel of
// src/benchmark.cpp:2978-3547) on the device.  One evaluation maps a real
// volume to a real volume:
//   r2c -> retained modes -> vertical forward (Nj x Nz per K^2 group)
//   -> real diagonal modal weights -> vertical inverse (Nz x Nj)
//   -> embed into the zero-padded half spectrum -> c2r
// with unnormalised transforms, as the reference's timed sequence
// (src/benchmark.cpp:3389-3403).
//
// Two boundary policies:
//   SKB_PIPELINE_WVM_DIRECT  the reference's wvm-direct policy
//       (src/benchmark.cpp:3204-3403): full WVM-order r2c, per-mode products
//       on the full spectrum, WVM modal work (k_x = 0 partners conjugated),
//       spectrum zeroed, inverse products, full c2r.
//   SKB_PIPELINE_COMPACT     the compact policies (plane-major-fused-split and
//       streaming-pruned-compact-split, src/benchmark.cpp:3405-3547): pruned
//       forward straight into the compact retained array, the modal round
//       trip on the compact columns, and the pruned inverse.  The round trip
//       is one kernel that keeps each tile's modal block in LDS; without an
//       instantiation for (Nz, Nj), or after skb_spectral_pipeline_set_fused
//       (plan, 0), it is three launches (forward GEMM, weights, inverse GEMM)
//       through a modal array.
// The stage-by-stage entry point writes the compact intermediates the
// reference checks (src/benchmark.cpp:3175-3198).
#include <memory>
#include <vector>

#include "common.hpp"
#include "internal.hpp"

struct skb_spectral_pipeline {
    skb_workload w{};
    int policy = SKB_PIPELINE_COMPACT;
    int precision = SKB_F64;
    uint64_t modes = 0, nj = 0;
    std::vector<skb_retained_mode> host_modes;  // mode tables of the WVM gathers
    skb_fft_plan* fft = nullptr;
    skb_pruned_plan* pruned = nullptr;
    skb_vertical_plan* vertical = nullptr;
    skb::DeviceBuffer weights;                     // plan precision, modal layout
    skb::DeviceBuffer compact, modal;              // compact policy
    skb::DeviceBuffer spectrum, modal_out, modal_in;  // WVM policy
    bool fused = false;
    void allocate_modal() {
        if (!modal.ptr) modal.allocate(modes * w.fields * nj * skb::complex_bytes(precision));
    }
    ~skb_spectral_pipeline() {
        if (fft) skb_fft_plan_destroy(fft);
        if (pruned) skb_pruned_plan_destroy(pruned);
        if (vertical) skb_vertical_plan_destroy(vertical);
    }
};

using namespace skb;

namespace {

void run_compact(skb_spectral_pipeline& p, const void* in, void* out, hipStream_t s) {
    check_status(skb_pruned_forward(p.pruned, in, p.compact.ptr, s));
    if (p.fused) {
        vertical_round_trip(p.vertical, p.weights.ptr, p.compact.ptr, p.compact.ptr, s);
    } else {
        check_status(skb_vertical_forward(p.vertical, p.compact.ptr, p.modal.ptr, s));
        modal_work(p.precision, p.modes * p.w.fields * p.nj, p.weights.ptr, p.modal.ptr, p.modal.ptr, s);
        check_status(skb_vertical_inverse(p.vertical, p.modal.ptr, p.compact.ptr, s));
    }
    check_status(skb_pruned_inverse(p.pruned, p.compact.ptr, out, s));
}

void run_wvm(skb_spectral_pipeline& p, const void* in, void* out, hipStream_t s) {
    check_status(skb_fft_forward(p.fft, in, p.spectrum.ptr, s));
    check_status(skb_vertical_wvm_forward(p.vertical, p.spectrum.ptr, p.modal_out.ptr, s));
    check_status(skb_vertical_wvm_modal_work(p.vertical, p.weights.ptr, p.modal_out.ptr, p.modal_in.ptr, s));
    check_status(skb_vertical_wvm_initialize_spectrum(p.vertical, p.spectrum.ptr, s));
    check_status(skb_vertical_wvm_inverse(p.vertical, p.modal_in.ptr, p.spectrum.ptr, s));
    check_status(skb_fft_inverse(p.fft, p.spectrum.ptr, out, s));
}

}  // namespace

extern "C" {

int skb_spectral_pipeline_create(const skb_workload* workload, const skb_retained_mode* modes, uint64_t mode_count,
                                 const skb_mode_group* groups, uint64_t group_count, const double* forward,
                                 const double* inverse, const double* weights, int policy, int precision,
                                 skb_spectral_pipeline** plan) {
    return guarded([&] {
        if (!plan) fail(SKB_INVALID_ARGUMENT, "null plan pointer");
        *plan = nullptr;
        check_precision(precision);
        validate_workload(workload);
        if (policy != SKB_PIPELINE_WVM_DIRECT && policy != SKB_PIPELINE_COMPACT)
            fail(SKB_INVALID_ARGUMENT,
                 "spectral-pipeline boundary-policy must be wvm-direct or a compact policy.");
        if (!modes || mode_count == 0) fail(SKB_INVALID_ARGUMENT, "the spectral pipeline requires retained modes");
        const uint64_t nj = workload->nz < 2 ? 0 : 2 * (workload->nz - 1) / 3;
        if (nj == 0) fail(SKB_INVALID_ARGUMENT, "the spectral pipeline requires Nz >= 3 (Nj >= 1)");
        auto p = std::make_unique<skb_spectral_pipeline>();
        p->w = *workload;
        p->policy = policy;
        p->precision = precision;
        p->modes = mode_count;
        p->nj = nj;
        const uint64_t F = workload->fields, P = workload->nz * F, cb = complex_bytes(precision);
        const uint64_t nw = mode_count * F * nj;
        std::vector<double> wd;
        if (!weights) {
            wd.resize(nw);
            check_status(skb_synthetic_modal_weights(workload, modes, mode_count, wd.data()));
            weights = wd.data();
        }
        if (precision == SKB_F64) {
            p->weights.upload(weights, nw * sizeof(double));
        } else {
            std::vector<float> wf(weights, weights + nw);
            p->weights.upload(wf.data(), nw * sizeof(float));
        }
        if (policy == SKB_PIPELINE_WVM_DIRECT) {
            p->host_modes.assign(modes, modes + mode_count);
            check_status(skb_fft_plan_create(workload, precision, &p->fft));
            check_status(skb_vertical_wvm_plan_create(workload, modes, mode_count, nj, groups, group_count, forward,
                                                      inverse, precision, &p->vertical));
            const uint64_t freqs = (workload->nx / 2 + 1) * workload->ny;
            p->spectrum.allocate(freqs * P * cb);
            p->modal_out.allocate(freqs * F * nj * cb);
            p->modal_in.allocate(freqs * F * nj * cb);
            // only retained frequencies (and k_x = 0 partners) are ever written
            check_status(skb_vertical_wvm_initialize_modal(p->vertical, p->modal_out.ptr, nullptr));
            check_status(skb_vertical_wvm_initialize_modal(p->vertical, p->modal_in.ptr, nullptr));
            check_hip(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
        } else {
            check_status(skb_pruned_plan_create(workload, modes, mode_count, precision, &p->pruned));
            check_status(skb_vertical_plan_create(workload, nj, groups, group_count, forward, inverse, precision,
                                                  &p->vertical));
            if (skb_vertical_plan_items(p->vertical) == 0) fail(SKB_INVALID_ARGUMENT, "empty vertical family");
            uint64_t covered = 0;
            for (uint64_t g = 0; g < group_count; ++g) covered += groups[g].mode_count;
            if (covered != mode_count)
                fail(SKB_INVALID_ARGUMENT, "the vertical groups must cover every retained mode");
            p->compact.allocate(mode_count * P * cb);
            p->fused = vertical_round_trip_supported(p->vertical);
            if (!p->fused) p->allocate_modal();
        }
        *plan = p.release();
    });
}

int skb_spectral_pipeline_destroy(skb_spectral_pipeline* plan) {
    return guarded([&] { delete plan; });
}

int skb_spectral_pipeline_set_fused(skb_spectral_pipeline* plan, int fused) {
    return guarded([&] {
        if (!plan) fail(SKB_INVALID_ARGUMENT, "null plan");
        if (plan->policy != SKB_PIPELINE_COMPACT) {
            if (fused) fail(SKB_UNSUPPORTED, "the WVM-direct policy has no fused modal round trip");
            return;
        }
        if (fused && !vertical_round_trip_supported(plan->vertical))
            fail(SKB_UNSUPPORTED, "modal round trip not instantiated for this vertical size");
        if (!fused) plan->allocate_modal();
        plan->fused = fused != 0;
    });
}

int skb_spectral_pipeline_fused(const skb_spectral_pipeline* plan) { return plan && plan->fused ? 1 : 0; }

uint64_t skb_spectral_pipeline_device_bytes(const skb_spectral_pipeline* plan) {
    if (!plan) return 0;
    uint64_t b = plan->weights.bytes + plan->compact.bytes + plan->modal.bytes + plan->spectrum.bytes +
                 plan->modal_out.bytes + plan->modal_in.bytes + 2 * skb_vertical_matrix_bytes(plan->vertical);
    if (plan->pruned) b += skb_pruned_scratch_bytes(plan->pruned);
    return b;
}

int skb_spectral_pipeline_execute(skb_spectral_pipeline* plan, const void* real_input, void* real_output,
                                  void* stream) {
    return guarded([&] {
        if (!plan || !real_input || !real_output) fail(SKB_INVALID_ARGUMENT, "null argument");
        const hipStream_t s = as_stream(stream);
        if (plan->policy == SKB_PIPELINE_WVM_DIRECT) run_wvm(*plan, real_input, real_output, s);
        else run_compact(*plan, real_input, real_output, s);
    });
}

int skb_spectral_pipeline_stages(skb_spectral_pipeline* plan, const void* real_input, void* real_output,
                                 void* retained, void* forward_modal, void* post_modal, void* inverse_retained,
                                 void* stream) {
    return guarded([&] {
        if (!plan || !real_input || !real_output || !retained || !forward_modal || !post_modal || !inverse_retained)
            fail(SKB_INVALID_ARGUMENT, "null argument");
        auto& p = *plan;
        const hipStream_t s = as_stream(stream);
        if (p.policy == SKB_PIPELINE_COMPACT) {
            check_status(skb_pruned_forward(p.pruned, real_input, retained, s));
            check_status(skb_vertical_forward(p.vertical, retained, forward_modal, s));
            modal_work(p.precision, p.modes * p.w.fields * p.nj, p.weights.ptr, forward_modal, post_modal, s);
            check_status(skb_vertical_inverse(p.vertical, post_modal, inverse_retained, s));
            check_status(skb_pruned_inverse(p.pruned, inverse_retained, real_output, s));
            return;
        }
        const skb_retained_mode* m = p.host_modes.data();
        check_status(skb_fft_forward(p.fft, real_input, p.spectrum.ptr, s));
        check_status(skb_fft_gather_retained(p.fft, m, p.modes, p.spectrum.ptr, retained, s));
        check_status(skb_vertical_wvm_forward(p.vertical, p.spectrum.ptr, p.modal_out.ptr, s));
        check_status(skb_fft_gather_retained_modal(p.fft, m, p.modes, p.nj, p.modal_out.ptr, forward_modal, s));
        check_status(skb_vertical_wvm_modal_work(p.vertical, p.weights.ptr, p.modal_out.ptr, p.modal_in.ptr, s));
        check_status(skb_fft_gather_retained_modal(p.fft, m, p.modes, p.nj, p.modal_in.ptr, post_modal, s));
        check_status(skb_vertical_wvm_initialize_spectrum(p.vertical, p.spectrum.ptr, s));
        check_status(skb_vertical_wvm_inverse(p.vertical, p.modal_in.ptr, p.spectrum.ptr, s));
        check_status(skb_fft_gather_retained(p.fft, m, p.modes, p.spectrum.ptr, inverse_retained, s));
        check_status(skb_fft_inverse(p.fft, p.spectrum.ptr, real_output, s));
    });
}

}  // extern "C"
