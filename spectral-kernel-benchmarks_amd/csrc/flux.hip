// Antialiased nonlinear flux graph: the production-lifetime flux of
// src/benchmark.cpp:3957-4234, composed from the provider's plans.
//
// Inputs are 3 + 3T modal fields [U, V, W, q0x, q0y, q0z, ..., q(T-1)z]
// (modal layout j + Nj*field + Nj*F*mode); outputs are T modal fields.  Per
// evaluation:
//   1. vertical inverse of all input fields     modal -> compact physical
//   2. pruned horizontal inverse                 compact -> real volumes
//   3. per target t: target = -(u*qtx + v*qty + w*qtz) / (Nx*Ny)^2
//      (both inverses are unnormalised; src/benchmark.cpp:4040-4041)
//   4. pruned horizontal forward (raw)           real -> compact
//   5. vertical forward of the T outputs         compact -> modal
// Steps 2-4 run either as separate kernels (real volumes in HBM) or fused
// into one row pass that keeps the physical fields in registers
// (skb_flux_set_fused).
//
// The advection flux of one tracer (skb_advection_flux_*) is the same graph
// with spectral inputs: fields (u, v, w, q) in the compact retained layout;
// d/dx q and d/dy q are spectral (i*2pi*k/lx, i*2pi*l/ly, applied inside the
// inverse column pass), d/dz q = D q through an Nz x Nz matrix on the
// vertical GEMM; one output field -(u q_x + v q_y + w q_z)/(Nx*Ny)^2,
// forward-transformed and truncated to the retained modes.
#include <algorithm>
#include <memory>

#include "common.hpp"
#include "internal.hpp"

struct skb_flux_plan {
    skb_workload w{};  // fields = targets (the output workload)
    int precision = SKB_F64;
    uint64_t targets = 0, modes = 0, nj = 0;
    skb_vertical_plan* vin = nullptr;   // 3 + 3T fields, inverse direction
    skb_vertical_plan* vout = nullptr;  // T fields, forward direction
    skb_pruned_plan* pin = nullptr;     // 3 + 3T fields
    skb_pruned_plan* pout = nullptr;    // T fields
    skb::DeviceBuffer phys_in, phys_out, real_in, real_out;
    bool fused = true;
    // Real volumes of the separate-kernel path, allocated only when that path
    // can run (at creation, or when skb_flux_set_fused selects it).
    void allocate_real() {
        if (real_in.ptr) return;
        const uint64_t volume = w.nx * w.ny * w.nz, rb = skb::real_bytes(precision);
        real_in.allocate(volume * (3 + 3 * targets) * rb);
        real_out.allocate(volume * targets * rb);
    }
    ~skb_flux_plan() {
        if (vin) skb_vertical_plan_destroy(vin);
        if (vout) skb_vertical_plan_destroy(vout);
        if (pin) skb_pruned_plan_destroy(pin);
        if (pout) skb_pruned_plan_destroy(pout);
    }
};

using namespace skb;

struct skb_advection_flux_plan {
    skb_workload w{};  // fields = 4: u, v, w, q
    int precision = SKB_F64;
    uint64_t modes = 0;
    skb_vertical_plan* dz = nullptr;  // one group holding D, fields = 1
    skb_pruned_plan* pin = nullptr;   // 6 fields: u, v, w, q_x, q_y, q_z
    skb_pruned_plan* pout = nullptr;  // 1 field
    skb::DeviceBuffer qz, real_in, real_out;
    bool fused = true;
    int workgroups = 0;  // launch cap of every kernel (0 = one workgroup per work item)
    void allocate_real() {
        if (real_in.ptr) return;
        const uint64_t volume = w.nx * w.ny * w.nz, rb = skb::real_bytes(precision);
        real_in.allocate(6 * volume * rb);
        real_out.allocate(volume * rb);
    }
    ~skb_advection_flux_plan() {
        if (dz) skb_vertical_plan_destroy(dz);
        if (pin) skb_pruned_plan_destroy(pin);
        if (pout) skb_pruned_plan_destroy(pout);
    }
};

extern "C" {

int skb_flux_plan_create(const skb_workload* workload, const skb_retained_mode* modes, uint64_t mode_count,
                         uint64_t nj, const skb_mode_group* groups, uint64_t group_count, const double* forward,
                         const double* inverse, int precision, skb_flux_plan** plan) {
    return guarded([&] {
        if (!plan) fail(SKB_INVALID_ARGUMENT, "null plan pointer");
        *plan = nullptr;
        check_precision(precision);
        validate_workload(workload);
        if (workload->nx != workload->ny)
            fail(SKB_INVALID_ARGUMENT, "production-lifetime-flux requires a square horizontal grid.");
        if (!modes || mode_count == 0) fail(SKB_INVALID_ARGUMENT, "the flux graph requires retained modes");
        const uint64_t targets = workload->fields;
        if (3 + 3 * targets > 48) fail(SKB_UNSUPPORTED, "at most 15 flux targets");
        auto p = std::make_unique<skb_flux_plan>();
        p->w = *workload;
        p->precision = precision;
        p->targets = targets;
        p->modes = mode_count;
        p->nj = nj;
        skb_workload win = *workload;
        win.fields = 3 + 3 * targets;
        check_status(skb_vertical_plan_create(&win, nj, groups, group_count, forward, inverse, precision, &p->vin));
        check_status(skb_vertical_plan_create(workload, nj, groups, group_count, forward, inverse, precision,
                                              &p->vout));
        if (skb_vertical_plan_items(p->vin) == 0) fail(SKB_INVALID_ARGUMENT, "empty vertical family");
        check_status(skb_pruned_plan_create(&win, modes, mode_count, precision, &p->pin));
        check_status(skb_pruned_plan_create(workload, modes, mode_count, precision, &p->pout));
        const uint64_t cb = complex_bytes(precision);
        p->phys_in.allocate(mode_count * workload->nz * win.fields * cb);
        p->phys_out.allocate(mode_count * workload->nz * targets * cb);
        if (!flux_rows_supported(p->pin, p->pout)) p->allocate_real();
        *plan = p.release();
    });
}

int skb_flux_plan_destroy(skb_flux_plan* plan) {
    return guarded([&] { delete plan; });
}

int skb_flux_set_fused(skb_flux_plan* plan, int fused) {
    return guarded([&] {
        if (!plan) fail(SKB_INVALID_ARGUMENT, "null plan");
        if (!fused) plan->allocate_real();
        plan->fused = fused != 0;
    });
}

int skb_flux_execute(skb_flux_plan* plan, const void* input_modal, void* output_modal, void* stream) {
    return guarded([&] {
        if (!plan || !input_modal || !output_modal) fail(SKB_INVALID_ARGUMENT, "null argument");
        auto& p = *plan;
        const auto& w = p.w;
        const double scale = 1.0 / (static_cast<double>(w.nx * w.ny) * static_cast<double>(w.nx * w.ny));
        check_status(skb_vertical_inverse(p.vin, input_modal, p.phys_in.ptr, stream));
        if (p.fused && flux_rows_supported(p.pin, p.pout)) {
            check_status(skb_pruned_inverse_columns(p.pin, p.phys_in.ptr, stream));
            flux_rows(p.pin, p.pout, static_cast<int>(p.targets), scale, stream);
            check_status(skb_pruned_forward_columns(p.pout, p.phys_out.ptr, stream));
        } else {
            const uint64_t volume = w.nx * w.ny * w.nz;
            const size_t rb = real_bytes(p.precision);
            check_status(skb_pruned_inverse(p.pin, p.phys_in.ptr, p.real_in.ptr, stream));
            for (uint64_t t = 0; t < p.targets; ++t) {
                const char* base = static_cast<const char*>(p.real_in.ptr);
                check_status(skb_pointwise_advection(volume, scale, base, base + (3 + 3 * t) * volume * rb,
                                                     static_cast<char*>(p.real_out.ptr) + t * volume * rb,
                                                     p.precision, stream));
            }
            check_status(skb_pruned_forward(p.pout, p.real_out.ptr, p.phys_out.ptr, stream));
        }
        check_status(skb_vertical_forward(p.vout, p.phys_out.ptr, output_modal, stream));
    });
}

}  // extern "C"

extern "C" {

int skb_vertical_derivative_matrix(uint64_t nz, double lz, double* d) {
    return guarded([&] {
        if (!d || nz < 3 || !(lz > 0.0)) fail(SKB_INVALID_ARGUMENT, "derivative matrix needs nz >= 3 and lz > 0");
        const double h = lz / static_cast<double>(nz - 1);
        std::fill_n(d, nz * nz, 0.0);
        d[0] = -1.5 / h;
        d[1] = 2.0 / h;
        d[2] = -0.5 / h;
        for (uint64_t k = 1; k + 1 < nz; ++k) {
            d[k * nz + k - 1] = -0.5 / h;
            d[k * nz + k + 1] = 0.5 / h;
        }
        d[(nz - 1) * nz + nz - 3] = 0.5 / h;
        d[(nz - 1) * nz + nz - 2] = -2.0 / h;
        d[(nz - 1) * nz + nz - 1] = 1.5 / h;
    });
}

int skb_advection_flux_plan_create(const skb_workload* workload, const skb_retained_mode* modes, uint64_t mode_count,
                                   const double* dz, int precision, skb_advection_flux_plan** plan) {
    return guarded([&] {
        if (!plan) fail(SKB_INVALID_ARGUMENT, "null plan pointer");
        *plan = nullptr;
        check_precision(precision);
        validate_workload(workload);
        if (workload->fields != 4) fail(SKB_INVALID_ARGUMENT, "the advection flux takes four fields: u, v, w, q");
        if (!modes || mode_count == 0) fail(SKB_INVALID_ARGUMENT, "the flux graph requires retained modes");
        if (!dz) fail(SKB_INVALID_ARGUMENT, "null vertical derivative matrix");
        auto p = std::make_unique<skb_advection_flux_plan>();
        p->w = *workload;
        p->precision = precision;
        p->modes = mode_count;
        skb_workload w1 = *workload, w6 = *workload;
        w1.fields = 1;
        w6.fields = 6;
        const skb_mode_group all{0, 0, mode_count};
        check_status(skb_vertical_plan_create(&w1, workload->nz, &all, 1, dz, dz, precision, &p->dz));
        check_status(skb_pruned_plan_create(&w6, modes, mode_count, precision, &p->pin));
        check_status(skb_pruned_plan_create(&w1, modes, mode_count, precision, &p->pout));
        p->qz.allocate(mode_count * workload->nz * complex_bytes(precision));
        if (!flux_rows_supported(p->pin, p->pout)) p->allocate_real();
        *plan = p.release();
    });
}

int skb_advection_flux_plan_destroy(skb_advection_flux_plan* plan) {
    return guarded([&] { delete plan; });
}

int skb_advection_flux_set_fused(skb_advection_flux_plan* plan, int fused) {
    return guarded([&] {
        if (!plan) fail(SKB_INVALID_ARGUMENT, "null plan");
        if (!fused) plan->allocate_real();
        plan->fused = fused != 0;
    });
}

// Caps every kernel of the plan at `workgroups` workgroups (persistent
// workgroups loop over the work items; 0 restores one per item).  Results
// are unchanged: each work item is computed the same way.
int skb_advection_flux_set_workgroups(skb_advection_flux_plan* plan, int workgroups) {
    return guarded([&] {
        if (!plan) fail(SKB_INVALID_ARGUMENT, "null plan");
        if (workgroups < 0) fail(SKB_INVALID_ARGUMENT, "workgroups must be >= 0");
        for (int stage = 0; stage < 4; ++stage) {
            check_status(skb_pruned_plan_set_grid(plan->pin, stage, workgroups));
            check_status(skb_pruned_plan_set_grid(plan->pout, stage, workgroups));
        }
        check_status(skb_vertical_plan_set_grid(plan->dz, workgroups));
        plan->workgroups = workgroups;
    });
}

// Device bytes owned by the plan (its sub-plans' scratch included).
uint64_t skb_advection_flux_device_bytes(const skb_advection_flux_plan* plan) {
    if (!plan) return 0;
    return plan->qz.bytes + plan->real_in.bytes + plan->real_out.bytes + skb_pruned_scratch_bytes(plan->pin) +
           skb_pruned_scratch_bytes(plan->pout);
}

int skb_advection_flux_execute(skb_advection_flux_plan* plan, const void* spectral, void* flux, void* stream) {
    return guarded([&] {
        if (!plan || !spectral || !flux) fail(SKB_INVALID_ARGUMENT, "null argument");
        auto& p = *plan;
        const auto& w = p.w;
        const double scale = 1.0 / (static_cast<double>(w.nx * w.ny) * static_cast<double>(w.nx * w.ny));
        vertical_forward_columns(p.dz, spectral, 4, 3, p.qz.ptr, stream);
        const FluxSource sources[6] = {{spectral, 4, 0, kFactorNone}, {spectral, 4, 1, kFactorNone},
                                       {spectral, 4, 2, kFactorNone}, {spectral, 4, 3, kFactorDx},
                                       {spectral, 4, 3, kFactorDy},   {p.qz.ptr, 1, 0, kFactorNone}};
        flux_inverse_columns(p.pin, sources, 6, stream);
        if (p.fused && flux_rows_supported(p.pin, p.pout)) {
            flux_rows(p.pin, p.pout, 1, scale, stream, p.workgroups);
            check_status(skb_pruned_forward_columns(p.pout, flux, stream));
        } else {
            const uint64_t volume = w.nx * w.ny * w.nz;
            const size_t rb = real_bytes(p.precision);
            check_status(skb_pruned_inverse_rows(p.pin, p.real_in.ptr, stream));
            const char* base = static_cast<const char*>(p.real_in.ptr);
            check_status(skb_pointwise_advection(volume, scale, base, base + 3 * volume * rb, p.real_out.ptr,
                                                 p.precision, stream));
            check_status(skb_pruned_forward(p.pout, p.real_out.ptr, flux, stream));
        }
    });
}

}  // extern "C"
