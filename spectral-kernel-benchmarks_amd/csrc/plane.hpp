// Plane-resident transforms of the retained operator (plane.hip): shapes,
// per-plan device tables and launchers, shared with the plan code of fft.hip
// and the operator of operator.hip.
#pragma once

#include <vector>

#include "common.hpp"

namespace skb {

// The one shape the plane-resident kernels are built for: 256 x 256 planes,
// active k_x band 86 (the 2/3 rule), FP32.
constexpr int kPlaneKc = 86;
constexpr int kPlaneMaxModes = 11440;  // retained values staged in one workgroup's LDS
constexpr int kPlaneMaxPlanes = 480;   // LDS tile of the layout transpose

// Modes of the chunked plane-major retained layout (modes rounded up to 16):
// mode m of plane p at 16*P*(m / 16) + 16*p + m % 16, for P planes.
uint64_t plane_np(uint64_t modes);
bool plane_supported(const skb_workload& w, uint64_t kc, uint64_t modes, int precision);
// The mode set leaves the k_y slots the plane kernels skip empty (k_y in
// [96, 159] for every active k_x): true for every 2/3-rule set at 256^2.
// fmap / imap: dense [kx * Ny + ky] -> code or -1 (see PlaneTables).
bool plane_band_empty(const std::vector<int>& fmap, const std::vector<int>& imap, uint64_t kc);

// Device tables of one mode set: twiddles, the packed 16-bit column codes
// (staged index per (kx, ky) slot, the k_x = 0 partners flagged) in the order
// each kernel's lanes read them, and the modes' conj flags.  fmap / imap:
// dense [kx * Ny + ky] -> 2*mode + conj or -1, forward (stored modes) and
// inverse (k_x = 0 partners added).
struct PlaneTables {
    DeviceBuffer tw, fcodes, icodes;
    DeviceBuffer cmask;  // per 16-mode chunk: bit m % 16 set when mode m is stored conjugated
    int modes = 0, np = 0;
    int workgroups = 0;  // persistent operator workgroups: the device's CUs (one plane workgroup fits a CU)
    void build(const skb_workload& w, uint64_t kc, uint64_t modes, const std::vector<int>& fmap,
               const std::vector<int>& imap);
};

// pm: chunked plane-major retained values (np * planes), real: x + Nx(y + Ny*plane).
void plane_inverse(const PlaneTables& t, const void* pm, void* real, int planes, hipStream_t s);
void plane_forward(const PlaneTables& t, const void* real, void* pm, int planes, hipStream_t s);
// Both transforms of each plane in one workgroup: pm_in -> real planes ->
// pm_out (pm_out may equal pm_in), for `count` evaluations whose arrays sit
// pm_stride complex / real_stride real elements apart.
void plane_operator(const PlaneTables& t, const void* pm_in, void* pm_out, void* real, int planes, hipStream_t s,
                    int count = 1, size_t pm_stride = 0, size_t real_stride = 0);
// compact (plane + P*mode, the caller's values) -> chunked plane-major (slot
// values) (to_pm) or back.
void pm_transpose(const PlaneTables& t, bool to_pm, const void* in, void* out, int planes, hipStream_t s);
// The compact inputs of the evaluations of one batched launch.
struct BatchInputs {
    const void* in[kBatchMax];
};
// compact -> chunked plane-major for `count` evaluations: in.in[e] -> out +
// e * out_stride (complex elements); 16-byte aligned arrays.
void pm_transpose_batch(const PlaneTables& t, const BatchInputs& in, int count, void* out, size_t out_stride,
                        int planes, hipStream_t s);

}  // namespace skb
