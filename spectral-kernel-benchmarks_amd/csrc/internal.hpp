// Entry points shared between the library's source files (not exported):
// the fused physical-space row stage of the flux graph (fft.hip, next to the
// pruned plans whose intermediates it reads and writes) and the vertical
// products the composed graphs call (vertical.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "skb_hip.h"

namespace skb {

// Split <-> interleaved complex arrays of n elements (layout.hip):
// out[i] = (re[i], im[i]); re[i] = scale * in[i].re, im[i] = scale * in[i].im.
// deinterleave skips a null re or im array (not both).
void interleave(int precision, uint64_t n, const void* re, const void* im, void* out, hipStream_t s);
void deinterleave(int precision, uint64_t n, const void* in, void* re, void* im, double scale, hipStream_t s);
// dst[p + dst_stride*f] = src[p + src_stride*f] for f < frequencies, p < planes.
void copy_planes(int precision, uint64_t frequencies, int planes, const void* src, uint64_t src_stride, void* dst,
                 uint64_t dst_stride, hipStream_t s);

struct PlaneTables;
// The pruned plan's plane-resident tables and its plane-major scratch
// (mode + np * plane), or null where that path is not built (fft.hip).
const PlaneTables* pruned_plane_tables(const skb_pruned_plan* p);
void* pruned_plane_scratch(const skb_pruned_plan* p);
int pruned_planes(const skb_pruned_plan* p);

// True when the two pruned plans (3 + 3T input fields, T output fields) share
// a grid the fused kernel handles (FP32, Nx >= 64, same Ny, Nz, k_x band).
bool flux_rows_supported(const skb_pruned_plan* pin, const skb_pruned_plan* pout);
// pout's intermediate <- forward rows of -scale * (u*qx + v*qy + w*qz) per
// target, from the inverse rows of pin's intermediate.
// workgroups > 0 caps the launch (persistent workgroups); 0 = one per item.
void flux_rows(skb_pruned_plan* pin, skb_pruned_plan* pout, int targets, double scale, void* stream,
               int workgroups = 0);

// Vertical forward product of the plan's columns read from a wider compact
// array: column c at in + K*(in_cs*c + in_co) (vertical.hip).
void vertical_forward_columns(const skb_vertical_plan* p, const void* in, int in_cs, int in_co, void* out,
                              void* stream);

// Forward product read from the chunked plane-major retained layout (element
// (mode, z, field) at pm[16*P*(mode/16) + 16*(z + Nz*field) + mode%16], np
// modes, a multiple of 16; slot values, conjugated for the modes flagged in
// cmask, bit m % 16 of word m / 16) into the plan's modal layout (the
// caller's values); square FP32 split-BF16 plans only (vertical.hip).
bool vertical_pm_supported(const skb_vertical_plan* p);
void vertical_forward_pm(const skb_vertical_plan* p, const void* pm, int np, const uint32_t* cmask, void* out,
                         void* stream);
// The same for `count` evaluations in one launch: evaluation e reads the
// block at pm + e * pm_stride (complex elements) and writes out[e].
struct BatchOutputs {
    void* out[kBatchMax];
};
void vertical_forward_pm_batch(const skb_vertical_plan* p, const void* pm, size_t pm_stride, int np,
                               const uint32_t* cmask, const BatchOutputs& out, int count, void* stream);

// Modal round trip in one kernel: out[:, c] = I_g (w[:, c] .* (F_g in[:, c]))
// over a compact plan's columns (weights in the plan's precision, modal
// layout j + Nj*c); out may equal in.  Supported when the kernel is
// instantiated for (Nz, Nj) (vertical.hip).
bool vertical_round_trip_supported(const skb_vertical_plan* p);
void vertical_round_trip(const skb_vertical_plan* p, const void* weights, const void* in, void* out, void* stream);
// out[e] = w[e] * in[e] over n complex elements (in place allowed).
void modal_work(int precision, uint64_t n, const void* weights, const void* in, void* out, void* stream);

// One source field of the inverse column pass of a flux graph: compact
// interleaved values of field `field` of an array with `fields` fields,
// optionally multiplied by a spectral derivative factor.
enum SpectralFactor { kFactorNone = 0, kFactorDx = 1, kFactorDy = 2 };
struct FluxSource {
    const void* base;  // compact array (plane + P*mode, P = nz*fields)
    int fields, field;
    int factor;        // SpectralFactor
};
// pin's intermediate <- inverse column transforms of the given sources, one
// per field of pin (at most 16), with factors i*2pi*kx/lx and i*2pi*l/ly.
void flux_inverse_columns(skb_pruned_plan* pin, const FluxSource* sources, int count, void* stream);

}  // namespace skb
