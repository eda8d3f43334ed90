// Fused physical-space row stage of the flux graph, implemented in fft.hip
// next to the pruned plans whose intermediates it reads and writes.
#pragma once

#include "skb_hip.h"

namespace skb {

// True when the two pruned plans (3 + 3T input fields, T output fields) share
// a grid the fused kernel handles (Nx >= 64, same Ny, Nz, k_x band, precision).
bool flux_rows_supported(const skb_pruned_plan* pin, const skb_pruned_plan* pout);
// pout's intermediate <- forward rows of -scale * (u*qx + v*qy + w*qz) per
// target, from the inverse rows of pin's intermediate.
void flux_rows(skb_pruned_plan* pin, skb_pruned_plan* pout, int targets, double scale, void* stream);

}  // namespace skb
