"""Flux graph (production-lifetime flux, src/benchmark.cpp:3957-4234)
through the C ABI against the oracle's restatement of the reference's
serial oracle (src/benchmark.cpp:4182-4234), which tests/test_oracle.py pins
(direct-DFT vs pocketfft composition, DC closed form).

Checks mirror the reference's record: per-target modal output against the
oracle (FP64 1e-12, FP32 1e-6 of the FP64 oracle) and the DC boundary
staying real.  The fused physical-space pass (inverse rows, products and
forward rows in registers) and the separate-kernel path must agree."""
import sys

import numpy as np
import pytest

import oracle as O
import skb_hip as S

pytestmark = pytest.mark.gpu


def note(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def host(t):
    import torch

    torch.cuda.synchronize()
    return t.detach().cpu().numpy().astype(np.complex128)


def build(shape, precision):
    w = S.Workload(*shape)
    ow = O.Workload(*shape)
    modes = S.retained_modes(w)
    groups = S.k2_groups(modes)
    nj = w.nj
    fwd, inv = S.k2_vertical_family(w.nz, nj, groups)
    plan = S.FluxPlan(w, modes, nj, groups, fwd, inv, precision)
    om = O.retained_modes(ow)
    x = O.flux_modal_fixture(ow, om, nj, 3 + 3 * w.fields, 2026)
    return w, ow, om, groups, nj, fwd, inv, plan, x


def run(plan, x, precision, device, fused):
    import torch

    cd = S.complex_dtype(precision)
    d_in = torch.from_numpy(x).to(device=device, dtype=cd)
    out = torch.empty(plan.output_elements, dtype=cd, device=device)
    plan.set_fused(fused)
    plan.execute(d_in, out)
    return host(out), host(d_in)


def dc_imag(out, om, w, nj):
    o = out.reshape(len(om), w.fields, nj)
    dc = (om["k"] == 0) & (om["l"] == 0)
    return np.abs(o[dc].imag).max()


@pytest.mark.parametrize("precision,shape", [(S.SKB_F64, (16, 16, 7, 2)), (S.SKB_F64, (64, 64, 9, 4)),
                                             (S.SKB_F32, (64, 64, 9, 4)), (S.SKB_F64, (128, 128, 17, 1))])
def test_flux_against_oracle_fused_and_separate(device, precision, shape):
    note("flux %s p=%d" % (shape, precision))
    w, ow, om, groups, nj, fwd, inv, plan, x = build(shape, precision)
    tol = 1e-12 if precision == S.SKB_F64 else 1e-6
    outs = {}
    for fused in (True, False):
        out, x_used = run(plan, x, precision, device, fused)
        ref = O.production_flux(ow, om, nj, groups["first_mode"], groups["mode_count"], fwd, inv, x_used, w.fields)
        for t in range(w.fields):
            got_t = out.reshape(len(om), w.fields, nj)[:, t]
            ref_t = ref.reshape(len(om), w.fields, nj)[:, t]
            assert O.max_relative_error(got_t, ref_t) <= tol, (fused, t)
        assert dc_imag(out, om, w, nj) == 0.0
        outs[fused] = out
    assert O.max_relative_error(outs[True], outs[False]) <= (1e-13 if precision == S.SKB_F64 else 1e-6)


@pytest.mark.parametrize("precision", [S.SKB_F32, S.SKB_F64])
def test_flux_256_nz65_four_targets(device, precision):
    note("flux 256^2 x 65, 4 targets p=%d" % precision)
    w, ow, om, groups, nj, fwd, inv, plan, x = build((256, 256, 65, 4), precision)
    out, x_used = run(plan, x, precision, device, True)
    ref = O.production_flux(ow, om, nj, groups["first_mode"], groups["mode_count"], fwd, inv, x_used, w.fields)
    tol = 1e-12 if precision == S.SKB_F64 else 1e-6
    assert O.max_relative_error(out, ref) <= tol
    assert dc_imag(out, om, w, nj) == 0.0


def test_flux_rejects_non_square_grid(device):
    w = S.Workload(32, 16, 5, 1)
    modes = S.retained_modes(w)
    groups = S.k2_groups(modes)
    fwd, inv = S.k2_vertical_family(w.nz, w.nj, groups)
    with pytest.raises(S.InvalidArgument, match="square"):
        S.FluxPlan(w, modes, w.nj, groups, fwd, inv, S.SKB_F64)
