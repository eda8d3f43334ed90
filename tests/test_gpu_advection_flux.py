"""Advection flux of one tracer (compact retained u, v, w, q -> compact
retained flux) through the C ABI against the oracle (oracle.advection_flux,
pinned in tests/test_oracle.py by closed forms for horizontal and vertical
advection and by the direct DFT).  FP64 within 1e-12, FP32 within 1e-6 of the
FP64 oracle; fused and separate physical-space stages must agree.  The
512^2 x Nz129 case is the full-size flux workload."""
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


def run_case(shape, precision, device, fused_modes=(True, False), lengths=(1.0, 1.0)):
    import torch

    w = S.Workload(*shape, lx=lengths[0], ly=lengths[1])
    ow = O.Workload(*shape, lx=lengths[0], ly=lengths[1])
    modes = S.retained_modes(w)
    om = O.retained_modes(ow)
    dz = S.vertical_derivative_matrix(w.nz, 1.0)
    assert np.array_equal(dz, O.vertical_derivative_matrix(w.nz, 1.0))
    plan = S.AdvectionFluxPlan(w, modes, dz, precision)
    spec = S.spectral_fixture(w, modes, 129, -5.0 / 6.0)  # ~|k|^(-5/3) amplitudes
    cd = S.complex_dtype(precision)
    d_spec = torch.from_numpy(spec).to(device=device, dtype=cd)
    outs = {}
    for fused in fused_modes:
        out = torch.empty(len(modes) * w.nz, dtype=cd, device=device)
        plan.set_fused(fused)
        plan.execute(d_spec, out)
        outs[fused] = host(out)
    ref = O.advection_flux(ow, om, dz, host(d_spec))
    return outs, ref


@pytest.mark.parametrize("precision,shape", [(S.SKB_F64, (16, 16, 9, 4)), (S.SKB_F64, (64, 64, 17, 4)),
                                             (S.SKB_F32, (64, 64, 17, 4)), (S.SKB_F32, (128, 128, 33, 4))])
def test_advection_flux_against_oracle(device, precision, shape):
    note("advection flux %s p=%d" % (shape, precision))
    outs, ref = run_case(shape, precision, device, lengths=(2.0, 0.75))
    tol = 1e-12 if precision == S.SKB_F64 else 1e-6
    for fused, out in outs.items():
        assert O.max_relative_error(out, ref) <= tol, fused
    assert O.max_relative_error(outs[True], outs[False]) <= tol


def test_advection_flux_512_nz129_fp32(device):
    note("advection flux 512^2 x 129 fp32 fused")
    outs, ref = run_case((512, 512, 129, 4), S.SKB_F32, device, fused_modes=(True,))
    assert O.max_relative_error(outs[True], ref) <= 1e-6


@pytest.mark.parametrize("precision,shape", [(S.SKB_F32, (256, 256, 65, 4)), (S.SKB_F64, (64, 64, 17, 4))])
def test_advection_flux_workgroup_cap_is_bit_identical(device, precision, shape):
    """A capped plan (persistent workgroups looping over the work items, the
    serving loop's CU reserve) computes every item as the uncapped plan does:
    outputs must be bit-identical, fused and separate stages, for caps below
    and above the CU count; the uncapped result is also held to the oracle."""
    import torch

    note("advection flux workgroup cap %s p=%d" % (shape, precision))
    w = S.Workload(*shape)
    modes = S.retained_modes(w)
    dz = S.vertical_derivative_matrix(w.nz, 1.0)
    plan = S.AdvectionFluxPlan(w, modes, dz, precision)
    spec = S.spectral_fixture(w, modes, 7, -5.0 / 6.0)
    cd = S.complex_dtype(precision)
    d_spec = torch.from_numpy(spec).to(device=device, dtype=cd)
    for fused in (True, False):
        plan.set_fused(fused)
        plan.set_workgroups(0)
        ref = torch.empty(len(modes) * w.nz, dtype=cd, device=device)
        plan.execute(d_spec, ref)
        for cap in (8, 64, 224, 4096):
            plan.set_workgroups(cap)
            out = torch.full_like(ref, float("nan"))
            plan.execute(d_spec, out)
            torch.cuda.synchronize()
            assert torch.equal(out, ref), (fused, cap)
    if shape[0] <= 64:
        oref = O.advection_flux(O.Workload(*shape), O.retained_modes(O.Workload(*shape)), dz, host(d_spec))
        assert O.max_relative_error(host(ref), oref) <= 1e-12
    with pytest.raises(S.SkbError):
        plan.set_workgroups(-1)
