"""WVM-direct vertical GEMM (the grouped products applied in place on full
WVM arrays) through the C ABI, against the oracle's restatement of
WvmDirectVerticalGemmProvider (src/vertical_gemm.cpp:1107-1157), which
tests/test_oracle.py pins to the compact path and the k_x = 0 partner
symmetry.  FP64 within the reference's 1e-12, FP32 within 1e-6 of the FP64
oracle; entries of other frequencies must stay as initialised."""
import sys

import numpy as np
import pytest

import oracle as O
import skb_hip as S

pytestmark = pytest.mark.gpu

SENTINEL = complex(3.5, -1.25)


def note(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def host(t):
    import torch

    torch.cuda.synchronize()
    return t.detach().cpu().numpy().astype(np.complex128)


def setup(shape, precision, device):
    import torch

    w = S.Workload(*shape)
    ow = O.Workload(*shape)
    modes = S.retained_modes(w)
    groups = S.k2_groups(modes)
    nj = w.nj
    fwd, inv = S.k2_vertical_family(w.nz, nj, groups)
    plan = S.WvmVerticalPlan(w, modes, nj, groups, fwd, inv, precision)
    rng = np.random.default_rng(17)
    full = rng.standard_normal(w.spectrum_elements) + 1j * rng.standard_normal(w.spectrum_elements)
    cd = S.complex_dtype(precision)
    d_full = torch.from_numpy(full).to(device=device, dtype=cd)
    return w, ow, modes, groups, nj, fwd, inv, plan, d_full, cd


@pytest.mark.parametrize("precision,shape", [(S.SKB_F64, (16, 16, 7, 2)), (S.SKB_F64, (256, 256, 65, 3)),
                                             (S.SKB_F32, (256, 256, 65, 3)), (S.SKB_F32, (64, 32, 33, 1))])
def test_wvm_direct_forward_and_inverse_against_oracle(device, precision, shape):
    import torch

    note("wvm vertical %s p=%d" % (shape, precision))
    w, ow, modes, groups, nj, fwd, inv, plan, d_full, cd = setup(shape, precision, device)
    tol = 1e-12 if precision == S.SKB_F64 else 1e-6
    om = O.retained_modes(ow)
    first, count = groups["first_mode"], groups["mode_count"]
    modal = torch.empty(plan.modal_elements, dtype=cd, device=device)
    plan.initialize_modal(modal)
    plan.forward(d_full, modal)
    ref = O.wvm_direct_vertical(ow, om, nj, first, count, fwd, host(d_full), True)
    assert O.max_relative_error(host(modal), ref) <= tol
    spec = torch.empty(w.spectrum_elements, dtype=cd, device=device)
    plan.initialize_spectrum(spec)
    plan.inverse(modal, spec)
    ref_back = O.wvm_direct_vertical(ow, om, nj, first, count, inv, host(modal), False)
    assert O.max_relative_error(host(spec), ref_back) <= tol


def test_wvm_direct_leaves_other_frequencies_and_checks_entry_points(device):
    import torch

    w, ow, modes, groups, nj, fwd, inv, plan, d_full, cd = setup((32, 32, 13, 3), S.SKB_F64, device)
    note("wvm vertical untouched entries")
    modal = torch.full((plan.modal_elements,), SENTINEL, dtype=cd, device=device)
    plan.forward(d_full, modal)
    om = O.retained_modes(ow)
    ref = O.wvm_direct_vertical(ow, om, nj, groups["first_mode"], groups["mode_count"], fwd, host(d_full), True)
    got = host(modal).reshape(-1, w.fields * nj)
    ref = ref.reshape(-1, w.fields * nj)
    nxh = w.nx // 2 + 1
    touched = np.zeros(nxh * w.ny, bool)
    for m in modes:
        touched[int(m["stored_kx"]) + nxh * int(m["stored_ky"])] = True
        if m["stored_kx"] == 0 and m["stored_ky"] != 0 and 2 * m["stored_ky"] != w.ny:
            touched[nxh * ((w.ny - int(m["stored_ky"])) % w.ny)] = True
    assert O.max_relative_error(got[touched], ref[touched]) < 1e-12
    assert np.all(got[~touched] == SENTINEL)
    # compact entry points reject a WVM-direct plan (std::logic_error)
    with pytest.raises(S.LogicError):
        S.lib and S.check(S.lib.skb_vertical_forward(plan.h, S.ctypes.c_void_p(d_full.data_ptr()),
                                                     S.ctypes.c_void_p(modal.data_ptr()), None))
    # the operator must be Nj = 2(Nz-1)/3 (src/vertical_gemm.cpp:965-968)
    fwd2, inv2 = S.qr_gaussian_family(w.nz, len(groups), 3)
    with pytest.raises(S.InvalidArgument):
        S.WvmVerticalPlan(w, modes, w.nz, groups, fwd2, inv2, S.SKB_F64)
    torch.cuda.synchronize()
