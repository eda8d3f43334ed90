"""Split-BF16 products of the FP32 vertical GEMM (three-term bf16 factors,
FP32 accumulation) against the FP32-product GEMM and the FP64 oracle, for the
split shapes (Nz = 33, 65, 129 square families): forward and inverse within
1e-6 max-scale-normalised of the oracle, and within a small multiple of the
FP32-product error (the dropped terms are below 2^-24 of each product, so the
split form must not be measurably less accurate).  Shapes without a split form
(the truncated Nj x Nz family) report FP32 products."""
import sys

import numpy as np
import pytest

import oracle as O
import skb_hip as S

pytestmark = pytest.mark.gpu


def note(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("shape", [(128, 128, 33, 3), (256, 256, 65, 3), (64, 64, 129, 2)])
def test_split_bf16_products_match_oracle(device, shape):
    import torch

    note("split products %s" % (shape,))
    w = S.Workload(*shape)
    modes = S.retained_modes(w)
    groups = S.k2_binned_groups(modes, 32)
    fwd, inv = S.qr_gaussian_family(w.nz, len(groups), 129)
    ow = O.Workload(w.nx, w.ny, w.nz, w.fields)
    first, count = groups["first_mode"].copy(), groups["mode_count"].copy()
    rng = np.random.default_rng(17)
    n = len(modes) * w.planes
    phys = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    want_f = O.vertical_forward(ow, w.nz, first, count, fwd, phys)
    want_i = O.vertical_inverse(ow, w.nz, first, count, inv, want_f)
    d_phys = torch.from_numpy(phys).to(device=device, dtype=torch.complex64)
    d_modal = torch.from_numpy(want_f).to(device=device, dtype=torch.complex64)
    plan = S.VerticalPlan(w, w.nz, groups, fwd, inv, S.SKB_F32)
    assert plan.products(0) == S.SKB_PRODUCTS_SPLIT_BF16
    assert plan.products(1) == S.SKB_PRODUCTS_SPLIT_BF16
    errs = {}
    for products in (S.SKB_PRODUCTS_SPLIT_BF16, S.SKB_PRODUCTS_FP32):
        plan.set_products(products)
        assert plan.products(0) == products
        out_f = torch.empty_like(d_phys)
        out_i = torch.empty_like(d_phys)
        plan.forward(d_phys, out_f)
        plan.inverse(d_modal, out_i)
        torch.cuda.synchronize()
        ef = O.max_relative_error(out_f.cpu().numpy().astype(np.complex128), want_f)
        ei = O.max_relative_error(out_i.cpu().numpy().astype(np.complex128), want_i)
        errs[products] = (ef, ei)
        assert ef <= 1e-6 and ei <= 1e-6, (products, ef, ei)
    split, fp32 = errs[S.SKB_PRODUCTS_SPLIT_BF16], errs[S.SKB_PRODUCTS_FP32]
    note("max-scale error split %s fp32 %s" % (split, fp32))
    assert split[0] <= 4 * fp32[0] + 1e-8 and split[1] <= 4 * fp32[1] + 1e-8


def test_truncated_family_has_no_split_form(device):
    w = S.Workload(64, 64, 65, 2)
    modes = S.retained_modes(w)
    groups = S.k2_groups(modes)
    fwd, inv = S.k2_vertical_family(w.nz, w.nj, groups)
    plan = S.VerticalPlan(w, w.nj, groups, fwd, inv, S.SKB_F32)
    assert plan.products(0) == S.SKB_PRODUCTS_FP32  # Nj x Nz: M = 42
    assert plan.products(1) == S.SKB_PRODUCTS_FP32  # Nz x Nj: K = 42
    with pytest.raises(S.InvalidArgument):
        plan.set_products(7)
