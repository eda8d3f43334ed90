"""Split-layout entry points of the C ABI: skb_interleave / skb_deinterleave /
skb_scale_split and the split vertical products (VerticalGemmLayout::split,
include/skbench/skbench.hpp:125-128, src/vertical_gemm.cpp:558-596, 771-819).

The split products run the interleaved kernels between an interleave and a
deinterleave, so they must equal the interleaved products bit for bit; the
interleaved products are held to the FP64 oracle (1e-12 FP64, 1e-6 FP32 of
the max scale, src/benchmark.cpp:47, 118-127).
"""
import sys

import numpy as np
import pytest

import oracle as O
import skb_hip as S

pytestmark = pytest.mark.gpu


def note(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("precision", [S.SKB_F64, S.SKB_F32])
def test_interleave_deinterleave_scale(device, precision):
    import torch

    note("interleave precision %d" % precision)
    rd, cd = S.real_dtype(precision), S.complex_dtype(precision)
    g = torch.Generator(device="cpu").manual_seed(5)
    n = 100003  # not a multiple of the block size
    re = torch.randn(n, generator=g, dtype=torch.float64).to(device=device, dtype=rd)
    im = torch.randn(n, generator=g, dtype=torch.float64).to(device=device, dtype=rd)
    z = torch.empty(n, dtype=cd, device=device)
    S.interleave(precision, re, im, z)
    torch.cuda.synchronize()
    assert torch.equal(z.real, re) and torch.equal(z.imag, im)
    a, b = torch.full_like(re, 7.0), torch.full_like(im, 7.0)
    S.deinterleave(precision, z, a, b, 0.5)
    torch.cuda.synchronize()
    assert torch.equal(a, 0.5 * re) and torch.equal(b, 0.5 * im)
    c = torch.full_like(im, 3.0)
    S.deinterleave(precision, z, None, c)
    torch.cuda.synchronize()
    assert torch.equal(c, im)
    S.scale_split(precision, a, b, 2.0)
    torch.cuda.synchronize()
    assert torch.equal(a, re) and torch.equal(b, im)
    empty = torch.empty(0, dtype=cd, device=device)
    S.interleave(precision, re[:0], im[:0], empty)  # zero elements: no launch


def _vertical(shape, precision, kind):
    w = S.Workload(*shape)
    modes = S.retained_modes(w)
    if kind == "k2":
        groups = S.k2_groups(modes)
        f, i = S.k2_vertical_family(w.nz, w.nj, groups)
        nj = w.nj
    else:
        groups = S.k2_binned_groups(modes, 32)
        f, i = S.qr_gaussian_family(w.nz, len(groups), 129)
        nj = w.nz
    return w, modes, groups, nj, f, i, S.VerticalPlan(w, nj, groups, f, i, precision)


CASES = [((32, 32, 17, 2), S.SKB_F64, "k2", 1e-12), ((256, 256, 65, 3), S.SKB_F32, "qr32", 1e-6)]


@pytest.mark.parametrize("shape,precision,kind,tol", CASES)
def test_split_vertical_equals_interleaved_and_oracle(device, shape, precision, kind, tol):
    import torch

    note("split vertical %s precision %d %s" % (shape, precision, kind))
    w, modes, groups, nj, f, i, plan = _vertical(shape, precision, kind)
    rd, cd = S.real_dtype(precision), S.complex_dtype(precision)
    n_phys, n_modal = len(modes) * w.planes, len(modes) * w.fields * nj
    rng = np.random.default_rng(17)
    x = rng.standard_normal(n_phys) + 1j * rng.standard_normal(n_phys)
    d_x = torch.from_numpy(x).to(device=device, dtype=cd)
    ref_modal = torch.empty(n_modal, dtype=cd, device=device)
    plan.forward(d_x, ref_modal)
    with pytest.raises(S.LogicError):
        plan.forward_split(d_x.real.contiguous(), d_x.imag.contiguous(), torch.empty(n_modal, dtype=rd, device=device),
                           torch.empty(n_modal, dtype=rd, device=device))
    plan.reserve_split()
    plan.reserve_split()  # idempotent
    xr, xi = d_x.real.contiguous(), d_x.imag.contiguous()
    mr, mi = torch.empty(n_modal, dtype=rd, device=device), torch.empty(n_modal, dtype=rd, device=device)
    plan.forward_split(xr, xi, mr, mi)
    torch.cuda.synchronize()
    assert torch.equal(mr, ref_modal.real) and torch.equal(mi, ref_modal.imag)
    expected = O.vertical_forward(O.Workload(*shape), nj, groups["first_mode"].copy(), groups["mode_count"].copy(),
                                  f, x)
    got = (mr.double() + 1j * mi.double()).cpu().numpy()
    assert O.max_relative_error(got, expected) <= tol
    # one part at a time: the other output array is not touched
    sentinel = torch.full_like(mi, 11.0)
    mr2 = torch.empty_like(mr)
    plan.forward_split(xr, None, mr2, None)
    plan.forward_split(None, xi, None, sentinel)
    torch.cuda.synchronize()
    assert torch.equal(mr2, mr) and torch.equal(sentinel, mi)
    with pytest.raises(S.InvalidArgument):
        plan.forward_split(xr, None, None, mi)
    # inverse
    ref_phys = torch.empty(n_phys, dtype=cd, device=device)
    plan.inverse(ref_modal, ref_phys)
    pr, pi = torch.empty(n_phys, dtype=rd, device=device), torch.empty(n_phys, dtype=rd, device=device)
    plan.inverse_split(mr, mi, pr, pi)
    torch.cuda.synchronize()
    assert torch.equal(pr, ref_phys.real) and torch.equal(pi, ref_phys.imag)
    plan.close()
