"""skb_vertical_forward_split_noalloc: the split forward product with its
staging in the library's static device scratch instead of buffers reserved on
the plan.  Through the C ABI, two calls in a row, the second on a smaller
plan, each equal bit for bit to skb_vertical_forward_split on the same plan
and inputs (the same kernels on the same values); the first is also held to
the FP64 oracle at the FP32 bound (1e-6 of the max scale,
src/benchmark.cpp:47, 118-127)."""
import sys

import numpy as np
import pytest

import oracle as O
import skb_hip as S

pytestmark = pytest.mark.gpu


def note(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def _plan(shape, precision):
    w = S.Workload(*shape)
    modes = S.retained_modes(w)
    groups = S.k2_binned_groups(modes, 32)
    f, i = S.qr_gaussian_family(w.nz, len(groups), 129)
    return w, modes, groups, f, S.VerticalPlan(w, w.nz, groups, f, i, precision)


def _one(device, shape, precision, seed):
    import torch

    w, modes, groups, f, plan = _plan(shape, precision)
    rd = S.real_dtype(precision)
    n_phys, n_modal = len(modes) * w.planes, len(modes) * w.fields * w.nz
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(n_phys) + 1j * rng.standard_normal(n_phys)
    xr = torch.from_numpy(x.real.copy()).to(device=device, dtype=rd)
    xi = torch.from_numpy(x.imag.copy()).to(device=device, dtype=rd)
    a_re, a_im = torch.full((n_modal,), 5.0, dtype=rd, device=device), torch.full((n_modal,), 5.0, dtype=rd,
                                                                                      device=device)
    plan.forward_split_noalloc(xr, xi, a_re, a_im)
    torch.cuda.synchronize()
    plan.reserve_split()
    b_re, b_im = torch.empty_like(a_re), torch.empty_like(a_im)
    plan.forward_split(xr, xi, b_re, b_im)
    torch.cuda.synchronize()
    assert torch.equal(a_re, b_re) and torch.equal(a_im, b_im)
    plan.close()
    return w, groups, f, x, a_re, a_im


def test_noalloc_split_forward_equals_reserved_on_two_calls(device):
    note("noalloc split forward 256x256x65x3 then 64x64x17x2")
    w, groups, f, x, a_re, a_im = _one(device, (256, 256, 65, 3), S.SKB_F32, 3)
    expected = O.vertical_forward(O.Workload(256, 256, 65, 3), w.nz, groups["first_mode"].copy(),
                                  groups["mode_count"].copy(), f, x)
    got = a_re.double().cpu().numpy() + 1j * a_im.double().cpu().numpy()
    assert O.max_relative_error(got, expected) <= 1e-6
    # a second call on a smaller plan reuses the same scratch
    _one(device, (64, 64, 17, 2), S.SKB_F32, 4)
    # and FP64
    _one(device, (32, 32, 9, 2), S.SKB_F64, 5)


def test_noalloc_split_forward_rejects_staging_above_the_scratch(device):
    import torch

    note("noalloc split forward rejects 512x512x129x4")
    w, modes, groups, f, plan = _plan((512, 512, 129, 4), S.SKB_F32)
    n_phys, n_modal = len(modes) * w.planes, len(modes) * w.fields * w.nz
    assert (n_phys + n_modal) * 8 > 64 << 20
    z = torch.zeros(n_phys, dtype=torch.float32, device=device)
    m = torch.zeros(n_modal, dtype=torch.float32, device=device)
    with pytest.raises(S.InvalidArgument, match="no-allocation scratch"):
        plan.forward_split_noalloc(z, z, m, m)
    plan.close()
