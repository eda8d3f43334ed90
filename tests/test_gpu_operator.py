"""The retained-operator plan and the plane-resident transforms, through the C ABI.

Tolerance: FP32 storage against the FP64 oracle on identical inputs, 1e-6 on
the max-scale-normalised error (and relative L2), the bound the FP32 form is
held to everywhere; FP64 plans 1e-12 (src/benchmark.cpp:47, 118-127).  The
plane path and the two-pass path are different kernels over the same
arithmetic, so they agree to rounding (checked at 1e-6 of the scale).
"""
import sys

import numpy as np
import pytest

import oracle as O
import skb_hip as S

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-6


def note(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def dev(a, dtype, device):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to(device=device, dtype=dtype)


def host(t):
    import torch

    torch.cuda.synchronize()
    return t.detach().cpu().numpy().astype(np.complex128 if t.is_complex() else np.float64)


def ow(w):
    return O.Workload(w.nx, w.ny, w.nz, w.fields)


def family(w, modes, kind):
    if kind == "k2":
        groups = S.k2_groups(modes)
        f, i = S.k2_vertical_family(w.nz, w.nj, groups)
        return groups, w.nj, f, i
    groups = S.k2_binned_groups(modes, 32)
    f, i = S.qr_gaussian_family(w.nz, len(groups), 129)
    return groups, w.nz, f, i


SHAPES = [(256, 256, 65, 3), (256, 256, 7, 1), (256, 256, 33, 5)]


@pytest.mark.parametrize("shape", SHAPES)
def test_plane_inverse_against_oracle_and_two_pass(device, shape):
    import torch

    w = S.Workload(*shape)
    note("plane inverse %s" % (shape,))
    modes = S.retained_modes(w)
    om = O.retained_modes(ow(w))
    x = O.spectral_gaussian(ow(w), om, 129)
    ref = O.retained_inverse(ow(w), om, x)
    plan = S.PrunedPlan(w, modes, S.SKB_F32)
    assert plan.path() == S.SKB_PATH_PLANE
    d_x = dev(x, torch.complex64, device)
    keep = d_x.clone()
    real = torch.empty(w.real_elements, dtype=torch.float32, device=device)
    plan.inverse(d_x, real)
    got = host(real)
    assert O.max_relative_error(got, ref) <= TOL32
    assert O.relative_l2_error(got, ref) <= TOL32
    assert torch.equal(d_x, keep), "caller input must be preserved"
    plan.set_path(S.SKB_PATH_TWO_PASS)
    assert plan.path() == S.SKB_PATH_TWO_PASS
    real2 = torch.empty_like(real)
    plan.inverse(d_x, real2)
    scale = np.abs(ref).max()
    assert np.abs(host(real2) - got).max() <= TOL32 * scale


@pytest.mark.parametrize("shape", SHAPES)
def test_plane_forward_against_oracle_and_two_pass(device, shape):
    import torch

    w = S.Workload(*shape)
    note("plane forward %s" % (shape,))
    modes = S.retained_modes(w)
    om = O.retained_modes(ow(w))
    real = np.random.default_rng(129).uniform(-1.0, 1.0, w.real_elements)
    ref = O.retained_forward(ow(w), om, real)
    plan = S.PrunedPlan(w, modes, S.SKB_F32)
    d_r = dev(real, torch.float32, device)
    keep = d_r.clone()
    out = torch.empty(len(modes) * w.planes, dtype=torch.complex64, device=device)
    plan.forward(d_r, out)
    got = host(out)
    assert O.max_relative_error(got, ref) <= TOL32
    assert O.relative_l2_error(got, ref) <= TOL32
    assert torch.equal(d_r, keep)
    plan.set_path(S.SKB_PATH_TWO_PASS)
    out2 = torch.empty_like(out)
    plan.forward(d_r, out2)
    assert np.abs(host(out2) - got).max() <= TOL32 * np.abs(ref).max()


def test_plane_round_trip_is_nxny_identity(device):
    """forward(inverse(x)) = Nx*Ny*x on the retained set (unnormalised
    transforms), for Hermitian-consistent x."""
    import torch

    w = S.Workload(256, 256, 65, 3)
    modes = S.retained_modes(w)
    om = O.retained_modes(ow(w))
    x = O.spectral_gaussian(ow(w), om, 7)
    plan = S.PrunedPlan(w, modes, S.SKB_F32)
    d_x = dev(x, torch.complex64, device)
    real = torch.empty(w.real_elements, dtype=torch.float32, device=device)
    back = torch.empty_like(d_x)
    plan.inverse(d_x, real)
    plan.forward(real, back)
    expected = O.retained_forward(ow(w), om, O.retained_inverse(ow(w), om, x))
    assert O.max_relative_error(host(back), expected) <= TOL32
    assert O.max_relative_error(host(back) / (w.nx * w.ny), expected / (w.nx * w.ny)) <= TOL32


@pytest.mark.parametrize("kind", ["qr32", "k2"])
@pytest.mark.parametrize("path", [S.SKB_PATH_AUTO, S.SKB_PATH_TWO_PASS])
def test_operator_against_fp64_oracle(device, kind, path):
    import torch

    w = S.Workload(256, 256, 65, 3)
    note("operator %s path %d" % (kind, path))
    modes = S.retained_modes(w)
    om = O.retained_modes(ow(w))
    groups, nj, f, i = family(w, modes, kind)
    x = O.spectral_gaussian(ow(w), om, 129)
    ref = O.retained_operator(ow(w), om, x, nj, groups["first_mode"].copy(), groups["mode_count"].copy(), f)
    op = S.OperatorPlan(w, modes, nj, groups, f, i, S.SKB_F32)
    op.set_path(path)
    assert op.path() == (S.SKB_PATH_PLANE if path == S.SKB_PATH_AUTO else S.SKB_PATH_TWO_PASS)
    d_x = dev(x, torch.complex64, device)
    keep = d_x.clone()
    modal = torch.empty(op.modal_elements, dtype=torch.complex64, device=device)
    op.execute(d_x, modal)
    got = host(modal)
    err = O.max_relative_error(got, ref)
    note("operator %s path %d max-scale error %.3e" % (kind, path, err))
    assert err <= TOL32
    assert O.relative_l2_error(got, ref) <= TOL32
    assert torch.equal(d_x, keep)
    # the physical-space boundary: the plan's real planes are the inverse transform of x
    real_ref = O.retained_inverse(ow(w), om, x)
    assert O.max_relative_error(host(op.real()), real_ref) <= TOL32


def test_operator_stages_equal_execute_and_repeat_bit_exact(device):
    import torch

    w = S.Workload(256, 256, 65, 3)
    modes = S.retained_modes(w)
    om = O.retained_modes(ow(w))
    groups, nj, f, i = family(w, modes, "qr32")
    op = S.OperatorPlan(w, modes, nj, groups, f, i, S.SKB_F32)
    d_x = dev(O.spectral_gaussian(ow(w), om, 3), torch.complex64, device)
    a = torch.empty(op.modal_elements, dtype=torch.complex64, device=device)
    b = torch.empty_like(a)
    c = torch.empty_like(a)
    assert op.fused()
    op.execute(d_x, a)
    real_fused = op.real().clone()
    for stage in S.OperatorPlan.STAGES:
        op.stage(stage, d_x, b)
    torch.cuda.synchronize()
    assert torch.equal(op.real(), real_fused)
    op.set_fused(False)
    op.bind_execute(d_x, c)()
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    assert torch.equal(a, c)


def test_operator_zero_after_nonzero_leaks_nothing(device):
    """An all-zero evaluation right after a non-zero one returns exact zeros
    (tests/skbench_tests.cpp:1558-1569 for the transforms; here through every
    buffer the plan reuses)."""
    import torch

    w = S.Workload(256, 256, 65, 3)
    modes = S.retained_modes(w)
    om = O.retained_modes(ow(w))
    groups, nj, f, i = family(w, modes, "qr32")
    for path in (S.SKB_PATH_AUTO, S.SKB_PATH_TWO_PASS):
        op = S.OperatorPlan(w, modes, nj, groups, f, i, S.SKB_F32)
        op.set_path(path)
        d_x = dev(O.spectral_gaussian(ow(w), om, 5), torch.complex64, device)
        modal = torch.empty(op.modal_elements, dtype=torch.complex64, device=device)
        op.execute(d_x, modal)
        op.execute(torch.zeros_like(d_x), modal)
        torch.cuda.synchronize()
        assert torch.count_nonzero(modal).item() == 0
        assert torch.count_nonzero(op.real()).item() == 0


def test_operator_fp64_and_unbuilt_shape_take_two_pass(device):
    import torch

    for shape, precision, tol in (((256, 256, 65, 3), S.SKB_F64, TOL64), ((128, 128, 33, 3), S.SKB_F32, TOL32)):
        w = S.Workload(*shape)
        note("operator two-pass %s %d" % (shape, precision))
        modes = S.retained_modes(w)
        om = O.retained_modes(ow(w))
        groups, nj, f, i = family(w, modes, "qr32")
        x = O.spectral_gaussian(ow(w), om, 129)
        ref = O.retained_operator(ow(w), om, x, nj, groups["first_mode"].copy(), groups["mode_count"].copy(), f)
        op = S.OperatorPlan(w, modes, nj, groups, f, i, precision)
        assert op.path() == S.SKB_PATH_TWO_PASS
        with pytest.raises(S.Unsupported):
            op.set_path(S.SKB_PATH_PLANE)
        d_x = dev(x, S.complex_dtype(precision), device)
        modal = torch.empty(op.modal_elements, dtype=S.complex_dtype(precision), device=device)
        op.execute(d_x, modal)
        assert O.max_relative_error(host(modal), ref) <= tol


def test_plane_path_with_8_byte_aligned_compact_arrays(device):
    """Compact arrays only 8-byte aligned (a view one complex value in) take
    the 8-byte layout copy and give the same results."""
    import torch

    w = S.Workload(256, 256, 7, 1)
    modes = S.retained_modes(w)
    om = O.retained_modes(ow(w))
    x = O.spectral_gaussian(ow(w), om, 11)
    plan = S.PrunedPlan(w, modes, S.SKB_F32)
    n = len(modes) * w.planes
    big = torch.zeros(n + 1, dtype=torch.complex64, device=device)
    big[1:] = dev(x, torch.complex64, device)
    aligned = big[1:].clone()
    real_a = torch.empty(w.real_elements, dtype=torch.float32, device=device)
    real_b = torch.empty_like(real_a)
    plan.inverse(aligned, real_a)
    plan.inverse(big[1:], real_b)
    torch.cuda.synchronize()
    assert torch.equal(real_a, real_b)
    out_big = torch.zeros(n + 1, dtype=torch.complex64, device=device)
    out_a = torch.empty(n, dtype=torch.complex64, device=device)
    plan.forward(real_a, out_a)
    plan.forward(real_a, out_big[1:])
    torch.cuda.synchronize()
    assert torch.equal(out_a, out_big[1:])
    assert out_big[0].item() == 0


def test_mode_set_outside_the_band_takes_two_pass(device):
    """A mode with k_y in the band the plane kernels skip (here l = 100) sends
    the plan to the two-pass kernels, which still match the oracle."""
    import torch

    w = S.Workload(256, 256, 7, 1)
    modes = S.retained_modes(w)
    extra = np.zeros(1, dtype=modes.dtype)
    extra["k"], extra["l"], extra["stored_kx"], extra["stored_ky"], extra["radial"] = 1, 100, 1, 100, 100.0
    modes = np.concatenate([modes, extra])
    plan = S.PrunedPlan(w, modes, S.SKB_F32)
    assert plan.path() == S.SKB_PATH_TWO_PASS
    om = O.retained_modes(ow(w))
    om = np.concatenate([om, extra.astype(om.dtype)])
    x = O.spectral_gaussian(ow(w), om, 17)
    real = torch.empty(w.real_elements, dtype=torch.float32, device=device)
    plan.inverse(dev(x, torch.complex64, device), real)
    assert O.max_relative_error(host(real), O.retained_inverse(ow(w), om, x)) <= TOL32
