"""Split compact forms of the pruned transforms (real and imaginary arrays
apart, optionally one strided view per field), through the C ABI.

Mirrors the reference's checks (tests/skbench_tests.cpp:1341-1460): the
split forward against the direct-DFT oracle gathered by mode, scaling by 0.25
against a separate scale, the split inverse against the direct c2r of the
embedded spectrum, and field views with mode stride Nz + 3; the gaps between
field-view columns must stay untouched.  The split and interleaved forms run
the same FFT source in separately compiled kernel instantiations, so they are
compared to rounding (1e-12 FP64, 1e-6 FP32 on the max-scale-normalised
error), not bit for bit.
"""
import sys

import numpy as np
import pytest

import oracle as O
import skb_hip as S

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
FIXTURES = ["impulse", "sinusoid", "random", "dc", "nyquist"]
SENTINEL = 7.25


def note(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def dev(a, dtype, device):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to(device=device, dtype=dtype)


def host(t):
    import torch

    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def ow(w):
    return O.Workload(w.nx, w.ny, w.nz, w.fields)


@pytest.mark.parametrize("kind", FIXTURES)
def test_split_forward_and_inverse_match_direct_dft(device, kind):
    import torch

    w = S.Workload(8, 8, 7, 2)
    note("split smoke %s" % kind)
    modes = S.retained_modes(w)
    om = O.retained_modes(ow(w))
    plan = S.PrunedPlan(w, modes, S.SKB_F64)
    x = O.make_fixture(ow(w), kind, 900 + FIXTURES.index(kind))
    oracle_ret = O.gather_retained(ow(w), om, O.direct_r2c(ow(w), x))
    d_x = dev(x, torch.float64, device)
    n = len(modes) * w.planes
    re = torch.empty(n, dtype=torch.float64, device=device)
    im = torch.empty(n, dtype=torch.float64, device=device)
    plan.forward_split(d_x, re, im)
    got = host(re) + 1j * host(im)
    assert O.max_relative_error(got, oracle_ret) < TOL64
    assert np.array_equal(host(d_x), x)
    # normalisation: scale 0.25 equals a separate multiply (exact: power of two)
    re4 = torch.empty_like(re)
    im4 = torch.empty_like(im)
    plan.forward_split(d_x, re4, im4, scale=0.25)
    assert np.array_equal(host(re4), 0.25 * host(re))
    assert np.array_equal(host(im4), 0.25 * host(im))
    # inverse from the oracle's split retained spectrum
    d_re = dev(oracle_ret.real, torch.float64, device)
    d_im = dev(oracle_ret.imag, torch.float64, device)
    real = torch.empty(w.real_elements, dtype=torch.float64, device=device)
    plan.inverse_split(d_re, d_im, real)
    oracle_real = O.direct_c2r(ow(w), O.embed_retained(ow(w), om, oracle_ret))
    assert O.max_relative_error(host(real), oracle_real) < TOL64
    assert np.array_equal(host(d_re), oracle_ret.real) and np.array_equal(host(d_im), oracle_ret.imag)


def _close(a, b, precision):
    tol = 1e-12 if precision == S.SKB_F64 else 1e-6
    return O.max_relative_error(np.asarray(a, np.complex128), np.asarray(b, np.complex128)) <= tol


def _field_views(torch, w, nmodes, stride, dtype, device):
    per = w.nz + stride * (nmodes - 1)
    views = []
    for _ in range(w.fields):
        views.append((torch.full((per,), SENTINEL, dtype=dtype, device=device),
                      torch.full((per,), SENTINEL, dtype=dtype, device=device), stride))
    return views


def _gather_views(views, w, nmodes):
    """(fields, nz, modes) complex array from the views, plus the gap values."""
    out = np.empty((w.fields, w.nz, nmodes), np.complex128)
    gaps = []
    for f, (re, im, stride) in enumerate(views):
        r, i = host(re).astype(np.float64), host(im).astype(np.float64)
        idx = np.arange(w.nz)[:, None] + stride * np.arange(nmodes)[None, :]
        out[f] = r[idx] + 1j * i[idx]
        mask = np.ones(r.size, bool)
        mask[idx.ravel()] = False
        gaps.append(r[mask])
        gaps.append(i[mask])
    return out, np.concatenate(gaps)


@pytest.mark.parametrize("precision", [S.SKB_F64, S.SKB_F32])
@pytest.mark.parametrize("shape", [(8, 8, 7, 2), (256, 256, 65, 3), (32, 128, 9, 4)])
def test_split_field_views_equal_interleaved(device, precision, shape):
    import torch

    w = S.Workload(*shape)
    note("split fields %s p=%d" % (shape, precision))
    modes = S.retained_modes(w)
    nm = len(modes)
    plan = S.PrunedPlan(w, modes, precision)
    rd, cd = S.real_dtype(precision), S.complex_dtype(precision)
    x = dev(O.make_fixture(ow(w), "random", 131), rd, device)
    ref = torch.empty(nm * w.planes, dtype=cd, device=device)
    plan.forward(x, ref)
    ref_h = host(ref).reshape(nm, w.fields, w.nz).astype(np.complex128)  # (mode, field, z)

    stride = w.nz + 3
    views = _field_views(torch, w, nm, stride, rd, device)
    plan.forward_split_fields(x, views)
    got, gaps = _gather_views(views, w, nm)
    assert _close(got, np.transpose(ref_h, (1, 2, 0)), precision)
    assert np.all(gaps == SENTINEL)

    # plain split layout equals the interleaved one
    re = torch.empty(nm * w.planes, dtype=rd, device=device)
    im = torch.empty_like(re)
    plan.forward_split(x, re, im)
    assert _close(host(re) + 1j * host(im).astype(np.complex128), host(ref), precision)

    # inverse: views holding the spectrum give the interleaved inverse
    spec = O.spectral_gaussian(ow(w), O.retained_modes(ow(w)), 129)
    d_spec = dev(spec, cd, device)
    real_ref = torch.empty(w.real_elements, dtype=rd, device=device)
    plan.inverse(d_spec, real_ref)
    sh = host(d_spec).reshape(nm, w.fields, w.nz)
    in_views = _field_views(torch, w, nm, stride, rd, device)
    for f, (vr, vi, st) in enumerate(in_views):
        idx = np.arange(w.nz)[:, None] + st * np.arange(nm)[None, :]
        hr = np.full(vr.numel(), SENTINEL)
        hi = np.full(vi.numel(), SENTINEL)
        hr[idx.ravel()] = sh[:, f, :].T.real.ravel()
        hi[idx.ravel()] = sh[:, f, :].T.imag.ravel()
        vr.copy_(torch.from_numpy(hr).to(rd))
        vi.copy_(torch.from_numpy(hi).to(rd))
    real = torch.empty_like(real_ref)
    plan.inverse_split_fields(in_views, real)
    assert _close(host(real), host(real_ref), precision)
    re_in = dev(spec.real, rd, device)
    im_in = dev(spec.imag, rd, device)
    real2 = torch.empty_like(real_ref)
    plan.inverse_split(re_in, im_in, real2)
    assert _close(host(real2), host(real_ref), precision)


def test_split_field_view_validation(device):
    import torch

    w = S.Workload(16, 16, 5, 3)
    note("split field validation")
    modes = S.retained_modes(w)
    plan = S.PrunedPlan(w, modes, S.SKB_F64)
    x = torch.zeros(w.real_elements, dtype=torch.float64, device=device)
    views = _field_views(torch, w, len(modes), w.nz, torch.float64, device)
    with pytest.raises(S.InvalidArgument, match="field count"):
        plan.forward_split_fields(x, views[:2])
    with pytest.raises(S.InvalidArgument, match="field count"):
        plan.inverse_split_fields(views[:2], x)
    with pytest.raises(S.InvalidArgument):
        plan.forward_split_fields(x, [(views[0][0], views[0][1], w.nz - 1)] + views[1:])
    torch.cuda.synchronize()
