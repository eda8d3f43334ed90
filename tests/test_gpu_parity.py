"""Parity of the HIP kernels (through the C ABI) against the oracle.

Tolerances: FP64 paths use the reference's 1e-12 on the max-scale-normalised
error (src/benchmark.cpp:47, 118-127); FP32-storage paths use the 1e-6 bound
on the same metric, measured against the FP64 oracle on identical inputs.
Full-size cases (256^2 x Nz65 x 3, 512^2 x Nz129 x 4) compare against the
pocketfft oracle; small grids against the direct DFT.
"""
import sys

import numpy as np
import pytest

import oracle as O
import skb_hip as S

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-6
FIXTURES = ["impulse", "sinusoid", "random", "dc", "nyquist"]


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


def tol(precision):
    return TOL64 if precision == S.SKB_F64 else TOL32


# ---- full r2c / c2r in WVM order ------------------------------------------------
@pytest.mark.parametrize("kind", FIXTURES)
def test_full_transforms_match_direct_dft_on_smoke_grid(device, kind):
    w = S.Workload(8, 8, 7, 2)
    note("full smoke %s" % kind)
    plan = S.FftPlan(w, S.SKB_F64)
    import torch

    x = O.make_fixture(ow(w), kind, 129)
    oracle_spec = O.direct_r2c(ow(w), x)
    spec = torch.empty(w.spectrum_elements, dtype=torch.complex128, device=device)
    plan.forward(dev(x, torch.float64, device), spec)
    assert O.max_relative_error(host(spec), oracle_spec) <= TOL64
    # inverse of the oracle spectrum recovers Nx*Ny*x (FFTW c2r destroys its input)
    real = torch.empty(w.real_elements, dtype=torch.float64, device=device)
    plan.inverse(dev(oracle_spec, torch.complex128, device), real)
    assert O.max_relative_error(host(real), x, 1.0 / (w.nx * w.ny)) <= TOL64


@pytest.mark.parametrize("precision", [S.SKB_F64, S.SKB_F32])
@pytest.mark.parametrize("nx,ny", [(4, 4), (8, 16), (16, 8), (32, 32), (64, 128), (128, 64), (256, 256),
                                   (512, 512), (1024, 1024), (1024, 4), (4, 1024)])
def test_full_transforms_every_size_against_pocketfft(device, precision, nx, ny):
    import torch

    w = S.Workload(nx, ny, 3, 1)  # odd plane count exercises the partial tile
    note("full sizes %dx%d p=%d" % (nx, ny, precision))
    plan = S.FftPlan(w, precision)
    x = O.make_fixture(ow(w), "random", 7 + nx + ny)
    spec = torch.empty(w.spectrum_elements, dtype=S.complex_dtype(precision), device=device)
    plan.forward(dev(x, S.real_dtype(precision), device), spec)
    xin = host(dev(x, S.real_dtype(precision), device))  # the exact input the GPU saw
    ref = O.fft_r2c_wvm(ow(w), xin)
    assert O.max_relative_error(host(spec), ref) <= tol(precision)
    rng = np.random.default_rng(nx * ny)
    inv_in = rng.standard_normal(w.spectrum_elements) + 1j * rng.standard_normal(w.spectrum_elements)
    d_in = dev(inv_in, S.complex_dtype(precision), device)
    inv_used = host(d_in)
    real = torch.empty(w.real_elements, dtype=S.real_dtype(precision), device=device)
    plan.inverse(d_in, real)
    assert O.max_relative_error(host(real), O.fft_c2r_wvm(ow(w), inv_used)) <= tol(precision)


@pytest.mark.parametrize("precision", [S.SKB_F64, S.SKB_F32])
def test_full_transforms_at_256_nz65_f3(device, precision):
    import torch

    w = S.Workload(256, 256, 65, 3)
    note("full 256x256x65x3 p=%d" % precision)
    plan = S.FftPlan(w, precision)
    x = O.make_fixture(ow(w), "random", 129)
    d_x = dev(x, S.real_dtype(precision), device)
    spec = torch.empty(w.spectrum_elements, dtype=S.complex_dtype(precision), device=device)
    plan.forward(d_x, spec)
    ref = O.fft_r2c_wvm(ow(w), host(d_x))
    err = O.max_relative_error(host(spec), ref)
    assert err <= tol(precision), err
    assert O.relative_l2_error(host(spec), ref) <= tol(precision)
    real = torch.empty(w.real_elements, dtype=S.real_dtype(precision), device=device)
    plan.inverse(spec, real)  # round trip: c2r(r2c(x)) = Nx*Ny*x
    assert O.max_relative_error(host(real), host(d_x), 1.0 / (w.nx * w.ny)) <= tol(precision)


def test_gather_embed_on_device_are_exact(device):
    import torch

    w = S.Workload(64, 64, 5, 3)
    note("gather/embed 64x64x5x3")
    plan = S.FftPlan(w, S.SKB_F64)
    modes = S.retained_modes(w)
    n = w.spectrum_elements
    wvm = np.arange(n) - 1j * np.arange(n) * 0.5
    ret = torch.empty(len(modes) * w.planes, dtype=torch.complex128, device=device)
    plan.gather_retained(modes, dev(wvm, torch.complex128, device), ret)
    expected = O.gather_retained(ow(w), O.retained_modes(ow(w)), wvm)
    assert np.array_equal(host(ret), expected)
    full = torch.full((n,), 7.0 + 7.0j, dtype=torch.complex128, device=device)
    plan.embed_retained(modes, ret, full)
    assert np.array_equal(host(full), O.embed_retained(ow(w), O.retained_modes(ow(w)), expected))


# ---- pruned retained transforms -------------------------------------------------
@pytest.mark.parametrize("kind", FIXTURES)
def test_pruned_matches_direct_dft_and_preserves_inputs(device, kind):
    import torch

    w = S.Workload(8, 8, 7, 2)
    note("pruned smoke %s" % kind)
    modes = S.retained_modes(w)
    plan = S.PrunedPlan(w, modes, S.SKB_F64)
    assert plan.active_kx() < w.nx_half
    om = O.retained_modes(ow(w))
    x = O.make_fixture(ow(w), kind, 700 + FIXTURES.index(kind))
    d_x = dev(x, torch.float64, device)
    ret = torch.empty(len(modes) * w.planes, dtype=torch.complex128, device=device)
    plan.forward(d_x, ret)
    oracle_ret = O.gather_retained(ow(w), om, O.direct_r2c(ow(w), x))
    assert O.max_relative_error(host(ret), oracle_ret) < TOL64
    assert np.array_equal(host(d_x), x)
    d_ret = dev(oracle_ret, torch.complex128, device)
    real = torch.empty(w.real_elements, dtype=torch.float64, device=device)
    plan.inverse(d_ret, real)
    oracle_real = O.direct_c2r(ow(w), O.embed_retained(ow(w), om, oracle_ret))
    assert O.max_relative_error(host(real), oracle_real) < TOL64
    assert np.array_equal(host(d_ret), oracle_ret)


@pytest.mark.parametrize("precision", [S.SKB_F64, S.SKB_F32])
@pytest.mark.parametrize("shape", [(256, 256, 65, 3), (128, 128, 33, 3), (512, 512, 129, 4), (32, 64, 3, 1)])
def test_pruned_against_pocketfft(device, precision, shape):
    import torch

    w = S.Workload(*shape)
    note("pruned %s p=%d" % (shape, precision))
    modes = S.retained_modes(w)
    om = O.retained_modes(ow(w))
    plan = S.PrunedPlan(w, modes, precision)
    x = O.make_fixture(ow(w), "random", 129)
    d_x = dev(x, S.real_dtype(precision), device)
    ret = torch.empty(len(modes) * w.planes, dtype=S.complex_dtype(precision), device=device)
    plan.forward(d_x, ret)
    ref = O.retained_forward(ow(w), om, host(d_x))
    assert O.max_relative_error(host(ret), ref) <= tol(precision)
    spec = O.spectral_gaussian(ow(w), om, 129)
    d_s = dev(spec, S.complex_dtype(precision), device)
    real = torch.empty(w.real_elements, dtype=S.real_dtype(precision), device=device)
    plan.inverse(d_s, real)
    ref_real = O.retained_inverse(ow(w), om, host(d_s))
    assert O.max_relative_error(host(real), ref_real) <= tol(precision)


# ---- vertical GEMM ----------------------------------------------------------------
def _family(w, modes, kind):
    if kind == "k2":
        groups = S.k2_groups(modes)
        f, i = S.k2_vertical_family(w.nz, w.nj, groups)
        return groups, w.nj, f, i
    groups = S.k2_binned_groups(modes, 32)
    f, i = S.qr_gaussian_family(w.nz, len(groups), 129)
    return groups, w.nz, f, i


@pytest.mark.parametrize("precision", [S.SKB_F64, S.SKB_F32])
@pytest.mark.parametrize("kind", ["k2", "qr32"])
@pytest.mark.parametrize("shape", [(256, 256, 65, 3), (16, 16, 7, 2), (64, 64, 33, 1)])
def test_vertical_gemm_against_oracle(device, precision, kind, shape):
    import torch

    w = S.Workload(*shape)
    note("vertical %s %s p=%d" % (kind, shape, precision))
    modes = S.retained_modes(w)
    groups, nj, f, i = _family(w, modes, kind)
    plan = S.VerticalPlan(w, nj, groups, f, i, precision)
    first = groups["first_mode"].copy()
    count = groups["mode_count"].copy()
    rng = np.random.default_rng(3)
    n = len(modes) * w.planes
    phys = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    d_p = dev(phys, S.complex_dtype(precision), device)
    modal = torch.empty(len(modes) * w.fields * nj, dtype=S.complex_dtype(precision), device=device)
    plan.forward(d_p, modal)
    ref = O.vertical_forward(ow(w), nj, first, count, f, host(d_p))
    assert O.max_relative_error(host(modal), ref) <= tol(precision)
    back = torch.empty_like(d_p)
    plan.inverse(modal, back)
    ref_back = O.vertical_inverse(ow(w), nj, first, count, i, host(modal))
    assert O.max_relative_error(host(back), ref_back) <= tol(precision)


# ---- pointwise advection ------------------------------------------------------------
@pytest.mark.parametrize("precision", [S.SKB_F64, S.SKB_F32])
@pytest.mark.parametrize("volume", [37, 1, 4099, 256 * 256 * 129])
def test_pointwise_advection_known_answer(device, precision, volume):
    note("advection v=%d p=%d" % (volume, precision))
    idx = np.arange(3 * volume)
    shared = (idx % 17) - 8.0
    deriv = ((3 * idx + 5) % 23) - 11.0
    scale = 0.03125
    rd = S.real_dtype(precision)
    d_s, d_d = dev(shared, rd, device), dev(deriv, rd, device)
    import torch

    out = torch.empty(volume, dtype=rd, device=device)
    S.pointwise_advection(volume, scale, d_s, d_d, out, precision)
    expected = O.pointwise_advection(volume, scale, shared, deriv)
    bound = 1e-15 if precision == S.SKB_F64 else 1e-7
    assert O.max_relative_error(host(out), expected) <= bound


@pytest.mark.parametrize("volume", [4096, 4099])
def test_pointwise_advection_fp32_sums_in_fp64(device, volume):
    """FP32 streams: u*qx + v*qy + w*qz is summed in FP64 and rounded once.
    With u = qx = 1 + 2^-12 (product 1 + 2^-11 + 2^-24), v*qy = -1 and
    w*qz = 2^-20, the exact sum 2^-11 + 2^-20 + 2^-24 is an FP32 value; a sum
    in FP32 drops the product's 2^-24 (relative error 2^-13 = 1.2e-4).  Each
    element scales u, v and w by 2^(i mod 8 - 4), which scales the sum
    exactly; the oracle's FP64 sum is exact on these inputs."""
    import torch

    note("advection fp64 sum v=%d" % volume)
    i = np.arange(volume)
    f = np.ldexp(1.0, (i % 8) - 4)
    a = 1.0 + 2.0 ** -12
    shared = np.concatenate([a * f, -1.0 * f, 2.0 ** -20 * f])
    deriv = np.concatenate([np.full(volume, a), np.ones(volume), np.ones(volume)])
    expected = O.pointwise_advection(volume, 1.0, shared, deriv)
    assert np.allclose(expected, -(2.0 ** -11 + 2.0 ** -20 + 2.0 ** -24) * f, rtol=0, atol=0)
    d_s, d_d = dev(shared, torch.float32, device), dev(deriv, torch.float32, device)
    out = torch.empty(volume, dtype=torch.float32, device=device)
    S.pointwise_advection(volume, 1.0, d_s, d_d, out, S.SKB_F32)
    assert np.array_equal(host(out).astype(np.float64), expected)


# ---- composed retained operator --------------------------------------------------------
@pytest.mark.parametrize("kind", ["qr32", "k2"])
def test_retained_operator_fp64_and_fp32_against_fp64_oracle(device, kind):
    import torch

    w = S.Workload(256, 256, 65, 3)
    note("retained operator %s" % kind)
    modes = S.retained_modes(w)
    om = O.retained_modes(ow(w))
    groups, nj, f, i = _family(w, modes, kind)
    x = O.spectral_gaussian(ow(w), om, 129)
    ref = O.retained_operator(ow(w), om, x, nj, groups["first_mode"].copy(), groups["mode_count"].copy(), f)
    for precision in (S.SKB_F64, S.SKB_F32):
        pruned = S.PrunedPlan(w, modes, precision)
        vert = S.VerticalPlan(w, nj, groups, f, i, precision)
        d_x = dev(x, S.complex_dtype(precision), device)
        real = torch.empty(w.real_elements, dtype=S.real_dtype(precision), device=device)
        back = torch.empty_like(d_x)
        modal = torch.empty(len(modes) * w.fields * nj, dtype=S.complex_dtype(precision), device=device)
        pruned.inverse(d_x, real)
        pruned.forward(real, back)
        vert.forward(back, modal)
        got = host(modal)
        if precision == S.SKB_F64:
            assert O.max_relative_error(got, ref) <= TOL64
        else:
            # FP32 storage of the input is part of the FP32 form; error is
            # against the FP64 oracle on the original FP64 input.
            err = O.max_relative_error(got, ref)
            assert err <= TOL32, err
            assert O.relative_l2_error(got, ref) <= TOL32
