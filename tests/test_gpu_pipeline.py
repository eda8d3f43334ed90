"""Synthetic spectral pipeline and modal work on the GPU, through the C ABI,
against the oracle (pinned in tests/test_pipeline_oracle.py).

Mirrors the reference's spectral-pipeline checks (src/benchmark.cpp:3175-3198;
tests/skbench_tests.cpp:2443-2520): the five correctness arrays (retained
horizontal coefficients, forward modal, post-work modal, inverse vertical
output, pipeline output) for each boundary policy, FP64 within the
reference's 1e-12 and FP32 within 1e-6 of the FP64 oracle (max-scale
normalised).  The modal work itself is one rounding per element, so it is
compared bit for bit.  Full-size shapes add the eigenfunction property
(output = Nx*Ny*weight*input for one horizontal mode times one column of
its group's inverse matrix).
"""
import sys

import numpy as np
import pytest

import oracle as O
import skb_hip as S

pytestmark = pytest.mark.gpu

TOL = {S.SKB_F64: 1e-12, S.SKB_F32: 1e-6}
POLICIES = {"wvm": S.SKB_PIPELINE_WVM_DIRECT, "compact": S.SKB_PIPELINE_COMPACT}


def note(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def host(t):
    import torch

    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def dev(a, dtype, device):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to(device=device, dtype=dtype)


def family(shape):
    w = S.Workload(*shape)
    modes = S.retained_modes(w)
    groups = S.k2_groups(modes)
    fwd, inv = S.k2_vertical_family(w.nz, w.nj, groups)
    return w, O.Workload(*shape), modes, groups, fwd, inv


def rounded(x, precision):
    """The input as the device sees it, back in FP64."""
    return np.asarray(x, np.float32).astype(np.float64) if precision == S.SKB_F32 else np.asarray(x, np.float64)


# ---- modal work -----------------------------------------------------------------
@pytest.mark.parametrize("precision", [S.SKB_F64, S.SKB_F32])
def test_modal_work_interleaved_and_split_bitwise(device, precision):
    import torch

    note("modal work p=%d" % precision)
    n = 100003
    rng = np.random.default_rng(2)
    rd, cd = S.real_dtype(precision), S.complex_dtype(precision)
    npr = np.float64 if precision == S.SKB_F64 else np.float32
    wt = (0.75 + 0.5 * rng.random(n)).astype(npr)
    re = rng.standard_normal(n).astype(npr)
    im = rng.standard_normal(n).astype(npr)
    d_w = dev(wt, rd, device)
    d_x = torch.complex(dev(re, rd, device), dev(im, rd, device)).contiguous()
    out = torch.empty_like(d_x)
    S.modal_work(d_w, d_x, out, precision)
    got = host(out)
    assert np.array_equal(got.real, wt * re) and np.array_equal(got.imag, wt * im)
    S.modal_work(d_w, d_x, d_x, precision)  # in place
    assert np.array_equal(host(d_x), got)
    o_re, o_im = torch.empty(n, dtype=rd, device=device), torch.empty(n, dtype=rd, device=device)
    S.modal_work_split(d_w, dev(re, rd, device), dev(im, rd, device), o_re, o_im, precision)
    assert np.array_equal(host(o_re), wt * re) and np.array_equal(host(o_im), wt * im)


@pytest.mark.parametrize("precision,shape", [(S.SKB_F64, (16, 16, 9, 3)), (S.SKB_F32, (32, 32, 17, 2)),
                                             (S.SKB_F64, (64, 32, 7, 1))])
def test_wvm_modal_work_and_modal_gather_bitwise(device, precision, shape):
    import torch

    note("wvm modal work %s p=%d" % (shape, precision))
    w, ow, modes, groups, fwd, inv = family(shape)
    om = O.retained_modes(ow)
    plan = S.WvmVerticalPlan(w, modes, w.nj, groups, fwd, inv, precision)
    cd, rd = S.complex_dtype(precision), S.real_dtype(precision)
    rng = np.random.default_rng(7)
    n = plan.modal_elements
    full = rounded(rng.standard_normal(n), precision) + 1j * rounded(rng.standard_normal(n), precision)
    wt = rounded(S.synthetic_modal_weights(w, modes), precision)
    d_in = dev(full, cd, device)
    sentinel = np.full(n, 3.5 - 1.25j)
    d_out = dev(sentinel, cd, device)
    S.wvm_modal_work(plan, dev(wt, rd, device), d_in, d_out)
    expect = O.apply_modal_work_wvm(ow, om, wt, full, sentinel)
    if precision == S.SKB_F32:
        expect = expect.astype(np.complex64)
    assert np.array_equal(host(d_out), expect)
    assert np.array_equal(host(d_in), full.astype(host(d_in).dtype))
    # gatherRetainedModal of the worked array
    fft = S.FftPlan(w, precision)
    compact = torch.empty(len(modes) * w.fields * w.nj, dtype=cd, device=device)
    fft.gather_retained_modal(modes, d_out, compact)
    g = O.gather_retained_modal(ow, om, expect.astype(np.complex128))
    assert np.array_equal(host(compact).astype(np.complex128), g)


# ---- the pipeline -------------------------------------------------------------------
def run_stages(pipe, w, x, device):
    import torch

    rd, cd = S.real_dtype(pipe.precision), S.complex_dtype(pipe.precision)
    nr, nm = pipe.modes * w.planes, pipe.modes * w.fields * w.nj
    d_x = dev(x, rd, device)
    arrays = {
        "output": torch.empty(w.real_elements, dtype=rd, device=device),
        "retained": torch.empty(nr, dtype=cd, device=device),
        "forward": torch.empty(nm, dtype=cd, device=device),
        "post": torch.empty(nm, dtype=cd, device=device),
        "inverse": torch.empty(nr, dtype=cd, device=device),
    }
    pipe.stages(d_x, arrays["output"], arrays["retained"], arrays["forward"], arrays["post"], arrays["inverse"])
    assert np.array_equal(host(d_x), np.asarray(x, host(d_x).dtype))  # input preserved
    return d_x, {k: host(v) for k, v in arrays.items()}


@pytest.mark.parametrize("policy", ["wvm", "compact"])
@pytest.mark.parametrize("precision", [S.SKB_F64, S.SKB_F32])
@pytest.mark.parametrize("shape,kind", [((8, 8, 7, 2), "impulse"), ((8, 8, 7, 2), "sinusoid"),
                                        ((8, 8, 7, 2), "random"), ((8, 8, 7, 2), "dc"), ((8, 8, 7, 2), "nyquist"),
                                        ((32, 32, 17, 3), "random"), ((64, 32, 33, 2), "random"),
                                        ((32, 64, 13, 2), "random")])
def test_pipeline_stages_and_execute_against_oracle(device, policy, precision, shape, kind):
    import torch

    note("pipeline %s %s p=%d %s" % (policy, shape, precision, kind))
    w, ow, modes, groups, fwd, inv = family(shape)
    om = O.retained_modes(ow)
    pipe = S.SpectralPipeline(w, modes, groups, fwd, inv, POLICIES[policy], precision)
    x = rounded(O.make_fixture(ow, kind, 5), precision)
    wt = O.synthetic_modal_weights(ow, om)
    ref = O.spectral_pipeline(ow, om, groups["first_mode"], groups["mode_count"], fwd, inv, wt, x)
    d_x, got = run_stages(pipe, w, x, device)
    tol = TOL[precision]
    for name in ("retained", "forward", "post", "inverse", "output"):
        err = O.max_relative_error(got[name].astype(ref[name].dtype), ref[name])
        assert err <= tol, (name, err)
    # execute: the default path equals the stage-by-stage run
    out = torch.empty(w.real_elements, dtype=S.real_dtype(precision), device=device)
    pipe.execute(d_x, out)
    if policy == "wvm":
        assert np.array_equal(host(out), got["output"])  # the same launches
    else:
        assert pipe.fused == (shape[2] != 13)  # Nz 13 (Nj 8) has no one-kernel round trip
        assert O.max_relative_error(host(out).astype(np.float64), ref["output"]) <= tol
        pipe.set_fused(False)
        pipe.execute(d_x, out)
        assert np.array_equal(host(out), got["output"])  # unfused = the stage launches


@pytest.mark.parametrize("precision,shape", [(S.SKB_F32, (256, 256, 65, 3)), (S.SKB_F64, (256, 256, 65, 3)),
                                             (S.SKB_F32, (256, 256, 129, 1)), (S.SKB_F64, (128, 128, 129, 2)),
                                             (S.SKB_F32, (128, 128, 257, 1))])
def test_pipeline_full_size_compact_fused(device, precision, shape):
    import torch

    note("pipeline full compact %s p=%d" % (shape, precision))
    w, ow, modes, groups, fwd, inv = family(shape)
    om = O.retained_modes(ow)
    pipe = S.SpectralPipeline(w, modes, groups, fwd, inv, S.SKB_PIPELINE_COMPACT, precision)
    assert pipe.fused
    x = rounded(O.make_fixture(ow, "random", 129), precision)
    wt = O.synthetic_modal_weights(ow, om)
    ref = O.spectral_pipeline(ow, om, groups["first_mode"], groups["mode_count"], fwd, inv, wt, x)["output"]
    rd = S.real_dtype(precision)
    d_x = dev(x, rd, device)
    out = torch.empty(w.real_elements, dtype=rd, device=device)
    pipe.execute(d_x, out)
    err = O.max_relative_error(host(out).astype(np.float64), ref)
    assert err <= TOL[precision], err
    # in place: real input and output may be the same array
    pipe.execute(d_x, d_x)
    assert np.array_equal(host(d_x), host(out))


def test_pipeline_full_size_wvm_direct(device):
    import torch

    shape = (256, 256, 65, 3)
    note("pipeline full wvm %s" % (shape,))
    w, ow, modes, groups, fwd, inv = family(shape)
    om = O.retained_modes(ow)
    pipe = S.SpectralPipeline(w, modes, groups, fwd, inv, S.SKB_PIPELINE_WVM_DIRECT, S.SKB_F64)
    x = O.make_fixture(ow, "random", 129)
    wt = O.synthetic_modal_weights(ow, om)
    ref = O.spectral_pipeline(ow, om, groups["first_mode"], groups["mode_count"], fwd, inv, wt, x)["output"]
    out = torch.empty(w.real_elements, dtype=torch.float64, device=device)
    pipe.execute(dev(x, torch.float64, device), out)
    assert O.max_relative_error(host(out), ref) <= 1e-12


@pytest.mark.parametrize("policy", ["wvm", "compact"])
def test_pipeline_eigenfunction_full_size_fp32(device, policy):
    """Size-independent property at 256^2 x Nz65 x 3: one horizontal mode
    times one column of its group's inverse matrix comes back multiplied by
    Nx*Ny*weight; custom weights make every (mode, j, field) distinct."""
    import torch

    shape = (256, 256, 65, 3)
    note("pipeline eigenfunction %s" % policy)
    w, ow, modes, groups, fwd, inv = family(shape)
    rng = np.random.default_rng(23)
    weights = 0.5 + rng.random(len(modes) * w.fields * w.nj)
    pipe = S.SpectralPipeline(w, modes, groups, fwd, inv, POLICIES[policy], S.SKB_F32, weights)
    for (k, l), j, f in [((-5, 7), 11, 2), ((0, 9), 3, 0), ((-40, 60), 30, 1), ((17, 0), 41, 2)]:
        m = int(np.nonzero((modes["k"] == k) & (modes["l"] == l))[0][0])
        g = int(np.searchsorted(groups["first_mode"], m, side="right") - 1)
        Q = inv.reshape(len(groups), w.nz, w.nj)[g]
        ix = np.arange(w.nx)[None, :]
        iy = np.arange(w.ny)[:, None]
        plane = np.cos(2 * np.pi * (k * ix / w.nx + l * iy / w.ny) + 0.3)
        x = np.zeros((w.fields, w.nz, w.ny, w.nx))
        x[f] = Q[:, j][:, None, None] * plane[None]
        x = x.reshape(-1)
        d_x = dev(x, torch.float32, device)
        out = torch.empty_like(d_x)
        pipe.execute(d_x, out)
        expect = w.nx * w.ny * weights[j + w.nj * (f + w.fields * m)] * rounded(x, S.SKB_F32)
        err = O.max_relative_error(host(out).astype(np.float64), expect)
        assert err <= 1e-6, ((k, l), j, f, err)


def test_pipeline_custom_weights_and_validation(device):
    import torch

    shape = (32, 32, 9, 2)
    note("pipeline custom weights / validation")
    w, ow, modes, groups, fwd, inv = family(shape)
    om = O.retained_modes(ow)
    rng = np.random.default_rng(9)
    weights = rng.standard_normal(len(modes) * w.fields * w.nj)
    x = O.make_fixture(ow, "random", 3)
    ref = O.spectral_pipeline(ow, om, groups["first_mode"], groups["mode_count"], fwd, inv, weights, x)["output"]
    for policy in POLICIES.values():
        pipe = S.SpectralPipeline(w, modes, groups, fwd, inv, policy, S.SKB_F64, weights)
        out = torch.empty(w.real_elements, dtype=torch.float64, device=device)
        pipe.execute(dev(x, torch.float64, device), out)
        assert O.max_relative_error(host(out), ref) <= 1e-12
        assert pipe.device_bytes() > 0
    wvm = S.SpectralPipeline(w, modes, groups, fwd, inv, S.SKB_PIPELINE_WVM_DIRECT, S.SKB_F64)
    with pytest.raises(S.Unsupported):
        wvm.set_fused(True)
    with pytest.raises(S.InvalidArgument, match="boundary-policy"):
        S.SpectralPipeline(w, modes, groups, fwd, inv, 7, S.SKB_F64)
    with pytest.raises(ValueError):
        S.SpectralPipeline(w, modes, groups, fwd, inv, S.SKB_PIPELINE_COMPACT, S.SKB_F64, weights[:-1])
    torch.cuda.synchronize()
