"""Constant-stratification type-I vertical transforms on the GPU, through
the C ABI, against the oracle (pinned in tests/test_type1_oracle.py).

Mirrors the reference's constant-stratification vertical checks
(src/constant_stratification.cpp:520-567): the normalised DCT-I and DST-I
against the direct oracle, and the full production schedule (15 inverse and
4 forward complex channel transforms) on full-half and retained rows, with
the retained rows equal to the full-half control.  FP64 within 1e-12, FP32
within 1e-6 of the FP64 oracle (max-scale normalised).
"""
import sys

import numpy as np
import pytest

import oracle as O
import skb_hip as S

pytestmark = pytest.mark.gpu

TOL = {S.SKB_F64: 1e-12, S.SKB_F32: 1e-6}


def note(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def host(t):
    import torch

    torch.cuda.synchronize()
    return t.detach().cpu().numpy().astype(np.complex128)


def dev(a, precision, device):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to(device=device, dtype=S.complex_dtype(precision))


def rounded(a, precision):
    return a.astype(np.complex64).astype(np.complex128) if precision == S.SKB_F32 else a


def err(a, b):
    return np.abs(a - b).max() / max(1.0, np.abs(b).max())


@pytest.mark.parametrize("precision", [S.SKB_F64, S.SKB_F32])
@pytest.mark.parametrize("nz", [7, 65, 129, 13])
def test_each_kind_on_channel_subsets(device, precision, nz):
    if precision == S.SKB_F64 and nz > 132:
        pytest.skip("FP64 type-I lengths stop at 132")
    note("type1 kinds nz=%d p=%d" % (nz, precision))
    rows, channels = 37, 3
    plan = S.Type1Plan(rows, nz, precision)
    rng = np.random.default_rng(nz)
    x = rounded(rng.standard_normal((rows, channels, nz)) + 1j * rng.standard_normal((rows, channels, nz)), precision)
    for kind in range(4):
        for first, count in ((0, 3), (0, 2), (2, 1), (1, 1)):
            d = dev(x, precision, device)
            plan.execute(kind, d, channels, first, count)
            got = host(d)
            ref = O.type1(kind, x, first, count)
            sel = slice(first, first + count)
            assert err(got[:, sel], ref[:, sel]) <= TOL[precision], (kind, first, count)
            other = [c for c in range(channels) if not (first <= c < first + count)]
            assert np.array_equal(got[:, other], x[:, other])  # other channels untouched
            if kind == S.SKB_TYPE1_DST:
                assert np.array_equal(got[:, sel, 0], x[:, sel, 0]) and np.array_equal(got[:, sel, -1], x[:, sel, -1])


@pytest.mark.parametrize("precision", [S.SKB_F64, S.SKB_F32])
@pytest.mark.parametrize("shape", [(8, 8, 7), (16, 16, 9), (256, 256, 65)])
def test_schedule_full_and_compact_against_oracle(device, precision, shape):
    note("type1 schedule %s p=%d" % (shape, precision))
    nx, ny, nz = shape
    w = S.Workload(nx, ny, nz, 4)
    modes = S.retained_modes(w)
    om = O.retained_modes(O.Workload(nx, ny, nz, 4))
    nxh, full_rows = w.nx_half, w.nx_half * ny
    results = {}
    for name, rows, m in (("full", full_rows, None), ("compact", len(modes), modes)):
        a3 = rounded(S.type1_fill(rows, 3, nz, m, nxh), precision)
        a1 = rounded(S.type1_fill(rows, 1, nz, m, nxh), precision)
        plan = S.Type1Plan(rows, nz, precision)
        d3, d1 = dev(a3, precision, device), dev(a1, precision, device)
        plan.schedule(d3, d1)
        r3, r1 = O.type1_schedule(a3, a1)
        g3, g1 = host(d3), host(d1)
        assert err(g3, r3) <= TOL[precision] and err(g1, r1) <= TOL[precision], name
        results[name] = (g3, g1)
    q = om["stored_kx"].astype(np.int64) + nxh * om["stored_ky"].astype(np.int64)
    cj = om["conj"] != 0
    for full, comp in zip(results["full"], results["compact"]):
        g = full[q].copy()
        g[cj] = np.conj(g[cj])
        assert err(comp, g) <= TOL[precision]


def test_normalised_kinds_match_reference_direct_oracle(device):
    """The reference's first two correctness metrics: row 0 of the retained
    fill through the normalised DCT-I and DST-I against the direct sums."""
    note("type1 direct oracle")
    w = S.Workload(256, 256, 65, 4)
    modes = S.retained_modes(w)
    x = S.type1_fill(len(modes), 1, w.nz, modes, w.nx_half)
    plan = S.Type1Plan(len(modes), w.nz, S.SKB_F64)
    for kind, direct in ((S.SKB_TYPE1_DCT_FORWARD, O.direct_forward_dct), (S.SKB_TYPE1_DST_FORWARD, O.direct_forward_dst)):
        d = dev(x, S.SKB_F64, device)
        plan.execute(kind, d, 1)
        assert err(host(d)[0, 0], direct(x[0, 0])) <= 1e-12


def test_type1_validation(device):
    note("type1 validation")
    with pytest.raises(S.InvalidArgument):
        S.Type1Plan(10, 2, S.SKB_F64)
    with pytest.raises(S.Unsupported):
        S.Type1Plan(10, 200, S.SKB_F64)
    plan = S.Type1Plan(10, 7, S.SKB_F64)
    import torch

    d = torch.zeros(10 * 3 * 7, dtype=torch.complex128, device=device)
    with pytest.raises(S.InvalidArgument):
        plan.execute(0, d, 3, 2, 2)
    with pytest.raises(S.InvalidArgument):
        plan.execute(9, d, 3, 0, 1)
    torch.cuda.synchronize()
