"""Native-stride WVM field views (FFTWInterleavedWvmFieldViewProvider,
include/skbench/skbench.hpp:423-465; src/fftw_provider.cpp:1175-1230)
through skb_fft_forward_view / skb_fft_inverse_view: each field of an F-field
interleaved WVM array transformed in place by a one-field plan with frequency
stride F*Nz, against the pocketfft oracle of the whole array.  Odd Nz checks
that a view's forward never writes the neighbouring field's planes; the 16^2
grid goes through the plan's contiguous stage.  Tolerances: 1e-12 FP64 and
1e-6 FP32 of the max scale (src/benchmark.cpp:47, 118-127)."""
import sys

import numpy as np
import pytest

import oracle as O
import skb_hip as S

pytestmark = pytest.mark.gpu

CASES = [((256, 256, 65), 3, S.SKB_F64), ((256, 256, 65), 3, S.SKB_F32), ((16, 16, 7), 3, S.SKB_F64),
         ((128, 64, 4), 2, S.SKB_F32)]


def note(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("shape,F,precision", CASES)
def test_field_views_forward_and_inverse(device, shape, F, precision):
    import torch

    nx, ny, nz = shape
    note("field views %s F=%d p=%d" % (shape, F, precision))
    tol = 1e-12 if precision == S.SKB_F64 else 1e-6
    rd, cd = S.real_dtype(precision), S.complex_dtype(precision)
    whole = O.Workload(nx, ny, nz, F)
    x = O.make_fixture(whole, "random", 31)
    ref = O.fft_r2c_wvm(whole, x)
    one = S.Workload(nx, ny, nz, 1)
    plan = S.FftPlan(one, precision)
    vol = nx * ny * nz
    stride = F * nz
    d_x = torch.from_numpy(x).to(device=device, dtype=rd)
    base = torch.full((whole.spectrum_elements,), 7.0 + 7.0j, dtype=cd, device=device)
    plan.forward_view(d_x[vol:2 * vol], base, nz, stride)  # field 1 only
    torch.cuda.synchronize()
    got = base.cpu().numpy().astype(np.complex128).reshape(-1, F, nz)
    assert np.all(got[:, 0, :] == 7.0 + 7.0j) and np.all(got[:, 2:, :] == 7.0 + 7.0j), "other fields untouched"
    for f in range(F):
        if f != 1:
            plan.forward_view(d_x[f * vol:(f + 1) * vol], base, f * nz, stride)
    torch.cuda.synchronize()
    got = base.cpu().numpy().astype(np.complex128)
    assert O.max_relative_error(got, ref) <= tol
    # inverse of each view: N x the field (the views may be overwritten)
    work = torch.from_numpy(ref).to(device=device, dtype=cd)
    out = torch.empty(vol, dtype=rd, device=device)
    for f in range(F):
        plan.inverse_view(work, f * nz, stride, out)
        torch.cuda.synchronize()
        back = out.cpu().numpy().astype(np.float64)
        assert O.max_relative_error(back, x[f * vol:(f + 1) * vol], 1.0 / (nx * ny)) <= tol
    with pytest.raises(S.InvalidArgument):
        plan.forward_view(d_x[:vol], base, 0, nz - 1)
