"""The reference's analytic fixtures at the full 256^2 x Nz65 x 3 size, FP64,
through the HIP full and pruned transforms, against their closed forms
(src/model.cpp:618-661 fixtures; the direct-DFT identities they satisfy):

  dc        amp_p                       -> N amp_p at (0, 0)
  nyquist   amp_p (-1)^(x+y)            -> N amp_p at (Nx/2, Ny/2)
  impulse   amp_p at ((1+p) % Nx, (2+3p) % Ny)
                                        -> amp_p exp(-2 pi i (kx x0/Nx + ky y0/Ny))
  sinusoid-5-17  amp_p cos(2 pi (5x/Nx + 17y/Ny) + 0.17 p)
                                        -> N amp_p exp(0.17 i p) / 2 at (5, 17)

with N = Nx Ny and amp_p = 1 + p/8.  Tolerance: the reference's 1e-12 on the
max-scale-normalised error (src/benchmark.cpp:47, 118-127).  The inverse of
each closed-form spectrum returns N times the fixture; the pruned inverse of
the retained closed-form values returns N times the fixture's band-limited
part (the whole fixture for dc and the (5, 17) sinusoid, zero for nyquist,
whose bin lies outside the 2/3-rule set).
"""
import sys

import numpy as np
import pytest

import oracle as O
import skb_hip as S

pytestmark = pytest.mark.gpu

TOL = 1e-12
SHAPE = (256, 256, 65, 3)


def note(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def closed_form(w, kind):
    """Half spectrum in WVM order (plane fastest, then kx, then ky)."""
    P, nxh, N = w.planes, w.nx // 2 + 1, w.nx * w.ny
    amp = 1.0 + 0.125 * np.arange(P)
    s = np.zeros((w.ny, nxh, P), np.complex128)
    if kind == "dc":
        s[0, 0] = N * amp
    elif kind == "nyquist":
        s[w.ny // 2, w.nx // 2] = N * amp
    elif kind == "sinusoid-5-17":
        s[17, 5] = 0.5 * N * amp * np.exp(1j * 0.17 * np.arange(P))
    elif kind == "impulse":
        p = np.arange(P)
        x0, y0 = (1 + p) % w.nx, (2 + 3 * p) % w.ny
        ky = np.arange(w.ny)[:, None, None]
        kx = np.arange(nxh)[None, :, None]
        s[:] = amp * np.exp(-2j * np.pi * (kx * x0 / w.nx + ky * y0 / w.ny))
    return s.reshape(-1)


@pytest.mark.parametrize("kind", ["impulse", "sinusoid-5-17", "dc", "nyquist"])
def test_fixture_closed_forms_at_256_nz65_f3(device, kind):
    import torch

    w = S.Workload(*SHAPE)
    ow = O.Workload(*SHAPE)
    note("closed form %s at %s" % (kind, SHAPE))
    x = S.make_fixture(w, kind, 1)
    assert np.array_equal(x, O.make_fixture(ow, kind, 1))
    spec_ref = closed_form(w, kind)
    N = w.nx * w.ny

    full = S.FftPlan(w, S.SKB_F64)
    d_x = torch.from_numpy(x).to(device)
    spec = torch.empty(w.spectrum_elements, dtype=torch.complex128, device=device)
    full.forward(d_x, spec)
    torch.cuda.synchronize()
    assert O.max_relative_error(spec.cpu().numpy(), spec_ref) <= TOL
    real = torch.empty(w.real_elements, dtype=torch.float64, device=device)
    full.inverse(torch.from_numpy(spec_ref).to(device), real)
    torch.cuda.synchronize()
    assert O.max_relative_error(real.cpu().numpy(), N * x) <= TOL

    modes = S.retained_modes(w)
    om = O.retained_modes(ow)
    retained_ref = O.gather_retained(ow, om, spec_ref)
    pruned = S.PrunedPlan(w, modes, S.SKB_F64)
    retained = torch.empty(len(modes) * w.planes, dtype=torch.complex128, device=device)
    pruned.forward(d_x, retained)
    torch.cuda.synchronize()
    got = retained.cpu().numpy()
    if kind == "nyquist":
        assert np.abs(got).max() <= TOL * N * (1.0 + 0.125 * (w.planes - 1))  # the bin is not retained
    else:
        assert O.max_relative_error(got, retained_ref) <= TOL
    back = torch.empty(w.real_elements, dtype=torch.float64, device=device)
    pruned.inverse(torch.from_numpy(retained_ref).to(device), back)
    torch.cuda.synchronize()
    back = back.cpu().numpy()
    if kind in ("dc", "sinusoid-5-17"):
        assert O.max_relative_error(back, N * x) <= TOL
    elif kind == "nyquist":
        assert np.abs(back).max() == 0.0
    else:  # the band-limited impulse: the full inverse of the zero-filled closed form
        assert O.max_relative_error(back, O.fft_c2r_wvm(ow, O.embed_retained(ow, om, retained_ref))) <= TOL
