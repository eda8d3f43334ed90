"""Pins the oracle of the constant-stratification type-I vertical transforms
(src/constant_stratification.cpp:95-292) and checks the product's host
matrices and fill against it.  CPU only.

Pins: scipy's DCT/DST type 1 (an independent library routine) against the
reference's own direct sums (directForwardDct / directForwardDst), closed
forms (a constant is a pure DC mode of DCT-I; a sine of index k0 is a pure
mode of DST-I), the involutions DCT-I o DCT-I = DST-I o DST-I = 2(Nz-1) I,
hand values of the fill, and the reference's compact-versus-full check of the
schedule.
"""
import numpy as np
import pytest

import oracle as O
import skb_hip as S


@pytest.mark.parametrize("nz", [3, 7, 17, 65, 129])
def test_forward_kinds_equal_reference_direct_sums(nz):
    rng = np.random.default_rng(nz)
    x = rng.standard_normal(nz) + 1j * rng.standard_normal(nz)
    dct = O.type1(O.TYPE1_DCT_FORWARD, x[None, None, :])[0, 0]
    dst = O.type1(O.TYPE1_DST_FORWARD, x[None, None, :])[0, 0]
    assert np.abs(dct - O.direct_forward_dct(x)).max() < 1e-13 * max(1.0, np.abs(dct).max())
    assert np.abs(dst - O.direct_forward_dst(x)).max() < 1e-13 * max(1.0, np.abs(dst).max())


@pytest.mark.parametrize("nz", [5, 9, 65])
def test_type1_closed_forms_and_involutions(nz):
    n1 = nz - 1
    c = 0.3 - 1.1j
    const = np.full((1, 1, nz), c)
    y = O.type1(O.TYPE1_DCT, const)[0, 0]
    assert abs(y[0] - 2 * n1 * c) < 1e-12 and np.abs(y[1:]).max() < 1e-12
    k0 = 2
    s = np.zeros((1, 1, nz), np.complex128)
    s[0, 0, 1:-1] = c * np.sin(np.pi * k0 * np.arange(1, n1) / n1)
    s[0, 0, 0], s[0, 0, -1] = 5.0, -7.0  # endpoints: untouched by the raw DST
    y = O.type1(O.TYPE1_DST, s)[0, 0]
    expect = np.zeros(nz, np.complex128)
    expect[k0] = n1 * c
    expect[0], expect[-1] = 5.0, -7.0
    assert np.abs(y - expect).max() < 1e-12
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2, 3, nz)) + 1j * rng.standard_normal((2, 3, nz))
    twice = O.type1(O.TYPE1_DCT, O.type1(O.TYPE1_DCT, x))
    assert np.abs(twice - 2 * n1 * x).max() < 1e-11
    twice = O.type1(O.TYPE1_DST, O.type1(O.TYPE1_DST, x, 1, 2), 1, 2)
    assert np.abs(twice[:, 1:, 1:-1] - 2 * n1 * x[:, 1:, 1:-1]).max() < 1e-11
    assert np.array_equal(twice[:, 0], x[:, 0])  # channel 0 untouched
    assert np.array_equal(twice[:, 1:, 0], x[:, 1:, 0]) and np.array_equal(twice[:, 1:, -1], x[:, 1:, -1])


@pytest.mark.parametrize("nz", [3, 7, 65, 129])
def test_host_matrices_equal_oracle(nz):
    eye = np.eye(nz).T.astype(np.complex128)  # columns of the identity as nz rows of one channel
    for kind in range(4):
        ref = O.type1(kind, eye[:, None, :])[:, 0, :].T  # column j = transform of e_j
        assert np.abs(S.type1_matrix(nz, kind) - ref.real).max() < 4e-15 * (nz - 1)


def test_fill_hand_values_and_host_equal_oracle():
    v = O.type1_fill(2, 3, 4)
    a = 0.017 + 0.113 + 0.071  # row 0, channel 0, z 0
    assert v[0, 0, 0] == pytest.approx(np.sin(a) + 1j * np.cos(1.7 * a), abs=1e-15)
    a = 0.017 * 2 + 0.113 * 3 + 0.071 * 4  # row 1, channel 2, z 3
    assert v[1, 2, 3] == pytest.approx(np.sin(a) + 1j * np.cos(1.7 * a), abs=1e-15)
    w = S.Workload(16, 16, 9, 1)
    modes = S.retained_modes(w)
    om = O.retained_modes(O.Workload(16, 16, 9, 1))
    assert np.abs(S.type1_fill(len(modes), 3, 9, modes, w.nx_half) - O.type1_fill(len(om), 3, 9, om, 9)).max() < 1e-15
    assert np.abs(S.type1_fill(100, 1, 9) - O.type1_fill(100, 1, 9)).max() < 1e-15
    # row index wraps modulo 97
    assert np.array_equal(O.type1_fill(98, 1, 3)[97], O.type1_fill(98, 1, 3)[0])


def test_schedule_compact_equals_full_control():
    """src/constant_stratification.cpp:558-567: the retained rows of the
    full-half schedule equal the schedule on the compact rows."""
    w = O.Workload(16, 16, 9, 4)
    modes = O.retained_modes(w)
    nxh, rows = w.nx_half, w.nx_half * w.ny
    f3, f1 = O.type1_schedule(O.type1_fill(rows, 3, w.nz), O.type1_fill(rows, 1, w.nz))
    c3, c1 = O.type1_schedule(O.type1_fill(len(modes), 3, w.nz, modes, nxh), O.type1_fill(len(modes), 1, w.nz, modes, nxh))
    q = modes["stored_kx"].astype(np.int64) + nxh * modes["stored_ky"].astype(np.int64)
    cj = modes["conj"] != 0
    for full, comp in ((f3, c3), (f1, c1)):
        g = full[q].copy()
        g[cj] = np.conj(g[cj])
        assert np.abs(g - comp).max() < 1e-12 * np.abs(g).max()
