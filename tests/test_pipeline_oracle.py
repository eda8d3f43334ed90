"""Pins the oracle of the synthetic spectral pipeline and its modal work
(src/model.cpp:246-379; src/benchmark.cpp:3085-3114), and checks the
product's host weights against it.  CPU only.

Pins:
  * weights: hand-computed values at chosen (mode, j, field), including a key
    that wraps modulo 97 and modes with negative k, plus the |k|, |l|
    symmetry;
  * modal work: the reference's own equivalences (split == interleaved and
    gather(WVM work(embed(x))) == interleaved work, exactly, as
    tests/skbench_tests.cpp:2443-2484 requires);
  * pipeline: a closed form -- a single horizontal mode times one column of
    its group's inverse matrix is an eigenfunction, returned multiplied by
    Nx*Ny*weight (unnormalised transforms) -- and agreement of the pocketfft
    composition with a direct-DFT composition on small grids.
"""
import numpy as np
import pytest

import oracle as O
import skb_hip as S

SMOKE = O.Workload(8, 8, 7, 2)


def _mode(modes, k, l):
    hit = np.nonzero((modes["k"] == k) & (modes["l"] == l))[0]
    assert len(hit) == 1, (k, l)
    return int(hit[0])


def test_weights_hand_values():
    modes = O.retained_modes(SMOKE)
    wt = O.synthetic_modal_weights(SMOKE, modes)
    nj, F = SMOKE.nj, SMOKE.fields
    assert wt.size == len(modes) * F * nj == 88

    def at(k, l, j, f):
        return wt[j + nj * (f + F * _mode(modes, k, l))]

    # DC, j = 0, field 0: key 0
    assert at(0, 0, 0, 0) == 0.75
    # (-2, 1), j 3, field 1: 34 + 29 + 21 + 11 = 95
    assert at(-2, 1, 3, 1) == pytest.approx(1.2447916666666667, abs=1e-15)
    # (-1, 2), j 3, field 1: 17 + 58 + 21 + 11 = 107 = 10 (mod 97)
    assert at(-1, 2, 3, 1) == pytest.approx(0.8020833333333334, abs=1e-15)
    # (1, 2), j 2, field 0: 17 + 58 + 14 = 89
    assert at(1, 2, 2, 0) == pytest.approx(1.2135416666666667, abs=1e-15)
    # |k| symmetry: (2, 1) and (-2, 1) carry the same weights
    a, b = _mode(modes, 2, 1), _mode(modes, -2, 1)
    blk = nj * F
    assert np.array_equal(wt[a * blk:(a + 1) * blk], wt[b * blk:(b + 1) * blk])
    # every weight is on the lattice 0.75 + n/192, n in [0, 96]
    n = (wt - 0.75) * 192
    assert np.allclose(n, np.round(n), atol=1e-9) and n.min() >= 0 and n.max() <= 96


@pytest.mark.parametrize("shape", [(8, 8, 7, 2), (64, 64, 33, 3), (256, 256, 65, 3)])
def test_host_weights_equal_oracle(shape):
    w = S.Workload(*shape)
    modes = S.retained_modes(w)
    got = S.synthetic_modal_weights(w, modes)
    ref = O.synthetic_modal_weights(O.Workload(*shape), O.retained_modes(O.Workload(*shape)))
    assert np.array_equal(got, ref)


def _reference_modal_input(count):
    # tests/skbench_tests.cpp:2454-2458
    i = np.arange(count)
    return (i % 31) / 31.0 - 1j * ((i % 29) / 29.0)


@pytest.mark.parametrize("shape", [(8, 8, 7, 2), (16, 16, 9, 3), (32, 16, 17, 1)])
def test_modal_work_equivalences(shape):
    w = O.Workload(*shape)
    modes = O.retained_modes(w)
    wt = O.synthetic_modal_weights(w, modes)
    x = _reference_modal_input(wt.size)
    interleaved = wt * x
    split = (wt * x.real) + 1j * (wt * x.imag)
    assert np.array_equal(split, interleaved)
    full = O.embed_retained_modal(w, modes, x)
    assert np.array_equal(O.gather_retained_modal(w, modes, full), x)
    worked = O.apply_modal_work_wvm(w, modes, wt, full, np.zeros_like(full))
    assert np.array_equal(O.gather_retained_modal(w, modes, worked), interleaved)
    # the WVM work rebuilds the embedded layout of the worked values
    assert np.array_equal(worked, O.embed_retained_modal(w, modes, interleaved))


def _family(w, modes):
    keys, first, count = O.k2_groups(modes)
    fwd, inv = O.k2_vertical_family(w.nz, w.nj, keys)
    return first, count, fwd, inv


@pytest.mark.parametrize("km", [(0, 0), (0, 2), (2, 1), (-2, 1), (-1, 3), (3, 0)])
def test_pipeline_eigenfunction_closed_form(km):
    w = O.Workload(16, 16, 9, 2)
    modes = O.retained_modes(w)
    first, count, fwd, inv = _family(w, modes)
    m = _mode(modes, *km)
    g = int(np.searchsorted(first, m, side="right") - 1)
    Q = inv.reshape(len(first), w.nz, w.nj)[g]
    rng = np.random.default_rng(11)
    weights = 0.5 + rng.random(len(modes) * w.fields * w.nj)
    j, f = 2, 1
    c = 0.7 - 0.4j if km != (0, 0) else 0.7
    ix = np.arange(w.nx)[None, :]
    iy = np.arange(w.ny)[:, None]
    plane = np.real(c * np.exp(2j * np.pi * (km[0] * ix / w.nx + km[1] * iy / w.ny)))
    x = np.zeros((w.fields, w.nz, w.ny, w.nx))
    x[f] = Q[:, j][:, None, None] * plane[None]
    res = O.spectral_pipeline(w, modes, first, count, fwd, inv, weights, x.reshape(-1))
    scale = w.nx * w.ny * weights[j + w.nj * (f + w.fields * m)]
    assert O.max_relative_error(res["output"], scale * x.reshape(-1)) < 1e-13
    # the modal intermediates hold the single coefficient at (m, j, f)
    post = res["post"].reshape(len(modes), w.fields, w.nj)
    mask = np.ones(post.shape, bool)
    mask[m, f, j] = False
    assert np.abs(post[mask]).max() < 1e-12 * np.abs(post[m, f, j])


@pytest.mark.parametrize("kind", ["impulse", "sinusoid", "random", "dc", "nyquist"])
def test_pipeline_fft_composition_equals_direct_dft(kind):
    w = SMOKE
    modes = O.retained_modes(w)
    first, count, fwd, inv = _family(w, modes)
    wt = O.synthetic_modal_weights(w, modes)
    x = O.make_fixture(w, kind, 41)
    res = O.spectral_pipeline(w, modes, first, count, fwd, inv, wt, x)
    # direct-DFT composition with the C loop products
    retained = O.gather_retained(w, modes, O.direct_r2c(w, x))
    modal = O.vertical_forward(w, w.nj, first, count, fwd, retained)
    back = O.vertical_inverse(w, w.nj, first, count, inv, wt * modal)
    out = O.direct_c2r(w, O.embed_retained(w, modes, back))
    assert O.max_relative_error(res["retained"], retained) < 1e-12
    assert O.max_relative_error(res["forward"], modal) < 1e-12
    assert O.max_relative_error(res["inverse"], back) < 1e-12
    assert O.max_relative_error(res["output"], out) < 1e-12


def test_blas_vertical_inverse_equals_loop():
    w = O.Workload(32, 32, 17, 3)
    modes = O.retained_modes(w)
    first, count, fwd, inv = _family(w, modes)
    rng = np.random.default_rng(4)
    n = len(modes) * w.fields * w.nj
    modal = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    a = O.vertical_inverse(w, w.nj, first, count, inv, modal)
    b = O.vertical_inverse_blas(w, w.nj, first, count, inv, modal)
    assert O.max_relative_error(b, a) <= 1e-14
