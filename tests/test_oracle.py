"""Pins the CPU oracle (oracle/) before anything is compared against it.

Each check is chosen so that a plausible slip (dropped term, wrong sign or
index, transposed operand) fails: published golden hashes for the integer
maps, closed forms for the DFT, an independent FFT (pocketfft) against the
direct DFT, exact permutation round trips, orthogonality invariants and the
reference's own known-answer test for the pointwise product.
"""
import math

import numpy as np
import pytest

import oracle as O

SMOKE = O.Workload(8, 8, 7, 2)


# ---- golden integer maps ------------------------------------------------------
def test_smoke_geometry():
    # tests/skbench_tests.cpp:1076-1079
    assert SMOKE.planes == 14 and SMOKE.nx_half == 5 and SMOKE.nj == 4


def test_smoke_k2_groups_golden():
    # tests/skbench_tests.cpp:1986-2000
    modes = O.retained_modes(SMOKE)
    assert modes[0]["k"] == 0 and modes[0]["l"] == 0
    key, first, count = O.k2_groups(modes)
    assert len(key) == 5 and key[0] == 0 and first[0] == 0
    assert int(count.sum()) == len(modes)
    assert O.group_hash(key, first, count) == "fnv1a64:69843c1f93cead00"


def _small_golden(golden_workloads):
    return {k: v for k, v in golden_workloads.items() if v["nx"] <= 512 and v["nz"] * v["fields"] <= 400}


def test_published_mode_and_group_hashes(golden_workloads):
    seen = set()
    for key, e in golden_workloads.items():
        if "retainedModeOrder" not in e:
            continue
        grid = (e["nx"], e["ny"])
        if grid in seen:
            continue
        seen.add(grid)
        w = O.Workload(e["nx"], e["ny"], e["nz"], e["fields"])
        modes = O.retained_modes(w)
        assert len(modes) == e["nkl"], key
        assert O.mode_order_hash(modes) == e["retainedModeOrder"], key
        if "k2GroupOrderHash" in e:
            k, f, c = O.k2_groups(modes)
            assert len(k) == e["k2GroupCount"], key
            assert O.group_hash(k, f, c) == e["k2GroupOrderHash"], key
            stats = e["k2GroupModes"]
            assert (int(c.min()), int(np.median(c)), int(c.max())) == (stats["minimum"], stats["median"], stats["maximum"])
    assert (256, 256) in seen and (512, 512) in seen and (1024, 1024) in seen


def test_published_wvm_order_hashes(golden_workloads):
    for key, e in _small_golden(golden_workloads).items():
        w = O.Workload(e["nx"], e["ny"], e["nz"], e["fields"])
        assert O.wvm_order_hash(w) == e["wvmFullSpectrumOrder"], key


def test_published_byte_counts(golden_workloads):
    e = golden_workloads["256x256x65x3"]
    w = O.Workload(256, 256, 65, 3)
    assert w.real_elements * 8 == e["bytes_real"] == 102236160
    assert w.spectrum_elements * 16 == e["bytes_fullSpectrum"] == 103034880
    assert e["nkl"] * w.planes * 16 == e["bytes_retainedSpectrum"] == 35689680
    assert e["nkl"] * w.fields * w.nj * 16 == e["bytes_modalSpectrum"] == 23061024


def test_grouping_rejects_noncontiguous_keys():
    modes = O.retained_modes(SMOKE).copy()
    # swap the first and last records: the DC key now reappears at the end
    modes[[0, -1]] = modes[[-1, 0]]
    with pytest.raises(ValueError):
        O.k2_groups(modes)


# ---- direct DFT closed forms --------------------------------------------------
def _wvm(w, arr):
    return arr.reshape(w.ny, w.nx_half, w.planes)


@pytest.mark.parametrize("shape", [(8, 8, 7, 2), (8, 4, 3, 2), (16, 8, 2, 1)])
def test_direct_r2c_closed_forms(shape):
    w = O.Workload(*shape)
    amps = 1.0 + 0.125 * np.arange(w.planes)
    n = w.nx * w.ny
    # dc: a single bin NxNy*amp at (0,0)
    s = _wvm(w, O.direct_r2c(w, O.make_fixture(w, "dc", 1)))
    expected = np.zeros_like(s)
    expected[0, 0, :] = n * amps
    assert np.abs(s - expected).max() <= 1e-12 * n * amps.max()
    # nyquist: (-1)^(x+y) -> a single bin at (Nx/2, Ny/2)
    s = _wvm(w, O.direct_r2c(w, O.make_fixture(w, "nyquist", 1)))
    expected = np.zeros_like(s)
    expected[w.ny // 2, w.nx // 2, :] = n * amps
    assert np.abs(s - expected).max() <= 1e-12 * n * amps.max()
    # impulse at ((1+p) mod Nx, (2+3p) mod Ny): pure phase ramp
    s = _wvm(w, O.direct_r2c(w, O.make_fixture(w, "impulse", 1)))
    ky, kx = np.meshgrid(np.arange(w.ny), np.arange(w.nx_half), indexing="ij")
    for p in range(w.planes):
        x0, y0 = (1 + p) % w.nx, (2 + 3 * p) % w.ny
        ramp = amps[p] * np.exp(-2j * np.pi * (kx * x0 / w.nx + ky * y0 / w.ny))
        assert np.abs(s[:, :, p] - ramp).max() <= 1e-12 * amps.max()
    # sinusoid cos(2pi(x/Nx + 2y/Ny) + 0.17p): one stored bin (kx=1, ky=2)
    s = _wvm(w, O.direct_r2c(w, O.make_fixture(w, "sinusoid", 1)))
    expected = np.zeros_like(s)
    expected[2 % w.ny, 1, :] = 0.5 * n * amps * np.exp(1j * 0.17 * np.arange(w.planes))
    assert np.abs(s - expected).max() <= 1e-11 * n * amps.max()


def test_direct_round_trip_scales_by_grid():
    w = SMOKE
    x = O.make_fixture(w, "random", 131)
    back = O.direct_c2r(w, O.direct_r2c(w, x))
    assert O.max_relative_error(back, x, 1.0 / (w.nx * w.ny)) <= 1e-12


# ---- independent FFT (pocketfft) versus direct DFT -----------------------------
@pytest.mark.parametrize("shape", [(8, 8, 7, 2), (16, 16, 3, 2), (32, 8, 2, 1)])
@pytest.mark.parametrize("kind", ["impulse", "sinusoid", "random", "dc", "nyquist"])
def test_pocketfft_r2c_matches_direct(shape, kind):
    w = O.Workload(*shape)
    x = O.make_fixture(w, kind, 129)
    assert O.max_relative_error(O.fft_r2c_wvm(w, x, 1), O.direct_r2c(w, x)) <= 1e-12


@pytest.mark.parametrize("shape", [(8, 8, 7, 2), (16, 16, 3, 2)])
def test_pocketfft_c2r_matches_direct_on_inconsistent_spectra(shape):
    # An arbitrary complex half spectrum (not Hermitian-consistent in the kx=0
    # and kx=Nx/2 columns) pins the c2r convention: Re of the full sum.
    w = O.Workload(*shape)
    rng = np.random.default_rng(5)
    spec = rng.standard_normal(w.spectrum_elements) + 1j * rng.standard_normal(w.spectrum_elements)
    a = O.fft_c2r_wvm(w, spec, 1)
    b = O.direct_c2r(w, spec)
    assert O.max_relative_error(a, b) <= 1e-12


# ---- retained gather / embed ---------------------------------------------------
def test_gather_embed_permutation_round_trip():
    w = SMOKE
    modes = O.retained_modes(w)
    n = w.spectrum_elements
    wvm = np.arange(n) - 1j * np.arange(n)
    ret = O.gather_retained(w, modes, wvm)
    again = O.gather_retained(w, modes, O.embed_retained(w, modes, ret))
    assert np.array_equal(again, ret)


def test_embed_repairs_kx0_partners_and_zero_fills():
    w = O.Workload(16, 16, 2, 1)
    modes = O.retained_modes(w)
    rng = np.random.default_rng(3)
    ret = rng.standard_normal(len(modes) * w.planes) + 1j * rng.standard_normal(len(modes) * w.planes)
    full = _wvm(w, O.embed_retained(w, modes, ret))
    r = ret.reshape(len(modes), w.planes)
    touched = np.zeros(full.shape[:2], bool)
    for m, md in enumerate(modes):
        skx, sky = int(md["stored_kx"]), int(md["stored_ky"])
        v = np.conj(r[m]) if md["conj"] else r[m]
        assert np.array_equal(full[sky, skx], v)
        touched[sky, skx] = True
        if skx == 0 and sky != 0 and 2 * sky != w.ny:
            assert np.array_equal(full[(w.ny - sky) % w.ny, 0], np.conj(v))
            touched[(w.ny - sky) % w.ny, 0] = True
    assert np.all(full[~touched] == 0)


def test_direct_retained_equals_gathered_direct_spectrum():
    w = SMOKE
    modes = O.retained_modes(w)
    x = O.make_fixture(w, "random", 77)
    a = O.direct_retained(w, modes, x)
    b = O.gather_retained(w, modes, O.direct_r2c(w, x))
    assert O.max_relative_error(a, b) <= 1e-12


# ---- vertical matrices and products ---------------------------------------------
def test_orthonormal_vertical_rows():
    for nz in (7, 65, 129):
        nj = 2 * (nz - 1) // 3
        f, i = O.orthonormal_vertical(nz, nj)
        F = f.reshape(nj, nz)
        assert np.abs(F @ F.T - np.eye(nj)).max() <= 1e-13
        assert np.array_equal(i.reshape(nz, nj), F.T)
        # row j is the DCT-II basis vector: its DCT against row 1 of the
        # full orthogonal DCT matrix (scipy) agrees
        import scipy.fft

        D = scipy.fft.dct(np.eye(nz), type=2, norm="ortho", axis=0)
        assert np.abs(D[:nj] - F).max() <= 1e-13


def test_k2_family_invariants():
    w = O.Workload(32, 32, 65, 3)
    modes = O.retained_modes(w)
    keys, first, count = O.k2_groups(modes)
    nj, nz = w.nj, w.nz
    f, i = O.k2_vertical_family(nz, nj, keys)
    base, _ = O.orthonormal_vertical(nz, nj)
    B = base.reshape(nj, nz)
    F = f.reshape(len(keys), nj, nz)
    I = i.reshape(len(keys), nz, nj)
    for g, key in enumerate(keys):
        assert np.abs(F[g] @ F[g].T - np.eye(nj)).max() <= 1e-13
        assert np.array_equal(I[g], F[g].T)
        theta = 0.5 * math.pi * ((int(key) % 4093) + 1) / 4094.0
        # each (j, j+1) pair spans the same plane as the base rows, rotated by theta
        for j in range(0, nj - 1, 2):
            assert abs(F[g, j] @ B[j] - math.cos(theta)) <= 1e-13
            assert abs(F[g, j] @ B[j + 1] - math.sin(theta)) <= 1e-13
            assert abs(F[g, j + 1] @ B[j] + math.sin(theta)) <= 1e-13
        if nj % 2 == 1:
            sign = -1.0 if int(key) & 1 else 1.0
            assert np.abs(F[g, nj - 1] - sign * B[nj - 1]).max() == 0.0


def test_vertical_products_against_brute_force_and_round_trip():
    w = O.Workload(16, 16, 9, 3)
    modes = O.retained_modes(w)
    keys, first, count = O.k2_groups(modes)
    f, i = O.k2_vertical_family(w.nz, w.nj, keys)
    rng = np.random.default_rng(11)
    n = len(modes) * w.planes
    phys = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    modal = O.vertical_forward(w, w.nj, first, count, f, phys)
    # independent brute force with numpy on the (mode, field, z) view
    P = phys.reshape(len(modes), w.fields, w.nz)
    M = modal.reshape(len(modes), w.fields, w.nj)
    F = f.reshape(len(keys), w.nj, w.nz)
    for g in range(len(keys)):
        sl = slice(int(first[g]), int(first[g] + count[g]))
        assert np.abs(M[sl] - P[sl] @ F[g].T).max() <= 1e-13
    # modal round trip (tests/skbench_tests.cpp:2075)
    back = O.vertical_forward(w, w.nj, first, count, f, O.vertical_inverse(w, w.nj, first, count, i, modal))
    assert O.max_relative_error(back, modal) <= 1e-12


def test_qr_gaussian_family_is_the_unique_qr_factor():
    n, groups, seed = 65, 3, 129
    fwd = O.qr_gaussian_family(n, groups, seed).reshape(groups, n, n)
    for g in range(groups):
        Q = fwd[g].T
        G = O.gaussians(seed + g, n * n).reshape(n, n)
        R = Q.T @ G
        assert np.abs(Q.T @ Q - np.eye(n)).max() <= 1e-13
        assert np.abs(np.tril(R, -1)).max() <= 1e-12
        assert np.all(np.diag(R) > 0)


def test_binned_groups_cover_modes_in_order():
    w = O.Workload(256, 256, 65, 3)
    modes = O.retained_modes(w)
    ids, first, count = O.k2_binned_groups(modes, 32)
    assert len(ids) == 32 and int(first[0]) == 0 and int(count.sum()) == len(modes)
    assert np.all(np.diff(ids.astype(np.int64)) > 0)
    assert np.array_equal(first[1:], (first + count)[:-1])


# ---- pointwise known-answer test (tests/skbench_tests.cpp:1136-1168) ----------
def test_pointwise_known_answer():
    volume, scale = 37, 0.03125
    idx = np.arange(3 * volume)
    shared = (idx % 17) - 8.0
    deriv = ((3 * idx + 5) % 23) - 11.0
    expected = np.array([
        -scale * (shared[p] * deriv[p] + shared[volume + p] * deriv[volume + p]
                  + shared[2 * volume + p] * deriv[2 * volume + p]) for p in range(volume)])
    got = O.pointwise_advection(volume, scale, shared, deriv)
    assert O.max_relative_error(got, expected) <= 1e-15


# ---- composed retained operator ---------------------------------------------------
def test_retained_operator_closed_form_and_direct_composition():
    w = O.Workload(16, 16, 5, 2)
    modes = O.retained_modes(w)
    keys, first, count = O.k2_groups(modes)
    f, _ = O.k2_vertical_family(w.nz, w.nj, keys)
    x = O.spectral_gaussian(w, modes, 129)
    got = O.retained_operator(w, modes, x, w.nj, first, count, f, 1)
    # inverse then forward is Nx*Ny times identity on Hermitian-consistent input
    closed = (w.nx * w.ny) * O.vertical_forward(w, w.nj, first, count, f, x)
    assert O.max_relative_error(got, closed) <= 1e-12
    # and the FFT inverse equals the direct inverse of the embedded spectrum
    real_fft = O.retained_inverse(w, modes, x, 1)
    real_dir = O.direct_c2r(w, O.embed_retained(w, modes, x))
    assert O.max_relative_error(real_fft, real_dir) <= 1e-12


def test_error_metric_definitions():
    # maximumRelativeError normalises by max(1, max|expected|); relative L2
    # is ||a-b|| / ||b||.
    a = np.array([1 + 1j, 2.0, 0.0])
    b = np.array([1.0, 2.0, 4.0])
    assert O.max_relative_error(a, b) == pytest.approx(4.0 / 4.0)
    assert O.relative_l2_error(a, b) == pytest.approx(math.sqrt(1 + 16) / math.sqrt(1 + 4 + 16))
    assert O.max_relative_error(np.array([0.5]), np.array([0.25])) == pytest.approx(0.25)


def test_blas_vertical_equals_loop_vertical():
    w = O.Workload(32, 32, 65, 3)
    modes = O.retained_modes(w)
    ids, first, count = O.k2_binned_groups(modes, 32)
    f = O.qr_gaussian_family(w.nz, len(ids), 129)
    x = O.spectral_gaussian(w, modes, 5)
    a = O.vertical_forward(w, w.nz, first, count, f, x)
    b = O.vertical_forward_blas(w, w.nz, first, count, f, x)
    assert O.max_relative_error(b, a) <= 1e-14
    c = O.retained_operator_fast(w, modes, x, w.nz, first, count, f, 2)
    d = O.retained_operator(w, modes, x, w.nz, first, count, f, 2)
    assert O.max_relative_error(c, d) <= 1e-14


def test_wvm_direct_vertical_matches_compact_path_and_partner_symmetry():
    """The WVM-direct product works on stored values; the compact path works on
    the gathered (k < 0: conjugated) values.  With real matrices
    A conj(x) = conj(A x), so gather(WVM-direct output) must equal the compact
    product of gather(input), and every k_x = 0 partner must hold the
    conjugate of its mode."""
    w = O.Workload(16, 16, 7, 2)
    modes = O.retained_modes(w)
    keys, first, count = O.k2_groups(modes)
    nj = 2 * (w.nz - 1) // 3
    fwd, inv = O.k2_vertical_family(w.nz, nj, keys)
    rng = np.random.default_rng(5)
    full = rng.standard_normal(w.spectrum_elements) + 1j * rng.standard_normal(w.spectrum_elements)
    wm = O.Workload(16, 16, nj, 2)  # modal arrays use the WVM layout with Nj in place of Nz
    out = O.wvm_direct_vertical(w, modes, nj, first, count, fwd, full, True)
    compact = O.vertical_forward(w, nj, first, count, fwd, O.gather_retained(w, modes, full))
    assert O.max_relative_error(O.gather_retained(wm, modes, out), compact) < 1e-13
    back = O.wvm_direct_vertical(w, modes, nj, first, count, inv, out, False)
    compact_back = O.vertical_inverse(w, nj, first, count, inv, compact)
    assert O.max_relative_error(O.gather_retained(w, modes, back), compact_back) < 1e-13
    nxh = w.nx // 2 + 1
    blocks = out.reshape(nxh * w.ny, w.fields, nj)
    seen = 0
    for m in modes:
        if m["stored_kx"] == 0 and m["stored_ky"] != 0 and 2 * m["stored_ky"] != w.ny:
            q = nxh * int(m["stored_ky"])
            qp = nxh * ((w.ny - int(m["stored_ky"])) % w.ny)
            assert np.array_equal(blocks[qp], np.conj(blocks[q]))
            seen += 1
    assert seen > 0
    # frequencies that are neither retained nor partners stay zero
    touched = set(int(m["stored_kx"]) + nxh * int(m["stored_ky"]) for m in modes)
    touched |= set(nxh * ((w.ny - int(m["stored_ky"])) % w.ny) for m in modes if m["stored_kx"] == 0)
    untouched = [f for f in range(nxh * w.ny) if f not in touched]
    assert untouched and np.all(blocks[untouched] == 0)


def test_flux_oracle_direct_and_fft_compositions_agree():
    w = O.Workload(8, 8, 7, 2)  # 2 targets: 9 input fields
    modes = O.retained_modes(w)
    keys, first, count = O.k2_groups(modes)
    nj = 2 * (w.nz - 1) // 3
    fwd, inv = O.k2_vertical_family(w.nz, nj, keys)
    x = O.flux_modal_fixture(w, modes, nj, 3 + 3 * w.fields, 41)
    a = O.production_flux(w, modes, nj, first, count, fwd, inv, x, w.fields, use_direct=True)
    b = O.production_flux(w, modes, nj, first, count, fwd, inv, x, w.fields, use_direct=False)
    assert O.max_relative_error(b, a) < 1e-12
    assert np.abs(a).max() > 0


def test_flux_oracle_dc_closed_form():
    """DC-only input: every physical field is constant on each plane, so the
    product is constant, the raw forward transform puts NxNy * product in the
    DC bin and nothing elsewhere, and the DC modal output is
    F_dc (NxNy * -(u qx + v qy + w qz) / (NxNy)^2) per target."""
    w = O.Workload(16, 16, 9, 2)
    modes = O.retained_modes(w)
    keys, first, count = O.k2_groups(modes)
    nj = 2 * (w.nz - 1) // 3
    fwd, inv = O.k2_vertical_family(w.nz, nj, keys)
    nin = 3 + 3 * w.fields
    rng = np.random.default_rng(9)
    x = np.zeros((len(modes), nin, nj), np.complex128)
    dc = int(np.nonzero((modes["k"] == 0) & (modes["l"] == 0))[0][0])
    x[dc] = rng.standard_normal((nin, nj))
    got = O.production_flux(w, modes, nj, first, count, fwd, inv, x.reshape(-1), w.fields)
    got = got.reshape(len(modes), w.fields, nj)
    g = int(np.searchsorted(first.astype(np.int64), dc, side="right") - 1)
    F = fwd.reshape(len(first), nj, w.nz)[g]
    I = inv.reshape(len(first), w.nz, nj)[g]
    phys = x[dc].real @ I.T  # (nin, nz): z profile of every field
    n2 = float(w.nx * w.ny)
    for t in range(w.fields):
        d = 3 + 3 * t
        prod = -(phys[0] * phys[d] + phys[1] * phys[d + 1] + phys[2] * phys[d + 2]) / n2 ** 2
        expect = F @ (n2 * prod)
        assert np.allclose(got[dc, t], expect, rtol=1e-12, atol=1e-14 * np.abs(expect).max())
    others = np.delete(got, dc, axis=0)
    assert np.abs(others).max() < 1e-12 * np.abs(got[dc]).max()


def test_vertical_derivative_matrix_exact_on_quadratics():
    nz, lz = 17, 3.0
    D = O.vertical_derivative_matrix(nz, lz).reshape(nz, nz)
    z = np.linspace(0.0, lz, nz)
    assert np.allclose(D @ np.ones(nz), 0.0, atol=1e-12)
    assert np.allclose(D @ z, np.ones(nz), atol=1e-12)
    assert np.allclose(D @ z ** 2, 2 * z, atol=1e-11)


def test_advection_flux_oracle_closed_forms_and_direct_dft():
    w = O.Workload(16, 16, 9, 4, 2.0, 3.0)
    modes = O.retained_modes(w)
    n, nz = len(modes), w.nz
    D = O.vertical_derivative_matrix(nz, 1.5)
    dc = int(np.nonzero((modes["k"] == 0) & (modes["l"] == 0))[0][0])
    # (-2, 3) is stored conjugated at (2, -3): exercises the conjugation
    m = int(np.nonzero((modes["k"] == -2) & (modes["l"] == 3))[0][0])
    n2 = float(w.nx * w.ny)
    # constant horizontal flow, q one Fourier mode constant in z (q_z = 0)
    U0, V0, A0 = 0.7, -1.3, 0.4 - 0.9j
    sp = np.zeros((n, 4, nz), np.complex128)
    sp[dc, 0, :], sp[dc, 1, :], sp[m, 3, :] = U0, V0, A0
    got = O.advection_flux(w, modes, D, sp.reshape(-1)).reshape(n, nz)
    kx, ly = 2 * np.pi * -2 / w.lx, 2 * np.pi * 3 / w.ly
    expect = -(U0 * 1j * kx + V0 * 1j * ly) * A0 / n2
    assert np.allclose(got[m], expect, rtol=1e-12)
    assert np.abs(np.delete(got, m, axis=0)).max() < 1e-12 * abs(expect)
    # vertical flow against q(z) = z^2 (the stencil is exact on quadratics)
    W0 = 1.1
    z = np.linspace(0.0, 1.5, nz)
    sp = np.zeros((n, 4, nz), np.complex128)
    sp[dc, 2, :] = W0
    sp[m, 3, :] = A0 * z ** 2
    got = O.advection_flux(w, modes, D, sp.reshape(-1)).reshape(n, nz)
    assert np.allclose(got[m], -W0 * A0 * 2 * z / n2, rtol=1e-11, atol=1e-14)
    # the FFT composition equals the direct DFT on random input
    rng = np.random.default_rng(3)
    sp = (rng.standard_normal((n, 4, nz)) + 1j * rng.standard_normal((n, 4, nz))).reshape(-1)
    a = O.advection_flux(w, modes, D, sp, use_direct=True)
    b = O.advection_flux(w, modes, D, sp)
    assert O.max_relative_error(b, a) < 1e-12
    # level chunks give the same values (q_z is formed before the chunking)
    for chunk in (1, 2, w.nz - 1):
        assert np.array_equal(O.advection_flux(w, modes, D, sp, z_chunk=chunk), b)
