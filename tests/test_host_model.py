"""The product's host model (libskb_hip.so, C++) against golden values and the
oracle, plus the C ABI surface.  CPU only: none of these calls touch a GPU."""
import os
import re

import numpy as np
import pytest

import oracle as O
import skb_hip as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "skb_hip.h"), encoding="utf-8").read()
    declared = set(re.findall(r"^\s*(?:[a-z_0-9]+\s*\*?\s+)+\**(skb_[a-z0-9_]+)\s*\(", header, re.M))
    assert len(declared) >= 30
    for name in sorted(declared):
        assert hasattr(S.lib, name), name
    assert set(S.EXPORTED) == declared


def test_version_identity():
    assert "gfx950" in S.version()


@pytest.mark.parametrize("grid", [(8, 8), (16, 8), (128, 128), (256, 256), (512, 512)])
def test_modes_match_oracle_exactly(grid):
    w = S.Workload(grid[0], grid[1], 3, 2)
    a = S.retained_modes(w)
    b = O.retained_modes(O.Workload(grid[0], grid[1], 3, 2))
    assert len(a) == len(b)
    for f in ("k", "l", "stored_kx", "stored_ky", "conj", "radial"):
        assert np.array_equal(a[f], b[f]), f


def test_hashes_match_published(golden_workloads):
    done = set()
    for key, e in golden_workloads.items():
        if "retainedModeOrder" not in e or (e["nx"], e["ny"]) in done:
            continue
        done.add((e["nx"], e["ny"]))
        w = S.Workload(e["nx"], e["ny"], e["nz"], e["fields"])
        modes = S.retained_modes(w)
        assert len(modes) == e["nkl"]
        assert S.mode_order_hash(modes) == e["retainedModeOrder"]
        groups = S.k2_groups(modes)
        assert len(groups) == e["k2GroupCount"]
        assert S.group_order_hash(groups) == e["k2GroupOrderHash"]
    e = golden_workloads["256x256x65x3"]
    assert S.wvm_order_hash(S.Workload(256, 256, 65, 3)) == e["wvmFullSpectrumOrder"]
    assert S.group_order_hash(S.k2_groups(S.retained_modes(S.Workload(8, 8, 7, 2)))) == "fnv1a64:69843c1f93cead00"


def test_binned_groups_match_oracle():
    w = S.Workload(256, 256, 65, 3)
    modes = S.retained_modes(w)
    g = S.k2_binned_groups(modes, 32)
    ids, first, count = O.k2_binned_groups(O.retained_modes(O.Workload(256, 256, 65, 3)), 32)
    assert np.array_equal(g["squared_key"], ids)
    assert np.array_equal(g["first_mode"], first)
    assert np.array_equal(g["mode_count"], count)


def test_vertical_families_match_oracle():
    w = S.Workload(64, 64, 65, 3)
    modes = S.retained_modes(w)
    groups = S.k2_groups(modes)
    f, i = S.k2_vertical_family(w.nz, w.nj, groups)
    of, oi = O.k2_vertical_family(w.nz, w.nj, groups["squared_key"].copy())
    assert np.abs(f - of).max() <= 1e-15 and np.abs(i - oi).max() <= 1e-15
    f, i = S.orthonormal_vertical(65, 42)
    of, oi = O.orthonormal_vertical(65, 42)
    assert np.abs(f - of).max() <= 1e-15 and np.abs(i - oi).max() <= 1e-15


def test_qr_family_matches_oracle_and_is_orthogonal():
    f, i = S.qr_gaussian_family(65, 2, 129)
    of = O.qr_gaussian_family(65, 2, 129)
    assert np.abs(f - of).max() <= 1e-12
    F = f.reshape(2, 65, 65)
    I = i.reshape(2, 65, 65)
    for g in range(2):
        assert np.abs(F[g] @ I[g] - np.eye(65)).max() <= 1e-13
        assert np.array_equal(I[g], F[g].T)


@pytest.mark.parametrize("kind", ["impulse", "sinusoid", "random", "dc", "nyquist", "sinusoid-5-17"])
def test_fixtures_match_oracle(kind):
    w = S.Workload(16, 8, 3, 2)
    a = S.make_fixture(w, kind, 129)
    b = O.make_fixture(O.Workload(16, 8, 3, 2), kind, 129)
    assert np.array_equal(a, b)


def test_spectral_fixture_matches_oracle():
    w = S.Workload(32, 32, 5, 3)
    modes = S.retained_modes(w)
    for exponent in (-1.0, -5.0 / 6.0):
        a = S.spectral_fixture(w, modes, 129, exponent)
        b = O.spectral_gaussian(O.Workload(32, 32, 5, 3), O.retained_modes(O.Workload(32, 32, 5, 3)), 129, exponent)
        assert np.abs(a - b).max() <= 1e-15
        assert a[: w.planes].imag.max() == 0.0  # DC mode is real


def test_invalid_workloads_are_rejected_before_any_device_work():
    with pytest.raises(S.InvalidArgument):
        S.retained_modes(S.Workload(1, 8, 3, 1))
    with pytest.raises(S.InvalidArgument):
        S.FftPlan(S.Workload(8, 8, 0, 1))
    with pytest.raises(S.Unsupported):
        S.FftPlan(S.Workload(12, 8, 3, 1))
    with pytest.raises(S.Unsupported):
        S.FftPlan(S.Workload(2048, 8, 1, 1))
    modes = S.retained_modes(S.Workload(8, 8, 2, 1))
    with pytest.raises(S.InvalidArgument):
        S.PrunedPlan(S.Workload(8, 8, 2, 1), modes[:0])
    bad = modes.copy()
    bad[1] = bad[0]  # two modes at the same stored slot
    with pytest.raises(S.InvalidArgument):
        S.PrunedPlan(S.Workload(8, 8, 2, 1), bad)
    groups = S.k2_groups(modes)
    f, i = S.k2_vertical_family(2, 0 or 1, groups)
    broken = groups.copy()
    broken["first_mode"][0] = 1
    with pytest.raises(S.InvalidArgument):
        S.VerticalPlan(S.Workload(8, 8, 2, 1), 1, broken, f, i)
