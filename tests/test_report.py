"""Result bundles in the reference's spectral-kernel-benchmark-v1 shape
(src/json.cpp:120-333), the samples CSV and `compare` (src/json.cpp:282-332).
CPU part: record assembly, the required-field check, CSV round trip.  GPU
part (marked): each kernel run end to end on the smoke profile."""
import io
import json

import numpy as np
import pytest

import report as R
import skb_hip as S


def _record(status_ok=True):
    w = S.Workload(8, 8, 7, 2)
    modes = S.retained_modes(w)
    groups = S.k2_groups(modes)
    corr = [R.metric("x versus x", np.ones(4), np.ones(4) * (1.0 if status_ok else 2.0), 1e-12)]
    timings = [R.timing("primitive", "raw FFT", "forward", 1024, [1e-5, 2e-5, 3e-5]),
               R.timing("adapter-component", "view", "forward", 0, [], state="elided")]
    d = R.execution_direction("out-of-place", False, "real-grid", "wvm", "real -> wvm", "x=1", 0)
    prov = R.provider_record(S, "hip", "alg", "wvm", timings, corr, [("raw forward FFT", "executed", "d")], 1, 0,
                             d, d)
    return R.assemble(S, "fft", "smoke", "float64", 129, 1, 3, w, modes, groups, "k2", [prov], "cpu-only")


def test_record_conforms_and_matches_published_workload_fields():
    rec = _record()
    assert R.check_record(rec) == []
    assert rec["status"] == "passed" and rec["numericType"] == {"id": "float64", "scalarBits": 64}
    wl = rec["workload"]
    # smoke geometry and hashes (tests/skbench_tests.cpp:1076-1079, 1986-2000)
    assert (wl["H"], wl["Nkl"], wl["Nj"], wl["planes"]) == (40, 11, 4, 14)
    assert wl["grouping"]["verticalGroupCount"] == 5
    assert wl["grouping"]["verticalGroupOrderHash"] == "fnv1a64:69843c1f93cead00"
    assert wl["stridesElements"]["fullSpectrum"] == {"z": 1, "field": 7, "kx": 14, "ky": 70}
    assert rec["providers"][0]["timings"][0]["medianSeconds"] == 2e-5
    json.dumps(rec)  # serialisable


def test_check_record_reports_problems():
    rec = _record()
    del rec["run"]["seed"]
    rec["providers"][0]["timings"][0]["state"] = "bogus"
    rec["providers"][0]["correctness"][0]["tolerance"] = 0.0
    del rec["providers"][0]["executionContract"]["inverse"]["aliasing"]
    problems = R.check_record(rec)
    assert any("run: missing seed" in p for p in problems)
    assert any("state invalid" in p for p in problems)
    assert any("tolerance must be positive" in p for p in problems)
    assert any("executionContract.inverse: missing aliasing" in p for p in problems)
    assert _record(status_ok=False)["status"] == "failed"


def test_csv_and_compare(tmp_path):
    rec = _record()
    path = tmp_path / "samples.csv"
    R.write_csv(rec, str(path))
    lines = path.read_text().splitlines()
    assert lines[0] == "run_id,provider,scope,stage,direction,state,bytes_moved,sample_index,seconds"
    assert len(lines) == 1 + 3 + 1  # three samples, one unsampled (elided) row
    out = io.StringIO()
    assert R.compare_csv(str(path), out) == 0
    rows = out.getvalue().splitlines()
    assert rows[0] == "provider\tscope\tstage\tdirection\tmedian_ms\tsamples"
    assert rows[1] == "hip\tprimitive\traw FFT\tforward\t0.020000\t3"


def test_validate_host_references_match_the_oracle():
    """The validate command's host DFT sums and gather equal the oracle's."""
    import oracle as O

    w = S.Workload(8, 8, 7, 2)
    ow = O.Workload(8, 8, 7, 2)
    for kind in R.VALIDATE_FIXTURES:
        x = O.make_fixture(ow, kind, 5)
        spec = R._host_r2c(w, x)
        assert O.max_relative_error(spec, O.direct_r2c(ow, x)) <= 1e-13
        assert np.array_equal(R._host_gather(w, S.retained_modes(w), spec),
                              O.gather_retained(ow, O.retained_modes(ow), spec))


def test_probe_columns_follow_reference_rule():
    assert R.probe_columns(1) == [0]
    assert R.probe_columns(5) == [0, 1, 2, 3, 4]
    p = R.probe_columns(34317)
    assert len(p) == 17 and p[0] == 0 and p[-1] == 34316


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", ["fft", "pruned-horizontal", "vertical-gemm", "spectral-pipeline",
                                    "production-lifetime-flux"])
@pytest.mark.parametrize("precision", ["float64", "float32"])
def test_report_runs_on_device(device, kernel, precision, tmp_path):
    print("[case] report %s %s" % (kernel, precision), flush=True)
    rec = R.run(kernel, "smoke", precision, warmups=1, samples=3)
    assert R.check_record(rec) == []
    assert rec["status"] == "passed", rec["providers"][0]["correctness"]
    assert all(t["medianSeconds"] > 0 for t in rec["providers"][0]["timings"])
    R.write_json(rec, str(tmp_path / "r.json"))
    R.write_csv(rec, str(tmp_path / "r.csv"))
    assert R.compare_csv(str(tmp_path / "r.csv"), io.StringIO()) == 0


@pytest.mark.gpu
def test_report_quick_profile_fp32_vertical(device):
    print("[case] report quick vertical-gemm float32", flush=True)
    rec = R.run("vertical-gemm", "quick", "float32", warmups=1, samples=5)
    assert rec["status"] == "passed" and R.check_record(rec) == []
    assert rec["workload"]["permutationHashes"]["retainedModeOrder"] == "fnv1a64:22e98f1dcda40ad5"


@pytest.mark.gpu
def test_validate_passes_and_prints_the_reference_lines(device):
    print("[case] report validate", flush=True)
    out = io.StringIO()
    assert R.validate("smoke", out) == 0
    lines = out.getvalue().splitlines()
    assert lines[-1] == "validation=passed"
    assert "retained-mode-count=11" in lines  # Nkl of the 8 x 8 smoke grid
    assert "retained-mode-order-hash=fnv1a64:" in " ".join(lines)
    for kind in R.VALIDATE_FIXTURES:
        for check in ("hip/%s/forward=" % kind, "hip/%s/inverse=" % kind, "operator/%s/retained=" % kind,
                      "operator/%s/horizontal-vertical=" % kind, "operator/%s/vertical-modal-round-trip=" % kind):
            assert any(line.startswith(check) for line in lines), check
