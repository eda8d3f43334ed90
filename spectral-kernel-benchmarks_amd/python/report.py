"""Result bundles of the hip provider in the reference's
spectral-kernel-benchmark-v1 shape (src/json.cpp:120-333; the field set of
schema/spectral-kernel-benchmark-v1.schema.json), with the reference's samples
CSV and its `compare` summary (src/json.cpp:282-332; src/main.cpp:229-238).

    python report.py run --kernel vertical-gemm --profile quick --precision float32 \\
        --json out.json --csv out.csv
    python report.py compare --input out.csv
    python report.py check out.json
    python report.py validate --profile smoke

Kernels: fft, pruned-horizontal, vertical-gemm, spectral-pipeline,
production-lifetime-flux.  `validate` is the reference's validateBenchmark
(src/benchmark.cpp:4898-5003, src/main.cpp:163-173) for the hip provider.  Each
record carries primitive-scope timings (device time of one call per sample,
after `warmups` untimed calls, median reported as in src/benchmark.cpp:91-111)
and correctness metrics computed by the harness itself -- two independent
device paths, or host sums on probe columns (src/benchmark.cpp:205-273) --
with the reference's error metrics (src/model.cpp:718-751) and tolerance
(1e-12 FP64; 1e-6 for FP32 storage).
"""
from __future__ import annotations

import argparse
import csv
import datetime
import json
import os
import platform
import socket
import subprocess
import sys

import numpy as np

SCHEMA = "spectral-kernel-benchmark-v1"
STATES = ("executed", "fused", "elided", "setup-only", "unsupported")

# Profiles of src/benchmark.cpp:1039-1057: (Nx, Ny, Nz, fields, warmups, samples).
PROFILES = {
    "smoke": (8, 8, 7, 2, 1, 2),
    "quick": (256, 256, 65, 3, 3, 15),
    "exhaustive": (512, 512, 129, 4, 3, 20),
    "wvm-historical-256-nz65-f3": (256, 256, 65, 3, 3, 15),
    "wvm-historical-256-nz65-f4": (256, 256, 65, 4, 3, 15),
    "wvm-historical-512-nz129-f3": (512, 512, 129, 3, 3, 20),
    "wvm-historical-512-nz129-f4": (512, 512, 129, 4, 3, 20),
    "wvm-current-256-nz129-f1": (256, 256, 129, 1, 3, 15),
    "wvm-current-256-nz129-f3": (256, 256, 129, 3, 3, 15),
    "wvm-current-256-nz129-f4": (256, 256, 129, 4, 3, 15),
    "wvm-current-512-nz257-f1": (512, 512, 257, 1, 3, 20),
    "wvm-current-512-nz257-f3": (512, 512, 257, 3, 3, 20),
    "wvm-current-512-nz257-f4": (512, 512, 257, 4, 3, 20),
}

# ---- error metrics (src/model.cpp:718-751) -----------------------------------------


def max_relative_error(actual, expected) -> float:
    a = np.asarray(actual, np.complex128).ravel()
    b = np.asarray(expected, np.complex128).ravel()
    return float(np.abs(a - b).max(initial=0.0) / max(1.0, float(np.abs(b).max(initial=0.0))))


def relative_l2_error(actual, expected) -> float:
    a = np.asarray(actual, np.complex128).ravel()
    b = np.asarray(expected, np.complex128).ravel()
    num = float(np.sum(np.abs(a - b) ** 2))
    den = float(np.sum(np.abs(b) ** 2))
    return float(np.sqrt(num / den)) if den > 0 else float(np.sqrt(num))


def metric(name: str, actual, expected, tolerance: float) -> dict:
    m, l2 = max_relative_error(actual, expected), relative_l2_error(actual, expected)
    return {"name": name, "maximumRelativeError": m, "relativeL2Error": l2, "tolerance": tolerance,
            "passed": bool(m <= tolerance and l2 <= tolerance)}


def timing(scope: str, stage: str, direction: str, bytes_moved: int, samples, state: str = "executed") -> dict:
    samples = [float(s) for s in samples]
    return {"scope": scope, "stage": stage, "direction": direction, "state": state, "bytesMoved": int(bytes_moved),
            "medianSeconds": float(np.median(samples)) if samples else 0.0, "samplesSeconds": samples}


def probe_columns(columns: int):
    """verticalProbeColumns (src/benchmark.cpp:205-215): up to 17 columns spread
    evenly over [0, columns)."""
    count = min(17, columns)
    out = []
    for i in range(count):
        c = 0 if count == 1 else i * (columns - 1) // (count - 1)
        if not out or out[-1] != c:
            out.append(c)
    return out


# ---- environment and workload sections ------------------------------------------------


def _git():
    root = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    try:
        commit = subprocess.run(["git", "-C", root, "rev-parse", "HEAD"], capture_output=True, text=True,
                                timeout=10).stdout.strip()
        dirty = bool(subprocess.run(["git", "-C", root, "status", "--porcelain"], capture_output=True, text=True,
                                    timeout=10).stdout.strip())
    except (OSError, subprocess.SubprocessError):
        commit, dirty = "", False
    return commit or "unknown", dirty


def environment(device_name: str) -> dict:
    cpu = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo", encoding="utf-8") as handle:
            for line in handle:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        memory = os.sysconf("SC_PAGE_SIZE") * os.sysconf("SC_PHYS_PAGES")
    except (ValueError, OSError):
        memory = 0
    commit, dirty = _git()
    cores = os.cpu_count() or 1
    return {"timestampUtc": datetime.datetime.now(datetime.timezone.utc).strftime("%Y-%m-%dT%H:%M:%SZ"),
            "hostname": socket.gethostname(), "operatingSystem": platform.platform(),
            "machineModel": device_name, "cpuBrand": cpu, "totalCores": cores, "performanceCores": cores,
            "efficiencyCores": 0, "physicalMemoryBytes": int(memory), "compiler": "hipcc (ROCm)",
            "compilerVersion": "gfx950", "compilerFlags": "--offload-arch=gfx950 -O3 -std=c++20",
            "buildType": "Release", "gitCommit": commit, "gitDirty": dirty}


def workload_section(S, w, modes, groups, family_id: str, scalar_bytes: int) -> dict:
    nj = w.nj
    cb = 2 * scalar_bytes
    counts = groups["mode_count"].astype(np.int64) if len(groups) else np.zeros(1, np.int64)
    return {
        "Nx": w.nx, "Ny": w.ny, "Nz": w.nz, "fields": w.fields, "H": w.nx_half * w.ny, "Nkl": int(len(modes)),
        "Nj": nj, "planes": w.planes,
        "antialias": {"enabled": True, "horizontalPolicy": "radial-two-thirds",
                      "horizontalCutoffFraction": 2.0 / 3.0, "verticalPolicy": "floor(2*(Nz-1)/3)",
                      "verticalRetainedFraction": nj / (w.nz - 1) if w.nz > 1 else 0.0},
        "grouping": {"verticalMatrixFamilyId": family_id, "verticalGroupCount": int(len(groups)),
                     "verticalGroupModes": {"minimum": int(counts.min()), "median": float(np.median(counts)),
                                            "maximum": int(counts.max())},
                     "verticalGroupColumns": {"minimum": int(counts.min() * w.fields),
                                              "median": float(np.median(counts) * w.fields),
                                              "maximum": int(counts.max() * w.fields)},
                     "verticalGroupOrderHash": S.group_order_hash(groups) if len(groups) else ""},
        "stridesElements": {"real": {"x": 1, "y": w.nx, "z": w.nx * w.ny, "field": w.nx * w.ny * w.nz},
                            "fullSpectrum": {"z": 1, "field": w.nz, "kx": w.planes, "ky": w.planes * w.nx_half},
                            "retainedSpectrum": {"z": 1, "field": w.nz, "mode": w.planes},
                            "modalSpectrum": {"j": 1, "field": nj, "mode": nj * w.fields}},
        "bytes": {"real": w.real_elements * scalar_bytes, "fullSpectrum": w.spectrum_elements * cb,
                  "retainedSpectrum": len(modes) * w.planes * cb, "modalSpectrum": len(modes) * w.fields * nj * cb,
                  "verticalMatrixFamilySource": int(2 * len(groups) * nj * w.nz * 8)},
        "permutationHashes": {"retainedModeOrder": S.mode_order_hash(modes), "wvmFullSpectrumOrder": S.wvm_order_hash(w)},
    }


def execution_direction(placement: str, destroys: bool, native_in: str, native_out: str, extents: str,
                        strides: str, reusable: int) -> dict:
    return {"nativePlacement": placement, "adapterPlacement": placement, "destroysNativeInput": destroys,
            "adapterPreservesCallerInput": not destroys, "requiresPreservationCopyForRepeatedExecution": destroys,
            "preservationIncludedInPrimitiveTiming": False, "preservationIncludedInAdapterTiming": False,
            "nativeInputRepresentationId": native_in, "nativeOutputRepresentationId": native_out,
            "adapterInputRepresentationId": native_in, "adapterOutputRepresentationId": native_out,
            "physicalExtents": extents, "stridesElements": strides, "paddingElements": 0,
            "minimumAlignmentBytes": 256, "aliasing": "caller buffers do not overlap",
            "reusableWorkBytes": int(reusable), "outputCanFeedOppositeDirection": True}


def provider_record(S, pid: str, algorithm: str, representation: str, timings, correctness, ledger,
                    persistent: int, scratch: int, forward: dict, inverse: dict, gemm_calls: int = 0) -> dict:
    return {
        "id": pid, "version": S.version(), "libraryIdentity": "libskb_hip.so (hand-written HIP, gfx950)",
        "algorithmId": algorithm, "nativeRepresentationId": representation,
        "modeOrderId": "logical-radial-retained-modes-keyed-by-k-l-j-field", "schedulingId": "hip-stream-1",
        "providerBuild": {"sourceIdentity": "spectral-kernel-benchmarks_amd/csrc", "sourceSha256": "",
                          "configureFlags": "make -C spectral-kernel-benchmarks_amd",
                          "compilerFlags": "--offload-arch=gfx950 -O3"},
        "workers": 1, "gemmCallsPerExecution": int(gemm_calls),
        "scheduling": {"internalWorkers": 1, "outerWorkers": 1, "totalLogicalWorkers": 1},
        "executionContract": {"forward": forward, "inverse": inverse},
        "setup": {"totalSeconds": 0.0, "otherSeconds": 0.0, "allocationSeconds": 0.0},
        "planning": {"seconds": 0.0, "configuration": "plans built at construction; no runtime planning",
                     "temporaryBytes": 0},
        "memory": {"persistentBytes": int(persistent), "scratchBytes": int(scratch), "opaqueProviderMemory": False},
        "componentLedger": [{"stage": s, "state": st, "detail": d} for s, st, d in ledger],
        "timings": timings, "correctness": correctness,
    }


def assemble(S, kernel: str, profile: str, precision_id: str, seed: int, warmups: int, samples: int, w, modes,
             groups, family_id: str, providers, device_name: str) -> dict:
    scalar_bytes = 4 if precision_id == "float32" else 8
    passed = all(c["passed"] for p in providers for c in p["correctness"])
    env = environment(device_name)
    stamp = env["timestampUtc"].replace("-", "").replace(":", "")
    return {
        "schema": SCHEMA, "status": "passed" if passed else "failed",
        "numericType": {"id": precision_id, "scalarBits": 8 * scalar_bytes},
        "run": {"id": "%s-%s-%s" % (stamp, kernel, env["hostname"]), "profile": profile, "seed": int(seed),
                "warmups": int(warmups), "samples": int(samples)},
        "provenance": {"benchmarkRepository": "JeffreyEarly/spectral-kernel-benchmarks",
                       "waveVortexModelRepository": "JeffreyEarly/wave-vortex-model", "waveVortexModelIssue": 129,
                       "auditedProductionBaselineCommit": "be0f78995c49a2bfe4c43d75827856e3812ac278",
                       "historicalHarnessCommit": "b63fef8f6a9e400c6ec560205906578b763d8298",
                       "historicalDecisionCommit": "34797085d1b642b9d9c822f30a6b5c18e139c2bf"},
        "logicalOperators": {"horizontalForward": "T_h^+=P_h F_xy", "horizontalInverse": "T_h^-=F_xy^-1 E_h",
                             "verticalForward": "T_z^+=V_f", "verticalInverse": "T_z^-=V_i",
                             "verticalFixtureId": family_id, "resultKey": "(k,l,j,field)"},
        "workload": workload_section(S, w, modes, groups, family_id, scalar_bytes),
        "environment": env,
        "providers": providers,
    }


# ---- required-field check of the v1 record ---------------------------------------------
# The required keys and value constraints of the v1 format, restated so a record
# can be checked without the schema file.
_REQUIRED = {
    "": ("schema", "status", "run", "provenance", "logicalOperators", "workload", "environment", "providers"),
    "run": ("id", "profile", "seed", "warmups", "samples"),
    "provenance": ("benchmarkRepository", "waveVortexModelRepository", "waveVortexModelIssue",
                   "auditedProductionBaselineCommit", "historicalHarnessCommit", "historicalDecisionCommit"),
    "logicalOperators": ("horizontalForward", "horizontalInverse", "verticalForward", "verticalInverse",
                         "verticalFixtureId", "resultKey"),
    "workload": ("Nx", "Ny", "Nz", "fields", "H", "Nkl", "Nj", "planes", "antialias", "grouping", "stridesElements",
                 "bytes", "permutationHashes"),
    "workload.antialias": ("enabled", "horizontalPolicy", "horizontalCutoffFraction", "verticalPolicy",
                           "verticalRetainedFraction"),
    "environment": ("timestampUtc", "hostname", "operatingSystem", "machineModel", "cpuBrand", "totalCores",
                    "performanceCores", "efficiencyCores", "physicalMemoryBytes", "compiler", "compilerVersion",
                    "compilerFlags", "buildType", "gitCommit", "gitDirty"),
    "provider": ("id", "version", "libraryIdentity", "algorithmId", "nativeRepresentationId", "modeOrderId",
                 "schedulingId", "providerBuild", "workers", "setup", "planning", "memory", "componentLedger",
                 "timings", "correctness"),
    "provider.providerBuild": ("sourceIdentity", "sourceSha256", "configureFlags", "compilerFlags"),
    "provider.setup": ("totalSeconds", "otherSeconds", "allocationSeconds"),
    "provider.planning": ("seconds", "configuration", "temporaryBytes"),
    "provider.memory": ("persistentBytes", "scratchBytes", "opaqueProviderMemory"),
    "ledger": ("stage", "state", "detail"),
    "timing": ("scope", "stage", "direction", "state", "bytesMoved", "medianSeconds", "samplesSeconds"),
    "correctness": ("name", "maximumRelativeError", "tolerance", "passed"),
    "direction": ("nativePlacement", "adapterPlacement", "destroysNativeInput", "adapterPreservesCallerInput",
                  "requiresPreservationCopyForRepeatedExecution", "preservationIncludedInPrimitiveTiming",
                  "preservationIncludedInAdapterTiming", "nativeInputRepresentationId",
                  "nativeOutputRepresentationId", "adapterInputRepresentationId", "adapterOutputRepresentationId",
                  "physicalExtents", "stridesElements", "paddingElements", "minimumAlignmentBytes", "aliasing",
                  "reusableWorkBytes", "outputCanFeedOppositeDirection"),
}


def check_record(rec: dict) -> list:
    """Problems found in a v1 record (empty when it conforms)."""
    bad = []

    def need(obj, key, where):
        if not isinstance(obj, dict):
            bad.append("%s is not an object" % where)
            return
        for k in _REQUIRED[key]:
            if k not in obj:
                bad.append("%s: missing %s" % (where or "record", k))

    need(rec, "", "")
    if rec.get("schema") != SCHEMA:
        bad.append("schema must be %s" % SCHEMA)
    if rec.get("status") not in ("passed", "failed"):
        bad.append("status must be passed or failed")
    nt = rec.get("numericType")
    if nt is not None and (nt.get("id") not in ("float64", "float32") or nt.get("scalarBits") not in (32, 64)):
        bad.append("numericType must be float64/64 or float32/32")
    for key in ("run", "provenance", "logicalOperators", "workload", "environment"):
        need(rec.get(key, {}), key, key)
    run = rec.get("run", {})
    if not (isinstance(run.get("samples"), int) and run["samples"] >= 1):
        bad.append("run.samples must be an integer >= 1")
    if rec.get("provenance", {}).get("waveVortexModelIssue") != 129:
        bad.append("provenance.waveVortexModelIssue must be 129")
    need(rec.get("workload", {}).get("antialias", {}), "workload.antialias", "workload.antialias")
    providers = rec.get("providers")
    if not isinstance(providers, list) or not providers:
        bad.append("providers must be a non-empty array")
        providers = []
    for i, p in enumerate(providers):
        where = "providers[%d]" % i
        need(p, "provider", where)
        for sub in ("providerBuild", "setup", "planning", "memory"):
            need(p.get(sub, {}), "provider." + sub, where + "." + sub)
        if not (isinstance(p.get("workers"), int) and p["workers"] >= 1):
            bad.append(where + ".workers must be an integer >= 1")
        for d in ("forward", "inverse"):
            ec = p.get("executionContract")
            if ec is not None:
                need(ec.get(d, {}), "direction", "%s.executionContract.%s" % (where, d))
        for j, e in enumerate(p.get("componentLedger", [])):
            need(e, "ledger", "%s.componentLedger[%d]" % (where, j))
            if e.get("state") not in STATES:
                bad.append("%s.componentLedger[%d].state invalid" % (where, j))
        for j, t in enumerate(p.get("timings", [])):
            need(t, "timing", "%s.timings[%d]" % (where, j))
            if t.get("state") not in STATES:
                bad.append("%s.timings[%d].state invalid" % (where, j))
            if any((not isinstance(s, (int, float))) or s < 0 for s in t.get("samplesSeconds", [])):
                bad.append("%s.timings[%d] has a negative or non-numeric sample" % (where, j))
        for j, c in enumerate(p.get("correctness", [])):
            need(c, "correctness", "%s.correctness[%d]" % (where, j))
            if not c.get("tolerance", 0) > 0:
                bad.append("%s.correctness[%d].tolerance must be positive" % (where, j))
    return bad


# ---- CSV and compare (src/json.cpp:282-332) ----------------------------------------------


def write_csv(rec: dict, path: str):
    with open(path, "w", newline="", encoding="utf-8") as handle:
        out = csv.writer(handle, lineterminator="\n")
        out.writerow(["run_id", "provider", "scope", "stage", "direction", "state", "bytes_moved", "sample_index",
                      "seconds"])
        for p in rec["providers"]:
            for t in p["timings"]:
                if not t["samplesSeconds"]:
                    out.writerow([rec["run"]["id"], p["id"], t["scope"], t["stage"], t["direction"], t["state"],
                                  t["bytesMoved"], "", ""])
                for i, s in enumerate(t["samplesSeconds"]):
                    out.writerow([rec["run"]["id"], p["id"], t["scope"], t["stage"], t["direction"], t["state"],
                                  t["bytesMoved"], i, repr(float(s))])


def compare_csv(path: str, stream=sys.stdout) -> int:
    groups = {}
    with open(path, newline="", encoding="utf-8") as handle:
        rows = list(csv.reader(handle))
    if not rows:
        raise ValueError("CSV result is empty.")
    for r in rows[1:]:
        if len(r) != 9 or not r[8]:
            continue
        groups.setdefault((r[1], r[2], r[3], r[4]), []).append(float(r[8]))
    stream.write("provider\tscope\tstage\tdirection\tmedian_ms\tsamples\n")
    for key in sorted(groups):
        v = groups[key]
        stream.write("%s\t%s\t%s\t%s\t%.6f\t%d\n" % (key + (1e3 * float(np.median(v)), len(v))))
    return 0 if groups else 1


def write_json(rec: dict, path: str):
    with open(path, "w", encoding="utf-8") as handle:
        json.dump(rec, handle, indent=1)
        handle.write("\n")


# ---- measuring the hip provider ----------------------------------------------------------


# ---- validate (src/benchmark.cpp:4898-5003) ------------------------------------------
# The command's own host references on the 8 x 8 validation grid, as the
# reference's validateBenchmark uses its host directR2C / gatherRetained /
# verticalForward (src/model.cpp:296-355, 663-716): explicit DFT sums, the
# mode permutation and the vertical products, in numpy.
VALIDATE_FIXTURES = ("impulse", "sinusoid", "random", "dc", "nyquist")


def _host_r2c(w, x) -> np.ndarray:
    """Direct DFT sums of every plane into WVM order (plane + P*(kx + nxh*ky))."""
    nxh, P = w.nx // 2 + 1, w.planes
    ex = np.exp(-2j * np.pi * np.outer(np.arange(nxh), np.arange(w.nx)) / w.nx)
    ey = np.exp(-2j * np.pi * np.outer(np.arange(w.ny), np.arange(w.ny)) / w.ny)
    planes = np.asarray(x, np.float64).reshape(P, w.ny, w.nx)
    spec = np.einsum("ky,pyx,qx->kqp", ey, planes, ex)  # [ky][kx][plane]
    return np.ascontiguousarray(spec.reshape(-1))


def _host_gather(w, modes, spec) -> np.ndarray:
    nxh, P = w.nx // 2 + 1, w.planes
    s = spec.reshape(w.ny, nxh, P)
    out = s[modes["stored_ky"].astype(np.int64), modes["stored_kx"].astype(np.int64)].copy()
    conj = modes["conj"] != 0
    out[conj] = np.conj(out[conj])
    return out.reshape(-1)


def validate(profile: str = "smoke", stream=sys.stdout) -> int:
    """validateBenchmark for the hip provider: full and pruned transforms of
    the five fixtures against direct DFT sums, the retained gather, the
    vertical projection of the retained modes and its modal round trip
    (FP64, 1e-12).  Prints the reference's message lines; returns 0 when
    every check passes, 1 otherwise."""
    import torch

    import skb_hip as S

    if profile not in PROFILES:
        raise ValueError("unknown profile %r" % profile)
    shape = PROFILES[profile][:4] if profile == "smoke" else (8, 8, 7, 2)
    w = S.Workload(*shape)
    device = torch.device("cuda:0")
    tol = 1e-12
    modes = S.retained_modes(w)
    nj = w.nj
    f_mat, i_mat = S.orthonormal_vertical(w.nz, nj)
    groups = np.zeros(1, S.GROUP_DTYPE)
    groups["first_mode"], groups["mode_count"] = 0, len(modes)
    full = S.FftPlan(w, S.SKB_F64)
    pruned = S.PrunedPlan(w, modes, S.SKB_F64)
    vplan = S.VerticalPlan(w, nj, groups, f_mat, i_mat, S.SKB_F64)
    F = f_mat.reshape(nj, w.nz)
    n = w.nx * w.ny
    cd = torch.complex128
    passed = True
    lines = []

    def dev(a, dtype):
        return torch.from_numpy(np.ascontiguousarray(a)).to(device=device, dtype=dtype)

    def host(t):
        torch.cuda.synchronize()
        return t.detach().cpu().numpy()

    for i, kind in enumerate(VALIDATE_FIXTURES):
        x = S.make_fixture(w, kind, 129 + i)
        oracle = _host_r2c(w, x)
        d_x = dev(x, torch.float64)
        spec = torch.empty(w.spectrum_elements, dtype=cd, device=device)
        full.forward(d_x, spec)
        err = max_relative_error(host(spec), oracle)
        passed &= err <= tol
        lines.append("hip/%s/forward=%f" % (kind, err))
        work = dev(oracle, cd)
        real = torch.empty(w.real_elements, dtype=torch.float64, device=device)
        full.inverse(work, real)
        err = max_relative_error(host(real) / n, x)
        passed &= err <= tol
        lines.append("hip/%s/inverse=%f" % (kind, err))
        retained_ref = _host_gather(w, modes, oracle)
        ret = torch.empty(len(modes) * w.planes, dtype=cd, device=device)
        pruned.forward(d_x, ret)
        err = max_relative_error(host(ret), retained_ref)
        passed &= err <= tol
        lines.append("operator/%s/retained=%f" % (kind, err))
        # pruned inverse against embed + full inverse on the device
        d_ref = dev(retained_ref, cd)
        back = torch.empty_like(real)
        pruned.inverse(d_ref, back)
        emb = torch.empty(w.spectrum_elements, dtype=cd, device=device)
        full.embed_retained(modes, d_ref, emb)
        back_full = torch.empty_like(real)
        full.inverse(emb, back_full)
        err = max_relative_error(host(back), host(back_full))
        passed &= err <= tol
        lines.append("hip-pruned/%s/inverse=%f" % (kind, err))
        modal = torch.empty(len(modes) * w.fields * nj, dtype=cd, device=device)
        vplan.forward(ret, modal)
        expect = (retained_ref.reshape(len(modes), w.fields, w.nz) @ F.T).reshape(-1)
        err = max_relative_error(host(modal), expect)
        passed &= err <= tol
        lines.append("operator/%s/horizontal-vertical=%f" % (kind, err))
        phys = torch.empty_like(ret)
        vplan.inverse(modal, phys)
        again = torch.empty_like(modal)
        vplan.forward(phys, again)
        err = max_relative_error(host(again), host(modal))
        passed &= err <= tol
        lines.append("operator/%s/vertical-modal-round-trip=%f" % (kind, err))
    lines.append("retained-mode-count=%d" % len(modes))
    lines.append("retained-mode-order-hash=%s" % S.mode_order_hash(modes))
    lines.append("validation=passed" if passed else "validation=failed")
    for line in lines:
        print(line, file=stream)
    return 0 if passed else 1


def _device_samples(torch, fn, warmups: int, samples: int):
    for _ in range(warmups):
        fn()
    torch.cuda.synchronize()
    out = []
    for _ in range(samples):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        b.synchronize()
        out.append(a.elapsed_time(b) * 1e-3)
    return out


def run(kernel: str, profile: str = "smoke", precision_id: str = "float64", seed: int = 129, warmups=None,
        samples=None) -> dict:
    import torch

    import skb_hip as S

    if profile not in PROFILES:
        raise ValueError("unknown profile %r" % profile)
    nx, ny, nz, fields, pw, ps_ = PROFILES[profile]
    warmups = pw if warmups is None else warmups
    samples = ps_ if samples is None else samples
    if samples < 1:
        raise ValueError("samples must be at least 1")
    precision = S.SKB_F32 if precision_id == "float32" else S.SKB_F64
    tol = 1e-6 if precision == S.SKB_F32 else 1e-12
    device = torch.device("cuda:0")
    rd, cd = S.real_dtype(precision), S.complex_dtype(precision)
    w = S.Workload(nx, ny, nz, fields)
    modes = S.retained_modes(w)
    groups = S.k2_groups(modes)
    fwd, inv = S.k2_vertical_family(nz, w.nj, groups)
    family_id = "k2-grouped-dct2-rotations"
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    x64 = 2.0 * torch.rand(w.real_elements, generator=g, device=device, dtype=torch.float64) - 1.0
    x = x64.to(rd)
    cb = torch.empty(0, dtype=cd).element_size()
    real_bytes, spec_bytes = w.real_elements * cb // 2, w.spectrum_elements * cb
    ret_bytes = len(modes) * w.planes * cb
    modal_bytes = len(modes) * w.fields * w.nj * cb

    def host(t):
        torch.cuda.synchronize()
        return t.detach().cpu().numpy().astype(np.complex128 if t.is_complex() else np.float64)

    if kernel == "fft":
        plan = S.FftPlan(w, precision)
        spec = torch.empty(w.spectrum_elements, dtype=cd, device=device)
        work = torch.empty_like(spec)
        out = torch.empty_like(x)
        plan.forward(x, spec)
        t_f = _device_samples(torch, lambda: plan.forward(x, spec), warmups, samples)
        t_i = _device_samples(torch, lambda: (work.copy_(spec), plan.inverse(work, out)), warmups, samples)
        work.copy_(spec)
        plan.inverse(work, out)
        corr = [metric("inverse(forward(x)) versus Nx*Ny*x", host(out), w.nx * w.ny * host(x), tol)]
        if precision == S.SKB_F32:
            p64, s64 = S.FftPlan(w, S.SKB_F64), torch.empty(w.spectrum_elements, dtype=torch.complex128,
                                                               device=device)
            p64.forward(x.to(torch.float64), s64)
            corr.append(metric("FP32 forward versus FP64 forward of the same input", host(spec), host(s64), tol))
        timings = [timing("primitive", "raw FFT", "forward", real_bytes + spec_bytes, t_f),
                   timing("primitive", "raw FFT", "inverse", spec_bytes + real_bytes, t_i)]
        ledger = [("raw forward FFT", "executed", "batched 2-D r2c straight into WVM order"),
                  ("raw inverse FFT", "executed", "batched 2-D c2r from WVM order; destroys its input")]
        fwd_dir = execution_direction("out-of-place", False, "real-grid", "wvm-frequency-major-interleaved",
                                      "real grid -> full WVM half-spectrum", "z=1,field=Nz,kx=P,ky=P*(Nx/2+1)", 0)
        inv_dir = execution_direction("out-of-place", True, "wvm-frequency-major-interleaved", "real-grid",
                                      "full WVM half-spectrum -> real grid", "x=1,y=Nx,z=Nx*Ny", 0)
        prov = provider_record(S, "hip", "two-pass-register-fft-v1", "wvm-frequency-major-interleaved", timings,
                               corr, ledger, spec_bytes, 0, fwd_dir, inv_dir)
    elif kernel == "pruned-horizontal":
        plan = S.PrunedPlan(w, modes, precision)
        full = S.FftPlan(w, precision)
        ret = torch.empty(len(modes) * w.planes, dtype=cd, device=device)
        out = torch.empty_like(x)
        t_f = _device_samples(torch, lambda: plan.forward(x, ret), warmups, samples)
        t_i = _device_samples(torch, lambda: plan.inverse(ret, out), warmups, samples)
        spec = torch.empty(w.spectrum_elements, dtype=cd, device=device)
        full.forward(x, spec)
        gathered = torch.empty_like(ret)
        full.gather_retained(modes, spec, gathered)
        full.embed_retained(modes, ret, spec)
        out_full = torch.empty_like(x)
        full.inverse(spec, out_full)
        corr = [metric("pruned forward versus full FFT + gather", host(ret), host(gathered), tol),
                metric("pruned inverse versus embed + full inverse FFT", host(out), host(out_full), tol)]
        timings = [timing("primitive", "pruned retained transform", "forward", real_bytes + ret_bytes, t_f),
                   timing("primitive", "pruned retained transform", "inverse", ret_bytes + real_bytes, t_i)]
        ledger = [("raw forward FFT", "fused", "rows, then only the active k_x columns"),
                  ("horizontal retention", "fused", "retained modes written by the column pass"),
                  ("horizontal embedding", "fused", "retained modes read by the inverse column pass")]
        fwd_dir = execution_direction("out-of-place", False, "real-grid", "retained-compact-interleaved",
                                      "real grid -> compact retained", "z=1,field=Nz,mode=P", plan.scratch_bytes())
        inv_dir = execution_direction("out-of-place", False, "retained-compact-interleaved", "real-grid",
                                      "compact retained -> real grid", "x=1,y=Nx,z=Nx*Ny", plan.scratch_bytes())
        prov = provider_record(S, "hip", "pruned-two-pass-fft-v1", "retained-compact-interleaved", timings, corr,
                               ledger, 0, plan.scratch_bytes(), fwd_dir, inv_dir)
    elif kernel == "vertical-gemm":
        plan = S.VerticalPlan(w, w.nj, groups, fwd, inv, precision)
        rng = np.random.default_rng(seed)
        n = len(modes) * w.planes
        phys_h = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        if precision == S.SKB_F32:
            phys_h = phys_h.astype(np.complex64).astype(np.complex128)
        phys = torch.from_numpy(phys_h).to(device=device, dtype=cd)
        modal = torch.empty(len(modes) * w.fields * w.nj, dtype=cd, device=device)
        back = torch.empty_like(phys)
        t_f = _device_samples(torch, lambda: plan.forward(phys, modal), warmups, samples)
        t_i = _device_samples(torch, lambda: plan.inverse(modal, back), warmups, samples)
        columns = len(modes) * w.fields
        probes = probe_columns(columns)
        first = groups["first_mode"].astype(np.int64)
        F = fwd.reshape(len(groups), w.nj, nz)
        I = inv.reshape(len(groups), nz, w.nj)
        m_h, b_h = host(modal).reshape(columns, w.nj), host(back).reshape(columns, nz)
        p_h = phys_h.reshape(columns, nz)
        grp = [int(np.searchsorted(first, c // w.fields, side="right") - 1) for c in probes]
        exp_f = np.stack([F[gi] @ p_h[c] for gi, c in zip(grp, probes)])
        exp_i = np.stack([I[gi] @ m_h[c] for gi, c in zip(grp, probes)])
        corr = [metric("forward vertical projection probes versus host sums", m_h[probes], exp_f, tol),
                metric("inverse vertical projection probes versus host sums", b_h[probes], exp_i, tol)]
        mb = int(plan.matrix_bytes())
        timings = [timing("primitive", "raw vertical MM", "forward", mb + ret_bytes + modal_bytes, t_f),
                   timing("primitive", "raw vertical MM", "inverse", mb + modal_bytes + ret_bytes, t_i)]
        ledger = [("raw forward vertical MM", "executed", "one MFMA launch over every K^2 group"),
                  ("raw inverse vertical MM", "executed", "one MFMA launch over every K^2 group")]
        fwd_dir = execution_direction("out-of-place", False, "retained-compact-interleaved", "modal-compact",
                                      "compact retained -> modal", "j=1,field=Nj,mode=Nj*F", 0)
        inv_dir = execution_direction("out-of-place", False, "modal-compact", "retained-compact-interleaved",
                                      "modal -> compact retained", "z=1,field=Nz,mode=P", 0)
        prov = provider_record(S, "hip", "k2-grouped-mfma-gemm-v1", "retained-compact-interleaved", timings, corr,
                               ledger, 2 * mb, 0, fwd_dir, inv_dir, gemm_calls=1)
    elif kernel == "spectral-pipeline":
        out = {}
        timings, corr = [], []
        for policy, name in ((S.SKB_PIPELINE_WVM_DIRECT, "wvm-direct"), (S.SKB_PIPELINE_COMPACT, "compact")):
            pipe = S.SpectralPipeline(w, modes, groups, fwd, inv, policy, precision)
            y = torch.empty_like(x)
            t = _device_samples(torch, lambda: pipe.execute(x, y), warmups, samples)
            timings.append(timing("uninstrumented-total", "synthetic antialiased spectral pipeline (%s)" % name,
                                  "round-trip", 2 * real_bytes + 2 * spec_bytes + 2 * modal_bytes, t))
            out[name] = host(y)
        corr.append(metric("compact policy versus wvm-direct policy", out["compact"], out["wvm-direct"], tol))
        if precision == S.SKB_F32:
            p64 = S.SpectralPipeline(w, modes, groups, fwd, inv, S.SKB_PIPELINE_COMPACT, S.SKB_F64)
            y64 = torch.empty(w.real_elements, dtype=torch.float64, device=device)
            p64.execute(x.to(torch.float64), y64)
            corr.append(metric("FP32 pipeline versus FP64 pipeline of the same input", out["compact"], host(y64),
                               tol))
        ledger = [("raw forward FFT", "executed", "pruned (compact) or full WVM (wvm-direct)"),
                  ("raw forward vertical MM", "executed", "MFMA over the K^2 groups"),
                  ("modal work", "fused", "weights applied inside the one-kernel modal round trip (compact)"),
                  ("raw inverse vertical MM", "executed", "MFMA over the K^2 groups"),
                  ("raw inverse FFT", "executed", "pruned (compact) or full WVM (wvm-direct)")]
        fwd_dir = execution_direction("out-of-place", False, "real-grid", "real-grid",
                                      "real grid -> real grid", "x=1,y=Nx,z=Nx*Ny", 0)
        prov = provider_record(S, "hip", "spectral-pipeline-v1", "real-grid", timings, corr, ledger, 0, 0, fwd_dir,
                               fwd_dir, gemm_calls=2)
    elif kernel == "production-lifetime-flux":
        # src/benchmark.cpp:3957-4234, 4440-4580: 15 modal inputs [U, V, W,
        # q_t x, q_t y, q_t z for four targets] -> four modal targets
        targets = 4
        wt = S.Workload(nx, ny, nz, targets)
        plan = S.FluxPlan(wt, modes, w.nj, groups, fwd, inv, precision)
        nin = plan.input_elements
        gin = torch.Generator(device=device)
        gin.manual_seed(seed)
        xin = torch.complex(torch.randn(nin, generator=gin, device=device, dtype=torch.float64),
                            torch.randn(nin, generator=gin, device=device, dtype=torch.float64))
        dc = int(np.nonzero((modes["k"] == 0) & (modes["l"] == 0))[0][0])
        xin.view(len(modes), 3 + 3 * targets, w.nj)[dc] = xin.view(len(modes), 3 + 3 * targets, w.nj)[dc].real.to(
            torch.complex128)
        d_in = xin.to(cd)
        keep = d_in.clone()
        out_f = torch.empty(plan.output_elements, dtype=cd, device=device)
        out_s = torch.empty_like(out_f)
        plan.set_fused(True)
        t_f = _device_samples(torch, lambda: plan.execute(d_in, out_f), warmups, samples)
        plan.set_fused(False)
        t_s = _device_samples(torch, lambda: plan.execute(d_in, out_s), warmups, samples)
        plan.set_fused(True)
        o_f, o_s = host(out_f), host(out_s)
        corr = [metric("fused physical-space pass versus separate-kernel composition", o_f, o_s, tol),
                metric("caller modal input preserved", host(d_in), host(keep), tol)]
        per = o_f.reshape(len(modes), targets, w.nj)
        corr.append({"name": "DC boundary modal outputs are real", "maximumRelativeError":
                     float(np.abs(per[dc].imag).max() / max(np.abs(per).max(), 1e-300)),
                     "relativeL2Error": 0.0, "tolerance": tol,
                     "passed": bool(np.abs(per[dc].imag).max() <= tol * max(np.abs(per).max(), 1e-300))})
        if precision == S.SKB_F32:
            p64 = S.FluxPlan(wt, modes, w.nj, groups, fwd, inv, S.SKB_F64)
            o64 = torch.empty(plan.output_elements, dtype=torch.complex128, device=device)
            p64.execute(xin, o64)
            corr.append(metric("FP32 flux versus FP64 flux of the same input", o_f, host(o64), tol))
        p64_ = per
        for t in range(targets):
            ref = o_s.reshape(len(modes), targets, w.nj)[:, t]
            corr.append(metric("target %d fused versus separate kernels" % t, p64_[:, t], ref, tol))
        in_bytes = nin * cb
        out_bytes = plan.output_elements * cb
        vol = wt.nx * wt.ny * wt.nz
        timings = [timing("uninstrumented-total", "production-lifetime four-target spectral-flux composition",
                          "forward", in_bytes + out_bytes, t_f),
                   timing("uninstrumented-total",
                          "production-lifetime flux, separate kernels (real volumes in memory)", "forward",
                          in_bytes + out_bytes + (15 + 4) * vol * cb, t_s)]
        ledger = [("raw inverse vertical MM (15 modal inputs)", "executed", "one MFMA launch"),
                  ("horizontal reconstruction", "fused", "inverse columns, then rows fused with the products"),
                  ("pointwise advection", "fused", "products formed in registers between the row transforms"),
                  ("horizontal forward and retention", "executed", "forward columns to compact retained"),
                  ("raw forward vertical MM (4 modal targets)", "executed", "one MFMA launch")]
        fwd_dir = execution_direction("out-of-place", False, "modal-compact", "modal-compact",
                                      "15 modal fields -> 4 modal targets", "j=1,field=Nj,mode=Nj*F", 0)
        prov = provider_record(S, "hip", "production-lifetime-flux-v1", "modal-compact", timings, corr, ledger, 0, 0,
                               fwd_dir, fwd_dir, gemm_calls=2)
    else:
        raise ValueError("kernel must be fft, pruned-horizontal, vertical-gemm, spectral-pipeline or "
                         "production-lifetime-flux")
    return assemble(S, kernel, profile, precision_id, seed, warmups, samples, w, modes, groups, family_id, [prov],
                    torch.cuda.get_device_name(0))


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="command", required=True)
    r = sub.add_parser("run")
    r.add_argument("--kernel", required=True)
    r.add_argument("--profile", default="smoke")
    r.add_argument("--precision", default="float64", choices=["float64", "float32"])
    r.add_argument("--seed", type=int, default=129)
    r.add_argument("--warmups", type=int, default=None)
    r.add_argument("--samples", type=int, default=None)
    r.add_argument("--json")
    r.add_argument("--csv")
    c = sub.add_parser("compare")
    c.add_argument("--input", required=True)
    k = sub.add_parser("check")
    k.add_argument("path")
    v = sub.add_parser("validate")
    v.add_argument("--profile", default="smoke")
    a = ap.parse_args(argv)
    if a.command == "validate":
        return validate(a.profile)
    if a.command == "compare":
        return compare_csv(a.input)
    if a.command == "check":
        with open(a.path, encoding="utf-8") as handle:
            problems = check_record(json.load(handle))
        for p in problems:
            print(p)
        return 1 if problems else 0
    rec = run(a.kernel, a.profile, a.precision, a.seed, a.warmups, a.samples)
    if a.json:
        write_json(rec, a.json)
    if a.csv:
        write_csv(rec, a.csv)
    print(json.dumps({"status": rec["status"], "run": rec["run"]["id"],
                      "correctness": [(c["name"], c["maximumRelativeError"]) for c in rec["providers"][0]["correctness"]]}))
    # exit codes of src/main.cpp: 0 passed, 1 correctness failed
    return 0 if rec["status"] == "passed" else 1


if __name__ == "__main__":
    sys.exit(main())
