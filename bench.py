#!/usr/bin/env python3
"""Retained-operator throughput of the hip provider on one MI355X.

One step = one batch of --batch (128) independent retained-operator
evaluations of the 256^2 x Nz65 x 3 workload, distinct inputs and outputs
resident in HBM, enqueued by ONE C-ABI call
(skb_retained_operator_execute_batch), each evaluation:
  1. pruned inverse c2r   compact retained (z, field, mode) -> real planes
  2. pruned forward r2c   real planes -> compact retained (2/3-rule truncated)
  3. K^2-grouped vertical GEMM on MFMA -> modal
The batch runs stage-batched (skb_retained_operator_reserve_batch): each
stage is one launch over all evaluations of the step (layout copies, plane
transforms, vertical products), so the 195 plane workgroups of successive
evaluations fill every CU and the vertical product runs in steady state
(--batch-mode lanes: evaluations alternate over streams on separate hardware
queues instead).  `value` = evaluations per second over the timed steps; the single-call rate (one
skb_retained_operator_execute per evaluation, in order) and per-evaluation
latencies are reported beside it.
The headline runs in FP32 storage (FP32 transforms, split-BF16 MFMA vertical
products); the plain FP64 GPU form is timed beside it and supplies the
reference for the FP32 error.  The CPU column times the FP64 oracle
(pocketfft + BLAS) on this machine's host cores on a bounded sample.

Prints ONE JSON line.  `--gpus N` with N > 1 exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "spectral-kernel-benchmarks_amd", "python"))

METRIC = "retained-operator evals/s @256²×Nz65×3, FP32 vs FP64 at 1e-6 err; %HBM; p99 ms"
HBM_PEAK_GBPS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_retained_operator_256_nz65_f3.json")


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40,
                    help="timed steps; one step is one batch of --batch evaluations")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--batch", type=int, default=128,
                    help="independent evaluations per step (one execute_batch call)")
    ap.add_argument("--batch-mode", choices=("stage", "lanes"), default="stage",
                    help="stage: one launch per stage over the batch; lanes: evaluations over --lanes streams")
    ap.add_argument("--lanes", type=int, default=3, help="lanes of --batch-mode lanes (1: in order)")
    ap.add_argument("--family", choices=["qr32", "k2"], default="qr32",
                    help="qr32: 32 K^2 bins of 65x65 QR-Gaussian matrices; k2: reference 2045x(42x65) family")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="budget of the CPU baseline sample")
    ap.add_argument("--quiet", action="store_true", help="no progress lines on stderr")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the secondary measurements (advection flux 512^2 x Nz129, batched uneven workload, "
                         "serving loop)")
    ap.add_argument("--serving-requests", type=int, default=1500)
    ap.add_argument("--serving-lanes", choices=("size", "class", "executor"), default="size",
                    help="serving streams: one lane for the large classes and one per small class, one per "
                         "class, or one per executor")
    ap.add_argument("--serving-reserve-cus", type=int, default=32,
                    help="CUs kept free of large-class kernels in the serving loop (0: no cap)")
    ap.add_argument("--profile-run", action="store_true",
                    help="short run for rocprofv3: warmup, then every stage in blocks of 50 launches rotating "
                         "over 8 plans (max(200, --steps) launches each, as the roofline timing), then one "
                         "256 MiB device copy (the calibration kernel for FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="turn two rocprofv3 --pmc counter_collection.csv files (FETCH_SIZE run, WRITE_SIZE run) "
                         "into " + os.path.relpath(TRAFFIC_FILE, ROOT) + " and exit")
    return ap.parse_args()


COPY_BYTES = 256 << 20
BATCH_CAP = 128  # evaluations per stage-batched launch (skb_retained_operator_reserve_batch's maximum)
KERNEL_NAMES = {  # substring of the device kernel name -> stage of the (stage-batched) evaluation
    "pm_to_pm_batch_kernel": "layout",
    "plane_operator_kernel<86>": "transforms",
    "gemm_pm_range_kernel<2, 65>": "vertical_gemm",
}
BATCH_STAGES = (("layout", 0), ("transforms", 4), ("vertical_gemm", 3))  # SKB_OP_STAGE_* of each


def summarize_pmc(fetch_csv, write_csv, family):
    """Per-kernel HBM bytes from rocprofv3 counter CSVs, scaled by what the same
    counters read for a plain device copy of COPY_BYTES (read COPY_BYTES,
    write COPY_BYTES)."""
    import csv

    def per_kernel(path, counter):
        """(median KiB per dispatch by kernel name, largest single copy dispatch
        in KiB).  The median keeps the batched launches' figure when a few
        single-evaluation dispatches of the same kernel share the run."""
        values, copy_max = {}, 0.0
        with open(path, newline="") as handle:
            for row in csv.DictReader(handle):
                if row.get("Counter_Name") != counter:
                    continue
                name = row.get("Kernel_Name", "")
                value = float(row["Counter_Value"])
                values.setdefault(name, []).append(value)
                if "copybuffer" in name.lower():
                    copy_max = max(copy_max, value)
        return {k: percentile(v, 0.5) for k, v in values.items()}, copy_max

    # The calibration dispatch is the COPY_BYTES device-to-device copy, by far
    # the largest copy of the run; the other copies are small host transfers.
    fetch, copy_f = per_kernel(fetch_csv, "FETCH_SIZE")
    write, copy_w = per_kernel(write_csv, "WRITE_SIZE")
    if copy_f <= 0 or copy_w <= 0:
        raise SystemExit("no copy kernel found in the counter files")
    f_scale = (COPY_BYTES / 1024.0) / copy_f
    w_scale = (COPY_BYTES / 1024.0) / copy_w
    out = {"family": family, "copy_bytes": COPY_BYTES, "fetch_scale": f_scale, "write_scale": w_scale,
           "kernels": {}}
    for name, stage in KERNEL_NAMES.items():
        fk = [v for k, v in fetch.items() if name in k]
        wk = [v for k, v in write.items() if name in k]
        if fk and wk:
            fb = fk[0] * 1024.0 * f_scale
            wb = wk[0] * 1024.0 * w_scale
            out["kernels"][stage] = {"fetch_bytes": fb, "write_bytes": wb, "hbm_bytes": fb + wb,
                                     "fetch_kib_raw": fk[0], "write_kib_raw": wk[0]}
    with open(TRAFFIC_FILE, "w") as handle:
        json.dump(out, handle, indent=1)
        handle.write("\n")
    print(json.dumps(out))


def log(args, msg):
    if not args.quiet:
        print("[bench] " + msg, file=sys.stderr, flush=True)


class Operator:
    """One retained-operator plan (skb_retained_operator: every kernel of an
    evaluation enqueued by one C-ABI call) with its input and output."""

    def __init__(self, S, torch, w, modes, groups, nj, fwd, inv, x64, precision, device):
        self.S, self.torch, self.w, self.nj = S, torch, w, nj
        self.op = S.OperatorPlan(w, modes, nj, groups, fwd, inv, precision)
        cd = S.complex_dtype(precision)
        self.x = torch.from_numpy(x64).to(device=device, dtype=cd)
        self.modal = torch.empty(self.op.modal_elements, dtype=cd, device=device)
        self.call = self.op.bind_execute(self.x, self.modal)
        self.plane = self.op.path() == S.SKB_PATH_PLANE
        # the kernels of one evaluation as separately launchable stages
        names = ((("layout", S.SKB_OP_STAGE_LAYOUT), ("transforms", S.SKB_OP_STAGE_TRANSFORMS),
                  ("vertical_gemm", S.SKB_OP_STAGE_VERTICAL)) if self.plane else
                 (("transforms", S.SKB_OP_STAGE_TRANSFORMS), ("vertical_gemm", S.SKB_OP_STAGE_VERTICAL)))
        self.stages = [(n, self.op.bind_stage(st, self.x, self.modal)) for n, st in names]
        # the two transforms as separate kernels (the unfused form), for reference
        self.unfused = [(n, self.op.bind_stage(st, self.x, self.modal))
                        for n, st in (("inverse", S.SKB_OP_STAGE_INVERSE), ("forward", S.SKB_OP_STAGE_FORWARD))]

    def step(self):
        self.call()


def timed(torch, fn, steps):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


LATENCY_EVALS = 1000
SINGLE_EVALS = 400


def per_eval_latency(torch, op, evals):
    """Device time of each evaluation: one HIP event pair per evaluation on
    the stream the plan launches on, over `evals` back-to-back evaluations."""
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(evals)]
    torch.cuda.synchronize()
    for a, b in ev:
        a.record()
        op.step()
        b.record()
    torch.cuda.synchronize()
    return [a.elapsed_time(b) for a, b in ev]


def per_stage_times(torch, ops, launches, block=50):
    """Average launch duration of every stage, HIP events around blocks of
    `block` back-to-back launches on the plans' stream.  The launches rotate
    over the plans in `ops` (each with its own buffers), so a stage's working
    set exceeds the 256 MB Infinity Cache and the rates are HBM rates.  Every
    stage rewrites its outputs from its inputs, so repeating one alone does
    the work it does inside an evaluation."""
    out = {}
    for idx, (name, _) in enumerate(ops[0].stages + ops[0].unfused):
        fns = [(op.stages + op.unfused)[idx][1] for op in ops]
        blocks = []
        for b in range(max(1, launches // block)):
            a, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            for i in range(block):
                fns[(b * block + i) % len(fns)]()
            e.record()
            blocks.append((a, e))
        torch.cuda.synchronize()
        out[name] = [a.elapsed_time(e) / block for a, e in blocks]
    return out


def batched_stage_times(torch, op, xs, outs, calls, block=10):
    """Average duration of one call of each stage of the stage-batched
    execute over the batch (xs -> outs: one launch per group of evaluations),
    HIP events around blocks of `block` back-to-back calls on the plan's
    stream.  The inputs and outputs are the headline's (distinct arrays), so
    the working set is HBM-resident."""
    out = {}
    for name, stage in BATCH_STAGES:
        fn = op.bind_execute_batch_stage(stage, xs, outs)
        fn()
        blocks = []
        for _ in range(max(1, calls // block)):
            a, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            for _ in range(block):
                fn()
            e.record()
            blocks.append((a, e))
        torch.cuda.synchronize()
        out[name] = [a.elapsed_time(e) / block for a, e in blocks]
    return out


def percentile(values, q):
    v = sorted(values)
    if not v:
        return None
    k = (len(v) - 1) * q
    lo, hi = int(k), min(int(k) + 1, len(v) - 1)
    return v[lo] + (v[hi] - v[lo]) * (k - lo)


# A GPU process on the MI355X box owns a 16-core share of the host.
CPU_SHARE = 16


def host_cpu():
    """lscpu model and topology, and the CPUs this process may run on."""
    import subprocess

    info = {}
    try:
        text = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in text.splitlines():
            key, _, val = line.partition(":")
            info[key.strip()] = val.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return {"model": info.get("Model name"), "sockets": info.get("Socket(s)"),
            "cores_per_socket": info.get("Core(s) per socket"), "threads_per_core": info.get("Thread(s) per core"),
            "logical_cpus": info.get("CPU(s)"), "affinity_cpus": affinity}


def cpu_baseline(args, w, x64, nj, groups, fwd):
    """FP64 oracle timed on the host cores; also returns its output, used only
    as the checker of the GPU results."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # checker / baseline only

    cpu = host_cpu()
    workers = max(1, min(CPU_SHARE, cpu["affinity_cpus"]))
    ow = O.Workload(w.nx, w.ny, w.nz, w.fields)
    om = O.retained_modes(ow)
    first = groups["first_mode"].copy()
    count = groups["mode_count"].copy()
    out = O.retained_operator_fast(ow, om, x64, nj, first, count, fwd, workers)  # warm (plans, pages)
    n, t0 = 0, time.perf_counter()
    while True:
        O.retained_operator_fast(ow, om, x64, nj, first, count, fwd, workers)
        n += 1
        el = time.perf_counter() - t0
        if el >= args.cpu_seconds or n >= 200:
            break
    rec = {
        "value": n / el,
        "unit": "evals/s",
        "cores": workers,
        "kind": "port",
        "sample": "%d full 256x256xNz65x3 FP64 retained-operator evals in %.1f s: scipy pocketfft "
                  "rfft2/irfft2 (workers=%d, not FFTW), C gather/embed, numpy BLAS vertical"
                  % (n, el, workers),
        "ms_per_eval": 1e3 * el / n,
        "host": cpu,
        "threads_note": "workers = the %d-core host share of one GPU on this machine (of %s affinity CPUs)"
                        % (CPU_SHARE, cpu["affinity_cpus"]),
    }
    return rec, out, O


def _timed_events(torch, fn, reps, stream=None):
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s = stream or torch.cuda.current_stream()
    a.record(s)
    for _ in range(reps):
        fn()
    b.record(s)
    b.synchronize()
    return a.elapsed_time(b) / reps  # ms


def device_spectrum(torch, S, w, modes, seed, device, dtype):
    """Seeded complex Gaussian x (1 + k^2 + l^2)^(-5/6) per mode (DC real),
    generated on the device."""
    import numpy as np

    g = torch.Generator(device=device)
    g.manual_seed(seed)
    n = len(modes)
    amp = torch.from_numpy((1.0 + modes["k"].astype(np.float64) ** 2 + modes["l"].astype(np.float64) ** 2)
                           ** (-5.0 / 6.0)).to(device)
    re = torch.randn((n, w.planes), generator=g, device=device, dtype=torch.float64)
    im = torch.randn((n, w.planes), generator=g, device=device, dtype=torch.float64)
    x = torch.complex(re, im) * amp.view(n, 1)
    dc = int(np.nonzero((modes["k"] == 0) & (modes["l"] == 0))[0][0])
    x[dc] = x[dc].real.to(torch.complex128)
    return x.reshape(-1).to(dtype)


def bench_flux(args, S, torch, device):
    """Advection flux of one tracer at 512^2 x Nz129 (u, v, w, q): FP32 fused
    physical stage, beside the FP64 form; error of FP32 vs FP64 on the GPU."""
    w = S.Workload(512, 512, 129, 4)
    modes = S.retained_modes(w)
    dz = S.vertical_derivative_matrix(w.nz, 1.0)
    x64 = device_spectrum(torch, S, w, modes, 129, device, torch.complex128)
    out = {}
    res = {}
    for prec, name in ((S.SKB_F32, "fp32"), (S.SKB_F64, "fp64")):
        plan = S.AdvectionFluxPlan(w, modes, dz, prec)
        cd = S.complex_dtype(prec)
        x = x64.to(cd)
        f = torch.empty(len(modes) * w.nz, dtype=cd, device=device)
        plan.execute(x, f)
        ms = _timed_events(torch, lambda: plan.execute(x, f), 20)
        res[name] = f.to(torch.complex128)
        out[name] = {"evals_per_s": 1e3 / ms, "ms_per_eval": ms}
        del plan
    d = (res["fp32"] - res["fp64"]).abs().max().item()
    out["fp32_vs_fp64_max_rel"] = d / res["fp64"].abs().max().item()
    out["workload"] = "advection flux 512x512 Nz129, fields u,v,w,q; spectral d/dx, d/dy; 129x129 d/dz on MFMA"
    return out


def bench_pipeline(args, S, torch, device):
    """Synthetic spectral pipeline (r2c -> retained -> vertical forward ->
    modal weights -> vertical inverse -> c2r) at 256^2 x Nz65 x 3 and
    512^2 x Nz129 x 4, per boundary policy and precision.  The reference's
    published M4 Max figures for the same round trips are quoted beside."""
    import numpy as np

    reference_ms = {(256, 65): {"wvm_direct": 13.258, "fused_split": 9.091},
                    (512, 129): {"wvm_direct": 134.531, "fused_split": 131.357}}
    out = {}
    for nx, nz, f in ((256, 65, 3), (512, 129, 4)):
        w = S.Workload(nx, nx, nz, f)
        modes = S.retained_modes(w)
        groups = S.k2_groups(modes)
        fwd, inv = S.k2_vertical_family(w.nz, w.nj, groups)
        g = torch.Generator(device=device)
        g.manual_seed(129)
        x64 = 2.0 * torch.rand(w.real_elements, generator=g, device=device, dtype=torch.float64) - 1.0
        res, outputs = {}, {}
        for policy, pname in ((S.SKB_PIPELINE_COMPACT, "compact"), (S.SKB_PIPELINE_WVM_DIRECT, "wvm_direct")):
            for prec, name in ((S.SKB_F32, "fp32"), (S.SKB_F64, "fp64")):
                pipe = S.SpectralPipeline(w, modes, groups, fwd, inv, policy, prec)
                rd = S.real_dtype(prec)
                x = x64.to(rd)
                y = torch.empty_like(x)
                # compact: the default, then the other form of the modal round
                # trip (one kernel or three launches) for comparison
                variants = [("", None)]
                if policy == S.SKB_PIPELINE_COMPACT:
                    variants.append(("_unfused", False) if pipe.fused else ("_fused", True))
                for suffix, fused in variants:
                    if fused is not None:
                        pipe.set_fused(fused)
                    for _ in range(10):  # warm: the first variant timed otherwise reads ~5% slow
                        pipe.execute(x, y)
                    ms = _timed_events(torch, lambda: pipe.execute(x, y), 10)
                    res["%s_%s%s" % (pname, name, suffix)] = {"ms_per_eval": ms, "evals_per_s": 1e3 / ms}
                outputs[(pname, name)] = y.to(torch.float64)
                del pipe
        ref = outputs[("compact", "fp64")]
        scale = ref.abs().max().item()
        res["fp32_vs_fp64_max_rel"] = max((outputs[(p, "fp32")] - ref).abs().max().item() / scale
                                          for p in ("compact", "wvm_direct"))
        res["wvm_vs_compact_fp64_max_rel"] = (outputs[("wvm_direct", "fp64")] - ref).abs().max().item() / scale
        res["reference_m4max_ms"] = reference_ms[(nx, nz)]
        out["%dx%d_nz%d_f%d" % (nx, nx, nz, f)] = res
        del x64, outputs, ref
        torch.cuda.empty_cache()
    out["workload"] = "synthetic spectral pipeline, K^2 family, synthetic modal weights; uniform[-1,1) real input"
    return out


def bench_type1(args, S, torch, device):
    """Constant-stratification vertical schedule (15 inverse + 4 forward
    complex channel transforms of type I) at 256^2 x Nz65, on full-half rows
    and on retained rows, FP32 and FP64."""
    import numpy as np

    w = S.Workload(256, 256, 65, 4)
    modes = S.retained_modes(w)
    out = {}
    for name, rows, m in (("full_half", w.nx_half * w.ny, None), ("retained", len(modes), modes)):
        a3 = S.type1_fill(rows, 3, w.nz, m, w.nx_half)
        a1 = S.type1_fill(rows, 1, w.nz, m, w.nx_half)
        res = {"rows": rows}
        for prec, pname in ((S.SKB_F32, "fp32"), (S.SKB_F64, "fp64")):
            cd = S.complex_dtype(prec)
            d3 = torch.from_numpy(a3.reshape(-1)).to(device=device, dtype=cd)
            d1 = torch.from_numpy(a1.reshape(-1)).to(device=device, dtype=cd)
            plan = S.Type1Plan(rows, w.nz, prec)
            plan.schedule(d3, d1)
            ms = _timed_events(torch, lambda: plan.schedule(d3, d1), 20)
            one = rows * w.nz * (8 if prec == S.SKB_F32 else 16)
            res[pname] = {"ms_per_schedule": ms, "schedules_per_s": 1e3 / ms,
                          "GBps_38_channel_passes": 38 * one / (ms * 1e-3) / 1e9}
            del plan, d3, d1
        out[name] = res
    out["workload"] = "constant-stratification vertical schedule 256x256 Nz65: 19 type-I channel transforms on MFMA"
    return out


def bench_batched(args, S, torch, device):
    """The uneven batch: 40 x 128^2xNz33x3, 40 x 256^2xNz65x3, 16 x 512^2xNz129x4
    retained-operator instances in one hipGraph, launch shapes autotuned."""
    import numpy as np

    mix = [((128, 128, 33, 3), 40), ((256, 256, 65, 3), 40), ((512, 512, 129, 4), 16)]
    # a class stacks at most 2^31 real elements (the kernels' per-array
    # limit): larger groups of instances become several classes
    chunks = []
    for shape, inst in mix:
        per = max(1, (1 << 31) // (shape[0] * shape[1] * shape[2] * shape[3]))
        while inst > 0:
            chunks.append((shape, min(per, inst)))
            inst -= min(per, inst)
    out = {"classes": [f"{n}x {s[0]}^2xNz{s[2]}x{s[3]}" for s, n in chunks]}
    for prec, name in ((S.SKB_F32, "fp32"), (S.SKB_F64, "fp64")):
        batch = S.Batch(prec)
        cd = S.complex_dtype(prec)
        keep = []
        for shape, inst in chunks:
            w = S.Workload(*shape)
            modes = S.retained_modes(w)
            groups = S.k2_binned_groups(modes, 32)
            fwd, inv = S.qr_gaussian_family(w.nz, len(groups), 129)
            cls = batch.add_class(w, inst, modes, w.nz, groups, fwd, inv)
            ws = S.Workload(shape[0], shape[1], shape[2], shape[3] * inst)
            x = device_spectrum(torch, S, ws, modes, 7 + cls, device, cd)
            y = torch.empty(batch.output_elements(cls), dtype=cd, device=device)
            batch.bind(cls, x, y)
            keep.append((x, y))
        side = torch.cuda.Stream()
        tune = batch.autotune(5, side)
        batch.capture(side)
        batch.launch(side)
        ms = _timed_events(torch, lambda: batch.launch(side), 10, stream=side)
        out[name] = {"batches_per_s": 1e3 / ms, "ms_per_batch": ms, "instance_evals_per_s": 96e3 / ms,
                     "autotuned_grids": [[st["grid"] for st in c["stages"]] for c in tune["classes"]],
                     "autotuned_products": [c["stages"][4].get("products") for c in tune["classes"]],
                     "tuning_spec": tune["spec"]}
        del batch, keep
        torch.cuda.synchronize()
    out["graph"] = "one hipGraph replay per batch (5 kernel launches per class, %d classes)" % len(chunks)
    return out


def bench_serving(args, S, torch, device):
    """Serving mix: 64 clients (50% 256^2xNz65x3, 30% 512^2xNz129x4,
    15% 128^2xNz33x3, 5% 1024^2xNz257x4), Poisson arrivals at 70% of the
    measured single-request capacity, FP32 advection-flux steps; sampled
    requests checked against the FP64 path (sizes up to Nz129)."""
    import numpy as np

    import serving as V

    classes = [V.make_class(shape, f, device) for shape, f, _ in V.DEFAULT_MIX]
    labels = V.assign_clients(V.DEFAULT_MIX, 64, 5)
    clients = [V.Client(i, classes[labels[i]], V.initial_state(classes[labels[i]], 500 + i, device))
               for i in range(64)]
    budget = 16 << 30  # binds: the 1024^2 class gets fewer executors than max_slots
    server = V.Server(classes, budget_bytes=budget, dt=1e-3, device=device, max_slots=4,
                      lanes=getattr(args, "serving_lanes", "size"),
                      reserve_cus=getattr(args, "serving_reserve_cus", 32))
    # capacity: mix-weighted single-request service time, measured without
    # the CU reserve (the arrival rate does not depend on it)
    server.set_reserve(False)
    svc = []
    for c, (_, _, share) in zip(classes, V.DEFAULT_MIX):
        cl = next(x for x in clients if x.cls is c)
        server.step_alone(cl)
        t0 = time.perf_counter()
        for _ in range(3):
            server.step_alone(cl)
        svc.append(share * (time.perf_counter() - t0) / 3)
        cl.steps = 0
    server.set_reserve(True)
    rate = 0.7 / sum(svc)
    arrivals = V.poisson_arrivals(rate, args.serving_requests, 64, 9)
    samples = []
    done = server.run(clients, arrivals, sample_every=50,
                      on_sample=lambda req, cl, x, f: samples.append((cl.cls, x, f)) if cl.cls.w.nz <= 132 else None)
    summ = V.latency_summary(done, clients)
    worst = 0.0
    plans64 = {}
    for cls, x, f in samples:
        if cls.key not in plans64:
            plans64[cls.key] = S.AdvectionFluxPlan(cls.w, cls.modes, cls.dz, S.SKB_F64)
        ref = torch.empty(len(cls.modes) * cls.w.nz, dtype=torch.complex128, device=device)
        plans64[cls.key].execute(x.to(torch.complex128), ref)
        worst = max(worst, (f.to(torch.complex128) - ref).abs().max().item() / ref.abs().max().item())
    summ.update({"arrival_rate_per_s": rate, "clients": 64, "slots": server.slots(),
                 "max_in_flight": server.max_inflight, "sampled_checks": len(samples),
                 "sampled_fp32_vs_fp64_max_rel": worst, "hbm_budget_bytes": budget,
                 "executor_bytes": server.used, "admission": server.admission()})
    for k, v in summ.get("per_class", {}).items():
        v["p99_over_p50"] = v["p99_ms"] / v["p50_ms"] if v["p50_ms"] > 0 else None
    return summ


def main():
    args = parse()
    if args.summarize_pmc:
        summarize_pmc(args.summarize_pmc[0], args.summarize_pmc[1], args.family)
        return 0
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (requested --gpus %d)" % args.gpus,
              file=sys.stderr)
        return 2
    import numpy as np
    import torch

    import skb_hip as S

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 2
    device = torch.device("cuda:0")
    # Everything runs on one created stream (default priority): launches on
    # the legacy null stream carry extra synchronisation per kernel.
    with torch.cuda.stream(torch.cuda.Stream(device=device)):
        return run(args, np, torch, S, device)


def run(args, np, torch, S, device):
    w = S.Workload(256, 256, 65, 3)
    log(args, "setup %s family=%s" % (w, args.family))
    modes = S.retained_modes(w)
    if args.family == "qr32":
        groups = S.k2_binned_groups(modes, 32)
        nj = w.nz
        fwd, inv = S.qr_gaussian_family(w.nz, len(groups), 129)
        family = "32 K^2 bins x 65x65 QR-Gaussian (seed 129)"
    else:
        groups = S.k2_groups(modes)
        nj = w.nj
        fwd, inv = S.k2_vertical_family(w.nz, nj, groups)
        family = "%d exact-K^2 groups x %dx%d pair-rotated DCT-II" % (len(groups), nj, w.nz)
    x64 = S.spectral_fixture(w, modes, 129, -1.0)

    ops = {p: Operator(S, torch, w, modes, groups, nj, fwd, inv, x64, p, device) for p in (S.SKB_F32, S.SKB_F64)}
    f32, f64 = ops[S.SKB_F32], ops[S.SKB_F64]
    log(args, "fp32 path: %s" % ("plane-resident transforms" if f32.plane else "two-pass transforms"))

    # FP32 versus the FP64 GPU form on identical inputs
    f32.step()
    f64.step()
    torch.cuda.synchronize()
    a = f32.modal.cpu().numpy().astype(np.complex128)
    b = f64.modal.cpu().numpy()
    err_max = float(np.abs(a - b).max() / max(1.0, np.abs(b).max()))
    err_l2 = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    log(args, "fp32 vs fp64 max-scale err %.3e, rel L2 %.3e" % (err_max, err_l2))

    if args.profile_run:
        # the launches the roofline is timed on: every stage of the
        # stage-batched execute over the headline's batch, in blocks
        B = args.batch
        f32.op.reserve_batch(min(B, BATCH_CAP))
        xs = [f32.x * (1.0 + 1e-3 * k) for k in range(B)]
        outs = [torch.empty_like(f32.modal) for _ in range(B)]
        batch_call = f32.op.bind_execute_batch(xs, outs)
        for _ in range(args.warmup):
            batch_call()
        batched_stage_times(torch, f32.op, xs, outs, max(20, args.steps))
        del xs, outs, batch_call
        f32.op.reserve_batch(0)
        src = torch.empty(COPY_BYTES // 4, dtype=torch.float32, device=device).fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(args, "profile run done")
        return 0

    # single-call rate: one evaluation per C-ABI call, in order on one stream
    n_single = max(SINGLE_EVALS, args.steps)
    log(args, "single-call evaluations: warmup %d, timed %d" % (args.warmup * 8, n_single))
    for _ in range(args.warmup * 8):
        f32.step()
    wall_single = timed(torch, f32.step, n_single)
    single = n_single / wall_single
    log(args, "fp32 single call %.1f evals/s (%.2f us/eval)" % (single, 1e6 * wall_single / n_single))
    # host cost of enqueueing one evaluation (the C-ABI call), measured while
    # the device works through the queued evaluations
    torch.cuda.synchronize()
    n_host = 200
    t0 = time.perf_counter()
    for _ in range(n_host):
        f32.step()
    host_us = 1e6 * (time.perf_counter() - t0) / n_host
    torch.cuda.synchronize()

    # headline: one step = one batch of independent evaluations (distinct
    # HBM-resident inputs and outputs), stage-batched (or over lanes)
    B = args.batch
    if args.batch_mode == "stage":
        f32.op.reserve_batch(min(B, BATCH_CAP))
        lanes, lanes_apart = 1, True
    else:
        f32.op.reserve_lanes(args.lanes)
        lanes, lanes_apart = f32.op.lanes()
    xs = [f32.x * (1.0 + 1e-3 * k) for k in range(B)]
    outs = [torch.empty_like(f32.modal) for _ in range(B)]
    batch_call = f32.op.bind_execute_batch(xs, outs)
    batch_call()
    for k in sorted({0, B // 2, B - 1}):  # every evaluation equals the same one run alone
        f32.op.execute(xs[k], f32.modal)
        torch.cuda.synchronize()
        if not torch.equal(outs[k], f32.modal):
            raise RuntimeError("batched evaluation %d differs from the same evaluation run alone" % k)
    log(args, "batched (%s): %d evaluations per step, %d lanes (apart %s); warmup %d, timed %d steps"
        % (args.batch_mode, B, lanes, lanes_apart, args.warmup, args.steps))
    for _ in range(args.warmup):
        batch_call()
    wall = timed(torch, batch_call, args.steps)
    value = B * args.steps / wall
    log(args, "fp32 batched %.1f evals/s (%.2f us/eval)" % (value, 1e6 * wall / (B * args.steps)))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(10):
        batch_call()
    host_batch_us = 1e6 * (time.perf_counter() - t0) / (10 * B)
    torch.cuda.synchronize()
    batched_stages = None
    if args.batch_mode == "stage":
        log(args, "per-stage times of the stage-batched execute")
        batched_stages = batched_stage_times(torch, f32.op, xs, outs, max(20, args.steps))
    del xs, outs, batch_call
    f32.op.reserve_batch(0)
    f32.op.reserve_lanes(1)
    torch.cuda.empty_cache()
    lat_evals = max(LATENCY_EVALS, args.steps)
    log(args, "latency over %d evaluations" % lat_evals)
    per_eval = per_eval_latency(torch, f32, lat_evals)
    log(args, "per-stage times over rotated plans")
    rot = [f32] + [Operator(S, torch, w, modes, groups, nj, fwd, inv, x64, S.SKB_F32, device) for _ in range(7)]
    per_stage = per_stage_times(torch, rot, max(200, args.steps))
    # the same stages on one plan (working set resident in the Infinity Cache)
    per_stage_resident = per_stage_times(torch, [f32], max(200, args.steps))
    log(args, "rotated evaluations (8 plans, %d MB of buffers)" % (8 * f32.op.device_bytes() >> 20))
    rot_calls = [o.call for o in rot]
    k = [0]

    def rot_step():
        rot_calls[k[0] % len(rot_calls)]()
        k[0] += 1

    for _ in range(16):
        rot_step()
    wall_rot = timed(torch, rot_step, n_single)
    del rot_calls
    rot = None
    torch.cuda.empty_cache()
    log(args, "fp64 form")
    for _ in range(max(2, args.warmup // 4)):
        f64.step()
    f64_steps = max(10, args.steps // 4)
    wall64 = timed(torch, f64.step, f64_steps)
    per_eval64 = per_eval_latency(torch, f64, f64_steps)

    # algorithmic (stage-boundary) bytes per launch of each stage, FP32 storage:
    # each stage reads its input once and writes its output once
    nkl, P = len(modes), w.planes
    ret_b, real_b = nkl * P * 8, w.real_elements * 4
    modal_b = nkl * w.fields * nj * 8
    mat_b = len(groups) * nj * w.nz * 4
    algo = {
        "layout": 2 * ret_b,                      # compact in, plane-major copy out (not in the yardstick)
        "transforms": 2 * ret_b + real_b,         # retained in, real planes out, retained out (no read-back)
        "inverse": ret_b + real_b,
        "forward": real_b + ret_b,
        "vertical_gemm": ret_b + modal_b + mat_b,
    }
    eval_bytes = ret_b + real_b + real_b + ret_b + ret_b + modal_b + mat_b  # the yardstick
    kernels = {}  # one evaluation per launch (single-call form)
    for n, ms in per_stage.items():
        avg = sum(ms) / len(ms)
        res = sum(per_stage_resident[n]) / len(per_stage_resident[n])
        kernels[n] = {"avg_us": 1e3 * avg, "algorithmic_bytes": algo[n], "GBps": algo[n] / (avg * 1e-3) / 1e9,
                      "resident_avg_us": 1e3 * res, "in_evaluation": n not in ("inverse", "forward")}
    kernels_batched = None
    if batched_stages:
        # one launch per group of `cap` evaluations; a call holds B / cap launches
        cap = min(B, BATCH_CAP)
        launches = (B + cap - 1) // cap
        kernels_batched = {}
        for n, ms in batched_stages.items():
            call_ms = sum(ms) / len(ms)
            kernels_batched[n] = {"evaluations_per_launch": cap, "avg_us_per_launch": 1e3 * call_ms / launches,
                                  "us_per_evaluation": 1e3 * call_ms / B,
                                  "algorithmic_bytes_per_launch": algo[n] * cap,
                                  "GBps": algo[n] * B / (call_ms * 1e-3) / 1e9}
    if kernels_batched:
        dominant = max(kernels_batched, key=lambda n: kernels_batched[n]["us_per_evaluation"])
    else:
        dominant = max((n for n in kernels if kernels[n]["in_evaluation"]), key=lambda n: kernels[n]["avg_us"])
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        try:
            tf = json.load(open(TRAFFIC_FILE))
            if tf.get("family") == args.family:
                traffic = tf.get("kernels", {}).get(dominant, {}).get("hbm_bytes")
        except ValueError:
            traffic = None
    if kernels_batched:
        dk = kernels_batched[dominant]
        roofline = {
            "bound": "hbm",
            "kernel": dominant,
            "achieved": dk["GBps"],
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": dk["GBps"] / HBM_PEAK_GBPS,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": dk["algorithmic_bytes_per_launch"],
            "evaluations_per_launch": dk["evaluations_per_launch"],
            "timing": "HIP events around blocks of 10 stage-batched calls over the headline's %d distinct inputs "
                      "and outputs (the launches of the timed loop)" % B,
        }
    else:
        dk = kernels[dominant]
        roofline = {
            "bound": "hbm",
            "kernel": dominant,
            "achieved": dk["GBps"],
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": dk["GBps"] / HBM_PEAK_GBPS,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": dk["algorithmic_bytes"],
            "timing": "HIP events around blocks of 50 launches rotating over 8 plans (HBM-resident working set)",
        }
    eval_avg_ms = sum(per_eval) / len(per_eval)
    eval_roofline = {
        "bytes_per_eval": eval_bytes,
        "achieved": eval_bytes / (eval_avg_ms * 1e-3) / 1e9,
        "frac": eval_bytes / (eval_avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS,
        "from": "mean of per-evaluation HIP event pairs",
        "wall_achieved": eval_bytes * value / 1e9,
        "wall_frac": eval_bytes * value / 1e9 / HBM_PEAK_GBPS,
        "wall_from": "the headline rate (batched evaluations)",
        "single_call_wall_frac": eval_bytes * single / 1e9 / HBM_PEAK_GBPS,
        "rotated_evals_per_s": n_single / wall_rot,
        "rotated_wall_frac": eval_bytes * n_single / wall_rot / 1e9 / HBM_PEAK_GBPS,
        "fused_minimum_bytes": ret_b + modal_b + mat_b,
    }

    cpu = None
    err_oracle = None
    if not args.no_cpu_baseline:
        log(args, "cpu baseline (oracle, bounded sample of %.0f s)" % args.cpu_seconds)
        cpu, ref, O = cpu_baseline(args, w, x64, nj, groups, fwd)
        err_oracle = {
            "fp32_max_scale": O.max_relative_error(a, ref),
            "fp32_relative_l2": O.relative_l2_error(a, ref),
            "fp64_max_scale": O.max_relative_error(b, ref),
        }

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "evals_per_step": B,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded complex-Gaussian retained spectra x (1+|k|^2)^-1 (DC real), " + family,
        "config": {
            "workload": "retained-mode operator 256x256 Nz65 fields3: pruned c2r -> r2c (2/3 radial) -> "
                        "K^2-grouped vertical GEMM",
            "nx": 256, "ny": 256, "nz": 65, "fields": 3, "retained_modes": nkl, "vertical_family": args.family,
            "vertical_groups": len(groups), "nj": nj,
            "launch": ("one prepared C-ABI call per step: skb_retained_operator_execute_batch of %d independent "
                       "evaluations (distinct inputs and outputs), " % B) +
                      ("stage-batched: one launch per stage over the batch" if args.batch_mode == "stage" else
                       "over %d lanes on separate hardware queues, forked from and joined to one stream" % lanes),
            "evals_per_step": B, "batch_mode": args.batch_mode, "lanes": lanes, "lanes_apart": lanes_apart,
            "path": "plane-resident transforms" if f32.plane else "two-pass transforms",
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "eval_roofline": eval_roofline,
        "latency_ms": {"p50": percentile(per_eval, 0.5), "p99": percentile(per_eval, 0.99),
                       "mean": eval_avg_ms, "evaluations": len(per_eval)},
        "host_enqueue_us_per_eval": host_us,
        "host_enqueue_us_per_eval_batched": host_batch_us,
        "single_call": {"evals_per_s": single, "ms_per_eval": 1e3 * wall_single / n_single, "evaluations": n_single,
                        "note": "one skb_retained_operator_execute per evaluation, in order on one stream"},
        "kernels": kernels,
        "kernels_batched": kernels_batched,
        "error_vs_fp64_gpu": {"max_scale_normalized": err_max, "relative_l2": err_l2, "tolerance": 1e-6,
                              "passed": err_max <= 1e-6 and err_l2 <= 1e-6},
        "error_vs_fp64_oracle": err_oracle,
        "fp64": {"value": f64_steps / wall64, "unit": "evals/s", "ms_per_step": 1e3 * wall64 / f64_steps,
                 "p50_ms": percentile(per_eval64, 0.5), "p99_ms": percentile(per_eval64, 0.99),
                 "bytes_per_eval": 2 * eval_bytes,
                 "path": "plane-resident transforms" if f64.plane else "two-pass transforms"},
        "device": torch.cuda.get_device_name(0),
        "library": S.version(),
    }
    if not args.no_extras:
        for key, fn in (("advection_flux", bench_flux), ("spectral_pipeline", bench_pipeline),
                        ("constant_stratification", bench_type1), ("batched", bench_batched),
                        ("serving", bench_serving)):
            log(args, "extra: %s" % key)
            try:
                line[key] = fn(args, S, torch, device)
            except Exception as exc:  # a failing extra must not hide the headline
                line[key] = {"error": "%s: %s" % (type(exc).__name__, exc)}
            torch.cuda.synchronize()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
