"""Zero device allocation in the steady state (docs/benchmark-contract.md:98,
196; the reference's tests/allocation_tracker.cpp:42-89 counts operator new
across repeated executions): after every entry point has run once, repeating
them leaves the library's device-allocation counter (skb_device_allocations)
unchanged.  A new retained-mode list under graph capture is refused, since its
tables would need an allocation."""
import sys

import numpy as np
import pytest

import skb_hip as S

pytestmark = pytest.mark.gpu


def note(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def test_repeated_executions_allocate_nothing(device):
    import torch

    note("steady-state allocations")
    f32, f64 = S.SKB_F32, S.SKB_F64
    w = S.Workload(256, 256, 65, 3)
    modes = S.retained_modes(w)
    groups = S.k2_binned_groups(modes, 32)
    fq, iq = S.qr_gaussian_family(w.nz, len(groups), 129)
    x = torch.from_numpy(S.spectral_fixture(w, modes, 3)).to(device=device, dtype=torch.complex64)
    calls = []

    # retained operator, both paths
    for path in (S.SKB_PATH_AUTO, S.SKB_PATH_TWO_PASS):
        op = S.OperatorPlan(w, modes, w.nz, groups, fq, iq, f32)
        op.set_path(path)
        modal = torch.empty(op.modal_elements, dtype=torch.complex64, device=device)
        calls.append(op.bind_execute(x, modal))
        for st in list(S.OperatorPlan.STAGES) + [S.SKB_OP_STAGE_TRANSFORMS]:
            calls.append(op.bind_stage(st, x, modal))
    # pruned plan: interleaved, split, split field views, stages
    ws = S.Workload(64, 64, 9, 2)
    ms = S.retained_modes(ws)
    pr = S.PrunedPlan(ws, ms, f64)
    real = torch.randn(ws.real_elements, dtype=torch.float64, device=device)
    ret = torch.empty(len(ms) * ws.planes, dtype=torch.complex128, device=device)
    re_, im_ = torch.empty(ret.numel(), dtype=torch.float64, device=device), torch.empty(ret.numel(),
                                                                                           dtype=torch.float64,
                                                                                           device=device)
    calls += [lambda: pr.forward(real, ret), lambda: pr.inverse(ret, real),
              lambda: pr.forward_split(real, re_, im_, 0.5), lambda: pr.inverse_split(re_, im_, real)]
    # full FFT with gather / embed (same mode list each time) and field views
    fp = S.FftPlan(ws, f64)
    spec = torch.empty(ws.spectrum_elements, dtype=torch.complex128, device=device)
    calls += [lambda: fp.forward(real, spec), lambda: fp.gather_retained(ms, spec, ret),
              lambda: fp.embed_retained(ms, ret, spec), lambda: fp.inverse(spec, real)]
    one = S.FftPlan(S.Workload(64, 64, 9, 1), f64)
    vol = 64 * 64 * 9
    calls += [lambda: one.forward_view(real[:vol], spec, 9, 18), lambda: one.inverse_view(spec, 9, 18, real[:vol])]
    # vertical products: interleaved, split (reserved once), no-allocation variant
    gk = S.k2_groups(ms)
    fk, ik = S.k2_vertical_family(ws.nz, ws.nj, gk)
    vp = S.VerticalPlan(ws, ws.nj, gk, fk, ik, f64)
    vp.reserve_split()
    md = torch.empty(len(ms) * ws.fields * ws.nj, dtype=torch.complex128, device=device)
    mre, mim = torch.empty(md.numel(), dtype=torch.float64, device=device), torch.empty(md.numel(),
                                                                                          dtype=torch.float64,
                                                                                          device=device)
    calls += [lambda: vp.forward(ret, md), lambda: vp.inverse(md, ret),
              lambda: vp.forward_split(re_, im_, mre, mim), lambda: vp.inverse_split(mre, mim, re_, im_),
              lambda: vp.forward_split_noalloc(re_, im_, mre, mim)]
    # flux graphs (the separate-kernel path's volumes exist once it has been selected)
    wf = S.Workload(64, 64, 9, 4)
    fl = S.FluxPlan(wf, ms, wf.nj, gk, fk, ik, f64)
    fin = torch.randn(fl.input_elements, dtype=torch.complex128, device=device)
    fout = torch.empty(fl.output_elements, dtype=torch.complex128, device=device)
    calls += [lambda: (fl.set_fused(True), fl.execute(fin, fout)), lambda: (fl.set_fused(False), fl.execute(fin, fout))]
    ad = S.AdvectionFluxPlan(wf, ms, S.vertical_derivative_matrix(wf.nz), f32)
    ain = torch.randn(len(ms) * 4 * wf.nz, dtype=torch.complex64, device=device)
    aout = torch.empty(len(ms) * wf.nz, dtype=torch.complex64, device=device)
    calls.append(lambda: ad.execute(ain, aout))
    # spectral pipeline, both policies and the unfused round trip
    for policy in (S.SKB_PIPELINE_WVM_DIRECT, S.SKB_PIPELINE_COMPACT):
        pp = S.SpectralPipeline(ws, ms, gk, fk, ik, policy, f64)
        yo = torch.empty_like(real)
        calls.append(lambda pp=pp, yo=yo: pp.execute(real, yo))
        if policy == S.SKB_PIPELINE_COMPACT:
            calls.append(lambda pp=pp, yo=yo: (pp.set_fused(False), pp.execute(real, yo), pp.set_fused(True)))
    # batch (stacked classes, one graph)
    b = S.Batch(f32)
    gb = S.k2_binned_groups(ms, 8)
    fb, ib = S.qr_gaussian_family(ws.nz, len(gb), 7)
    cls = b.add_class(ws, 2, ms, ws.nz, gb, fb, ib)
    bin_ = torch.randn(b.input_elements(cls), dtype=torch.complex64, device=device)
    bout = torch.empty(b.output_elements(cls), dtype=torch.complex64, device=device)
    b.bind(cls, bin_, bout)
    side = torch.cuda.Stream()
    b.capture(side)
    calls += [lambda: b.run(side), lambda: b.launch(side)]

    for c in calls:  # warm-up: every entry point once
        c()
    torch.cuda.synchronize()
    before = S.device_allocations()
    for _ in range(3):
        for c in calls:
            c()
    torch.cuda.synchronize()
    after = S.device_allocations()
    note("allocations %s -> %s over %d repeated calls" % (before, after, 3 * len(calls)))
    assert after == before


def test_new_mode_list_under_capture_is_refused(device):
    import torch

    note("mode tables under capture")
    w = S.Workload(32, 32, 5, 1)
    modes = S.retained_modes(w)
    fp = S.FftPlan(w, S.SKB_F64)
    spec = torch.zeros(w.spectrum_elements, dtype=torch.complex128, device=device)
    ret = torch.empty(len(modes) * w.planes, dtype=torch.complex128, device=device)
    g = torch.cuda.CUDAGraph()
    with pytest.raises(S.LogicError):
        with torch.cuda.graph(g):
            fp.gather_retained(modes, spec, ret)
    torch.cuda.synchronize()
    fp.gather_retained(modes, spec, ret)  # outside the capture it builds the tables
    torch.cuda.synchronize()
