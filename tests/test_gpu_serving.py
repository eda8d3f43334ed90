"""Serving loop (continuous batching of advection-flux requests): each
client's result in the mix must equal the same steps run alone, within the
FP32 tolerance 1e-6 (max |a - b| / max |b|), and the run must actually overlap
requests on the device.  Sampled requests are checked against the FP64 path
of the same plan (max |a - b| / max |b| <= 1e-6)."""
import sys

import numpy as np
import pytest

import skb_hip as S
import serving as V

pytestmark = pytest.mark.gpu

TOL = 1e-6


def note(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a = a.astype(np.complex128)
    b = b.astype(np.complex128)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


MIX = [((64, 64, 9), 4, 0.5), ((32, 32, 5), 3, 0.3), ((128, 128, 17), 4, 0.2)]


def build(device, nclients, seed):
    classes = [V.make_class(shape, f, device) for shape, f, _ in MIX]
    labels = V.assign_clients(MIX, nclients, seed)
    clients = [V.Client(i, classes[labels[i]], V.initial_state(classes[labels[i]], 1000 + i, device))
               for i in range(nclients)]
    return classes, clients


def test_serving_mix_equals_alone_and_overlaps(device):
    import torch

    note("serving mix vs alone")
    classes, clients = build(device, 8, 7)
    init = [c.state.clone() for c in clients]
    server = V.Server(classes, budget_bytes=2 << 30, dt=0.05, device=device, max_slots=3, work_cap_ms=1e9)
    assert all(n >= 1 for n in server.slots().values())
    arrivals = V.poisson_arrivals(rate=50000.0, count=40, clients=len(clients), seed=11)
    done = server.run(clients, arrivals)
    assert len(done) == 40 and server.max_inflight > 1
    for r in done:
        assert r.arrival <= r.start <= r.finish
    steps = [c.steps for c in clients]
    assert sum(steps) == 40
    for i, c in enumerate(clients):
        alone = V.Client(i, c.cls, init[i].clone())
        for _ in range(steps[i]):
            server.step_alone(alone)
        torch.cuda.synchronize()
        if steps[i]:
            assert rel(c.state.cpu().numpy(), alone.state.cpu().numpy()) <= TOL, i
    summary = V.latency_summary(done)
    assert summary["p99_ms"] >= summary["p50_ms"] > 0


def test_serving_samples_against_fp64(device):
    import torch

    note("serving sampled vs fp64")
    classes, clients = build(device, 5, 3)
    server = V.Server(classes, budget_bytes=1 << 30, dt=0.05, device=device, max_slots=2)
    samples = []
    server.run(clients, V.poisson_arrivals(20000.0, 15, len(clients), 5), sample_every=2,
               on_sample=lambda req, cl, x, f: samples.append((cl.cls, x, f)))
    assert len(samples) >= 5
    plans64 = {}
    for cls, x, f in samples:
        if cls.key not in plans64:
            plans64[cls.key] = S.AdvectionFluxPlan(cls.w, cls.modes, cls.dz, S.SKB_F64)
        out = torch.empty(len(cls.modes) * cls.w.nz, dtype=torch.complex128, device=device)
        plans64[cls.key].execute(x.to(torch.complex128), out)
        assert rel(f.cpu().numpy(), out.cpu().numpy()) <= TOL


def test_serving_admission_caps_large_work_and_budget_binds(device):
    """Size-aware admission: large requests in flight never exceed the work
    cap (one of the largest class at a time), small ones are not held back,
    results still equal the steps run alone (the large classes run capped to
    leave CUs free, which must not change a bit), and a tight HBM budget
    leaves a class below max_slots."""
    import torch

    note("serving admission")
    classes, clients = build(device, 8, 7)
    init = [c.state.clone() for c in clients]
    probe = V.Server(classes, budget_bytes=8 << 30, dt=0.05, device=device, max_slots=3, work_cap_ms=1e9)
    per_exec = {k: v[0].plan.device_bytes() for k, v in probe.pools.items()}
    del probe
    tight = sum(per_exec.values()) + min(per_exec.values()) // 2  # one executor each, then nothing fits
    server = V.Server(classes, budget_bytes=tight, dt=0.05, device=device, max_slots=3)
    adm = server.admission()
    assert adm["budget_limited_classes"], adm
    # size lanes: every small-class lane sits on a hardware queue of its own
    assert adm["lanes"] == "size" and adm["lane_check"] == "distinct", adm
    assert all(n == 1 for n in server.slots().values())
    done = server.run(clients, V.poisson_arrivals(rate=50000.0, count=30, clients=len(clients), seed=13))
    assert len(done) == 30
    largest = max(server.cost.values())
    assert server.max_large_work <= max(server.work_cap, largest) + 1e-9
    for i, c in enumerate(clients):
        if not c.steps:
            continue
        alone = V.Client(i, c.cls, init[i].clone())
        for _ in range(c.steps):
            server.step_alone(alone)
        torch.cuda.synchronize()
        assert rel(c.state.cpu().numpy(), alone.state.cpu().numpy()) <= TOL, i
    summary = V.latency_summary(done, clients)
    assert set(summary["per_class"]) == {str(c.key) for c in classes if any(x.cls is c and x.steps for x in clients)}
