"""Batched uneven workload (size classes of stacked instances, one hipGraph)
through the C ABI.  Every instance in the mix must equal the same instance
run alone through the single-workload plans (FP32 within 1e-6, FP64 within
1e-12, max-scale-normalised: pairing planes of neighbouring instances in the
row transforms changes rounding, not results).  The graph is replayed several
times with eager work and synchronisations between replays, and every replay
must equal the eager result bit for bit.  After the launch autotuner changes
the launch shapes (persistent kernels are separate instantiations that may
round differently), the result must agree to the same tolerances."""
import sys

import numpy as np
import pytest

import oracle as O
import skb_hip as S

pytestmark = pytest.mark.gpu


def note(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def host(t):
    import torch

    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def family(w, modes):
    groups = S.k2_binned_groups(modes, 32)
    fwd, inv = S.qr_gaussian_family(w.nz, len(groups), 129)
    return groups, fwd, inv


def alone(w, modes, groups, fwd, inv, x, precision, device):
    import torch

    cd = S.complex_dtype(precision)
    pr = S.PrunedPlan(w, modes, precision)
    vp = S.VerticalPlan(w, w.nz, groups, fwd, inv, precision)
    d_x = torch.from_numpy(x).to(device=device, dtype=cd)
    real = torch.empty(w.real_elements, dtype=S.real_dtype(precision), device=device)
    back = torch.empty(len(modes) * w.planes, dtype=cd, device=device)
    modal = torch.empty(len(modes) * w.fields * w.nz, dtype=cd, device=device)
    pr.inverse(d_x, real)
    pr.forward(real, back)
    vp.forward(back, modal)
    return host(modal).astype(np.complex128)


@pytest.mark.parametrize("precision", [S.SKB_F32, S.SKB_F64])
def test_batch_mix_equals_alone_and_graph_replays(device, precision, tmp_path):
    import torch

    shapes = [((128, 128, 33, 3), 3), ((256, 256, 65, 3), 2), ((64, 64, 17, 4), 2)]
    note("batch mix p=%d" % precision)
    cd = S.complex_dtype(precision)
    batch = S.Batch(precision)
    cases = []
    for shape, inst in shapes:
        w = S.Workload(*shape)
        modes = S.retained_modes(w)
        groups, fwd, inv = family(w, modes)
        cls = batch.add_class(w, inst, modes, w.nz, groups, fwd, inv)
        xs = [S.spectral_fixture(w, modes, 300 + 17 * i + shape[0], -1.0) for i in range(inst)]
        stacked = np.stack([x.reshape(len(modes), w.fields, w.nz) for x in xs], axis=1).reshape(-1)
        d_in = torch.from_numpy(stacked).to(device=device, dtype=cd)
        d_out = torch.empty(batch.output_elements(cls), dtype=cd, device=device)
        batch.bind(cls, d_in, d_out)
        cases.append((w, modes, groups, fwd, inv, inst, xs, d_out))
    side = torch.cuda.Stream()
    batch.run(side)
    side.synchronize()
    eager = [host(c[-1]).copy() for c in cases]
    tol = 1e-6 if precision == S.SKB_F32 else 1e-12
    for (w, modes, groups, fwd, inv, inst, xs, d_out), out in zip(cases, eager):
        got = out.astype(np.complex128).reshape(len(modes), inst, w.fields, w.nz)
        for i in range(inst):
            ref = alone(w, modes, groups, fwd, inv, xs[i], precision, device).reshape(len(modes), w.fields, w.nz)
            assert O.max_relative_error(got[:, i], ref) <= tol, (w, i)
    # graph: capture once, replay with eager work in between
    batch.capture(side)
    for r in range(3):
        for c in cases:
            c[-1].fill_(0)  # eager work on the outputs between replays
        torch.cuda.synchronize()
        batch.launch(side)
        side.synchronize()
        for c, ref in zip(cases, eager):
            assert np.array_equal(host(c[-1]), ref), r
    # autotuned launch shapes (and, for FP32 square families, product forms)
    # give the same numbers to the tolerance; the tuning is kept in a
    # manifest and reapplied without timing
    cache = str(tmp_path / "tuning.json")
    tune = batch.autotune(5, side, cache=cache)
    report = tune["classes"]
    assert not tune["cached"] and len(report) == len(shapes) and all(len(c["stages"]) == 5 for c in report)
    if precision == S.SKB_F32:  # the 33- and 65-level square families offer two product forms
        assert [c["stages"][4]["product_forms"] for c in report] == [2, 2, 1]
    assert tune["spec"] == batch.tuning()
    batch.capture(side)
    batch.launch(side)
    side.synchronize()
    for c, ref in zip(cases, eager):
        assert O.max_relative_error(host(c[-1]).astype(np.complex128), ref.astype(np.complex128)) <= tol
    again = batch.autotune(5, side, cache=cache)
    assert again["cached"] and again["spec"] == tune["spec"] and again["classes"] == report
    # a spec applies whole or not at all
    with pytest.raises(ValueError):
        batch.set_tuning("c0.s1=256;c9.s0=0;")
    with pytest.raises(ValueError):
        batch.set_tuning("c0.s7=0;")
    assert batch.tuning() == tune["spec"]
    # a partial spec changes only its entries
    batch.set_tuning("c0.s0=256;")
    changed = batch.tuning()
    assert "c0.s0=256;" in changed
    assert changed.replace("c0.s0=256;", "") == tune["spec"].replace(tune["spec"].split(";")[0] + ";", "")
    batch.set_tuning(tune["spec"])
    assert batch.tuning() == tune["spec"]


def test_batch_requires_bound_arrays(device):
    import torch

    w = S.Workload(32, 32, 5, 3)
    modes = S.retained_modes(w)
    groups, fwd, inv = family(w, modes)
    batch = S.Batch(S.SKB_F32)
    batch.add_class(w, 2, modes, w.nz, groups, fwd, inv)
    with pytest.raises(S.LogicError):
        batch.run(torch.cuda.Stream())
