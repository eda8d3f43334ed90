"""Batched retained-operator evaluations (skb_retained_operator_execute_batch):
independent evaluations spread over the plan's lanes (own streams and
intermediates) must each equal the same evaluation run alone by execute,
bit for bit (the same kernels on the same inputs), whatever the lane count,
batch size or order; a batch allocates nothing once the lanes are reserved;
plans off the plane path keep one lane and run the batch in order.  One
batch evaluation is also held to the FP64 oracle path (FP32 tolerance 1e-6
of the max-scale error against the FP64 plan on identical inputs)."""
import sys

import numpy as np
import pytest

import skb_hip as S

pytestmark = pytest.mark.gpu


def note(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def make(w, precision, device, n_inputs):
    import torch

    modes = S.retained_modes(w)
    groups = S.k2_binned_groups(modes, 32)
    f, i = S.qr_gaussian_family(w.nz, len(groups), 129)
    op = S.OperatorPlan(w, modes, w.nz, groups, f, i, precision)
    cd = S.complex_dtype(precision)
    xs = [torch.from_numpy(S.spectral_fixture(w, modes, 200 + k, -1.0)).to(device=device, dtype=cd)
          for k in range(n_inputs)]
    return op, modes, groups, f, i, xs


def alone(op, xs, device):
    import torch

    outs = []
    for x in xs:
        m = torch.empty(op.modal_elements, dtype=x.dtype, device=device)
        op.execute(x, m)
        outs.append(m)
    torch.cuda.synchronize()
    return outs


@pytest.mark.parametrize("lanes", [2, 3])
def test_batch_equals_alone_bit_for_bit(device, lanes):
    import torch

    note("operator batch lanes=%d" % lanes)
    w = S.Workload(256, 256, 65, 3)
    op, *_, xs = make(w, S.SKB_F32, device, 7)
    assert op.path() == S.SKB_PATH_PLANE
    ref = alone(op, xs, device)
    op.reserve_lanes(lanes)
    n, apart = op.lanes()
    assert n == lanes
    note("lanes apart: %s" % apart)
    for count in (1, 2, 5, 7):
        outs = [torch.full((op.modal_elements,), float("nan"), dtype=torch.complex64, device=device)
                for _ in range(count)]
        op.execute_batch(xs[:count], outs)
        torch.cuda.synchronize()
        for k in range(count):
            assert torch.equal(outs[k], ref[k]), (count, k)
    # reversed order: a lane's intermediates carry nothing between evaluations
    outs = [torch.empty(op.modal_elements, dtype=torch.complex64, device=device) for _ in range(6)]
    op.execute_batch(xs[5::-1], outs)
    torch.cuda.synchronize()
    for k in range(6):
        assert torch.equal(outs[k], ref[5 - k]), k
    # the caller's stream is joined: work queued after the batch sees every output
    side = torch.cuda.Stream(device=device)
    with torch.cuda.stream(side):
        outs = [torch.zeros(op.modal_elements, dtype=torch.complex64, device=device) for _ in range(4)]
        op.execute_batch(xs[:4], outs, stream=side)
        total = sum(o.abs().sum() for o in outs)
    side.synchronize()
    assert total.item() == pytest.approx(sum(r.abs().sum().item() for r in ref[:4]), rel=1e-6)
    op.reserve_lanes(1)
    assert op.lanes()[0] == 1
    outs = [torch.empty(op.modal_elements, dtype=torch.complex64, device=device) for _ in range(3)]
    op.execute_batch(xs[:3], outs)
    torch.cuda.synchronize()
    for k in range(3):
        assert torch.equal(outs[k], ref[k])


def test_batch_allocates_nothing_and_matches_fp64(device):
    import torch

    note("operator batch steady state + fp64")
    w = S.Workload(256, 256, 65, 3)
    op, modes, groups, f, i, xs = make(w, S.SKB_F32, device, 4)
    op.reserve_lanes(2)
    outs = [torch.empty(op.modal_elements, dtype=torch.complex64, device=device) for _ in range(4)]
    call = op.bind_execute_batch(xs, outs)
    call()
    torch.cuda.synchronize()
    before = S.device_allocations()
    for _ in range(5):
        call()
    torch.cuda.synchronize()
    assert S.device_allocations() == before
    op64 = S.OperatorPlan(w, modes, w.nz, groups, f, i, S.SKB_F64)
    op64.reserve_lanes(2)  # not on the plane path: one lane, the batch runs in order
    assert op64.lanes()[0] == 1
    x64 = [x.to(torch.complex128) for x in xs[:2]]
    o64 = [torch.empty(op64.modal_elements, dtype=torch.complex128, device=device) for _ in range(2)]
    op64.execute_batch(x64, o64)
    torch.cuda.synchronize()
    for k in range(2):
        a = outs[k].to(torch.complex128)
        err = (a - o64[k]).abs().max().item() / max(1.0, o64[k].abs().max().item())
        assert err <= 1e-6, (k, err)
    with pytest.raises(ValueError):
        op.execute_batch(xs[:2], [outs[0], outs[0]])
    with pytest.raises(S.SkbError):
        op.reserve_lanes(0)


@pytest.mark.parametrize("nz,cap", [(65, 4), (65, 128), (33, 8)])
def test_stage_batch_equals_alone_bit_for_bit(device, nz, cap):
    """reserve_batch: every stage one launch over a group of evaluations."""
    import torch

    note("operator stage batch nz=%d cap=%d" % (nz, cap))
    w = S.Workload(256, 256, nz, 3)
    op, *_, xs = make(w, S.SKB_F32, device, 9)
    ref = alone(op, xs, device)
    op.reserve_batch(cap)
    assert op.batch() == cap
    for count in (1, 3, cap, 9):
        if count > 9:
            continue
        outs = [torch.full((op.modal_elements,), float("nan"), dtype=torch.complex64, device=device)
                for _ in range(count)]
        op.execute_batch(xs[:count], outs)
        torch.cuda.synchronize()
        for k in range(count):
            assert torch.equal(outs[k], ref[k]), (count, k)
    # reversed order, several groups when cap < 9: the intermediates carry nothing over
    outs = [torch.empty(op.modal_elements, dtype=torch.complex64, device=device) for _ in range(9)]
    op.execute_batch(xs[::-1], outs)
    torch.cuda.synchronize()
    for k in range(9):
        assert torch.equal(outs[k], ref[8 - k]), k
    # steady state allocates nothing
    call = op.bind_execute_batch(xs, outs)
    call()
    torch.cuda.synchronize()
    before = S.device_allocations()
    for _ in range(3):
        call()
    torch.cuda.synchronize()
    assert S.device_allocations() == before
    # the unfused transforms take the lane / in-order path, same results
    op.set_fused(False)
    outs = [torch.empty(op.modal_elements, dtype=torch.complex64, device=device) for _ in range(3)]
    op.execute_batch(xs[:3], outs)
    torch.cuda.synchronize()
    for k in range(3):
        assert torch.equal(outs[k], ref[k]), k
    op.reserve_batch(0)
    assert op.batch() == 0


def test_stage_batch_stages_compose(device):
    """execute_batch_stage LAYOUT, TRANSFORMS, VERTICAL in order over one
    group equal the batch (and each evaluation alone) bit for bit."""
    import torch

    note("operator stage batch by stage")
    w = S.Workload(256, 256, 65, 3)
    op, *_, xs = make(w, S.SKB_F32, device, 5)
    ref = alone(op, xs, device)
    with pytest.raises(S.SkbError):
        op.bind_execute_batch_stage(S.SKB_OP_STAGE_LAYOUT, xs, [torch.empty_like(r) for r in ref])()
    op.reserve_batch(8)
    outs = [torch.full((op.modal_elements,), float("nan"), dtype=torch.complex64, device=device) for _ in range(5)]
    for stage in (S.SKB_OP_STAGE_LAYOUT, S.SKB_OP_STAGE_TRANSFORMS, S.SKB_OP_STAGE_VERTICAL):
        op.bind_execute_batch_stage(stage, xs, outs)()
    torch.cuda.synchronize()
    for k in range(5):
        assert torch.equal(outs[k], ref[k]), k
    with pytest.raises(ValueError):
        op.bind_execute_batch_stage(S.SKB_OP_STAGE_INVERSE, xs, outs)()
    op.reserve_batch(0)


def test_stage_batch_reservation_limits(device):
    note("operator stage batch limits")
    w = S.Workload(256, 256, 65, 3)
    op, modes, groups, f, i, xs = make(w, S.SKB_F32, device, 1)
    with pytest.raises(ValueError):
        op.reserve_batch(129)
    with pytest.raises(ValueError):
        op.reserve_batch(-1)
    op64 = S.OperatorPlan(w, modes, w.nz, groups, f, i, S.SKB_F64)
    with pytest.raises(S.Unsupported):
        op64.reserve_batch(4)
    assert op64.batch() == 0
