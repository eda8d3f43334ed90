"""The full batched uneven workload: 96 instances (40 x 128^2 x Nz33 x 3,
40 x 256^2 x Nz65 x 3, 16 x 512^2 x Nz129 x 4) in one batch, FP32, K^2-binned
QR-Gaussian families.  A class stacks at most 2^31 real elements, so the 16
instances at 512^2 form two classes (15 + 1), as the batched bench runs them.

Every instance must equal the same instance evaluated alone through the
single-workload retained-operator plan, within the FP32 bound (1e-6 of the
max scale: the batch pairs planes of neighbouring instances in its row
transforms, which changes rounding, not results).  One instance per size,
filled with the seeded spectral fixture, is also held to the FP64 oracle
(pocketfft transforms, BLAS vertical product) at 1e-6 (src/benchmark.cpp:47,
118-127).  The other instances are device-random spectra: equality with the
alone evaluation needs no particular input."""
import sys

import numpy as np
import pytest

import oracle as O
import skb_hip as S

pytestmark = pytest.mark.gpu

TOL32 = 1e-6
MIX = [((128, 128, 33, 3), 40), ((256, 256, 65, 3), 40), ((512, 512, 129, 4), 16)]


def note(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def chunks(mix):
    out = []
    for shape, inst in mix:
        per = max(1, (1 << 31) // (shape[0] * shape[1] * shape[2] * shape[3]))
        while inst > 0:
            out.append((shape, min(per, inst)))
            inst -= min(per, inst)
    return out


def test_full_96_instance_batch_equals_alone_and_oracle(device):
    import torch

    parts = chunks(MIX)
    assert sum(n for _, n in parts) == 96
    note("96-instance batch: %s" % parts)
    gen = torch.Generator(device=device).manual_seed(96)
    batch = S.Batch(S.SKB_F32)
    cases = []
    fixture_done = set()
    for shape, inst in parts:
        w = S.Workload(*shape)
        modes = S.retained_modes(w)
        groups = S.k2_binned_groups(modes, 32)
        fwd, inv = S.qr_gaussian_family(w.nz, len(groups), 129)
        cls = batch.add_class(w, inst, modes, w.nz, groups, fwd, inv)
        x = torch.randn((len(modes), inst, w.fields, w.nz), dtype=torch.complex64, device=device, generator=gen)
        seeded = None
        if shape not in fixture_done:
            fixture_done.add(shape)
            seeded = S.spectral_fixture(w, modes, 129)
            x[:, 0] = torch.from_numpy(seeded).to(device=device, dtype=torch.complex64).view(len(modes), w.fields,
                                                                                             w.nz)
        y = torch.empty(batch.output_elements(cls), dtype=torch.complex64, device=device)
        batch.bind(cls, x.view(-1), y)
        cases.append((shape, w, modes, groups, fwd, inv, inst, x, y, seeded))
    side = torch.cuda.Stream()
    batch.run(side)
    side.synchronize()
    note("batch ran; comparing instances with their alone evaluations")
    for shape, w, modes, groups, fwd, inv, inst, x, y, seeded in cases:
        op = S.OperatorPlan(w, modes, w.nz, groups, fwd, inv, S.SKB_F32)
        out = y.view(len(modes), inst, w.fields, w.nz)
        modal = torch.empty(op.modal_elements, dtype=torch.complex64, device=device)
        worst = 0.0
        for i in range(inst):
            op.execute(x[:, i].contiguous().view(-1), modal)
            ref = modal.view(len(modes), w.fields, w.nz)
            err = ((out[:, i] - ref).abs().max() / ref.abs().max()).item()
            worst = max(worst, err)
            assert err <= TOL32, (shape, i, err)
        note("%s x%d: worst batch-vs-alone error %.3e" % (shape, inst, worst))
        if seeded is not None:
            ow = O.Workload(*shape)
            expected = O.retained_operator_fast(ow, O.retained_modes(ow), seeded, w.nz, groups["first_mode"].copy(),
                                                groups["mode_count"].copy(), fwd, workers=16)
            got = out[:, 0].contiguous().cpu().numpy().astype(np.complex128).reshape(-1)
            err = O.max_relative_error(got, expected)
            note("%s instance 0 vs FP64 oracle: %.3e" % (shape, err))
            assert err <= TOL32
            assert O.relative_l2_error(got, expected) <= TOL32
        op.close()
    batch.close()
