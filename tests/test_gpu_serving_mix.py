"""The serving mix at its four sizes (256^2 x Nz65 x 3, 512^2 x Nz129 x 4,
128^2 x Nz33 x 3, 1024^2 x Nz257 x 4 advection-flux clients): every client's
state after the mixed run equals the same steps run alone (FP32 tolerance
1e-6 of the max scale), and one 1024^2 x Nz257 request's FP32 flux is held to
the FP64 oracle (oracle.advection_flux: pocketfft transforms, taken 16 levels
at a time to bound host memory) at 1e-6 of the max scale and in relative L2
(src/benchmark.cpp:47, 118-127)."""
import sys

import numpy as np
import pytest

import oracle as O
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


@pytest.mark.timeout(300)
def test_serving_four_sizes_equal_alone_and_1024_against_oracle(device):
    import torch

    note("serving mix at the four sizes")
    classes = [V.make_class(shape, f, device) for shape, f, _ in V.DEFAULT_MIX]
    order = [0, 1, 2, 3, 0, 2]  # every class has a client; the 1024^2 class one
    clients = [V.Client(i, classes[c], V.initial_state(classes[c], 2000 + i, device)) for i, c in enumerate(order)]
    init = [c.state.clone() for c in clients]
    server = V.Server(classes, budget_bytes=48 << 30, dt=1e-3, device=device, max_slots=2)
    arrivals = V.poisson_arrivals(rate=2000.0, count=14, clients=len(clients), seed=21)
    big = classes[3]
    samples = []

    def keep(req, cl, x, f):
        if cl.cls is big and not samples:
            samples.append((x, f))

    done = server.run(clients, arrivals, sample_every=1, on_sample=keep)
    assert len(done) == 14
    steps = [c.steps for c in clients]
    note("steps per client %s, max in flight %d" % (steps, server.max_inflight))
    for i, c in enumerate(clients):
        if not steps[i]:
            continue
        alone = V.Client(i, c.cls, init[i].clone())
        for _ in range(steps[i]):
            server.step_alone(alone)
        torch.cuda.synchronize()
        err = rel(c.state.cpu().numpy(), alone.state.cpu().numpy())
        note("client %d %s: mix vs alone %.3e" % (i, c.cls.shape, err))
        assert err <= TOL, i
    del init
    if not samples:  # the 1024^2 client got no request in this arrival draw: run one alone
        cl = clients[3]
        x = cl.state.clone()
        server.step_alone(cl)
        samples.append((x, server.pools[big.key][0].flux.clone()))
    x, f = samples[0]
    note("1024^2 x Nz257 request against the FP64 oracle")
    ow = O.Workload(1024, 1024, 257, 4)
    expected = O.advection_flux(ow, O.retained_modes(ow), big.dz, x.cpu().numpy().astype(np.complex128),
                                z_chunk=16)
    got = f.cpu().numpy().astype(np.complex128)
    err = O.max_relative_error(got, expected)
    note("1024^2 x Nz257 FP32 flux vs FP64 oracle: %.3e" % err)
    assert err <= TOL
    assert O.relative_l2_error(got, expected) <= TOL
