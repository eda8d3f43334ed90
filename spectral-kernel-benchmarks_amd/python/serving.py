"""Serving loop for advection-flux requests on one MI355X, with continuous
batching.

Each client owns a resident spectral state on the device: compact retained
(u, v, w, q) of one grid size (4 fields, plane z + Nz*field).  A "3-field"
client (u, v, q) keeps w = 0.  A request advances the client by one
explicit step
    q <- q + dt * F(u, v, w, q),   F = F_xy[-(u q_x + v q_y + w q_z)] / (Nx Ny)^2
computed by the advection-flux plan (spectral d/dx, d/dy; d/dz through an
Nz x Nz matrix; fused physical-space stage; FP32).

Scheduling:
  * Requests arrive as a Poisson process; each names a client.
  * Executors (a plan and its own HIP stream) are pooled per size class.
    The pools are sized by an HBM budget on executor device memory (the
    plans' intermediates and scratch); client states are resident on top of
    it.
  * A request is admitted as soon as an executor of its class is free and the
    client has nothing in flight (a client's steps are sequential).
    Requests that arrive while others run are admitted as executors free up,
    and their work overlaps on the device through the executors' streams.
  * Latency is completion minus arrival, both on the host clock; completion is
    observed by polling each executor's event.

No reference counterpart: the reference runs one workload per process.  No
collective or multi-GPU path.
"""
from __future__ import annotations

import collections
import math
import time
from dataclasses import dataclass, field

import numpy as np

import skb_hip as S

# (shape (Nx, Ny, Nz), state fields, share of requests)
DEFAULT_MIX = [((256, 256, 65), 3, 0.50), ((512, 512, 129), 4, 0.30), ((128, 128, 33), 3, 0.15),
               ((1024, 1024, 257), 4, 0.05)]


@dataclass
class SizeClass:
    shape: tuple
    fields: int  # 3: (u, v, q) with w = 0; 4: (u, v, w, q)
    w: S.Workload = None
    modes: np.ndarray = None
    dz: np.ndarray = None
    amplitude: object = None  # device tensor: per-mode spectral amplitude (1 + k^2 + l^2)^(-5/6)

    @property
    def key(self):
        return (self.shape, self.fields)


@dataclass
class Executor:
    cls: SizeClass
    plan: S.AdvectionFluxPlan
    stream: object
    flux: object
    done: object = None
    request: object = None


@dataclass
class Request:
    rid: int
    client: int
    arrival: float
    start: float = math.nan
    finish: float = math.nan
    sampled: bool = False


@dataclass
class Client:
    cid: int
    cls: SizeClass
    state: object  # device tensor, compact retained (u, v, w, q)
    queue: collections.deque = field(default_factory=collections.deque)
    busy: bool = False
    steps: int = 0


def make_class(shape, fields, device, dz_length=1.0) -> SizeClass:
    import torch

    nx, ny, nz = shape
    c = SizeClass(shape, fields)
    c.w = S.Workload(nx, ny, nz, 4)
    c.modes = S.retained_modes(c.w)
    c.dz = S.vertical_derivative_matrix(nz, dz_length)
    k2 = c.modes["k"].astype(np.float64) ** 2 + c.modes["l"].astype(np.float64) ** 2
    c.amplitude = torch.from_numpy((1.0 + k2) ** (-5.0 / 6.0)).to(device=device, dtype=torch.float32)
    return c


def initial_state(c: SizeClass, seed: int, device):
    """Seeded complex Gaussian x |k|^(-5/3)-like amplitude per mode (DC real);
    generated on the device (the large grids hold hundreds of MB)."""
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(seed)
    n, nz = len(c.modes), c.w.nz
    re = torch.randn((n, 4, nz), generator=g, device=device, dtype=torch.float32)
    im = torch.randn((n, 4, nz), generator=g, device=device, dtype=torch.float32)
    st = torch.complex(re, im) * c.amplitude.view(n, 1, 1)
    dc = int(np.nonzero((c.modes["k"] == 0) & (c.modes["l"] == 0))[0][0])
    st[dc] = st[dc].real.to(torch.complex64)
    if c.fields == 3:
        st[:, 2, :] = 0  # no vertical velocity
    return st.reshape(-1).contiguous()


class Server:
    """Continuous-batching request loop (see the module docstring)."""

    def __init__(self, classes, budget_bytes: float, dt: float, device="cuda", max_slots: int = 8):
        import torch

        self.torch = torch
        self.device = device
        self.dt = dt
        self.classes = {c.key: c for c in classes}
        self.pools = {}
        self.budget = budget_bytes
        self.used = 0
        # one executor per class first, then round-robin additions while the
        # budget allows (bounded by max_slots per class)
        for c in classes:
            self._add_executor(c)
        grew = True
        while grew:
            grew = False
            for c in classes:
                if len(self.pools[c.key]) < max_slots and self.used + self._cost(c) <= self.budget:
                    self._add_executor(c)
                    grew = True

    def _cost(self, c):
        pool = self.pools.get(c.key)
        return pool[0].plan.device_bytes() + len(c.modes) * c.w.nz * 8 if pool else 0

    def _add_executor(self, c):
        torch = self.torch
        plan = S.AdvectionFluxPlan(c.w, c.modes, c.dz, S.SKB_F32)
        ex = Executor(c, plan, torch.cuda.Stream(),
                      torch.empty(len(c.modes) * c.w.nz, dtype=torch.complex64, device=self.device))
        self.pools.setdefault(c.key, []).append(ex)
        self.used += plan.device_bytes() + ex.flux.numel() * 8

    def slots(self):
        return {str(k): len(v) for k, v in self.pools.items()}

    def step_alone(self, client: Client):
        """One request for `client` on the first executor of its class, waited
        for on the host (the reference path for the mix test)."""
        ex = self.pools[client.cls.key][0]
        self._launch(ex, client)
        ex.stream.synchronize()
        client.steps += 1

    def _launch(self, ex: Executor, client: Client):
        torch = self.torch
        c = client.cls
        ex.stream.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(ex.stream):
            ex.plan.execute(client.state, ex.flux, stream=ex.stream)
            q = client.state.view(len(c.modes), 4, c.w.nz)[:, 3, :]
            q.add_(ex.flux.view(len(c.modes), c.w.nz), alpha=self.dt)
            ex.done = torch.cuda.Event()
            ex.done.record(ex.stream)

    def run(self, clients, arrivals, sample_every: int = 0, on_sample=None):
        """Serve `arrivals` = [(t_arrival_seconds, client_index)] in real time.
        Returns the finished requests.  `on_sample(request, client, input_state,
        flux)` receives copies for every `sample_every`-th request (taken on the
        executor's stream, after the evaluation)."""
        torch = self.torch
        torch.cuda.synchronize()
        reqs = [Request(i, c, t) for i, (t, c) in enumerate(arrivals)]
        if sample_every:
            for r in reqs[::sample_every]:
                r.sampled = True
        free = {k: list(v) for k, v in self.pools.items()}
        inflight = []
        nxt = 0
        done = []
        max_inflight = 0
        t0 = time.perf_counter()
        while len(done) < len(reqs):
            now = time.perf_counter() - t0
            while nxt < len(reqs) and reqs[nxt].arrival <= now:
                clients[reqs[nxt].client].queue.append(reqs[nxt])
                nxt += 1
            # completions
            still = []
            for ex, req, snap in inflight:
                if ex.done.query():
                    req.finish = time.perf_counter() - t0
                    cl = clients[req.client]
                    cl.busy = False
                    cl.steps += 1
                    if snap is not None and on_sample is not None:
                        on_sample(req, cl, snap, ex.flux.clone())
                    free[ex.cls.key].append(ex)
                    done.append(req)
                else:
                    still.append((ex, req, snap))
            inflight = still
            # admission: oldest waiting request first
            waiting = sorted((cl.queue[0].arrival, cl.cid) for cl in clients if cl.queue and not cl.busy)
            for _, cid in waiting:
                cl = clients[cid]
                pool = free[cl.cls.key]
                if not pool:
                    continue
                ex = pool.pop()
                req = cl.queue.popleft()
                snap = cl.state.clone() if (req.sampled and on_sample is not None) else None
                req.start = time.perf_counter() - t0
                cl.busy = True
                self._launch(ex, cl)
                inflight.append((ex, req, snap))
            max_inflight = max(max_inflight, len(inflight))
            if not inflight and nxt < len(reqs):
                wait = reqs[nxt].arrival - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(min(wait, 1e-3))
        torch.cuda.synchronize()
        self.max_inflight = max_inflight
        return done


def poisson_arrivals(rate: float, count: int, clients: int, seed: int):
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, count))
    who = rng.integers(0, clients, count)
    return list(zip(t.tolist(), who.tolist()))


def assign_clients(mix, clients: int, seed: int):
    """Client -> class index by the mix's shares (largest remainder), shuffled."""
    shares = np.array([m[2] for m in mix], np.float64)
    raw = shares / shares.sum() * clients
    counts = np.floor(raw).astype(int)
    for i in np.argsort(raw - counts)[::-1][: clients - counts.sum()]:
        counts[i] += 1
    labels = np.concatenate([np.full(n, i) for i, n in enumerate(counts)])
    np.random.default_rng(seed).shuffle(labels)
    return labels.tolist()


def latency_summary(done, clients=None):
    lat = np.array([r.finish - r.arrival for r in done]) * 1e3
    span = max(r.finish for r in done) - min(r.arrival for r in done)
    out = {"requests": len(done), "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
           "mean_ms": float(lat.mean()), "evals_per_s": len(done) / span}
    if clients is not None:
        per = {}
        for r in done:
            per.setdefault(str(clients[r.client].cls.key), []).append(1e3 * (r.finish - r.arrival))
        out["per_class"] = {k: {"requests": len(v), "p50_ms": float(np.percentile(v, 50)),
                                "p99_ms": float(np.percentile(v, 99))} for k, v in per.items()}
    return out
