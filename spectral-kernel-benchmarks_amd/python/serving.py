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
  * Executors (a plan, its flux buffer and a HIP stream) are pooled per size
    class.  The pools are sized by an HBM budget on executor device memory
    (the plans' intermediates and scratch); client states are resident on
    top of it.
  * Lanes.  A process has few hardware queues (HIP's default is 4); streams
    are spread over them round-robin in creation order, and streams that
    share a queue run in FIFO order, so a small request on a queue behind a
    large request's multi-millisecond kernels waits for all of them.
    lanes="size" (the default): the large classes (below) share ONE stream,
    and the small classes get lanes of their own (one per class, at most
    queues - 1 of them), each checked at set-up to sit on a hardware queue
    apart from the large lane and the other small lanes (a bounded spin
    kernel on one, a tiny kernel on the other), so no small request shares
    a queue with large work and at most one large kernel runs at a time.  lanes="class": one stream per class;
    lanes="executor": one stream per executor.  Executors on one lane run
    their requests back to back in admission order.
  * A request is admitted as soon as an executor of its class is free and the
    client has nothing in flight (a client's steps are sequential).
    Requests that arrive while others run are admitted as executors free up,
    and their work overlaps on the device through the executors' streams.
  * Size-aware admission: every class has a cost, its measured service time
    alone.  Classes costing more than `small_fraction` of the largest cost
    are "large" and share an in-flight work cap (`work_cap_ms`, by default
    the largest cost): a large request waits until the large work in flight
    plus its own fits the cap.  Small requests are never held back by it,
    so a stream of big flux steps cannot fill the device in front of them.
  * CU reserve (reserve_cus, lanes="size" only): a kernel on another
    hardware queue gets no CU while a large kernel's workgroups hold them
    all, so the large classes' plans run capped at (CUs - reserve_cus)
    persistent workgroups, about one per CU; with one large kernel at a time
    on the large lane, reserve_cus CUs stay free for the small classes.
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

    def __init__(self, classes, budget_bytes: float, dt: float, device="cuda", max_slots: int = 8,
                 work_cap_ms: float = None, small_fraction: float = 0.125, lanes: str = "size",
                 reserve_cus: int = 32):
        import os

        import torch

        if lanes not in ("size", "class", "executor"):
            raise ValueError("lanes must be 'size', 'class' or 'executor'")
        self.torch = torch
        self.lanes = lanes
        self.lane = {}
        self.device = device
        self.dt = dt
        self.classes = {c.key: c for c in classes}
        self.pools = {}
        self.budget = budget_bytes
        self.used = 0
        # lanes="size": executors start on one set-up stream; the lanes are
        # made once the classes are sorted into large and small
        self.setup_stream = torch.cuda.Stream(device=device) if lanes == "size" else None
        # one executor per class first, then round-robin additions while the
        # budget allows (bounded by max_slots per class); a class whose pool
        # stopped below max_slots for lack of budget is "budget-limited"
        for c in classes:
            self._add_executor(c)
        self.budget_limited = set()
        grew = True
        while grew:
            grew = False
            for c in classes:
                if len(self.pools[c.key]) >= max_slots:
                    continue
                if self.used + self._cost(c) <= self.budget:
                    self._add_executor(c)
                    grew = True
                else:
                    self.budget_limited.add(c.key)
        self.budget_limited -= {k for k, v in self.pools.items() if len(v) >= max_slots}
        # service time alone of each class (ms), measured on its first executor
        self.cost = {c.key: self._measure(c) for c in classes}
        largest = max(self.cost.values())
        self.large = {k for k, v in self.cost.items() if v > small_fraction * largest}
        self.lane_check = "not run"
        if lanes == "size":
            queues = int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or 4)
            big = torch.cuda.Stream(device=device)
            small = [c.key for c in classes if c.key not in self.large]
            lanes_small = self._distinct_lanes(big, max(1, min(len(small), queues - 1)))
            for k in self.large:
                self.lane[k] = big
            for j, k in enumerate(small):
                self.lane[k] = lanes_small[j % len(lanes_small)]
            for k, pool in self.pools.items():
                for ex in pool:
                    ex.stream = self.lane[k]
        # CU reserve: the large classes' kernels are capped below the CU count
        # (each workgroup loops over work items), which leaves reserve_cus CUs
        # free for the small classes; their cost is then re-measured capped.
        cus = torch.cuda.get_device_properties(device).multi_processor_count
        self.reserve_cus = max(0, min(int(reserve_cus), cus - 8)) if lanes == "size" else 0
        self.large_workgroups = cus - self.reserve_cus if self.reserve_cus else 0
        self.set_reserve(True)
        if self.reserve_cus:
            for k in self.large:
                self.cost[k] = self._measure(self.classes[k])
        largest = max(self.cost.values())
        self.work_cap = largest if work_cap_ms is None else work_cap_ms
        self.max_large_work = 0.0

    def _shares_queue(self, a, b, cycles: int = 4_000_000) -> bool:
        """True when work on stream b waits for work on stream a (the two
        streams sit on one hardware queue): a bounded spin kernel on a, a tiny
        kernel on b, and a look at whether the spin had ended by the time b's
        kernel completed."""
        torch = self.torch
        probe = torch.zeros(64, device=self.device)
        torch.cuda.synchronize()
        with torch.cuda.stream(a):
            torch.cuda._sleep(cycles)
            ea = torch.cuda.Event()
            ea.record(a)
        with torch.cuda.stream(b):
            probe.add_(1.0)
            eb = torch.cuda.Event()
            eb.record(b)
        eb.synchronize()
        shared = ea.query()
        torch.cuda.synchronize()
        return shared

    def _distinct_lanes(self, big, count: int, attempts: int = 16):
        """`count` new streams, each on a hardware queue that neither `big`
        nor another chosen lane uses, where the process has such queues
        (streams are drawn until one passes); lanes that could not be kept
        apart are recorded in lane_check."""
        torch = self.torch
        chosen = []
        for _ in range(count):
            for _ in range(attempts):
                s = torch.cuda.Stream(device=self.device)
                if not any(self._shares_queue(o, s) for o in [big] + chosen):
                    chosen.append(s)
                    break
            else:
                break
        self.lane_check = "distinct" if len(chosen) == count else "only %d of %d lanes apart" % (len(chosen), count)
        while len(chosen) < count:
            chosen.append(chosen[-1] if chosen else torch.cuda.Stream(device=self.device))
        return chosen

    def set_reserve(self, on: bool):
        """Applies (or lifts) the workgroup cap of the large classes' plans."""
        for k in self.large:
            for ex in self.pools[k]:
                ex.plan.set_workgroups(self.large_workgroups if on else 0)

    def _cost(self, c):
        pool = self.pools.get(c.key)
        return pool[0].plan.device_bytes() + len(c.modes) * c.w.nz * 8 if pool else 0

    def _add_executor(self, c):
        torch = self.torch
        plan = S.AdvectionFluxPlan(c.w, c.modes, c.dz, S.SKB_F32)
        if self.lanes == "size":
            stream = self.setup_stream
        elif self.lanes == "class":
            if c.key not in self.lane:
                self.lane[c.key] = torch.cuda.Stream(device=self.device)
            stream = self.lane[c.key]
        else:
            stream = torch.cuda.Stream(device=self.device)
        ex = Executor(c, plan, stream, torch.empty(len(c.modes) * c.w.nz, dtype=torch.complex64, device=self.device))
        self.pools.setdefault(c.key, []).append(ex)
        self.used += plan.device_bytes() + ex.flux.numel() * 8

    def _measure(self, c, reps: int = 3) -> float:
        """Service time alone (ms) of one request of class c, on a scratch state."""
        torch = self.torch
        ex = self.pools[c.key][0]
        state = initial_state(c, 1, self.device)
        ex.plan.execute(state, ex.flux, stream=ex.stream)  # warm
        ex.stream.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            ex.plan.execute(state, ex.flux, stream=ex.stream)
        ex.stream.synchronize()
        del state
        torch.cuda.synchronize()
        return 1e3 * (time.perf_counter() - t0) / reps

    def slots(self):
        return {str(k): len(v) for k, v in self.pools.items()}

    def admission(self):
        """The admission settings, for reports."""
        return {"work_cap_ms": self.work_cap, "lanes": self.lanes, "lane_check": self.lane_check,
                "reserved_cus": self.reserve_cus,
                "large_class_workgroups": self.large_workgroups,
                "class_cost_ms": {str(k): v for k, v in self.cost.items()},
                "large_classes": sorted(str(k) for k in self.large),
                "budget_limited_classes": sorted(str(k) for k in self.budget_limited),
                "max_large_work_in_flight_ms": self.max_large_work}

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
        large_work = 0.0
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
                    if ex.cls.key in self.large:
                        large_work -= self.cost[ex.cls.key]
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
                key = cl.cls.key
                if key in self.large:
                    if large_work > 0 and large_work + self.cost[key] > self.work_cap + 1e-9:
                        continue  # waits for large work in flight to drain
                    large_work += self.cost[key]
                    self.max_large_work = max(self.max_large_work, large_work)
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
