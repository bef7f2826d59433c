"""Continuous-batching request loop for simulations (no reference counterpart).

Requests (a lattice size, a step count, swimmer and tracer counts, a storage
form, a seed) arrive over time.  The Server admits each arrived request as soon
as its estimated device memory fits the HBM budget and a slot is free, builds
its state on the device, and adds it to the ensemble of its storage form.  The
ensembles are stepped in chunks: every chunk runs the smallest number of steps
any member still needs (capped), so members with different step counts each
stop exactly at their own count; finished members are copied back, removed
and their memory released, and the next arrivals are admitted -- requests join
and leave while others keep running on the GPU.  Every result is
bit-identical to running the request alone (ensemble members do not interact).

Reported: per-request time to result (arrival to result on the host), its p50
and p99, and aggregate MLUPS over the wall time of the run.
"""
import collections
import concurrent.futures
import dataclasses
import heapq
import math
import threading
import time

import numpy as np

from . import _core
from .ensemble import Ensemble


@dataclasses.dataclass
class Request:
    rid: int
    size: int
    steps: int
    swimmers: int
    tracers: int
    storage: str = "f64"
    seed: int = 0
    tau: float = 0.25
    alpha: float = 0.01
    arrival: float = 0.0  # seconds after the start of the run
    # filled in by the server
    admitted: float = float("nan")
    finished: float = float("nan")
    result: dict = None

    @property
    def sites(self):
        return self.size ** 3

    def latency(self):
        return self.finished - self.arrival


def poisson_requests(n, rate, seed=1234, sizes=(24, 96), steps=(200, 5000), swimmers=(16, 4096),
                     tracers=(0, 100000)):
    """n requests with exponential inter-arrival times at `rate` per second:
    cube size uniform in `sizes`, step count log-uniform in `steps`, swimmer
    and tracer counts uniform, FP32 or FP64 storage 50/50."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        out.append(Request(
            rid=i, size=int(rng.integers(sizes[0], sizes[1] + 1)),
            steps=int(round(math.exp(rng.uniform(math.log(steps[0]), math.log(steps[1]))))),
            swimmers=int(rng.integers(swimmers[0], swimmers[1] + 1)),
            tracers=int(rng.integers(tracers[0], tracers[1] + 1)),
            storage="f32" if rng.random() < 0.5 else "f64", seed=seed + 1 + i, arrival=t))
    return out


def estimate_bytes(req):
    """Device memory of a request: two distribution buffers, rho, u (in the
    reference layout and interleaved) and force, the swimmer and tracer
    stores (fields, staging and sort scratch), plus the spreading workspace
    it adds to its ensemble."""
    e = 8 if req.storage == "f64" else 4
    lattice = req.sites * (2 * 15 + 1 + 3 + 3 + 3) * e + 64 * 1024
    swim = req.swimmers * (5 * 24 + 6 * 4 + 8) + 2 * req.swimmers * (176 + 64 * 4 * 2)
    trac = req.tracers * (4 * 24 + 6 * 4 + 8)
    tiles = req.sites // 512 + 1
    return lattice + swim + trac + tiles * 16 + (1 << 20)


def initial_state(req):
    """Synthetic initial fields in the reference layout: rho = 1 + 1e-3 N(0,1),
    u = 1e-3 N(0,1); swimmers uniform with isotropic orientations; tracers
    uniform (numpy PCG64 from the request seed, drawn in that order)."""
    rng = np.random.default_rng(req.seed)
    n = req.sites
    rho = 1.0 + 1e-3 * rng.standard_normal(n)
    u = 1e-3 * rng.standard_normal(3 * n)
    sw_r = rng.uniform(0.0, req.size, (3, req.swimmers))
    nv = rng.standard_normal((3, req.swimmers))
    sw_n = nv / np.linalg.norm(nv, axis=0)
    tr_r = rng.uniform(0.0, req.size, (3, req.tracers))
    return rho, u, sw_r, sw_n, tr_r


def common_params(req):
    """Swimmer parameters of a request (CommonParams fields): P = 1e-3,
    l = 0.1, the request's alpha and seed, and the mobility of README.md's
    example, (1/a - 1/L_h)/(6 pi eta) with a = 0.1, L_h = 1.5, eta = tau/3."""
    eta = req.tau / 3
    return dict(P=1e-3, l=0.1, alpha=req.alpha, mobility=(1.0 / 0.1 - 1.0 / 1.5) / (6.0 * math.pi * eta),
                translational_advection_off=False, rotational_advection_off=False, seed=req.seed)


def build(req):
    """Device objects of a request, initialised as InitFromHydro from its state."""
    from . import lb, swimmers, tracers

    rho, u, sw_r, sw_n, tr_r = initial_state(req)
    L = req.size
    lat = lb.Lattice(L, L, L, req.tau, req.tau, storage=req.storage)
    lat.rho.data.ravel()[:] = rho
    lat.rho.H2D()
    lat.u.data.ravel()[:] = u
    lat.u.H2D()
    lat.ZeroForce()
    things = []
    if req.swimmers:
        cp = swimmers.CommonParams()
        for k, v in common_params(req).items():
            setattr(cp, k, v)
        sw = swimmers.Array(req.swimmers, cp)
        sw.r[:] = sw_r
        sw.n[:] = sw_n
        sw.H2D()
        things.append(sw)
    if req.tracers:
        tr = tracers.Array(req.tracers)
        tr.r[:] = tr_r
        tr.H2D()
        things.append(tr)
    lat.InitFromHydro()
    return lat, things


def collect(lat, things):
    """Final state of a request on the host."""
    out = {}
    for name in ("fOld", "rho", "u"):
        f = getattr(lat, name)
        f.D2H()
        out[name] = np.array(f.data).ravel()
    for th in things:
        th.D2H()
        if isinstance(th, _core.SwimmerArray):
            out["sw_r"], out["sw_n"] = np.array(th.r.data), np.array(th.n.data)
        else:
            out["tr_r"] = np.array(th.r.data)
    return out


def run_alone(req):
    """The request on its own through System.Step (the reference's driver order)."""
    from .System import System

    lat, things = build(req)
    system = System(lat)
    for th in things:
        system.AddThing(th)
    system.Step(req.steps)
    return collect(lat, things)


class Server:
    """Continuous batching of requests on one GPU under an HBM budget.

    The loop thread only enqueues device work: it adds prepared requests to
    the ensemble of their storage form, steps the ensembles in chunks (graph
    replays), and takes finished members out, running at most `ahead` chunks
    ahead of the device (Mark).  Preparing a request (host fields, device
    objects, copies, InitFromHydro) and collecting a result (copies back,
    freeing its memory) run on worker threads; their device calls wait on the
    library stream with the GIL released, so the device has the next chunks
    queued while they run."""

    def __init__(self, hbm_budget=32 << 30, max_members=64, chunk=100, graph=True, keep_results=True, workers=2,
                 ahead=2):
        self.budget = int(hbm_budget)
        self.max_members = max_members
        self.chunk = chunk
        self.keep_results = keep_results
        self.ahead = ahead
        self.ens = {"f64": Ensemble("f64"), "f32": Ensemble("f32")}
        for e in self.ens.values():
            e.set_graph(graph)
        self.pool = concurrent.futures.ThreadPoolExecutor(max_workers=workers)
        self.lock = threading.Lock()
        self.used = 0
        self.active = {}     # rid -> [req, lattice, things, ensemble key, member, remaining]
        self.preparing = {}  # rid -> (req, future of build)
        self.finishing = {}  # rid -> future of the collection
        self.marks = collections.deque()
        self.done = []
        self.rejected = []
        self.cancelled = []
        self._clock = time.perf_counter
        self._on_done = None

    def fits(self, req):
        busy = len(self.active) + len(self.preparing)
        with self.lock:
            return busy < self.max_members and self.used + estimate_bytes(req) <= self.budget

    def submit(self, req, now):
        """Reserve the request's memory and prepare it on a worker thread."""
        with self.lock:
            self.used += estimate_bytes(req)
        req.admitted = now
        self.preparing[req.rid] = (req, self.pool.submit(build, req))

    def admit_ready(self):
        """Add every prepared request to its ensemble."""
        for rid, (req, fut) in list(self.preparing.items()):
            if not fut.done():
                continue
            del self.preparing[rid]
            lat, things = fut.result()
            member = self.ens[req.storage].add(lat, things)
            self.active[rid] = [req, lat, things, req.storage, member, req.steps]

    def _collect(self, req, lat, things, keep):
        try:
            if keep:
                req.result = collect(lat, things) if self.keep_results else {}
                req.finished = self._clock()
        finally:
            del lat, things  # the device memory is freed here (after the stream has passed it)
            with self.lock:
                self.used -= estimate_bytes(req)
                if keep:
                    self.done.append(req)
            if keep and self._on_done is not None:
                self._on_done(req)

    def release(self, rid, keep):
        """Take a member out (no wait) and collect it on a worker thread."""
        req, lat, things, key, member, _ = self.active.pop(rid)
        self.ens[key].remove(member)
        self.finishing[rid] = self.pool.submit(self._collect, req, lat, things, keep)

    def cancel(self, rid, now=0.0):
        """Drop a running or preparing request: its slot and memory are
        freed, the others continue unchanged."""
        if rid in self.preparing:
            req, fut = self.preparing.pop(rid)
            lat, things = fut.result()
            self.finishing[rid] = self.pool.submit(self._collect, req, lat, things, False)
        else:
            self.release(rid, keep=False)
        self.cancelled.append(rid)

    def step_chunk(self):
        """Advance every ensemble by the steps its closest-to-done member needs
        (<= chunk), then wait until the device is at most `ahead` chunks behind."""
        for key, ens in self.ens.items():
            rem = [a[5] for a in self.active.values() if a[3] == key]
            if not rem:
                continue
            k = min(self.chunk, min(rem))
            ens.step(k)
            for a in self.active.values():
                if a[3] == key:
                    a[5] -= k
        self.marks.append(_core.Mark())
        while len(self.marks) > self.ahead:
            self.marks.popleft().wait()

    def _reap(self):
        for rid, fut in list(self.finishing.items()):
            if fut.done():
                del self.finishing[rid]
                fut.result()  # a failed collection raises here

    def run(self, requests, realtime=True, on_chunk=None, on_done=None):
        """Serve `requests` (arrival times in seconds from the start).  With
        realtime=False arrivals are ignored and everything is admitted as
        budget allows.  on_done(request) is called (from a worker thread) as
        each result arrives.  Returns summary statistics."""
        queue = [(r.arrival if realtime else 0.0, r.rid, r) for r in requests]
        heapq.heapify(queue)
        t0 = time.perf_counter()
        self._clock = lambda: time.perf_counter() - t0  # noqa: E731
        self._on_done = on_done
        clock = self._clock
        while queue or self.active or self.preparing or self.finishing:
            now = clock()
            while queue and queue[0][0] <= now:
                req = queue[0][2]
                if estimate_bytes(req) > self.budget:
                    heapq.heappop(queue)
                    self.rejected.append(req.rid)
                    continue
                if not self.fits(req):
                    break
                heapq.heappop(queue)
                self.submit(req, now)
            self.admit_ready()
            self._reap()
            if not self.active:
                # nothing to step: wait for a preparation, a collection or the next arrival
                waiting = [f for _, f in self.preparing.values()] + list(self.finishing.values())
                timeout = 0.01 if not queue else max(0.0, min(0.01, queue[0][0] - clock()))
                if waiting:
                    concurrent.futures.wait(waiting, timeout=timeout, return_when=concurrent.futures.FIRST_COMPLETED)
                elif queue:
                    time.sleep(timeout)
                continue
            self.step_chunk()
            if on_chunk is not None:
                on_chunk(self)
            for rid in [rid for rid, a in self.active.items() if a[5] == 0]:
                self.release(rid, keep=True)
        for m in self.marks:
            m.wait()
        self.marks.clear()
        wall = clock()
        done = list(self.done)
        lat = np.array([r.latency() for r in done]) if done else np.zeros(0)
        updates = sum(r.sites * r.steps for r in done)
        return {
            "requests": len(requests), "completed": len(done), "rejected": len(self.rejected),
            "cancelled": len(self.cancelled), "wall_s": wall,
            "p50_s": float(np.percentile(lat, 50)) if lat.size else float("nan"),
            "p99_s": float(np.percentile(lat, 99)) if lat.size else float("nan"),
            "mean_s": float(lat.mean()) if lat.size else float("nan"),
            "mlups": updates / wall / 1e6 if wall > 0 else 0.0,
        }

    def close(self):
        self.pool.shutdown(wait=True)


def run_sequential(requests, realtime=True, on_done=None):
    """Baseline: one request at a time in arrival order, each alone with the
    plain per-system step (no batching); on_done(request) after each.

    Also returns busy_s, the summed service times (build + steps + sync of
    each request alone), and floor_s, the latest arrival + service time over
    the requests: no schedule of the same stream can finish sooner if a
    request runs no faster beside others than alone."""
    from .System import System

    t0 = time.perf_counter()
    lat_s = []
    updates = 0
    busy = 0.0
    floor = 0.0
    for req in sorted(requests, key=lambda r: r.arrival):
        if realtime:
            wait = req.arrival - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
        start = time.perf_counter() - t0
        lat, things = build(req)
        system = System(lat)
        for th in things:
            system.AddThing(th)
        system.Step(req.steps)
        _core.synchronize()
        done = time.perf_counter() - t0
        busy += done - start
        floor = max(floor, (req.arrival if realtime else 0.0) + done - start)
        lat_s.append(done - (req.arrival if realtime else 0.0))
        updates += req.sites * req.steps
        if on_done is not None:
            on_done(req)
        del lat, things, system
    wall = time.perf_counter() - t0
    lat_s = np.array(lat_s)
    return {"requests": len(requests), "wall_s": wall, "p50_s": float(np.percentile(lat_s, 50)),
            "p99_s": float(np.percentile(lat_s, 99)), "mean_s": float(lat_s.mean()), "mlups": updates / wall / 1e6,
            "busy_s": busy, "floor_s": floor}
