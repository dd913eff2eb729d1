"""Served form: a request loop with continuous batching (no reference
counterpart; the reference driver runs one solve to completion per process,
euler3d_set2c.jl:1049-1160).

A request names an equation set / problem, a mesh size, N and an end time
(a number of LSRK54 steps of the problem's own dt).  The server

  * prepares each request's mesh, connectivity and maps on host worker
    threads as soon as it arrives, and its geometry, initial and background
    states and time step on the GPU when it is admitted,
  * admits prepared requests in arrival order while the HBM footprint of the
    resident solves stays within a budget and a slot is free,
  * keeps every admitted state resident on the GPU in one batch per
    dimension (Batch32: FP32 perturbation form; or Batch: plain FP64), so
    the requests on the GPU advance together, the 2D and 3D batches on their
    own streams (k steps per call of the library's batch loop,
    canary_batch_run, or, with graph=True, k replays of the batch's captured
    step graph, re-captured when requests join or leave),
  * retires a request when it reaches its end time (its state is copied to
    the host) and admits waiting ones in the freed room.

Latency of a request = time its result is on the host - its arrival time.
"""
import concurrent.futures as cf
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import problems as P
from .batch import Batch, Batch32
from ._lib import FP64
from .streams import OwnedStream, reap


@dataclass
class Request:
    rid: int
    kind: str            # "vortex2d" or "bubble3d"
    Ne: int
    N: int
    nsteps: int          # end time in LSRK54 steps of the problem's dt
    arrival: float       # seconds after the stream starts
    seed: int = 0
    # filled in by the server
    prepared: float = None
    admitted: float = None
    done: float = None
    result: np.ndarray = field(default=None, repr=False)
    setup: object = field(default=None, repr=False)  # host setup when the server keeps them
    dof: int = 0
    dt: float = 0.0

    @property
    def latency(self):
        return None if self.done is None else self.done - self.arrival


def poisson_stream(n=64, rate=200.0, seed=2024, vortex_ne=(16, 64), vortex_n=(3, 4, 5, 6, 7),
                   bubble_ne=(8, 24), bubble_n=(3, 4, 5), steps=(50, 500)):
    """Seeded Poisson arrivals (mean `rate` per second): half 2D isentropic
    vortices (vortex_ne[0]^2..vortex_ne[1]^2 quads, N in vortex_n), half 3D
    rising bubbles (bubble_ne[0]^3..bubble_ne[1]^3 hexes, N in bubble_n),
    each to an end time of steps[0]..steps[1] steps."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    out = []
    for i in range(n):
        if rng.random() < 0.5:
            kind, ne, N = "vortex2d", int(rng.integers(vortex_ne[0], vortex_ne[1] + 1)), \
                int(rng.choice(vortex_n))
        else:
            kind, ne, N = "bubble3d", int(rng.integers(bubble_ne[0], bubble_ne[1] + 1)), \
                int(rng.choice(bubble_n))
        out.append(Request(rid=i, kind=kind, Ne=ne, N=N,
                           nsteps=int(rng.integers(steps[0], steps[1] + 1)),
                           arrival=float(t[i]), seed=int(rng.integers(1, 1 << 30))))
    return out


def make_setup_host(req):
    """Host part of a request's preparation (mesh, connectivity, maps, 1D
    operators; no geometry, no GPU work): runs on the server's worker
    threads."""
    if req.kind == "vortex2d":
        return P.vortex2d_mesh(Ne=req.Ne, N=req.N)
    if req.kind == "bubble3d":
        return P.bubble3d_mesh(Ne=req.Ne, N=req.N)
    raise ValueError("unknown request kind %r" % req.kind)


def make_setup_device(req, s):
    """Device part: geometry, initial and background states on the GPU and the
    time step from a device reduction; returns (setup, dt)."""
    if req.kind == "vortex2d":
        P.vortex2d_on_device(s)
        dt = P.euler_courant_device(s, s.Qd)
    else:
        P.bubble3d_on_device(s, dtheta=2.0, jitter=1e-3, seed=req.seed)
        dt = P.cfl_device(s, s.Qd) / s.N ** np.sqrt(2)  # the reference's step (euler3d_set2c.jl:1104)
    return s, dt


def make_setup(req):
    """Both parts on the calling thread."""
    return make_setup_device(req, make_setup_host(req))


class Server:
    def __init__(self, precision="fp32", hbm_budget=64 << 30, max_solves=256, step_chunk=8,
                 prep_workers=8, graph=False, epw=None, keep_setups=False, device="cuda",
                 tier_work=None, tier_ratio=2):
        if precision not in ("fp32", "fp64"):
            raise ValueError("precision is 'fp32' (perturbation form) or 'fp64'")
        self.precision = precision
        # tier_work (DOF x steps): requests with at most this much work go to
        # a batch of their own, advanced every iteration, the others every
        # tier_ratio-th iteration while small ones are resident, so small
        # requests do not wait a whole-batch step for each of their steps
        # (None: one batch per dimension, every request in lock step)
        self.tier_work = None if tier_work is None else float(tier_work)
        self.tier_ratio = max(1, int(tier_ratio))
        self.hbm_budget = int(hbm_budget)
        self.max_solves = int(max_solves)
        self.step_chunk = int(step_chunk)
        self.prep_workers = int(prep_workers)
        # graph: replay each batch's step graphs (Batch.run(graph=True)); the
        # worker threads preparing meshes meanwhile make no HIP call that
        # could race with the launches (a collected stream is queued, not
        # destroyed, and reaped from this thread: streams.reap)
        self.graph = bool(graph)
        self.keep_setups = keep_setups
        self.device = torch.device(device)
        self.epw = dict(epw or {})   # (dim, N) -> elements per workgroup
        self.batches = {}
        self.streams = {}
        self.trace = []              # (time, active solves, resident bytes)

    def _batch(self, dim, tier=0):
        # the tiers of one dimension share its stream: every change of a
        # batch's tables stays ordered after that stream's launches
        b = self.batches.get((dim, tier))
        if b is None:
            b = Batch32(dim, device=self.device) if self.precision == "fp32" else \
                Batch(P._lib.EULER, dim, FP64, device=self.device)
            self.batches[(dim, tier)] = b
            if dim not in self.streams:
                self.streams[dim] = OwnedStream(self.device)
        return b

    def _tier(self, r, s):
        if self.tier_work is None:
            return 0
        return 0 if s.ndof() * r.nsteps <= self.tier_work else 1

    def footprint(self, s):
        return self._batch(s.dim).footprint(s)

    def serve(self, requests, setups=None, clock=time.perf_counter):
        """Run the stream; returns the requests with timings and results.
        setups: optional {rid: host setup (make_setup_host)} prepared
        beforehand (then the host part is not in the latency)."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        reap()
        pool = cf.ThreadPoolExecutor(max_workers=self.prep_workers)
        futures, ready, active = {}, [], {}
        resident = 0
        nxt = 0
        t0 = clock()
        done = 0
        try:
            while done < len(reqs):
                now = clock() - t0
                while nxt < len(reqs) and reqs[nxt].arrival <= now:
                    r = reqs[nxt]
                    if setups is not None and r.rid in setups:
                        fut = cf.Future()
                        fut.set_result(setups[r.rid])
                    else:
                        fut = pool.submit(make_setup_host, r)
                    futures[r.rid] = (r, fut)
                    nxt += 1
                for rid in list(futures):
                    r, fut = futures[rid]
                    if fut.done():
                        s = fut.result()
                        r.prepared = clock() - t0
                        r.dof = s.ndof()
                        if self.keep_setups:
                            r.setup = s
                        ready.append((r, s))
                        del futures[rid]
                ready.sort(key=lambda x: x[0].arrival)
                # admission: arrival order, by HBM footprint and free slots
                admitted = False
                while ready and len(active) < self.max_solves:
                    r, s = ready[0]
                    need = self.footprint(s)
                    if active and resident + need > self.hbm_budget:
                        break
                    ready.pop(0)
                    s, dt = make_setup_device(r, s)
                    r.dt = dt
                    b = self._batch(s.dim, self._tier(r, s))
                    sid = b.add(s, dt)
                    s.dev = s.Qd = s.Q0d = None  # the batch holds what it needs
                    active[r.rid] = (r, b, sid, need)
                    resident += need
                    r.admitted = clock() - t0
                    admitted = True
                if admitted:
                    for (dim, _), b in self.batches.items():
                        if b.solves and b.dirty:
                            self._finalize(dim, b)
                if not active:
                    if futures:
                        cf.wait([f for _, f in futures.values()], timeout=0.002,
                                return_when=cf.FIRST_COMPLETED)
                    elif nxt < len(reqs):
                        time.sleep(max(0.0, min(0.002, reqs[nxt].arrival - (clock() - t0))))
                    continue
                # advance the resident solves of each batch by k steps (k <= the
                # nearest end time in that batch); with tiers, the large-work
                # batch only every tier_ratio-th iteration while small ones wait
                rem = {}
                for r, b, sid, _ in active.values():
                    left = r.nsteps - b.solves[sid].steps
                    rem[id(b)] = min(rem.get(id(b), left), left)
                small = any(b.solves for (_, t), b in self.batches.items() if t == 0)
                it = len(self.trace)
                cur = torch.cuda.current_stream(self.device)
                for (dim, tier), b in self.batches.items():
                    if not b.solves:
                        continue
                    if tier > 0 and small and it % self.tier_ratio:
                        continue
                    st = self.streams[dim].stream
                    st.wait_stream(cur)
                    with torch.cuda.stream(st):
                        b.run(min(self.step_chunk, rem[id(b)]), graph=self.graph)
                for dim, st in self.streams.items():
                    cur.wait_stream(st.stream)
                self.trace.append((clock() - t0, len(active), resident))
                # retire finished requests
                fin = [rid for rid, (r, b, sid, _) in active.items()
                       if b.solves[sid].steps >= r.nsteps]
                for rid in fin:
                    r, b, sid, need = active.pop(rid)
                    r.result = b.remove(sid)
                    r.done = clock() - t0
                    resident -= need
                    done += 1
                if fin:
                    for (dim, _), b in self.batches.items():
                        if b.solves and b.dirty:
                            self._finalize(dim, b)
            for b in self.batches.values():
                b.check()
        finally:
            pool.shutdown(wait=True)
            reap()
        self.makespan = clock() - t0
        return reqs

    def _finalize(self, dim, b):
        Ns = sorted({v.N for v in b.solves.values()})
        known = {N: self.epw[(dim, N)] for N in Ns if (dim, N) in self.epw}
        tune = [N for N in Ns if N not in known]
        b.finalize(epw=known, autotune=bool(tune))
        for N in Ns:
            self.epw[(dim, N)] = b.epw[N]


def summary(reqs, makespan):
    """p50 / p99 latency, throughput figures of a served stream."""
    lat = np.array([r.latency for r in reqs])
    dof_stages = sum(5.0 * r.dof * r.nsteps for r in reqs)
    return {
        "requests": len(reqs),
        "p50_latency_s": float(np.percentile(lat, 50)),
        "p99_latency_s": float(np.percentile(lat, 99)),
        "max_latency_s": float(lat.max()),
        "mean_queue_s": float(np.mean([r.admitted - r.arrival for r in reqs])),
        "makespan_s": float(makespan),
        "aggregate_gdof_stage_s": dof_stages / makespan / 1e9,
        "dof_stages": dof_stages,
    }
