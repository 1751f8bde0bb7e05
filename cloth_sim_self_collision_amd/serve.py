"""Serving many cloth simulations on one GPU with continuous batching.

Requests (a cloth and a number of steps) arrive over time.  The server keeps
every admitted simulation resident on the device (mesh, matrix, factor, BVH,
state) and runs a tick loop: admit the requests that have arrived while the
resident device footprint stays under the HBM budget, step all active cloths
together with one ``cs_batch_step`` (shared kernel launches, one grid row per
cloth), retire the finished ones and free their memory.  A request that
arrives while others are on the GPU joins at the next tick; its steps overlap
with theirs on the device.

Each cloth's trajectory is bit-identical to stepping it alone (the batched
kernels run each cloth's arithmetic unchanged), so "the same alone and in the
mix" holds exactly, not only within a tolerance.

Reported per request: the final state, the latency of each of its steps (the
wall time of the tick that advanced it) and its queueing delay; per run:
aggregate vertex-steps/s and p50/p99 step latency.
"""
from __future__ import annotations

import dataclasses
import time
from typing import Callable, List, Optional

import numpy as np

from . import scene
from .scene import Cloth, SplitMix64
from .sim import DeviceBatch, DeviceCloth

# admission estimate before a cloth exists on the device: bytes per vertex of
# the resident footprint (measured ~5.5 KB/vertex at 65k vertices, rounded up)
BYTES_PER_VERTEX_ESTIMATE = 8192


@dataclasses.dataclass
class Request:
    rid: int
    cloth: Cloth
    steps: int
    arrival: float = 0.0  # seconds after the start of the run


@dataclasses.dataclass
class Result:
    rid: int
    xyz: np.ndarray
    uvw: np.ndarray
    step_ms: List[float]
    queued_s: float
    admitted_s: float
    finished_s: float
    vertices: int
    pcg_iters: List[int]


@dataclasses.dataclass
class _Active:
    req: Request
    dev: DeviceCloth
    done: int = 0
    step_ms: List[float] = dataclasses.field(default_factory=list)
    pcg_iters: List[int] = dataclasses.field(default_factory=list)
    admitted_s: float = 0.0
    memory: int = 0


class Server:
    """Continuous-batching tick loop over resident simulations."""

    def __init__(self, hbm_budget_bytes: int, clock: Callable[[], float] = time.perf_counter):
        self.budget = int(hbm_budget_bytes)
        self.clock = clock

    def run(self, requests: List[Request], *, realtime: bool = True) -> dict:
        """Serve `requests` (arrival times in seconds from now).  realtime:
        a request is admitted only once its arrival time has passed; if not,
        every request counts as arrived at once (throughput runs)."""
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        for r in pending:
            if r.cloth.np_ * BYTES_PER_VERTEX_ESTIMATE > self.budget:
                raise ValueError(f"request {r.rid} can never fit the HBM budget")
        active: List[_Active] = []
        results: List[Result] = []
        t0 = self.clock()
        resident = 0
        batch: Optional[DeviceBatch] = None
        members: tuple = ()
        ticks = 0
        try:
            while pending or active:
                now = self.clock() - t0
                # admission: arrived, in order, while the footprint fits
                while pending and (not realtime or pending[0].arrival <= now):
                    need = pending[0].cloth.np_ * BYTES_PER_VERTEX_ESTIMATE
                    if active and resident + need > self.budget:
                        break
                    r = pending.pop(0)
                    dev = DeviceCloth(r.cloth)
                    a = _Active(req=r, dev=dev, admitted_s=self.clock() - t0)
                    a.memory = max(dev.memory(), need)
                    resident += a.memory
                    active.append(a)
                if not active:
                    # idle until the next arrival
                    if realtime and pending:
                        time.sleep(max(0.0, min(0.05, pending[0].arrival - now)))
                    continue
                key = tuple(id(a.dev) for a in active)
                if key != members:
                    if batch is not None:
                        batch.close()
                    batch = DeviceBatch([a.dev for a in active])
                    members = key
                ts = self.clock()
                stats = batch.step(1)
                dt_ms = (self.clock() - ts) * 1e3
                ticks += 1
                still = []
                for a, st in zip(active, stats):
                    a.done += 1
                    a.step_ms.append(dt_ms)
                    a.pcg_iters.append(st["pcg_iters"])
                    if a.done >= a.req.steps:
                        x, v = a.dev.get_state()
                        results.append(Result(
                            rid=a.req.rid, xyz=x, uvw=v, step_ms=a.step_ms,
                            queued_s=a.admitted_s - a.req.arrival if realtime else 0.0,
                            admitted_s=a.admitted_s, finished_s=self.clock() - t0,
                            vertices=a.req.cloth.np_, pcg_iters=a.pcg_iters))
                        resident -= a.memory
                        a.dev.close()
                    else:
                        still.append(a)
                if len(still) != len(active):
                    active = still
                    if batch is not None:
                        batch.close()
                    batch = None
                    members = ()
        finally:
            if batch is not None:
                batch.close()
            for a in active:
                a.dev.close()
        wall = self.clock() - t0
        lat = np.array([m for r in results for m in r.step_ms])
        vsteps = sum(r.vertices * len(r.step_ms) for r in results)
        return {"results": sorted(results, key=lambda r: r.rid), "wall_s": wall, "ticks": ticks,
                "vertex_steps_per_s": vsteps / wall if wall > 0 else 0.0,
                "p50_step_ms": float(np.percentile(lat, 50)) if len(lat) else 0.0,
                "p99_step_ms": float(np.percentile(lat, 99)) if len(lat) else 0.0}


def run_plain(requests: List[Request]) -> dict:
    """The same requests one after another, each cloth stepped alone (the
    plain GPU step), for comparison."""
    t0 = time.perf_counter()
    results, lat = [], []
    vsteps = 0
    for r in sorted(requests, key=lambda r: (r.arrival, r.rid)):
        with DeviceCloth(r.cloth) as d:
            ms = []
            for _ in range(r.steps):
                t1 = time.perf_counter()
                d.step(1)
                ms.append((time.perf_counter() - t1) * 1e3)
            x, v = d.get_state()
        lat += ms
        vsteps += r.cloth.np_ * r.steps
        results.append((r.rid, x, v))
    wall = time.perf_counter() - t0
    lat = np.array(lat)
    return {"results": results, "wall_s": wall, "vertex_steps_per_s": vsteps / wall,
            "p50_step_ms": float(np.percentile(lat, 50)), "p99_step_ms": float(np.percentile(lat, 99))}


def poisson_stream(count: int, rate_per_s: float, *, seed: int = 0, tri_range=(1.0e3, 2.0e5),
                   aspect_range=(1.0, 4.0), step_range=(50, 300), pcg_maxit: int = 1000) -> List[Request]:
    """Requests of the serving workload: Poisson arrivals at `rate_per_s`;
    rectangular grids whose triangle count is log-uniform in `tri_range` and
    aspect ratio uniform in `aspect_range`; steps uniform in `step_range`;
    each dropped from a height in [0.3, 0.8] m onto its own sphere or box,
    with the batch workload's physics (dt 0.01)."""
    rng = SplitMix64(seed)
    t = 0.0
    out = []
    for k in range(count):
        t += -np.log(1.0 - rng.uniform()) / rate_per_s
        tri = float(np.exp(np.log(tri_range[0]) +
                           (np.log(tri_range[1]) - np.log(tri_range[0])) * rng.uniform()))
        asp = aspect_range[0] + (aspect_range[1] - aspect_range[0]) * rng.uniform()
        # 2 (nx-1)(ny-1) triangles with nx = aspect * ny
        ny = max(3, int(round(np.sqrt(tri / (2.0 * asp)))) + 1)
        nx = max(3, int(round(asp * (ny - 1))) + 1)
        steps = step_range[0] + int(rng.uniform() * (step_range[1] - step_range[0] + 1))
        h = 0.3 + 0.5 * rng.uniform()
        ob = scene.sphere((0.0, 0.0, 0.0), 0.3) if rng.uniform() < 0.5 else \
            scene.box((0.0, 0.0, 0.0), (0.25, 0.25, 0.25))
        js = int(rng.next_u64() & 0x7FFFFFFF)
        cl = scene.drop_on_obstacle(nx, h, ob, ny=ny, seed=js, pcg_maxit=pcg_maxit,
                                    name=f"req{k}_{nx}x{ny}")
        out.append(Request(rid=k, cloth=cl, steps=steps, arrival=t))
    return out
