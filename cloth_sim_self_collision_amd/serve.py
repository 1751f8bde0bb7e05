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

Wire format (JSON lines, `python -m cloth_sim_self_collision_amd.serve`):

  request  {"id": any JSON value, "scene": a scene configuration
            (scene.from_config; one cloth), "steps": int >= 1,
            "arrival": seconds after the start (optional, default 0),
            "precision": "f64" (optional; the only precision served),
            "return": subset of ["x", "v"] (optional, default ["x"])}
  result   {"id", "status": "ok", "vertices", "triangles", "steps",
            "arrival", "queued_s", "admitted_s", "finished_s",
            "step_ms": [...], "pcg_iters": [...], "impulse_applied": [...],
            "n_prox": [...], "n_ccd": [[...] per step],
            "x": {"dtype": "<f8", "shape": [np, 3], "data": base64}, "v": ...}
           or {"id", "status": "error", "error": message} for a request
           that cannot be served (bad record, unknown scene, precision).

Results are written as requests finish, one line each, in completion order.
"""
from __future__ import annotations

import base64
import dataclasses
import json
import sys
import time
from typing import Callable, Iterable, List, Optional, TextIO

import numpy as np

from . import scene
from .scene import Cloth, SplitMix64
from .sim import DeviceBatch, DeviceCloth

# admission estimate before a cloth exists on the device: bytes per vertex of
# the resident footprint (measured ~5.5 KB/vertex at 65k vertices, rounded up)
BYTES_PER_VERTEX_ESTIMATE = 8192


@dataclasses.dataclass
class Request:
    rid: object  # any JSON value on the wire
    cloth: Cloth
    steps: int
    arrival: float = 0.0  # seconds after the start of the run
    returns: tuple = ("x",)  # arrays of the final state in the wire result


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
    impulse_applied: List[bool] = dataclasses.field(default_factory=list)
    n_prox: List[int] = dataclasses.field(default_factory=list)
    n_ccd: List[list] = dataclasses.field(default_factory=list)


@dataclasses.dataclass
class _Active:
    req: Request
    dev: DeviceCloth
    done: int = 0
    step_ms: List[float] = dataclasses.field(default_factory=list)
    pcg_iters: List[int] = dataclasses.field(default_factory=list)
    impulse_applied: List[bool] = dataclasses.field(default_factory=list)
    n_prox: List[int] = dataclasses.field(default_factory=list)
    n_ccd: List[list] = dataclasses.field(default_factory=list)
    admitted_s: float = 0.0
    memory: int = 0


class Server:
    """Continuous-batching tick loop over resident simulations."""

    def __init__(self, hbm_budget_bytes: int, clock: Callable[[], float] = time.perf_counter,
                 on_result: Optional[Callable[["Result"], None]] = None):
        self.budget = int(hbm_budget_bytes)
        self.clock = clock
        self.on_result = on_result  # called as each request finishes

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
                    a.impulse_applied.append(bool(st["impulse_applied"]))
                    a.n_prox.append(int(st["n_prox"]))
                    a.n_ccd.append([int(c) for c in st["n_ccd"][:st["n_ccd_pass"]]])
                    if a.done >= a.req.steps:
                        x, v = a.dev.get_state()
                        results.append(Result(
                            rid=a.req.rid, xyz=x, uvw=v, step_ms=a.step_ms,
                            queued_s=a.admitted_s - a.req.arrival if realtime else 0.0,
                            admitted_s=a.admitted_s, finished_s=self.clock() - t0,
                            vertices=a.req.cloth.np_, pcg_iters=a.pcg_iters,
                            impulse_applied=a.impulse_applied, n_prox=a.n_prox,
                            n_ccd=a.n_ccd))
                        if self.on_result is not None:
                            self.on_result(results[-1])
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


def poisson_requests(count: int, rate_per_s: float, *, seed: int = 0, tri_range=(1.0e3, 2.0e5),
                     aspect_range=(1.0, 4.0), step_range=(50, 300), pcg_maxit: int = 1000) -> list:
    """The serving workload as wire requests: Poisson arrivals at `rate_per_s`;
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
        ob = ({"kind": "sphere", "center": [0.0, 0.0, 0.0], "radius": 0.3} if rng.uniform() < 0.5
              else {"kind": "box", "center": [0.0, 0.0, 0.0], "half": [0.25, 0.25, 0.25]})
        js = int(rng.next_u64() & 0x7FFFFFFF)
        out.append({"id": k, "steps": steps, "arrival": t,
                    "scene": {"scene": "drop", "n": nx, "ny": ny, "height": h, "obstacle": ob,
                              "seed": js, "pcg_maxit": pcg_maxit, "name": f"req{k}_{nx}x{ny}"}})
    return out


def poisson_stream(count: int, rate_per_s: float, **kw) -> List[Request]:
    """poisson_requests parsed into Requests."""
    return [parse_request(r, k) for k, r in enumerate(poisson_requests(count, rate_per_s, **kw))]


# ---------------------------------------------------------------------------
# wire format

PRECISIONS = ("f64",)


def encode_array(a: np.ndarray) -> dict:
    """A float64 array as {"dtype": "<f8", "shape", "data": base64 of its bytes}."""
    a = np.ascontiguousarray(a, dtype="<f8")
    return {"dtype": "<f8", "shape": list(a.shape), "data": base64.b64encode(a.tobytes()).decode()}


def decode_array(obj: dict) -> np.ndarray:
    if obj.get("dtype") != "<f8":
        raise ValueError("only <f8 arrays")
    a = np.frombuffer(base64.b64decode(obj["data"]), dtype="<f8")
    return a.reshape(obj["shape"]).copy()


def parse_request(obj: dict, index: int = 0) -> Request:
    """A wire request (see the module docstring) as a Request; ValueError
    with the reason if it cannot be served."""
    if not isinstance(obj, dict):
        raise ValueError("a request is a JSON object")
    rid = obj.get("id", index)
    if "scene" not in obj:
        raise ValueError("missing 'scene'")
    steps = obj.get("steps")
    if not isinstance(steps, int) or isinstance(steps, bool) or steps < 1:
        raise ValueError("'steps' must be an integer >= 1")
    prec = obj.get("precision", "f64")
    if prec not in PRECISIONS:
        raise ValueError(f"precision {prec!r} is not served (served: {', '.join(PRECISIONS)})")
    arrival = float(obj.get("arrival", 0.0))
    if not arrival >= 0.0:
        raise ValueError("'arrival' must be >= 0")
    returns = tuple(obj.get("return", ["x"]))
    if not set(returns) <= {"x", "v"}:
        raise ValueError("'return' takes 'x' and/or 'v'")
    cloth = scene.from_config(obj["scene"])
    if isinstance(cloth, list):
        raise ValueError("a request carries one cloth (use one request per cloth)")
    return Request(rid=rid, cloth=cloth, steps=steps, arrival=arrival, returns=returns)


def result_record(r: Result, req: Optional[Request] = None) -> dict:
    out = {"id": r.rid, "status": "ok", "vertices": r.vertices,
           "triangles": req.cloth.nt if req is not None else None, "steps": len(r.step_ms),
           "arrival": req.arrival if req is not None else None,
           "queued_s": r.queued_s, "admitted_s": r.admitted_s, "finished_s": r.finished_s,
           "step_ms": r.step_ms, "pcg_iters": r.pcg_iters,
           "impulse_applied": r.impulse_applied, "n_prox": r.n_prox, "n_ccd": r.n_ccd}
    returns = req.returns if req is not None else ("x",)
    if "x" in returns:
        out["x"] = encode_array(r.xyz.reshape(-1, 3))
    if "v" in returns:
        out["v"] = encode_array(r.uvw.reshape(-1, 3))
    return out


def error_record(rid, message: str) -> dict:
    return {"id": rid, "status": "error", "error": message}


def serve_jsonl(lines: Iterable[str], out: TextIO, *, hbm_budget_bytes: int,
                realtime: bool = True) -> dict:
    """Serve the JSON-lines requests in `lines`, writing one result line per
    request to `out` as each finishes (error lines first, for requests that
    cannot be served).  Returns the run summary of Server.run."""
    reqs, by_rid = [], {}
    for k, line in enumerate(lines):
        line = line.strip()
        if not line:
            continue
        rid = k
        try:
            obj = json.loads(line)
            rid = obj.get("id", k) if isinstance(obj, dict) else k
            req = parse_request(obj, k)
        except (ValueError, KeyError, TypeError) as e:
            out.write(json.dumps(error_record(rid, str(e))) + "\n")
            out.flush()
            continue
        key = len(reqs)  # internal id: wire ids need not be unique or hashable
        by_rid[key] = req
        reqs.append(dataclasses.replace(req, rid=key))

    def emit(r: Result):
        req = by_rid[r.rid]
        rec = result_record(dataclasses.replace(r, rid=req.rid), req)
        out.write(json.dumps(rec) + "\n")
        out.flush()

    if not reqs:
        return {"results": [], "wall_s": 0.0, "ticks": 0, "vertex_steps_per_s": 0.0,
                "p50_step_ms": 0.0, "p99_step_ms": 0.0}
    return Server(hbm_budget_bytes, on_result=emit).run(reqs, realtime=realtime)


def main(argv=None) -> int:
    import argparse
    ap = argparse.ArgumentParser(description="Serve cloth simulations on one GPU (JSON lines).")
    ap.add_argument("--requests", default="-", help="request lines (default: stdin)")
    ap.add_argument("--results", default="-", help="result lines (default: stdout)")
    ap.add_argument("--budget-gb", type=float, default=64.0, help="HBM budget for resident cloths")
    ap.add_argument("--no-realtime", action="store_true",
                    help="admit every request at once (ignore arrival times)")
    ap.add_argument("--setup-cache", default=None, help="directory of cached per-mesh setups")
    ap.add_argument("--make-poisson", type=int, default=0, metavar="COUNT",
                    help="write COUNT requests of the Poisson serving workload and exit")
    ap.add_argument("--rate", type=float, default=4.0, help="--make-poisson arrival rate, 1/s")
    ap.add_argument("--seed", type=int, default=0, help="--make-poisson seed")
    args = ap.parse_args(argv)
    if args.make_poisson > 0:
        fout = sys.stdout if args.results == "-" else open(args.results, "w")
        for r in poisson_requests(args.make_poisson, args.rate, seed=args.seed):
            fout.write(json.dumps(r) + "\n")
        if fout is not sys.stdout:
            fout.close()
        return 0
    if args.setup_cache:
        from . import _lib
        _lib.set_setup_cache(args.setup_cache)
    fin = sys.stdin if args.requests == "-" else open(args.requests)
    fout = sys.stdout if args.results == "-" else open(args.results, "w")
    try:
        summary = serve_jsonl(fin, fout, hbm_budget_bytes=int(args.budget_gb * 2**30),
                              realtime=not args.no_realtime)
    finally:
        if fin is not sys.stdin:
            fin.close()
        if fout is not sys.stdout:
            fout.close()
    print(json.dumps({k: v for k, v in summary.items() if k != "results"}), file=sys.stderr)
    return 0


if __name__ == "__main__":
    sys.exit(main())
