"""Benchmark of the implicit cloth step with self-contact on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--requests N]

Headline: the serving stream.  N requests arrive as a Poisson process; each
is a rectangular cloth whose triangle count is log-uniform in [1e3, 2e5]
(aspect 1..4), dropped from a seeded height onto its own sphere or box, with
dt 0.01 and ILU(0)-PCG at tolerance 1e-6 / cap 1000.  Each request runs
min(drawn steps, K) steps (drawn uniformly in 50..300; the cap bounds the
run time and is written into `config`).  One step is one StepTime
(self_contact_sparse/main.cpp:213-263): assembly, ILU(0), ILU-PCG, proximity
+ continuous contact pass, post-contact velocity update, in FP64 and
bit-identical to the reference.

The stream is served with continuous batching (cloth_sim_self_collision_amd.
serve.Server: HBM admission, size lanes stepping their resident cloths
together), then the same requests run one after another with the plain
single-cloth step; `value` is the served whole-job vertex-steps/s (inputs
resident once admitted; cloth creation on admission is inside the timed
run, for both legs), `plain` the plain one, with per-step p50/p99 latencies.
Every served result is compared bit for bit with the plain one, and the
first step of each request in the CPU sample with the reference's own code
(`pos_err_m`).

`roofline`: the kernel class with the most device time in the served run
(HIP events around the shared launches on each lane's stream), its
algorithmic bytes (unique bytes its launches must move, summed over the
cloths each launch works on) over its measured launch time.  `cpu_baseline`:
the reference's step compiled from its sources (oracle/_ref, else the C
restatement), one process per host core, each running one step of one
request (largest first), throughput = vertex-steps / wall time.

Also reported: the single 256x256 folding cloth (plain step), the batch of 64
cloths (sides 24..128) stepped by one cs_batch_step, and the 32x32 drape,
each with its CPU baseline.
"""
from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import os
import queue
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from cloth_sim_self_collision_amd import scene  # noqa: E402

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch over algorithmic bytes per launch, keyed by (workload,
# kernel class): FETCH_SIZE + WRITE_SIZE from two rocprofv3 --pmc passes, each
# scaled by what the counter reads for k_stream_copy of known size; the passes
# and the scaling are in profiles/
TRAFFIC_RATIO: dict = {
    # profiles/fold256_hbm_traffic_pmc.md: (34065.4 * 2.000 + 4608.0) KiB / 59.98 MB
    ("fold256", "ilu_sweep"): 1.242,
    # profiles/served_stream_hbm_traffic_pmc.md: batched sweeps, 1.166 + 0.079
    ("served", "ilu_sweep"): 1.245,
}
METRIC = "vertex-steps/s + p99 step latency, served vs plain GPU, at ≤1e-5 m pos err"


def algorithmic_bytes(kind: str, np_: int, nnz: int) -> float:
    """Unique bytes one pass of a kernel class over one cloth must move (FP64
    values, int32 indices); re-reads and caches are not counted."""
    mat = 72.0 * (nnz + np_)          # off-diagonal + diagonal 3x3 blocks
    idx = 4.0 * nnz + 4.0 * (np_ + 1)  # column indices + row offsets
    vec = 24.0 * np_                  # one 3-vector per vertex
    if kind == "ilu_sweep":           # L, U, D^-1 + diagonal index + level rows; r in, z and r.*z out
        return mat + idx + 12.0 * np_ + 3 * vec
    if kind == "spmv":                # A; z, p_old in; p, Ap, p.*Ap out
        return mat + idx + 5 * vec
    if kind == "pcg_reduce":          # the product vector
        return vec
    if kind == "vector_update":       # r, Ap, x, p in; r, x, r.*r out
        return 7 * vec
    if kind == "ilu_factor":          # A in, factor out, indices
        return 2 * mat + idx + 12.0 * np_
    return 0.0


def step_bytes(kind: str, np_: int, nnz: int, iters: int) -> float:
    """Algorithmic bytes of one cloth's step in a kernel class, given its PCG
    iterations: one factorisation; iters + 1 preconditioner sweeps, matrix
    products and vector updates; three reductions per iteration; the
    one-workgroup solve does all of these in one launch."""
    passes = iters + 1
    if kind == "ilu_factor":
        return algorithmic_bytes(kind, np_, nnz)
    if kind in ("ilu_sweep", "spmv", "vector_update"):
        return passes * algorithmic_bytes(kind, np_, nnz)
    if kind == "pcg_reduce":
        return 3 * passes * algorithmic_bytes(kind, np_, nnz)
    if kind == "pcg_workgroup":
        return sum(step_bytes(k, np_, nnz, iters) for k in ("ilu_sweep", "spmv", "vector_update",
                                                            "pcg_reduce"))
    return 0.0


def roofline_from_ticks(ticks: list, copy_gbs, workload=None) -> dict:
    """Dominant kernel class of a batched run from per-tick records
    ({"prof": {class: (ms, launches)}, "members": [(np, nnz, iters, status)]}):
    its algorithmic bytes summed over every member's work ÷ its total launch
    time."""
    ms, n, by = {}, {}, {}
    for t in ticks:
        for k, (m, c) in t["prof"].items():
            if c <= 0:
                continue
            ms[k] = ms.get(k, 0.0) + m
            n[k] = n.get(k, 0) + c
            by[k] = by.get(k, 0.0) + sum(step_bytes(k, a, b, it) for a, b, it, st in t["members"]
                                         if st == 0)
    if not ms:
        return None
    dom = max(ms, key=ms.get)
    share = ms[dom] / sum(ms.values())
    ach = by[dom] / (ms[dom] * 1e-3) / 1e9 if by[dom] > 0 else None
    return {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": (ach / HBM_PEAK_GBS) if ach else None,
            "traffic": (TRAFFIC_RATIO[(workload, dom)] * by[dom] / n[dom]
                        if (workload, dom) in TRAFFIC_RATIO else None),
            "bytes_per_launch": by[dom] / n[dom], "avg_launch_us": ms[dom] / n[dom] * 1e3,
            "launches": n[dom], "share_of_profiled_gpu_time": share,
            "class_ms": {k: round(v, 2) for k, v in sorted(ms.items(), key=lambda kv: -kv[1])},
            "classes": {k: {"launches": n[k], "avg_launch_us": ms[k] / n[k] * 1e3,
                            "bytes_per_launch": by[k] / n[k]} for k in ms},
            "stream_copy_gbs": copy_gbs,
            "frac_of_stream_copy": (ach / copy_gbs) if (ach and copy_gbs) else None}


# ---------------------------------------------------------------------------
# CPU baseline: one process per core, each stepping whole cloths


def _cpu_worker(cpu, kind, jobs, out):
    try:
        os.sched_setaffinity(0, {cpu})
    except (AttributeError, OSError):
        pass
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    while True:
        job = jobs.get()
        if job is None:
            return
        key, cl, steps = job
        o = pyoracle.Oracle(cl, kind)
        x, v, t = cl.xyz.copy(), cl.uvw.copy(), 0.0
        times, x1 = [], None
        t_begin = time.time()
        for k in range(steps):
            scene.apply_sweep(cl, x, t + cl.dt)
            t0 = time.perf_counter()
            x, v, *_ = o.step(x, v)
            times.append(time.perf_counter() - t0)
            if k == 0:
                x1 = x.copy()
            t += cl.dt
        t_end = time.time()
        o.close()
        out.put((key, cl.np_, times, x1, t_begin, t_end, x))


def cpu_baseline(jobs: list, processes: int, what: str) -> tuple:
    """Run `jobs` [(key, cloth, steps)] on `processes` pinned single-core
    processes (largest cloth first); returns (baseline dict, {key: x after
    the first step}, {key: x after the last step})."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    kind = "reference" if pyoracle.available("reference") else "port"
    cpus = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else list(
        range(os.cpu_count() or 1))
    procs_n = max(1, min(processes, len(cpus), len(jobs)))
    ctx = mp.get_context("spawn")
    q_in, q_out = ctx.Queue(), ctx.Queue()
    for j in sorted(jobs, key=lambda j: -j[1].np_ * j[2]):
        q_in.put(j)
    for _ in range(procs_n):
        q_in.put(None)
    procs = [ctx.Process(target=_cpu_worker, args=(cpus[k], kind, q_in, q_out), daemon=True)
             for k in range(procs_n)]
    for p in procs:
        p.start()
    res = []
    while len(res) < len(jobs):
        try:
            res.append(q_out.get(timeout=1.0))
        except queue.Empty:
            if any(p.exitcode not in (None, 0) for p in procs):
                raise RuntimeError("a CPU baseline process failed")
    for p in procs:
        p.join()
    wall = max(r[5] for r in res) - min(r[4] for r in res)
    vsteps = sum(r[1] * len(r[2]) for r in res)
    lat = np.array([t for r in res for t in r[2]]) * 1e3
    base = {"value": vsteps / wall, "unit": "vertex-steps/s", "cores": procs_n, "kind": kind,
            "sample": what, "wall_s": round(wall, 2),
            "p50_step_ms": float(np.percentile(lat, 50)), "p99_step_ms": float(np.percentile(lat, 99)),
            "steps_timed": int(len(lat))}
    return base, {r[0]: r[3] for r in res}, {r[0]: r[6] for r in res}


# ---------------------------------------------------------------------------
# GPU legs


def serve_leg(args, copy_gbs) -> dict:
    """The headline: the served stream and the same requests plain."""
    from cloth_sim_self_collision_amd import serve
    wire = serve.poisson_requests(args.requests, args.rate, seed=args.seed, tri_range=(1.0e3, 2.0e5),
                                  aspect_range=(1.0, 4.0), step_range=(50, 300),
                                  pcg_maxit=args.pcg_maxit)
    drawn = [w["steps"] for w in wire]
    for w in wire:
        w["steps"] = min(w["steps"], args.steps)
    reqs = [serve.parse_request(w, k) for k, w in enumerate(wire)]
    kw = {"solo_from": None if args.solo_from == 0 else args.solo_from} if args.solo_from >= 0 else {}
    if args.lanes:
        kw["lanes"] = tuple(int(v) for v in args.lanes.split(","))
    if args.lane_pcg:
        kw["lane_pcg"] = tuple(int(v) for v in args.lane_pcg.split(","))
    if args.autotune:
        kw["autotune"] = True
    if args.lane_instances:
        kw["lane_instances"] = tuple(int(v) for v in args.lane_instances.split(","))
    if args.lane_window:
        kw["lane_window"] = tuple(float(v) for v in args.lane_window.split(","))
    # warm-up: the three smallest requests, served and plain, W steps each
    small = sorted(reqs, key=lambda r: r.cloth.np_)[:3]
    warm = [serve.Request(f"warm{k}", r.cloth, steps=max(1, args.warmup)) for k, r in enumerate(small)]
    if args.warmup > 0:
        serve.Server(64 << 30, **kw).run(warm, realtime=False)
        serve.run_plain(warm[:1])
    print(f"[bench] serve: {len(reqs)} requests, {sum(r.cloth.nt for r in reqs)} triangles, "
          f"step cap {args.steps}", file=sys.stderr, flush=True)
    srv = serve.Server(64 << 30, profile=not args.no_profile, **kw)
    served = srv.run(reqs, realtime=True)
    print(f"[bench] served in {served['wall_s']:.2f} s ({served['ticks']} ticks)", file=sys.stderr,
          flush=True)
    plain, same = None, None
    if not args.serve_only:
        plain = serve.run_plain(reqs, first_states=True)
        print(f"[bench] plain in {plain['wall_s']:.2f} s", file=sys.stderr, flush=True)
        by_rid = {r.rid: r for r in served["results"]}
        same = (not served["failures"] and not plain["failures"] and len(by_rid) == len(reqs) and
                all(np.array_equal(by_rid[rid].xyz, x) and np.array_equal(by_rid[rid].uvw, v)
                    for rid, x, v in plain["results"]))
    roof = roofline_from_ticks(srv.tick_log, copy_gbs, "served") if srv.tick_log else None
    # the served run's timeline: per tick, its lane, start (s), duration (ms),
    # its members' vertex counts and the device time of each kernel class (ms)
    timeline = [[r["lane"], round(r["t0_s"], 3), round(r["ms"], 1), [m[0] for m in r["members"]],
                 {k: round(v[0], 1) for k, v in r["prof"].items() if v[1]}]
                for r in sorted(srv.tick_log, key=lambda r: r["t0_s"])]
    # per request: id, vertices, arrival, admission and finish (s)
    arr = {r.rid: r.arrival for r in reqs}
    req_times = [[x.rid, x.vertices, round(arr[x.rid], 3), round(x.admitted_s, 3),
                  round(x.finished_s, 3)] for x in sorted(served["results"], key=lambda x: x.rid)]
    return {"reqs": reqs, "drawn_steps": drawn, "served": served, "plain": plain, "same": same,
            "roofline": roof, "timeline": timeline, "req_times": req_times,
            "lane_pcg": srv.lane_pcg, "lane_instances": srv.lane_instances,
            "lane_window": srv.lane_window,
            "lanes": list(srv.bounds),
            "solo_from": srv.solo_from}


def fold_leg(steps: int, copy_gbs, maxit: int) -> dict:
    """The single 256x256 folding cloth, plain step, with the ILU sweep's
    launch time from HIP events on the cloth's stream, priced also against
    two ceilings of its one-workgroup shape: one CU's read bandwidth
    (cs_read_probe) and the issue time of the wave-load instructions one
    launch issues (cs_sweep_shape x cs_issue_probe)."""
    from cloth_sim_self_collision_amd.sim import DeviceCloth, issue_probe, read_probe
    cl = scene.fold_sweep(256, pcg_maxit=maxit)
    with DeviceCloth(cl) as d:
        shape = d.sweep_shape()
        d.step(1)
        d.profile(True, sweep_only=True)
        lat, its = [], []
        t0 = time.perf_counter()
        for _ in range(steps):
            t1 = time.perf_counter()
            its.append(d.step(1)["pcg_iters"])
            lat.append(time.perf_counter() - t1)
        wall = time.perf_counter() - t0
        ms, n = d.profile_read()["ilu_sweep"]
        d.profile(False)
        b = algorithmic_bytes("ilu_sweep", d.np_, d.nnz)
        np_ = d.np_
    ach = b / (ms / n * 1e-3) / 1e9 if n else None
    avg_us = ms / n * 1e3 if n else None
    one_cu = read_probe(1 << 23, 5, 1)
    issue_ns = issue_probe(shape["waves"]) if shape["quad"] else None
    floor_us = shape["wave_loads"] * issue_ns * 1e-3 if issue_ns else None
    print(f"[bench] fold256: {steps} steps in {wall:.2f} s", file=sys.stderr, flush=True)
    return {"workload": "fold256", "np": np_, "nt": cl.nt, "steps": steps,
            "vertex_steps_per_s": np_ * steps / wall, "ms_per_step": wall / steps * 1e3,
            "p50_step_ms": float(np.percentile(np.array(lat) * 1e3, 50)),
            "p99_step_ms": float(np.percentile(np.array(lat) * 1e3, 99)), "pcg_iters": its,
            "roofline": {"kernel": "ilu_sweep", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": ach / HBM_PEAK_GBS if ach else None, "bytes_per_launch": b,
                         "avg_launch_us": avg_us, "launches": n,
                         "traffic": TRAFFIC_RATIO[("fold256", "ilu_sweep")] * b,
                         "frac_of_stream_copy": ach / copy_gbs if (ach and copy_gbs) else None,
                         "one_cu_read_gbs": one_cu,
                         "frac_of_one_cu_read": ach / one_cu if (ach and one_cu) else None,
                         "sweep_shape": shape, "ns_per_wave_load": issue_ns,
                         "issue_floor_us": floor_us,
                         "frac_of_issue_floor": floor_us / avg_us if (floor_us and avg_us) else None}}


def stack_leg(steps: int, maxit: int) -> dict:
    """Four stacked 512x512 sheets (1,048,576 vertices, clearance 5 mm, over a
    bar and the ground; configs/stack4x512.json): whole steps from the
    configuration's initial state, with the per-class device time of each
    step from HIP events (the contact pass is one class)."""
    from cloth_sim_self_collision_amd.sim import DeviceCloth
    cl = scene.load_config(os.path.join(ROOT, "cloth_sim_self_collision_amd", "configs",
                                        "stack4x512.json"))
    cl.pcg_maxit = maxit
    with DeviceCloth(cl) as d:
        d.profile(True)
        lat, rec = [], []
        for _ in range(steps):
            t1 = time.perf_counter()
            st = d.step(1)
            lat.append(time.perf_counter() - t1)
            rec.append({k: st[k] for k in ("pcg_iters", "n_prox", "n_ccd_pass", "n_ccd",
                                           "n_riz_pass", "impulse_applied")})
        prof = d.profile_read()
        d.profile(False)
        np_ = d.np_
    print(f"[bench] stack4x512: {steps} steps in {sum(lat):.2f} s", file=sys.stderr, flush=True)
    return {"workload": "stack4x512", "np": np_, "nt": cl.nt, "steps": steps, "pcg_maxit": maxit,
            "ms_per_step": sum(lat) / steps * 1e3, "step_ms": [t * 1e3 for t in lat],
            "per_step": rec,
            "class_ms_per_step": {k: v[0] / steps for k, v in prof.items() if v[1]}}


def copies_leg(counts, maxit: int) -> list:
    """Batches of N copies of the batch workload's largest cloth (128x128):
    one step each, with the per-launch time of each kernel class, for every
    N in `counts`.  Every member does the same work, so what grows with N is
    what batching itself costs."""
    from cloth_sim_self_collision_amd.sim import DeviceBatch, DeviceCloth
    cls = scene.batch_drop(64, seed=0)
    big = max(cls, key=lambda c: c.np_)
    big.pcg_maxit = maxit
    out = []
    for n in counts:
        devs = [DeviceCloth(big) for _ in range(n)]
        try:
            with DeviceBatch(devs) as b:
                b.step(1)
                b.profile(True)
                t0 = time.perf_counter()
                b.step(1)
                ms = (time.perf_counter() - t0) * 1e3
                prof = b.profile_read()
                b.profile(False)
        finally:
            for d in devs:
                d.close()
        out.append({"copies": n, "np": big.np_, "ms_per_step": ms,
                    "per_launch_us": {k: v[0] / v[1] * 1e3 for k, v in prof.items() if v[1]}})
        print(f"[bench] {n} copies: {ms:.1f} ms per step", file=sys.stderr, flush=True)
    return out


def batch_leg(count: int, steps: int, copy_gbs, maxit: int) -> dict:
    """The batch of `count` cloths (sides 24..128, seed 0, each dropped onto
    its own sphere or box): one plain round (each cloth one step alone), then
    `steps` steps of all of them in one cs_batch_step each."""
    from cloth_sim_self_collision_amd.sim import DeviceBatch, DeviceCloth
    cls = scene.batch_drop(count, seed=0)
    for c in cls:
        c.pcg_maxit = maxit
    nv = sum(c.np_ for c in cls)
    devs = [DeviceCloth(c) for c in cls]
    try:
        t0 = time.perf_counter()
        for d in devs:
            d.step(1)
        t_plain = time.perf_counter() - t0
        print(f"[bench] batch: plain round {t_plain:.2f} s", file=sys.stderr, flush=True)
        ticks = []
        with DeviceBatch(devs) as b:
            b.step(1)  # warm-up of the batched kernels
            b.profile(True)
            lat = []
            t0 = time.perf_counter()
            for _ in range(steps):
                t1 = time.perf_counter()
                st = b.step(1)
                lat.append(time.perf_counter() - t1)
                ticks.append({"prof": b.profile_read(),
                              "members": [(d.np_, d.nnz, s["pcg_iters"], s["status"])
                                          for d, s in zip(devs, st)]})
                b.profile(True)
            t_batch = time.perf_counter() - t0
            b.profile(False)
        print(f"[bench] batch: {steps} batched steps {t_batch:.2f} s", file=sys.stderr, flush=True)
    finally:
        for d in devs:
            d.close()
    lat_ms = np.array(lat) * 1e3
    plain = nv / t_plain
    batched = nv * steps / t_batch
    return {"workload": f"batch_drop{count}", "cloths": count, "vertices": nv, "steps": steps,
            "plain_vertex_steps_per_s": plain, "batched_vertex_steps_per_s": batched,
            "batched_vs_plain": batched / plain, "batched_p50_step_ms": float(np.percentile(lat_ms, 50)),
            "batched_p99_step_ms": float(np.percentile(lat_ms, 99)),
            "pcg_iters_max": max(s["pcg_iters"] for s in st),
            "roofline": roofline_from_ticks(ticks, copy_gbs)}, cls


def drape_leg(steps: int, maxit: int) -> dict:
    """The 32x32 drape over a sphere (1,024 vertices), plain step, from its
    initial state: FP64 and mixed precision (FP32 sweep coefficients); the
    final positions are kept for the comparison with the reference."""
    from cloth_sim_self_collision_amd.sim import DeviceCloth
    cl = scene.drape_over_sphere(32)
    cl.pcg_maxit = maxit
    out = {"workload": "drape32", "np": cl.np_, "steps": steps}
    for leg, prec in (("f64", 0), ("mixed", 1)):
        with DeviceCloth(cl) as d:
            d.set_option(DeviceCloth.OPT_PRECISION, prec)
            lat = []
            t0 = time.perf_counter()
            for _ in range(steps):
                t1 = time.perf_counter()
                d.step(1)
                lat.append(time.perf_counter() - t1)
            wall = time.perf_counter() - t0
            x, _ = d.get_state()
        lat_ms = np.array(lat) * 1e3
        out[leg] = {"ms_per_step": wall / steps * 1e3, "vertex_steps_per_s": cl.np_ * steps / wall,
                    "p50_step_ms": float(np.percentile(lat_ms, 50)),
                    "p99_step_ms": float(np.percentile(lat_ms, 99)), "x": x}
    # the workload's own figures are the FP64 ones
    out.update({k: out["f64"][k] for k in ("ms_per_step", "vertex_steps_per_s", "p50_step_ms",
                                            "p99_step_ms")})
    return out, cl


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2,
                    help="steps per request of the served stream (cap on the drawn 50..300)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--requests", type=int, default=64, help="requests of the served stream")
    ap.add_argument("--rate", type=float, default=20.0, help="Poisson arrival rate, 1/s")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--lanes", default="", help="lane vertex-count bounds, e.g. 2500,21846 "
                                                "(default: the server's)")
    ap.add_argument("--lane-pcg", default="", help="PCG variant per lane, e.g. 2,2,0")
    ap.add_argument("--lane-instances", default="",
                    help="instances per lane that may step at once, e.g. 1,1,1,1,1,2")
    ap.add_argument("--lane-window", default="",
                    help="seconds an idle lane instance gathers members, e.g. 0,0,0,0,0,1")
    ap.add_argument("--autotune", action="store_true",
                    help="lanes pick their PCG variant from measured ticks")
    ap.add_argument("--solo-from", type=int, default=-1,
                    help="vertex count from which a request runs alone (-1: the server's default, "
                         "0: never)")
    ap.add_argument("--hw-queues", type=int, default=16,
                    help="GPU_MAX_HW_QUEUES for this process (HIP's default is 4; at most 32)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--cpu-procs", type=int, default=16,
                    help="CPU baseline processes (at most the cores this process may use)")
    ap.add_argument("--no-profile", action="store_true", help="no per-class HIP events")
    ap.add_argument("--fold-steps", type=int, default=3, help="steps of fold256 (0: skip)")
    ap.add_argument("--batch-count", type=int, default=64, help="cloths of the batch (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--drape-steps", type=int, default=50, help="steps of drape32 (0: skip)")
    ap.add_argument("--copies", default="",
                    help="batches of N copies of the 128x128 cloth, e.g. 1,8,16,32,64 (default: skip)")
    ap.add_argument("--stack-steps", type=int, default=0,
                    help="whole steps of the four stacked 512^2 sheets (0: skip, the default: "
                         "about 1 s of setup and 13 s per step at the PCG cap of 1000)")
    ap.add_argument("--pcg-maxit", type=int, default=1000,
                    help="PCG iteration cap of every workload (1000; lower only for profiling "
                         "runs, whose line is then not the benchmark)")
    ap.add_argument("--serve-only", action="store_true",
                    help="the served stream alone (profiling runs): no plain leg, no other "
                         "workloads, no CPU baselines")
    ap.add_argument("--copy-mb", type=float, default=512.0,
                    help="size of the streaming-copy probe that measures the copy peak")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (every box this "
              "runs on has one GPU; the sharded design is described but not built)", file=sys.stderr)
        return 2

    # hardware queues: every lane and solo request steps on streams of its own
    if args.hw_queues > 0:
        os.environ["GPU_MAX_HW_QUEUES"] = str(min(args.hw_queues, 32))
    from cloth_sim_self_collision_amd import _lib
    if _lib.device_count() < 1:
        print("no HIP device", file=sys.stderr)
        return 1
    from cloth_sim_self_collision_amd.sim import bandwidth_probe
    n_copy = int(args.copy_mb * 2**20 / 8) // 2 * 2
    copy_gbs = bandwidth_probe(n_copy, 10) if n_copy > 0 else None

    if args.requests <= 0:  # profiling runs of the other workloads
        print("[bench] no served stream (--requests 0): no headline value", file=sys.stderr)
    sv = serve_leg(args, copy_gbs) if args.requests > 0 else None
    if args.serve_only:
        args.fold_steps = args.batch_count = args.drape_steps = 0
        args.no_cpu = True
    fold = fold_leg(args.fold_steps, copy_gbs, args.pcg_maxit) if args.fold_steps > 0 else None
    batch, batch_cls = batch_leg(args.batch_count, args.batch_steps, copy_gbs, args.pcg_maxit) \
        if args.batch_count > 0 else (None, None)
    drape, drape_cl = drape_leg(args.drape_steps, args.pcg_maxit) if args.drape_steps > 0 \
        else (None, None)
    stack = stack_leg(args.stack_steps, args.pcg_maxit) if args.stack_steps > 0 else None
    copies = copies_leg([int(v) for v in args.copies.split(",")], args.pcg_maxit) \
        if args.copies else None

    cpu, err = None, None
    if not args.no_cpu:
        t0 = time.perf_counter()
        if sv is not None:
            reqs, plain = sv["reqs"], sv["plain"]
            cpu, first, _ = cpu_baseline([(r.rid, r.cloth, 1) for r in reqs], args.cpu_procs,
                                      f"the first step of each of the {len(reqs)} requests, one "
                                      "process per core, largest first")
            err = max(float(np.abs(first[rid] - x).max()) for rid, x in plain["first"].items()
                      if rid in first)
            print(f"[bench] cpu served sample {time.perf_counter() - t0:.1f} s", file=sys.stderr,
                  flush=True)
        if batch is not None:
            batch["cpu_baseline"], _, _ = cpu_baseline(
                [(k, c, 1) for k, c in enumerate(batch_cls)], args.cpu_procs,
                f"one step of each of the {len(batch_cls)} cloths, one process per core")
        if drape is not None:
            drape["cpu_baseline"], _, last = cpu_baseline(
                [("drape32", drape_cl, args.drape_steps)], 1,
                f"{args.drape_steps} steps of drape32 on one core")
            # position error after all the steps against the reference's own run
            for leg in ("f64", "mixed"):
                drape[leg]["pos_err_m"] = float(np.abs(drape[leg].pop("x") - last["drape32"]).max())
        print(f"[bench] cpu baselines done {time.perf_counter() - t0:.1f} s", file=sys.stderr,
              flush=True)

    out = headline(args, sv, err) if sv is not None else {
        "metric": METRIC, "value": None, "unit": "vertex-steps/s", "n_gpus": 1,
        "note": "no served stream (--requests 0): a profiling run of the other workloads"}
    if drape is not None:
        for leg in ("f64", "mixed"):
            drape[leg].pop("x", None)
    out.update({"cpu_baseline": cpu, "fold256": fold, "batch64": batch, "drape32": drape})
    if stack is not None:
        out["stack4x512"] = stack
    if copies is not None:
        out["batch_copies"] = copies
    # vs_baseline stays null: the reference publishes no number for this metric
    if cpu and out.get("value"):
        out["value_vs_cpu_baseline"] = out["value"] / cpu["value"]
    print(json.dumps(out))
    return 0


def headline(args, sv, err) -> dict:
    """The served stream's part of the bench line."""
    served, plain, reqs = sv["served"], sv["plain"], sv["reqs"]
    nt = [r.cloth.nt for r in reqs]
    by_rid = {r.rid: r for r in served["results"]}
    return {
        "metric": METRIC, "value": served["vertex_steps_per_s"], "unit": "vertex-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": served["wall_s"] / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "served Poisson stream of drop scenes", "requests": len(reqs),
                   "arrival_rate_per_s": args.rate, "seed": args.seed,
                   "tri_range": [1.0e3, 2.0e5], "aspect_range": [1.0, 4.0],
                   "triangles_drawn": [min(nt), max(nt)], "triangles_total": int(sum(nt)),
                   "vertices_total": int(sum(r.cloth.np_ for r in reqs)),
                   "steps_drawn_range": [50, 300], "step_cap": args.steps,
                   "steps_per_request": [r.steps for r in reqs],
                   "dt": reqs[0].cloth.dt, "pcg_tol": reqs[0].cloth.pcg_tol,
                   "pcg_maxit": reqs[0].cloth.pcg_maxit, "lanes_vertex_bounds": sv["lanes"],
                   "lane_pcg": list(sv["lane_pcg"]), "lane_autotune": bool(args.autotune),
                   "lane_instances": list(sv["lane_instances"]),
                   "lane_window_s": list(sv["lane_window"]),
                   "solo_from_vertices": sv["solo_from"],
                   "gpu_max_hw_queues": os.environ.get("GPU_MAX_HW_QUEUES"),
                   "precision": "f64, bit-exact"},
        "p50_step_ms": served["p50_step_ms"], "p99_step_ms": served["p99_step_ms"],
        "plain": {"vertex_steps_per_s": plain["vertex_steps_per_s"], "wall_s": plain["wall_s"],
                  "p50_step_ms": plain["p50_step_ms"], "p99_step_ms": plain["p99_step_ms"]}
        if plain else None,
        "served_wall_s": served["wall_s"],
        "served_vs_plain": served["vertex_steps_per_s"] / plain["vertex_steps_per_s"] if plain else None,
        "served_equals_plain_bitwise": sv["same"],
        "pos_err_m": err,
        "pos_err_scope": "first step of every request: GPU plain step vs the CPU baseline's step",
        "pcg_iters": {str(r.rid): by_rid[r.rid].pcg_iters for r in reqs if r.rid in by_rid},
        "lane_ticks": served["lanes"], "solo_ticks": served["solo"],
        "timeline": sv["timeline"], "request_times": sv["req_times"],
        "roofline": sv["roofline"],
    }


if __name__ == "__main__":
    sys.exit(main())
