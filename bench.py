"""Benchmark of the implicit cloth step with self-contact on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload fold256]

A step is one StepTime (self_contact_sparse/main.cpp:213-263) of the whole
cloth: assembly, ILU(0), ILU-PCG, proximity + continuous contact pass,
post-contact velocity update, all device-resident in FP64 and bit-identical
to the reference.  The default workload is the single large cloth: a
256x256-vertex 2 m sheet (130,050 triangles), one edge pinned, the opposite
edge swept to fold it onto itself, dt = 0.005, PCG tolerance 1e-6 / cap 1000.

Prints one JSON line: whole-job vertex-steps/s, per-step latency percentiles,
the roofline of the dominant kernel class (algorithmic bytes / measured
average launch time, HIP events on the cloth's stream over the timed
region), the position error of the first step against the CPU reference and
the CPU baseline (the reference's own code, compiled from its sources,
timed on one host core for one step of the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from cloth_sim_self_collision_amd import scene  # noqa: E402

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch (FETCH_SIZE + WRITE_SIZE from rocprofv3 --pmc, each
# scaled by what the counter reads for k_stream_copy of known size), keyed by
# (workload, kernel class); the passes and the scaling are in profiles/
TRAFFIC_BYTES: dict = {
    # profiles/fold256_hbm_traffic_pmc.md: (34064.5 * 2.000 + 4608.0 * 1.000) KiB
    ("fold256", "ilu_sweep"): 74.48e6,
}


def workload(name: str) -> scene.Cloth:
    """fold256 (default), foldN, drapeN, or a scene configuration file (.json)."""
    if name.endswith(".json"):
        cl = scene.load_config(name)
        if isinstance(cl, list):
            raise SystemExit("the bench steps one cloth; the batch workload is --batch-count")
        return cl
    if name == "fold256":
        return scene.fold_sweep(256)
    if name.startswith("fold"):
        return scene.fold_sweep(int(name[4:]))
    if name.startswith("drape"):
        return scene.drape_over_sphere(int(name[5:]))
    raise SystemExit(f"unknown workload {name}")


def algorithmic_bytes(kind: str, np_: int, nnz: int) -> float:
    """Unique bytes one launch of a kernel class must move (FP64 values,
    int32 indices); re-reads and caches are not counted."""
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


def cpu_baseline(cloth: scene.Cloth, gx1: np.ndarray):
    """Time one step of the same workload on the host (1 thread) with the
    reference's compiled code when available, else the restatement."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    kind = "reference" if pyoracle.available("reference") else "port"
    backend = "reference" if kind == "reference" else "port"
    o = pyoracle.Oracle(cloth, backend)
    x = cloth.xyz.copy()
    v = cloth.uvw.copy()
    scene.apply_sweep(cloth, x, cloth.dt)
    t0 = time.perf_counter()
    x1, v1, *_ = o.step(x, v)
    dt_s = time.perf_counter() - t0
    o.close()
    err = float(np.abs(x1 - gx1).max()) if gx1 is not None else None
    return {"value": cloth.np_ / dt_s, "unit": "vertex-steps/s", "cores": 1, "kind": kind,
            "sample": f"1 step of {cloth.name} from the initial state ({dt_s:.1f} s)"}, err


def batch_measure(count: int, steps: int, seed: int = 0) -> dict:
    """The batch workload (independent cloths of seeded sizes 24..128 dropped
    onto their own sphere or box): the same cloths stepped one after another
    with their own launches ("plain") and together by cs_batch_step."""
    from cloth_sim_self_collision_amd.sim import DeviceBatch, DeviceCloth
    cls = scene.batch_drop(count, seed=seed)
    nv = sum(c.np_ for c in cls)
    devs = [DeviceCloth(c) for c in cls]
    try:
        for d in devs:  # warm-up: one step each
            d.step(1)
        t0 = time.perf_counter()
        for d in devs:
            d.step(1)
        t_plain = time.perf_counter() - t0
        print(f"[bench] batch: plain round {t_plain:.2f} s", file=sys.stderr, flush=True)
        with DeviceBatch(devs) as b:
            b.step(1)  # warm-up of the batched kernels
            lat = []
            t0 = time.perf_counter()
            for _ in range(steps):
                t1 = time.perf_counter()
                st = b.step(1)
                lat.append(time.perf_counter() - t1)
            t_batch = time.perf_counter() - t0
        print(f"[bench] batch: {steps} batched steps {t_batch:.2f} s", file=sys.stderr, flush=True)
        lat_ms = np.array(lat) * 1e3
        plain = nv / t_plain
        batched = nv * steps / t_batch
        return {"workload": f"batch_drop{count}", "cloths": count, "vertices": nv,
                "plain_vertex_steps_per_s": plain, "batched_vertex_steps_per_s": batched,
                "speedup": batched / plain, "batched_p50_step_ms": float(np.percentile(lat_ms, 50)),
                "batched_p99_step_ms": float(np.percentile(lat_ms, 99)),
                "pcg_iters_max": max(s["pcg_iters"] for s in st)}
    finally:
        for d in devs:
            d.close()


def serve_measure(count: int, rate: float, seed: int = 0) -> dict:
    """A bounded sample of the serving workload (Poisson arrivals, rectangular
    cloths of log-uniform triangle counts and aspect 1..4, sphere or box):
    served with continuous batching, then the same requests one after another
    with the plain step."""
    from cloth_sim_self_collision_amd import serve
    reqs = serve.poisson_stream(count, rate, seed=seed, tri_range=(1.0e3, 6.0e3),
                                step_range=(5, 15))
    served = serve.Server(hbm_budget_bytes=64 << 30).run(reqs, realtime=True)
    print(f"[bench] served {count} requests in {served['wall_s']:.2f} s", file=sys.stderr, flush=True)
    plain = serve.run_plain(reqs)
    print(f"[bench] plain {count} requests in {plain['wall_s']:.2f} s", file=sys.stderr, flush=True)
    same = all(np.array_equal(a.xyz, b[1]) and np.array_equal(a.uvw, b[2])
               for a, b in zip(served["results"], plain["results"]))
    return {"workload": f"poisson{count}", "requests": count, "rate_per_s": rate,
            "triangles": [int(r.cloth.nt) for r in reqs], "steps": [r.steps for r in reqs],
            "served_vertex_steps_per_s": served["vertex_steps_per_s"],
            "served_p50_step_ms": served["p50_step_ms"], "served_p99_step_ms": served["p99_step_ms"],
            "served_wall_s": served["wall_s"],
            "plain_vertex_steps_per_s": plain["vertex_steps_per_s"],
            "plain_p50_step_ms": plain["p50_step_ms"], "plain_p99_step_ms": plain["p99_step_ms"],
            "plain_wall_s": plain["wall_s"],
            "served_vs_plain": served["vertex_steps_per_s"] / plain["vertex_steps_per_s"],
            "results_identical": bool(same)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="fold256")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-profile", action="store_true", help="time without per-kernel events")
    ap.add_argument("--maxit", type=int, default=None,
                    help="PCG iteration cap override (short profiling runs only)")
    ap.add_argument("--sweep-variant", type=int, default=0,
                    help="ILU sweep implementation (CS_OPT_SWEEP; results are identical)")
    ap.add_argument("--batch-count", type=int, default=64,
                    help="cloths of the batch workload measured beside the headline (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=2)
    ap.add_argument("--serve-count", type=int, default=12,
                    help="requests of the serving sample measured beside the headline (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=4.0, help="Poisson arrival rate, 1/s")
    ap.add_argument("--copy-mb", type=float, default=512.0,
                    help="size of the streaming-copy probe that measures the copy peak")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(every box here has one GPU)", file=sys.stderr)
        return 2

    from cloth_sim_self_collision_amd import _lib
    from cloth_sim_self_collision_amd.sim import DeviceCloth
    if _lib.device_count() < 1:
        print("no HIP device", file=sys.stderr)
        return 1

    from cloth_sim_self_collision_amd.sim import bandwidth_probe, issue_probe, read_probe
    n_copy = int(args.copy_mb * 2**20 / 8) // 2 * 2
    copy_gbs = bandwidth_probe(n_copy, 10) if n_copy > 0 else None
    # the one-workgroup sweep streams its factor through one CU: that CU's
    # read ceiling, over a buffer of the factor's size
    one_cu_gbs = read_probe(1 << 23, 5, 1) if n_copy > 0 else None
    cloth = workload(args.workload)
    if args.maxit is not None:
        cloth.pcg_maxit = args.maxit
    d = DeviceCloth(cloth)
    d.set_option(1, args.sweep_variant)
    # the sweep runs on one CU as a chain of dependent sub-levels, each paced by
    # its waves' vector-load instructions: the issue floor prices them at the
    # cost one load instruction has in a workgroup of the same shape
    shape = d.sweep_shape()
    issue_ns = issue_probe(shape["waves"]) if (shape["quad"] and n_copy > 0) else None
    # warmup; the first step is also compared against the CPU reference
    first_x = None
    for k in range(max(args.warmup, 1)):
        t0 = time.perf_counter()
        d.step(1)
        print(f"[bench] warmup step {k}: {time.perf_counter() - t0:.3f} s", file=sys.stderr,
              flush=True)
        if k == 0:
            first_x, _ = d.get_state()
    # HIP events around each launch of the ILU sweep (the dominant kernel) on
    # the cloth's stream over the timed steps: two per launch.  Events around
    # every kernel class would cost the step about 4%, so the per-class
    # breakdown comes from one more, untimed step below.
    if not args.no_profile:
        d.profile(True, sweep_only=True)
    lat = []
    stats = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        stats.append(d.step(1))
        lat.append(time.perf_counter() - t0)
        print(f"[bench] step {len(lat) - 1}: {lat[-1]:.3f} s, pcg {stats[-1]['pcg_iters']}",
              file=sys.stderr, flush=True)
    t_total = time.perf_counter() - t_start
    sweep_timed = d.profile_read()["ilu_sweep"] if not args.no_profile else None
    prof, red = {}, None
    if not args.no_profile:
        d.profile(True)
        t0 = time.perf_counter()
        d.step(1)
        print(f"[bench] profiled step (every kernel class): {time.perf_counter() - t0:.3f} s",
              file=sys.stderr, flush=True)
        prof = d.profile_read()
        red = d.reduction_stats()
    d.profile(False)

    value = cloth.np_ * args.steps / t_total
    lat_ms = np.array(lat) * 1e3
    roof = None
    if prof:
        kinds = {k: v for k, v in prof.items() if v[1] > 0}
        dom = max(kinds, key=lambda k: kinds[k][0])
        share = kinds[dom][0] / sum(v[0] for v in kinds.values())
        if dom == "ilu_sweep" and sweep_timed and sweep_timed[1] > 0:
            ms, n = sweep_timed
            over = "timed steps"
        else:
            ms, n = kinds[dom]
            over = "profiled step"
        b = algorithmic_bytes(dom, d.np_, d.nnz)
        avg_s = ms / n * 1e-3
        ach = b / avg_s / 1e9 if b > 0 else None
        roof = {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": (ach / HBM_PEAK_GBS) if ach else None,
                "traffic": TRAFFIC_BYTES.get((cloth.name, dom)),
                "bytes_per_launch": b, "avg_launch_us": avg_s * 1e6, "launches": n,
                "launches_timed_over": over, "share_of_gpu_time": share,
                "stream_copy_gbs": copy_gbs,
                "frac_of_stream_copy": (ach / copy_gbs) if (ach and copy_gbs) else None,
                "one_cu_read_gbs": one_cu_gbs,
                "frac_of_one_cu_read": (ach / one_cu_gbs) if (ach and one_cu_gbs) else None}
        if dom == "ilu_sweep" and issue_ns:
            loads = shape["wave_loads"]
            floor_us = loads * issue_ns * 1e-3
            roof.update({"sweep_shape": shape, "ns_per_wave_load": issue_ns,
                         "wave_loads_per_launch": loads, "issue_floor_us": floor_us,
                         "frac_of_issue_floor": floor_us / (avg_s * 1e6)})
    d.close()
    batch = batch_measure(args.batch_count, args.batch_steps) if args.batch_count > 0 else None
    served = serve_measure(args.serve_count, args.serve_rate) if args.serve_count > 0 else None
    cpu, err = (None, None) if args.no_cpu else cpu_baseline(workload(args.workload), first_x)
    out = {
        "metric": "vertex-steps/s + p99 step latency, served vs plain GPU, at <=1e-5 m pos err",
        "value": value, "unit": "vertex-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": t_total / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": cloth.name, "np": cloth.np_, "nt": cloth.nt, "nq": cloth.nq,
                   "dt": cloth.dt, "pcg_tol": cloth.pcg_tol, "pcg_maxit": cloth.pcg_maxit,
                   "path": "plain GPU step, FP64 bit-exact"},
        "p50_step_ms": float(np.percentile(lat_ms, 50)),
        "p99_step_ms": float(np.percentile(lat_ms, 99)),
        "pos_err_m": err,
        "pcg_iters": [s["pcg_iters"] for s in stats],
        "contacts": {"n_prox": stats[-1]["n_prox"], "n_ccd": stats[-1]["n_ccd"],
                     "n_riz_pass": stats[-1]["n_riz_pass"]},
        "kernel_ms_profiled_step": {k: round(v[0], 3) for k, v in prof.items()} if prof else None,
        "reduction_stats": red,
        "roofline": roof,
        "cpu_baseline": cpu,
        "batch": batch,
        "served": served,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
