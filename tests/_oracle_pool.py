"""Oracle trajectories computed in worker processes (test infrastructure).

The GPU tests of the batch and serving workloads compare dozens of cloths
with the CPU oracle; each cloth's oracle trajectory is independent, so they
run in a pool of spawned processes (no GPU use in the workers).
"""
from __future__ import annotations

import concurrent.futures as cf
import multiprocessing as mp
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _paths():
    for p in (ROOT, os.path.join(ROOT, "oracle")):
        if p not in sys.path:
            sys.path.insert(0, p)


def trajectory(job):
    """(cloth, steps) -> (x, v, pcg iterations per step, impulse flag per step)
    after `steps` oracle steps from the cloth's initial state, the kinematic
    sweep applied before each step as the device does."""
    _paths()
    import pyoracle
    from cloth_sim_self_collision_amd import scene
    cl, steps = job
    o = pyoracle.Oracle(cl, "port")
    x, v, t = cl.xyz.copy(), cl.uvw.copy(), 0.0
    its, imps = [], []
    for _ in range(steps):
        scene.apply_sweep(cl, x, t + cl.dt)
        x, v, imp, it, conv = o.step(x, v)
        its.append(int(it))
        imps.append(bool(imp))
        t += cl.dt
    o.close()
    return x, v, its, imps


def trajectories(jobs, workers: int = 8):
    """trajectory() of every job, in order, on up to `workers` processes."""
    workers = max(1, min(workers, len(jobs), os.cpu_count() or 1))
    if workers == 1:
        return [trajectory(j) for j in jobs]
    with cf.ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn")) as ex:
        return list(ex.map(trajectory, jobs))
