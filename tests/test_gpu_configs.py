"""The batch and serving workloads at their full size distributions on the
GPU, every cloth bit for bit against the CPU oracle (computed in worker
processes, tests/_oracle_pool.py), and the stacked-sheet workload's whole
step against a digest of the oracle's result (tests/golden/)."""
import hashlib
import io
import json
import os

import numpy as np
import pytest

from cloth_sim_self_collision_amd import scene, serve

import _oracle_pool

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def sim_mod():
    from cloth_sim_self_collision_amd import _lib, sim
    if _lib.device_count() < 1:
        pytest.skip("no HIP device")
    return sim


@pytest.mark.gpu
def test_batch_of_64_cloths_full_size(sim_mod):
    """64 cloths with sides drawn from 24..128 (seed 0), each dropped onto its
    own sphere or box, stepped together for 3 steps with PCG capped at 20
    iterations: once with shared launches per PCG operation (automatic) and
    once with one workgroup per cloth running its whole solve; every cloth's
    positions, velocities and iteration counts equal the oracle's."""
    cls = scene.batch_drop(64, seed=0)
    steps = 3
    for c in cls:
        c.pcg_maxit = 20
    want = _oracle_pool.trajectories([(c, steps) for c in cls])
    for pcg in (0, 2):
        devs = [sim_mod.DeviceCloth(c) for c in cls]
        try:
            for d in devs:
                d.set_option(2, pcg)
            with sim_mod.DeviceBatch(devs) as b:
                st = b.step(steps)
            for k, (d, s, (x, v, its, imps)) in enumerate(zip(devs, st, want)):
                gx, gv = d.get_state()
                assert s["status"] == 0 and s["pcg_iters"] == its[-1], (pcg, k, s)
                assert np.array_equal(gx, x) and np.array_equal(gv, v), (
                    pcg, k, cls[k].name, np.abs(gx - x).max())
        finally:
            for d in devs:
                d.close()


@pytest.mark.gpu
def test_served_stream_full_size_range(sim_mod):
    """The serving workload's distribution at full size: 10 requests (seed 1:
    2,244 to 197,400 triangles, aspect 1..4, Poisson arrivals at 20/s, sphere
    or box), steps capped at 2 and PCG at 20 iterations, served over the
    JSON-lines wire format by the lane server; every decoded final state
    equals the oracle's bit for bit."""
    reqs = serve.poisson_requests(10, 20.0, seed=1, pcg_maxit=20)
    for r in reqs:
        r["steps"] = min(r["steps"], 2)
        r["return"] = ["x", "v"]
    want = _oracle_pool.trajectories([(scene.from_config(r["scene"]), r["steps"]) for r in reqs])
    out = io.StringIO()
    summary = serve.serve_jsonl([json.dumps(r) for r in reqs], out, hbm_budget_bytes=64 << 30,
                                realtime=True)
    recs = {rec["id"]: rec for rec in map(json.loads, out.getvalue().splitlines())}
    assert len(recs) == len(reqs) and not summary["failures"]
    for r, (x, v, its, imps) in zip(reqs, want):
        rec = recs[r["id"]]
        assert rec["status"] == "ok" and rec["steps"] == r["steps"], rec.get("error")
        assert rec["pcg_iters"] == its and rec["impulse_applied"] == imps, r["scene"]["name"]
        gx = serve.decode_array(rec["x"]).reshape(-1)
        gv = serve.decode_array(rec["v"]).reshape(-1)
        assert np.array_equal(gx, x) and np.array_equal(gv, v), r["scene"]["name"]


def _digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, np.float64).tobytes()).hexdigest()


@pytest.mark.gpu
def test_stacked_sheets_whole_step_against_golden_digest(sim_mod):
    """Four stacked 512^2 sheets (1,048,576 vertices, clearance 5 mm) over a
    bar and the ground: one whole StepTime, contact pass included, PCG capped
    at 3 iterations.  The oracle's step at this size takes minutes, so its
    result was computed once by tests/golden/make_fullsize.py and is held as
    SHA-256 digests of the positions and velocities plus the contact counts;
    the device's step must reproduce them exactly."""
    with open(os.path.join(HERE, "golden", "stack4x512_step.json")) as f:
        gold = json.load(f)
    cl = scene.load_config(os.path.join(ROOT, "cloth_sim_self_collision_amd", "configs",
                                        "stack4x512.json"))
    cl.pcg_maxit = gold["pcg_maxit"]
    with sim_mod.DeviceCloth(cl) as d:
        st = d.step(1)
        gx, gv = d.get_state()
    assert st["pcg_iters"] == gold["pcg_iters"] and st["impulse_applied"] == gold["impulse_applied"]
    assert st["n_prox"] == gold["n_prox"]
    assert _digest(gx) == gold["x_sha256"] and _digest(gv) == gold["v_sha256"]
