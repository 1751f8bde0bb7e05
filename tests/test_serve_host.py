"""The serving loop's host logic (admission by HBM budget, lanes, per-request
errors, cancel, evict and resume) with the device replaced by a stand-in
whose state after k steps is a known function of k, so that every result is
checkable without a GPU.  The device-backed server is tested in
tests/test_gpu_parity.py."""
import io
import json
import threading
import time

import numpy as np
import pytest

from cloth_sim_self_collision_amd import _lib, scene, serve


class FakeCloth:
    """Device stand-in: its state after k steps is (x0 + k, v0 - k)."""
    created = []
    fail_create = set()   # cloth names whose creation fails
    oom_once = set()      # cloth names whose first creation runs out of memory
    nonfinite_at = {}     # cloth name -> step index whose solve turns non-finite
    lock = threading.Lock()

    def __init__(self, cloth):
        name = cloth.name
        with FakeCloth.lock:
            if name in FakeCloth.fail_create:
                raise _lib.ClothSimError(1, "bad mesh")
            if name in FakeCloth.oom_once:
                FakeCloth.oom_once.discard(name)
                raise _lib.ClothSimError(3, "out of memory")
            FakeCloth.created.append(name)
        self.cloth = cloth
        self.np_ = cloth.np_
        self.k = 0
        self.time = 0.0
        self.closed = False

    def memory(self):
        return self.np_ * 4096

    def set_option(self, option, value):
        self.option = (option, value)
        if option == 3:
            self.precision = value

    nnz = 0

    def step(self, n):
        """The plain step of one cloth (solo requests)."""
        return FakeBatch([self]).step(n)[0]

    def get_state(self):
        return self.cloth.xyz + self.k, self.cloth.uvw - self.k

    def set_state(self, x, v):
        self.k = int(round(float(x[0] - self.cloth.xyz[0])))

    def close(self):
        self.closed = True


class FakeBatch:
    def __init__(self, devs):
        assert all(not d.closed for d in devs)
        # as cs_batch_create / cs_batch_step require
        if len({(d.cloth.pcg_tol, d.cloth.pcg_maxit, getattr(d, "precision", 0))
                for d in devs}) > 1:
            raise _lib.ClothSimError(1, "the cloths must share the PCG settings and precision")
        self.devs = list(devs)

    tick_s = 0.001   # a tick takes time, so that control actions land mid-run
    ticks = []       # (start, end, member names) of every tick

    def step(self, n):
        t0 = time.perf_counter()
        time.sleep(FakeBatch.tick_s)
        with FakeCloth.lock:
            FakeBatch.ticks.append((t0, time.perf_counter(), [d.cloth.name for d in self.devs]))
        out = []
        for d in self.devs:
            assert not d.closed
            bad = FakeCloth.nonfinite_at.get(d.cloth.name)
            st = {"status": 0, "pcg_iters": 3, "impulse_applied": False, "n_prox": 0,
                  "n_ccd_pass": 0, "n_ccd": [0] * 5}
            if bad is not None and d.k >= bad:
                st["status"] = 6
            else:
                d.k += 1
                d.time += d.cloth.dt
            out.append(st)
        return out

    def close(self):
        pass


@pytest.fixture
def fake(monkeypatch):
    monkeypatch.setattr(serve, "DeviceCloth", FakeCloth)
    monkeypatch.setattr(serve, "DeviceBatch", FakeBatch)
    FakeCloth.created = []
    FakeCloth.fail_create = set()
    FakeCloth.oom_once = set()
    FakeCloth.nonfinite_at = {}
    FakeBatch.tick_s = 0.001
    FakeBatch.ticks = []
    return FakeCloth


def _req(rid, n, steps, arrival=0.0, name=None):
    cl = scene.drape_over_sphere(n)
    cl.name = name or f"c{rid}"
    return serve.Request(rid, cl, steps=steps, arrival=arrival)


def _check(res, req):
    assert len(res.step_ms) == req.steps
    assert np.array_equal(res.xyz, req.cloth.xyz + req.steps)
    assert np.array_equal(res.uvw, req.cloth.uvw - req.steps)


def test_requests_in_several_lanes_all_finish(fake):
    reqs = [_req(0, 8, 5), _req(1, 60, 3, 0.01), _req(2, 20, 7, 0.02), _req(3, 160, 2, 0.03)]
    out = serve.Server(1 << 40, lanes=(500, 10000), solo_from=None).run(reqs, realtime=True)
    assert not out["failures"]
    by = {r.rid: r for r in out["results"]}
    for q in reqs:
        _check(by[q.rid], q)
    assert sum(1 for t in out["lanes"] if t > 0) == 3  # three lanes ticked


def test_solo_requests_step_alone_beside_the_lanes(fake):
    fake.nonfinite_at["c4"] = 1
    reqs = [_req(0, 8, 5), _req(1, 60, 4, 0.01), _req(2, 20, 7, 0.02), _req(3, 70, 3, 0.03),
            _req(4, 65, 3), _req(5, 80, 300)]
    state = {"cancelled": False}

    def control(srv, now):
        if {r.rid for r in srv.results} >= {0, 1, 2, 3} and not state["cancelled"]:
            srv.cancel(5)
            state["cancelled"] = True

    out = serve.Server(1 << 40, lanes=(500,), solo_from=3600).run(reqs, realtime=True,
                                                                 control=control)
    by = {r.rid: r for r in out["results"]}
    for q in reqs[:4]:
        _check(by[q.rid], q)
    fails = {f.rid: f.error for f in out["failures"]}
    assert set(fails) == {4, 5} and "NaN" in fails[4] and fails[5] == "cancelled"
    assert len(out["solo"]) == 4  # requests 1, 3, 4, 5 (>= 3600 vertices) ran alone


def test_mixed_precision_requests_fail_where_pcg_does_not_converge(fake, monkeypatch):
    """Mixed requests run in lanes of their own; one whose step stops at the
    PCG cap gets an error record, the others finish."""
    orig = FakeBatch.step
    members = []

    def step(self, n):
        members.append(tuple(sorted(d.cloth.name for d in self.devs)))
        out = orig(self, n)
        for d, st in zip(self.devs, out):
            st["pcg_conv"] = 0.5 if d.cloth.name == "stuck" else 1e-9
            st["pcg_iters"] = d.cloth.pcg_maxit if d.cloth.name == "stuck" else 7
        return out

    monkeypatch.setattr(FakeBatch, "step", step)
    reqs = [_req(0, 8, 3), _req(1, 8, 3, name="stuck"), _req(2, 9, 4)]
    reqs[1].precision = "mixed"
    reqs[2].precision = "mixed"
    out = serve.Server(1 << 40, lanes=()).run(reqs, realtime=False)
    fails = {f.rid: f.error for f in out["failures"]}
    assert set(fails) == {1} and "converges" in fails[1]
    by = {r.rid: r for r in out["results"]}
    _check(by[0], reqs[0])
    _check(by[2], reqs[2])
    # FP64 and mixed cloths never share a batch
    assert not any("c0" in m and ("stuck" in m or "c2" in m) for m in members)
    assert [c for c in FakeCloth.created]  # created through the pool


def test_requests_with_other_pcg_settings_get_their_own_lane(fake):
    reqs = [_req(k, 8 + k, 4) for k in range(4)]
    reqs[1].cloth.pcg_maxit = 50
    reqs[3].cloth.pcg_tol = 1e-4
    out = serve.Server(1 << 40, lanes=()).run(reqs, realtime=False)
    assert not out["failures"]
    for q in reqs:
        _check({r.rid: r for r in out["results"]}[q.rid], q)
    assert len(out["lanes"]) == 3


def test_variant_tuner_tries_both_then_keeps_the_faster():
    t = serve._VariantTuner(explore=4)
    seq = []
    for k in range(9):
        v = t.pick()
        seq.append(v)
        t.record(v, 1.0 if v == 0 else 3.0)  # the multi-launch variant is faster here
    assert seq[:2] == [2, 0]
    assert seq[2:] == [0, 2, 0, 0, 0, 2, 0]  # every 4th pick explores the other


def test_autotuned_lanes_give_the_same_results(fake):
    reqs = [_req(k, 8 + k, 6) for k in range(4)]
    out = serve.Server(1 << 40, lanes=(), autotune=True).run(reqs, realtime=False)
    assert not out["failures"]
    by = {r.rid: r for r in out["results"]}
    for q in reqs:
        _check(by[q.rid], q)


def test_budget_and_creation_failures_are_per_request(fake):
    fake.fail_create.add("broken")
    reqs = [_req(0, 8, 4), _req(1, 300, 2, name="huge"), _req(2, 8, 3, name="broken"),
            _req(3, 10, 2)]
    budget = 200 * 200 * serve.BYTES_PER_VERTEX_ESTIMATE  # request 1 (90,000 vertices) cannot fit
    out = serve.Server(budget).run(reqs, realtime=False)
    fails = {f.rid: f.error for f in out["failures"]}
    assert set(fails) == {1, 2}
    assert "budget" in fails[1] and "bad mesh" in fails[2]
    assert sorted(r.rid for r in out["results"]) == [0, 3]


def test_tight_budget_queues_and_out_of_memory_retries(fake):
    fake.oom_once.add("c1")
    reqs = [_req(k, 12, 3 + k) for k in range(4)]
    budget = 2 * 12 * 12 * serve.BYTES_PER_VERTEX_ESTIMATE + 1  # two resident cloths at a time
    out = serve.Server(budget, lanes=()).run(reqs, realtime=False)
    assert not out["failures"]
    for q in reqs:
        _check({r.rid: r for r in out["results"]}[q.rid], q)
    assert fake.created.count("c1") == 1  # the out-of-memory creation was retried


def test_nonfinite_step_fails_only_that_request(fake):
    fake.nonfinite_at["c1"] = 2
    reqs = [_req(0, 10, 5), _req(1, 10, 5), _req(2, 10, 5)]
    out = serve.Server(1 << 40, lanes=()).run(reqs, realtime=False)
    assert [f.rid for f in out["failures"]] == [1] and "NaN" in out["failures"][0].error
    assert sorted(r.rid for r in out["results"]) == [0, 2]


def test_cancel_and_evict_resume(fake):
    reqs = [_req(0, 10, 400), _req(1, 10, 300), _req(2, 10, 6)]
    state = {"evicted": False, "cancelled": False}

    def control(srv, now):
        done = {r.rid for r in srv.results}
        if 2 in done and not state["evicted"]:
            srv.evict(1)
            state["evicted"] = True
        if 2 in done and not state["cancelled"]:
            srv.cancel(0)
            state["cancelled"] = True

    out = serve.Server(1 << 40, lanes=()).run(reqs, realtime=False, control=control)
    assert [(f.rid, f.error) for f in out["failures"]] == [(0, "cancelled")]
    by = {r.rid: r for r in out["results"]}
    _check(by[1], reqs[1])  # evicted, admitted again, continued from its saved state
    _check(by[2], reqs[2])
    assert fake.created.count("c1") == 2


def test_wire_run_mixes_error_and_result_lines(fake):
    good = {"scene": "drape", "n": 8}
    lines = [json.dumps({"id": "a", "scene": good, "steps": 3}),
             json.dumps({"id": "big", "scene": {"scene": "drape", "n": 400}, "steps": 2}),
             json.dumps({"id": "b", "scene": good, "steps": 2, "return": ["x", "v"]})]
    out = io.StringIO()
    serve.serve_jsonl(lines, out, hbm_budget_bytes=64 * 64 * serve.BYTES_PER_VERTEX_ESTIMATE,
                      realtime=False)
    recs = {r["id"]: r for r in map(json.loads, out.getvalue().splitlines())}
    assert recs["big"]["status"] == "error" and "budget" in recs["big"]["error"]
    assert recs["a"]["status"] == "ok" and recs["a"]["steps"] == 3
    assert recs["b"]["status"] == "ok" and "v" in recs["b"]
    x = serve.decode_array(recs["b"]["x"]).reshape(-1)
    assert np.array_equal(x, scene.from_config(good).xyz + 2)


def test_poisson_requests_keep_the_long_side_on_the_fast_index():
    for r in serve.poisson_requests(40, 5.0, seed=1):
        sc = r["scene"]
        assert sc["ny"] >= sc["n"] and sc["side"] == 1.0
        nt = 2 * (sc["n"] - 1) * (sc["ny"] - 1)
        assert 700 <= nt <= 2.2e5


def _first_tick(name):
    return min((t for t in FakeBatch.ticks if name in t[2]), key=lambda t: t[0])


@pytest.mark.parametrize("instances", [1, 2])
def test_lane_instances_start_while_the_lane_steps(fake, instances):
    """A request arriving while its lane is in a tick waits for the next tick
    with one instance, and starts a second instance at once with two; every
    result is still the stand-in's known state."""
    FakeBatch.tick_s = 0.3
    reqs = [_req(0, 8, 2, 0.0, "a"), _req(1, 8, 2, 0.1, "b")]
    out = serve.Server(1 << 40, lanes=(), lane_instances=(instances,),
                       lane_window=(0.0,)).run(reqs, realtime=True)
    assert not out["failures"]
    by = {r.rid: r for r in out["results"]}
    for r in reqs:
        _check(by[r.rid], r)
    a0, b0 = _first_tick("a"), _first_tick("b")
    if instances == 2:
        assert b0[0] < a0[1] and b0[2] == ["b"] and a0[2] == ["a"]
    else:
        assert b0[0] >= a0[1]


@pytest.mark.parametrize("window", [0.0, 0.25])
def test_lane_window_gathers_the_first_tick(fake, window):
    """An idle lane with a window waits that long before its first tick, so a
    request arriving just after the first joins that tick."""
    FakeBatch.tick_s = 0.05
    reqs = [_req(0, 8, 1, 0.0, "a"), _req(1, 8, 1, 0.1, "b")]
    out = serve.Server(1 << 40, lanes=(), lane_instances=(1,),
                       lane_window=(window,)).run(reqs, realtime=True)
    assert not out["failures"]
    by = {r.rid: r for r in out["results"]}
    for r in reqs:
        _check(by[r.rid], r)
    first = min(FakeBatch.ticks, key=lambda t: t[0])
    assert sorted(first[2]) == (["a", "b"] if window > 0 else ["a"])


def test_lane_settings_are_checked(fake):
    with pytest.raises(ValueError):
        serve.Server(1 << 40, lanes=(1000,), lane_instances=(1,))
    with pytest.raises(ValueError):
        serve.Server(1 << 40, lanes=(), lane_instances=(0,))
    with pytest.raises(ValueError):
        serve.Server(1 << 40, lanes=(), lane_window=(-1.0,))
