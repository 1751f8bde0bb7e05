"""The serving wire format (JSON lines) on the host: requests are validated,
arrays round-trip exactly, and requests that cannot be served get error
records without touching a device."""
import io
import json

import numpy as np
import pytest

from cloth_sim_self_collision_amd import scene, serve


def test_array_encoding_round_trips_exactly():
    rng = np.random.default_rng(3)
    a = rng.standard_normal((7, 3)) * np.exp2(rng.integers(-300, 300, (7, 3)))
    a[0, 0], a[1, 1] = -0.0, np.inf
    enc = serve.encode_array(a)
    assert enc["dtype"] == "<f8" and enc["shape"] == [7, 3]
    b = serve.decode_array(json.loads(json.dumps(enc)))
    assert a.tobytes() == b.tobytes()


def test_requests_are_parsed_and_validated():
    cfg = {"scene": "drop", "n": 12, "height": 0.3, "seed": 2,
           "obstacle": {"kind": "bar", "a": [-1, 0, 0], "b": [1, 0, 0], "radius": 0.05}}
    r = serve.parse_request({"id": "a-7", "scene": cfg, "steps": 5, "arrival": 0.25,
                             "return": ["x", "v"]})
    assert (r.rid, r.steps, r.arrival, r.returns) == ("a-7", 5, 0.25, ("x", "v"))
    want = scene.from_config(cfg)
    assert np.array_equal(r.cloth.xyz, want.xyz) and r.cloth.obstacle.kind == scene.OBST_BAR
    bad = [
        ({"steps": 3}, "scene"),
        ({"scene": cfg, "steps": 0}, "steps"),
        ({"scene": cfg, "steps": 2.5}, "steps"),
        ({"scene": cfg, "steps": 3, "precision": "bf16"}, "precision"),
        ({"scene": cfg, "steps": 3, "return": ["q"]}, "return"),
        ({"scene": cfg, "steps": 3, "arrival": -1}, "arrival"),
        ({"scene": {"scene": "batch", "count": 2}, "steps": 3}, "one cloth"),
        ({"scene": {"scene": "teapot"}, "steps": 3}, "teapot"),
    ]
    for obj, word in bad:
        with pytest.raises(ValueError, match=word):
            serve.parse_request(obj)


def test_unservable_requests_get_error_lines():
    lines = ['{"id": 1, "steps": 3}', "not json", "",
             '{"id": [2, "x"], "scene": {"scene": "drape", "n": 8}, "steps": 1, "precision": "fp8"}']
    out = io.StringIO()
    summary = serve.serve_jsonl(lines, out, hbm_budget_bytes=1 << 30)
    recs = [json.loads(s) for s in out.getvalue().splitlines()]
    assert [r["status"] for r in recs] == ["error"] * 3
    assert [r["id"] for r in recs] == [1, 1, [2, "x"]]  # a line that is not JSON: its line index
    assert "precision" in recs[2]["error"] and summary["results"] == []
