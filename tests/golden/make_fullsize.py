"""Writes tests/golden/stack4x512_step.json: the oracle's result for one whole
step (internal solve and contact pass) of the four stacked 512^2 sheets
(cloth_sim_self_collision_amd/configs/stack4x512.json, 1,048,576 vertices)
with PCG capped at 3 iterations, as SHA-256 digests of the positions and
velocities (float64 bytes) plus the step's counts.  The oracle needs minutes
for this step, so the GPU test compares against these digests instead of
running it.  Run from the repository root:

    python tests/golden/make_fullsize.py
"""
import hashlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402
import pyoracle  # noqa: E402

from cloth_sim_self_collision_amd import scene  # noqa: E402

PCG_MAXIT = 3


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, np.float64).tobytes()).hexdigest()


def main() -> int:
    cl = scene.load_config(os.path.join(ROOT, "cloth_sim_self_collision_amd", "configs",
                                        "stack4x512.json"))
    cl.pcg_maxit = PCG_MAXIT
    o = pyoracle.Oracle(cl, "port")
    x, v = cl.xyz.copy(), cl.uvw.copy()
    scene.apply_sweep(cl, x, cl.dt)
    t0 = time.time()
    n_prox = len(o.prox_set(x, cl.clearance))  # the step's proximity set: the start positions
    x1, v1, imp, it, conv = o.step(x, v)
    out = {"config": "stack4x512", "np": cl.np_, "pcg_maxit": PCG_MAXIT, "pcg_iters": int(it),
           "impulse_applied": bool(imp), "n_prox": int(n_prox),
           "x_sha256": digest(x1), "v_sha256": digest(v1),
           "oracle_seconds": round(time.time() - t0, 1)}
    with open(os.path.join(HERE, "stack4x512_step.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
