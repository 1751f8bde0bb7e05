"""Writes the golden fixtures of tests/golden/ from the compiled reference.

    make -C oracle ref && python tests/golden/make_golden.py

The reference (nobuyuki83/cloth_sim_self_collision) is compiled from its own
sources by oracle/Makefile into oracle/_ref/libclothref.so; this script
drives it through oracle/pyoracle.py on scenes built by
cloth_sim_self_collision_amd.scene and stores inputs-derived outputs only
(numbers, no code).  The fixtures pin the CPU oracle where the reference
tree is absent.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyoracle  # noqa: E402
from cloth_sim_self_collision_amd import scene  # noqa: E402


def trajectory(o, cl, steps, keep):
    x, v, t = cl.xyz.copy(), cl.uvw.copy(), 0.0
    out = {}
    for k in range(steps):
        scene.apply_sweep(cl, x, t + cl.dt)
        if k in keep:
            out[k] = (x.copy(), v.copy())
        x, v, imp, it, conv = o.step(x, v)
        t += cl.dt
        out[("after", k + 1)] = (x.copy(), v.copy(), imp, it, conv)
    return out


def scene_fixture(name, cl, steps, probe_steps, traj_marks):
    o = pyoracle.Oracle(cl, "reference")
    f = {}
    rp, ci, di = o.pattern()
    f["rowptr"], f["colind"], f["diaind"] = rp, ci, di
    f["bvh"] = o.bvh()
    f["ring_ptr"], f["ring"] = o.edge()
    tr = trajectory(o, cl, steps, probe_steps)
    for k in probe_steps:
        x, v = tr[k]
        p = f"s{k}_"
        f[p + "x"], f[p + "v"] = x, v
        dia, crs, b = o.assemble(x, v)
        f[p + "dia"], f[p + "crs"], f[p + "b"] = dia, crs, b
        fd, fc = o.ilu(dia, crs)
        f[p + "fdia"], f[p + "fcrs"] = fd, fc
        xs, rs, conv, it = o.pcg(dia, crs, fd, fc, b, cl.pcg_tol, cl.pcg_maxit)
        f[p + "pcg_x"], f[p + "pcg_r"] = xs, rs
        f[p + "pcg_conv"], f[p + "pcg_iters"] = np.array(conv), np.array(it)
        f[p + "prox"] = o.prox_set(x, cl.clearance)
        f[p + "ccd"] = o.ccd_set(x, 3.0 * v, cl.dt, cl.clearance)
        f[p + "bb_prx"] = o.refit_prx(x, cl.clearance)
        vr, applied = o.contact_resolve(x, 3.0 * v)
        f[p + "resolved_v"], f[p + "resolved_applied"] = vr, np.array(applied)
    for k in traj_marks:
        x, v, imp, it, conv = tr[("after", k)]
        f[f"t{k}_x"], f[f"t{k}_v"] = x, v
        f[f"t{k}_imp"], f[f"t{k}_iters"], f[f"t{k}_conv"] = np.array(imp), np.array(it), np.array(conv)
    o.close()
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, **f)
    print(path, os.path.getsize(path))


def element_fixture():
    rng = np.random.default_rng(20131107)
    n = 48
    C = rng.standard_normal((n, 9))
    c = C + 0.2 * rng.standard_normal((n, 9))
    Cq = rng.standard_normal((n, 12))
    cq = Cq + 0.2 * rng.standard_normal((n, 12))
    f = {"cst_C": C, "cst_c": c, "bend_C": Cq, "bend_c": cq}
    cst = [pyoracle.element_cst(C[i], c[i], 576.923, 384.615, "reference") for i in range(n)]
    f["cst_W"] = np.array([r[0] for r in cst])
    f["cst_dW"] = np.array([r[1] for r in cst])
    f["cst_ddW"] = np.array([r[2] for r in cst])
    bend = [pyoracle.element_bend(Cq[i], cq[i], 1e-3, "reference") for i in range(n)]
    f["bend_dW"] = np.array([r[0] for r in bend])
    f["bend_ddW"] = np.array([r[1] for r in bend])
    s2 = rng.standard_normal((n, 3))
    f["psd_in"] = s2
    f["psd_out"] = np.array([pyoracle.psd22(s, "reference") for s in s2])
    S = rng.standard_normal((n, 12))
    E = S + 0.5 * rng.standard_normal((n, 12))
    f["cop_s"], f["cop_e"] = S, E
    f["cop_t"] = np.array([pyoracle.coplanar(S[i], E[i], "reference") for i in range(n)])
    P = rng.standard_normal((n, 12))
    fv = [pyoracle.dist_fv(P[i], "reference") for i in range(n)]
    ee = [pyoracle.dist_ee(P[i], "reference") for i in range(n)]
    f["dist_p"] = P
    f["dist_fv"] = np.array([r[0] for r in fv])
    f["dist_fv_w"] = np.array([r[1] for r in fv])
    f["dist_ee"] = np.array([r[0] for r in ee])
    f["dist_ee_w"] = np.array([r[1] for r in ee])
    pts = rng.standard_normal((n, 3)) * 0.4
    f["contact_p"] = pts
    for name, ob in (("sphere", scene.sphere((0.0, 0.0, 0.0), 0.3)),
                     ("plane", scene.plane((0.0, 0.0, 1.0), -0.1))):
        res = [pyoracle.element_contact(p, 1e3, 0.01, ob.as_array(), "reference") for p in pts]
        f[f"contact_{name}_W"] = np.array([r[0] for r in res])
        f[f"contact_{name}_dW"] = np.array([r[1] for r in res])
        f[f"contact_{name}_ddW"] = np.array([r[2] for r in res])
    path = os.path.join(HERE, "elements.npz")
    np.savez_compressed(path, **f)
    print(path, os.path.getsize(path))


if __name__ == "__main__":
    if not pyoracle.available("reference"):
        raise SystemExit("build the reference first: make -C oracle ref")
    element_fixture()
    scene_fixture("drape16", scene.drape_over_sphere(16), 60, (0, 45), (1, 10, 60))
    scene_fixture("refdemo0", scene.reference_demo(0), 30, (0, 29), (1, 10, 30))
    scene_fixture("fold20", scene.fold_sweep(20, size=0.25), 40, (20,), (1, 40))
