"""Python front-end of one device-resident cloth.

``DeviceCloth`` mirrors the reference's step loop (StepTime,
self_contact_sparse/main.cpp:213-263) and the stage functions it calls:

=============================================  ===================================
reference                                      DeviceCloth
=============================================  ===================================
StepTime_InternalDynamicsILU                   step_internal_dynamics_ilu()
GetIntermidiateVelocityContactResolved         intermediate_velocity_contact_resolved()
x = x1 + v dt; UpdateIntermidiateVelocity      update_intermediate_velocity()
StepTime                                       step()
CMatrixSquareSparse::MatVec                    matvec()
CPreconditionerILU::SetValueILU+DoILUDecomp    ilu_factor()
CPreconditionerILU::Solve                      ilu_solve()
Solve_PCG                                      pcg_solve()
BuildBoundingBoxChild_Prx / _CCD               bvh_refit_prx() / bvh_refit_ccd()
GetContactElement_Proximity / _CCD             contacts_proximity() / contacts_ccd()
=============================================  ===================================

All compute runs in libclothsim.so on the GPU; this module only moves arrays.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib
from ._lib import check, dp, f64, i32, ip
from .scene import Cloth


def make_params(cloth: Cloth) -> _lib.Params:
    m = cloth.material
    p = _lib.Params()
    p.lam = m.lam
    p.myu = m.mu
    p.stiff_bend = m.stiff_bend
    for k in range(3):
        p.gravity[k] = m.gravity[k]
    p.mass_point = cloth.mass_point
    p.stiff_contact = m.stiff_contact
    p.contact_clearance = cloth.clearance
    p.dt = cloth.dt
    p.pcg_tol = cloth.pcg_tol
    p.pcg_maxit = cloth.pcg_maxit
    p.obstacle_kind = cloth.obstacle.kind
    for k, v in enumerate(cloth.obstacle.params[:8]):
        p.obstacle[k] = v
    return p


class DeviceCloth:
    """One cloth resident on the GPU."""

    def __init__(self, cloth: Cloth):
        lib = _lib.load()
        self.cloth = cloth
        self._arrays = [f64(cloth.xyz0), f64(cloth.xyz), f64(cloth.uvw), i32(cloth.tri),
                        i32(cloth.quad), i32(cloth.bc)]
        a = self._arrays
        mesh = _lib.Mesh(cloth.np_, cloth.nt, cloth.nq, dp(a[0]), dp(a[1]), dp(a[2]), ip(a[3]),
                         ip(a[4]), ip(a[5]))
        self._params = make_params(cloth)
        h = C.c_void_p()
        check(lib.cs_cloth_create(C.byref(mesh), C.byref(self._params), C.byref(h)))
        self.h = h
        sz = [C.c_int() for _ in range(5)]
        check(lib.cs_cloth_sizes(self.h, *[C.byref(s) for s in sz]))
        self.np_, self.nt, self.nq, self.nnz, self.nnode = [s.value for s in sz]
        if cloth.sweep is not None:
            verts = i32(cloth.sweep.vertices)
            delta = f64(cloth.sweep.delta)
            check(lib.cs_set_sweep(self.h, len(verts), ip(verts), dp(delta),
                                   float(cloth.sweep.duration)))

    def memory(self) -> int:
        """Device bytes held by this cloth (arena + contact buffers)."""
        b = C.c_longlong()
        check(_lib.load().cs_cloth_memory(self.h, C.byref(b)))
        return b.value

    # -- lifetime
    def close(self):
        if getattr(self, "h", None):
            _lib.load().cs_cloth_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- state
    def set_state(self, xyz=None, uvw=None):
        x = f64(xyz) if xyz is not None else None
        v = f64(uvw) if uvw is not None else None
        check(_lib.load().cs_set_state(self.h, dp(x) if x is not None else None,
                                       dp(v) if v is not None else None))

    def get_state(self):
        x = np.zeros(3 * self.np_)
        v = np.zeros(3 * self.np_)
        check(_lib.load().cs_get_state(self.h, dp(x), dp(v)))
        return x, v

    def set_step_start(self, xyz1):
        check(_lib.load().cs_set_step_start(self.h, dp(f64(xyz1))))

    def set_params(self, cloth: Optional[Cloth] = None):
        self._params = make_params(cloth or self.cloth)
        check(_lib.load().cs_cloth_set_params(self.h, C.byref(self._params)))

    def set_obstacle_values(self, pd=None, n=None):
        if pd is None:
            check(_lib.load().cs_set_obstacle_values(self.h, None, None))
        else:
            check(_lib.load().cs_set_obstacle_values(self.h, dp(f64(pd)), dp(f64(n))))

    # cs_set_option options (results are identical for every value of the first two)
    OPT_SWEEP = 1  # 0 automatic, 1 row-by-row CSR sweep
    OPT_PCG = 2    # 0 automatic, 1 one launch per PCG operation, 2 one-workgroup solve
    OPT_PRECISION = 3  # 0 FP64 (bit-exact), 1 mixed: FP32 sweep coefficients (not bit-exact)

    def set_option(self, option: int, value: int):
        check(_lib.load().cs_set_option(self.h, option, value))

    # -- the three stages of StepTime
    def step_internal_dynamics_ilu(self):
        conv = C.c_double()
        it = C.c_int()
        check(_lib.load().cs_step_internal(self.h, C.byref(conv), C.byref(it)))
        return conv.value, it.value

    def intermediate_velocity_contact_resolved(self):
        applied = C.c_int()
        st = _lib.StepStats()
        check(_lib.load().cs_contact_resolve(self.h, C.byref(applied), C.byref(st)))
        return bool(applied.value), st.as_dict()

    def update_intermediate_velocity(self):
        """x = x_start + v dt, then UpdateIntermidiateVelocity."""
        check(_lib.load().cs_post_contact_update(self.h))

    def update_velocity(self, uvw, xyz):
        """UpdateIntermidiateVelocity alone at positions xyz; returns v."""
        self.set_state(xyz, uvw)
        check(_lib.load().cs_update_velocity(self.h))
        return self.get_state()[1]

    @property
    def time(self) -> float:
        t = C.c_double()
        check(_lib.load().cs_get_time(self.h, C.byref(t)))
        return t.value

    @time.setter
    def time(self, t: float) -> None:
        check(_lib.load().cs_set_time(self.h, float(t)))

    def step(self, n: int = 1) -> dict:
        """n StepTime iterations on the device (the kinematic sweep, if any, is
        applied on the device before each step).  Returns the last step's stats."""
        st = _lib.StepStats()
        check(_lib.load().cs_step(self.h, n, C.byref(st)))
        return st.as_dict()

    # -- timing
    PROFILE_KINDS = ("ilu_sweep", "pcg_reduce", "spmv", "vector_update", "ilu_factor",
                     "assemble", "contact", "other", "pcg_workgroup")

    def profile(self, on: bool = True, sweep_only: bool = False) -> None:
        """Per-kernel-class HIP-event timing (cs_profile_enable): every class
        and the reduction counters, or with sweep_only the ILU sweeps alone."""
        mode = (2 if sweep_only else 1) if on else 0
        check(_lib.load().cs_profile_enable(self.h, mode))

    def reduction_stats(self) -> dict:
        out = (C.c_longlong * 8)()
        check(_lib.load().cs_reduction_stats(self.h, out))
        return dict(zip(("sums", "tiles", "fast_tiles", "walker_steps", "burst_elements",
                         "chain_ns", "walker_ns", "burst_ns"), out))

    def sweep_shape(self) -> dict:
        """Shape of the ILU sweeps (cs_sweep_shape): the longest component's
        forward/backward sub-levels, waves per workgroup, vector loads per wave
        per forward/backward sub-level, components, whether the packed quad
        sweep runs, and the wave-load instructions one launch issues."""
        out = (C.c_int * 8)()
        check(_lib.load().cs_sweep_shape(self.h, out))
        return dict(zip(("fwd_sublevels", "bwd_sublevels", "waves", "loads_fwd", "loads_bwd",
                         "components", "quad", "wave_loads"), out))

    def profile_read(self) -> dict:
        out = {}
        for k, name in enumerate(self.PROFILE_KINDS):
            ms = C.c_double()
            n = C.c_longlong()
            check(_lib.load().cs_profile_read(self.h, k, C.byref(ms), C.byref(n)))
            out[name] = (ms.value, n.value)
        return out

    # -- lower-level operations
    def pattern(self):
        rp = np.zeros(self.np_ + 1, np.int32)
        ci = np.zeros(max(self.nnz, 1), np.int32)
        di = np.zeros(self.np_, np.int32)
        check(_lib.load().cs_get_pattern(self.h, ip(rp), ip(ci), ip(di)))
        return rp, ci[: self.nnz], di

    def assemble(self, xyz, uvw):
        check(_lib.load().cs_assemble(self.h, dp(f64(xyz)), dp(f64(uvw))))
        return self.get_system()

    def get_system(self):
        dia = np.zeros(9 * self.np_)
        crs = np.zeros(9 * max(self.nnz, 1))
        b = np.zeros(3 * self.np_)
        check(_lib.load().cs_get_system(self.h, dp(dia), dp(crs), dp(b)))
        return dia, crs[: 9 * self.nnz], b

    def set_system(self, dia=None, crs=None, b=None):
        args = [dp(f64(a)) if a is not None and len(a) else None for a in (dia, crs, b)]
        check(_lib.load().cs_set_system(self.h, *args))

    def ilu_factor(self):
        ns = C.c_int()
        check(_lib.load().cs_ilu_factor(self.h, C.byref(ns)))
        return ns.value

    def get_factor(self):
        fd = np.zeros(9 * self.np_)
        fc = np.zeros(9 * max(self.nnz, 1))
        check(_lib.load().cs_get_factor(self.h, dp(fd), dp(fc)))
        return fd, fc[: 9 * self.nnz]

    def set_factor(self, fdia, fcrs):
        check(_lib.load().cs_set_factor(self.h, dp(f64(fdia)),
                                        dp(f64(fcrs)) if len(fcrs) else None))

    def ilu_solve(self, vec):
        v = np.array(vec, dtype=np.float64)
        check(_lib.load().cs_ilu_solve(self.h, dp(v)))
        return v

    def matvec(self, alpha, x, beta, y):
        yy = np.array(y, dtype=np.float64)
        check(_lib.load().cs_matvec(self.h, alpha, dp(f64(x)), beta, dp(yy)))
        return yy

    def pcg_solve(self, b, tol, maxit):
        x = np.zeros(3 * self.np_)
        r = np.zeros(3 * self.np_)
        conv = C.c_double()
        it = C.c_int()
        check(_lib.load().cs_pcg_solve(self.h, dp(f64(b)), tol, maxit, dp(x), dp(r),
                                       C.byref(conv), C.byref(it)))
        return x, r, conv.value, it.value

    def bvh(self):
        nodes = np.zeros(3 * self.nnode, np.int32)
        check(_lib.load().cs_get_bvh(self.h, ip(nodes)))
        return nodes.reshape(-1, 3)

    def bvh_refit_prx(self, xyz, delta):
        bb = np.zeros(6 * self.nnode)
        check(_lib.load().cs_bvh_refit_prx(self.h, dp(f64(xyz)), delta, dp(bb)))
        return bb.reshape(-1, 6)

    def bvh_refit_ccd(self, xyz, uvw, dt):
        bb = np.zeros(6 * self.nnode)
        check(_lib.load().cs_bvh_refit_ccd(self.h, dp(f64(xyz)), dp(f64(uvw)), dt, dp(bb)))
        return bb.reshape(-1, 6)

    def _contacts(self, fn, *args):
        cap = 1 << 16
        while True:
            out = np.zeros(5 * cap, np.int32)
            n = C.c_int()
            check(fn(self.h, *args, ip(out), cap, C.byref(n)))
            if n.value <= cap:
                return out[: 5 * n.value].reshape(-1, 5)
            cap = n.value

    def contacts_proximity(self, xyz, delta):
        x = f64(xyz)
        return self._contacts(_lib.load().cs_query_proximity, dp(x), delta)

    def contacts_ccd(self, xyz, uvw, dt, delta):
        x = f64(xyz)
        v = f64(uvw)
        return self._contacts(_lib.load().cs_query_ccd, dp(x), dp(v), dt, delta)


def element_cst(C_, c_, lam, mu):
    """WdWddW_CST on the device for arrays of triangles: (dW[n,9], ddW[n,81])."""
    C_ = f64(C_).reshape(-1)
    c_ = f64(c_).reshape(-1)
    n = len(C_) // 9
    dW = np.zeros(9 * n)
    ddW = np.zeros(81 * n)
    check(_lib.load().cs_element_cst(n, dp(C_), dp(c_), lam, mu, dp(dW), dp(ddW)))
    return dW.reshape(n, 9), ddW.reshape(n, 81)


def element_bend(C_, c_, stiff):
    C_ = f64(C_).reshape(-1)
    c_ = f64(c_).reshape(-1)
    n = len(C_) // 12
    dW = np.zeros(12 * n)
    ddW = np.zeros(144 * n)
    check(_lib.load().cs_element_bend(n, dp(C_), dp(c_), stiff, dp(dW), dp(ddW)))
    return dW.reshape(n, 12), ddW.reshape(n, 144)


def element_contact(p, stiff, clearance, obstacle):
    p = f64(p).reshape(-1)
    n = len(p) // 3
    o = f64(list(obstacle.params[:8]) + [0.0] * (8 - len(obstacle.params[:8])))
    dW = np.zeros(3 * n)
    ddW = np.zeros(9 * n)
    check(_lib.load().cs_element_contact(n, dp(p), stiff, clearance, obstacle.kind, dp(o),
                                         dp(dW), dp(ddW)))
    return dW.reshape(n, 3), ddW.reshape(n, 9)


def element_cst_energy(C_, c_, lam, mu):
    """The energy W of WdWddW_CST on the device, per triangle."""
    C_ = f64(C_).reshape(-1)
    c_ = f64(c_).reshape(-1)
    n = len(C_) // 9
    W = np.zeros(n)
    check(_lib.load().cs_element_cst_energy(n, dp(C_), dp(c_), lam, mu, dp(W)))
    return W


def element_contact_depth(pd, nrm, stiff, clearance):
    """WdWddW_Contact on the device from callback depths and directions:
    (W[n], dW[n,3], ddW[n,9])."""
    pd = f64(pd).reshape(-1)
    nrm = f64(nrm).reshape(-1)
    n = len(pd)
    W = np.zeros(n)
    dW = np.zeros(3 * n)
    ddW = np.zeros(9 * n)
    check(_lib.load().cs_element_contact_depth(n, dp(pd), dp(nrm), stiff, clearance, dp(W),
                                               dp(dW), dp(ddW)))
    return W, dW.reshape(n, 3), ddW.reshape(n, 9)


class DeviceBatch:
    """Several DeviceCloths stepped together by shared launches
    (cs_batch_*).  Each cloth's result equals stepping it alone."""

    def __init__(self, cloths):
        self.cloths = list(cloths)
        arr = (C.c_void_p * len(self.cloths))(*[c.h.value if isinstance(c.h, C.c_void_p) else c.h
                                                for c in self.cloths])
        h = C.c_void_p()
        check(_lib.load().cs_batch_create(len(self.cloths), arr, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            _lib.load().cs_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def step(self, n: int = 1) -> list:
        """n StepTime iterations of every cloth; per-cloth stats of the last step."""
        st = (_lib.StepStats * len(self.cloths))()
        check(_lib.load().cs_batch_step(self.h, n, C.cast(st, C.c_void_p)))
        return [s.as_dict() for s in st]

    def profile(self, on: bool = True) -> None:
        """Per-kernel-class HIP-event timing of the batch's shared launches."""
        check(_lib.load().cs_batch_profile_enable(self.h, 1 if on else 0))

    def profile_read(self) -> dict:
        """{class: (total ms, launches)} since profiling was enabled."""
        out = {}
        for k, name in enumerate(DeviceCloth.PROFILE_KINDS):
            ms = C.c_double()
            n = C.c_longlong()
            check(_lib.load().cs_batch_profile_read(self.h, k, C.byref(ms), C.byref(n)))
            out[name] = (ms.value, n.value)
        return out


def bandwidth_probe(n_doubles: int = 1 << 26, reps: int = 20) -> float:
    """Measured device copy bandwidth in GB/s (read + write bytes)."""
    g = C.c_double()
    check(_lib.load().cs_bandwidth_probe(int(n_doubles), int(reps), C.byref(g)))
    return g.value


def read_probe(n_doubles: int = 1 << 23, reps: int = 5, blocks: int = 1) -> float:
    """Measured read bandwidth in GB/s of `blocks` workgroups of 512 threads
    (1: one CU, the ceiling of the one-workgroup ILU sweep)."""
    unit = 16384 * blocks
    n = max(unit, int(n_doubles) // unit * unit)
    g = C.c_double()
    check(_lib.load().cs_read_probe(n, int(reps), int(blocks), C.byref(g)))
    return g.value


def issue_probe(waves: int = 8, steps: int = 3000) -> float:
    """ns per vector-load instruction of one wave in a sweep-shaped workgroup
    of `waves` waves (cs_issue_probe)."""
    g = C.c_double()
    check(_lib.load().cs_issue_probe(int(waves), int(steps), C.byref(g)))
    return g.value


def selftest_divsqrt(a, b):
    a = f64(a)
    b = f64(b)
    out = np.zeros(2 * len(a))
    check(_lib.load().cs_selftest_divsqrt(len(a), dp(a), dp(b), dp(out)))
    return out[0::2], out[1::2]


def ordered_sum(t) -> float:
    """The device's left-to-right reduction of t (the PCG dot-product sum)."""
    t = f64(t)
    out = np.zeros(1)
    check(_lib.load().cs_selftest_ordered_sum(len(t), dp(t) if len(t) else None, dp(out)))
    return float(out[0])
