"""ctypes front-end to the CPU oracle (test infrastructure only).

Two back-ends share one API:

* ``Oracle(cloth, backend="port")``  -> oracle/liboracle.so, the C restatement
  of the reference's step (oracle/cloth_oracle.c);
* ``Oracle(cloth, backend="reference")`` -> oracle/_ref/libclothref.so, the
  reference's own sources compiled by oracle/Makefile (present only where the
  reference tree was available at build time).

Only tests, the smoke check and bench.py's CPU-baseline leg import this.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PORT_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libclothref.so")

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_libs = {}


def _dp(a):
    return a.ctypes.data_as(_D)


def _ip(a):
    return a.ctypes.data_as(_I)


def available(backend: str) -> bool:
    return os.path.exists(PORT_LIB if backend == "port" else REF_LIB)


def _load(backend: str):
    if backend in _libs:
        return _libs[backend]
    path = PORT_LIB if backend == "port" else REF_LIB
    if not os.path.exists(path):
        raise FileNotFoundError(f"oracle back-end '{backend}' not built: {path}")
    lib = C.CDLL(path)
    pre = "orc_" if backend == "port" else "ref_"

    def f(name, res, args):
        fn = getattr(lib, pre + name)
        fn.restype = res
        fn.argtypes = args
        return fn

    V = C.c_void_p
    api = {
        "create": f("create", V, [C.c_int, _D, _D, C.c_int, _I, C.c_int, _I, _I, _D, _D]),
        "destroy": f("destroy", None, [V]),
        "nnz": f("nnz", C.c_int, [V]),
        "nnode": f("nnode", C.c_int, [V]),
        "root": f("root", C.c_int, [V]),
        "nedge": f("nedge", C.c_int, [V]),
        "get_pattern": f("get_pattern", None, [V, _I, _I, _I]),
        "get_bvh": f("get_bvh", None, [V, _I]),
        "get_edge": f("get_edge", None, [V, _I, _I]),
        "assemble": f("assemble", None, [V, _D, _D, _D, _D, _D]),
        "ilu": f("ilu", None, [V, _D, _D, _D, _D]),
        "pcg": f("pcg", None, [V, _D, _D, _D, _D, _D, _D, C.c_double, C.c_int, _D, _I]),
        "ilu_solve": f("ilu_solve", None, [V, _D, _D, _D]),
        "matvec": f("matvec", None, [V, _D, _D, C.c_double, _D, C.c_double, _D]),
        "step_internal_tol": f("step_internal_tol", None, [V, _D, _D, C.c_double, C.c_int, _D, _I]),
        "refit_prx": f("refit_prx", None, [V, _D, C.c_double, _D]),
        "refit_ccd": f("refit_ccd", None, [V, _D, _D, C.c_double, _D]),
        "prox_set": f("prox_set", C.c_int, [V, _D, C.c_double, _I, C.c_int]),
        "ccd_set": f("ccd_set", C.c_int, [V, _D, _D, C.c_double, C.c_double, _I, C.c_int]),
        "contact_resolve": f("contact_resolve", C.c_int, [V, _D, _D]),
        "update_velocity": f("update_velocity", None, [V, _D, _D]),
        "step": f("step", None, [V, _D, _D, _I, _D]),
        "cst": f("cst", None, [_D, _D, C.c_double, C.c_double, _D, _D, _D]),
        "bend": f("bend", None, [_D, _D, C.c_double, _D, _D]),
        "contact_elem": f("contact_elem", None, [_D, C.c_double, C.c_double, _D, _D, _D, _D]),
        "psd22": f("psd22", None, [_D, _D]),
        "coplanar": f("coplanar", C.c_double, [_D, _D]),
        "dist_fv": f("dist_fv", C.c_double, [_D, _D]),
        "dist_ee": f("dist_ee", C.c_double, [_D, _D]),
    }
    if backend == "port":
        api["set_dot_mode"] = f("set_dot_mode", None, [V, C.c_int])
    if backend == "reference":
        api["step_internal_native"] = f("step_internal_native", None, [V, _D, _D])
    _libs[backend] = api
    return api


def cloth_params(cloth) -> np.ndarray:
    m = cloth.material
    return np.array([m.lam, m.mu, m.stiff_bend, m.gravity[0], m.gravity[1], m.gravity[2],
                     cloth.mass_point, m.stiff_contact, cloth.clearance, cloth.dt,
                     cloth.pcg_tol, float(cloth.pcg_maxit)], dtype=np.float64)


class Oracle:
    """One cloth simulated on the CPU by the restatement or the reference."""

    def __init__(self, cloth, backend: str = "port"):
        self.backend = backend
        self.api = _load(backend)
        self.cloth = cloth
        self.np_ = cloth.np_
        self._keep = [np.ascontiguousarray(cloth.xyz0, np.float64),
                      np.ascontiguousarray(cloth.xyz, np.float64),
                      np.ascontiguousarray(cloth.tri, np.int32),
                      np.ascontiguousarray(cloth.quad, np.int32),
                      np.ascontiguousarray(cloth.bc, np.int32),
                      cloth_params(cloth), _obst9(cloth.obstacle.as_array())]
        k = self._keep
        self.h = self.api["create"](self.np_, _dp(k[0]), _dp(k[1]), cloth.nt, _ip(k[2]), cloth.nq,
                                    _ip(k[3]), _ip(k[4]), _dp(k[5]), _dp(k[6]))
        self.nnz = self.api["nnz"](self.h)

    def set_dot_mode(self, mode: int):
        """PCG reduction order of the restatement: 0 sequential (the reference), 1 pairwise."""
        self.api["set_dot_mode"](self.h, mode)

    def close(self):
        if self.h:
            self.api["destroy"](self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- setup products
    def pattern(self):
        rp = np.zeros(self.np_ + 1, np.int32)
        ci = np.zeros(max(self.nnz, 1), np.int32)
        di = np.zeros(self.np_, np.int32)
        self.api["get_pattern"](self.h, _ip(rp), _ip(ci), _ip(di))
        return rp, ci[: self.nnz], di

    def bvh(self):
        n = self.api["nnode"](self.h)
        nodes = np.zeros(n * 3, np.int32)
        self.api["get_bvh"](self.h, _ip(nodes))
        return nodes.reshape(-1, 3)

    def edge(self):
        ne = self.api["nedge"](self.h)
        idx = np.zeros(self.np_ + 1, np.int32)
        arr = np.zeros(max(ne, 1), np.int32)
        self.api["get_edge"](self.h, _ip(idx), _ip(arr))
        return idx, arr[:ne]

    # -- stages
    def assemble(self, xyz, uvw):
        dia = np.zeros(self.np_ * 9)
        crs = np.zeros(max(self.nnz, 1) * 9)
        b = np.zeros(self.np_ * 3)
        self.api["assemble"](self.h, _dp(np.ascontiguousarray(xyz, np.float64)),
                             _dp(np.ascontiguousarray(uvw, np.float64)), _dp(dia), _dp(crs), _dp(b))
        return dia, crs[: self.nnz * 9], b

    def ilu(self, dia, crs):
        fd = np.zeros_like(dia)
        fc = np.zeros(max(len(crs), 9))
        self.api["ilu"](self.h, _dp(np.ascontiguousarray(dia)), _dp(_pad(crs)), _dp(fd), _dp(fc))
        return fd, fc[: len(crs)]

    def pcg(self, dia, crs, fd, fc, b, tol, maxit):
        r = np.array(b, dtype=np.float64)
        x = np.zeros_like(r)
        conv = np.zeros(1)
        it = np.zeros(1, np.int32)
        self.api["pcg"](self.h, _dp(np.ascontiguousarray(dia)), _dp(_pad(crs)),
                        _dp(np.ascontiguousarray(fd)), _dp(_pad(fc)), _dp(r), _dp(x), tol, maxit,
                        _dp(conv), _ip(it))
        return x, r, float(conv[0]), int(it[0])

    def ilu_solve(self, fd, fc, vec):
        v = np.array(vec, dtype=np.float64)
        self.api["ilu_solve"](self.h, _dp(np.ascontiguousarray(fd)), _dp(_pad(fc)), _dp(v))
        return v

    def matvec(self, dia, crs, alpha, x, beta, y):
        yy = np.array(y, dtype=np.float64)
        self.api["matvec"](self.h, _dp(np.ascontiguousarray(dia)), _dp(_pad(crs)), alpha,
                           _dp(np.ascontiguousarray(x, np.float64)), beta, _dp(yy))
        return yy

    def step_internal(self, xyz, uvw, tol=None, maxit=None):
        x = np.array(xyz, dtype=np.float64)
        v = np.array(uvw, dtype=np.float64)
        conv = np.zeros(1)
        it = np.zeros(1, np.int32)
        self.api["step_internal_tol"](self.h, _dp(x), _dp(v),
                                      self.cloth.pcg_tol if tol is None else tol,
                                      self.cloth.pcg_maxit if maxit is None else maxit,
                                      _dp(conv), _ip(it))
        return x, v, float(conv[0]), int(it[0])

    def refit_prx(self, xyz, delta):
        n = self.api["nnode"](self.h)
        bb = np.zeros(n * 6)
        self.api["refit_prx"](self.h, _dp(np.ascontiguousarray(xyz, np.float64)), delta, _dp(bb))
        return bb.reshape(-1, 6)

    def refit_ccd(self, xyz, uvw, dt):
        n = self.api["nnode"](self.h)
        bb = np.zeros(n * 6)
        self.api["refit_ccd"](self.h, _dp(np.ascontiguousarray(xyz, np.float64)),
                              _dp(np.ascontiguousarray(uvw, np.float64)), dt, _dp(bb))
        return bb.reshape(-1, 6)

    def _set(self, fn, *args):
        cap = 1 << 16
        while True:
            out = np.zeros(cap * 5, np.int32)
            n = fn(self.h, *args, _ip(out), cap)
            if n <= cap:
                return out[: n * 5].reshape(-1, 5)
            cap = n

    def prox_set(self, xyz, delta):
        x = np.ascontiguousarray(xyz, np.float64)
        return self._set(self.api["prox_set"], _dp(x), delta)

    def ccd_set(self, xyz, uvw, dt, delta):
        x = np.ascontiguousarray(xyz, np.float64)
        v = np.ascontiguousarray(uvw, np.float64)
        return self._set(self.api["ccd_set"], _dp(x), _dp(v), dt, delta)

    def contact_resolve(self, xyz1, uvw):
        v = np.array(uvw, dtype=np.float64)
        applied = self.api["contact_resolve"](self.h, _dp(np.ascontiguousarray(xyz1, np.float64)),
                                              _dp(v))
        return v, bool(applied)

    def update_velocity(self, uvw, xyz):
        v = np.array(uvw, dtype=np.float64)
        self.api["update_velocity"](self.h, _dp(v), _dp(np.ascontiguousarray(xyz, np.float64)))
        return v

    def step(self, xyz, uvw):
        """One StepTime; returns (xyz, uvw, impulse_applied, pcg_iters, conv)."""
        x = np.array(xyz, dtype=np.float64)
        v = np.array(uvw, dtype=np.float64)
        st = np.zeros(2, np.int32)
        conv = np.zeros(1)
        self.api["step"](self.h, _dp(x), _dp(v), _ip(st), _dp(conv))
        return x, v, bool(st[0]), int(st[1]), float(conv[0])


def _pad(a):
    a = np.ascontiguousarray(a, np.float64)
    if len(a) == 0:
        return np.zeros(9)
    return a


def element_cst(C_, c_, lam, mu, backend="port"):
    api = _load(backend)
    W = np.zeros(1)
    dW = np.zeros(9)
    ddW = np.zeros(81)
    api["cst"](_dp(np.ascontiguousarray(C_, np.float64)), _dp(np.ascontiguousarray(c_, np.float64)),
               lam, mu, _dp(W), _dp(dW), _dp(ddW))
    return float(W[0]), dW, ddW


def element_bend(C_, c_, stiff, backend="port"):
    api = _load(backend)
    dW = np.zeros(12)
    ddW = np.zeros(144)
    api["bend"](_dp(np.ascontiguousarray(C_, np.float64)), _dp(np.ascontiguousarray(c_, np.float64)),
                stiff, _dp(dW), _dp(ddW))
    return dW, ddW


def _obst9(a):
    """The 9-value obstacle record the C code reads (kind, 8 parameters)."""
    out = np.zeros(9, np.float64)
    a = np.asarray(a, np.float64).ravel()
    out[:min(9, len(a))] = a[:9]
    return out


def element_contact(c_, k, clearance, obst_array, backend="port"):
    api = _load(backend)
    W = np.zeros(1)
    dW = np.zeros(3)
    ddW = np.zeros(9)
    api["contact_elem"](_dp(np.ascontiguousarray(c_, np.float64)), k, clearance,
                        _dp(_obst9(obst_array)), _dp(W), _dp(dW), _dp(ddW))
    return float(W[0]), dW, ddW


def psd22(s2, backend="port"):
    api = _load(backend)
    out = np.zeros(3)
    api["psd22"](_dp(np.ascontiguousarray(s2, np.float64)), _dp(out))
    return out


def coplanar(s, e, backend="port"):
    api = _load(backend)
    return api["coplanar"](_dp(np.ascontiguousarray(s, np.float64)),
                           _dp(np.ascontiguousarray(e, np.float64)))


def dist_fv(p, backend="port"):
    api = _load(backend)
    w = np.zeros(2)
    d = api["dist_fv"](_dp(np.ascontiguousarray(p, np.float64)), _dp(w))
    return d, w


def dist_ee(p, backend="port"):
    api = _load(backend)
    w = np.zeros(2)
    d = api["dist_ee"](_dp(np.ascontiguousarray(p, np.float64)), _dp(w))
    return d, w
