"""Scene layer: meshes, bending stencils, obstacles, material and schedules.

Everything the step loop consumes is built here as plain numpy arrays
(float64 positions, int32 indices) so that the HIP library, the CPU oracle
and the compiled reference all receive byte-identical inputs.

Reference behaviour that this module restates:

* ``reference_square`` reproduces ``SetClothShape_Square``
  (self_contact_sparse/main.cpp:31-106): ix-major vertex numbering, two
  triangles per cell with one uniform diagonal, and the three families of
  bending quads in the demo's order.
* ``grid_cloth`` is the general grid builder used by the benchmark
  workloads: the same ix-major numbering, an optional alternating diagonal,
  and bending quads derived from interior-edge adjacency in the layout
  ``[opp0, opp1, e0, e1]`` that ``WdWddW_Bend`` expects
  (cloth_internal_physics.h:180-198: C[2], C[3] span the shared edge).
* mass per vertex = area * density / np (self_contact_sparse/main.cpp:568).
* obstacles follow the ``penetrationDepth(pd, n, p)`` convention
  (cloth_internal_physics.h:251-255): pd > 0 inside, n = push-out direction.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

# obstacle kinds understood by the C-ABI (kept in sync with include/clothsim.h)
OBST_NONE = 0
OBST_PLANE = 1    # pd = d - dot(n, p), normal n (unit)
OBST_SPHERE = 2   # pd = r - |p - c|, n = (p - c)/|p - c|
OBST_SPHERE_INV = 3  # self_contact_sparse/main.cpp:117-133 sign-flipped sphere
OBST_BOX = 4         # axis-aligned box (center, half-extents); not in the reference

_MASK64 = (1 << 64) - 1


class SplitMix64:
    """Repo-owned RNG so that every consumer sees identical jitter.

    SplitMix64 (Steele, Lea, Flood 2014); doubles are the top 53 bits / 2^53.
    """

    def __init__(self, seed: int):
        self.state = seed & _MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)

    def symmetric(self, a: float) -> float:
        """U(-a, a) as a*(2u-1)."""
        return a * (2.0 * self.uniform() - 1.0)


@dataclasses.dataclass
class Obstacle:
    kind: int = OBST_NONE
    params: tuple = ()  # plane: (nx,ny,nz,d); sphere: (cx,cy,cz,r)

    def as_array(self) -> np.ndarray:
        a = np.zeros(8, dtype=np.float64)
        a[0] = float(self.kind)
        for i, v in enumerate(self.params):
            a[1 + i] = float(v)
        return a

    def depth(self, p: np.ndarray):
        """Vectorised penetration depth and push-out normal (numpy, for tests)."""
        p = np.asarray(p, dtype=np.float64).reshape(-1, 3)
        if self.kind == OBST_NONE:
            return np.full(len(p), -100.0), np.zeros_like(p)
        if self.kind == OBST_PLANE:
            n = np.array(self.params[:3], dtype=np.float64)
            d = float(self.params[3])
            pd = d - (p[:, 0] * n[0] + p[:, 1] * n[1] + p[:, 2] * n[2])
            return pd, np.broadcast_to(n, p.shape).copy()
        if self.kind == OBST_BOX:
            s = p - np.array(self.params[:3], dtype=np.float64)
            d = np.abs(s) - np.array(self.params[3:6], dtype=np.float64)
            inside = (d <= 0.0).all(axis=1)
            m = np.argmax(d, axis=1)  # first axis on ties
            pd = np.empty(len(p))
            n = np.zeros_like(p)
            rows = np.arange(len(p))
            pd[inside] = -d[rows, m][inside]
            n[rows[inside], m[inside]] = np.where(s[rows, m] < 0.0, -1.0, 1.0)[inside]
            o = np.maximum(d, 0.0)
            ln = np.sqrt(o[:, 0] * o[:, 0] + o[:, 1] * o[:, 1] + o[:, 2] * o[:, 2])
            out = ~inside
            pd[out] = -ln[out]
            n[out] = (np.where(s < 0.0, -o, o) / np.where(ln > 0, ln, 1.0)[:, None])[out]
            return pd, n
        c = np.array(self.params[:3], dtype=np.float64)
        r = float(self.params[3])
        v = p - c
        ln = np.sqrt(v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1] + v[:, 2] * v[:, 2])
        n = v / ln[:, None]
        pd = r - ln
        if self.kind == OBST_SPHERE_INV:
            return -pd, -n
        return pd, n


def sphere(center, radius) -> Obstacle:
    return Obstacle(OBST_SPHERE, (center[0], center[1], center[2], radius))


def box(center, half) -> Obstacle:
    """Axis-aligned box; inside pd = distance to the nearest face."""
    return Obstacle(OBST_BOX, (center[0], center[1], center[2], half[0], half[1], half[2]))


def plane(normal, offset) -> Obstacle:
    """Half space dot(n,p) < offset is inside (pd = offset - dot(n,p))."""
    return Obstacle(OBST_PLANE, (normal[0], normal[1], normal[2], offset))


NO_OBSTACLE = Obstacle()


@dataclasses.dataclass
class Material:
    lam: float
    mu: float
    stiff_bend: float
    density: float = 1.0
    gravity: tuple = (0.0, 0.0, -10.0)
    stiff_contact: float = 1.0e3

    @staticmethod
    def from_young(E: float, nu: float, **kw) -> "Material":
        """3-D Lame parameters: lambda = E nu / ((1+nu)(1-2nu)), mu = E / (2(1+nu))."""
        lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
        mu = E / (2.0 * (1.0 + nu))
        return Material(lam=lam, mu=mu, **kw)


@dataclasses.dataclass
class Sweep:
    """Prescribed kinematic path for a set of constrained vertices.

    Before step k (time t_k -> t_k + dt) the listed vertices are placed at
    rest + offset(t_k + dt); offset moves linearly from 0 to ``delta`` over
    ``duration`` seconds and is held afterwards.  The vertices carry bc=1, so
    the implicit solve keeps them where they are placed.
    """
    vertices: np.ndarray
    delta: tuple
    duration: float

    def offset(self, t: float) -> np.ndarray:
        s = min(max(t / self.duration, 0.0), 1.0)
        return np.array(self.delta, dtype=np.float64) * s


@dataclasses.dataclass
class Cloth:
    """One cloth: mesh, rest/initial state, constraints and physics."""
    xyz0: np.ndarray          # rest positions (np*3) f64
    xyz: np.ndarray           # initial positions (np*3) f64
    uvw: np.ndarray           # initial velocities (np*3) f64
    tri: np.ndarray           # (nt*3) int32
    quad: np.ndarray          # (nq*4) int32
    bc: np.ndarray            # (np) int32, 1 = fixed
    area: float
    material: Material
    obstacle: Obstacle
    dt: float
    clearance: float
    pcg_tol: float = 1.0e-6
    pcg_maxit: int = 1000
    sweep: Optional[Sweep] = None
    name: str = ""

    @property
    def np_(self) -> int:
        return len(self.xyz0) // 3

    @property
    def nt(self) -> int:
        return len(self.tri) // 3

    @property
    def nq(self) -> int:
        return len(self.quad) // 4

    @property
    def mass_point(self) -> float:
        # self_contact_sparse/main.cpp:568
        return self.area * self.material.density / float(self.np_)

    def copy(self) -> "Cloth":
        return dataclasses.replace(self, xyz0=self.xyz0.copy(), xyz=self.xyz.copy(),
                                   uvw=self.uvw.copy(), tri=self.tri.copy(),
                                   quad=self.quad.copy(), bc=self.bc.copy())


# ---------------------------------------------------------------------------
# meshes

def bending_quads(tri: np.ndarray) -> np.ndarray:
    """Bending stencils from interior-edge adjacency.

    For every edge shared by two triangles, emits ``[opp0, opp1, e0, e1]``
    with opp0 the vertex of the lower-numbered triangle opposite the edge,
    opp1 the other opposite vertex and (e0, e1) the edge in the lower
    triangle's winding.  Order: by lower triangle index, then its local edge
    (edges (1,2), (2,0), (0,1)).
    """
    tri = np.asarray(tri, dtype=np.int64).reshape(-1, 3)
    e2n = ((1, 2), (2, 0), (0, 1))
    owner = {}
    quads: List[tuple] = []
    for it, t in enumerate(tri):
        for ie, (a, b) in enumerate(e2n):
            key = (min(t[a], t[b]), max(t[a], t[b]))
            owner.setdefault(key, []).append((it, ie))
    for it, t in enumerate(tri):
        for ie, (a, b) in enumerate(e2n):
            key = (min(t[a], t[b]), max(t[a], t[b]))
            lst = owner[key]
            if len(lst) != 2 or lst[0][0] != it:
                continue
            jt, je = lst[1]
            opp0 = int(t[ie])
            opp1 = int(tri[jt][je])
            quads.append((opp0, opp1, int(t[a]), int(t[b])))
    return np.asarray(quads, dtype=np.int32).reshape(-1)


def grid_cloth(nx: int, ny: int, size_x: float, size_y: float, *,
               center=(0.0, 0.0), z: float = 0.0, alternating: bool = True,
               jitter: float = 0.0, seed: int = 0):
    """Flat nx-by-ny vertex grid in the xy-plane at height z.

    Vertex (ix, iy) has index ix*ny + iy (ix-major, as main.cpp:47-53).
    Cell (ix, iy) is split along v10-v01 (the reference diagonal,
    main.cpp:58-68); with ``alternating`` the cells with (ix+iy) odd use
    v00-v11 instead.  z-jitter: SplitMix64(seed), U(-jitter, jitter), drawn in
    vertex order.
    Returns (xyz, tri, quad, area).
    """
    hx = size_x / (nx - 1)
    hy = size_y / (ny - 1)
    x0 = center[0] - 0.5 * size_x
    y0 = center[1] - 0.5 * size_y
    rng = SplitMix64(seed)
    xyz = np.empty((nx * ny, 3), dtype=np.float64)
    for ix in range(nx):
        for iy in range(ny):
            i = ix * ny + iy
            xyz[i, 0] = x0 + ix * hx
            xyz[i, 1] = y0 + iy * hy
            xyz[i, 2] = z
    if jitter > 0.0:
        for i in range(nx * ny):
            xyz[i, 2] = z + rng.symmetric(jitter)
    tri = np.empty((2 * (nx - 1) * (ny - 1), 3), dtype=np.int32)
    k = 0
    for ix in range(nx - 1):
        for iy in range(ny - 1):
            v00 = ix * ny + iy
            v10 = (ix + 1) * ny + iy
            v01 = ix * ny + iy + 1
            v11 = (ix + 1) * ny + iy + 1
            if alternating and ((ix + iy) & 1):
                tri[k] = (v00, v10, v11)
                tri[k + 1] = (v00, v11, v01)
            else:
                tri[k] = (v00, v10, v01)
                tri[k + 1] = (v11, v01, v10)
            k += 2
    tri = tri.reshape(-1)
    quad = bending_quads(tri)
    return xyz.reshape(-1), tri, quad, size_x * size_y


def reference_square(elem_length: float, size_x: float, size_z: float):
    """The demo mesh of self_contact_sparse/main.cpp:31-106, array for array.

    Vertices in the xz-plane (y = 0), ix-major; triangles with one uniform
    diagonal; quads emitted as the three loops at main.cpp:71-94.
    Returns (xyz, tri, quad, area).
    """
    ndx = int(size_x / elem_length)
    ndz = int(size_z / elem_length)
    xyz = []
    for ix in range(ndx + 1):
        for iz in range(ndz + 1):
            xyz += [ix * elem_length, 0.0, iz * elem_length]
    tri = []
    for ix in range(ndx):
        for iz in range(ndz):
            tri += [ix * (ndz + 1) + iz, (ix + 1) * (ndz + 1) + iz, ix * (ndz + 1) + iz + 1]
            tri += [(ix + 1) * (ndz + 1) + iz + 1, ix * (ndz + 1) + iz + 1, (ix + 1) * (ndz + 1) + iz]
    quad = []
    for ix in range(ndx):
        for iz in range(ndz):
            quad += [(ix + 0) * (ndz + 1) + iz, (ix + 1) * (ndz + 1) + iz + 1,
                     (ix + 1) * (ndz + 1) + iz, (ix + 0) * (ndz + 1) + iz + 1]
    for ix in range(ndx):
        for iz in range(ndz - 1):
            quad += [(ix + 1) * (ndz + 1) + iz, (ix + 0) * (ndz + 1) + iz + 2,
                     (ix + 1) * (ndz + 1) + iz + 1, (ix + 0) * (ndz + 1) + iz + 1]
    for ix in range(ndx - 1):
        for iz in range(ndz):
            quad += [(ix + 0) * (ndz + 1) + iz + 1, (ix + 2) * (ndz + 1) + iz,
                     (ix + 1) * (ndz + 1) + iz, (ix + 1) * (ndz + 1) + iz + 1]
    return (np.asarray(xyz, dtype=np.float64), np.asarray(tri, dtype=np.int32),
            np.asarray(quad, dtype=np.int32), size_x * size_z)


# ---------------------------------------------------------------------------
# workloads

def material_default() -> Material:
    """E=1e3, nu=0.3, bending 1e-3, density 1, gravity -10 z, k_contact 1e3."""
    return Material.from_young(1.0e3, 0.3, stiff_bend=1.0e-3, density=1.0,
                               gravity=(0.0, 0.0, -10.0), stiff_contact=1.0e3)


def drape_over_sphere(n: int = 32, *, seed: int = 0, jitter: float = 1.0e-3,
                      pcg_tol: float = 1.0e-6, pcg_maxit: int = 1000) -> Cloth:
    """Square 1 m cloth, n x n vertices, alternating diagonals, flat at z=0.5
    with U(+-jitter) z-jitter, dropped on a fixed sphere r=0.3 at the origin;
    dt=0.01, clearance 0.01 (the 32x32 case is the reference CPU workload)."""
    xyz, tri, quad, area = grid_cloth(n, n, 1.0, 1.0, z=0.5, alternating=True,
                                      jitter=jitter, seed=seed)
    np_ = len(xyz) // 3
    return Cloth(xyz0=xyz.copy(), xyz=xyz.copy(), uvw=np.zeros_like(xyz), tri=tri, quad=quad,
                 bc=np.zeros(np_, dtype=np.int32), area=area, material=material_default(),
                 obstacle=sphere((0.0, 0.0, 0.0), 0.3), dt=0.01, clearance=0.01,
                 pcg_tol=pcg_tol, pcg_maxit=pcg_maxit, name=f"drape{n}")


def drop_on_obstacle(n: int, height: float, obstacle: Obstacle, *, ny: Optional[int] = None,
                     seed: int = 0, jitter: float = 1.0e-3, dt: float = 0.01,
                     pcg_tol: float = 1.0e-6, pcg_maxit: int = 1000, name: str = "") -> Cloth:
    """Cloth of n x ny vertices (ny = n: square), 1 m along x with square
    cells, alternating diagonals, flat at z=height with U(+-jitter) z-jitter,
    falling onto `obstacle`; clearance 0.01."""
    ny = n if ny is None else ny
    xyz, tri, quad, area = grid_cloth(n, ny, 1.0, (ny - 1) / (n - 1), z=height, alternating=True,
                                      jitter=jitter, seed=seed)
    np_ = len(xyz) // 3
    return Cloth(xyz0=xyz.copy(), xyz=xyz.copy(), uvw=np.zeros_like(xyz), tri=tri, quad=quad,
                 bc=np.zeros(np_, dtype=np.int32), area=area, material=material_default(),
                 obstacle=obstacle, dt=dt, clearance=0.01, pcg_tol=pcg_tol, pcg_maxit=pcg_maxit,
                 name=name or f"drop{n}")


def batch_drop(count: int = 64, *, seed: int = 0, side_range=(24, 128),
               height_range=(0.3, 0.8), pcg_tol: float = 1.0e-6, pcg_maxit: int = 1000) -> list:
    """Independent cloths of seeded sizes dropped from seeded heights, each
    onto its own sphere (r 0.3) or box (half-extents 0.25) at the origin;
    dt 0.01.  Draw order per cloth: side, height, obstacle kind, jitter seed."""
    rng = SplitMix64(seed)
    out = []
    lo, hi = side_range
    for k in range(count):
        n = lo + int(rng.uniform() * (hi - lo + 1))
        h = height_range[0] + (height_range[1] - height_range[0]) * rng.uniform()
        ob = sphere((0.0, 0.0, 0.0), 0.3) if rng.uniform() < 0.5 else \
            box((0.0, 0.0, 0.0), (0.25, 0.25, 0.25))
        js = int(rng.next_u64() & 0x7FFFFFFF)
        out.append(drop_on_obstacle(n, h, ob, seed=js, pcg_tol=pcg_tol, pcg_maxit=pcg_maxit,
                                    name=f"drop{n}_{k}"))
    return out


def fold_sweep(n: int = 256, *, size: float = 2.0, seed: int = 0, jitter: float = 1.0e-3,
               pcg_tol: float = 1.0e-6, pcg_maxit: int = 1000) -> Cloth:
    """Single large cloth folded onto itself.

    n x n vertices over ``size`` metres in the xy-plane at z=0 with seeded
    z-jitter; the edge ix=0 is pinned, the edge ix=n-1 is swept by
    (-0.9*size, 0, +0.05*size) over 1 s so that the free half folds over the
    pinned half; no obstacle; dt=0.005, clearance 0.01.
    """
    xyz, tri, quad, area = grid_cloth(n, n, size, size, z=0.0, alternating=True,
                                      jitter=jitter, seed=seed)
    np_ = len(xyz) // 3
    bc = np.zeros(np_, dtype=np.int32)
    bc[0:n] = 1                      # ix = 0 line
    swept = np.arange((n - 1) * n, n * n, dtype=np.int32)  # ix = n-1 line
    bc[swept] = 1
    sw = Sweep(vertices=swept, delta=(-0.9 * size, 0.0, 0.05 * size), duration=1.0)
    return Cloth(xyz0=xyz.copy(), xyz=xyz.copy(), uvw=np.zeros_like(xyz), tri=tri, quad=quad,
                 bc=bc, area=area, material=material_default(), obstacle=NO_OBSTACLE,
                 dt=0.005, clearance=0.01, pcg_tol=pcg_tol, pcg_maxit=pcg_maxit,
                 sweep=sw, name=f"fold{n}")


def reference_demo(imode_contact: int = 0) -> Cloth:
    """The self_contact_sparse demo scene (main.cpp:110-156, 561-583):
    0.4 x 5.0 m strip, elem 0.1, lambda 0, mu 30, bend 1e-2, density 0.04,
    gravity (0, 0.05, -10), dt 0.01, k_contact 300, clearance 0.01; obstacle 0
    is the plane z = -0.5, obstacle 1 the inverted sphere of main.cpp:117-133.
    PCG settings are the hard-coded 1e-4 / 100 of solve_internal_sparse.h:318-319."""
    xyz, tri, quad, area = reference_square(0.1, 0.4, 5.0)
    np_ = len(xyz) // 3
    mat = Material(lam=0.0, mu=30.0, stiff_bend=1.0e-2, density=0.04,
                   gravity=(0.0, 0.05, -10.0), stiff_contact=300.0)
    if imode_contact == 0:
        obst = plane((0.0, 0.0, 1.0), -0.5)
    else:
        obst = Obstacle(OBST_SPHERE_INV, (0.5, 0.0, 2.0, 3.0))
    return Cloth(xyz0=xyz.copy(), xyz=xyz.copy(), uvw=np.zeros_like(xyz), tri=tri, quad=quad,
                 bc=np.zeros(np_, dtype=np.int32), area=area, material=mat, obstacle=obst,
                 dt=0.01, clearance=0.01, pcg_tol=1.0e-4, pcg_maxit=100, name="refdemo")


def apply_sweep(cloth: Cloth, xyz: np.ndarray, t_next: float) -> None:
    """Place swept vertices at rest + offset(t_next) (in place)."""
    if cloth.sweep is None:
        return
    off = cloth.sweep.offset(t_next)
    v = cloth.sweep.vertices
    x = xyz.reshape(-1, 3)
    x0 = cloth.xyz0.reshape(-1, 3)
    x[v] = x0[v] + off


__all__ = [
    "SplitMix64", "Obstacle", "Material", "Sweep", "Cloth", "sphere", "plane", "NO_OBSTACLE",
    "bending_quads", "grid_cloth", "reference_square", "material_default",
    "drape_over_sphere", "fold_sweep", "reference_demo", "apply_sweep",
    "OBST_NONE", "OBST_PLANE", "OBST_SPHERE", "OBST_SPHERE_INV", "OBST_BOX", "box",
    "drop_on_obstacle", "batch_drop",
]
