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
from typing import Optional

import numpy as np

# obstacle kinds understood by the C-ABI (kept in sync with include/clothsim.h)
OBST_NONE = 0
OBST_PLANE = 1    # pd = d - dot(n, p), normal n (unit)
OBST_SPHERE = 2   # pd = r - |p - c|, n = (p - c)/|p - c|
OBST_SPHERE_INV = 3  # self_contact_sparse/main.cpp:117-133 sign-flipped sphere
OBST_BOX = 4         # axis-aligned box (center, half-extents); not in the reference
OBST_BAR = 5         # capsule around a segment (a, b, radius); not in the reference
OBST_BAR_GROUND = 6  # the bar and the ground z = h, the deeper of the two

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

    def symmetric_array(self, a: float, count: int) -> np.ndarray:
        """The next `count` values of symmetric(a), computed together (the
        same doubles: wrapping uint64 arithmetic, the same float operations)."""
        k = np.arange(1, count + 1, dtype=np.uint64)
        with np.errstate(over="ignore"):
            z = np.uint64(self.state) + k * np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
        self.state = (self.state + count * 0x9E3779B97F4A7C15) & _MASK64
        u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return a * (2.0 * u - 1.0)


@dataclasses.dataclass
class Obstacle:
    kind: int = OBST_NONE
    # plane: (nx,ny,nz,d); sphere: (cx,cy,cz,r); box: (cx,cy,cz,hx,hy,hz);
    # bar: (ax,ay,az,bx,by,bz,r); bar over ground: (ax,ay,az,bx,by,bz,r,h)
    params: tuple = ()

    def as_array(self) -> np.ndarray:
        """[kind, params..., 0 padding]: the oracle's obstacle record (9 values)."""
        a = np.zeros(9, dtype=np.float64)
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
        if self.kind in (OBST_BAR, OBST_BAR_GROUND):
            a = np.array(self.params[:3], dtype=np.float64)
            e = np.array(self.params[3:6], dtype=np.float64) - a
            w = p - a
            ee = e[0] * e[0] + e[1] * e[1] + e[2] * e[2]
            t = np.clip((w[:, 0] * e[0] + w[:, 1] * e[1] + w[:, 2] * e[2]) / ee, 0.0, 1.0)
            v = w - t[:, None] * e
            ln = np.sqrt(v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1] + v[:, 2] * v[:, 2])
            n = v / ln[:, None]
            pd = float(self.params[6]) - ln
            if self.kind == OBST_BAR_GROUND:
                pg = float(self.params[7]) - p[:, 2]
                g = pg > pd
                pd = np.where(g, pg, pd)
                n[g] = (0.0, 0.0, 1.0)
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


def bar(a, b, radius) -> Obstacle:
    """Capsule of `radius` around the segment a-b."""
    return Obstacle(OBST_BAR, (a[0], a[1], a[2], b[0], b[1], b[2], radius))


def bar_over_ground(a, b, radius, ground_z) -> Obstacle:
    """The bar and the ground plane z = ground_z (normal +z)."""
    return Obstacle(OBST_BAR_GROUND, (a[0], a[1], a[2], b[0], b[1], b[2], radius, ground_z))


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
    if len(tri) == 0:
        return np.zeros(0, dtype=np.int32)
    # every (triangle, local edge) in that order; an edge key occurs once per
    # triangle that has it, in order (a stable sort keeps that order)
    ie = np.tile(np.arange(3), len(tri))
    it = np.repeat(np.arange(len(tri)), 3)
    va = tri[it, np.array([1, 2, 0])[ie]]
    vb = tri[it, np.array([2, 0, 1])[ie]]
    key = np.minimum(va, vb) * (int(tri.max()) + 1) + np.maximum(va, vb)
    order = np.argsort(key, kind="stable")
    ks = key[order]
    start = np.flatnonzero(np.r_[True, ks[1:] != ks[:-1]])
    count = np.diff(np.r_[start, len(ks)])
    first = order[start[count == 2]]       # the lower triangle's (triangle, edge)
    second = order[start[count == 2] + 1]  # the other triangle's
    sel = np.argsort(first)                # by lower triangle, then its local edge
    first, second = first[sel], second[sel]
    quads = np.stack([tri[first // 3, first % 3], tri[second // 3, second % 3],
                      va[first], vb[first]], axis=1)
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
    ix = np.repeat(np.arange(nx), ny)
    iy = np.tile(np.arange(ny), nx)
    xyz = np.empty((nx * ny, 3), dtype=np.float64)
    xyz[:, 0] = x0 + ix * hx
    xyz[:, 1] = y0 + iy * hy
    xyz[:, 2] = z
    if jitter > 0.0:
        xyz[:, 2] = z + SplitMix64(seed).symmetric_array(jitter, nx * ny)
    cx = np.repeat(np.arange(nx - 1), ny - 1)
    cy = np.tile(np.arange(ny - 1), nx - 1)
    v00 = cx * ny + cy
    v10 = (cx + 1) * ny + cy
    v01 = cx * ny + cy + 1
    v11 = (cx + 1) * ny + cy + 1
    alt = ((cx + cy) & 1).astype(bool) if alternating else np.zeros(len(cx), dtype=bool)
    t = np.empty((len(cx), 2, 3), dtype=np.int32)
    t[:, 0] = np.where(alt[:, None], np.stack([v00, v10, v11], 1), np.stack([v00, v10, v01], 1))
    t[:, 1] = np.where(alt[:, None], np.stack([v00, v11, v01], 1), np.stack([v11, v01, v10], 1))
    tri = t.reshape(-1, 3)
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
                     pcg_tol: float = 1.0e-6, pcg_maxit: int = 1000, name: str = "",
                     side: Optional[float] = None) -> Cloth:
    """Cloth of n x ny vertices (ny = n: square) with square cells,
    alternating diagonals, flat at z=height with U(+-jitter) z-jitter, falling
    onto `obstacle`; clearance 0.01.  Size: 1 m along x, or with `side` the
    longer of the two sides that long."""
    ny = n if ny is None else ny
    h = (1.0 / (n - 1)) if side is None else side / (max(n, ny) - 1)
    xyz, tri, quad, area = grid_cloth(n, ny, h * (n - 1), h * (ny - 1), z=height,
                                      alternating=True, jitter=jitter, seed=seed)
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


def stacked_sheets(count: int = 4, n: int = 512, *, gap: float = 0.02, size: float = 1.0,
                   z0: float = 0.1, seed: int = 0, jitter: float = 2.0e-3, dt: float = 0.002,
                   clearance: float = 0.005, obstacle: Optional[Obstacle] = None,
                   pcg_tol: float = 1.0e-6, pcg_maxit: int = 1000, name: str = "") -> Cloth:
    """`count` n x n sheets of side `size` (alternating diagonals), sheet i
    flat at z0 + i*gap with U(+-jitter) z-jitter from SplitMix64(seed + i),
    as one cloth whose vertices and triangles are the sheets' in order (sheet
    on sheet contact is self-contact).  Default obstacle: a bar of radius 0.05
    along y under the stack's middle (axis at z = 0) over the ground z = -0.3.
    With count 4, n 512: 1,048,576 vertices, 2,088,968 triangles."""
    if obstacle is None:
        h = 0.75 * size
        obstacle = bar_over_ground((0.0, -h, 0.0), (0.0, h, 0.0), 0.05, -0.3)
    xs, ts, qs = [], [], []
    for i in range(count):
        xyz, tri, quad, _ = grid_cloth(n, n, size, size, z=z0 + i * gap, alternating=True,
                                       jitter=jitter, seed=seed + i)
        off = i * n * n
        xs.append(xyz)
        ts.append(tri + off)
        qs.append(quad + off)
    xyz = np.concatenate(xs)
    np_ = len(xyz) // 3
    return Cloth(xyz0=xyz.copy(), xyz=xyz.copy(), uvw=np.zeros_like(xyz),
                 tri=np.concatenate(ts).astype(np.int32), quad=np.concatenate(qs).astype(np.int32),
                 bc=np.zeros(np_, dtype=np.int32), area=count * size * size,
                 material=material_default(), obstacle=obstacle, dt=dt, clearance=clearance,
                 pcg_tol=pcg_tol, pcg_maxit=pcg_maxit, name=name or f"stack{count}x{n}")


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


# ---------------------------------------------------------------------------
# scene configuration files (JSON objects)

_OBSTACLE_KINDS = {
    "none": lambda spec: NO_OBSTACLE,
    "plane": lambda spec: plane(spec["normal"], spec["offset"]),
    "sphere": lambda spec: sphere(spec["center"], spec["radius"]),
    "box": lambda spec: box(spec["center"], spec["half"]),
    "bar": lambda spec: bar(spec["a"], spec["b"], spec["radius"]),
    "bar_over_ground": lambda spec: bar_over_ground(spec["a"], spec["b"], spec["radius"],
                                                    spec["ground_z"]),
}


def obstacle_from_config(spec) -> Obstacle:
    """{"kind": "sphere", "center": [0, 0, 0], "radius": 0.3} and the like:
    plane (normal, offset), sphere (center, radius), box (center, half), bar
    (a, b, radius), bar_over_ground (a, b, radius, ground_z), none."""
    if spec is None:
        return NO_OBSTACLE
    kind = spec.get("kind", "none")
    if kind not in _OBSTACLE_KINDS:
        raise ValueError(f"unknown obstacle kind {kind!r} (known: {sorted(_OBSTACLE_KINDS)})")
    return _OBSTACLE_KINDS[kind](spec)


def obstacle_to_config(ob: Obstacle) -> dict:
    p = [float(v) for v in ob.params]
    if ob.kind == OBST_NONE:
        return {"kind": "none"}
    if ob.kind == OBST_PLANE:
        return {"kind": "plane", "normal": p[:3], "offset": p[3]}
    if ob.kind == OBST_SPHERE:
        return {"kind": "sphere", "center": p[:3], "radius": p[3]}
    if ob.kind == OBST_BOX:
        return {"kind": "box", "center": p[:3], "half": p[3:6]}
    if ob.kind == OBST_BAR:
        return {"kind": "bar", "a": p[:3], "b": p[3:6], "radius": p[6]}
    if ob.kind == OBST_BAR_GROUND:
        return {"kind": "bar_over_ground", "a": p[:3], "b": p[3:6], "radius": p[6],
                "ground_z": p[7]}
    raise ValueError(f"obstacle kind {ob.kind} has no configuration form")


def _mesh_cloth(cfg) -> Cloth:
    xyz = np.asarray(cfg["xyz"], dtype=np.float64).reshape(-1)
    tri = np.asarray(cfg["tri"], dtype=np.int32).reshape(-1)
    np_ = len(xyz) // 3
    if len(xyz) != 3 * np_ or len(tri) % 3 or len(tri) == 0:
        raise ValueError("mesh: xyz needs 3 values per vertex and tri 3 indices per triangle")
    if tri.min() < 0 or tri.max() >= np_:
        raise ValueError("mesh: triangle index out of range")
    bc = np.zeros(np_, dtype=np.int32)
    for i in cfg.get("fixed", []):
        bc[int(i)] = 1
    if "area" in cfg:
        area = float(cfg["area"])
    else:  # total rest area of the triangles
        p = xyz.reshape(-1, 3)
        t = tri.reshape(-1, 3)
        area = float(0.5 * np.linalg.norm(np.cross(p[t[:, 1]] - p[t[:, 0]],
                                                   p[t[:, 2]] - p[t[:, 0]]), axis=1).sum())
    return Cloth(xyz0=xyz.copy(), xyz=xyz.copy(), uvw=np.zeros_like(xyz), tri=tri,
                 quad=bending_quads(tri), bc=bc, area=area, material=material_default(),
                 obstacle=obstacle_from_config(cfg.get("obstacle")), dt=float(cfg.get("dt", 0.01)),
                 clearance=float(cfg.get("clearance", 0.01)), name=cfg.get("name", "mesh"))


def _material_from_config(m: dict, base: Material) -> Material:
    kw = dict(lam=base.lam, mu=base.mu, stiff_bend=base.stiff_bend, density=base.density,
              gravity=base.gravity, stiff_contact=base.stiff_contact)
    if "young" in m:
        lm = Material.from_young(float(m["young"]), float(m.get("poisson", 0.3)), stiff_bend=0.0)
        kw["lam"], kw["mu"] = lm.lam, lm.mu
    for k in ("lam", "mu", "stiff_bend", "density", "stiff_contact"):
        if k in m:
            kw[k] = float(m[k])
    if "gravity" in m:
        kw["gravity"] = tuple(float(g) for g in m["gravity"])
    return Material(**kw)


def from_config(cfg: dict):
    """A cloth (or, for "batch", a list of cloths) from a JSON scene object.

    "scene": "drape" (n, seed, jitter), "fold" (n, size, seed, jitter),
    "drop" (n, ny, height, obstacle, seed, jitter, dt, side), "stack" (count, n,
    gap, size, z0, seed, jitter, dt, clearance, obstacle), "batch" (count,
    seed, side_range, height_range), "mesh" (xyz, tri, fixed vertex ids,
    area, obstacle, dt, clearance) for any triangle mesh (bending stencils
    from its interior edges).  Optional for all: pcg_tol, pcg_maxit,
    material {lam, mu (or young, poisson), stiff_bend, density, gravity,
    stiff_contact}, name.
    """
    kind = cfg.get("scene")
    opt = {k: cfg[k] for k in ("seed", "jitter") if k in cfg}
    if kind == "drape":
        out = drape_over_sphere(int(cfg.get("n", 32)), **opt)
    elif kind == "fold":
        out = fold_sweep(int(cfg.get("n", 256)), size=float(cfg.get("size", 2.0)), **opt)
    elif kind == "drop":
        out = drop_on_obstacle(int(cfg["n"]), float(cfg.get("height", 0.5)),
                               obstacle_from_config(cfg.get("obstacle",
                                                            {"kind": "sphere", "center": [0, 0, 0],
                                                             "radius": 0.3})),
                               ny=cfg.get("ny"), dt=float(cfg.get("dt", 0.01)),
                               side=(float(cfg["side"]) if "side" in cfg else None), **opt)
    elif kind == "stack":
        extra = {k: float(cfg[k]) for k in ("gap", "size", "z0", "dt", "clearance") if k in cfg}
        ob = obstacle_from_config(cfg["obstacle"]) if "obstacle" in cfg else None
        out = stacked_sheets(int(cfg.get("count", 4)), int(cfg.get("n", 512)), obstacle=ob,
                             **extra, **opt)
    elif kind == "batch":
        out = batch_drop(int(cfg.get("count", 64)), seed=int(cfg.get("seed", 0)),
                         side_range=tuple(cfg.get("side_range", (24, 128))),
                         height_range=tuple(cfg.get("height_range", (0.3, 0.8))))
    elif kind == "mesh":
        out = _mesh_cloth(cfg)
    else:
        raise ValueError(f"unknown scene {kind!r} (drape, fold, drop, stack, batch, mesh)")
    cloths = out if isinstance(out, list) else [out]
    for cl in cloths:
        if "pcg_tol" in cfg:
            cl.pcg_tol = float(cfg["pcg_tol"])
        if "pcg_maxit" in cfg:
            cl.pcg_maxit = int(cfg["pcg_maxit"])
        if "material" in cfg:
            cl.material = _material_from_config(cfg["material"], cl.material)
        if "name" in cfg and not isinstance(out, list):
            cl.name = str(cfg["name"])
    return out


def load_config(path):
    """from_config of a JSON file."""
    import json
    with open(path) as f:
        return from_config(json.load(f))


__all__ = [
    "SplitMix64", "Obstacle", "Material", "Sweep", "Cloth", "sphere", "plane", "NO_OBSTACLE",
    "bending_quads", "grid_cloth", "reference_square", "material_default",
    "drape_over_sphere", "fold_sweep", "reference_demo", "apply_sweep",
    "OBST_NONE", "OBST_PLANE", "OBST_SPHERE", "OBST_SPHERE_INV", "OBST_BOX", "OBST_BAR",
    "OBST_BAR_GROUND", "box", "bar", "bar_over_ground",
    "drop_on_obstacle", "batch_drop", "stacked_sheets", "from_config", "load_config",
    "obstacle_from_config", "obstacle_to_config",
]
