"""Seeded synthetic scenes: the inputs every solver path and the oracle consume.

The generators restate the reference's input side so that the synthetic
workloads have the reference's topology and parameter semantics:

* grid mesh and triangle order       -- Velvet/Scene.hpp:131-168 (GenerateClothMesh)
* stretch = structural + both shears -- Velvet/VtClothObjectGPU.hpp:75-116 (GenerateStretch)
* per-quad bending, rest angle 0     -- Velvet/VtClothObjectGPU.hpp:118-132 (GenerateBending)
* long-range attachments             -- Velvet/VtClothObjectGPU.hpp:134-148 (GenerateAttach)
* particle diameter = |v0 - v1| * 1.5 -- Velvet/VtClothObjectGPU.hpp:49
* solver defaults                     -- Velvet/Common.hpp:19-48 (VtSimParams)
* collider transform = T * Ry * Rz * Rx * S -- Velvet/Transform.hpp:22-29, Velvet/Helper.cpp:8-14

Vertex numbering follows the reference's ``VertexIndexAt(x, y) = x * (res + 1) + y``:
vertex ``k = a * n + b`` sits at row ``a`` and column ``b`` of an ``n x n`` grid.
Cloths are laid out horizontally (the reference rotates its vertical mesh by 90 degrees
about x in every cloth scene, e.g. Velvet/main.cpp:141).

All randomness comes from SplitMix64 over (seed, stream, counter), so the same seed gives the
same scene on every machine and numpy version.
"""
from __future__ import annotations

import functools
import math
from dataclasses import dataclass, field, replace
from typing import List, Optional, Sequence

import numpy as np

F32 = np.float32
FRAME_DT = F32(1.0) / F32(60.0)          # Velvet/Timer.hpp:235

SPHERE, PLANE, CUBE = 0, 1, 2             # Velvet/Common.hpp:113-118 (ColliderType order)

# ---------------------------------------------------------------------------------------------
# Counter-based random numbers
# ---------------------------------------------------------------------------------------------
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(z: np.ndarray) -> np.ndarray:
    z = z.astype(np.uint64, copy=True)
    with np.errstate(over="ignore"):
        z += np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, stream: int, count: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
    """``count`` float64 draws in [lo, hi) from SplitMix64(seed, stream, 0..count-1)."""
    base = np.uint64((seed * 0x100000001B3 + stream * 0x9E3779B1) & 0xFFFFFFFFFFFFFFFF)
    with np.errstate(over="ignore"):
        ctr = np.arange(count, dtype=np.uint64) + base * np.uint64(0x2545F4914F6CDD1D)
    u = (_splitmix64(ctr) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + u * (hi - lo)


# ---------------------------------------------------------------------------------------------
# Parameters
# ---------------------------------------------------------------------------------------------
@dataclass
class Params:
    """Solver knobs; field names and defaults follow VtSimParams (Velvet/Common.hpp:19-48).

    ``hash_table_size`` (0 = 2 x particles, Velvet/SpatialHashGPU.hpp:22),
    ``stretch_stiffness`` (a multiplier on the stretch lambda; 1.0 is the reference) and
    ``tune_hash_grid`` (1 lets the tuner shift the hash grid and enlarge the table) are
    extensions.
    """
    num_substeps: int = 2
    num_iterations: int = 4
    max_num_neighbors: int = 64
    max_speed: float = 50.0
    gravity: Sequence[float] = (0.0, -9.8, 0.0)
    bend_compliance: float = 0.0
    damping: float = 0.25
    relaxation_factor: float = 1.0
    long_range_stretchiness: float = 1.2
    collision_margin: float = 0.06
    friction: float = 0.1
    enable_self_collision: bool = True
    interleaved_hash: int = 3
    particle_diameter: float = 0.0
    particle_diameter_scalar: float = 1.5
    hash_cell_size_scalar: float = 1.5
    hash_table_size: int = 0
    stretch_stiffness: float = 1.0
    tune_hash_grid: int = 0


# ---------------------------------------------------------------------------------------------
# Colliders
# ---------------------------------------------------------------------------------------------
def _rot(axis: int, deg: float) -> np.ndarray:
    r = math.radians(deg)
    c, s = math.cos(r), math.sin(r)
    m = np.eye(4)
    i, j = [(1, 2), (2, 0), (0, 1)][axis]
    m[i, i] = c
    m[j, j] = c
    m[i, j] = -s
    m[j, i] = s
    return m


def transform_matrix(position, rotation_deg, scale) -> np.ndarray:
    """T * Ry * Rz * Rx * S as a float64 4x4 row-major (math) matrix.

    Velvet/Transform.hpp:22-29 with Velvet/Helper.cpp:8-14 (rotate y, then z, then x).
    """
    t = np.eye(4)
    t[:3, 3] = position
    s = np.diag([scale[0], scale[1], scale[2], 1.0])
    return t @ _rot(1, rotation_deg[1]) @ _rot(2, rotation_deg[2]) @ _rot(0, rotation_deg[0]) @ s


@dataclass
class Collider:
    """A collider as the solver sees it (Velvet/VtClothSolverGPU.cuh:8-18).

    ``cur``/``last`` are the current and previous frame's transforms
    (Velvet/Collider.hpp:33-41); a static collider has ``last == cur``.
    """
    type: int
    position: Sequence[float]
    scale: Sequence[float]
    rotation: Sequence[float] = (0.0, 0.0, 0.0)
    last_position: Optional[Sequence[float]] = None
    last_rotation: Optional[Sequence[float]] = None
    last_scale: Optional[Sequence[float]] = None

    def cur_matrix(self) -> np.ndarray:
        return transform_matrix(self.position, self.rotation, self.scale)

    def last_matrix(self) -> np.ndarray:
        return transform_matrix(
            self.last_position if self.last_position is not None else self.position,
            self.last_rotation if self.last_rotation is not None else self.rotation,
            self.last_scale if self.last_scale is not None else self.scale)


def sphere(center, radius) -> Collider:
    return Collider(SPHERE, tuple(center), (radius, radius, radius))


def plane(height=0.0) -> Collider:
    return Collider(PLANE, (0.0, height, 0.0), (1.0, 1.0, 1.0))


# ---------------------------------------------------------------------------------------------
# Cloth generation
# ---------------------------------------------------------------------------------------------
@dataclass
class Cloth:
    """One cloth's registration data, with indices local to the cloth."""
    x: np.ndarray                 # (n, 3) float32 world positions
    tris: np.ndarray              # (T, 3) uint32
    stretch: np.ndarray           # (S, 2) int32
    stretch_rest: np.ndarray      # (S,) float32
    bend: np.ndarray              # (B, 4) uint32
    bend_rest: np.ndarray         # (B,) float32
    slots: np.ndarray             # (L, 3) float32
    attach_pid: np.ndarray        # (A,) int32
    attach_slot: np.ndarray       # (A,) int32
    attach_dist: np.ndarray       # (A,) float32
    diameter: np.float32
    side: int = 0

    @property
    def n(self) -> int:
        return int(self.x.shape[0])


def _length_f32(d: np.ndarray) -> np.ndarray:
    """glm::length in binary32 without contraction: sqrt((x*x + y*y) + z*z)."""
    d = d.astype(F32)
    s = d[..., 0] * d[..., 0]
    s = s + d[..., 1] * d[..., 1]
    s = s + d[..., 2] * d[..., 2]
    return np.sqrt(s).astype(F32)


def grid_mesh(n: int):
    """Triangles of an n x n vertex grid in the reference's order (Velvet/Scene.hpp:153-165)."""
    r = n - 1
    a, b = np.meshgrid(np.arange(r), np.arange(r), indexing="ij")
    a = a.ravel()
    b = b.ravel()
    A = a * n + b
    Bv = (a + 1) * n + b
    C = a * n + (b + 1)
    D = (a + 1) * n + (b + 1)
    tris = np.empty((2 * r * r, 3), dtype=np.uint32)
    tris[0::2] = np.stack([A, Bv, C], axis=1)
    tris[1::2] = np.stack([C, Bv, D], axis=1)
    return tris


def grid_stretch(n: int):
    """Structural + shear pairs in the reference's loop order (Velvet/VtClothObjectGPU.hpp:75-116)."""
    r = n - 1
    out = []
    for x in range(n):
        for y in range(n):
            if y != r:
                out.append((x * n + y, x * n + y + 1))
            if x != r:
                out.append((x * n + y, (x + 1) * n + y))
            if y != r and x != r:
                out.append((x * n + y, (x + 1) * n + y + 1))
                out.append((x * n + y + 1, (x + 1) * n + y))
    return np.asarray(out, dtype=np.int32).reshape(-1, 2)


def _grid_stretch_fast(n: int) -> np.ndarray:
    """Vectorised grid_stretch (identical order) for large grids."""
    r = n - 1
    x, y = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    x = x.ravel()
    y = y.ravel()
    k = x * n + y
    cols = []
    # four candidate constraints per vertex, in the reference's order; mask out the missing ones
    c0 = np.stack([k, k + 1], 1)
    m0 = y != r
    c1 = np.stack([k, k + n], 1)
    m1 = x != r
    c2 = np.stack([k, k + n + 1], 1)
    m2 = (y != r) & (x != r)
    c3 = np.stack([k + 1, k + n], 1)
    m3 = m2
    allc = np.stack([c0, c1, c2, c3], 1)          # (n*n, 4, 2)
    allm = np.stack([m0, m1, m2, m3], 1)          # (n*n, 4)
    cols = allc[allm]
    return cols.astype(np.int32)


def bend_per_quad(tris: np.ndarray) -> np.ndarray:
    """Reference bending: one constraint per quad, (i0, i5, i2, i1) of each 6-index run.

    Velvet/VtClothObjectGPU.hpp:118-132.
    """
    flat = tris.reshape(-1)
    q = flat.reshape(-1, 6)
    return np.stack([q[:, 0], q[:, 5], q[:, 2], q[:, 1]], axis=1).astype(np.uint32)


def bend_pairs(tris: np.ndarray) -> np.ndarray:
    """One bending constraint per interior edge (adjacent-triangle pairs).

    Index order follows the reference's per-quad constraint: (o1, o2, e0, e1) with o1/o2 the
    vertices opposite the shared edge (e0, e1), such that for triangle (o1, e0, e1) the
    reference's n1 = (e0 - o1) x (e1 - o1) and n2 = (e1 - o2) x (e0 - o2) agree on a flat mesh
    (Velvet/VtClothSolverGPU.cu:150-151). Edges are visited in first-seen triangle order.
    """
    edges = {}
    order = []
    for t, (i, j, k) in enumerate(tris.tolist()):
        for (u, v, o) in ((i, j, k), (j, k, i), (k, i, j)):
            key = (u, v) if u < v else (v, u)
            if key in edges:
                edges[key].append((u, v, o))
            else:
                edges[key] = [(u, v, o)]
                order.append(key)
    out = []
    for key in order:
        lst = edges[key]
        if len(lst) != 2:
            continue
        (u, v, o1), (_, _, o2) = lst
        # triangle 1 visits the edge as u -> v, with winding (u, v, o1); the reference's n1 is
        # (e0 - o1) x (e1 - o1) = (v - o1) x (u - o1) for e0 = v, e1 = u ... choose e0/e1 so that
        # n1 has triangle 1's winding: (e0, e1, o1) cyclic == (u, v, o1) cyclic -> e0 = u, e1 = v
        # gives (u - o1) x (v - o1) = winding normal of (o1, u, v) == (u, v, o1).
        out.append((o1, o2, u, v))
    return np.asarray(out, dtype=np.uint32).reshape(-1, 4)


def grid_cloth(n: int, spacing: float, center=(0.0, 1.5, 0.0), jitter: float = 1e-4,
               seed: int = 0, stream: int = 0, attached: Sequence[int] = (),
               bend_mode: str = "pairs", diameter_scalar: float = 1.5) -> Cloth:
    """An n x n horizontal cloth; see the module docstring for the reference behaviour kept."""
    a, b = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    off = (n - 1) * spacing * 0.5
    x = np.zeros((n * n, 3), dtype=np.float64)
    x[:, 0] = center[0] - off + b.ravel() * spacing
    x[:, 1] = center[1]
    x[:, 2] = center[2] - off + a.ravel() * spacing
    if jitter:
        x += uniform(seed, stream, n * n * 3, -jitter, jitter).reshape(-1, 3)
    x = x.astype(F32)

    tris, st, bd = grid_topology(n, bend_mode)
    rest = _length_f32(x[st[:, 0]] - x[st[:, 1]])
    brest = np.zeros(bd.shape[0], dtype=F32)

    slots = np.zeros((len(attached), 3), dtype=F32)
    pid, sid, dist = [], [], []
    for s, p in enumerate(attached):
        slots[s] = x[p]
        pid.append(np.arange(n * n, dtype=np.int32))
        sid.append(np.full(n * n, s, dtype=np.int32))
        dist.append(_length_f32(slots[s][None, :] - x))
    if attached:
        pid = np.concatenate(pid)
        sid = np.concatenate(sid)
        dist = np.concatenate(dist).astype(F32)
    else:
        pid = np.zeros(0, np.int32)
        sid = np.zeros(0, np.int32)
        dist = np.zeros(0, F32)
    diam = F32(_length_f32(x[0:1] - x[1:2])[0] * F32(diameter_scalar))
    return Cloth(x=x, tris=tris, stretch=st, stretch_rest=rest, bend=bd, bend_rest=brest,
                 slots=slots, attach_pid=pid, attach_slot=sid, attach_dist=dist,
                 diameter=diam, side=n)


def mesh_constraints(x: np.ndarray, tris: np.ndarray):
    """Stretch (every unique edge) and bending (every edge of two triangles, rest angle from ``x``)
    of a triangle mesh, from the library's generator (``vtx_mesh_constraints``).
    Returns (stretch (S, 2) int32, stretch_rest (S,) float32, bend (B, 4) uint32, bend_rest (B,))."""
    import ctypes as C
    from ._lib import check, lib
    x = np.ascontiguousarray(x, dtype=F32).reshape(-1, 3)
    t = np.ascontiguousarray(tris, dtype=np.uint32).reshape(-1, 3)
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    up = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint32))
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
    ns, nb = C.c_uint32(), C.c_uint32()
    check(lib().vtx_mesh_constraints(fp(x), x.shape[0], up(t), t.shape[0], None, None, 0, C.byref(ns),
                                     None, None, 0, C.byref(nb)))
    st = np.empty((ns.value, 2), np.int32)
    sr = np.empty(ns.value, F32)
    bd = np.empty((nb.value, 4), np.uint32)
    br = np.empty(nb.value, F32)
    check(lib().vtx_mesh_constraints(fp(x), x.shape[0], up(t), t.shape[0], ip(st), fp(sr), ns.value,
                                     C.byref(ns), up(bd), fp(br), nb.value, C.byref(nb)))
    return st, sr, bd, br


def irregular_mesh(res: int, seed: int = 0, stream: int = 0):
    """GenerateClothMeshIrregular (Velvet/Scene.hpp:170-248): a (res+1)^2 grid of size 2.0 in the
    xy plane, interior vertices moved by U[0, 0.4/res)^2, each quad split along the diagonal the
    reference's angle test picks. Kept as in the reference: vertices are pushed y-major while
    triangles index x * (res+1) + y, and the test compares acos(dot) of unnormalised edge vectors.
    The reference draws its noise with rand(); the draws here are SplitMix64(seed, stream)."""
    m = res + 1
    noise = uniform(seed, stream, 2 * m * m, 0.0, 1.0).astype(F32).reshape(m, m, 2)   # [y, x]
    ns = F32(1.0 / res * 0.4)
    x = np.zeros((m * m, 3), dtype=F32)
    k = 0
    for yy in range(m):
        for xx in range(m):
            bnd = xx == 0 or yy == 0 or xx == res or yy == res
            nz = np.zeros(2, F32) if bnd else ns * noise[yy, xx]
            u = nz[0] + F32(xx) / F32(res)
            v = nz[1] + F32(yy) / F32(res)
            x[k] = F32(2.0) * np.array([u - F32(0.5), -v, 0.0], dtype=F32)
            k += 1

    def at(xx, yy):
        return xx * m + yy

    def angle(l, mid, r):
        a, b = (l - mid).astype(F32), (r - mid).astype(F32)
        return np.arccos(np.clip(F32(a[0] * b[0] + a[1] * b[1] + a[2] * b[2]), -1, 1), dtype=np.float64)

    tris = []
    for yy in range(res):
        for xx in range(res):
            p1, p2, p3, p4 = x[at(xx, yy)], x[at(xx + 1, yy)], x[at(xx, yy + 1)], x[at(xx + 1, yy + 1)]
            a1, a2, a3, a4 = angle(p3, p1, p2), angle(p1, p2, p4), angle(p1, p3, p4), angle(p3, p4, p2)
            if a1 + a4 > a2 + a3:
                tris += [(at(xx, yy), at(xx + 1, yy + 1), at(xx, yy + 1)),
                         (at(xx, yy), at(xx + 1, yy), at(xx + 1, yy + 1))]
            else:
                tris += [(at(xx, yy), at(xx + 1, yy), at(xx, yy + 1)),
                         (at(xx, yy + 1), at(xx + 1, yy), at(xx + 1, yy + 1))]
    return x, np.asarray(tris, dtype=np.uint32)


def mesh_cloth(x: np.ndarray, tris: np.ndarray, diameter: float, attached: Sequence[int] = ()) -> Cloth:
    """A cloth from any triangle mesh: generated stretch / bending constraints (rest lengths and
    rest angles of the given shape) and long-range attachments to ``attached`` vertices."""
    x = np.ascontiguousarray(x, dtype=F32).reshape(-1, 3)
    st, sr, bd, br = mesh_constraints(x, tris)
    n = x.shape[0]
    slots = np.zeros((len(attached), 3), dtype=F32)
    pid, sid, dist = [], [], []
    for s_, p in enumerate(attached):
        slots[s_] = x[p]
        pid.append(np.arange(n, dtype=np.int32))
        sid.append(np.full(n, s_, dtype=np.int32))
        dist.append(_length_f32(slots[s_][None, :] - x))
    cat = lambda a, t: np.concatenate(a).astype(t) if a else np.zeros(0, t)
    return Cloth(x=x, tris=np.ascontiguousarray(tris, dtype=np.uint32), stretch=st, stretch_rest=sr, bend=bd,
                 bend_rest=br, slots=slots, attach_pid=cat(pid, np.int32), attach_slot=cat(sid, np.int32),
                 attach_dist=cat(dist, F32), diameter=F32(diameter))


@functools.lru_cache(maxsize=64)
def grid_topology(n: int, bend_mode: str):
    """(triangles, stretch pairs, bends) of an n x n grid; shared read-only between cloths."""
    tris = grid_mesh(n)
    st = _grid_stretch_fast(n)
    if bend_mode == "quad":
        bd = bend_per_quad(tris)
    elif bend_mode == "pairs":
        bd = bend_pairs_grid(n)
    elif bend_mode == "none":
        bd = np.zeros((0, 4), dtype=np.uint32)
    else:
        raise ValueError(bend_mode)
    for a in (tris, st, bd):
        a.setflags(write=False)
    return tris, st, bd


def bend_pairs_grid(n: int) -> np.ndarray:
    """bend_pairs(grid_mesh(n)) computed in closed form (same constraints, same order)."""
    if n <= 40:
        return bend_pairs(grid_mesh(n))
    # generic path is O(T) python; for large grids build the same list vectorised by replaying
    # the edge discovery order: triangles are visited in order, each contributes edges
    # (i,j),(j,k),(k,i); an interior edge is emitted at its first visit with o1 from that
    # triangle and o2 from the other triangle sharing it.
    tris = grid_mesh(n).astype(np.int64)
    T = tris.shape[0]
    u = np.stack([tris[:, 0], tris[:, 1], tris[:, 2]], 1).reshape(-1)
    v = np.stack([tris[:, 1], tris[:, 2], tris[:, 0]], 1).reshape(-1)
    o = np.stack([tris[:, 2], tris[:, 0], tris[:, 1]], 1).reshape(-1)
    key = np.minimum(u, v) * (n * n) + np.maximum(u, v)
    order = np.argsort(key, kind="stable")
    ks = key[order]
    first = np.ones(ks.shape[0], dtype=bool)
    first[1:] = ks[1:] != ks[:-1]
    grp_start = np.nonzero(first)[0]
    grp_size = np.diff(np.append(grp_start, ks.shape[0]))
    interior = grp_size == 2
    e1 = order[grp_start[interior]]          # first visit (lower flat index = earlier triangle)
    e2 = order[grp_start[interior] + 1]
    # emit in first-seen order
    emit = np.argsort(e1, kind="stable")
    e1 = e1[emit]
    e2 = e2[emit]
    out = np.stack([o[e1], o[e2], u[e1], v[e1]], axis=1)
    return out.astype(np.uint32)


# ---------------------------------------------------------------------------------------------
# Scenes (one independent simulation each)
# ---------------------------------------------------------------------------------------------
@dataclass
class Scene:
    params: Params
    cloths: List[Cloth]
    colliders: List[Collider]
    frames: int = 1
    name: str = ""

    @property
    def n(self) -> int:
        return sum(c.n for c in self.cloths)

    def extent(self) -> float:
        x = np.concatenate([c.x for c in self.cloths])
        return float(np.max(x.max(0) - x.min(0)))

    def counts(self) -> dict:
        return dict(N=self.n,
                    S=sum(c.stretch.shape[0] for c in self.cloths),
                    B=sum(c.bend.shape[0] for c in self.cloths),
                    A=sum(c.attach_pid.shape[0] for c in self.cloths),
                    L=sum(c.slots.shape[0] for c in self.cloths),
                    T=sum(c.tris.shape[0] for c in self.cloths),
                    C=len(self.colliders))


def max_speed_for(diameter: float, substeps: int) -> float:
    """maxSpeed = 2 d / frame_dt * substeps in binary32 (Velvet/VtClothSolverGPU.hpp:125)."""
    return float(F32(F32(F32(2.0) * F32(diameter)) / FRAME_DT) * F32(substeps))


def finish_params(p: Params, cloths: Sequence[Cloth]) -> Params:
    """What AddCloth does to the global params: last cloth's diameter wins, maxSpeed from it."""
    d = float(cloths[-1].diameter)
    return replace(p, particle_diameter=d, max_speed=max_speed_for(d, p.num_substeps))


def config1(seed: int = 0, frames: int = 60, bend_mode: str = "pairs", jitter: float = 1e-4,
            **kw) -> Scene:
    """40x40 cloth (spacing 0.05) pinned at two top corners over a sphere (r 0.5) 1.0 below."""
    n = 40
    c = grid_cloth(n, 0.05, center=(0.0, 1.5, 0.0), jitter=jitter, seed=seed, stream=1,
                   attached=(0, n - 1), bend_mode=bend_mode)
    p = Params(num_substeps=5, num_iterations=10, relaxation_factor=1.5, **kw)
    return Scene(finish_params(p, [c]), [c], [sphere((0.0, 0.5, 0.0), 0.5)], frames, "config1")


def config2(seed: int = 0, frames: int = 300, bend_mode: str = "pairs") -> Scene:
    """256x256 cloth (spacing 0.01), one edge long-range attached, sphere r 0.6 + ground plane."""
    n = 256
    c = grid_cloth(n, 0.01, center=(0.0, 1.5, 0.0), jitter=1e-4, seed=seed, stream=2,
                   attached=tuple(range(n)), bend_mode=bend_mode)
    p = Params(num_substeps=5, num_iterations=10, hash_cell_size_scalar=1.0,
               hash_table_size=1 << 18)
    cols = [sphere((0.0, 0.6, 0.3), 0.6), plane(0.0)]
    return Scene(finish_params(p, [c]), [c], cols, frames, "config2")


def config3(seed: int = 0, frames: int = 120, n: int = 1024, bend_mode: str = "pairs") -> Scene:
    """1024x1024 cloth draped over three seeded spheres, radii U(0.2, 0.5); 4x8; 2^21 cells."""
    spacing = 2.0 / 1023.0
    c = grid_cloth(n, spacing, center=(0.0, 1.5, 0.0), jitter=2e-5, seed=seed, stream=3,
                   attached=(), bend_mode=bend_mode)
    r = uniform(seed, 31, 3, 0.2, 0.5)
    cx = uniform(seed, 32, 3, -0.6, 0.6)
    cz = uniform(seed, 33, 3, -0.6, 0.6)
    cy = uniform(seed, 34, 3, 0.3, 0.8)
    cols = [sphere((cx[i], cy[i], cz[i]), r[i]) for i in range(3)]
    p = Params(num_substeps=4, num_iterations=8, hash_table_size=1 << 21 if n == 1024 else 0)
    return Scene(finish_params(p, [c]), [c], cols, frames, "config3")


def batch_cloth_params(count: int = 512, seed: int = 0, smin: int = 16, smax: int = 128):
    sides = np.rint(np.exp(uniform(seed, 41, count, math.log(smin), math.log(smax)))).astype(int)
    spacing = uniform(seed, 42, count, 0.01, 0.05)
    stiff = uniform(seed, 43, count, 0.5, 1.0)
    nattach = np.minimum((uniform(seed, 44, count) * 5).astype(int), 4)
    nsph = 1 + np.minimum((uniform(seed, 45, count) * 3).astype(int), 2)
    return sides, spacing, stiff, nattach, nsph


def config4(seed: int = 0, frames: int = 120, count: int = 512, bend_mode: str = "pairs",
            smin: int = 16, smax: int = 128) -> List[Scene]:
    """512 independent cloths, sides log-uniform 16..128; one Scene (segment) per cloth."""
    sides, spacing, stiff, nattach, nsph = batch_cloth_params(count, seed, smin, smax)
    scenes = []
    for k in range(count):
        n = int(sides[k])
        corners = (0, n - 1, n * n - 1, n * (n - 1))[: int(nattach[k])]
        c = grid_cloth(n, float(spacing[k]), center=(0.0, 1.0, 0.0), jitter=1e-4, seed=seed,
                       stream=1000 + k, attached=corners, bend_mode=bend_mode)
        ext = (n - 1) * float(spacing[k])
        rr = uniform(seed, 2000 + k, 3 * 4)
        cols = []
        for s in range(int(nsph[k])):
            rad = ext * (0.15 + 0.2 * rr[4 * s])
            cx = ext * (rr[4 * s + 1] - 0.5) * 0.6
            cz = ext * (rr[4 * s + 2] - 0.5) * 0.6
            cy = 1.0 - rad - 0.05 - 0.3 * ext * rr[4 * s + 3]
            cols.append(sphere((cx, cy, cz), rad))
        p = Params(num_substeps=5, num_iterations=10, stretch_stiffness=float(F32(stiff[k])))
        scenes.append(Scene(finish_params(p, [c]), [c], cols, frames, f"config4[{k}]"))
    return scenes


def serving_job(seed: int, job: int, smin: int = 32, smax: int = 512, bend_mode: str = "pairs"):
    """One serving request: (Scene, horizon frames). Sizes log-uniform smin..smax."""
    u = uniform(seed, 5000 + job, 16)
    n = int(round(math.exp(math.log(smin) + u[0] * (math.log(smax) - math.log(smin)))))
    horizon = int(30 + u[1] * (600 - 30))
    spacing = 2.0 / (n - 1)
    nat = min(int(u[2] * 5), 4)
    corners = (0, n - 1, n * n - 1, n * (n - 1))[:nat]
    c = grid_cloth(n, spacing, center=(0.0, 1.5, 0.0), jitter=1e-4, seed=seed,
                   stream=6000 + job, attached=corners, bend_mode=bend_mode)
    ncol = 1 + min(int(u[3] * 4), 3)
    cols = []
    for s in range(ncol):
        if s == ncol - 1 and u[4] < 0.5:
            cols.append(plane(0.0))
        else:
            rad = 0.2 + 0.3 * u[5 + s]
            cols.append(sphere((0.8 * (u[8 + s] - 0.5), 0.5 + 0.3 * u[11 + s], 0.0), rad))
    p = Params(num_substeps=5, num_iterations=10)
    return Scene(finish_params(p, [c]), [c], cols, horizon, f"job{job}"), horizon


def algorithmic_bytes_per_iteration(scene_counts: dict) -> int:
    """Reference-layout bytes of one Jacobi iteration: 12S + 20B + 12A + 12L + 28N."""
    c = scene_counts
    return 12 * c["S"] + 20 * c["B"] + 12 * c["A"] + 12 * c["L"] + 28 * c["N"]


# ---------------------------------------------------------------------------------------------
# Flattening: one Scene -> the global arrays a single solver (or the oracle) sees
# ---------------------------------------------------------------------------------------------
def collider_matrices(col: Collider):
    """(cur, last) as binary32 column-major 4x4 (glm layout), flattened to 16 floats."""
    cur = col.cur_matrix().astype(F32)
    last = col.last_matrix().astype(F32)
    return cur.T.reshape(-1).copy(), last.T.reshape(-1).copy()


def flatten(scene: Scene) -> dict:
    """Concatenate the scene's cloths the way repeated AddCloth calls do.

    Particle indices are offset by the cloth's first particle (Velvet/VtClothSolverGPU.hpp:131-134)
    and slot indices by the slots registered before the cloth. Inverse masses start at 1 and an
    attachment with distance 0 pins its particle (Velvet/VtClothSolverGPU.hpp:172).
    """
    xs, tris, st, sr, bd, br, sl, ap, asl, ad = [], [], [], [], [], [], [], [], [], []
    off = 0
    soff = 0
    for c in scene.cloths:
        xs.append(c.x)
        tris.append(c.tris.astype(np.int64) + off)
        st.append(c.stretch.astype(np.int64) + off)
        sr.append(c.stretch_rest)
        bd.append(c.bend.astype(np.int64) + off)
        br.append(c.bend_rest)
        sl.append(c.slots)
        ap.append(c.attach_pid.astype(np.int64) + off)
        asl.append(c.attach_slot.astype(np.int64) + soff)
        ad.append(c.attach_dist)
        off += c.n
        soff += c.slots.shape[0]
    x = np.concatenate(xs).astype(F32)
    w = np.ones(x.shape[0], dtype=F32)
    ap = np.concatenate(ap).astype(np.int32)
    ad = np.concatenate(ad).astype(F32)
    w[ap[ad == 0]] = 0.0
    return dict(
        x=x, w=w,
        tris=np.concatenate(tris).astype(np.uint32).reshape(-1, 3),
        stretch=np.concatenate(st).astype(np.int32).reshape(-1, 2),
        stretch_rest=np.concatenate(sr).astype(F32),
        bend=np.concatenate(bd).astype(np.uint32).reshape(-1, 4),
        bend_rest=np.concatenate(br).astype(F32),
        slots=np.concatenate(sl).astype(F32).reshape(-1, 3),
        attach_pid=ap, attach_slot=np.concatenate(asl).astype(np.int32), attach_dist=ad)


def hash_cell_size(p: Params) -> np.float32:
    """cell = particleDiameter * hashCellSizeScalar in binary32 (Velvet/SpatialHashGPU.hpp:21)."""
    return F32(F32(p.particle_diameter) * F32(p.hash_cell_size_scalar))


def hash_table_size(p: Params, n: int) -> int:
    return int(p.hash_table_size) if p.hash_table_size > 0 else 2 * n
