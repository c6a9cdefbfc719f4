"""Python face of the solver, mirroring VtClothSolverGPU's API (Velvet/VtClothSolverGPU.hpp).

``ClothSolver`` wraps one ``vtx_solver`` handle. Method names follow the reference
(``AddCloth``, ``AddStretch``, ``AddAttachSlot``, ``AddAttach``, ``AddBend``,
``UpdateColliders``, ``Simulate``) and accept an optional segment index for batching; segment
0 is the reference's single solver. ``load_scene`` registers a seeded scene the way
VtClothObjectGPU::Start does (Velvet/VtClothObjectGPU.hpp:43-58).
"""
from __future__ import annotations

import ctypes as C
from typing import Iterable, List, Optional, Sequence

import numpy as np

from . import scenes as S
from ._lib import VtxCollider, VtxParams, VtxTuning, check, lib

F32 = np.float32


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _up(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def to_vtx_params(p: S.Params) -> VtxParams:
    v = VtxParams()
    lib().vtx_params_default(C.byref(v))
    for name in ("num_substeps", "num_iterations", "max_num_neighbors", "interleaved_hash",
                 "hash_table_size", "tune_hash_grid"):
        setattr(v, name, int(getattr(p, name)))
    v.enable_self_collision = int(bool(p.enable_self_collision))
    for name in ("max_speed", "bend_compliance", "damping", "relaxation_factor",
                 "long_range_stretchiness", "collision_margin", "friction", "particle_diameter",
                 "particle_diameter_scalar", "hash_cell_size_scalar", "stretch_stiffness"):
        setattr(v, name, float(getattr(p, name)))
    for k in range(3):
        v.gravity[k] = float(p.gravity[k])
    return v


def to_vtx_colliders(cols: Sequence[S.Collider], frame_dt=None):
    arr = (VtxCollider * max(1, len(cols)))()
    for i, c in enumerate(cols):
        cur, last = S.collider_matrices(c)
        arr[i].type = int(c.type)
        for k in range(3):
            arr[i].position[k] = float(c.position[k])
            arr[i].scale[k] = float(c.scale[k])
        for k in range(16):
            arr[i].cur[k] = float(cur[k])
            arr[i].last[k] = float(last[k])
        arr[i].delta_time = float(S.FRAME_DT if frame_dt is None else frame_dt)
    return arr


class ClothSolver:
    """One vtx_solver handle; segments are independent simulations packed together."""

    def __init__(self, params: Optional[S.Params] = None, use_graph: bool = True, stencil: bool = True):
        h = C.c_void_p()
        vp = to_vtx_params(params or S.Params())
        check(lib().vtx_solver_create(C.byref(vp), C.byref(h)))
        self.h = h
        self.set_use_graph(use_graph)
        if not stencil:
            self.set_stencil_path(False)

    @classmethod
    def borrow(cls, handle):
        """Wrap a solver handle owned elsewhere (e.g. by a server); close() leaves it alone."""
        obj = cls.__new__(cls)
        obj.h = handle
        obj._borrowed = True
        return obj

    def close(self):
        if getattr(self, "h", None):
            if not getattr(self, "_borrowed", False):
                lib().vtx_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- segments / params
    def add_segment(self, params: S.Params) -> int:
        out = C.c_int32()
        check(lib().vtx_add_segment(self.h, C.byref(to_vtx_params(params)), C.byref(out)))
        return out.value

    def remove_segment(self, seg: int):
        check(lib().vtx_remove_segment(self.h, seg))

    def set_segment_active(self, seg: int, active: bool):
        check(lib().vtx_set_segment_active(self.h, seg, int(bool(active))))

    def num_segments(self) -> int:
        out = C.c_int32()
        check(lib().vtx_num_segments(self.h, C.byref(out)))
        return out.value

    def set_params(self, params: S.Params, seg: int = 0):
        check(lib().vtx_set_params(self.h, seg, C.byref(to_vtx_params(params))))

    def get_params(self, seg: int = 0) -> VtxParams:
        out = VtxParams()
        check(lib().vtx_get_params(self.h, seg, C.byref(out)))
        return out

    # -- registration, reference names (Velvet/VtClothSolverGPU.hpp:114-205)
    def AddCloth(self, xyz, tris, diameter, model=None, seg: int = 0) -> int:
        """Register a cloth (Velvet/VtClothSolverGPU.hpp:114-156); returns its first particle.

        ``model`` transforms the positions (InitializePositions, Velvet/VtClothSolverGPU.cu:30-40):
        a 4x4 array is a matrix in the usual math layout (translation in ``model[:3, 3]``) and is
        passed to the library column-major; a flat 16-vector is taken as already column-major
        (glm's memory layout). None is the identity.
        """
        x = np.ascontiguousarray(xyz, dtype=F32).reshape(-1)
        t = np.ascontiguousarray(tris, dtype=np.uint32).reshape(-1)
        m = None
        if model is not None:
            ma = np.asarray(model, dtype=F32)
            m = np.ascontiguousarray(ma.T if ma.shape == (4, 4) else ma, dtype=F32).reshape(16)
        off = C.c_int32()
        check(lib().vtx_add_cloth(self.h, seg, _fp(x), x.size // 3, _up(t), t.size // 3,
                                  _fp(m) if m is not None else None, float(diameter), C.byref(off)))
        return off.value

    def AddStretch(self, i, j, rest, seg: int = 0):
        check(lib().vtx_add_stretch(self.h, seg, int(i), int(j), float(rest)))

    def AddStretchN(self, ij, rest, seg: int = 0):
        ij = np.ascontiguousarray(ij, dtype=np.int32).reshape(-1)
        rest = np.ascontiguousarray(rest, dtype=F32).reshape(-1)
        check(lib().vtx_add_stretch_n(self.h, seg, _ip(ij), _fp(rest), rest.size))

    def add_mesh_constraints(self, offset: int, tris, seg: int = 0):
        """Stretch and bending constraints of the cloth added at ``offset``, generated from its
        current positions (rest angles included)."""
        t = np.ascontiguousarray(tris, dtype=np.uint32).reshape(-1)
        check(lib().vtx_add_mesh_constraints(self.h, seg, int(offset), t.ctypes.data_as(C.POINTER(C.c_uint32)),
                                             t.size // 3))

    def AddAttachSlot(self, pos, seg: int = 0) -> int:
        p = np.ascontiguousarray(pos, dtype=F32).reshape(3)
        out = C.c_int32()
        check(lib().vtx_add_attach_slot(self.h, seg, _fp(p), C.byref(out)))
        return out.value

    def SetAttachSlotPosition(self, slot, pos, seg: int = 0):
        p = np.ascontiguousarray(pos, dtype=F32).reshape(3)
        check(lib().vtx_set_attach_slot_position(self.h, seg, int(slot), _fp(p)))

    def AddAttach(self, particle, slot, dist, seg: int = 0):
        check(lib().vtx_add_attach(self.h, seg, int(particle), int(slot), float(dist)))

    def AddAttachN(self, pid, slot, dist, seg: int = 0):
        pid = np.ascontiguousarray(pid, dtype=np.int32)
        slot = np.ascontiguousarray(slot, dtype=np.int32)
        dist = np.ascontiguousarray(dist, dtype=F32)
        check(lib().vtx_add_attach_n(self.h, seg, _ip(pid), _ip(slot), _fp(dist), dist.size))

    def AddBend(self, i0, i1, i2, i3, angle, seg: int = 0):
        check(lib().vtx_add_bend(self.h, seg, int(i0), int(i1), int(i2), int(i3), float(angle)))

    def AddBendN(self, idx, angle, seg: int = 0):
        idx = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1)
        angle = np.ascontiguousarray(angle, dtype=F32)
        check(lib().vtx_add_bend_n(self.h, seg, _up(idx), _fp(angle), angle.size))

    def UpdateColliders(self, colliders: Sequence[S.Collider], seg: int = 0, frame_dt=None):
        arr = to_vtx_colliders(colliders, frame_dt)
        check(lib().vtx_set_colliders(self.h, seg, arr, len(colliders)))

    def Simulate(self, frame_dt=None):
        check(lib().vtx_simulate(self.h, float(S.FRAME_DT if frame_dt is None else frame_dt)))

    def synchronize(self):
        check(lib().vtx_synchronize(self.h))

    def set_use_graph(self, on: bool):
        check(lib().vtx_set_use_graph(self.h, int(bool(on))))

    # Launch variants (all give bit-identical results), the hash grid of segments that opted in
    # (tune_hash_grid), and the tuner that picks among them.
    _TUNING_KEYS = ("use_graph", "neighbor_order", "lra_batch", "xcd_remap", "gather_block", "streams", "lra_split",
                    "hash_lds", "hash_grid")

    def tuning(self) -> dict:
        t = VtxTuning()
        check(lib().vtx_get_tuning(self.h, C.byref(t)))
        return {k: int(getattr(t, k)) for k in self._TUNING_KEYS}

    def set_tuning(self, **kw):
        cur = self.tuning()
        unknown = set(kw) - set(cur)
        if unknown:
            raise KeyError(f"unknown tuning keys {sorted(unknown)}")
        cur.update({k: int(v) for k, v in kw.items()})
        t = VtxTuning(**cur, frame_ms=0.0)
        check(lib().vtx_set_tuning(self.h, C.byref(t)))

    def autotune(self, frames: int = 3, frame_dt: Optional[float] = None, use_cache: bool = False) -> dict:
        t = VtxTuning()
        check(lib().vtx_autotune(self.h, int(frames), float(S.FRAME_DT if frame_dt is None else frame_dt),
                                 int(bool(use_cache)), C.byref(t)))
        out = {k: int(getattr(t, k)) for k in self._TUNING_KEYS}
        out["frame_ms"] = float(t.frame_ms)
        return out

    def readback_request(self, seg: int = 0):
        check(lib().vtx_readback_request(self.h, seg))

    def readback_latest(self, seg: int = 0):
        """(positions (n, 4), normals (n, 4), frame) of the newest completed readback, copied out
        of the pinned buffer, or None."""
        p, q = C.POINTER(C.c_float)(), C.POINTER(C.c_float)()
        n, f = C.c_uint32(), C.c_uint64()
        check(lib().vtx_readback_latest(self.h, seg, C.byref(p), C.byref(q), C.byref(n), C.byref(f)))
        if not p:
            return None
        pos = np.ctypeslib.as_array(p, shape=(n.value, 4)).copy()
        nrm = np.ctypeslib.as_array(q, shape=(n.value, 4)).copy()
        return pos, nrm, int(f.value)

    def snapshot(self, seg: int = 0) -> bytes:
        """Positions (+ inverse masses), velocities, initial and slot positions of a segment."""
        n = C.c_uint64()
        check(lib().vtx_snapshot_size(self.h, seg, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        check(lib().vtx_snapshot(self.h, seg, buf, n.value))
        return buf.raw

    def restore(self, blob: bytes, seg: int = 0):
        buf = C.create_string_buffer(bytes(blob), len(blob))
        check(lib().vtx_restore(self.h, seg, buf, len(blob)))

    def set_stencil_path(self, on: bool):
        check(lib().vtx_set_stencil_path(self.h, int(bool(on))))

    def stencil_segments(self) -> int:
        out = C.c_int32()
        check(lib().vtx_stencil_segments(self.h, C.byref(out)))
        return out.value

    def set_stream(self, stream_ptr):
        check(lib().vtx_set_stream(self.h, C.c_void_p(stream_ptr) if stream_ptr else None))

    # -- state
    def num_particles(self, seg: int = 0) -> int:
        n = C.c_uint32()
        check(lib().vtx_num_particles(self.h, seg, C.byref(n)))
        return n.value

    def _read3(self, fn, seg):
        out = np.empty((self.num_particles(seg), 3), dtype=F32)
        check(fn(self.h, seg, _fp(out)))
        return out

    def positions(self, seg: int = 0) -> np.ndarray:
        return self._read3(lib().vtx_read_positions, seg)

    def velocities(self, seg: int = 0) -> np.ndarray:
        return self._read3(lib().vtx_read_velocities, seg)

    def normals(self, seg: int = 0) -> np.ndarray:
        return self._read3(lib().vtx_read_normals, seg)

    def inv_masses(self, seg: int = 0) -> np.ndarray:
        out = np.empty(self.num_particles(seg), dtype=F32)
        check(lib().vtx_read_inv_masses(self.h, seg, _fp(out)))
        return out

    def device_positions(self, seg: int = 0):
        ptr = C.c_void_p()
        n = C.c_uint32()
        check(lib().vtx_device_positions(self.h, seg, C.byref(ptr), C.byref(n)))
        return ptr.value, n.value

    def set_particle(self, i, pos=None, vel=None, inv_mass=-1.0, seg: int = 0):
        p = None if pos is None else np.ascontiguousarray(pos, dtype=F32)
        v = None if vel is None else np.ascontiguousarray(vel, dtype=F32)
        check(lib().vtx_set_particle(self.h, seg, int(i), _fp(p) if p is not None else None,
                                     _fp(v) if v is not None else None, float(inv_mass)))

    # -- spatial hash inspection
    def write_predicted(self, xyz, seg: int = 0):
        x = np.ascontiguousarray(xyz, dtype=F32).reshape(-1)
        check(lib().vtx_write_predicted(self.h, seg, _fp(x)))

    def run_hash(self):
        check(lib().vtx_run_hash(self.h))

    def read_hash(self, seg: int = 0):
        M = C.c_uint32()
        check(lib().vtx_hash_table_size(self.h, seg, C.byref(M)))
        n = self.num_particles(seg)
        K = self.get_params(seg).max_num_neighbors
        cs = np.empty(M.value, np.uint32)
        ce = np.empty(M.value, np.uint32)
        sh = np.empty(n, np.uint32)
        si = np.empty(n, np.uint32)
        nb = np.empty(K * n, np.uint32)
        check(lib().vtx_read_hash(self.h, seg, _up(cs), _up(ce), _up(sh), _up(si), _up(nb)))
        return cs, ce, sh, si, nb.reshape(K, n)

    def set_timing(self, on: bool):
        check(lib().vtx_set_timing(self.h, int(bool(on))))

    def timing(self) -> str:
        buf = C.create_string_buffer(4096)
        check(lib().vtx_read_timing(self.h, buf, len(buf)))
        return buf.value.decode()

    # -- scenes
    def load_scene(self, scene: S.Scene, seg: Optional[int] = None) -> int:
        """Register a scene into segment ``seg``, or into the next unused segment.

        Cloth by cloth, in VtClothObjectGPU::Start order: AddCloth, stretch, attach slots and
        attachments, bending (Velvet/VtClothObjectGPU.hpp:43-58); then the colliders.
        """
        if seg is None:
            if not getattr(self, "_seg0_used", False) and self.num_particles(0) == 0:
                seg = 0
            else:
                seg = self.add_segment(scene.params)
        if seg == 0:
            self._seg0_used = True
        self.set_params(scene.params, seg)
        slot_base = 0
        def shifted(a, k):    # index arrays offset by k (no copy for k = 0, the usual single cloth)
            return a if k == 0 else a.astype(np.int64) + k

        for c in scene.cloths:
            off = self.AddCloth(c.x, c.tris, c.diameter, seg=seg)
            self.AddStretchN(shifted(c.stretch, off), c.stretch_rest, seg=seg)
            for s in range(c.slots.shape[0]):
                self.AddAttachSlot(c.slots[s], seg=seg)
            if c.attach_pid.size:
                self.AddAttachN(shifted(c.attach_pid, off), shifted(c.attach_slot, slot_base), c.attach_dist, seg=seg)
            self.AddBendN(shifted(c.bend, off), c.bend_rest, seg=seg)
            slot_base += c.slots.shape[0]
        # AddCloth set diameter / max speed to the values scene.params already holds
        self.set_params(scene.params, seg)
        self.UpdateColliders(scene.colliders, seg=seg)
        return seg
