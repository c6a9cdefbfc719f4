"""ctypes binding of velvet_amd/libvtx.so (the C ABI of include/vtx.h).

The library is the only compute path: if it is missing this module raises, there is no
fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvtx.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "vtx.h")

i32, u32, f32 = C.c_int32, C.c_uint32, C.c_float
P = C.POINTER


ABI_VERSION = 5   # VTX_ABI_VERSION of include/vtx.h for the structures below


class VtxParams(C.Structure):
    _fields_ = [("num_substeps", i32), ("num_iterations", i32), ("max_num_neighbors", i32),
                ("max_speed", f32), ("gravity", f32 * 3), ("bend_compliance", f32),
                ("damping", f32), ("relaxation_factor", f32), ("long_range_stretchiness", f32),
                ("collision_margin", f32), ("friction", f32), ("enable_self_collision", i32),
                ("interleaved_hash", i32), ("particle_diameter", f32),
                ("particle_diameter_scalar", f32), ("hash_cell_size_scalar", f32),
                ("hash_table_size", i32), ("stretch_stiffness", f32), ("tune_hash_grid", i32)]


class VtxCollider(C.Structure):
    _fields_ = [("type", i32), ("position", f32 * 3), ("scale", f32 * 3), ("cur", f32 * 16),
                ("last", f32 * 16), ("delta_time", f32)]


class VtxServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("hbm_reserve_bytes", C.c_uint64),
                ("max_live_particles", C.c_uint64), ("max_live_jobs", u32), ("use_graph", i32),
                ("check_every", i32), ("step_budget_ms", f32), ("reserve_particles", C.c_uint64),
                ("pipelined", i32)]


class VtxTuning(C.Structure):
    _fields_ = [("use_graph", i32), ("neighbor_order", i32), ("lra_batch", i32), ("xcd_remap", i32),
                ("frame_ms", f32), ("gather_block", i32), ("streams", i32), ("lra_split", i32),
                ("hash_lds", i32), ("hash_grid", i32)]


class VtxHashParams(C.Structure):
    _fields_ = [("num_objects", u32), ("max_num_neighbors", u32), ("cell_spacing", f32), ("cell_spacing2", f32),
                ("table_size", i32), ("particle_diameter2", f32)]


class VtxServerStats(C.Structure):
    _fields_ = [("live_jobs", u32), ("queued_jobs", u32), ("admitted", u32), ("completed", u32),
                ("failed", u32), ("pad", u32), ("live_particles", C.c_uint64),
                ("particle_substeps", C.c_uint64), ("step_ms", C.c_double), ("admit_ms", C.c_double),
                ("frame_ms", C.c_double), ("retire_ms", C.c_double)]


_SIGS = {
    "vtx_params_default": (None, [P(VtxParams)]),
    "vtx_last_error": (C.c_char_p, []),
    "vtx_version": (C.c_char_p, []),
    "vtx_abi_version": (i32, []),
    "vtx_solver_create": (i32, [P(VtxParams), P(C.c_void_p)]),
    "vtx_solver_destroy": (i32, [C.c_void_p]),
    "vtx_set_params": (i32, [C.c_void_p, i32, P(VtxParams)]),
    "vtx_get_params": (i32, [C.c_void_p, i32, P(VtxParams)]),
    "vtx_add_segment": (i32, [C.c_void_p, P(VtxParams), P(i32)]),
    "vtx_remove_segment": (i32, [C.c_void_p, i32]),
    "vtx_set_segment_active": (i32, [C.c_void_p, i32, i32]),
    "vtx_num_segments": (i32, [C.c_void_p, P(i32)]),
    "vtx_add_cloth": (i32, [C.c_void_p, i32, P(f32), u32, P(u32), u32, P(f32), f32, P(i32)]),
    "vtx_add_stretch": (i32, [C.c_void_p, i32, i32, i32, f32]),
    "vtx_add_stretch_n": (i32, [C.c_void_p, i32, P(i32), P(f32), u32]),
    "vtx_add_attach_slot": (i32, [C.c_void_p, i32, P(f32), P(i32)]),
    "vtx_set_attach_slot_position": (i32, [C.c_void_p, i32, i32, P(f32)]),
    "vtx_add_attach": (i32, [C.c_void_p, i32, i32, i32, f32]),
    "vtx_add_attach_n": (i32, [C.c_void_p, i32, P(i32), P(i32), P(f32), u32]),
    "vtx_add_bend": (i32, [C.c_void_p, i32, u32, u32, u32, u32, f32]),
    "vtx_add_bend_n": (i32, [C.c_void_p, i32, P(u32), P(f32), u32]),
    "vtx_set_colliders": (i32, [C.c_void_p, i32, P(VtxCollider), u32]),
    "vtx_simulate": (i32, [C.c_void_p, f32]),
    "vtx_synchronize": (i32, [C.c_void_p]),
    "vtx_set_use_graph": (i32, [C.c_void_p, i32]),
    "vtx_get_tuning": (i32, [C.c_void_p, P(VtxTuning)]),
    "vtx_mesh_constraints": (i32, [P(C.c_float), u32, P(u32), u32, P(i32), P(C.c_float), u32, P(u32),
                                   P(u32), P(C.c_float), u32, P(u32)]),
    "vtx_add_mesh_constraints": (i32, [C.c_void_p, i32, i32, P(u32), u32]),
    "vtx_snapshot_size": (i32, [C.c_void_p, i32, P(C.c_uint64)]),
    "vtx_readback_request": (i32, [C.c_void_p, i32]),
    "vtx_readback_latest": (i32, [C.c_void_p, i32, P(P(C.c_float)), P(P(C.c_float)), P(u32), P(C.c_uint64)]),
    "vtx_snapshot": (i32, [C.c_void_p, i32, C.c_void_p, C.c_uint64]),
    "vtx_restore": (i32, [C.c_void_p, i32, C.c_void_p, C.c_uint64]),
    "vtx_set_tuning": (i32, [C.c_void_p, P(VtxTuning)]),
    "vtx_autotune": (i32, [C.c_void_p, i32, f32, i32, P(VtxTuning)]),
    "vtx_set_stream": (i32, [C.c_void_p, C.c_void_p]),
    "vtx_set_stencil_path": (i32, [C.c_void_p, i32]),
    "vtx_stencil_segments": (i32, [C.c_void_p, P(i32)]),
    "vtx_num_particles": (i32, [C.c_void_p, i32, P(u32)]),
    "vtx_read_positions": (i32, [C.c_void_p, i32, P(f32)]),
    "vtx_read_velocities": (i32, [C.c_void_p, i32, P(f32)]),
    "vtx_read_normals": (i32, [C.c_void_p, i32, P(f32)]),
    "vtx_read_inv_masses": (i32, [C.c_void_p, i32, P(f32)]),
    "vtx_device_positions": (i32, [C.c_void_p, i32, P(C.c_void_p), P(u32)]),
    "vtx_device_normals": (i32, [C.c_void_p, i32, P(C.c_void_p), P(u32)]),
    "vtx_set_particle": (i32, [C.c_void_p, i32, i32, P(f32), P(f32), f32]),
    "vtx_write_predicted": (i32, [C.c_void_p, i32, P(f32)]),
    "vtx_run_hash": (i32, [C.c_void_p]),
    "vtx_hash_table_size": (i32, [C.c_void_p, i32, P(u32)]),
    "vtx_read_hash": (i32, [C.c_void_p, i32, P(u32), P(u32), P(u32), P(u32), P(u32)]),
    "vtx_hash_objects": (i32, [C.c_void_p] * 7 + [P(VtxHashParams), C.c_void_p]),
    "vtx_hash_objects_noalloc": (i32, [C.c_void_p] * 7 + [P(VtxHashParams), C.c_void_p]),
    "vtx_set_timing": (i32, [C.c_void_p, i32]),
    "vtx_read_timing": (i32, [C.c_void_p, C.c_char_p, C.c_size_t]),
    "vtx_server_create": (i32, [P(VtxServerConfig), P(C.c_void_p)]),
    "vtx_server_preempt": (i32, [C.c_void_p, i32]),
    "vtx_server_drain": (i32, [C.c_void_p]),
    "vtx_server_destroy": (i32, [C.c_void_p]),
    "vtx_server_solver": (i32, [C.c_void_p, P(C.c_void_p)]),
    "vtx_server_new_job": (i32, [C.c_void_p, P(VtxParams), i32, P(i32), P(i32)]),
    "vtx_server_submit": (i32, [C.c_void_p, i32]),
    "vtx_server_step": (i32, [C.c_void_p, f32, P(VtxServerStats)]),
    "vtx_server_job_status": (i32, [C.c_void_p, i32, P(i32), P(i32)]),
    "vtx_server_read_result": (i32, [C.c_void_p, i32, P(f32), u32]),
    "vtx_server_release_job": (i32, [C.c_void_p, i32]),
}

_LIB = None


class VtxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"vtx error {code}: {msg}")
        self.code = code


def declared_symbols():
    """Function names declared in include/vtx.h."""
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(vtx_\w+)\s*\(", txt, re.M)))


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python velvet_amd/build.py`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        if L.vtx_abi_version() != ABI_VERSION:   # the structures below follow this layout version
            raise RuntimeError(f"{LIB_PATH} has ABI version {L.vtx_abi_version()}, this binding expects "
                               f"{ABI_VERSION}: rebuild with `python velvet_amd/build.py`")
        _LIB = L
    return _LIB


def check(rc):
    if rc != 0:
        raise VtxError(rc, lib().vtx_last_error().decode())
    return rc
