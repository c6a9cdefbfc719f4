"""Continuous-batching server (vtx_server_* of include/vtx.h) from Python.

A job is one scene (an independent simulation) with a frame horizon. ``submit(scene, frames)``
registers the scene on a pending segment of the server's solver and queues it; ``step()``
admits queued jobs while the device budget and the live-particle cap allow, advances every live
job one frame in the same launches and retires finished jobs.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import scenes as S
from ._lib import VtxServerConfig, VtxServerStats, check, lib
from .solver import ClothSolver, to_vtx_params

QUEUED, RUNNING, DONE, FAILED, REJECTED = 1, 2, 3, 4, 5
STATUS = {0: "pending", 1: "queued", 2: "running", 3: "done", 4: "failed", 5: "rejected"}


class Server:
    def __init__(self, hbm_budget_bytes: int = 0, hbm_reserve_bytes: int = 4 << 30,
                 max_live_particles: int = 0, max_live_jobs: int = 0, use_graph: bool = False,
                 check_every: int = 0, step_budget_ms: float = 0.0, reserve_particles: int = 0,
                 pipelined: bool = False):
        cfg = VtxServerConfig(hbm_budget_bytes=hbm_budget_bytes, hbm_reserve_bytes=hbm_reserve_bytes,
                              max_live_particles=max_live_particles, max_live_jobs=max_live_jobs,
                              use_graph=int(bool(use_graph)), check_every=check_every,
                              step_budget_ms=step_budget_ms, reserve_particles=reserve_particles,
                              pipelined=int(bool(pipelined)))
        h = C.c_void_p()
        check(lib().vtx_server_create(C.byref(cfg), C.byref(h)))
        self.h = h
        sh = C.c_void_p()
        check(lib().vtx_server_solver(self.h, C.byref(sh)))
        self.solver = ClothSolver.borrow(sh)
        self.n = {}

    def close(self):
        if getattr(self, "h", None):
            lib().vtx_server_destroy(self.h)
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

    def submit(self, scene: S.Scene, frames: Optional[int] = None) -> int:
        frames = scene.frames if frames is None else frames
        job, seg = C.c_int32(), C.c_int32()
        check(lib().vtx_server_new_job(self.h, C.byref(to_vtx_params(scene.params)), int(frames),
                                       C.byref(job), C.byref(seg)))
        self.solver.load_scene(scene, seg=seg.value)
        check(lib().vtx_server_submit(self.h, job.value))
        self.n[job.value] = scene.n
        return job.value

    def step(self, frame_dt=None) -> VtxServerStats:
        st = VtxServerStats()
        check(lib().vtx_server_step(self.h, float(S.FRAME_DT if frame_dt is None else frame_dt), C.byref(st)))
        return st

    def drain(self):
        """Pipelined servers: wait for the frame in flight and retire what it finished."""
        check(lib().vtx_server_drain(self.h))

    def status(self, job: int):
        s, d = C.c_int32(), C.c_int32()
        check(lib().vtx_server_job_status(self.h, job, C.byref(s), C.byref(d)))
        return s.value, d.value

    def result(self, job: int) -> np.ndarray:
        out = np.empty((self.n[job], 3), dtype=np.float32)
        check(lib().vtx_server_read_result(self.h, job, out.ctypes.data_as(C.POINTER(C.c_float)),
                                           self.n[job]))
        return out

    def preempt(self, job: int):
        check(lib().vtx_server_preempt(self.h, job))

    def release(self, job: int):
        check(lib().vtx_server_release_job(self.h, job))
        self.n.pop(job, None)
