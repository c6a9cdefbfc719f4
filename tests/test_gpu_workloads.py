"""The batch and serving workloads at their full sizes on the HIP path.

* the 512-cloth batch (sides log-uniform 16..128, 1,926,260 particles) packed into one solver and
  run on two stream groups: every cloth within 1e-3 x its extent of the FP64 oracle after two
  frames (the oracle runs the cloths on host threads; ctypes releases the GIL);
* serving at the request sizes (32^2..512^2 cloths, jobs of up to 249,001 particles arriving while
  others run): each result bit-identical to the job's solo run, one job against the oracle.
"""
import os
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
from velvet_amd import scenes as S
from velvet_amd.server import DONE, Server
from velvet_amd.solver import ClothSolver

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def _oracle_many(scenes, frames):
    threads = min(16, os.cpu_count() or 1)
    with ThreadPoolExecutor(threads) as ex:
        return list(ex.map(lambda sc: oracle.simulate(sc, frames=frames).x.reshape(-1, 3), scenes))


def test_full_batch_vs_fp64():
    scenes = S.config4(count=512, frames=2)
    say(f"full batch: {len(scenes)} cloths, {sum(sc.n for sc in scenes)} particles, 2 frames, 2 streams")
    with ClothSolver() as s:
        for sc in scenes:
            s.load_scene(sc)
        s.set_tuning(streams=2)
        for _ in range(2):
            s.Simulate()
        s.synchronize()
        xs = [s.positions(k) for k in range(len(scenes))]
    refs = _oracle_many(scenes, 2)
    worst = 0.0
    for k, (sc, x, r) in enumerate(zip(scenes, xs, refs)):
        err = float(np.max(np.abs(x.astype(np.float64) - r)))
        worst = max(worst, err / sc.extent())
        assert err <= 1e-3 * sc.extent(), (k, err)
    say(f"  worst max err / extent {worst:.3e}")


def test_serving_at_request_sizes():
    picks = [4, 0, 9, 12, 7, 14]            # 499^2, 53^2, 274^2, 204^2, 311^2, 86^2
    horizons = [4, 3, 3, 5, 3, 2]
    jobs = []
    for k, hz in zip(picks, horizons):
        sc, _ = S.serving_job(0, k)
        jobs.append((sc, hz))
    say("serving at request sizes: " + ", ".join(f"{sc.cloths[0].side}^2 x {hz}" for sc, hz in jobs))
    arrivals = {0: [0, 1], 1: [2], 2: [3, 4], 4: [5]}
    with Server(reserve_particles=1_000_000) as sv:
        ids = {}
        for step in range(30):
            for k in arrivals.get(step, []):
                ids[k] = sv.submit(*jobs[k])
            sv.step()
            if len(ids) == len(jobs) and all(sv.status(j)[0] == DONE for j in ids.values()):
                break
        res = {k: sv.result(j) for k, j in ids.items()}
        for k, j in ids.items():
            assert sv.status(j) == (DONE, jobs[k][1]), k
    for k, (sc, hz) in enumerate(jobs):
        with ClothSolver() as s:
            s.load_scene(sc)
            for _ in range(hz):
                s.Simulate()
            s.synchronize()
            assert np.array_equal(res[k], s.positions()), k
    sc, hz = jobs[2]
    ref = oracle.simulate(sc, frames=hz).x.reshape(-1, 3)
    err = float(np.max(np.abs(res[2].astype(np.float64) - ref)))
    say(f"  {sc.cloths[0].side}^2 job vs fp64 after {hz} frames: {err:.3e}")
    assert err <= 1e-3 * sc.extent()
