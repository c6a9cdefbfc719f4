"""Serving (continuous batching) on the GPU through the C ABI.

* each job's result is bit-identical alone and in a mix where other jobs arrive and leave;
* admission respects the device-memory budget and the live-particle cap, queues the rest and
  drains the queue as jobs finish; a job larger than the budget is rejected;
* a job whose state turns non-finite (forced from the host) is evicted as failed while the
  other jobs' results stay identical to their solo runs;
* a preempted job (state parked on the host, device memory freed, queued again) ends
  bit-identical to its solo run.
"""
import sys

import numpy as np
import pytest

import oracle
from velvet_amd import scenes as S
from velvet_amd.server import DONE, FAILED, QUEUED, REJECTED, RUNNING, Server
from velvet_amd.solver import ClothSolver

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def jobs(count=5, smin=16, smax=40):
    out = []
    for k in range(count):
        sc, _ = S.serving_job(7, k, smin=smin, smax=smax)
        out.append((sc, 4 + 3 * k))
    return out


def alone(scene, frames):
    with ClothSolver() as s:
        s.load_scene(scene)
        for _ in range(frames):
            s.Simulate()
        s.synchronize()
        return s.positions()


@pytest.mark.parametrize("use_graph,reserve,pipelined", [(False, 0, False), (True, 0, False), (False, 200_000, False),
                                                        (False, 0, True), (True, 200_000, True)])
def test_mix_equals_alone(use_graph, reserve, pipelined):
    """Jobs that arrive while others run are uploaded during a frame (staged) and join at the
    next step; every result equals the job's solo run, with or without pre-sized pools, and with
    steps that return while their frame runs (pipelined)."""
    say(f"serving mix vs alone graph={use_graph} reserve={reserve} pipelined={pipelined}")
    js = jobs(5)
    with Server(use_graph=use_graph, reserve_particles=reserve, pipelined=pipelined) as sv:
        ids = []
        # staggered arrivals: two now, one after 2 steps, two after 5
        arrivals = {0: [0, 1], 2: [2], 5: [3, 4]}
        step = 0
        while step < 40:
            for k in arrivals.get(step, []):
                ids.append((k, sv.submit(js[k][0], js[k][1])))
            st = sv.step()
            step += 1
            assert min(st.admit_ms, st.frame_ms, st.retire_ms) >= 0.0
            assert st.frame_ms <= st.step_ms and st.retire_ms <= st.step_ms
            if len(ids) == 5 and all(sv.status(j)[0] == DONE for _, j in ids):
                break
        for k, j in ids:
            status, done = sv.status(j)
            assert status == DONE and done == js[k][1]
            ref = alone(*js[k])
            assert np.array_equal(sv.result(j), ref), k
            sv.release(j)


def test_admission_budget_and_queue():
    say("serving admission")
    js = jobs(6)
    cap = max(sc.n for sc, _ in js) + 10
    with Server(max_live_particles=cap) as sv:
        ids = [sv.submit(sc, f) for sc, f in js]
        seen_live = []
        for _ in range(200):
            st = sv.step()
            seen_live.append(st.live_particles)
            assert st.live_particles <= cap
            if all(sv.status(j)[0] == DONE for j in ids):
                break
        assert all(sv.status(j)[0] == DONE for j in ids)
        assert max(seen_live) <= cap
        for (sc, f), j in zip(js, ids):
            assert np.array_equal(sv.result(j), alone(sc, f))
    # a job larger than the whole budget is rejected, the others still run
    with Server(max_live_particles=500) as sv:
        small = S.Scene(js[0][0].params, [S.grid_cloth(12, 0.05, seed=1, stream=3)], js[0][0].colliders, 3)
        big = S.Scene(js[0][0].params, [S.grid_cloth(30, 0.05, seed=1, stream=4)], js[0][0].colliders, 3)
        jb = sv.submit(big, 3)
        js_ = sv.submit(small, 3)
        assert sv.status(jb)[0] == REJECTED
        for _ in range(5):
            sv.step()
        assert sv.status(js_)[0] == DONE


def test_step_time_budget_admission():
    """A budget below any frame's cost admits one job per empty server: jobs run one at a time
    and each result still equals its solo run; a roomy budget runs them together."""
    say("serving step-time budget")
    js = jobs(4)
    with Server(step_budget_ms=1e-6) as sv:
        ids = [sv.submit(sc, f) for sc, f in js]
        most = 0
        for _ in range(200):
            st = sv.step()
            most = max(most, st.live_jobs)
            if all(sv.status(j)[0] == DONE for j in ids):
                break
        assert most == 1
        for (sc, f), j in zip(js, ids):
            assert sv.status(j)[0] == DONE
            assert np.array_equal(sv.result(j), alone(sc, f))
    with Server(step_budget_ms=1e4) as sv:
        ids = [sv.submit(sc, f) for sc, f in js]
        most = 0
        for _ in range(200):
            st = sv.step()
            most = max(most, st.live_jobs)
            if all(sv.status(j)[0] == DONE for j in ids):
                break
        assert most == len(js)      # one job on the first step (no measurement yet), then the rest
        assert all(sv.status(j)[0] == DONE for j in ids)


@pytest.mark.parametrize("pipelined", [False, True])
def test_preempt_resume_equals_alone(pipelined):
    """A running job taken off the device and re-admitted later ends bit-identical to its solo run
    (pipelined: preemption first completes the frame in flight)."""
    say(f"serving preemption pipelined={pipelined}")
    js = jobs(3)
    with Server(pipelined=pipelined) as sv:
        ids = [sv.submit(sc, f) for sc, f in js]
        sv.step()
        sv.step()
        sv.preempt(ids[1])
        status, done = sv.status(ids[1])
        assert status == QUEUED and done == 2
        st = sv.step()
        assert st.admitted == 1
        for _ in range(60):
            sv.step()
            if all(sv.status(j)[0] == DONE for j in ids):
                break
        for (sc, f), j in zip(js, ids):
            assert sv.status(j) == (DONE, f)
            assert np.array_equal(sv.result(j), alone(sc, f))
        # the preempted job's result is the simulation's: the FP64 oracle over its horizon
        sc, f = js[1]
        ref = oracle.simulate(sc, frames=f).x.reshape(-1, 3)
        err = float(np.max(np.abs(sv.result(ids[1]).astype(np.float64) - ref)))
        say(f"  preempted job vs fp64 after {f} frames: {err:.3e}")
        assert err <= 1e-3 * sc.extent()
        with pytest.raises(Exception):
            sv.preempt(ids[0])          # done jobs cannot be preempted


def test_preempt_staged_job():
    """A job admitted while another runs is uploaded during that frame and joins at the next step;
    preempted before it has run a frame, it goes back to the queue with 0 frames done and still
    ends bit-identical to its solo run."""
    say("serving preemption of a staged job")
    js = jobs(2)
    with Server() as sv:
        a = sv.submit(*js[0])
        sv.step()                                   # empty server: a starts at once
        assert sv.status(a) == (RUNNING, 1)
        b = sv.submit(*js[1])
        st = sv.step()                              # b is admitted during a's frame
        assert st.admitted == 1 and st.live_jobs == 1 and sv.status(b) == (RUNNING, 0)
        sv.preempt(b)
        assert sv.status(b) == (QUEUED, 0)
        for _ in range(60):
            sv.step()
            if all(sv.status(j)[0] == DONE for j in (a, b)):
                break
        for (sc, f), j in zip(js, (a, b)):
            assert sv.status(j) == (DONE, f)
            assert np.array_equal(sv.result(j), alone(sc, f))


def test_nonfinite_job_is_evicted():
    say("serving failure detection")
    js = jobs(3)
    with Server(check_every=1) as sv:
        ids = [sv.submit(sc, f) for sc, f in js]
        sv.step()
        # force a non-finite position into job 1 from the host
        seg = None
        # job 1's segment: the server's solver segments are created in submission order
        segs = [k for k in range(1, 10) if _has_segment(sv.solver, k)]
        seg = segs[1]
        sv.solver.set_particle(5, pos=np.array([np.nan, 0.0, 0.0], np.float32), seg=seg)
        for _ in range(40):
            sv.step()
            if all(sv.status(j)[0] in (DONE, FAILED) for j in ids):
                break
        assert sv.status(ids[1])[0] == FAILED
        for k in (0, 2):
            assert sv.status(ids[k])[0] == DONE
            assert np.array_equal(sv.result(ids[k]), alone(*js[k]))


def _has_segment(solver, seg):
    try:
        solver.num_particles(seg)
        return True
    except Exception:
        return False


def test_pipelined_steps_lag_by_one_frame_and_drain():
    """A pipelined step returns with its frame still running: the frame is counted (frames done,
    completed jobs, frame time) by the next step or by drain(), and the results match a
    synchronous server's bit for bit."""
    say("serving pipelined step accounting")
    js = jobs(2)
    with Server(pipelined=True) as sv:
        a = sv.submit(js[0][0], 3)
        st = sv.step()                       # frame 1 launched
        assert st.frame_ms == 0.0 and st.particle_substeps == 0 and sv.status(a) == (RUNNING, 0)
        st = sv.step()                       # frame 1 counted, frame 2 launched
        assert st.frame_ms > 0.0 and st.particle_substeps > 0 and sv.status(a) == (RUNNING, 1)
        st = sv.step()                       # frame 2 counted, frame 3 (the last) launched
        assert sv.status(a) == (RUNNING, 2)
        sv.drain()                           # frame 3 counted, the job retired
        assert sv.status(a) == (DONE, 3)
        st = sv.step()                       # the drained frame's figures arrive here
        assert st.completed == 1 and st.frame_ms > 0.0 and st.live_jobs == 0
        res = sv.result(a)
    assert np.array_equal(res, alone(js[0][0], 3))
