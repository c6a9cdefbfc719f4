"""Parity in the draped state at full size, through the C ABI: configurations 2 and 3 are run on
the GPU until the cloth rests on its colliders with thousands of particle contacts, and from that
state (positions and velocities read back, exact in binary64)

* the hash tables and neighbour lists the GPU builds for the draped positions are bit-exact
  against the binary32 restatement (config 2, 2^18 cells);
* one substep on the GPU agrees with one substep of the FP64 oracle started from the same state
  within 1e-5 x extent (the one-substep bound), and one frame within 1e-4 x extent (config 2);
  measured on MI355X: config 2 (8,992 contact pairs after 60 frames) 1.2e-6 (substep) and 7.0e-6
  (frame) x extent; config 3 (148,795 contact pairs after 40 frames) 9.1e-6 x extent for the
  substep, the contacts at the edge of the diameter being where binary32 and binary64 decide a
  few pairs differently.

The oracle never runs the 60 frames itself: it starts from the GPU's state, so the test covers the
contact-heavy phase of the horizon at the cost of one oracle frame.
"""
import dataclasses
import sys

import numpy as np
import pytest

import oracle
from tests.test_gpu_parity import _check_hash_equal
from velvet_amd import scenes as S
from velvet_amd.solver import ClothSolver

pytestmark = pytest.mark.gpu
F32 = np.float32
SENT = 0xFFFFFFFF


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def _gpu_tables(sc, x):
    """The GPU's hash tables for predicted positions x (a fresh solver of the same scene)."""
    with ClothSolver(use_graph=False) as s:
        s.load_scene(sc)
        s.write_predicted(x.astype(F32))
        s.run_hash()
        return s.read_hash(0)


def _contacts(nb, x, d):
    """Listed pairs closer than the diameter, and the particles in them (vectorised)."""
    live = np.cumprod(nb != SENT, axis=0).astype(bool)
    ii = np.broadcast_to(np.arange(nb.shape[1]), nb.shape)[live]
    jj = nb[live].astype(np.int64)
    hit = np.linalg.norm(x[ii] - x[jj], axis=1) < d
    return int(hit.sum()), int(np.unique(ii[hit]).size)


def _draped(sc, frames):
    """(positions, velocities) after `frames` frames on the GPU, and the solver kept open."""
    s = ClothSolver()
    s.load_scene(sc)
    for _ in range(frames):
        s.Simulate()
    s.synchronize()
    return s, s.positions().astype(np.float64), s.velocities().astype(np.float64)


def _oracle_from(sc, x, v, frame_dt=None):
    sim = oracle.OracleSim(sc, frame_dt)
    sim.x.reshape(-1, 3)[:] = x
    sim.v.reshape(-1, 3)[:] = v
    return sim


def _one_substep(s, sc, x, v):
    """One substep on the GPU solver s (continuing from x, v) and on the oracle from x, v."""
    sub = dataclasses.replace(sc.params, num_substeps=1)
    h = F32(S.FRAME_DT / F32(sc.params.num_substeps))
    s.set_params(sub, 0)
    s.Simulate(h)
    s.synchronize()
    xg = s.positions().astype(np.float64)
    sim = _oracle_from(S.Scene(sub, sc.cloths, sc.colliders, 1, sc.name), x, v, h)
    sim.run(1)
    return float(np.max(np.abs(xg - sim.x.reshape(-1, 3))))


def test_config2_draped_state():
    sc = S.config2(frames=61)
    d = float(sc.params.particle_diameter)
    say("config2: 60 frames on the GPU")
    s, x, v = _draped(sc, 60)
    try:
        tables = _gpu_tables(sc, x)
        pairs, parts = _contacts(tables[4], x, d)
        say(f"  draped: {pairs} contact pairs, {parts} particles in contact")
        assert pairs >= 1000
        # hash stage on the draped positions
        ref = oracle.hash32(x.astype(F32), S.flatten(sc)["x"], float(S.hash_cell_size(sc.params)), d,
                            S.hash_table_size(sc.params, sc.n))
        _check_hash_equal(tables, ref, sc.n)
        err = _one_substep(s, sc, x, v)
    finally:
        s.close()
    say(f"  one substep from frame 60: max err {err:.3e} ({err / sc.extent():.2e} x extent)")
    assert err <= 1e-5 * sc.extent()
    # one whole frame from the same state
    s, x2, _ = _draped(sc, 61)
    s.close()
    sim = _oracle_from(sc, x, v)
    sim.run(1)
    err = float(np.max(np.abs(x2 - sim.x.reshape(-1, 3))))
    say(f"  one frame from frame 60: max err {err:.3e} ({err / sc.extent():.2e} x extent)")
    assert err <= 1e-4 * sc.extent()


def test_config3_draped_substep():
    """The 1024 x 1024 cloth resting on its three spheres after 40 frames: one substep against
    the oracle from the same state (its hash stage at this size is covered bit-exactly in
    test_gpu_parity)."""
    sc = S.config3(frames=41)
    d = float(sc.params.particle_diameter)
    say("config3: 40 frames on the GPU")
    s, x, v = _draped(sc, 40)
    try:
        pairs, parts = _contacts(_gpu_tables(sc, x)[4], x, d)
        say(f"  draped: {pairs} contact pairs, {parts} particles in contact")
        assert pairs >= 1000
        err = _one_substep(s, sc, x, v)
    finally:
        s.close()
    say(f"  one substep from frame 40: max err {err:.3e} ({err / sc.extent():.2e} x extent)")
    assert err <= 1e-5 * sc.extent()
