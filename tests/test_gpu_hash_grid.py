"""The hash-grid resolution (vtx_tuning.hash_grid) of segments that opt in (tune_hash_grid = 1),
through the C ABI, against the oracle run on the same grid:

* cell tables, sorted order and neighbour lists bit-exact against the binary32 restatement with
  the grid's shift and table size, for every hash_grid value, on the one-workgroup and the
  device-wide hash build, with a segment that did not opt in packed beside them (it keeps the
  reference grid);
* trajectories of a contact-heavy scene within 1e-5 x extent of the FP64 oracle on the same grid
  after one substep and within 1e-3 x extent after 10 frames, and a batch within 1e-3 x extent;
* vtx_autotune tries the grid when a segment opted in, leaves the state where it was, and the
  run under its choice matches the oracle on the chosen grid.
"""
import sys

import numpy as np
import pytest

import oracle
from tests.test_gpu_contacts import _contact_pairs, _start_velocities, stacked
from velvet_amd import scenes as S
from velvet_amd.solver import ClothSolver

pytestmark = pytest.mark.gpu
F32 = np.float32
SENT = 0xFFFFFFFF
SHIFT = 1000003          # VTX_HASH_GRID_SHIFT


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def grid_of(base_M, tune, hash_grid):
    """(table size, shift) a segment hashes with: the reference grid unless it opted in."""
    if not tune or hash_grid == 0:
        return base_M, 0
    m = base_M << max(0, hash_grid - 1)
    while m >= (1 << 28) and m > base_M:
        m >>= 1
    return m, SHIFT


def _folded(n, seed, lift=0.04):
    c = S.grid_cloth(n, 0.05, center=(0.0, 1.0, 0.0), jitter=1e-4, seed=seed, stream=7)
    x0 = c.x.astype(np.float64)
    p = x0.copy()
    half = (n // 2) * 0.05
    zc = x0[:, 2] - x0[:, 2].min()
    fold = zc > half
    p[fold, 2] = x0[:, 2].min() + 2 * half - zc[fold]
    p[fold, 1] += lift
    return c, p.astype(F32)


@pytest.mark.parametrize("hash_grid", [0, 1, 2, 3])
@pytest.mark.parametrize("hash_lds", [1, 0])
def test_tables_bit_exact_on_every_grid(hash_grid, hash_lds):
    # 128 x 128 = 16384 particles: one-workgroup build up to a 2^16-cell table (hash_grid <= 2),
    # device-wide beyond; 131 x 131 always device-wide; the 24 x 24 cloth did not opt in
    items = [_folded(16, 1), _folded(33, 2), _folded(24, 3), _folded(128, 6), _folded(131, 5)]
    tune = [1, 1, 0, 1, 1]
    sizes = [0, 3000, 0, 0, 0]
    say(f"hash grid {hash_grid}, hash_lds {hash_lds}")
    s = ClothSolver(use_graph=False)
    for k, (c, p) in enumerate(items):
        prm = S.Params(hash_table_size=sizes[k], particle_diameter=float(c.diameter), tune_hash_grid=tune[k])
        seg = 0 if k == 0 else s.add_segment(prm)
        s.set_params(prm, seg)
        s.AddCloth(c.x, c.tris, c.diameter, seg=seg)
        s.set_params(prm, seg)
    s.set_tuning(hash_grid=hash_grid, hash_lds=hash_lds)
    for k, (c, p) in enumerate(items):
        s.write_predicted(p, seg=k)
    s.run_hash()
    for k, (c, p) in enumerate(items):
        M, shift = grid_of(sizes[k] or 2 * c.n, tune[k], hash_grid)
        cell = float(F32(F32(c.diameter) * F32(1.5)))
        ph, rsh, rsi, rcs, rce, rnb = oracle.hash32(p, c.x, cell, float(c.diameter), M, shift=shift)
        cs, ce, sh, si, nb = s.read_hash(k)
        assert cs.shape[0] == M, (k, cs.shape, M)
        np.testing.assert_array_equal(cs, rcs)
        np.testing.assert_array_equal(ce[rcs != SENT], rce[rcs != SENT])
        np.testing.assert_array_equal(sh, rsh)
        np.testing.assert_array_equal(si, rsi)
        assert oracle.neighbor_lists(nb) == oracle.neighbor_lists(rnb), k
    s.close()


def _run(sc, v0, frames, hash_grid, frame_dt=None):
    with ClothSolver() as s:
        s.load_scene(sc)
        s.set_tuning(hash_grid=hash_grid)
        for i in np.nonzero(np.any(v0 != 0, axis=1))[0]:
            s.set_particle(int(i), vel=v0[i])
        for _ in range(frames):
            s.Simulate(frame_dt)
        s.synchronize()
        M = int(s.read_hash(0)[0].shape[0])
        return s.positions().astype(np.float64), M


@pytest.mark.parametrize("hash_grid", [1, 3])
def test_contacts_on_tuned_grid_vs_fp64(hash_grid):
    sc, moving = stacked()
    sc.params.tune_hash_grid = 1
    d = float(sc.params.particle_diameter)
    M_exp, shift = grid_of(2 * sc.n, 1, hash_grid)
    # one substep (tangential start velocities: see test_gpu_contacts)
    sc.params.num_substeps = 1
    h = F32(S.FRAME_DT / F32(5))
    v0 = _start_velocities(sc, moving, -0.07 * d / float(h), amp=1e-2)
    pairs, parts = _contact_pairs(sc, S.flatten(sc)["x"].astype(np.float64), v0.astype(np.float64), float(h))
    assert parts >= 0.1 * sc.n
    xg, M = _run(sc, v0, 1, hash_grid, frame_dt=h)
    assert M == M_exp
    sim = oracle.OracleSim(sc, h, hash_grid=(M, shift))
    sim.v.reshape(-1, 3)[:] = v0.astype(np.float64)
    sim.run(1)
    err = float(np.max(np.abs(xg - sim.x.reshape(-1, 3))))
    say(f"grid {hash_grid} (M={M}): {pairs} contact pairs, one substep max err {err:.3e}")
    assert err <= 1e-5 * sc.extent()
    # ten frames
    sc, moving = stacked()
    sc.params.tune_hash_grid = 1
    v0 = _start_velocities(sc, moving, -0.07 * d / (float(S.FRAME_DT) / sc.params.num_substeps))
    xg, M = _run(sc, v0, 10, hash_grid)
    sim = oracle.OracleSim(sc, hash_grid=(M, shift))
    sim.v.reshape(-1, 3)[:] = v0.astype(np.float64)
    sim.run(10)
    err = float(np.max(np.abs(xg - sim.x.reshape(-1, 3))))
    say(f"  ten frames max err {err:.3e}")
    assert err <= 1e-3 * sc.extent()


def _batch():
    scenes = S.config4(count=6, frames=3, smax=48)
    for sc in scenes:
        sc.params.tune_hash_grid = 1
    return scenes


def _check_batch(s, scenes, hash_grid, frames):
    for k, sc in enumerate(scenes):
        M = int(s.read_hash(k)[0].shape[0])
        M_exp, shift = grid_of(2 * sc.n, 1, hash_grid)
        assert M == M_exp, (k, M, M_exp)
        ref = oracle.simulate(sc, frames, hash_grid=(M, shift)).x.reshape(-1, 3)
        err = float(np.max(np.abs(s.positions(k).astype(np.float64) - ref)))
        assert err <= 1e-3 * sc.extent(), (k, err)


def test_batch_on_tuned_grid_vs_fp64():
    scenes = _batch()
    with ClothSolver() as s:
        for sc in scenes:
            s.load_scene(sc)
        s.set_tuning(hash_grid=2)
        for _ in range(3):
            s.Simulate()
        s.synchronize()
        say("batch of 6 on grid 2, 3 frames")
        _check_batch(s, scenes, 2, 3)


def test_autotune_tries_the_grid_and_matches_the_oracle():
    scenes = _batch()
    with ClothSolver() as s:
        for sc in scenes:
            s.load_scene(sc)
        before = [s.positions(k).copy() for k in range(len(scenes))]
        t = s.autotune(frames=2)
        say(f"autotune picked {t}")
        assert 0 <= t["hash_grid"] <= 3
        for k in range(len(scenes)):
            assert np.array_equal(s.positions(k), before[k])      # trials were undone
        for _ in range(3):
            s.Simulate()
        s.synchronize()
        _check_batch(s, scenes, t["hash_grid"], 3)


def test_segments_that_did_not_opt_in_ignore_the_grid():
    sc = S.config4(count=3, frames=3, smax=32)
    out = []
    for g in (0, 3):
        with ClothSolver() as s:
            for x in sc:
                s.load_scene(x)
            s.set_tuning(hash_grid=g)
            for _ in range(3):
                s.Simulate()
            s.synchronize()
            out.append([(s.positions(k), s.velocities(k)) for k in range(len(sc))])
    for (a, va), (b, vb) in zip(*out):
        assert np.array_equal(a, b) and np.array_equal(va, vb)
