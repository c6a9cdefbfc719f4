"""Parity of the HIP path (through the C ABI) with the oracle.

* hash stage: cell tables, sorted order and neighbour lists bit-exact against the binary32
  restatement, for single and packed segments, with duplicates and caps;
* trajectories: within 1e-5 x extent of the FP64 oracle after one substep and within
  1e-3 x extent after 60 frames (config 1);
* closed-form cases through the whole pipeline;
* packing invariance: a segment gives bit-identical results alone or packed with others,
  graph or no graph, run after run.
"""
import dataclasses
import sys

import numpy as np
import pytest

import oracle
from velvet_amd import scenes as S
from velvet_amd.solver import ClothSolver

pytestmark = pytest.mark.gpu
F32 = np.float32
SENT = 0xFFFFFFFF


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def _check_hash_equal(gpu, ref, n):
    cs, ce, sh, si, nb = gpu
    ph, rsh, rsi, rcs, rce, rnb = ref
    np.testing.assert_array_equal(cs, rcs)
    valid = rcs != SENT
    np.testing.assert_array_equal(ce[valid], rce[valid])
    np.testing.assert_array_equal(sh, rsh)
    np.testing.assert_array_equal(si, rsi)
    assert oracle.neighbor_lists(nb) == oracle.neighbor_lists(rnb)


def _folded(n=16, seed=0, lift=0.04):
    c = S.grid_cloth(n, 0.05, center=(0.0, 1.0, 0.0), jitter=1e-4, seed=seed, stream=7)
    x0 = c.x.astype(np.float64)
    p = x0.copy()
    half = (n // 2) * 0.05
    zc = x0[:, 2] - x0[:, 2].min()
    fold = zc > half
    p[fold, 2] = x0[:, 2].min() + 2 * half - zc[fold]
    p[fold, 1] += lift
    return c, p.astype(F32)


def _hash_solver(cloths_and_p, table_sizes=None):
    s = ClothSolver(use_graph=False)
    for k, (c, p) in enumerate(cloths_and_p):
        prm = S.Params(hash_table_size=(table_sizes[k] if table_sizes else 0),
                       particle_diameter=float(c.diameter))
        seg = 0 if k == 0 else s.add_segment(prm)
        if k == 0:
            s.set_params(prm, 0)
        s.AddCloth(c.x, c.tris, c.diameter, seg=seg)
        s.set_params(prm, seg)
    for k, (c, p) in enumerate(cloths_and_p):
        s.write_predicted(p, seg=k)
    s.run_hash()
    return s


@pytest.mark.parametrize("n,M", [(16, 0), (24, 0), (16, 1), (16, 97), (40, 1 << 12)])
def test_hash_stage_bit_exact_single(n, M):
    say(f"hash single n={n} M={M}")
    c, p = _folded(n, seed=n)
    s = _hash_solver([(c, p)], [M])
    Mx = M if M else 2 * c.n
    cell = float(F32(F32(c.diameter) * F32(1.5)))
    ref = oracle.hash32(p, c.x, cell, float(c.diameter), Mx)
    _check_hash_equal(s.read_hash(0), ref, c.n)
    s.close()


def test_hash_stage_bit_exact_packed_segments():
    say("hash packed segments")
    items = [_folded(16, seed=1), _folded(33, seed=2), _folded(9, seed=3, lift=0.02), _folded(50, seed=4)]
    sizes = [0, 3000, 0, 1 << 13]
    s = _hash_solver(items, sizes)
    for k, (c, p) in enumerate(items):
        Mx = sizes[k] if sizes[k] else 2 * c.n
        cell = float(F32(F32(c.diameter) * F32(1.5)))
        ref = oracle.hash32(p, c.x, cell, float(c.diameter), Mx)
        _check_hash_equal(s.read_hash(k), ref, c.n)
    s.close()


def test_hash_nbr_entry_width_boundary():
    """Neighbour tables at the 16-bit entry limit: a 65,535-particle segment (the largest with 16-bit
    entries: indices up to 65,534, the sentinel 0xffff equal to n) packed with a 65,536-particle one
    (32-bit entries); both read back bit-exact against the oracle, in every thread order."""
    say("hash entry-width boundary n=65535 / 65536")
    c, p = _folded(256, seed=11)
    keep = 65535
    tris = c.tris[np.all(c.tris < keep, axis=1)]
    small = (dataclasses.replace(c, x=c.x[:keep].copy(), tris=tris), p[:keep].copy())
    items = [small, (c, p)]
    s = _hash_solver(items)
    try:
        for order in (1, 2, 0):
            s.set_tuning(neighbor_order=order)
            for k, (ck, pk) in enumerate(items):
                s.write_predicted(pk, seg=k)
            s.run_hash()
            for k, (ck, pk) in enumerate(items):
                n = ck.x.shape[0]
                cell = float(F32(F32(ck.diameter) * F32(1.5)))
                ref = oracle.hash32(pk, ck.x, cell, float(ck.diameter), 2 * n)
                _check_hash_equal(s.read_hash(k), ref, n)
    finally:
        s.close()


@pytest.mark.parametrize("hash_lds,order", [(1, 2), (0, 2), (1, 1), (0, 1), (1, 0)])
def test_hash_stage_bit_exact_mixed_paths(hash_lds, order):
    """Segments hashed in one workgroup each (at most 16384 particles, at most 2^16 cells; 128 x 128
    is the largest) packed with segments that take the device-wide sort (131 x 131 = 17161
    particles; a small cloth with a 2^17-cell table): every segment's tables and lists match the
    oracle, with the small-segment path on or off and every neighbour-search thread order."""
    say(f"hash mixed paths hash_lds={hash_lds} neighbor_order={order}")
    items = [_folded(16, seed=1), _folded(131, seed=2), _folded(40, seed=3), _folded(20, seed=4),
             _folded(128, seed=6)]                  # 16384 particles: the largest one-workgroup build
    sizes = [0, 0, 1 << 17, 97, 0]
    s = _hash_solver(items, sizes)
    s.set_tuning(hash_lds=hash_lds, neighbor_order=order)
    for k, (c, p) in enumerate(items):
        s.write_predicted(p, seg=k)
    s.run_hash()
    for k, (c, p) in enumerate(items):
        Mx = sizes[k] if sizes[k] else 2 * c.n
        cell = float(F32(F32(c.diameter) * F32(1.5)))
        ref = oracle.hash32(p, c.x, cell, float(c.diameter), Mx)
        _check_hash_equal(s.read_hash(k), ref, c.n)
    s.close()


def test_hash_table_size_limit():
    """Cell tables of 2^28 or more cells are refused before anything is allocated (the neighbour
    search addresses a table with 32-bit byte offsets); 2^28 - 1 is accepted."""
    from velvet_amd._lib import VtxError
    c, p = _folded(8, seed=1)
    for M, ok in ((1 << 28, False), ((1 << 28) - 1, True)):
        say(f"hash table size {M}")
        s = ClothSolver(use_graph=False)
        try:
            prm = S.Params(hash_table_size=M, particle_diameter=float(c.diameter))
            s.set_params(prm, 0)
            s.AddCloth(c.x, c.tris, c.diameter)
            s.set_params(prm, 0)
            if ok:
                s.Simulate()
                s.synchronize()
            else:
                with pytest.raises(VtxError, match="2\\^28"):
                    s.Simulate()
        finally:
            s.close()


@pytest.mark.parametrize("shift", [117966.5, -117966.2])
def test_hash_stage_bit_exact_far_from_origin(shift):
    """A folded cloth about 2^20 cells from the origin on x (positions with a 1/128 m binary32
    spacing, cell coordinates beyond 20 bits): cell tables and lists stay bit-exact."""
    say(f"hash far from origin x+{shift}")
    c, p = _folded(24, seed=5)
    off = np.array([shift, 0.0, 0.0], F32)
    c = S.Cloth(**{**c.__dict__, "x": (c.x + off).astype(F32)})
    p = (p + off).astype(F32)
    s = _hash_solver([(c, p)], [0])
    cell = float(F32(F32(c.diameter) * F32(1.5)))
    ix = np.floor(p[:, 0] / F32(cell))
    assert ix.min() < 2 ** 20 <= ix.max() or ix.min() < -(2 ** 20) <= ix.max()
    ref = oracle.hash32(p, c.x, cell, float(c.diameter), 2 * c.n)
    _check_hash_equal(s.read_hash(0), ref, c.n)
    s.close()


def test_hash_stage_matches_golden_fixture():
    say("hash golden fixture")
    import os
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hash_folded16.npz"))
    c, p = _folded(16, seed=0)
    c = S.grid_cloth(16, 0.05, center=(0.0, 1.0, 0.0), jitter=1e-4, seed=0, stream=7)
    assert np.array_equal(c.x, g["x0"])
    s = _hash_solver([(c, g["p"])], [int(g["M"])])
    cs, ce, sh, si, nb = s.read_hash(0)
    np.testing.assert_array_equal(cs, g["cell_start"])
    valid = cs != SENT
    np.testing.assert_array_equal(ce[valid], g["cell_end"][valid])
    np.testing.assert_array_equal(si, g["sorted_index"])
    assert oracle.neighbor_lists(nb) == oracle.neighbor_lists(g["nbr"])
    s.close()


@pytest.mark.parametrize("K", [64, 13, 7, 1])
def test_hash_caps_and_duplicates(K):
    """80 particles: the first 50 share one predicted position (49 candidates each, lists cut at
    the cap), the other 30 are isolated (empty lists) or in a group of 10 (9 entries and a
    sentinel when the cap is above 9); caps that are not a multiple of four end in a partial
    16-byte chunk of the device table."""
    say(f"hash caps/duplicates, max_num_neighbors {K}")
    n = 80
    x0 = np.stack([np.arange(n) * 1.0, np.zeros(n), np.zeros(n)], 1).astype(F32)
    p = np.tile(np.array([[0.01, 0.01, 0.01]], F32), (n, 1))
    p[50:60] = (5.01, 0.01, 0.01)
    p[60:] = np.stack([10.0 + np.arange(20) * 3.0, np.zeros(20), np.zeros(20)], 1)
    c = S.Cloth(x=x0, tris=np.zeros((0, 3), np.uint32), stretch=np.zeros((0, 2), np.int32),
                stretch_rest=np.zeros(0, F32), bend=np.zeros((0, 4), np.uint32), bend_rest=np.zeros(0, F32),
                slots=np.zeros((0, 3), F32), attach_pid=np.zeros(0, np.int32), attach_slot=np.zeros(0, np.int32),
                attach_dist=np.zeros(0, F32), diameter=F32(0.05))
    for M in (2 * n, 1, 5):
        s = ClothSolver(use_graph=False)
        prm = S.Params(hash_table_size=M, particle_diameter=0.05, hash_cell_size_scalar=2.0, max_num_neighbors=K)
        s.set_params(prm)
        s.AddCloth(c.x, c.tris, c.diameter)
        s.set_params(prm)
        s.write_predicted(p)
        s.run_hash()
        ref = oracle.hash32(p, x0, 0.1, 0.05, M, max_nbr=K)
        got = s.read_hash(0)
        assert got[4].shape == (K, n)
        _check_hash_equal(got, ref, n)
        s.close()


def _run_gpu(scene, frames, frame_dt=None, use_graph=True, stencil=True, expect_stencil=None):
    with ClothSolver(use_graph=use_graph, stencil=stencil) as s:
        s.load_scene(scene)
        if expect_stencil is not None:
            assert s.stencil_segments() == expect_stencil
        for _ in range(frames):
            s.Simulate(frame_dt)
        s.synchronize()
        return s.positions(), s.velocities(), s.normals()


@pytest.mark.parametrize("stencil", [True, False])
def test_config1_one_substep_vs_fp64(stencil):
    say(f"config1 one substep stencil={stencil}")
    sc = S.config1(frames=1)
    sc.params.num_substeps = 1
    dt = F32(S.FRAME_DT / F32(5))
    x, v, nrm = _run_gpu(sc, 1, frame_dt=dt, stencil=stencil, expect_stencil=int(stencil))
    ref = oracle.simulate(sc, frames=1, frame_dt=dt)
    err = np.max(np.abs(x.astype(np.float64) - ref.x.reshape(-1, 3)))
    say(f"  max err {err:.3e}")
    assert err <= 1e-5 * sc.extent()


@pytest.mark.parametrize("stencil", [True, False])
def test_config1_sixty_frames_vs_fp64(stencil):
    say(f"config1 60 frames stencil={stencil}")
    sc = S.config1(frames=60)
    x, v, nrm = _run_gpu(sc, 60, stencil=stencil, expect_stencil=int(stencil))
    import os
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "config1_oracle.npz"))
    assert np.max(np.abs(x.astype(np.float64) - g["sixty_frames"])) <= 1e-3 * sc.extent()
    ref = oracle.simulate(sc, frames=60)
    d = np.abs(x.astype(np.float64) - ref.x.reshape(-1, 3))
    err = float(np.max(d))
    say(f"  max err {err:.3e} (extent {sc.extent():.3f})")
    assert np.all(np.isfinite(x))
    assert err <= 1e-3 * sc.extent()
    # normals agree too (unit vectors; measured on MI355X: 2.5e-4 after one substep and after 60
    # frames, on both Jacobi paths)
    nref = ref.normals.reshape(-1, 3)
    nerr = float(np.max(np.abs(nrm - nref)))
    say(f"  normals max err {nerr:.3e}")
    assert nerr < 2e-3


@pytest.mark.parametrize("bend_mode", ["quad", "pairs", "none"])
@pytest.mark.parametrize("stencil", [True, False])
def test_config1_frames_reference_topologies(bend_mode, stencil):
    say(f"config1 10 frames {bend_mode} stencil={stencil}")
    sc = S.config1(frames=10, bend_mode=bend_mode)
    x, _, _ = _run_gpu(sc, 10, stencil=stencil, expect_stencil=int(stencil))
    ref = oracle.simulate(sc, frames=10)
    err = np.max(np.abs(x.astype(np.float64) - ref.x.reshape(-1, 3)))
    assert err <= 1e-3 * sc.extent()


def _tiny(x, **kw):
    from tests.test_oracle_pins import tiny_scene
    return tiny_scene(x, **kw)


def test_closed_forms_through_pipeline():
    say("closed forms")
    # free fall with damping
    sc = _tiny([[0, 1, 0]], num_substeps=5, damping=0.25, gravity=(0.0, -9.8, 0.0))
    x, v, _ = _run_gpu(sc, 3)
    ref = oracle.simulate(sc, frames=3)
    np.testing.assert_allclose(x.reshape(-1), ref.x.reshape(-1), atol=2e-6)
    np.testing.assert_allclose(v.reshape(-1), ref.v.reshape(-1), atol=2e-5)
    # stretch projection: two particles 2 apart, rest 1
    sc = _tiny([[0, 0, 0], [2, 0, 0]], stretch=[(0, 1, 1.0)])
    x, _, _ = _run_gpu(sc, 1)
    assert abs((x[1, 0] - x[0, 0]) - 1.0) < 2e-6
    # attachment to 1.2 r and pinning
    sc = _tiny([[3, 0, 0]], slots=[[0, 0, 0]], attach=[(0, 0, 1.0)])
    x, _, _ = _run_gpu(sc, 1)
    assert abs(np.linalg.norm(x[0]) - 1.2) < 2e-6
    sc = _tiny([[1.0, 0, 0], [1.0, 5.0, 0]], stretch=[(0, 1, 10.0)], slots=[[0, 0, 0]],
               attach=[(0, 0, 0.0)], num_iterations=4, relaxation_factor=1.5)
    x, _, _ = _run_gpu(sc, 1)
    assert abs(x[0, 0] - (1 - 1.5 / 2) ** 4) < 1e-6
    # sphere projection
    sc = _tiny([[0.1, 0.2, -0.05]], colliders=[S.sphere((0.0, 0.0, 0.0), 0.5)], collision_margin=0.06)
    x, _, _ = _run_gpu(sc, 1)
    assert abs(np.linalg.norm(x[0]) - 0.56) < 2e-6


def test_packing_invariance_and_determinism():
    say("batch vs alone")
    scenes = S.config4(count=6, frames=5, smax=48)
    alone = []
    for sc in scenes:
        x, v, nrm = _run_gpu(sc, 5)
        alone.append((x, v))
    with ClothSolver() as s:
        for sc in scenes:
            s.load_scene(sc)
        assert s.num_segments() == len(scenes)
        for _ in range(5):
            s.Simulate()
        s.synchronize()
        for k in range(len(scenes)):
            assert np.array_equal(s.positions(k), alone[k][0]), k
            assert np.array_equal(s.velocities(k), alone[k][1]), k
    # graph vs direct launches, and a second run: bit-identical
    x1, _, _ = _run_gpu(scenes[0], 5, use_graph=False)
    x2, _, _ = _run_gpu(scenes[0], 5, use_graph=True)
    assert np.array_equal(x1, alone[0][0]) and np.array_equal(x2, alone[0][0])


def test_batch_segments_vs_oracle():
    say("batch segments vs oracle")
    scenes = S.config4(count=4, frames=10, smax=40)
    with ClothSolver() as s:
        for sc in scenes:
            s.load_scene(sc)
        for _ in range(10):
            s.Simulate()
        s.synchronize()
        for k, sc in enumerate(scenes):
            ref = oracle.simulate(sc, frames=10).x.reshape(-1, 3)
            err = np.max(np.abs(s.positions(k).astype(np.float64) - ref))
            assert err <= 1e-3 * sc.extent(), (k, err)


def test_normals_flat_and_host_edits():
    say("normals / host edits")
    sc = S.config1(frames=1, jitter=0.0)
    sc.params.gravity = (0.0, 0.0, 0.0)
    sc.colliders = []
    with ClothSolver() as s:
        s.load_scene(sc)
        s.Simulate()
        s.synchronize()
        nr = s.normals()
        assert np.all(np.abs(np.abs(nr[:, 1]) - 1) < 1e-5)
        # MouseGrabber-style edit: pin particle 100 at a new place
        target = np.array([0.1, 2.0, 0.1], F32)
        s.set_particle(100, pos=target, vel=np.zeros(3, F32), inv_mass=0.0)
        assert s.inv_masses()[100] == 0.0
        s.Simulate()
        s.synchronize()
        x = s.positions()
        assert np.linalg.norm(x[100] - target) < 0.05
        # slot animation moves the pinned corner with it
        s.SetAttachSlotPosition(0, sc.cloths[0].slots[0] + np.array([0, 0.05, 0], F32))
        for _ in range(5):
            s.Simulate()
        s.synchronize()
        assert s.positions()[0, 1] > sc.cloths[0].x[0, 1] + 0.02


def test_stencil_detection_rejects_non_grid():
    say("stencil detection")
    sc = S.config1(frames=1)
    c = sc.cloths[0]
    # an extra long-range stretch pair is not a grid neighbour: general path
    c2 = S.Cloth(**{**c.__dict__, "stretch": np.concatenate([c.stretch, [[0, 45]]]).astype(np.int32),
                   "stretch_rest": np.concatenate([c.stretch_rest, [np.float32(0.3)]]).astype(F32)})
    sc2 = S.Scene(sc.params, [c2], sc.colliders, 1)
    with ClothSolver() as s:
        s.load_scene(sc2)
        assert s.stencil_segments() == 0
        s.Simulate()
        s.synchronize()
        x = s.positions()
    ref = oracle.simulate(sc2, frames=1).x.reshape(-1, 3)
    assert np.max(np.abs(x - ref)) <= 1e-3 * sc.extent()
    # two cloths in one segment (the reference's shared solver): general path, shared hash
    sc3 = S.Scene(sc.params, [c, S.grid_cloth(12, 0.05, center=(0.0, 1.8, 0.0), seed=5, stream=9)],
                  sc.colliders, 1)
    with ClothSolver() as s:
        s.load_scene(sc3)
        assert s.stencil_segments() == 0 and s.num_particles() == c.n + 144
        for _ in range(3):
            s.Simulate()
        s.synchronize()
        x = s.positions()
    ref = oracle.simulate(sc3, frames=3).x.reshape(-1, 3)
    assert np.max(np.abs(x - ref)) <= 1e-3 * sc.extent()


def test_config2_lra_frames_vs_fp64():
    say("config2 3 frames")
    sc = S.config2(frames=3)
    x, _, _ = _run_gpu(sc, 3, expect_stencil=1)
    ref = oracle.simulate(sc, frames=3).x.reshape(-1, 3)
    err = np.max(np.abs(x.astype(np.float64) - ref))
    say(f"  max err {err:.3e}")
    assert err <= 1e-3 * sc.extent()


def _lra40(frames=4):
    """40 x 40 cloth whose first row (40 particles) is long-range attached: 40 dense slots, enough
    for the four-threads-per-particle Jacobi pass on the stencil path."""
    sc = S.config1(frames=frames)
    c = S.grid_cloth(40, 0.05, center=(0.0, 1.5, 0.0), seed=0, stream=1, attached=tuple(range(40)))
    return S.Scene(sc.params, [c], sc.colliders, frames, "lra40")


@pytest.mark.parametrize("stencil,slots", [(True, 2), (False, 2), (True, 40), (False, 40)])
def test_dense_and_listed_attachments_agree(stencil, slots):
    """Attachments registered slot-major over every particle are stored as distances only; the
    same attachments registered particle-major go through the per-particle lists. Each particle
    sums its terms in four parts by slot index, each part in slot order, either way, so the
    trajectories are bit-identical -- also when the dense form runs four threads per particle."""
    say(f"dense vs listed attachments stencil={stencil} slots={slots}")
    sc = S.config1(frames=4) if slots == 2 else _lra40()
    c = sc.cloths[0]
    order = np.lexsort((c.attach_slot, c.attach_pid))       # particle-major, slot order within
    c2 = S.Cloth(**{**c.__dict__, "attach_pid": c.attach_pid[order], "attach_slot": c.attach_slot[order],
                   "attach_dist": c.attach_dist[order]})
    sc2 = S.Scene(sc.params, [c2], sc.colliders, 4)
    xa, va, _ = _run_gpu(sc, 4, stencil=stencil)
    xb, vb, _ = _run_gpu(sc2, 4, stencil=stencil)
    assert np.array_equal(xa, xb) and np.array_equal(va, vb)


def test_snapshot_restore_roundtrip():
    """A segment's snapshot taken mid-run replays bit-identically, in place and in a fresh solver
    with the same registration; mismatched blobs are rejected."""
    say("snapshot / restore")
    scenes = S.config4(count=4, frames=6, smax=40)
    with ClothSolver() as s:
        for sc in scenes:
            s.load_scene(sc)
        for _ in range(3):
            s.Simulate()
        s.synchronize()
        snap = s.snapshot(seg=2)
        for _ in range(3):
            s.Simulate()
        s.synchronize()
        xa, va = s.positions(2), s.velocities(2)
        s.restore(snap, seg=2)
        for _ in range(3):
            s.Simulate()
        s.synchronize()
        assert np.array_equal(s.positions(2), xa) and np.array_equal(s.velocities(2), va)
        # the replayed state is the simulation's: the FP64 oracle after 6 frames
        ref = oracle.simulate(scenes[2], frames=6)
        err = float(np.max(np.abs(xa.astype(np.float64) - ref.x.reshape(-1, 3))))
        say(f"  restored run vs fp64 after 6 frames: {err:.3e}")
        assert err <= 1e-3 * scenes[2].extent()
        with pytest.raises(Exception):
            s.restore(snap[:-16], seg=2)
        if scenes[1].n != scenes[2].n:
            with pytest.raises(Exception):
                s.restore(snap, seg=1)
    with ClothSolver() as s2:
        s2.load_scene(scenes[2])
        s2.restore(snap)
        for _ in range(3):
            s2.Simulate()
        s2.synchronize()
        assert np.array_equal(s2.positions(), xa) and np.array_equal(s2.velocities(), va)


def _irregular_horizontal(res, seed, y=1.4, scale=0.5):
    x, tris = S.irregular_mesh(res, seed=seed, stream=11)
    h = np.empty_like(x)
    h[:, 0] = x[:, 0] * scale
    h[:, 1] = y
    h[:, 2] = (-x[:, 1] - 1.0) * scale
    return h.astype(F32), tris


def test_mesh_rest_angles_hold_a_curved_shape():
    """An irregular mesh rolled into a cylinder arc keeps its shape with generated rest angles
    (no gravity, nothing touching it); with the rest angles zeroed it unrolls."""
    say("mesh rest angles")
    x, tris = _irregular_horizontal(20, seed=4)
    R = 0.4
    th = x[:, 0] / R
    curved = np.stack([R * np.sin(th), 1.4 + R * (np.cos(th) - 1.0), x[:, 2]], axis=1).astype(F32)
    spacing = 1.0 / 20
    c = S.mesh_cloth(curved, tris, diameter=spacing * 1.5)
    assert c.bend.shape[0] > 0 and float(np.max(c.bend_rest)) > 0.05
    p = S.Params(num_substeps=5, num_iterations=10, gravity=(0.0, 0.0, 0.0))
    sc = S.Scene(S.finish_params(p, [c]), [c], [], 20)
    xg, _, _ = _run_gpu(sc, 20, stencil=True, expect_stencil=0)
    moved = float(np.max(np.abs(xg - curved)))
    flat = S.Cloth(**{**c.__dict__, "bend_rest": np.zeros_like(c.bend_rest)})
    sc0 = S.Scene(sc.params, [flat], [], 20)
    x0, _, _ = _run_gpu(sc0, 20)
    moved0 = float(np.max(np.abs(x0 - curved)))
    say(f"  moved {moved:.2e} with rest angles, {moved0:.2e} without")
    assert moved <= 1e-4 * sc.extent()
    assert moved0 > 100 * moved


def test_irregular_mesh_vs_fp64():
    say("irregular mesh vs fp64")
    x, tris = _irregular_horizontal(24, seed=2)
    c = S.mesh_cloth(x, tris, diameter=1.5 / 24 * 0.5, attached=(0, 24))
    p = S.Params(num_substeps=5, num_iterations=10)
    sc = S.Scene(S.finish_params(p, [c]), [c], [S.sphere((0.0, 0.9, -0.5), 0.35)], 10)
    xg, _, _ = _run_gpu(sc, 10)
    ref = oracle.simulate(sc, frames=10).x.reshape(-1, 3)
    err = float(np.max(np.abs(xg.astype(np.float64) - ref)))
    say(f"  max err {err:.3e}")
    assert err <= 1e-3 * sc.extent()


def test_add_mesh_constraints_matches_registered_arrays():
    say("add_mesh_constraints")
    x, tris = _irregular_horizontal(16, seed=6)
    c = S.mesh_cloth(x, tris, diameter=0.05)
    p = S.finish_params(S.Params(num_substeps=3, num_iterations=6), [c])
    sc = S.Scene(p, [c], [S.sphere((0.0, 1.0, -0.5), 0.3)], 5)
    xa, va, _ = _run_gpu(sc, 5)
    with ClothSolver() as s:
        s.set_params(p)
        off = s.AddCloth(x, tris, diameter=float(c.diameter))
        s.add_mesh_constraints(off, tris)
        s.UpdateColliders(sc.colliders)
        for _ in range(5):
            s.Simulate()
        s.synchronize()
        assert np.array_equal(s.positions(), xa) and np.array_equal(s.velocities(), va)


def test_viewer_readback_buffers():
    """Readbacks queued after frames return the positions and normals of exactly that frame, the
    newest completed one first, with the frame count; nothing before the first request; the
    buffer handed out is not written by later requests."""
    say("viewer readback")
    sc = S.config1(frames=4)
    with ClothSolver() as s:
        s.load_scene(sc)
        assert s.readback_latest() is None
        seen = []
        for f in range(1, 4):
            s.Simulate()
            s.readback_request()
            s.synchronize()
            seen.append((s.positions(), s.normals()))
            pos, nrm, frame = s.readback_latest()
            assert frame == f
            assert np.array_equal(pos[:, :3], seen[-1][0]) and np.array_equal(nrm[:, :3], seen[-1][1])
            assert np.array_equal(pos[:, 3], s.inv_masses())
        # two requests in a row: the newest wins; the older buffer kept the previous frame
        s.Simulate()
        s.readback_request()
        s.readback_request()
        s.synchronize()
        pos, _, frame = s.readback_latest()
        assert frame == 4 and np.array_equal(pos[:, :3], s.positions())
        # the buffer handed out stays untouched by later requests until the next readback_latest
        import ctypes as C
        from velvet_amd._lib import lib
        pp, qq = C.POINTER(C.c_float)(), C.POINTER(C.c_float)()
        nn, ff = C.c_uint32(), C.c_uint64()
        assert lib().vtx_readback_latest(s.h, 0, C.byref(pp), C.byref(qq), C.byref(nn), C.byref(ff)) == 0
        held = np.ctypeslib.as_array(pp, shape=(nn.value, 4))
        kept = held.copy()
        for _ in range(3):
            s.Simulate()
            s.readback_request()
            s.readback_request()
        s.synchronize()
        assert ff.value == 4 and np.array_equal(held, kept)
        assert s.readback_latest()[2] == 7
    # what the viewer received is the simulation's state: the FP64 oracle after 4 frames
    ref = oracle.simulate(sc, frames=4)
    err = float(np.max(np.abs(pos[:, :3].astype(np.float64) - ref.x.reshape(-1, 3))))
    say(f"  readback vs fp64 after 4 frames: {err:.3e}")
    assert err <= 1e-3 * sc.extent()


def _lists_equal(nb, rnb):
    """Column-major neighbour tables agree up to and including each list's first sentinel
    (rows after it are not part of the list)."""
    live_r = np.cumprod(rnb != SENT, axis=0).astype(bool)
    live_g = np.cumprod(nb != SENT, axis=0).astype(bool)
    assert np.array_equal(live_r, live_g)
    assert np.array_equal(nb[live_r], rnb[live_r])


def test_hash_stage_bit_exact_full_size():
    """The whole hash stage at the 1024 x 1024 size (one segment, 2^21 cells), perturbed so the
    lists are not the rest lattice: cell tables, sorted order and lists bit-exact."""
    say("hash full size 1024^2")
    sc = S.config3(frames=1)
    c = sc.cloths[0]
    sp = 2.0 / 1023.0
    p = (c.x + S.uniform(5, 77, 3 * c.n, -0.3 * sp, 0.3 * sp).reshape(-1, 3)).astype(F32)
    s = _hash_solver([(c, p)], [1 << 21])
    cell = float(F32(F32(c.diameter) * F32(1.5)))
    ph, rsh, rsi, rcs, rce, rnb = oracle.hash32(p, c.x, cell, float(c.diameter), 1 << 21)
    cs, ce, sh, si, nb = s.read_hash(0)
    s.close()
    np.testing.assert_array_equal(cs, rcs)
    valid = rcs != SENT
    np.testing.assert_array_equal(ce[valid], rce[valid])
    np.testing.assert_array_equal(sh, rsh)
    np.testing.assert_array_equal(si, rsi)
    _lists_equal(nb, rnb)
    say(f"  {int(valid.sum())} occupied cells, {int((rnb != SENT).sum())} list entries")


def test_empty_and_ragged_segments():
    """An empty solver steps; a batch mixing an empty segment, a one-particle segment, a cloth
    with no constraints, an inactive segment and cloths of different sizes gives each active
    segment its solo result, and leaves the inactive one where it was."""
    say("empty and ragged segments")
    with ClothSolver() as s:
        s.Simulate()
        s.synchronize()
        assert s.num_particles(0) == 0 and s.positions(0).shape == (0, 3)
    from tests.test_oracle_pins import tiny_scene
    # active segments share the substep loop: same substeps x iterations (and hash interleave)
    one = tiny_scene([[0.1, 1.0, 0.2]], gravity=(0.0, -9.8, 0.0), enable_self_collision=True, num_substeps=5,
                     num_iterations=10)
    base = S.config4(count=3, frames=3, smax=30)
    c = base[0].cloths[0]
    bare_c = S.Cloth(**{**c.__dict__, "stretch": np.zeros((0, 2), np.int32), "stretch_rest": np.zeros(0, F32),
                        "bend": np.zeros((0, 4), np.uint32), "bend_rest": np.zeros(0, F32),
                        "slots": np.zeros((0, 3), F32), "attach_pid": np.zeros(0, np.int32),
                        "attach_slot": np.zeros(0, np.int32), "attach_dist": np.zeros(0, F32)})
    bare = S.Scene(base[0].params, [bare_c], base[0].colliders, 3)
    scenes = [base[1], one, bare, base[2], base[0]]
    alone = [_run_gpu(sc, 3, use_graph=False)[:2] for sc in scenes[:4]]
    with ClothSolver() as s:
        segs = [s.load_scene(sc) for sc in scenes]
        empty = s.add_segment(base[0].params)
        s.set_segment_active(segs[4], False)
        x_parked = s.positions(segs[4])
        for _ in range(3):
            s.Simulate()
        s.synchronize()
        for k in range(4):
            assert np.array_equal(s.positions(segs[k]), alone[k][0]), k
            assert np.array_equal(s.velocities(segs[k]), alone[k][1]), k
        assert s.num_particles(empty) == 0
        assert np.array_equal(s.positions(segs[4]), x_parked)


def test_config3_full_size_vs_fp64():
    """The 1024 x 1024 cloth (1,048,576 particles, self-collision, 3 spheres, 4 x 8) for two frames
    against the FP64 oracle (the oracle takes ~20 s)."""
    say("config3 full size 2 frames vs fp64")
    sc = S.config3(frames=2)
    x, _, _ = _run_gpu(sc, 2, expect_stencil=1)
    ref = oracle.simulate(sc, frames=2).x.reshape(-1, 3)
    err = float(np.max(np.abs(x.astype(np.float64) - ref)))
    say(f"  max err {err:.3e} (extent {sc.extent():.3f})")
    assert err <= 1e-3 * sc.extent()
