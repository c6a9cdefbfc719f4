"""Launch variants and the tuner on the GPU through the C ABI.

* every value of the tuning knobs (graph replay, neighbour-search thread order, dense attachment
  batch, XCD remap, gather-kernel block size), alone and mixed, gives bit-identical positions and
  velocities, on a batch that has
  grid cloths (patch order) and on one whose segments are general meshes (two cloths sharing a
  segment);
* vtx_autotune leaves the simulation state where it was (trials are undone) and a later run
  under the tuned variant equals an untuned run bit for bit;
* a second solver of the same shape class takes the cached tuning without timing;
* host operations queued between frames (slot moves, colliders, particle edits, readback,
  snapshot / restore) give the same results with 1 to 4 stream groups;
* on a caller-bound stream, a copy queued right after a frame sees every stream group's work.
"""
import sys

import numpy as np
import pytest

from velvet_amd import scenes as S
from velvet_amd.solver import ClothSolver

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def _batch():
    # self-collision, 0-4 dense attachment slots, 1-3 spheres per cloth
    return S.config4(count=6, frames=4, smax=48)


def _general():
    # two cloths in one segment: a shared hash and the general (incidence-list) path
    sc = S.config1(frames=4)
    c2 = S.grid_cloth(12, 0.05, center=(0.0, 1.55, 0.0), seed=5, stream=9)
    return [S.Scene(sc.params, [sc.cloths[0], c2], sc.colliders, 4)]


def _run(scenes, frames, **tuning):
    with ClothSolver() as s:
        for sc in scenes:
            s.load_scene(sc)
        if tuning:
            s.set_tuning(**tuning)
        for _ in range(frames):
            s.Simulate()
        s.synchronize()
        return [(s.positions(k), s.velocities(k)) for k in range(len(scenes))]


# every value of every knob on its own, plus mixed combinations
VARIANTS = ([dict(use_graph=g) for g in (0, 1)] + [dict(neighbor_order=o) for o in (0, 1, 2)]
            + [dict(lra_batch=b) for b in (4, 16, 64)] + [dict(xcd_remap=r) for r in (0, 1)]
            + [dict(gather_block=b) for b in (64, 128, 256)] + [dict(streams=k) for k in (1, 2, 3, 4)] + [dict(hash_lds=h) for h in (0, 1)]
            + [dict(use_graph=1, neighbor_order=2, lra_batch=4, xcd_remap=0, gather_block=64, streams=2),
               dict(use_graph=0, neighbor_order=0, lra_batch=64, xcd_remap=1, gather_block=128, streams=4),
               dict(use_graph=1, neighbor_order=1, lra_batch=16, xcd_remap=0, gather_block=64, streams=1, hash_lds=0)])


@pytest.mark.parametrize("which", ["grid batch", "general"])
def test_variants_bit_identical(which):
    scenes = _batch() if which == "grid batch" else _general()
    ref = _run(scenes, 4)
    for v in VARIANTS:
        say(f"{which}: {v}")
        got = _run(scenes, 4, **v)
        for k, ((x, vel), (xr, vr)) in enumerate(zip(got, ref)):
            assert np.array_equal(x, xr) and np.array_equal(vel, vr), (k, v)


def test_variants_bit_identical_long_range_attachments():
    say("variants on config 2 (256 dense slots)")
    sc = S.config2(frames=2)
    ref = _run([sc], 2)
    for lb, split in ((4, 1), (64, 1), (4, 0), (16, 0), (64, 0)):
        say(f"  lra_batch={lb} lra_split={split}")
        got = _run([sc], 2, lra_batch=lb, lra_split=split, neighbor_order=0, xcd_remap=0)
        assert np.array_equal(got[0][0], ref[0][0]) and np.array_equal(got[0][1], ref[0][1])


def test_lra_split_in_a_mixed_batch():
    """A grid cloth with 40 dense slots (four threads per particle) batched with small cloths
    (one thread per particle), on one or two streams: each cloth matches its own solo run."""
    say("lra_split in a mixed batch")
    big = S.grid_cloth(40, 0.05, center=(0.0, 1.5, 0.0), seed=0, stream=1, attached=tuple(range(40)))
    small = _batch()[:3]
    lra = S.Scene(small[0].params, [big], small[0].colliders, 3, "lra40")
    scenes = [small[0], lra, small[1], small[2]]
    alone = [_run([sc], 3)[0] for sc in scenes]
    for v in (dict(), dict(lra_split=0), dict(streams=2), dict(streams=3, lra_split=0, lra_batch=4)):
        say(f"  {v}")
        got = _run(scenes, 3, **v)
        for k in range(len(scenes)):
            assert np.array_equal(got[k][0], alone[k][0]) and np.array_equal(got[k][1], alone[k][1]), (k, v)


def _run_with_host_ops(scenes, **tuning):
    """Frames without a synchronize between them, with every kind of host operation that
    touches device state queued between two frames."""
    with ClothSolver(use_graph=False) as s:
        segs = [s.load_scene(sc) for sc in scenes]
        s.set_tuning(**tuning)
        snap = None
        for f in range(6):
            s.Simulate()
            if f == 0:
                for k, sc in enumerate(scenes):
                    if sc.cloths[0].slots.shape[0]:
                        s.SetAttachSlotPosition(0, sc.cloths[0].slots[0] + np.float32([0.05, 0.0, 0.0]), seg=segs[k])
            elif f == 1:
                s.UpdateColliders(list(scenes[-1].colliders) + [S.sphere((0.0, 0.6, 0.0), 0.3)], seg=segs[-1])
            elif f == 2:
                s.set_particle(3, pos=[0.0, 1.2, 0.0], vel=[0.0, 0.0, 0.0], seg=segs[-1])
            elif f == 3:
                s.readback_request(segs[2])
                snap = s.snapshot(segs[4])
            elif f == 4:
                s.restore(snap, seg=segs[4])
        s.synchronize()
        rb = s.readback_latest(segs[2])
        return [(s.positions(k), s.velocities(k)) for k in segs], snap, rb


def test_host_operations_between_frames_with_stream_groups():
    """On the solver's own stream the stream groups are joined lazily; every host operation
    between frames must still see (and be seen by) all of them, as with a single stream."""
    scenes = _batch()
    ref, snap_ref, rb_ref = _run_with_host_ops(scenes, streams=1)
    for k in (2, 3, 4):
        say(f"host operations between frames, streams={k}")
        got, snap, rb = _run_with_host_ops(scenes, streams=k)
        assert snap == snap_ref, k
        assert rb is not None and rb_ref is not None and rb[2] == rb_ref[2]
        assert np.array_equal(rb[0], rb_ref[0]) and np.array_equal(rb[1], rb_ref[1]), k
        for j, ((x, v), (xr, vr)) in enumerate(zip(got, ref)):
            assert np.array_equal(x, xr) and np.array_equal(v, vr), (k, j)


def test_bound_stream_sees_every_group():
    """On a stream the caller bound, work queued after vtx_simulate sees the whole frame, also
    the segments of the extra stream groups (no lazy join there)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    stream = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(stream)) == 0
    scenes = _batch()
    try:
        with ClothSolver(use_graph=False) as s:
            segs = [s.load_scene(sc) for sc in scenes]
            s.set_tuning(streams=2)
            s.set_stream(stream.value)
            last = segs[-1]                      # in the second group
            ptr, n = s.device_positions(last)    # uploads; the address stays valid (no registration follows)
            host = C.c_void_p()
            assert hip.hipHostMalloc(C.byref(host), C.c_size_t(16 * n), 0) == 0
            try:
                got = []
                for f in range(3):
                    s.Simulate()
                    # a copy queued on the bound stream right after the frame, no vtx_synchronize
                    assert hip.hipMemcpyAsync(host, C.c_void_p(ptr), C.c_size_t(16 * n), 2, stream) == 0
                    assert hip.hipStreamSynchronize(stream) == 0
                    got.append(np.ctypeslib.as_array(C.cast(host, C.POINTER(C.c_float)), shape=(n, 4))[:, :3].copy())
                    s.synchronize()
                    assert np.array_equal(got[-1], s.positions(last)), f
            finally:
                hip.hipHostFree(host)
            s.set_stream(None)
    finally:
        hip.hipStreamDestroy(stream)
    ref = _run(scenes, 3)
    assert np.array_equal(got[-1], ref[-1][0])


def test_autotune_keeps_state_and_caches():
    say("autotune")
    scenes = _batch()
    ref = _run(scenes, 3)
    with ClothSolver() as s:
        for sc in scenes:
            s.load_scene(sc)
        s.Simulate()
        s.synchronize()
        x0 = [s.positions(k) for k in range(len(scenes))]
        n0 = [s.normals(k) for k in range(len(scenes))]
        s.readback_request(0)
        s.synchronize()
        t = s.autotune(frames=2)
        say(f"  tuned {t}")
        assert t["frame_ms"] > 0 and t["lra_batch"] in (4, 16, 64)
        for k in range(len(scenes)):
            assert np.array_equal(s.positions(k), x0[k]) and np.array_equal(s.normals(k), n0[k])
        s.readback_request(0)
        s.synchronize()
        assert s.readback_latest(0)[2] == 1          # tuning frames are not counted
        for _ in range(2):
            s.Simulate()
        s.synchronize()
        for k in range(len(scenes)):
            assert np.array_equal(s.positions(k), ref[k][0]) and np.array_equal(s.velocities(k), ref[k][1])
    with ClothSolver() as s:
        for sc in scenes:
            s.load_scene(sc)
        c = s.autotune(frames=2, use_cache=True)
        assert c["frame_ms"] == 0.0
        assert {k: c[k] for k in ClothSolver._TUNING_KEYS} == {k: t[k] for k in ClothSolver._TUNING_KEYS}


def test_tuning_rejects_bad_values():
    with ClothSolver() as s:
        with pytest.raises(Exception):
            s.set_tuning(lra_batch=7)
        with pytest.raises(Exception):
            s.set_tuning(neighbor_order=3)
        with pytest.raises(Exception):
            s.set_tuning(gather_block=96)
        with pytest.raises(Exception):
            s.set_tuning(streams=5)
        for bad in (-1, 4):
            with pytest.raises(Exception):
                s.set_tuning(hash_grid=bad)
        s.set_tuning(lra_split=0, hash_grid=3)
        assert s.tuning()["lra_split"] == 0 and s.tuning()["hash_grid"] == 3
        with pytest.raises(Exception):   # tune_hash_grid is 0 or 1
            s.add_segment(S.Params(tune_hash_grid=2))
