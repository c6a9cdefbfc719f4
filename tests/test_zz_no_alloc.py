"""vtx_hash_objects_noalloc (workspace in a fixed per-device array, no allocation) against
vtx_hash_objects (workspace from hipMallocAsync) and the binary32 restatement, on two calls in a
row through the C ABI: the largest documented size first, then a smaller input that must not see
anything the first call left in the shared array."""
import numpy as np
import pytest

from tests.test_gpu_hash_objects import _folded, check_against_restatement, hash_objects, say
from velvet_amd import scenes as S

pytestmark = pytest.mark.gpu
F32 = np.float32
SENT = 0xFFFFFFFF


def same_tables(a, b):
    for x, y in zip(a[:4], b[:4]):
        np.testing.assert_array_equal(x, y)
    live_a = np.cumprod(a[4] != SENT, axis=0).astype(bool)
    live_b = np.cumprod(b[4] != SENT, axis=0).astype(bool)
    assert np.array_equal(live_a, live_b)
    assert np.array_equal(a[4][live_a], b[4][live_b])


def test_noalloc_matches_allocating_entry_point_twice():
    # call 1: 1024 x 1024 objects with 2^21 cells (56.5 MiB of the 64 MiB workspace)
    sc = S.config3(frames=1)
    c = sc.cloths[0]
    sp = 2.0 / 1023.0
    p = (c.x + S.uniform(5, 78, 3 * c.n, -0.3 * sp, 0.3 * sp).reshape(-1, 3)).astype(F32)
    d = float(c.diameter)
    cell = float(F32(F32(d) * F32(1.5)))
    M = 1 << 21
    say("hash_objects_noalloc call 1: 1024^2 objects, 2^21 cells")
    ref = hash_objects(p, c.x, cell, d, M)
    got = hash_objects(p, c.x, cell, d, M, entry="vtx_hash_objects_noalloc")
    same_tables(got, ref)
    entries = check_against_restatement(got, p, c.x, cell, d, M)
    say(f"  {entries} list entries")

    # call 2: a smaller input with a smaller, aliasing table and a lower neighbour cap
    x0, p2, d2 = _folded(33, seed=33)
    cell2 = float(F32(F32(d2) * F32(1.5)))
    say("hash_objects_noalloc call 2: folded 33^2 objects, 3000 cells, K=16")
    ref2 = hash_objects(p2, x0, cell2, d2, 3000, K=16)
    got2 = hash_objects(p2, x0, cell2, d2, 3000, K=16, entry="vtx_hash_objects_noalloc")
    same_tables(got2, ref2)
    assert check_against_restatement(got2, p2, x0, cell2, d2, 3000, K=16) > 0
