"""vtx_hash_objects, the kernel-launch tier of the spatial hash (HashObjects on raw device arrays,
Velvet/SpatialHashGPU.cu:159-196), bit-exact against the binary32 restatement: sorted hashes and
indices, cell tables and neighbour lists, for a folded cloth, caps and bucket aliasing, the
1024 x 1024 size and no objects at all."""
import ctypes as C
import sys

import numpy as np
import pytest

import oracle
from tests.hipmem import DeviceArray
from velvet_amd import scenes as S
from velvet_amd._lib import VtxHashParams, check, lib

pytestmark = pytest.mark.gpu
F32 = np.float32
SENT = 0xFFFFFFFF


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def hash_objects(p, x0, cell, diameter, M, K=64, entry="vtx_hash_objects", keep=None):
    """Run an entry point of the HashObjects form on device copies; returns host arrays
    (sorted hash, sorted index, cell start, cell end, neighbours (K, n))."""
    p = np.ascontiguousarray(p, F32)
    x0 = np.ascontiguousarray(x0, F32)
    n = p.shape[0]
    dp, dx = DeviceArray(p), DeviceArray(x0)
    dh = DeviceArray(shape=(n,), dtype=np.uint32, fill=0x5A)
    di = DeviceArray(shape=(n,), dtype=np.uint32, fill=0x5A)
    ds = DeviceArray(shape=(M,), dtype=np.uint32, fill=0x5A)
    de = DeviceArray(shape=(M,), dtype=np.uint32, fill=0x5A)
    dn = DeviceArray(shape=(K, n), dtype=np.uint32, fill=0x5A)
    prm = VtxHashParams(num_objects=n, max_num_neighbors=K, cell_spacing=F32(cell),
                        cell_spacing2=F32(F32(cell) * F32(cell)), table_size=M,
                        particle_diameter2=F32(F32(diameter) * F32(diameter)))
    fn = getattr(lib(), entry)
    check(fn(dh.ptr, di.ptr, ds.ptr, de.ptr, dn.ptr, dp.ptr, dx.ptr, C.byref(prm), None))
    out = dh.get(), di.get(), ds.get(), de.get(), dn.get()
    if keep is not None:
        keep.extend([dp, dx, dh, di, ds, de, dn])
    return out


def check_against_restatement(out, p, x0, cell, diameter, M, K=64):
    sh, si, cs, ce, nb = out
    ph, rsh, rsi, rcs, rce, rnb = oracle.hash32(p, x0, cell, diameter, M, K)
    np.testing.assert_array_equal(sh, rsh)
    np.testing.assert_array_equal(si, rsi)
    np.testing.assert_array_equal(cs, rcs)
    empty = rcs == SENT
    np.testing.assert_array_equal(ce[~empty], rce[~empty])
    assert np.all(ce[empty] == SENT)
    live = np.cumprod(rnb != SENT, axis=0).astype(bool)
    assert np.array_equal(np.cumprod(nb != SENT, axis=0).astype(bool), live)
    assert np.array_equal(nb[live], rnb[live])
    return int(live.sum())


def _folded(n, seed, lift=0.04):
    c = S.grid_cloth(n, 0.05, center=(0.0, 1.0, 0.0), jitter=1e-4, seed=seed, stream=7)
    x0 = c.x.astype(np.float64)
    p = x0.copy()
    half = (n // 2) * 0.05
    zc = x0[:, 2] - x0[:, 2].min()
    fold = zc > half
    p[fold, 2] = x0[:, 2].min() + 2 * half - zc[fold]
    p[fold, 1] += lift
    return c.x, p.astype(F32), float(c.diameter)


@pytest.mark.parametrize("n,M,K", [(16, 0, 64), (33, 3000, 64), (40, 1 << 16, 32), (24, 97, 8)])
def test_hash_objects_folded(n, M, K):
    x0, p, d = _folded(n, seed=n)
    M = M or 2 * n * n
    cell = float(F32(F32(d) * F32(1.5)))
    say(f"hash_objects folded n={n * n} M={M} K={K}")
    entries = check_against_restatement(hash_objects(p, x0, cell, d, M, K), p, x0, cell, d, M, K)
    assert entries > 0


def test_hash_objects_caps_and_aliasing():
    n = 80
    x0 = np.stack([np.arange(n) * 1.0, np.zeros(n), np.zeros(n)], 1).astype(F32)
    p = np.tile(np.array([[0.01, 0.01, 0.01]], F32), (n, 1))
    for M in (2 * n, 1, 5):
        say(f"hash_objects caps M={M}")
        check_against_restatement(hash_objects(p, x0, 0.1, 0.05, M), p, x0, 0.1, 0.05, M)


def test_hash_objects_full_size():
    sc = S.config3(frames=1)
    c = sc.cloths[0]
    sp = 2.0 / 1023.0
    p = (c.x + S.uniform(5, 78, 3 * c.n, -0.3 * sp, 0.3 * sp).reshape(-1, 3)).astype(F32)
    cell = float(F32(F32(c.diameter) * F32(1.5)))
    say("hash_objects 1024^2, 2^21 cells")
    entries = check_against_restatement(hash_objects(p, c.x, cell, float(c.diameter), 1 << 21), p, c.x, cell,
                                        float(c.diameter), 1 << 21)
    say(f"  {entries} list entries")


def test_hash_objects_no_objects():
    say("hash_objects n=0")
    p = np.zeros((0, 3), F32)
    sh, si, cs, ce, nb = hash_objects(p, p, 0.1, 0.05, 7)
    assert np.all(cs == SENT) and np.all(ce == SENT)
