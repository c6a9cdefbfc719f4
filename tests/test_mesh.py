"""Constraint generation for arbitrary triangle meshes (host only: runs without a GPU).

Pinned by closed forms: edge and interior-edge counts of a grid; rest angle 0 on any flat,
consistently wound mesh; a hinge folded by a known angle; a prism whose axial edges all bend by
2*pi/N; no bending on an edge of three triangles; rest lengths are the edge lengths.
"""
import math

import numpy as np
import pytest

from velvet_amd import scenes as S

F32 = np.float32


def test_grid_counts_and_flat_rest():
    for n in (2, 3, 7, 16):
        r = n - 1
        x = np.zeros((n * n, 3), F32)
        a, b = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        x[:, 0] = b.ravel() * 0.1
        x[:, 2] = a.ravel() * 0.1
        tris = S.grid_mesh(n)
        st, sr, bd, br = S.mesh_constraints(x, tris)
        assert st.shape[0] == 3 * r * r + 2 * r                  # horizontal + vertical + diagonals
        assert bd.shape[0] == 3 * r * r + 2 * r - 4 * r          # minus the boundary edges
        assert np.all(np.abs(br) < 1e-6)
        d = np.linalg.norm(x[st[:, 0]].astype(np.float64) - x[st[:, 1]], axis=1)
        np.testing.assert_allclose(sr, d, rtol=1e-6)
        # no duplicate edges, and every bend's (p2, p3) is a mesh edge
        keys = {tuple(sorted(e)) for e in st.tolist()}
        assert len(keys) == st.shape[0]
        assert all(tuple(sorted((int(q[2]), int(q[3])))) in keys for q in bd)


@pytest.mark.parametrize("alpha", [0.0, 0.3, 1.0, 2.0, 3.0])
def test_hinge_rest_angle(alpha):
    # edge (2, 3) along x; triangle (2, 3, 0) in the xy plane, triangle (3, 2, 1) rotated about x
    x = np.array([[0.5, 1.0, 0.0], [0.5, -math.cos(alpha), math.sin(alpha)], [0, 0, 0], [1, 0, 0]], F32)
    tris = np.array([[2, 3, 0], [3, 2, 1]], np.uint32)
    st, sr, bd, br = S.mesh_constraints(x, tris)
    assert bd.tolist() == [[0, 1, 2, 3]]
    assert abs(float(br[0]) - alpha) < 2e-6


def test_prism_rest_angles():
    """A closed N-gon prism of L rings: axial edges join two faces at the exterior angle 2*pi/N;
    ring edges and quad diagonals join coplanar triangles (rest 0)."""
    N, L = 8, 4
    ang = 2 * math.pi / N
    x = np.array([[math.cos(k * ang), math.sin(k * ang), z] for z in range(L) for k in range(N)], F32)
    tris = []
    for z in range(L - 1):
        for k in range(N):
            a, b, c, d = z * N + k, z * N + (k + 1) % N, (z + 1) * N + k, (z + 1) * N + (k + 1) % N
            tris += [(a, b, d), (a, d, c)]
    st, sr, bd, br = S.mesh_constraints(x, np.array(tris, np.uint32))
    kinds = {"axial": [], "ring": [], "diag": []}
    for q, r in zip(bd, br):
        (z2, k2), (z3, k3) = divmod(int(q[2]), N), divmod(int(q[3]), N)
        kind = "axial" if k2 == k3 else "ring" if z2 == z3 else "diag"
        kinds[kind].append(float(r))
    assert len(kinds["axial"]) == N * (L - 1)
    assert len(kinds["ring"]) == N * (L - 2)
    assert len(kinds["diag"]) == N * (L - 1)
    np.testing.assert_allclose(kinds["axial"], ang, atol=1e-5)
    assert np.all(np.abs(kinds["ring"] + kinds["diag"]) < 1e-5)


def test_three_triangles_on_one_edge_get_no_bend():
    x = np.array([[0, 0, 0], [1, 0, 0], [0.5, 1, 0], [0.5, -1, 0], [0.5, 0, 1]], F32)
    tris = np.array([[0, 1, 2], [1, 0, 3], [0, 1, 4]], np.uint32)
    st, sr, bd, br = S.mesh_constraints(x, tris)
    assert st.shape[0] == 7
    assert all(sorted((int(q[2]), int(q[3]))) != [0, 1] for q in bd)


def test_capacity_and_bad_index_errors():
    import ctypes as C
    from velvet_amd._lib import lib
    x = np.zeros((4, 3), F32)
    x[1, 0] = x[2, 1] = 1.0
    x[3] = (1, 1, 0)
    tris = np.array([[0, 1, 2], [2, 1, 3]], np.uint32)
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    up = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint32))
    ns, nb = C.c_uint32(), C.c_uint32()
    small = np.empty((1, 2), np.int32)
    rc = lib().vtx_mesh_constraints(fp(x), 4, up(tris), 2, small.ctypes.data_as(C.POINTER(C.c_int32)), None, 1,
                                    C.byref(ns), None, None, 0, C.byref(nb))
    assert rc != 0 and ns.value == 5 and nb.value == 1
    bad = np.array([[0, 1, 9]], np.uint32)
    rc = lib().vtx_mesh_constraints(fp(x), 4, up(bad), 1, None, None, 0, C.byref(ns), None, None, 0, C.byref(nb))
    assert rc != 0


def test_irregular_mesh_restatement():
    x, tris = S.irregular_mesh(12, seed=3, stream=7)
    assert x.shape == (169, 3) and tris.shape == (2 * 144, 3)
    x2, tris2 = S.irregular_mesh(12, seed=3, stream=7)
    assert np.array_equal(x, x2) and np.array_equal(tris, tris2)
    # boundary vertices stay on the regular grid, interior ones move by less than 0.4 / res * 2
    m = 13
    for k in range(m * m):
        yy, xx = divmod(k, m)
        reg = np.array([2 * (xx / 12 - 0.5), -2 * yy / 12, 0])
        dev = np.abs(x[k] - reg).max()
        if xx in (0, 12) or yy in (0, 12):
            assert dev < 1e-6
        else:
            assert dev < 2 * 0.4 / 12 + 1e-6
    # one winding for every triangle, so a flat mesh has all bending rest angles 0
    e1 = x[tris[:, 1]] - x[tris[:, 0]]
    e2 = x[tris[:, 2]] - x[tris[:, 0]]
    nz = np.cross(e1, e2)[:, 2]
    assert np.all(nz > 0) or np.all(nz < 0)
    st, sr, bd, br = S.mesh_constraints(x, tris)
    assert np.all(np.abs(br) < 1e-5)
