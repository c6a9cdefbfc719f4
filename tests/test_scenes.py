"""The seeded scene generators reproduce the reference's topology and counts."""
import numpy as np
import pytest

from velvet_amd import scenes as S


@pytest.mark.parametrize("n", [2, 3, 16, 40, 41])
def test_grid_counts_closed_form(n):
    c = S.grid_cloth(n, 0.05, attached=(0, n - 1), bend_mode="quad")
    assert c.stretch.shape[0] == 2 * n * (n - 1) + 2 * (n - 1) ** 2
    assert c.tris.shape[0] == 2 * (n - 1) ** 2
    assert c.bend.shape[0] == (n - 1) ** 2
    assert c.attach_pid.shape[0] == 2 * n * n
    p = S.grid_cloth(n, 0.05, bend_mode="pairs")
    # interior edges of the triangulated grid: 3 (n-1)^2 + 2 (n-1) edges in total, minus the
    # 4 (n-1) boundary edges
    assert p.bend.shape[0] == 3 * (n - 1) ** 2 - 2 * (n - 1)


def test_stretch_order_matches_reference_loops():
    for n in (3, 7, 12):
        np.testing.assert_array_equal(S._grid_stretch_fast(n), S.grid_stretch(n))


@pytest.mark.parametrize("n", [41, 64])
def test_bend_pairs_vectorised_matches_generic(n):
    np.testing.assert_array_equal(S.bend_pairs_grid(n), S.bend_pairs(S.grid_mesh(n)))


def test_bend_pairs_flat_normals_agree():
    c = S.grid_cloth(9, 0.1, jitter=0.0, bend_mode="pairs")
    x = c.x.astype(np.float64)
    b = c.bend.astype(int)
    p0, p1, p2, p3 = (x[b[:, k]] for k in range(4))
    n1 = np.cross(p2 - p0, p3 - p0)
    n2 = np.cross(p3 - p1, p2 - p1)
    cos = (n1 * n2).sum(1) / np.linalg.norm(n1, axis=1) / np.linalg.norm(n2, axis=1)
    np.testing.assert_allclose(cos, 1.0, atol=1e-12)
    q = S.grid_cloth(9, 0.1, jitter=0.0, bend_mode="quad")
    b = q.bend.astype(int)
    p0, p1, p2, p3 = (x[b[:, k]] for k in range(4))
    n1 = np.cross(p2 - p0, p3 - p0)
    n2 = np.cross(p3 - p1, p2 - p1)
    cos = (n1 * n2).sum(1) / np.linalg.norm(n1, axis=1) / np.linalg.norm(n2, axis=1)
    np.testing.assert_allclose(cos, 1.0, atol=1e-12)


def test_quad_bend_is_reference_index_pattern():
    t = S.grid_mesh(3).reshape(-1)
    b = S.bend_per_quad(S.grid_mesh(3))
    np.testing.assert_array_equal(b[0], [t[0], t[5], t[2], t[1]])


def test_seeded_and_diameter():
    a = S.config1()
    b = S.config1()
    np.testing.assert_array_equal(a.cloths[0].x, b.cloths[0].x)
    c = a.cloths[0]
    d = np.float32(np.sqrt(np.float32(((c.x[0] - c.x[1]).astype(np.float32) ** 2).sum()))) * np.float32(1.5)
    assert abs(float(c.diameter) - float(d)) < 1e-8
    assert a.params.max_speed == pytest.approx(2 * float(c.diameter) * 60 * 5, rel=1e-6)
    j = c.x - S.grid_cloth(40, 0.05, jitter=0.0).x
    assert np.abs(j).max() <= 1.0001e-4 and np.abs(j).max() > 5e-5


def test_config_counts():
    c1 = S.config1().counts()
    assert c1 == dict(N=1600, S=6162, B=4485, A=3200, L=2, T=3042, C=1)
    sides, spacing, stiff, nat, nsph = S.batch_cloth_params(512)
    assert sides.min() >= 16 and sides.max() <= 128
    assert stiff.min() >= 0.5 and stiff.max() <= 1.0
    assert set(nat.tolist()) <= {0, 1, 2, 3, 4} and set(nsph.tolist()) <= {1, 2, 3}
