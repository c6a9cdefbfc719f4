"""The oracle reproduces the committed golden fixtures (regression guard)."""
import os

import numpy as np

import oracle
from tests.golden.make_golden import config1_runs, folded16

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_config1_oracle_golden():
    g = np.load(os.path.join(HERE, "config1_oracle.npz"))
    one_sub, one_frame, sixty = config1_runs()
    np.testing.assert_allclose(one_sub, g["one_substep"], atol=1e-12)
    np.testing.assert_allclose(one_frame, g["one_frame"], atol=1e-12)
    np.testing.assert_allclose(sixty, g["sixty_frames"], atol=1e-9)


def test_hash_golden():
    g = np.load(os.path.join(HERE, "hash_folded16.npz"))
    c, p = folded16()
    np.testing.assert_array_equal(p, g["p"])
    ph, sh, si, cs, ce, nb = oracle.hash32(g["p"], g["x0"], float(g["cell"]), float(g["diameter"]), int(g["M"]))
    np.testing.assert_array_equal(cs, g["cell_start"])
    valid = cs != 0xFFFFFFFF
    np.testing.assert_array_equal(ce[valid], g["cell_end"][valid])
    np.testing.assert_array_equal(si, g["sorted_index"])
    assert oracle.neighbor_lists(nb) == oracle.neighbor_lists(g["nbr"])
