"""Writes the golden fixtures of this directory from the oracle (FP64 / binary32 restatements).

    python tests/golden/make_golden.py

config1_oracle.npz   config-1 cloth (seed 0): positions after one substep (1 substep of 1/300 s),
                     after 1 frame and after 60 frames of the FP64 oracle.
hash_folded16.npz    binary32 hash stage of a folded 16x16 cloth: inputs (p, x0, cell, diameter,
                     M) and outputs (cell_start, cell_end, sorted_index, neighbour table).
These are regression fixtures produced by this repository's oracle, not reference outputs.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from velvet_amd import scenes as S  # noqa: E402


def folded16():
    n = 16
    c = S.grid_cloth(n, 0.05, center=(0.0, 1.0, 0.0), jitter=1e-4, seed=0, stream=7)
    x0 = c.x.astype(np.float64)
    p = x0.copy()
    half = (n // 2) * 0.05
    zc = x0[:, 2] - x0[:, 2].min()
    fold = zc > half
    p[fold, 2] = x0[:, 2].min() + 2 * half - zc[fold]
    p[fold, 1] += 0.04
    return c, p.astype(np.float32)


def config1_runs():
    sc = S.config1(frames=1)
    sc.params.num_substeps = 1
    dt = np.float32(S.FRAME_DT / np.float32(5))
    one_sub = oracle.simulate(sc, frames=1, frame_dt=dt).x.reshape(-1, 3).copy()
    one_frame = oracle.simulate(S.config1(frames=1), frames=1).x.reshape(-1, 3).copy()
    sixty = oracle.simulate(S.config1(frames=60), frames=60).x.reshape(-1, 3).copy()
    return one_sub, one_frame, sixty


def main():
    one_sub, one_frame, sixty = config1_runs()
    np.savez_compressed(os.path.join(HERE, "config1_oracle.npz"), one_substep=one_sub,
                        one_frame=one_frame, sixty_frames=sixty)
    c, p = folded16()
    cell = np.float32(np.float32(c.diameter) * np.float32(1.5))
    M = 2 * c.n
    ph, sh, si, cs, ce, nb = oracle.hash32(p, c.x, cell, c.diameter, M)
    np.savez_compressed(os.path.join(HERE, "hash_folded16.npz"), p=p, x0=c.x, cell=cell,
                        diameter=np.float32(c.diameter), M=np.int64(M), cell_start=cs,
                        cell_end=ce, sorted_index=si, sorted_hash=sh, nbr=nb)
    print("wrote", os.listdir(HERE))


if __name__ == "__main__":
    main()
