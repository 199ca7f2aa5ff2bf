"""Writes tests/golden/tiny_tile.npz: the oracle's outputs on the seeded tiny
tile (2 exposures of 8x8, 8x8 output), so that later changes to the host
set-up, the oracle or the GPU path are checked against fixed numbers.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from imcom_amd import synth  # noqa: E402
from oracle import oracle as O  # noqa: E402


def main():
    t = synth.tiny(seed=7, nexp=2, npix_in=8, nout=8)
    r = O.solve(t)
    rng = np.random.default_rng(0)
    ai = rng.integers(0, t.n, (64, 2))
    bi = np.stack([rng.integers(0, t.n, 64), rng.integers(0, t.m, 64)], axis=1)
    np.savez_compressed(
        os.path.join(HERE, "tiny_tile.npz"),
        seed=7, n=t.n, m=t.m, C_a=t.C_a,
        A_idx=ai, A_val=r["A"][ai[:, 0], ai[:, 1]],
        B_idx=bi, B_val=r["B"][bi[:, 0], bi[:, 1]],
        L=r["L"], K=r["K"], U=r["U"], S=r["S"], H=r["H"], evals=r["evals"])
    print("wrote tiny_tile.npz", t.n, t.m)


if __name__ == "__main__":
    main()
