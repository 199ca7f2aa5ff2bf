"""The oracle against the committed golden outputs of the seeded tiny tile
(tests/golden/tiny_tile.npz, written by tests/golden/make_golden.py)."""
import os

import numpy as np

from imcom_amd import synth
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tiny_tile.npz")


def test_oracle_matches_golden():
    g = np.load(GOLD, allow_pickle=False)
    t = synth.tiny(seed=int(g["seed"]), nexp=2, npix_in=8, nout=8)
    assert (t.n, t.m) == (int(g["n"]), int(g["m"]))
    assert abs(t.C_a - float(g["C_a"])) <= 1e-15 * t.C_a
    r = O.solve(t)
    ai, bi = g["A_idx"], g["B_idx"]
    assert np.allclose(r["A"][ai[:, 0], ai[:, 1]], g["A_val"], rtol=1e-13, atol=1e-18)
    assert np.allclose(r["B"][bi[:, 0], bi[:, 1]], g["B_val"], rtol=1e-13, atol=1e-18)
    scale = np.max(np.abs(g["L"]))
    assert np.max(np.abs(r["L"] - g["L"])) <= 1e-13 * scale
    assert np.allclose(r["K"], g["K"], rtol=1e-12)
    assert np.max(np.abs(r["U"] - g["U"])) <= 1e-13
    assert np.allclose(r["S"], g["S"], rtol=1e-10)
    assert np.allclose(r["H"], g["H"], rtol=1e-9, atol=1e-15)
    assert np.array_equal(r["evals"], g["evals"])
