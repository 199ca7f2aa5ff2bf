"""Batched tiles (uneven n and m) through one batched solve, with and
without HIP-graph replay, against the single-tile path and the oracle.

Tolerances: kappa within 1e-12 relative (same bisection decisions) in >= 98%
of pixels, U within 1e-11 absolute, H within 1e-8 relative of the oracle;
graph replays bitwise equal to the eager batched run (fixed-order sums)."""
import numpy as np
import pytest

from imcom_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _tiles():
    specs = [(2, 8, 8, 1.1), (3, 9, 10, 1.3), (2, 12, 12, 1.5), (4, 8, 9, 1.2), (2, 10, 14, 1.0),
             (3, 11, 8, 1.6)]
    return [synth.tiny(seed=100 + i, nexp=e, npix_in=p, nout=o, psf_sigma=s, gamma_sigma=1.8)
            for i, (e, p, o, s) in enumerate(specs)]


@pytest.fixture(scope="module")
def batch_tiles():
    ts = _tiles()
    refs = [O.solve(t) for t in ts]
    return ts, refs


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_matches_oracle(gpu_ctx, batch_tiles, use_graph):
    print(f"[case] batch use_graph={use_graph}", flush=True)
    ts, refs = batch_tiles
    gpu_ctx.batch_set(ts)
    gpu_ctx.batch_solve(use_graph=use_graph)
    first = [gpu_ctx.batch_result(i) for i in range(len(ts))]
    gpu_ctx.batch_solve(use_graph=use_graph)  # replay
    for i, (t, r) in enumerate(zip(ts, refs)):
        out = gpu_ctx.batch_result(i)
        assert np.array_equal(out["K"], first[i]["K"]) and np.array_equal(out["H"], first[i]["H"])
        close = np.abs(out["K"] / r["K"] - 1.0) <= 1e-12
        assert close.mean() >= 0.98, (i, close.mean())
        assert np.max(np.abs(out["U"] - r["U"])) <= 1e-11
        assert np.max(np.abs(out["H"] - r["H"])) <= 1e-8 * np.max(np.abs(r["H"]))
        assert np.max(np.abs(out["S"] - r["S"]) / r["S"]) <= 1e-8


def test_batch_equals_single_tile_path(gpu_ctx, batch_tiles):
    print("[case] batch vs single", flush=True)
    ts, _ = batch_tiles
    gpu_ctx.batch_set(ts)
    gpu_ctx.batch_solve(use_graph=True)
    for i, t in enumerate(ts):
        b = gpu_ctx.batch_result(i, want_T=True)
        s = gpu_ctx.solve(t)
        close = np.abs(b["K"] / s["K"] - 1.0) <= 1e-12
        assert close.mean() >= 0.98
        assert np.max(np.abs(b["U"] - s["U"])) <= 1e-12
        assert np.max(np.abs(b["T"][:, close] - s["T"][:, close])) <= 1e-8 * np.max(np.abs(s["T"]))
