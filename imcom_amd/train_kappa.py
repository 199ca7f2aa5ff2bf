"""Train the learned kappa-bracket network on the GPU and store its weights.

  python -m imcom_amd.train_kappa [--small-tiles 100] [--large-tiles 4]
                                  [--batch-tiles 32] [--epochs 30]

Training pairs come from the plain bisection on seeded variants: small tiles
(4 x 16^2 inputs, PSF sigma in [1, 2]), tiles of the batched workload's
shape family (2-6 exposures of 16-32 px squares, PSF sigma in [1.0, 1.6],
noise log-uniform in [1e-4, 1e-2]; seeds disjoint from the benchmark's) and,
for the large workload, variants of the large tile (4 x 64^2, Airy x pixel
PSF, lambda/D in [0.9, 1.1]), each at several leakage targets log-uniform in
[1e-8, 1e-4].  Held-out tiles measure
the gain (U evaluations per pixel, bracket hit rate) and check that every
pixel passes the same leakage test with and without the bracket.  Writes
imcom_amd/data/kappa_mlp.npy and a JSON summary.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import kappa_net, synth
from .gpu import ImcomGPU


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--small-tiles", type=int, default=100)
    ap.add_argument("--large-tiles", type=int, default=4)
    ap.add_argument("--batch-tiles", type=int, default=32)
    ap.add_argument("--targets", type=int, default=8)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default=kappa_net.WEIGHTS)
    ap.add_argument("--summary", default=None)
    a = ap.parse_args(argv)
    g = ImcomGPU(0)
    t0 = time.time()
    rng = np.random.default_rng(20261210)
    Xs, ys = [], []
    for k in range(a.small_tiles):
        t = kappa_net.variant_tile(30000 + k)
        X, y = kappa_net.bisection_pairs(g, t, 10.0 ** rng.uniform(-8, -4, a.targets))
        Xs.append(X)
        ys.append(y)
        if k % 10 == 0:
            print(f"small tile {k}: {sum(len(v) for v in ys)} pairs, {time.time() - t0:.0f}s", flush=True)
    for k, t in enumerate(synth.batch(ntiles=a.batch_tiles, seed=20261300) if a.batch_tiles else []):
        X, y = kappa_net.bisection_pairs(g, t, 10.0 ** rng.uniform(-8, -4, max(2, a.targets // 2)))
        Xs.append(X)
        ys.append(y)
        if k % 8 == 0:
            print(f"batch-shape tile {k} (n={t.n}): {sum(len(v) for v in ys)} pairs, "
                  f"{time.time() - t0:.0f}s", flush=True)
    for k in range(a.large_tiles):
        lod = float(rng.uniform(0.9, 1.1))
        t = synth.large(seed=40000 + k, lam_over_d=lod)
        X, y = kappa_net.bisection_pairs(g, t, 10.0 ** rng.uniform(-8, -4, max(2, a.targets // 2)))
        Xs.append(X)
        ys.append(y)
        print(f"large tile {k}: {sum(len(v) for v in ys)} pairs, {time.time() - t0:.0f}s", flush=True)
    X = np.concatenate(Xs)
    y = np.concatenate(ys)
    print(f"dataset: {len(y)} pairs in {time.time() - t0:.0f}s", flush=True)
    t1 = time.time()
    p, losses = kappa_net.train(g, X, y, epochs=a.epochs, batch=a.batch, lr=a.lr, seed=1)
    print(f"trained {a.epochs} epochs in {time.time() - t1:.1f}s: loss {losses[0]:.3f} -> "
          f"{losses[-1]:.3f}", flush=True)
    kappa_net.save_weights(p, a.out)
    evals = []
    for s, U_t in ((9001, 1e-6), (9002, 1e-7), (9003, 1e-5), (9004, 3e-8)):
        g.mlp_set_params(p)
        c = kappa_net.compare(g, kappa_net.variant_tile(s), U_t)
        c.update(tile=f"small variant {s}", U_target=U_t)
        evals.append(c)
        print(c, flush=True)
    for i, t in enumerate(synth.batch(ntiles=4, seed=20261400)):
        g.mlp_set_params(p)
        c = kappa_net.compare(g, t, 1e-6)
        c.update(tile=f"batch-shape held-out {i} (n={t.n})", U_target=1e-6)
        evals.append(c)
        print(c, flush=True)
    g.mlp_set_params(p)
    c = kappa_net.compare(g, synth.large(), 1e-6)
    c.update(tile="large", U_target=1e-6)
    evals.append(c)
    print(c, flush=True)
    summary = dict(pairs=int(len(y)), epochs=a.epochs, loss_first=float(losses[0]),
                   loss_last=float(losses[-1]), heldout=evals,
                   seconds=round(time.time() - t0, 1))
    if a.summary:
        with open(a.summary, "w") as f:
            json.dump(summary, f, indent=1)
    g.close()
    return summary


if __name__ == "__main__":
    main(sys.argv[1:])
