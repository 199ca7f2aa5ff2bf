"""Learned kappa bracket: training data from the plain bisection, training on
the GPU through the library's own forward/backward/Adam kernels, and the
evaluation of what the bracket saves.

Training pairs are (features, log10 kappa) produced by the plain full-range
bisection itself (imcom_bisect.f90:174-222 restated on the GPU) on seeded
variants of a tile shape, with the PSF width and the leakage target varied.
The network only narrows the start of that same bisection; every pixel still
ends under the same U (or S) test, and a bracket that misses falls back to
the full range inside the bisection kernel.
"""
from __future__ import annotations

import os

import numpy as np

from . import synth

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "data", "kappa_mlp.npy")

NF, NH, NO = 32, 64, 2
N_TRAIN = NF * NH + NH + NH * NH + NH + NH * NO + NO
N_PARAMS = N_TRAIN + 2 * NF + 2


def init_params(X, y, seed=0):
    """Glorot-uniform weights, zero biases, output layer scaled down so the
    untrained network predicts the mean with unit spread; standardisation of
    features and target from the data."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, dtype=np.float64)
    o = 0
    for fi, fo, scale in ((NF, NH, 1.0), (NH, NH, 1.0), (NH, NO, 0.1)):
        lim = scale * np.sqrt(6.0 / (fi + fo))
        p[o:o + fi * fo] = rng.uniform(-lim, lim, fi * fo)
        o += fi * fo + fo
    assert o == N_TRAIN
    xm = X.mean(0)
    xs = np.maximum(X.std(0), 1e-3)
    p[N_TRAIN:N_TRAIN + NF] = xm
    p[N_TRAIN + NF:N_TRAIN + 2 * NF] = xs
    p[N_TRAIN + 2 * NF] = y.mean()
    p[N_TRAIN + 2 * NF + 1] = max(float(y.std()), 1e-3)
    return p.astype(np.float32)


def variant_tile(seed, shape="small", psf_sigma=None):
    rng = np.random.default_rng(seed)
    sig = float(rng.uniform(1.0, 2.0)) if psf_sigma is None else psf_sigma
    if shape == "small":
        return synth.small(seed=seed, psf_sigma=sig, gamma_sigma=max(1.5, sig + 0.3))
    if shape == "tiny":
        return synth.tiny(seed=seed, psf_sigma=sig, gamma_sigma=max(1.5, sig + 0.3))
    raise ValueError(shape)


def bisection_pairs(g, tile, targets, with_plain_stats=False):
    """(features, log10 kappa) pairs of one tile from the plain bisection at
    each leakage target.  Pixels whose leakage cannot reach the target even
    at kappa_min are skipped."""
    g.set_tile(tile)
    g.lookup_A(tile.npoly, tile.npad)
    g.eigen_A(out=False)
    g.lookup_B(tile.npoly, tile.npad)
    g.build_P()
    Xs, ys, evs = [], [], []
    for U_t in targets:
        X = g.kappa_features(U_t)
        K, U, ev = g.bisect_kappa(US=U_t, tol=1e-3 * U_t)
        ok = (U < U_t) & (K > tile.kappa_min)
        Xs.append(X[ok])
        ys.append(np.log10(K[ok]).astype(np.float32))
        evs.append(ev[ok])
    X = np.concatenate(Xs)
    y = np.concatenate(ys)
    if with_plain_stats:
        return X, y, np.concatenate(evs)
    return X, y


def make_dataset(g, ntiles, targets_per_tile=8, seed0=1000, shape="small"):
    rng = np.random.default_rng(seed0)
    Xs, ys = [], []
    for k in range(ntiles):
        t = variant_tile(seed0 + k, shape)
        targets = 10.0 ** rng.uniform(-8.0, -4.0, targets_per_tile)
        X, y = bisection_pairs(g, t, targets)
        Xs.append(X)
        ys.append(y)
    return np.concatenate(Xs), np.concatenate(ys)


def train(g, X, y, epochs=30, batch=2048, lr=2e-3, seed=0, params=None):
    p = init_params(X, y, seed) if params is None else np.asarray(params, dtype=np.float32)
    g.mlp_set_params(p)
    losses = g.mlp_train(X, y, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return g.mlp_get_params(), losses


def compare(g, tile, U_t, zsig=3.0):
    """Plain vs learned bisection on one tile and target: mean U/S
    evaluations per pixel, bracket hit rate, and the leakage test of every
    pixel under both."""
    g.set_tile(tile)
    g.lookup_A(tile.npoly, tile.npad)
    g.eigen_A(out=False)
    g.lookup_B(tile.npoly, tile.npad)
    g.build_P()
    tol = 1e-3 * U_t
    K0, U0, e0 = g.bisect_kappa(US=U_t, tol=tol)
    K1, U1, e1, hits = g.bisect_kappa_learned(zsig=zsig, US=U_t, tol=tol)
    reach = U0 < U_t
    return dict(evals_plain=float(e0.mean()), evals_learned=float(e1.mean()),
                hit_rate=float(hits.mean()), all_U_ok_plain=bool(np.all(U0[reach] < U_t)),
                all_U_ok_learned=bool(np.all(U1[reach] < U_t)),
                max_dlogK=float(np.max(np.abs(np.log10(K1[reach]) - np.log10(K0[reach]))))
                if reach.any() else 0.0)


def load_weights(path=WEIGHTS):
    if not os.path.exists(path):
        return None
    return np.load(path, allow_pickle=False).astype(np.float32)


def save_weights(p, path=WEIGHTS):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.save(path, np.asarray(p, dtype=np.float32), allow_pickle=False)
