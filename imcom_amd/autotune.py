"""Autotuner for the eigensolver's blocking: the tridiagonal-reduction panel
width nb and the back-transform block nbb.

For each size class it times eigen_A (device time of the reduction and the
back-transform, HIP events) on a seeded SPD matrix of the class's size for
every candidate (nb, nbb), keeps the fastest, and stores the table in
imcom_amd/data/tuning.json.  ImcomGPU users apply it with apply(g, n).

  python -m imcom_amd.autotune [--classes 1024,2048,4096,8192,16384] [--out PATH]
"""
from __future__ import annotations

import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "data", "tuning.json")

CLASSES = [1024, 2048, 4096, 8192, 16384]          # upper bounds of the n classes
CANDIDATES = [(nb, nbb) for nb in (16, 32, 64) for nbb in (64, 128, 256, 512)]
DEFAULT = (32, 128)


def size_class(n: int) -> int:
    for c in CLASSES:
        if n <= c:
            return c
    return CLASSES[-1]


def load_table(path: str = TABLE) -> dict:
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def choice(n: int, table: dict | None = None):
    table = load_table() if table is None else table
    e = table.get(str(size_class(n)))
    return (int(e["nb"]), int(e["nbb"])) if e else DEFAULT


def apply(g, n: int, table: dict | None = None):
    nb, nbb = choice(n, table)
    g.set_tuning(nb, nbb)
    return nb, nbb


def _spd(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, n))
    return np.asfortranarray(X @ X.T / n)


def tune(g, classes=(1024, 2048, 4096), candidates=CANDIDATES, repeats=2, seed=0):
    """Time every candidate on one SPD matrix per class; returns the table."""
    from . import synth
    table = {}
    for n in classes:
        A = _spd(n, seed + n)
        t = synth.tiny()
        t.n, t.m = n, 4
        t.grid = None  # the coordinates below are not on the tile's grids
        t.x_i, t.y_i = np.zeros(n), np.zeros(n)
        t.exp_i = np.ones(n, dtype=np.int32)
        t.X_a, t.Y_a = np.zeros(4), np.zeros(4)
        t.Ndiag = np.ones(n)
        g.set_tile(t)
        timings = {}
        for nb, nbb in candidates:
            g.set_tuning(nb, nbb)
            best = None
            for _ in range(repeats):
                g.set_A(A)
                g.set_profiling(1)
                g.eigen_A(out=False)
                st = g.stage_times()
                g.set_profiling(0)
                # nb and nbb act on the reduction and the back-transform only
                # (the tridiagonal eigensolver between them does not see them)
                ms = st["sytrd"][0] + st["ormtr"][0]
                best = ms if best is None else min(best, ms)
            timings[f"{nb},{nbb}"] = round(best, 3)
        nb, nbb = (int(v) for v in min(timings, key=timings.get).split(","))
        table[str(n)] = {"nb": nb, "nbb": nbb, "ms": timings}
    g.set_tuning(*DEFAULT)
    return table


def save_table(table: dict, path: str = TABLE):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)


def main(argv=None):
    import argparse
    from .gpu import ImcomGPU
    ap = argparse.ArgumentParser()
    ap.add_argument("--classes", default=",".join(str(c) for c in CLASSES))
    ap.add_argument("--out", default=TABLE)
    ap.add_argument("--repeats", type=int, default=2)
    a = ap.parse_args(argv)
    classes = [int(v) for v in a.classes.split(",")]
    with ImcomGPU(0) as g:
        table = {}
        for n in classes:  # one class at a time, printed as it finishes
            table.update(tune(g, classes=(n,), repeats=a.repeats))
            e = table[str(n)]
            print(f"n={n}: nb={e['nb']} nbb={e['nbb']} {e['ms']}", flush=True)
    save_table(table, a.out)
    return table


if __name__ == "__main__":
    main()
