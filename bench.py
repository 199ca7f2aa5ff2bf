"""Benchmark of the IMCOM solve path on one MI355X.

A "step" is one pass of the whole solve path over one tile whose inputs
(coordinates, exposure ids, noise, lookup tables) are already resident in
HBM: lookup_A -> eigen_A -> lookup_B -> build_P -> bisect_kappa -> build_T ->
build_S -> build_image, exactly the driver's sequence (imcom.f90:151-169).
The default workload is the large tile: 4 exposures of 64x64 (n = 16384
input pixels), Airy (obscuration 0.3) x pixel PSF, 96x96 output pixels at
0.5 px (m = 9216), leakage target U_max = 1e-6.  The metric is output
pixels per second; every pixel's U is checked against U_max after the timed
steps.

The JSON line carries a roofline entry for the dominant kernel (the
matrix-vector kernel of the tridiagonal reduction, HBM-bound; its time is
measured with HIP events on the context's stream in one extra profiled step)
and a CPU baseline: the oracle's C/LAPACK restatement timed on bounded
samples of the same tile on this host's cores, scaled stage by stage to the
whole tile.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6
METRIC = "output pixels/sec at U≤1e-6 (A/B+eig+κ+T), learned+split vs plain FP64"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def symv_alg_bytes(n: int) -> float:
    """Algorithmic bytes of all matrix-vector launches of the one-stage
    reduction: column c reads the lower triangle of the (n-c-1)^2 trailing
    matrix once (8 bytes per element); n-1 launches."""
    s = np.arange(1, n, dtype=np.float64)  # trailing sizes n-1 .. 1
    return float(np.sum(s * (s + 1) / 2.0) * 8.0)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="large", choices=["large", "small"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--stages", action="store_true", help="print the per-stage breakdown")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); tiles shard "
              "across GPUs with no communication, see the design notes", file=sys.stderr)
        sys.exit(2)

    import torch  # device plumbing only: barrier/synchronize around the timed region
    from imcom_amd import synth
    from imcom_amd.gpu import ImcomGPU

    t0 = time.time()
    tile = synth.large() if args.workload == "large" else synth.small()
    log(f"workload {args.workload}: n={tile.n} m={tile.m} table={tile.Alookup.shape} "
        f"({time.time() - t0:.1f}s host set-up)")
    g = ImcomGPU(0)
    g.set_tile(tile)

    def step():
        g.lookup_A(tile.npoly, tile.npad)
        g.eigen_A(out=False)
        g.lookup_B(tile.npoly, tile.npad)
        g.build_P()
        K, U, ev = g.bisect_kappa()
        g.build_T(out=False)
        S = g.build_S()
        H = g.build_image()
        return K, U, S, H, ev

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    g.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        K, U, S, H, ev = step()
    g.synchronize()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    sec = (t2 - t1) / args.steps
    value = tile.m / sec
    n_bad = int(np.sum(~(U < tile.US_max)))
    log(f"{args.steps} steps: {sec * 1e3:.1f} ms/step, {value:.1f} px/s, "
        f"pixels with U >= U_max: {n_bad}, mean evals {ev.mean():.1f}")

    # one profiled step: per-stage and per-launch device times; keeps L and P
    # of this tile for the CPU baseline's bisection sample
    g.set_profiling(2)
    g.lookup_A(tile.npoly, tile.npad)
    _, Lh, _ = g.eigen_A(out_Q=False, out_L=True)
    g.lookup_B(tile.npoly, tile.npad)
    Ph = g.build_P(out=True)
    g.bisect_kappa()
    g.build_T(out=False)
    g.build_S()
    g.build_image()
    st = g.stage_times()
    g.set_profiling(0)
    log("stage times (ms, launches): " + ", ".join(f"{k}={v[0]:.1f}/{v[1]}" for k, v in st.items()))
    symv_ms, symv_n = st["symv"]
    alg = symv_alg_bytes(tile.n)
    per_launch_bytes = alg / max(symv_n, 1)
    avg_ms = symv_ms / max(symv_n, 1)
    achieved = per_launch_bytes / (avg_ms * 1e-3) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"sy_symv_traffic_{args.workload}.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    roofline = {"bound": "hbm", "kernel": "sy_symv", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic, "alg_bytes_per_launch": round(per_launch_bytes),
                "avg_launch_ms": round(avg_ms, 5), "launches": symv_n}

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        O.build()
        tc = time.time()
        tot, stages = O.sampled_cpu_time(tile, L=Lh, P=Ph)
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu = {"value": round(tile.m / tot, 4), "unit": "output pixels/sec", "cores": cores,
               "kind": "port",
               "sample": ("oracle C+LAPACK(OpenBLAS) on bounded samples of this tile, scaled "
                          "per stage: A rows 0-95, B rows 0-95, DSYEVR on the leading 2048^2 "
                          "block (x (n/2048)^3), DGEMM on 128 output columns, bisection on 48 "
                          "and T on 8 output pixels; "
                          f"{time.time() - tc:.0f}s of CPU work"),
               "stages_s": {k: round(v, 2) for k, v in stages.items()}}
        log(f"cpu baseline: {tot:.0f} s/tile -> {tile.m / tot:.3f} px/s ({cores} threads)")

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "output pixels/sec", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(sec * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {tile.meta.get('name')} tile, n={tile.n} "
                               f"input px, m={tile.m} output px, U_max={tile.US_max}",
                   "path": "plain FP64 (full-range bisection, FP64 GEMMs)",
                   "pixels_U_ge_Umax": n_bad},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    g.close()


if __name__ == "__main__":
    main()
