"""Benchmark of the IMCOM solve path on one MI355X.

A "step" is one pass of the whole solve path with the inputs (coordinates,
exposure ids, noise, lookup tables) already resident in HBM:
lookup_A -> eigen_A -> lookup_B -> build_P -> bisect_kappa -> build_T ->
build_S -> build_image, the driver's sequence (imcom.f90:151-169).

Workloads:
  large  (default) one tile: 4 exposures of 64x64 (n = 16384 input pixels),
         Airy (obscuration 0.3, lambda/D = 1 input px) x unit-pixel PSF on a
         256^2 image at 0.25 px (lookup tables 1536^2), 96x96 output pixels at
         0.25 px (m = 9216, a 24 px square centred in the 64 px footprint),
         U_max = 1e-6.  The pitch and lambda/D are this benchmark's choice:
         at lambda/D = 0.5 px the aliased Airy power leaves a leakage floor
         near 2e-4, and a 0.5 px pitch puts the grid's edge 8 px from the
         footprint's, too close for the Airy wings to reach U = 1e-6.
  small  one tile: 4 x 16x16 (n = 1024), Gaussian PSF sigma 1.2 px, 24x24
         output at 0.5 px (m = 576)
  batch  64 tiles with n in [1024, 6144] and m in [256, 2304]; the fast path
         is three concurrent batched solves per step (HIP-graph replay)

Two paths are timed: "plain" (full-range bisection, FP64 GEMMs) and "fast"
(learned kappa bracket when trained weights exist, BF16x6 split GEMM for T;
P stays FP64 because U ~ 1e-6 is a difference of O(1) terms).  `value` is
the fast path's output pixels per second; the plain path's number is in
config.  After the timed steps every pixel's U is checked against U_max.
The default (large) run also measures the 64-tile mosaic (config.batch),
where the batched fast path (one batched solve of all 64 tiles) is compared
with the plain single-tile path.  --batch-streams 3 splits the tiles over
three concurrent solves from three host threads (faster, but that form
stalled silently in two test runs with no cause found, so it is not the
default).

The JSON line carries a roofline entry for the dominant kernel, timed by
HIP events that its dispatch records on the context's stream in one extra
profiled step: with the two-stage reduction (the default from n = 7168 up,
so for the large tile) q2_apply_kernel, the bulge back-transform (FP64
MFMA-bound, 4 n ((n-1) n / 2 - 1) algorithmic flops); with the one-stage
reduction sy_symv, its matrix-vector kernel (HBM-bound), or for the small
tile the persistent tail kernel that reduces every column; with
the per-stage rooflines of that step beside it, and a CPU baseline: the
oracle's C/LAPACK restatement on this host's cores, on the whole tile for the
small workload; for the large one, lookup_A and DSYEVR on the whole tile and
the other stages on bounded samples scaled stage by stage (with the scale
factors checked by a second sample).
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
FP64_PEAK_TFS = 78.6     # dense FP64 (vector and MFMA)
BF16_PEAK_TFS = 2500.0   # dense BF16 MFMA
METRIC = "output pixels/sec at U≤1e-6 (A/B+eig+κ+T), learned+split vs plain FP64"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def symv_alg_bytes(n: int, launches: int | None = None) -> float:
    """Algorithmic bytes of the one-stage reduction's matrix-vector launches:
    column c reads the lower triangle of the (n-c-1)^2 trailing matrix once
    (8 bytes per element).  The blocked kernels run the first `launches`
    columns (all n-1 by default); the persistent tail kernel, which keeps its
    trailing matrix on chip, reduces the rest."""
    k = n - 1 if launches is None else launches
    s = np.arange(n - 1, n - 1 - k, -1, dtype=np.float64)
    return float(np.sum(s * (s + 1) / 2.0) * 8.0)


def q2_alg_flops(n: int) -> float:
    """Algorithmic flops of the two-stage reduction's bulge back-transform
    (one launch of q2_apply_kernel): the bulge chase leaves one reflector of
    length L for every task, and the tasks of sweep s cover rows s+1..n-1 in
    pieces of at most 64, so the lengths of sweep s add up to n-1-s; each
    entry is applied to the n columns of Z with 4 flops (a dot product and an
    update): 4 n sum_{s=0}^{n-3} (n-1-s) = 4 n ((n-1) n / 2 - 1), about 2 n^3."""
    return 4.0 * n * ((n - 1) * n / 2.0 - 1.0)


def describe(workload: str, tiles) -> str:
    """The workload's shape and sizes, as the JSON line states them."""
    t = tiles[0]
    N = t.npad * int(t.meta.get("n1pad", 0))
    if workload == "large":
        return (f"large: 1 tile, 4 x 64^2 inputs (n={t.n}), Airy (obscuration 0.3, lambda/D = 1 "
                f"input px) x unit-pixel PSF on a 256^2 image at 0.25 px ({N}^2 lookup tables), "
                f"96^2 outputs at 0.25 px pitch (m={t.m}), Gaussian target sigma 1.8 px, "
                f"U_max={t.US_max}")
    if workload == "small":
        return (f"small: 1 tile, 4 x 16^2 inputs (n={t.n}), Gaussian PSF sigma 1.2 px on a 128^2 "
                f"image at 0.25 px ({N}^2 lookup tables), 24^2 outputs at 0.5 px pitch "
                f"(m={t.m}), Gaussian target sigma 1.5 px, U_max={t.US_max}")
    return (f"batch: {len(tiles)} tiles, n in [{min(x.n for x in tiles)}, "
            f"{max(x.n for x in tiles)}], sum m={sum(x.m for x in tiles)}, Gaussian PSFs sigma "
            f"in [1.0, 1.6] px, U_max={t.US_max}")


def stage_rooflines(n: int, m: int, st: dict, t_bf16_ms: float) -> list:
    """Roofline of each stage of one profiled plain step (HIP-event device
    times, no downloads inside them) from its algorithmic work:
      lookup_A  8 n^2 bytes written (A, both triangles)     HBM
      lookup_B  8 n m bytes written                          HBM
      eigen_A   (10/3) n^3 flops (reduction + D&C + back-transform; the
                one-stage reduction is itself HBM-bound: see roofline)  FP64
      build_P   2 n^2 m flops                                FP64 MFMA
      bisect    8 n m bytes read (P once)                    HBM
      build_T   2 n^2 m flops (plus W = P / (L + K))         FP64 MFMA
      build_T BF16x6  2 n^2 m FP64-equivalent flops, and 6 x 2 n^2 m flops
                of BF16 MFMA products against the BF16 dense peak"""
    def row(stage, ms, work, unit, peak, bound, what):
        if ms <= 0:
            return None
        ach = work / (ms * 1e-3) / (1e9 if unit == "GB/s" else 1e12)
        return {"stage": stage, "bound": bound, "work": what, "ms": round(ms, 3),
                "achieved": round(ach, 2), "peak": peak, "unit": unit,
                "frac": round(ach / peak, 4)}
    eig_ms = st["sytrd"][0] + st["stedc"][0] + st["ormtr"][0]
    rows = [
        row("lookup_A", st["lookup_A"][0], 8.0 * n * n, "GB/s", HBM_PEAK_GBS, "hbm", "8 n^2 B written"),
        row("lookup_B", st["lookup_B"][0], 8.0 * n * m, "GB/s", HBM_PEAK_GBS, "hbm", "8 n m B written"),
        row("eigen_A", eig_ms, 10.0 / 3.0 * n ** 3, "TFLOP/s", FP64_PEAK_TFS, "fp64",
            "(10/3) n^3 flops"),
        row("build_P", st["build_P"][0], 2.0 * n * n * m, "TFLOP/s", FP64_PEAK_TFS, "mfma",
            "2 n^2 m flops"),
        row("bisect_kappa", st["bisect"][0], 8.0 * n * m, "GB/s", HBM_PEAK_GBS, "hbm",
            "8 n m B read"),
        row("build_T", st["build_T"][0], 2.0 * n * n * m, "TFLOP/s", FP64_PEAK_TFS, "mfma",
            "2 n^2 m flops"),
        row("build_T_bf16x6", t_bf16_ms, 12.0 * n * n * m, "TFLOP/s", BF16_PEAK_TFS, "mfma",
            "6 x 2 n^2 m BF16 flops (2 n^2 m FP64-equivalent)"),
    ]
    return [r for r in rows if r is not None]


def batch_block(steps: int, warmup: int, weights, skip_plain: bool, nstreams: int = 1) -> dict:
    """The 64-tile mosaic (synth.batch): the fast path as batched solves
    replayed from HIP graphs (learned bracket, BF16x6 T): one solve of all the
    tiles on one context and stream (nstreams = 1), or the tiles dealt over
    nstreams concurrent solves from as many host threads (imcom_amd.batch);
    the plain path as the single-tile FP64 path tile after tile, each tile's
    inputs resident in its own context before the timed region.  Every
    pixel's U is checked against U_max after the timed steps."""
    import torch
    from imcom_amd import autotune, synth
    from imcom_amd.batch import StreamedBatch
    from imcom_amd.gpu import IMCOM_PREC_BF16X6, ImcomGPU
    tiles = synth.batch()
    m_total = sum(t.m for t in tiles)
    Umax = np.concatenate([np.full(t.m, t.US_max) for t in tiles])
    learned = weights is not None

    def timed(step, k, w):
        for _ in range(w):
            step()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(k):
            U, ev = step()
        torch.cuda.synchronize()
        return (time.perf_counter() - t1) / k, int(np.sum(~(U < Umax))), float(ev.mean())

    # fast: opened before any other context (an idle context beside
    # concurrent solves slows them)
    if nstreams > 1:
        sb = StreamedBatch(tiles, nstreams, weights=weights)
        solve, result, close = sb.solve, sb.result, sb.close
    else:
        gb = ImcomGPU(0)
        autotune.apply(gb, max(t.n for t in tiles))
        if weights is not None:
            gb.mlp_set_params(weights)
        gb.batch_set(tiles)
        solve, result, close = gb.batch_solve, gb.batch_result, gb.close
    try:
        def fast():
            solve(precision_T=IMCOM_PREC_BF16X6, learned=learned, use_graph=True)
            res = [result(i) for i in range(len(tiles))]
            return (np.concatenate([r["U"] for r in res]),
                    np.concatenate([r["evals"] for r in res]))
        f_sec, f_bad, f_ev = timed(fast, steps, warmup)
    finally:
        close()
    log(f"batch fast: {f_sec * 1e3:.1f} ms/step, {m_total / f_sec:.1f} px/s, pixels with "
        f"U >= U_max: {f_bad}, mean U evaluations {f_ev:.2f}")
    how = ("one batched solve of all the tiles (one context and stream, HIP-graph replay)"
           if nstreams == 1 else
           f"{nstreams} concurrent batched solves (one stream and host thread each, HIP-graph "
           "replay)")
    out = {"workload": (f"{len(tiles)} tiles, n in [{min(t.n for t in tiles)}, "
                        f"{max(t.n for t in tiles)}], sum m = {m_total}, U_max = 1e-6"),
           "fast_path": (how + ", " + ("learned kappa bracket + " if learned else "")
                         + "BF16x6 split GEMM for T; P and the bisection FP64"),
           "fast_px_per_s": round(m_total / f_sec, 2), "fast_ms_per_step": round(f_sec * 1e3, 2),
           "fast_pixels_U_ge_Umax": f_bad, "fast_mean_U_evaluations": round(f_ev, 2),
           "steps": steps, "warmup": warmup}
    if skip_plain:
        return out
    singles = []
    try:
        for t in tiles:
            gs = ImcomGPU(0)
            singles.append(gs)
            autotune.apply(gs, t.n)
            gs.set_tile(t)

        def plain():
            Us, evs = [], []
            for gs, t in zip(singles, tiles):
                gs.lookup_A(t.npoly, t.npad)
                gs.eigen_A(out=False)
                gs.lookup_B(t.npoly, t.npad)
                gs.build_P()
                _, U, ev = gs.bisect_kappa()
                gs.build_T(out=False)
                gs.build_S()
                gs.build_image()
                Us.append(U)
                evs.append(ev)
            return np.concatenate(Us), np.concatenate(evs)
        p_sec, p_bad, p_ev = timed(plain, 1, 1)
    finally:
        for gs in singles:
            gs.close()
    log(f"batch plain: {p_sec * 1e3:.1f} ms/step, {m_total / p_sec:.1f} px/s, pixels with "
        f"U >= U_max: {p_bad}, mean U evaluations {p_ev:.2f}")
    out.update({"plain_path": "single-tile FP64 path, tile after tile, full-range bisection",
                "plain_px_per_s": round(m_total / p_sec, 2),
                "plain_ms_per_step": round(p_sec * 1e3, 2), "plain_pixels_U_ge_Umax": p_bad,
                "plain_mean_U_evaluations": round(p_ev, 2), "plain_steps": 1,
                "fast_over_plain": round(p_sec / f_sec, 3)})
    return out


def vendor_eig(g, tile):
    """rocSOLVER's dsyevd (one-stage reduction, divide and conquer,
    back-transform; ROCm's FP64 symmetric eigensolver) on the same A as
    this library's eigen_A, both timed on the device-resident matrix."""
    import ctypes
    import torch
    rb = ctypes.CDLL("/opt/rocm/lib/librocblas.so")
    rs = ctypes.CDLL("/opt/rocm/lib/librocsolver.so")
    vp, ci = ctypes.c_void_p, ctypes.c_int
    rs.rocsolver_dsyevd.argtypes = [vp, ci, ci, ci, vp, ci, vp, vp, vp]
    rs.rocsolver_dsyevd.restype = ci
    h = vp()
    if rb.rocblas_create_handle(ctypes.byref(h)) != 0:
        raise RuntimeError("rocblas_create_handle failed")
    EVECT_ORIGINAL, FILL_LOWER = 211, 122
    n = tile.n
    A = g.lookup_A(tile.npoly, tile.npad, out=True)
    g.synchronize()
    t1 = time.perf_counter()
    _, L, _ = g.eigen_A(out_Q=False, out_L=True)
    g.synchronize()
    ours = time.perf_counter() - t1

    def syevd(Ad, nn):
        D = torch.empty(nn, dtype=torch.float64, device="cuda:0")
        E = torch.empty_like(D)
        info = torch.zeros(1, dtype=torch.int32, device="cuda:0")
        st = rs.rocsolver_dsyevd(h, EVECT_ORIGINAL, FILL_LOWER, nn, Ad.data_ptr(), nn,
                                 D.data_ptr(), E.data_ptr(), info.data_ptr())
        torch.cuda.synchronize()
        if st != 0 or int(info.item()) != 0:
            raise RuntimeError(f"rocsolver_dsyevd: status {st}, info {int(info.item())}")
        return D
    syevd(torch.eye(512, dtype=torch.float64, device="cuda:0").reshape(-1).clone(), 512)  # code load
    Ad = torch.from_numpy(np.ascontiguousarray(A.ravel(order="F"))).to("cuda:0")
    del A
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    D = syevd(Ad, n)
    vendor = time.perf_counter() - t1
    Dv = D.cpu().numpy()[::-1]
    rb.rocblas_destroy_handle(h)
    diff = float(np.max(np.abs(Dv - L)) / np.max(np.abs(L)))
    log(f"eigendecomposition n={n}: this library {ours * 1e3:.1f} ms, rocSOLVER dsyevd "
        f"{vendor * 1e3:.1f} ms; max eigenvalue difference {diff:.2e} of lambda_max")
    return {"library": "rocSOLVER dsyevd (ROCm 7.2), evect=original, lower", "n": n,
            "vendor_ms": round(vendor * 1e3, 1), "this_library_ms": round(ours * 1e3, 1),
            "max_eigenvalue_diff_rel": diff}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="large", choices=["large", "small", "batch"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--no-batch", action="store_true",
                    help="large workload: leave out the 64-tile mosaic block (config.batch)")
    ap.add_argument("--batch-streams", type=int, default=1,
                    help="the mosaic's fast path as this many concurrent batched solves (one "
                         "host thread each); 1 (default): one batched solve of all the tiles.  "
                         "The three-solve form is faster but stalled silently in two test runs "
                         "with no cause found")
    ap.add_argument("--vendor-eig", action="store_true",
                    help="also time rocSOLVER's dsyevd on the same A (single-tile workloads)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB 5 times (scales the "
                         "FETCH_SIZE / WRITE_SIZE counters to bytes) and exit")
    ap.add_argument("--tail-rows", type=int, default=2048,
                    help="single-tile eigendecompositions reduce their last columns (up to this "
                         "trailing size, at most 2048, the default) in the persistent kernel; "
                         "0: every column on the blocked kernels")
    ap.add_argument("--stages", type=int, default=0, choices=[0, 1, 2],
                    help="reduction to tridiagonal form: 1 one-stage, 2 two-stage (for every n), "
                         "0 the library's default (two-stage from n = 7168 up)")
    ap.add_argument("--eigen-only", action="store_true",
                    help="build A and run one eigendecomposition (no timing events): the run "
                         "q2_apply_kernel's FETCH_SIZE / WRITE_SIZE counters are collected on")
    ap.add_argument("--reduction-only", action="store_true",
                    help="single-tile workloads: build A and run its tridiagonal reduction once "
                         "(no timing events), print its device time and exit; the run that "
                         "sy_symv's FETCH_SIZE / WRITE_SIZE counters are collected on")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        nbytes = args.copy_calibration << 20
        src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes,
                          "launches": 5}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); a mosaic's "
              "tiles would shard across GPUs with no communication, one tile's solve stays "
              "on one GPU", file=sys.stderr)
        sys.exit(2)

    import torch  # device plumbing only: barrier/synchronize around the timed region
    from imcom_amd import autotune, kappa_net, synth
    from imcom_amd.gpu import IMCOM_PREC_BF16X6, IMCOM_PREC_FP64, TWOSTAGE_MIN, ImcomGPU

    t0 = time.time()
    if args.workload == "batch":
        tiles = synth.batch()
    else:
        tiles = [synth.large() if args.workload == "large" else synth.small()]
    tile = tiles[0]
    m_total = sum(t.m for t in tiles)
    log(f"workload {args.workload}: {len(tiles)} tile(s), n in [{min(t.n for t in tiles)}, "
        f"{max(t.n for t in tiles)}], sum m = {m_total} ({time.time() - t0:.1f}s host set-up)")
    weights = kappa_net.load_weights()
    learned = weights is not None
    g = None  # the single-tile context (the batch workload makes its own)
    if args.workload != "batch":
        g = ImcomGPU(0)
        autotune.apply(g, max(t.n for t in tiles))
        g.set_reduction_tail(args.tail_rows)
        if args.stages:
            g.set_reduction_stages(args.stages, 3)
        if weights is not None:
            g.mlp_set_params(weights)
        if args.eigen_only:
            g.set_tile(tile)
            g.lookup_A(tile.npoly, tile.npad)
            torch.cuda.synchronize()
            tr = time.perf_counter()
            g.eigen_A(out=False)
            g.synchronize()
            print(json.dumps({"eigen": args.workload, "n": tile.n,
                              "seconds": round(time.perf_counter() - tr, 4)}), flush=True)
            g.close()
            return
        if args.reduction_only:
            g.set_tile(tile)
            g.lookup_A(tile.npoly, tile.npad)
            torch.cuda.synchronize()
            tr = time.perf_counter()
            d, e = g.tridiagonalize()
            # sy_symv runs the columns before the first panel boundary from
            # which the persistent tail kernel takes over
            nb = autotune.choice(tile.n)[0]
            k_tail = tile.n - 1
            if args.tail_rows >= 2:
                k_tail = 0
                while tile.n - k_tail > min(args.tail_rows, 2048):
                    k_tail += nb
            print(json.dumps({"reduction": args.workload, "n": tile.n, "symv_launches": k_tail,
                              "seconds": round(time.perf_counter() - tr, 4)}), flush=True)
            g.close()
            return

    if args.workload == "batch":
        # fast: one batched solve of all the tiles (lockstep launches replayed
        # from a HIP graph), or with --batch-streams N the size-sorted tiles
        # cut into N runs of about equal n^3 work, one batched solve each on
        # its own stream, solved concurrently (imcom_amd.batch); plain: the
        # single-tile FP64 path tile after tile, each tile's inputs resident in
        # its own context before the timed region (the plain path's contexts
        # are released before the fast path's are set up: streams beyond the
        # process's hardware queues share them)
        from imcom_amd.batch import StreamedBatch
        sbs = []
        singles = []
        if not args.skip_plain:
            for t in tiles:
                gs = ImcomGPU(0)
                autotune.apply(gs, t.n)
                gs.set_tile(t)
                singles.append(gs)
        Umax_all = np.concatenate([np.full(t.m, t.US_max) for t in tiles])

        def single_step(gs, t):
            gs.lookup_A(t.npoly, t.npad)
            gs.eigen_A(out=False)
            gs.lookup_B(t.npoly, t.npad)
            gs.build_P()
            K, U, ev = gs.bisect_kappa()
            gs.build_T(IMCOM_PREC_FP64, out=False)
            gs.build_S()
            gs.build_image()
            return U, ev

        def prepare_fast():
            for gs in singles:
                gs.close()
            singles.clear()
            if args.batch_streams > 1:
                sbs.append(StreamedBatch(tiles, args.batch_streams, weights=weights))
            else:
                gb = ImcomGPU(0)
                autotune.apply(gb, max(t.n for t in tiles))
                if weights is not None:
                    gb.mlp_set_params(weights)
                gb.batch_set(tiles)
                sbs.append(gb)

        def step(fast):
            if not fast:
                res = [single_step(gs, t) for gs, t in zip(singles, tiles)]
                return (np.concatenate([r[0] for r in res]), np.concatenate([r[1] for r in res]),
                        Umax_all)
            sb = sbs[0]
            if args.batch_streams > 1:
                sb.solve(precision_T=IMCOM_PREC_BF16X6, learned=learned, use_graph=True)
                res = [sb.result(i) for i in range(len(tiles))]
            else:
                sb.batch_solve(precision_T=IMCOM_PREC_BF16X6, learned=learned, use_graph=True)
                res = [sb.batch_result(i) for i in range(len(tiles))]
            U = np.concatenate([r["U"] for r in res])
            ev = np.concatenate([r["evals"] for r in res])
            return U, ev, Umax_all
    else:
        g.set_tile(tile)
        prepare_fast = lambda: None  # noqa: E731

        def step(fast):
            g.lookup_A(tile.npoly, tile.npad)
            g.eigen_A(out=False)
            g.lookup_B(tile.npoly, tile.npad)
            g.build_P()
            if fast and learned:
                K, U, ev, _ = g.bisect_kappa_learned()
            else:
                K, U, ev = g.bisect_kappa()
            g.build_T(IMCOM_PREC_BF16X6 if fast else IMCOM_PREC_FP64, out=False)
            g.build_S()
            g.build_image()
            return U, ev, np.full(tile.m, tile.US_max)

    def timed(fast):
        for _ in range(args.warmup):
            step(fast)
        torch.cuda.synchronize()
        if g is not None:
            g.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            U, ev, Umax = step(fast)
        if g is not None:
            g.synchronize()
        torch.cuda.synchronize()
        sec = (time.perf_counter() - t1) / args.steps
        bad = int(np.sum(~(U < Umax)))
        log(f"{'fast' if fast else 'plain'}: {sec * 1e3:.1f} ms/step, {m_total / sec:.1f} px/s, "
            f"pixels with U >= U_max: {bad}, mean U evaluations {ev.mean():.2f}")
        return sec, bad, float(ev.mean())

    plain = None if args.skip_plain else timed(False)
    prepare_fast()
    fast = timed(True)
    sec, n_bad, mean_ev = fast

    # the same fast step with the host hand-over around it: inputs uploaded
    # from host memory first, T and the maps copied back (not the metric).
    # T goes to a page-locked host buffer allocated beforehand (fetch_async),
    # copied while S and H are formed.
    pcie = None
    if args.workload != "batch":
        from imcom_amd.gpu import host_empty
        T_host = host_empty((tile.n, tile.m))
        T_host[:] = 0.0  # page-table population outside the timed region
        g.synchronize()
        t1 = time.perf_counter()
        g.set_tile(tile)
        g.lookup_A(tile.npoly, tile.npad)
        g.eigen_A(out=False)
        g.lookup_B(tile.npoly, tile.npad)
        g.build_P()
        if learned:
            g.bisect_kappa_learned()
        else:
            g.bisect_kappa()
        g.build_T(IMCOM_PREC_BF16X6, out=False)
        g.fetch_async("T", out=T_host)
        g.build_S()
        g.build_image()
        g.fetch_wait("T")
        g.synchronize()
        pcie = time.perf_counter() - t1
        log(f"fast step with host hand-over (inputs up, T/K/U/S/H down): {pcie * 1e3:.1f} ms")

    # one profiled plain step (single tile): per-stage and per-launch device
    # times (no stage output leaves the device inside it), then the BF16x6 T
    roofline = None
    Lh = Ph = None
    if args.workload != "batch":
        g.set_profiling(2)
        g.lookup_A(tile.npoly, tile.npad)
        _, Lh, _ = g.eigen_A(out_Q=False, out_L=True)
        g.lookup_B(tile.npoly, tile.npad)
        g.build_P()
        g.bisect_kappa()
        g.build_T(out=False)
        g.build_S()
        g.build_image()
        st = g.stage_times()
        g.set_profiling(1)
        g.build_T(IMCOM_PREC_BF16X6, out=False)
        t_bf16 = g.stage_times()["build_T"][0]
        g.set_profiling(0)
        Ph = g.fetch_async("P")
        g.fetch_wait("P")
        log("stage times (ms, launches): " + ", ".join(f"{k}={v[0]:.1f}/{v[1]}" for k, v in st.items()))
        symv_ms, symv_n = st["symv"]
        if symv_n > 0:
            per_launch = symv_alg_bytes(tile.n, symv_n) / symv_n
            avg_ms = symv_ms / symv_n
            achieved = per_launch / (avg_ms * 1e-3) / 1e9
            traffic = None
            tfile = os.path.join(ROOT, "profiles", f"sy_symv_traffic_{args.workload}.json")
            if os.path.exists(tfile):
                with open(tfile) as f:
                    traffic = json.load(f).get("hbm_bytes_per_launch")
            roofline = {"bound": "hbm", "kernel": "sy_symv", "achieved": round(achieved, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                        "alg_bytes_per_launch": round(per_launch),
                        "avg_launch_ms": round(avg_ms, 5), "launches": symv_n}
        elif st["q2"][1] > 0:
            # the two-stage reduction: its dominant kernel is the bulge
            # back-transform's one launch, bound by the FP64 matrix cores
            q2_ms, q2_n = st["q2"]
            flops = q2_alg_flops(tile.n)
            achieved = flops / (q2_ms * 1e-3) / 1e12
            traffic = None
            tfile = os.path.join(ROOT, "profiles", f"q2_apply_traffic_{args.workload}.json")
            if os.path.exists(tfile):
                with open(tfile) as f:
                    traffic = json.load(f).get("hbm_bytes_per_launch")
            roofline = {"bound": "mfma", "kernel": "q2_apply_kernel", "achieved": round(achieved, 2),
                        "peak": FP64_PEAK_TFS, "unit": "TFLOP/s",
                        "frac": round(achieved / FP64_PEAK_TFS, 4), "traffic": traffic,
                        "alg_flops_per_launch": round(flops), "avg_launch_ms": round(q2_ms / q2_n, 3),
                        "launches": q2_n,
                        "work": "4 n ((n-1) n / 2 - 1) flops: the bulge chase's reflectors "
                                "applied to the n columns of the eigenvectors",
                        "bulge_chase_ms": round(st["bulge"][0], 2)}
        else:
            # the persistent tail reduced every column: its one launch reads
            # the trailing matrix once (8 n^2 B) and is bound by its per-column
            # barrier, not by HBM
            red_ms = st["sytrd"][0]
            achieved = 8.0 * tile.n ** 2 / (red_ms * 1e-3) / 1e9
            roofline = {"bound": "hbm", "kernel": "sy_tail_kernel", "achieved": round(achieved, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": None,
                        "alg_bytes_per_launch": round(8.0 * tile.n ** 2),
                        "avg_launch_ms": round(red_ms, 4), "launches": 1,
                        "note": "latency-bound: one grid barrier per column"}
        roofline["stage_ms"] = {k: round(v[0], 2) for k, v in st.items()}
        roofline["stages"] = stage_rooflines(tile.n, tile.m, st, t_bf16)

    # the lookup-table set-up on the device (build_Alookup/Blookup from the
    # tile's spectra) beside the host's numpy set-up; not part of the step
    tables = None
    if args.workload != "batch" and tile.Gt_rot is not None:
        from imcom_amd import tile as TL

        def dev_tables():
            g.build_Alookup(tile.psfconst, tile.Gt_rot, tile.npad)
            g.build_Blookup(tile.psfconst, tile.Gammat, tile.Gt_rot, tile.npad)
            g.synchronize()
        dev_tables()
        t1 = time.perf_counter()
        dev_tables()
        t_dev = time.perf_counter() - t1
        t1 = time.perf_counter()
        TL.build_Alookup(tile.Gt_rot, tile.npad)
        TL.build_Blookup(tile.Gammat, tile.Gt_rot, tile.npad)
        t_host = time.perf_counter() - t1
        N = tile.npad * tile.Gt_rot[0].shape[0]
        tables = {"grid": f"{N}x{N}", "n_tables": tile.nexppsf * (tile.nexppsf + 3) // 2,
                  "device_ms": round(t_dev * 1e3, 2), "host_numpy_ms": round(t_host * 1e3, 2)}
        log(f"table set-up: device {t_dev * 1e3:.1f} ms, host numpy {t_host * 1e3:.1f} ms ({N}^2)")

    vendor = None
    if args.vendor_eig and args.workload != "batch":
        vendor = vendor_eig(g, tile)

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        O.build()
        tc = time.time()
        tot = 0.0
        stages = {}
        checks = {}
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        if args.workload == "small":
            # the whole tile: every stage of the oracle's solve, timed
            t1 = time.perf_counter()
            O.solve(tile)
            tot = time.perf_counter() - t1
            m_sampled = m_total
            sample = (f"oracle C+LAPACK(OpenBLAS) solve of the whole tile (n={tile.n}, "
                      f"m={tile.m}): lookup_A, DSYEVR, lookup_B, DGEMM, bisection, T, S, H")
        else:
            for t in (tiles if args.workload != "batch" else tiles[:4]):
                s_t, st_t, ck = O.sampled_cpu_time(t, L=Lh if t is tile else None,
                                                   P=Ph if t is tile else None,
                                                   checks=args.workload == "large",
                                                   full_eig=args.workload == "large")
                tot += s_t
                checks.update(ck)
                for k, v in st_t.items():
                    stages[k] = stages.get(k, 0.0) + v
            m_sampled = m_total if args.workload != "batch" else sum(t.m for t in tiles[:4])
            if args.workload == "large":
                sample = ("oracle C+LAPACK(OpenBLAS) on this tile: lookup_A and DSYEVR on the "
                          "whole tile (timed, not scaled); the rest on bounded samples scaled per "
                          "stage: B rows 0-95, DGEMM on 128 output columns, bisection on 48 and T "
                          "(one DGEMV per pixel) on 32 output pixels")
            else:
                sample = ("oracle C+LAPACK(OpenBLAS) on bounded samples of the first 4 tiles, "
                          "scaled per stage: A rows 0-95, B rows 0-95, DSYEVR on the leading "
                          "2048^2 block (x (n/2048)^3), DGEMM on 128 output columns, bisection "
                          "on 48 and T on 16 output pixels")
        cpu = {"value": round(m_sampled / tot, 4), "unit": "output pixels/sec", "cores": cores,
               "kind": "port", "sample": sample + f"; {time.time() - tc:.0f}s of CPU work",
               "stages_s": {k: round(v, 2) for k, v in stages.items()}}
        if checks:
            cpu["scale_checks"] = checks
        log(f"cpu baseline: {tot:.0f} s for {m_sampled} px -> {cpu['value']} px/s ({cores} threads)")

    cfg = {"workload": describe(args.workload, tiles),
           "path": ((("batched tiles (one batched solve, replayed from a HIP graph) + "
                      if args.batch_streams <= 1 else
                      f"batched tiles ({args.batch_streams} concurrent batched solves, one stream "
                      "each, each replayed from HIP graphs) + ")
                     if args.workload == "batch" else "")
                    + ("learned kappa bracket + BF16x6 split GEMM for T" if learned
                       else "BF16x6 split GEMM for T (no trained bracket weights found)")
                    + "; P and the bisection FP64"),
           "plain_path": ("single-tile FP64 path, tile after tile" if args.workload == "batch"
                          else "single-tile FP64 path, full-range bisection"),
           "pixels_U_ge_Umax": n_bad, "mean_U_evaluations": round(mean_ev, 2)}
    if pcie is not None:
        cfg["pcie_inclusive_px_per_s"] = round(m_total / pcie, 2)
    if args.workload != "batch":
        two = args.stages == 2 or (args.stages == 0 and tile.n >= TWOSTAGE_MIN)
        cfg["reduction"] = ("two-stage (dense -> band 64 -> tridiagonal)" if two
                            else "one-stage blocked")
        cfg["reduction_tail_rows"] = None if two else args.tail_rows
    if tables is not None:
        cfg["table_setup"] = tables
    if vendor is not None:
        cfg["vendor_eigensolver"] = vendor
    if plain is not None:
        cfg["plain_fp64_px_per_s"] = round(m_total / plain[0], 2)
        cfg["plain_fp64_ms_per_step"] = round(plain[0] * 1e3, 2)
        cfg["plain_mean_U_evaluations"] = round(plain[2], 2)
        cfg["fast_over_plain"] = round(plain[0] / sec, 3)
    if g is not None:
        g.close()
        g = None
    if args.workload == "large" and not args.no_batch:
        # the 64-tile mosaic beside the large tile (fewer steps: a plain
        # step solves the 64 tiles one after another, about 4 s)
        cfg["batch"] = batch_block(2, 1, weights, args.skip_plain, args.batch_streams)
    out = {
        "metric": METRIC, "value": round(m_total / sec, 2), "unit": "output pixels/sec",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(sec * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic", "config": cfg, "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    if args.workload == "batch":
        for sb in sbs:
            sb.close()
        for gs in singles:
            gs.close()
    if g is not None:
        g.close()


if __name__ == "__main__":
    main()
