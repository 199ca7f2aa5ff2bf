"""Split-precision GEMMs (FP32 x3 and BF16 x6 on MFMA with FP32 accumulation
and FP64 recombination) against FP64.

Tolerances: entrywise |P_split - P_fp64| <= 1e-6 * (|Q|^T |B|), the
accumulation bound of FP32 sums recombined every 256 k.  The end-to-end error
in T, U and S of each scheme on the small tile is measured against the FP64
oracle and recorded (printed); the test bounds it loosely (T and S relative
1e-4, U absolute 1e-6).
"""
import numpy as np
import pytest

from imcom_amd import synth
from imcom_amd.gpu import IMCOM_PREC_BF16X6, IMCOM_PREC_FP32X3
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scheme", [IMCOM_PREC_FP32X3, IMCOM_PREC_BF16X6])
@pytest.mark.parametrize("n,m", [(50, 30), (1024, 576), (2500, 700)])
def test_split_P_random(gpu_ctx, scheme, n, m):
    print(f"[case] split P scheme={scheme} n={n} m={m}", flush=True)
    rng = np.random.default_rng(n * 7 + m + scheme)
    t = synth.tiny()
    t.n, t.m = n, m
    t.grid = None  # the coordinates below are not on the tile's grids
    t.x_i, t.y_i = np.zeros(n), np.zeros(n)
    t.exp_i = np.ones(n, dtype=np.int32)
    t.X_a, t.Y_a = np.zeros(m), np.zeros(m)
    t.Ndiag = np.full(n, 1e-6)
    gpu_ctx.set_tile(t)
    Q = np.asfortranarray(rng.standard_normal((n, n)) / np.sqrt(n))
    B = np.asfortranarray(rng.standard_normal((n, m)) * 10.0 ** rng.uniform(-8, 0, (1, m)))
    gpu_ctx.set_QL(Q, np.ones(n))
    gpu_ctx.set_B(B)
    P = gpu_ctx.build_P(precision=scheme, out=True)
    bound = np.abs(Q).T @ np.abs(B)
    err = np.abs(P - Q.T @ B)
    assert np.all(err <= 1e-6 * bound), float(np.max(err / bound))


@pytest.mark.parametrize("scheme", [IMCOM_PREC_FP32X3, IMCOM_PREC_BF16X6])
def test_split_T_U_S_error_small_tile(gpu_ctx, scheme):
    """Both GEMMs split (P = Q^T B and T = Q W), the bisection on the split P:
    max error of T, U, S and H against the FP64 oracle on the small tile.
    A relative error d in P moves U by about d (U = 1 - (~1)), and T's
    1/(lambda + kappa) factor amplifies P's error along the small eigenvalues,
    so the bounds are loose (measured values are printed)."""
    print(f"[case] split T/U/S error scheme={scheme}", flush=True)
    t = synth.small()
    r = O.solve(t)
    gpu_ctx.set_tile(t)
    gpu_ctx.set_QL(r["Q"], r["L"])
    gpu_ctx.lookup_B(t.npoly, t.npad)
    gpu_ctx.build_P(precision=scheme)
    K, U, ev = gpu_ctx.bisect_kappa()
    T = gpu_ctx.build_T(precision=scheme)
    S = gpu_ctx.build_S()
    H = gpu_ctx.build_image()
    relT = float(np.max(np.abs(T - r["T"])) / np.max(np.abs(r["T"])))
    dU = float(np.max(np.abs(U - r["U"])))
    relS = float(np.max(np.abs(S - r["S"]) / r["S"]))
    relH = float(np.max(np.abs(H - r["H"])) / np.max(np.abs(r["H"])))
    # leakage of the split T itself, in FP64: U = (C + T^T A T + B^T T) / C
    A, B = r["A"], r["B"]
    Utrue = (t.C_a + np.einsum("ia,ia->a", T, A @ T) + np.einsum("ia,ia->a", B, T)) / t.C_a
    print(f"[split-error] scheme={scheme} relT={relT:.3e} absU={dU:.3e} relS={relS:.3e} "
          f"relH={relH:.3e} max U_true={Utrue.max():.6e} "
          f"kappa_changed={(np.abs(K / r['K'] - 1) > 1e-12).mean():.3f}", flush=True)
    # the bounds guard the recorded errors (FP32x3 / BF16x6 on the small tile:
    # T 5.9e-3 / 2.4e-3, U 9.4e-10 / 9.7e-10, S 1.6e-3 / 6.5e-4, H 1.6e-4 /
    # 6.4e-5); with P split, the true leakage of T overshoots U_max by up to
    # ~30%, which is why the fast path keeps P and the bisection in FP64
    assert relT <= 1e-2
    assert dU <= 1e-8
    assert relS <= 1e-2
    assert relH <= 1e-2
    assert Utrue.max() <= 2.0 * t.US_max


@pytest.mark.parametrize("scheme", [IMCOM_PREC_FP32X3, IMCOM_PREC_BF16X6])
def test_split_T_after_fp64_bisection(gpu_ctx, scheme):
    """The benchmark's fast path: P and the bisection stay FP64 (so kappa is
    the FP64 path's), only T = Q W is split.  Every pixel's leakage,
    recomputed in FP64 from the split T, stays within 1e-10 of the FP64
    path's U to within the bisection tolerance US_tol (T's error enters U at
    second order plus a kappa-weighted first-order term), so every pixel keeps
    U <= U_max + US_tol."""
    print(f"[case] split T after FP64 bisection scheme={scheme}", flush=True)
    t = synth.small()
    r = O.solve(t)
    gpu_ctx.set_tile(t)
    gpu_ctx.set_QL(r["Q"], r["L"])
    gpu_ctx.lookup_B(t.npoly, t.npad)
    gpu_ctx.build_P()
    K, U, ev = gpu_ctx.bisect_kappa()
    T64 = gpu_ctx.build_T()
    T = gpu_ctx.build_T(precision=scheme)
    A, B = r["A"], r["B"]
    Utrue = (t.C_a + np.einsum("ia,ia->a", T, A @ T) + np.einsum("ia,ia->a", B, T)) / t.C_a
    relT = float(np.max(np.abs(T - T64)) / np.max(np.abs(T64)))
    dU = float(np.max(np.abs(Utrue - U)))
    print(f"[split-T] scheme={scheme} relT={relT:.3e} max|U_true - U|={dU:.3e} "
          f"max U_true={Utrue.max():.6e}", flush=True)
    assert relT <= 1e-5
    assert dU <= t.US_tol  # within the bisection's own tolerance
    assert np.all(Utrue <= t.US_max + t.US_tol)
