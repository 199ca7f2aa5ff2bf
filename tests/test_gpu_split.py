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
    """Max error of T, U, S (and H) of the split schemes against the FP64
    oracle on the small tile.  P = Q^T B and T = Q W both run split; the
    bisection runs on the split P.  T is compared at the same kappa (GEMM
    error only); U, S, H include the effect of the split P on kappa."""
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
    relT = float(np.max(np.abs(T - O.build_T(r["Q"], r["P"], r["L"], K))) / np.max(np.abs(r["T"])))
    dU = float(np.max(np.abs(U - r["U"])))
    relS = float(np.max(np.abs(S - r["S"]) / r["S"]))
    relH = float(np.max(np.abs(H - r["H"])) / np.max(np.abs(r["H"])))
    print(f"[split-error] scheme={scheme} relT={relT:.3e} absU={dU:.3e} relS={relS:.3e} "
          f"relH={relH:.3e} kappa_changed={(np.abs(K / r['K'] - 1) > 1e-12).mean():.3f}", flush=True)
    assert relT <= 1e-4
    assert dU <= 1e-6
    assert relS <= 1e-4
