"""The persistent tail of the tridiagonal reduction (sytrd_tail.hip) through
the C ABI: with imcom_gpu_set_reduction_tail(rows) (2048 by default), the columns from the first
panel boundary with at most `rows` remaining are reduced in one launch (the
whole matrix when n <= rows, read from column-major A; otherwise the trailing
block layout left by the blocked kernels).

  eigenvalues      of the tridiagonal form, within 1e3 n eps ||A|| of numpy's
                   eigvalsh(A), with and without the tail
  eigen_A          ||A Q - Q L|| and ||Q^T Q - I|| within 1e3 n eps (||A||),
                   the back-transform reading the tail's reflectors
  reproducible     a second reduction gives bitwise the same d, e
"""
import numpy as np
import pytest

from imcom_amd import synth

pytestmark = pytest.mark.gpu
EPS = 2.0 ** -53


def _tile(n):
    t = synth.tiny()
    t.n, t.m = n, 4
    t.grid = None  # the coordinates below are not on the tile's grids
    t.x_i, t.y_i = np.zeros(n), np.zeros(n)
    t.exp_i = np.ones(n, dtype=np.int32)
    t.X_a, t.Y_a = np.zeros(4), np.zeros(4)
    t.Ndiag = np.ones(n)
    return t


def _matrix(n, kind, rng):
    if kind == "spd":
        X = rng.standard_normal((n, n))
        return np.asfortranarray(X @ X.T / n)
    # graded: eigenvalues over 12 decades, as the system matrices' spectra
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = 10.0 ** rng.uniform(-12, 0, n)
    return np.asfortranarray((Q * lam) @ Q.T)


def _tridiag_eigs(d, e):
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    return np.linalg.eigvalsh(T)


@pytest.mark.parametrize("n,nb,rows,kind", [(2, 32, 2048, "spd"), (3, 32, 2048, "spd"),
                                            (65, 32, 2048, "spd"), (1024, 32, 2048, "graded"),
                                            (1920, 64, 2048, "spd"), (2048, 64, 2048, "graded"),
                                            (2049, 64, 2048, "spd"), (3000, 32, 2048, "spd"),
                                            (3000, 64, 1000, "graded"), (4100, 16, 640, "spd")])
def test_tail_tridiagonal_form(gpu_ctx, n, nb, rows, kind):
    print(f"[case] reduction tail n={n} nb={nb} rows={rows} {kind}", flush=True)
    rng = np.random.default_rng(n + rows)
    A = _matrix(n, kind, rng)
    ref = np.linalg.eigvalsh(A)
    nrm = np.abs(A).max()
    gpu_ctx.set_tile(_tile(n))
    gpu_ctx.set_tuning(nb, 128)
    try:
        gpu_ctx.set_reduction_tail(0)
        gpu_ctx.set_A(A)
        d0, e0 = gpu_ctx.tridiagonalize()
        gpu_ctx.set_reduction_tail(rows)
        gpu_ctx.set_A(A)
        d1, e1 = gpu_ctx.tridiagonalize()
        gpu_ctx.set_A(A)
        d2, e2 = gpu_ctx.tridiagonalize()
    finally:
        gpu_ctx.set_reduction_tail(2048)
        gpu_ctx.set_tuning(32, 128)
    l0, l1 = _tridiag_eigs(d0, e0), _tridiag_eigs(d1, e1)
    err0 = np.abs(l0 - ref).max() / nrm
    err1 = np.abs(l1 - ref).max() / nrm
    print(f"[tail] n={n}: eigenvalue error / ||A||: blocked {err0:.2e}, tail {err1:.2e}; "
          f"max|d1-d0| {np.abs(d1 - d0).max():.2e}", flush=True)
    assert err1 <= 1e3 * n * EPS
    assert err0 <= 1e3 * n * EPS
    assert np.array_equal(d1, d2) and np.array_equal(e1, e2)


@pytest.mark.parametrize("n,nb,nbb", [(700, 32, 128), (1024, 64, 512), (2500, 32, 256)])
def test_tail_eigen_A_residuals(gpu_ctx, n, nb, nbb):
    print(f"[case] eigen_A with the reduction tail n={n} nb={nb} nbb={nbb}", flush=True)
    rng = np.random.default_rng(7 * n)
    A = _matrix(n, "graded", rng)
    gpu_ctx.set_tile(_tile(n))
    gpu_ctx.set_tuning(nb, nbb)
    gpu_ctx.set_reduction_tail(2048)
    try:
        gpu_ctx.set_A(A)
        Q, L, _ = gpu_ctx.eigen_A(out=True)
    finally:
        gpu_ctx.set_tuning(32, 128)
    nrm = np.abs(A).max()
    assert np.all(np.diff(L) <= 0)
    assert np.abs(A @ Q - Q * L).max() <= 1e3 * n * EPS * nrm
    assert np.abs(Q.T @ Q - np.eye(n)).max() <= 1e3 * n * EPS
    assert np.abs(np.sort(L) - np.linalg.eigvalsh(A)).max() <= 1e3 * n * EPS * nrm
