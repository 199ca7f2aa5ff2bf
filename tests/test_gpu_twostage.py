"""The two-stage reduction to tridiagonal form (sytrd2.hip) through the C
ABI, selected with imcom_gpu_set_reduction_stages(2, min_n) (the default from
n = 7168 up; the tests select it for every size): A to a band of
width 64 by QR panels, the band to tridiagonal form by bulge chasing, and the
eigenvectors back-transformed through both stages (the DSYTRD / DORMTR stages
of DSYEVR, imcom_matrix.f90:642).

  band_reduce      the lower triangle is a band of width 64 (exact zeros
                   below it) whose eigenvalues are A's, within 1e3 n eps ||A||
                   of numpy's eigvalsh (an orthogonal similarity)
  tridiagonalize   eigenvalues of the tridiagonal form, same bound; a second
                   reduction gives bitwise the same d and e
  eigen_A          ||A Q - Q L|| and ||Q^T Q - I|| within 1e3 n eps (||A||),
                   L descending and equal to the one-stage solver's to the
                   same bound; eigenvectors of well-separated eigenvalues
                   equal the one-stage ones up to sign
  reproducible     a second solve gives bitwise the same Q and L

Sizes cover no panel at all (n <= 65), one panel, a partial last panel and
task, and several hundred panels and sweeps; spectra random, graded over 12
decades (as the system matrices'), and clustered (many near-equal
eigenvalues, where the tridiagonal solver deflates).
"""
import numpy as np
import pytest

from imcom_amd import synth
from imcom_amd.gpu import REDUCTION_STAGES, TWOSTAGE_MIN

pytestmark = pytest.mark.gpu
DEFAULT = (REDUCTION_STAGES, TWOSTAGE_MIN)
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
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    if kind == "graded":
        lam = 10.0 ** rng.uniform(-12, 0, n)
    else:  # clustered: four values, each repeated, slightly perturbed
        lam = rng.choice([1.0, 2.0, 5.0, 7.0], n) + 1e-10 * rng.standard_normal(n)
    return np.asfortranarray((Q * lam) @ Q.T)


def _tridiag_eigs(d, e):
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    return np.linalg.eigvalsh(T)


@pytest.fixture
def twostage(gpu_ctx):
    gpu_ctx.set_reduction_stages(2, 3)
    yield gpu_ctx
    gpu_ctx.set_reduction_stages(*DEFAULT)


@pytest.mark.parametrize("n,kind", [(66, "spd"), (67, "spd"), (130, "graded"), (1000, "spd"),
                                    (2500, "graded")])
def test_band_reduce(twostage, n, kind):
    print(f"[case] band reduction n={n} {kind}", flush=True)
    rng = np.random.default_rng(n)
    A = _matrix(n, kind, rng)
    ref = np.linalg.eigvalsh(A)
    nrm = np.abs(A).max()
    twostage.set_tile(_tile(n))
    twostage.set_A(A)
    Ab = twostage.band_reduce()
    L = np.tril(Ab)
    i, j = np.indices((n, n))
    assert np.all(L[(i - j) > 64] == 0.0)
    Bm = L + np.tril(L, -1).T
    err = np.abs(np.linalg.eigvalsh(Bm) - ref).max() / nrm
    print(f"[band] n={n}: eigenvalue error / ||A|| {err:.2e}", flush=True)
    assert err <= 1e3 * n * EPS


@pytest.mark.parametrize("n,kind", [(3, "spd"), (4, "spd"), (65, "spd"), (66, "graded"),
                                    (129, "spd"), (1000, "graded"), (3000, "spd"),
                                    (4100, "clustered")])
def test_twostage_tridiagonal_form(twostage, n, kind):
    print(f"[case] two-stage tridiagonal form n={n} {kind}", flush=True)
    rng = np.random.default_rng(n + 7)
    A = _matrix(n, kind, rng)
    ref = np.linalg.eigvalsh(A)
    nrm = np.abs(A).max()
    twostage.set_tile(_tile(n))
    twostage.set_A(A)
    d, e = twostage.tridiagonalize()
    err = np.abs(_tridiag_eigs(d, e) - ref).max() / nrm
    # a second reduction of the same matrix gives bitwise the same d, e
    twostage.set_A(A)
    d2, e2 = twostage.tridiagonalize()
    same = np.array_equal(d, d2) and np.array_equal(e, e2)
    err2 = err if same else np.abs(_tridiag_eigs(d2, e2) - ref).max() / nrm
    print(f"[tridiagonal] n={n}: eigenvalue error / ||A|| {err:.2e} (second run "
          f"{'identical' if same else f'differs, error {err2:.2e}'})", flush=True)
    assert err <= 1e3 * n * EPS and err2 <= 1e3 * n * EPS
    assert same


@pytest.mark.parametrize("n,kind", [(3, "spd"), (65, "graded"), (67, "spd"), (128, "clustered"),
                                    (129, "spd"), (700, "graded"), (1000, "spd"),
                                    (2049, "clustered"), (3000, "graded")])
def test_twostage_eigen(gpu_ctx, n, kind):
    print(f"[case] two-stage eigen_A n={n} {kind}", flush=True)
    rng = np.random.default_rng(n + 11)
    A = _matrix(n, kind, rng)
    nrm = np.abs(A).max()
    gpu_ctx.set_tile(_tile(n))
    try:
        gpu_ctx.set_reduction_stages(1, TWOSTAGE_MIN)
        gpu_ctx.set_A(A)
        Q1, L1, _ = gpu_ctx.eigen_A()
        gpu_ctx.set_reduction_stages(2, 3)
        gpu_ctx.set_A(A)
        Q2, L2, _ = gpu_ctx.eigen_A()
        gpu_ctx.set_A(A)
        Q3, L3, _ = gpu_ctx.eigen_A()
    finally:
        gpu_ctx.set_reduction_stages(*DEFAULT)
    res = np.abs(A @ Q2 - Q2 * L2).max() / nrm
    orth = np.abs(Q2.T @ Q2 - np.eye(n)).max()
    ref = np.sort(np.linalg.eigvalsh(A))[::-1]
    lerr = np.abs(L2 - ref).max() / nrm
    ldiff = np.abs(L2 - L1).max() / nrm
    print(f"[two-stage] n={n}: residual {res:.2e}, orthogonality {orth:.2e}, eigenvalues "
          f"{lerr:.2e} (one-stage difference {ldiff:.2e})", flush=True)
    assert res <= 1e3 * n * EPS
    assert orth <= 1e3 * n * EPS
    assert lerr <= 1e3 * n * EPS and ldiff <= 1e3 * n * EPS
    assert np.all(np.diff(L2) <= 0.0)
    # eigenvectors of eigenvalues separated from their neighbours by more
    # than 1e-3 ||A|| agree with the one-stage solver's up to sign
    gap = np.minimum(np.abs(np.diff(L2, prepend=np.inf)), np.abs(np.diff(L2, append=-np.inf)))
    sep = np.nonzero(gap > 1e-3 * nrm)[0]
    if sep.size:
        dots = np.abs(np.sum(Q1[:, sep] * Q2[:, sep], axis=0))
        print(f"[two-stage] {sep.size} separated eigenvectors, min |q1.q2| {dots.min():.15f}",
              flush=True)
        assert np.all(np.abs(dots - 1.0) <= 1e6 * n * EPS)
    assert np.array_equal(Q2, Q3) and np.array_equal(L2, L3)


def test_twostage_below_threshold_is_one_stage(gpu_ctx):
    """Sizes below min_n take the one-stage path: bitwise its results (the
    default threshold leaves every tile below n = 7168 on it)."""
    n = 300
    print(f"[case] two-stage threshold n={n}", flush=True)
    rng = np.random.default_rng(5)
    A = _matrix(n, "spd", rng)
    gpu_ctx.set_tile(_tile(n))
    try:
        gpu_ctx.set_reduction_stages(1, TWOSTAGE_MIN)
        gpu_ctx.set_A(A)
        Q1, L1, _ = gpu_ctx.eigen_A()
        gpu_ctx.set_reduction_stages(2, n + 1)
        gpu_ctx.set_A(A)
        Q2, L2, _ = gpu_ctx.eigen_A()
    finally:
        gpu_ctx.set_reduction_stages(*DEFAULT)
    assert np.array_equal(Q1, Q2) and np.array_equal(L1, L2)
