"""GPU parity of each solve-path stage against the CPU oracle, through the C ABI.

Tolerances (FP64 path):
  A, B           |gpu - oracle| <= 1e-13 * max|oracle|   (Lagrange vs Neville form)
  P = Q^T B      |gpu - numpy| <= 1e-13 * (|Q|^T |B|) entrywise bound
  kappa          identical to the oracle's in every pixel where no bisection
                 decision is within rounding of U_max (checked: >= 99% equal,
                 the rest within one bisection step and |U - U_max| < tol)
  U, S           |gpu - oracle| <= 1e-12 (U is ~1e-6; S relative 1e-10)
  T, H           relative 1e-10 at the same (Q, L, P, kappa)
  eigen          ||AQ - QL|| <= 1e3 n eps ||A||, ||Q^T Q - I|| <= 1e3 n eps,
                 eigenvalues within 1e3 n eps ||A|| of LAPACK DSYEVR
"""
import sys

import numpy as np
import pytest

from imcom_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu
EPS = 2.0 ** -53


def _say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def small_tile():
    return synth.small()


@pytest.fixture(scope="module")
def small_ref(small_tile):
    return O.solve(small_tile)


@pytest.mark.parametrize("n,m", [(37, 23), (300, 200), (1024, 576), (4096, 2048)])
def test_gemm_P_random(gpu_ctx, n, m):
    """P = Q^T B on the library's MFMA kernel against numpy, entrywise within
    1e-13 of the |Q|^T |B| bound."""
    _say(f"gemm P=Q^T B n={n} m={m}")
    rng = np.random.default_rng(n + m)
    t = synth.tiny()
    t.n, t.m = n, m
    t.grid = None  # the coordinates below are not on the tile's grids
    t.x_i = rng.uniform(0, 8, n)
    t.y_i = rng.uniform(0, 8, n)
    t.exp_i = np.ones(n, dtype=np.int32)
    t.X_a = rng.uniform(0, 8, m)
    t.Y_a = rng.uniform(0, 8, m)
    t.Ndiag = np.full(n, 1e-6)
    gpu_ctx.set_tile(t)
    Q = np.asfortranarray(rng.standard_normal((n, n)))
    B = np.asfortranarray(rng.standard_normal((n, m)))
    L = np.sort(rng.uniform(0.1, 1.0, n))[::-1].copy()
    gpu_ctx.set_QL(Q, L)
    gpu_ctx.set_B(B)
    P = gpu_ctx.build_P(out=True)
    ref = Q.T @ B
    bound = np.abs(Q).T @ np.abs(B)
    assert np.all(np.abs(P - ref) <= 1e-13 * bound + 1e-300)


def test_lookup_A_B_tiny(gpu_ctx):
    _say("lookup A/B tiny")
    t = synth.tiny()
    gpu_ctx.set_tile(t)
    A = gpu_ctx.lookup_A(out=True)
    B = gpu_ctx.lookup_B(out=True)
    Ar = O.lookup_A(t)
    Br = O.lookup_B(t)
    assert np.array_equal(A, A.T)
    assert np.max(np.abs(A - Ar)) <= 1e-13 * np.max(np.abs(Ar))
    assert np.max(np.abs(B - Br)) <= 1e-13 * np.max(np.abs(Br))


def test_lookup_A_B_displaced_psf(gpu_ctx):
    """The device lookups on a tile whose second exposure's PSF is displaced
    by s = (0.75, -0.5) px (different PSF per exposure, psfconst = 0): equal
    to the oracle to 1e-13 and to the closed forms with the shifted arguments
    (oracle/closed_forms.py), which fix the sign and axis of x_j - x_i and
    x_i - X_a."""
    from oracle import closed_forms as CF
    _say("lookup A/B displaced PSF")
    t = synth.displaced_psf_tile()
    mt = t.meta
    shifts = [(0.0, 0.0), mt["shift"]]
    gpu_ctx.set_tile(t)
    A = gpu_ctx.lookup_A(out=True)
    B = gpu_ctx.lookup_B(out=True)
    Ar, Br = O.lookup_A(t), O.lookup_B(t)
    assert np.max(np.abs(A - Ar)) <= 1e-13 * np.max(np.abs(Ar))
    assert np.max(np.abs(B - Br)) <= 1e-13 * np.max(np.abs(Br))
    Ac = CF.A_closed_form(t, shifts, mt["psf_sigma"], mt["psfscale"])
    Bc = CF.B_closed_form(t, shifts, mt["psf_sigma"], mt["gamma_sigma"], mt["psfscale"])
    assert np.abs(A - Ac).max() <= 1e-8 * Ac.max()
    assert np.abs(B - Bc).max() <= 1e-6 * np.abs(Bc).max()


def test_lookup_A_B_small(gpu_ctx, small_tile, small_ref):
    _say("lookup A/B small n=1024 m=576")
    gpu_ctx.set_tile(small_tile)
    A = gpu_ctx.lookup_A(out=True)
    B = gpu_ctx.lookup_B(out=True)
    assert np.array_equal(A, A.T)
    assert np.max(np.abs(A - small_ref["A"])) <= 1e-13 * np.max(np.abs(small_ref["A"]))
    assert np.max(np.abs(B - small_ref["B"])) <= 1e-13 * np.max(np.abs(small_ref["B"]))


def test_bisect_at_oracle_inputs(gpu_ctx, small_tile, small_ref):
    _say("bisect U-mode at oracle (L, P)")
    t, r = small_tile, small_ref
    gpu_ctx.set_tile(t)
    gpu_ctx.set_QL(r["Q"], r["L"])
    gpu_ctx.set_P(r["P"])
    K, U, ev = gpu_ctx.bisect_kappa()
    # same decision path (evaluation count) and kappa up to the last bits of
    # 10**root (device exp10 vs host pow)
    same = (ev == r["evals"]) & (np.abs(K / r["K"] - 1.0) <= 1e-14)
    assert same.mean() >= 0.99, f"kappa equal in {same.mean():.3f} of pixels"
    assert np.max(np.abs(U - r["U"])) <= 1e-12
    assert np.all(U < t.US_max)
    # T, S, H at identical inputs
    T = gpu_ctx.build_T()
    Tr = O.build_T(r["Q"], r["P"], r["L"], K)
    assert np.max(np.abs(T - Tr)) <= 1e-10 * np.max(np.abs(Tr))
    S = gpu_ctx.build_S()
    Sr = O.build_S(Tr, t.Ndiag)
    assert np.max(np.abs(S - Sr) / np.abs(Sr)) <= 1e-10
    H = gpu_ctx.build_image()
    Hr = O.build_image(Tr, t.I_i)
    assert np.max(np.abs(H - Hr)) <= 1e-10 * np.max(np.abs(Hr))


def test_bisect_S_mode(gpu_ctx, small_tile, small_ref):
    _say("bisect S-mode at oracle (L, P)")
    t, r = small_tile, small_ref
    gpu_ctx.set_tile(t)
    gpu_ctx.set_QL(r["Q"], r["L"])
    gpu_ctx.set_P(r["P"])
    Smax = 2.0 * np.median(r["S"])
    K, S, ev = gpu_ctx.bisect_kappa(UnotS=False, US=Smax, tol=1e-3 * Smax)
    Kr, Sr, evr = O.bisect_kappa(r["L"], r["P2"], t.C_a, Smax, t.kappa_min, t.kappa_max,
                                 1e-3 * Smax, 53, UnotS=False, Ndiag=t.Ndiag)
    same = (ev == evr) & (np.abs(K / Kr - 1.0) <= 1e-14)
    assert same.mean() >= 0.99, f"kappa equal in {same.mean():.3f} of pixels"
    assert np.max(np.abs(S - Sr) / Sr) <= 1e-10
    U = gpu_ctx.build_U()
    Ur = O.build_U(r["L"], r["P2"], t.C_a, K)
    assert np.max(np.abs(U - Ur)) <= 1e-12


def _tridiag(d, e):
    return np.diag(d) + np.diag(e, 1) + np.diag(e, -1)


@pytest.mark.parametrize("kind,n", [("random", 1), ("random", 2), ("random", 7), ("random", 100),
                                    ("random", 1000), ("clustered", 513), ("graded", 800),
                                    ("wilkinson", 201), ("zero_e", 64)])
def test_tridiag_eig(gpu_ctx, kind, n):
    _say(f"tridiag_eig {kind} n={n}")
    rng = np.random.default_rng(n)
    if kind == "random":
        d, e = rng.standard_normal(n), rng.standard_normal(max(n - 1, 0))
    elif kind == "clustered":
        d = np.concatenate([np.full(n // 2, 1.0), rng.uniform(0, 1e-12, n - n // 2)])
        e = rng.uniform(0, 1e-9, n - 1)
    elif kind == "graded":
        d = 10.0 ** rng.uniform(-16, 0, n)
        e = 10.0 ** rng.uniform(-17, -1, n - 1)
    elif kind == "wilkinson":
        m = (n - 1) // 2
        d = np.abs(np.arange(-m, m + 1, dtype=float))
        e = np.ones(n - 1)
    else:
        d, e = rng.standard_normal(n), np.zeros(n - 1)
    lam, Z = gpu_ctx.tridiag_eig(d, e)
    Tm = _tridiag(d, e)
    nrm = max(np.max(np.abs(Tm)), 1e-300)
    ref = np.linalg.eigvalsh(Tm)
    assert np.all(np.diff(lam) >= 0)
    assert np.max(np.abs(lam - ref)) <= 1e3 * n * EPS * nrm
    assert np.max(np.abs(Z.T @ Z - np.eye(n))) <= 1e3 * n * EPS
    assert np.max(np.abs(Tm @ Z - Z * lam)) <= 1e3 * n * EPS * nrm


def test_tridiagonalize_small(gpu_ctx, small_tile, small_ref):
    _say("tridiagonalize small")
    gpu_ctx.set_tile(small_tile)
    gpu_ctx.set_A(small_ref["A"])
    d, e = gpu_ctx.tridiagonalize()
    lam = np.linalg.eigvalsh(_tridiag(d, e))
    ref = np.sort(small_ref["L"])
    n = small_tile.n
    nrm = np.max(np.abs(small_ref["A"]))
    assert np.max(np.abs(lam - ref)) <= 1e3 * n * EPS * nrm


@pytest.mark.parametrize("n,nb,nbb", [(5, 32, 128), (130, 32, 128), (700, 32, 128),
                                       (700, 64, 256), (1100, 16, 256), (700, 64, 512),
                                       (1300, 32, 512)])
def test_eigen_random_spd(gpu_ctx, n, nb, nbb):
    _say(f"eigen_A random n={n} nb={nb} nbb={nbb}")
    rng = np.random.default_rng(3 * n)
    X = rng.standard_normal((n, n))
    A = np.asfortranarray(X @ X.T / n)
    t = synth.tiny()
    t.n, t.m = n, 4
    t.grid = None  # the coordinates below are not on the tile's grids
    t.x_i, t.y_i = np.zeros(n), np.zeros(n)
    t.exp_i = np.ones(n, dtype=np.int32)
    t.X_a, t.Y_a = np.zeros(4), np.zeros(4)
    t.Ndiag = np.ones(n)
    gpu_ctx.set_tile(t)
    gpu_ctx.set_A(A)
    gpu_ctx.set_tuning(nb, nbb)
    try:
        Q, L, cond = gpu_ctx.eigen_A()
    finally:
        gpu_ctx.set_tuning(32, 128)
    nrm = np.max(np.abs(A))
    assert np.all(np.diff(L) <= 0)
    assert np.max(np.abs(A @ Q - Q * L)) <= 1e3 * n * EPS * nrm
    assert np.max(np.abs(Q.T @ Q - np.eye(n))) <= 1e3 * n * EPS
    _, Lr, _ = O.eigen_A(A)
    assert np.max(np.abs(L - Lr)) <= 1e3 * n * EPS * nrm


def test_eigen_maximum_size(gpu_ctx):
    """n = 32768, the largest tile the C ABI accepts: A = H D H with the
    reflector H = I - 2 u u^T has the eigenvalues diag(D) exactly and the
    eigenvectors H e_i.  Checks every eigenvalue, the descending order and
    four eigenpairs' residuals."""
    n = 32768
    _say(f"eigen_A reflected diagonal n={n}")
    rng = np.random.default_rng(32768)
    d = 1.0 + rng.permutation(n) / n           # distinct, in [1, 2)
    u = rng.standard_normal(n)
    u /= np.linalg.norm(u)
    du = d * u
    A = np.multiply.outer(u, 4.0 * float(u @ du) * u - 2.0 * du)
    A -= np.multiply.outer(2.0 * du, u)
    A[np.diag_indices(n)] += d
    t = synth.tiny()
    t.n, t.m = n, 4
    t.grid = None  # the coordinates below are not on the tile's grids
    t.x_i, t.y_i = np.zeros(n), np.zeros(n)
    t.exp_i = np.ones(n, dtype=np.int32)
    t.X_a, t.Y_a = np.zeros(4), np.zeros(4)
    t.Ndiag = np.ones(n)
    gpu_ctx.set_tile(t)
    gpu_ctx.set_A(A.T)                          # symmetric: the transpose view is Fortran-ordered
    gpu_ctx.set_tuning(64, 256)
    try:
        Q, L, _ = gpu_ctx.eigen_A()
    finally:
        gpu_ctx.set_tuning(32, 128)
    nrm = 2.0
    assert np.all(np.diff(L) <= 0)
    err = np.max(np.abs(L - np.sort(d)[::-1]))
    print(f"[eigen n={n}] max |lambda - d| = {err:.3e} = {err / (n * EPS * nrm):.3f} n eps |A|")
    assert err <= 10 * n * EPS * nrm
    for j in (0, 1, n // 2, n - 1):
        q = Q[:, j]
        r = A @ q - L[j] * q
        assert np.max(np.abs(r)) <= 10 * n * EPS * nrm
        assert abs(np.linalg.norm(q) - 1.0) <= 1e3 * n * EPS
    del A, Q


@pytest.mark.parametrize("nb,nbb", [(32, 128), (16, 64), (64, 128), (32, 256)])
def test_eigen_accuracy_report(gpu_ctx, nb, nbb):
    """Backward error, orthogonality and eigenvalue error of the GPU
    eigensolver, in units of n*eps (printed), on a tile whose pixels sit near
    the leakage floor (n=128, condition ~5e11), for each panel width and
    back-transform block the tuner may pick.  LAPACK's DSYEVR reaches a few
    units on this matrix."""
    _say(f"eigen accuracy nb={nb} nbb={nbb}")
    t = synth.tiny(seed=100, nexp=2, npix_in=8, nout=8, psf_sigma=1.1, gamma_sigma=1.8)
    A = O.lookup_A(t)
    n = t.n
    gpu_ctx.set_tuning(nb, nbb)
    try:
        gpu_ctx.set_tile(t)
        gpu_ctx.set_A(A)
        d, e = gpu_ctx.tridiagonalize()
        gpu_ctx.set_A(A)
        Q, L, _ = gpu_ctx.eigen_A()
    finally:
        gpu_ctx.set_tuning(32, 128)
    nrm = np.linalg.norm(A, 2)
    unit = n * EPS
    res = np.max(np.abs(A @ Q - Q * L)) / (unit * nrm)
    orth = np.max(np.abs(Q.T @ Q - np.eye(n))) / unit
    Qr, Lr, _ = O.eigen_A(A)
    lam = np.max(np.abs(L - Lr)) / (unit * nrm)
    tri = np.max(np.abs(np.linalg.eigvalsh(_tridiag(d, e)) - np.sort(Lr))) / (unit * nrm)
    res_r = np.max(np.abs(A @ Qr - Qr * Lr)) / (unit * nrm)
    orth_r = np.max(np.abs(Qr.T @ Qr - np.eye(n))) / unit
    print(f"[eig-accuracy] nb={nb} nbb={nbb} n={n} residual={res:.2f} orth={orth:.2f} "
          f"eigenvalues={lam:.2f} tridiagonal={tri:.2f} (n*eps units; LAPACK residual={res_r:.2f} "
          f"orth={orth_r:.2f})", flush=True)
    assert res <= 100 and orth <= 100 and lam <= 100 and tri <= 100


def test_eigen_small(gpu_ctx, small_tile, small_ref):
    _say("eigen_A small n=1024")
    gpu_ctx.set_tile(small_tile)
    A = small_ref["A"]
    gpu_ctx.set_A(A)
    Q, L, cond = gpu_ctx.eigen_A()
    n = small_tile.n
    nrm = np.max(np.abs(A))
    assert np.max(np.abs(A @ Q - Q * L)) <= 1e3 * n * EPS * nrm
    assert np.max(np.abs(Q.T @ Q - np.eye(n))) <= 1e3 * n * EPS
    assert np.max(np.abs(L - small_ref["L"])) <= 1e3 * n * EPS * nrm


def test_solve_small_end_to_end(gpu_ctx, small_tile, small_ref):
    """Whole path on the GPU against the oracle (LAPACK eigensolver): T, U, S,
    H compared through eigenvector-free quantities."""
    _say("solve small end-to-end")
    from oracle import tolerances as TOL
    t, r = small_tile, small_ref
    out = gpu_ctx.solve(t)
    assert np.all(out["U"] < t.US_max)
    ok = r["evals"] > 1
    assert np.all((np.abs(out["U"] - r["U"]) <= TOL.u_tol(r["L"], r["P2"], r["K"], t.C_a, t.n))[ok])
    # kappa: same bisection decisions in almost every pixel
    close = np.abs(out["K"] / r["K"] - 1.0) <= 1e-12
    assert close.mean() >= 0.95
    # leakage identity from T itself: U = (C + T^T A T + B^T T) / C
    A, B, T = r["A"], r["B"], out["T"]
    Uid = (t.C_a + np.einsum("ia,ia->a", T, A @ T) + np.einsum("ia,ia->a", B, T)) / t.C_a
    assert np.max(np.abs(Uid - out["U"])) <= 1e-10
    # T, S, H per pixel within the conditioning bounds where kappa agrees
    same = close & ok
    tT = TOL.t_tol(r["L"], r["P"], r["K"], t.n, np.abs(r["T"]).max())
    dT = np.max(np.abs(T - r["T"]), axis=0)
    print(f"[small e2e] kappa equal {close.mean():.3f}, max dT/t_tol {np.max((dT / tT)[same]):.3e}",
          flush=True)
    assert np.all((dT <= tT)[same])
    assert np.all((np.abs(out["S"] - r["S"]) <= TOL.S_tol(tT, r["T"], t.Ndiag, r["S"]))[same])
    assert np.all((np.abs(out["H"] - r["H"]) <= TOL.H_tol(tT, r["H"], t.I_i))[same])
