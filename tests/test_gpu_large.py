"""Parity on the large tile: 4 exposures of 64x64 inputs (n = 16384), 96x96
outputs (m = 9216), Airy (obscuration 0.3) x pixel PSF, U_max = 1e-6, the
whole FP64 path through the C ABI.  The oracle cannot solve this tile in
seconds, so every stage is checked where it can be, at the GPU's own inputs:

  every pixel         U < U_max, kappa inside [kappa_min, kappa_max]
  A, B                rows 0-95 against the oracle's Neville lookup, 1e-13 of max;
                      the lattice lookups (regular grids) against the per-entry
                      kernels over the whole matrices, 1e-13 of max
  P = Q^T B           8 columns against numpy, entrywise 1e-13 of |Q|^T |B|
  eigenpairs          the 4 largest and 4 smallest: |A q - lambda q| <= 100 n eps
                      lambda_max, |q_i . q_j - delta_ij| <= 100 n eps
  bisection           64 pixels: the oracle's bisection at the GPU's (L, P) takes
                      the same decisions (evaluation count) and kappa to 1e-14 in
                      >= 95% of them; the others end within tol of U_max
  T                   8 columns against the oracle's DGEMV form at the GPU's
                      (Q, L, P, kappa), 1e-10 of the columns' max
  leakage identity    16 pixels: U = (C + T^T A T + B^T T) / C from the GPU's A,
                      B, T to 1e-9 (U itself is ~1e-6)
  S, H                the same pixels: sum_i T_ia^2 N_i and T^T I from the GPU's T,
                      1e-12 relative
"""
import numpy as np
import pytest

from imcom_amd import autotune, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu
EPS = 2.0 ** -53


@pytest.fixture(scope="module")
def large():
    from imcom_amd.gpu import ImcomGPU
    t = synth.large()
    print(f"[case] large tile full FP64 solve n={t.n} m={t.m}", flush=True)
    with ImcomGPU(0) as g:
        autotune.apply(g, t.n)
        g.set_tile(t)
        A = g.lookup_A(t.npoly, t.npad, out=True)
        Q, L, cond = g.eigen_A(out=True)
        B = g.lookup_B(t.npoly, t.npad, out=True)
        P = g.build_P(out=True)
        K, U, ev = g.bisect_kappa()
        T = g.build_T(out=True)
        S = g.build_S()
        H = g.build_image()
        lattice = g.lattice_state()
        g.clear_grid()  # the per-entry lookups, for the lattice comparison
        A_pe = g.lookup_A(t.npoly, t.npad, out=True)
        B_pe = g.lookup_B(t.npoly, t.npad, out=True)
    rng = np.random.default_rng(16384)
    return dict(t=t, A=A, Q=Q, L=L, cond=cond, B=B, P=P, K=K, U=U, ev=ev, T=T, S=S, H=H, rng=rng,
                lattice=lattice, A_pe=A_pe, B_pe=B_pe)


def test_large_lattice_lookups_match_per_entry(large):
    """The solve's A and B came from the grid lattices (one interpolation per
    pixel separation); the per-entry kernels give the same matrices to
    1e-13 of their largest entry, and A is exactly symmetric."""
    print("[case] large lattice A, B vs per-entry", flush=True)
    assert large["lattice"] == (True, True)
    A, B = large["A"], large["B"]
    dA = np.abs(A - large["A_pe"]).max()
    dB = np.abs(B - large["B_pe"]).max()
    print(f"[large] lattice vs per-entry: max|dA| {dA:.3e} (max|A| {np.abs(A).max():.3e}), "
          f"max|dB| {dB:.3e} (max|B| {np.abs(B).max():.3e})", flush=True)
    assert dA <= 1e-13 * np.abs(large["A_pe"]).max()
    assert dB <= 1e-13 * np.abs(large["B_pe"]).max()
    assert np.array_equal(A, A.T)


def test_large_every_pixel_meets_target(large):
    t, U, K = large["t"], large["U"], large["K"]
    print(f"[large] cond={large['cond']:.3e} U in [{U.min():.3e}, {U.max():.3e}] "
          f"mean evaluations {large['ev'].mean():.2f}", flush=True)
    assert U.shape == (t.m,)
    assert np.all(U < t.US_max)
    assert np.all((K >= t.kappa_min) & (K <= t.kappa_max))


def test_large_A_B_rows_match_oracle(large):
    t = large["t"]
    print("[case] large A, B rows 0-95 vs oracle", flush=True)
    r = 96
    Ar = O.lookup_A_rows(t, 0, r)
    upper = np.triu(np.ones((r, t.n), dtype=bool))
    dA = np.abs(large["A"][:r, :] - Ar)[upper]
    assert dA.max() <= 1e-13 * np.abs(Ar).max(), dA.max()
    assert np.array_equal(large["A"][:r, :], large["A"][:, :r].T)   # mirrored exactly
    sub = t.__class__(**{**t.__dict__})
    sub.n = r
    sub.x_i, sub.y_i, sub.exp_i = t.x_i[:r], t.y_i[:r], t.exp_i[:r]
    Br = O.lookup_B(sub)
    assert np.abs(large["B"][:r, :] - Br).max() <= 1e-13 * np.abs(Br).max()


def test_large_P_columns(large):
    idx = large["rng"].choice(large["t"].m, 8, replace=False)
    Q, B = large["Q"], large["B"][:, idx]
    ref = Q.T @ B
    bound = np.abs(Q).T @ np.abs(B)
    assert np.all(np.abs(large["P"][:, idx] - ref) <= 1e-13 * bound + 1e-300)


def test_large_extreme_eigenpairs(large):
    A, Q, L, n = large["A"], large["Q"], large["L"], large["t"].n
    assert np.all(np.diff(L) <= 0)
    j = np.r_[0:4, n - 4:n]
    R = A @ Q[:, j] - Q[:, j] * L[j]
    res = np.abs(R).max() / (n * EPS * L[0])
    G = Q[:, j].T @ Q[:, j]
    orth = np.abs(G - np.eye(len(j))).max() / (n * EPS)
    print(f"[large] extreme eigenpairs: residual {res:.3f}, orthogonality {orth:.3f} "
          f"(n eps lambda_max units)", flush=True)
    assert res <= 100 and orth <= 100


def test_large_bisection_at_gpu_inputs(large):
    t, L, P = large["t"], large["L"], large["P"]
    idx = np.sort(large["rng"].choice(t.m, 64, replace=False))
    Kr, Ur, evr = O.bisect_kappa(L, P[:, idx] ** 2, t.C_a, t.US_max, t.kappa_min, t.kappa_max,
                                 t.US_tol, t.maxNbis, True)
    same = (large["ev"][idx] == evr) & (np.abs(large["K"][idx] / Kr - 1.0) <= 1e-14)
    print(f"[large] bisection: same decisions in {same.mean():.3f} of 64 pixels", flush=True)
    assert same.mean() >= 0.95
    assert np.all(Ur < t.US_max)
    assert np.all(np.abs(large["U"][idx][~same] - t.US_max) <= t.US_tol)
    assert np.all(np.abs(Ur[~same] - t.US_max) <= t.US_tol)
    assert np.max(np.abs(large["U"][idx] - Ur)[same]) <= 1e-12


def test_large_T_columns(large):
    idx = np.sort(large["rng"].choice(large["t"].m, 8, replace=False))
    Tr = O.build_T(large["Q"], large["P"][:, idx], large["L"], large["K"][idx])
    dT = np.abs(large["T"][:, idx] - Tr).max(axis=0) / np.abs(Tr).max(axis=0)
    print(f"[large] T columns: max relative difference {dT.max():.3e}", flush=True)
    assert dT.max() <= 1e-10


def test_large_leakage_identity_and_maps(large):
    t, A, B, T = large["t"], large["A"], large["B"], large["T"]
    idx = np.sort(large["rng"].choice(t.m, 16, replace=False))
    Ti = T[:, idx]
    Uid = (t.C_a + np.sum(Ti * (A @ Ti), axis=0) + np.sum(B[:, idx] * Ti, axis=0)) / t.C_a
    dU = np.abs(Uid - large["U"][idx]).max()
    print(f"[large] leakage identity: max |U_id - U| = {dU:.3e}", flush=True)
    assert dU <= 1e-9
    S = np.sum(Ti * Ti * t.Ndiag[:, None], axis=0)
    H = Ti.T @ t.I_i
    assert np.max(np.abs(large["S"][idx] / S - 1.0)) <= 1e-12
    assert np.max(np.abs(large["H"][idx] - H)) <= 1e-12 * (np.abs(Ti).T @ np.abs(t.I_i)).max()
