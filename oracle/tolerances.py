"""Per-pixel parity bounds for the FP64 solve path (test infrastructure:
used by tests/ and __graft_entry__.smoke() only).

A's spectrum spans many decades, so two correct eigensolvers (LAPACK's
DSYEVR and the GPU's) differ by eigenvalue errors of about n eps lambda_max;
the outputs move by the first-order change that such errors cause, and that
change is large exactly where a pixel leans on the smallest eigenvalues.
The bounds below are 100 x that first-order change plus a floor:

  U_a   dU/dlambda_k = P_ka^2 (lambda_k + 3 kappa) / (4 C (lambda_k + kappa)^3)
        u_tol = 1e-12 + 100 sum_k |dU/dlambda_k| n eps lambda_max
  T_.a  T = -1/2 Q P / (L + kappa): |dT_ia| <= 1/2 sum_k |P_ka| d / (lambda_k + kappa_a)^2
        with |Q| <= 1 and d = n eps lambda_max;
        t_tol = 1e-10 max|T| + 100 x that sum
  S_a   sum_i T_ia^2 N_i:  1e-10 S + 2 t_tol sum_i |T_ia| N_i
  H_a   sum_i T_ia I_i:    1e-10 max|H| + t_tol sum_i |I_i|
"""
import numpy as np

EPS = np.finfo(np.float64).eps


def u_tol(L, P2, K, C_a, n):
    L = np.asarray(L)
    w = np.abs(L[:, None] + 3 * K[None, :]) / np.abs(L[:, None] + K[None, :]) ** 3
    d = n * EPS * np.max(np.abs(L))
    return 1e-12 + 100.0 * d / (4 * C_a) * np.sum(P2 * w, axis=0)


def t_tol(L, P, K, n, Tmax):
    L = np.asarray(L)
    d = n * EPS * np.max(np.abs(L))
    g = 0.5 * np.sum(np.abs(P) * d / (L[:, None] + K[None, :]) ** 2, axis=0)
    return 1e-10 * Tmax + 100.0 * g


def S_tol(tT, T, Ndiag, S):
    return 1e-10 * np.abs(S) + 2 * tT * (np.abs(T).T @ np.asarray(Ndiag))


def H_tol(tT, H, I):
    return 1e-10 * np.max(np.abs(H)) + tT * np.sum(np.abs(np.asarray(I)))
