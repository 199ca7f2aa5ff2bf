"""CPU oracle of the IMCOM solve path -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg, as the reference the GPU results are compared with.  The GPU library
never calls it.

Pieces and what follows the reference (file:line):
  lookup_A / lookup_B   C restatement (imcom_oracle.c) of imcom_lookup_A/B and
                        the Neville interpolation imcom_interp_lookup ->
                        imcom_polin2 -> imcom_polint
                        (imcom_matrix.f90:206-357,497-542; imcom_proc.f90:351-429)
  eigen_A               LAPACK DSYEVR('V','A','U', abstol=DLAMCH('S')) through
                        scipy's OpenBLAS 0.3.29, reversed to descending order
                        (imcom_matrix.f90:586-678) -- the reference's own routine
  build_P               DGEMM('T','N') P = Q^T B, P2 = P*P (imcom_matrix.f90:682-712)
  bisect_kappa          C restatement of imcom_bisect_kappa/by_U/by_S
                        (imcom_bisect.f90:78-267) over imcom_U/S_from_kappa
                        (imcom_matrix.f90:904-970)
  build_T/U/S/image     C restatements (imcom_matrix.f90:716-877)

How each piece is pinned (tests/test_oracle.py, tests/test_tables_oracle.py):
  interp        exact reproduction of degree<=7 polynomials sampled on a table
                (8-point and edge-clipped stencils), exact node values, and
                the Gaussian closed form A(D) = s^2/(4 pi sig^2) exp(-D^2/4sig^2)
  lookup tables Gaussian closed forms for A, B = -2 x cross-correlation, and
                Parseval for C
  eigen_A       ||AQ - QL|| and ||Q^TQ - I|| at n*eps; numpy eigvalsh
  build_P       brute-force sum
  U/S(kappa)    the identity U = (C + T^T A T + B^T T)/C with T from the
                direct solve T = -1/2 (A + kappa I)^-1 B; S = sum T^2 N
  bisection     bracket invariants U(k) < U_max <= U(k*10^dx) or k = k_min,
                and an independent Python restatement on small inputs
  build_T       T = -1/2 (A + kappa I)^-1 B by LAPACK solve
Summation order: DDOT/DGEMV/DGEMM sums are restated as ordered sums; the
reference's OpenBLAS kernels block their sums differently (last-bit level).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


def build(force=False):
    """Compile the C restatement into oracle/liboracle.so."""
    so = os.path.join(_HERE, "liboracle.so")
    src = os.path.join(_HERE, "imcom_oracle.c")
    if force or not os.path.exists(so) or os.path.getmtime(so) < os.path.getmtime(src):
        subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle.so"])
    return so


def lib():
    global _LIB
    if _LIB is None:
        so = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(so):
            build()
        L = ctypes.CDLL(so)
        L.or_interp_lookup.restype = ctypes.c_double
        L.or_interp_lookup.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_int]
        L.or_U_from_kappa.restype = ctypes.c_double
        L.or_S_from_kappa.restype = ctypes.c_double
        L.or_bisect_by_U.restype = ctypes.c_int
        L.or_bisect_by_S.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(_dp)


def _f(a, dtype=np.float64):
    return np.asfortranarray(a, dtype=dtype)


def interp_lookup(table, xarg, yarg, npoly=7):
    t = _f(table)
    return lib().or_interp_lookup(t.shape[0], t.shape[1], _p(t), float(xarg), float(yarg), npoly)


def lookup_A(tile):
    n = tile.n
    A = np.zeros((n, n), order="F")
    tab = _f(tile.Alookup)
    ex = np.ascontiguousarray(tile.exp_i, dtype=np.int32)
    x, y = np.ascontiguousarray(tile.x_i), np.ascontiguousarray(tile.y_i)
    lib().or_lookup_A(ctypes.c_int64(n), _p(x), _p(y), ex.ctypes.data_as(_ip),
                      ctypes.c_int(tile.psfconst), ctypes.c_int(tab.shape[0]),
                      ctypes.c_int(tab.shape[1]), _p(tab), ctypes.c_int(tile.npad),
                      ctypes.c_int(tile.npoly), ctypes.c_double(tile.psfxscale),
                      ctypes.c_double(tile.psfyscale), _p(A))
    return A


def lookup_A_rows(tile, i0, i1):
    """Rows [i0, i1) of A (entries j >= i, as imcom_lookup_A fills them,
    imcom_matrix.f90:223-257), returned as an (i1 - i0) x n array."""
    n = tile.n
    A = np.zeros((n, n), order="F")
    tab = _f(tile.Alookup)
    ex = np.ascontiguousarray(tile.exp_i, dtype=np.int32)
    x, y = np.ascontiguousarray(tile.x_i), np.ascontiguousarray(tile.y_i)
    lib().or_lookup_A_rows(ctypes.c_int64(n), ctypes.c_int64(i0), ctypes.c_int64(i1), _p(x), _p(y),
                           ex.ctypes.data_as(_ip), ctypes.c_int(tile.psfconst),
                           ctypes.c_int(tab.shape[0]), ctypes.c_int(tab.shape[1]), _p(tab),
                           ctypes.c_int(tile.npad), ctypes.c_int(tile.npoly),
                           ctypes.c_double(tile.psfxscale), ctypes.c_double(tile.psfyscale), _p(A))
    return np.array(A[i0:i1, :])


def lookup_B(tile):
    n, m = tile.n, tile.m
    B = np.zeros((n, m), order="F")
    tab = _f(tile.Blookup)
    ex = np.ascontiguousarray(tile.exp_i, dtype=np.int32)
    arrs = [np.ascontiguousarray(v) for v in (tile.x_i, tile.y_i, tile.X_a, tile.Y_a)]
    lib().or_lookup_B(ctypes.c_int64(n), ctypes.c_int64(m), _p(arrs[0]), _p(arrs[1]),
                      ex.ctypes.data_as(_ip), _p(arrs[2]), _p(arrs[3]),
                      ctypes.c_int(tile.psfconst), ctypes.c_int(tab.shape[0]),
                      ctypes.c_int(tab.shape[1]), _p(tab), ctypes.c_int(tile.npad),
                      ctypes.c_int(tile.npoly), ctypes.c_double(tile.psfxscale),
                      ctypes.c_double(tile.psfyscale), _p(B))
    return B


def eigen_A(A):
    """DSYEVR('V','A','U') with abstol = DLAMCH('Safe minimum'), reversed to
    descending eigenvalues (imcom_matrix.f90:586-678).  Returns (Q, L, cond)."""
    from scipy.linalg import lapack
    w, z, mfound, _isuppz, info = lapack.dsyevr(_f(A), compute_v=1, range="A", lower=0,
                                                abstol=np.finfo(np.float64).tiny)
    if info != 0:
        raise RuntimeError(f"DSYEVR info={info}")
    L = w[::-1].copy()
    Q = np.asfortranarray(z[:, ::-1])
    return Q, L, float(L.max() / L.min())


def build_P(Q, B):
    P = np.asfortranarray(Q.T @ B)
    return P, P * P


def bisect_kappa(L, P2, C_a, US, k_min, k_max, tol, maxnits=53, UnotS=True, Ndiag=None):
    n, m = P2.shape
    P2 = _f(P2)
    L = np.ascontiguousarray(L, dtype=np.float64)
    Nd = np.ascontiguousarray(Ndiag if Ndiag is not None else np.zeros(n), dtype=np.float64)
    K = np.zeros(m)
    US_a = np.zeros(m)
    ev = np.zeros(m, dtype=np.int32)
    lib().or_bisect_kappa(ctypes.c_int(1 if UnotS else 0), ctypes.c_int64(n),
                          ctypes.c_int64(m), _p(L), _p(P2), _p(Nd), ctypes.c_double(C_a),
                          ctypes.c_double(US), ctypes.c_double(k_min), ctypes.c_double(k_max),
                          ctypes.c_double(tol), ctypes.c_int(maxnits), _p(K), _p(US_a),
                          ev.ctypes.data_as(_ip))
    return K, US_a, ev


def U_from_kappa(L, P2col, C_a, k):
    L = np.ascontiguousarray(L, dtype=np.float64)
    p = np.ascontiguousarray(P2col, dtype=np.float64)
    return lib().or_U_from_kappa(ctypes.c_int64(L.shape[0]), _p(L), _p(p),
                                 ctypes.c_double(C_a), ctypes.c_double(k))


def S_from_kappa(L, P2col, Ndiag, k):
    L = np.ascontiguousarray(L, dtype=np.float64)
    p = np.ascontiguousarray(P2col, dtype=np.float64)
    Nd = np.ascontiguousarray(Ndiag, dtype=np.float64)
    return lib().or_S_from_kappa(ctypes.c_int64(L.shape[0]), _p(L), _p(p), _p(Nd),
                                 ctypes.c_double(k))


def build_T(Q, P, L, K):
    n, m = P.shape
    T = np.zeros((n, m), order="F")
    Q, P = _f(Q), _f(P)
    L = np.ascontiguousarray(L, dtype=np.float64)
    K = np.ascontiguousarray(K, dtype=np.float64)
    lib().or_build_T(ctypes.c_int64(n), ctypes.c_int64(m), _p(Q), _p(P), _p(L), _p(K), _p(T))
    return T


def build_image(T, I_i):
    n, m = T.shape
    H = np.zeros(m)
    T = _f(T)
    I_i = np.ascontiguousarray(I_i, dtype=np.float64)
    lib().or_build_image(ctypes.c_int64(n), ctypes.c_int64(m), _p(T), _p(I_i), _p(H))
    return H


def build_U(L, P2, C_a, K):
    n, m = P2.shape
    U = np.zeros(m)
    P2 = _f(P2)
    L = np.ascontiguousarray(L, dtype=np.float64)
    K = np.ascontiguousarray(K, dtype=np.float64)
    lib().or_build_U(ctypes.c_int64(n), ctypes.c_int64(m), _p(L), _p(P2), ctypes.c_double(C_a),
                     _p(K), _p(U))
    return U


def build_S(T, Ndiag):
    n, m = T.shape
    S = np.zeros(m)
    T = _f(T)
    Nd = np.ascontiguousarray(Ndiag, dtype=np.float64)
    lib().or_build_S(ctypes.c_int64(n), ctypes.c_int64(m), _p(T), _p(Nd), _p(S))
    return S


def sampled_cpu_time(tile, L=None, P=None, rows_A=96, rows_B=96, n_eig=2048, cols_P=128,
                     pix_bisect=48, pix_T=16, checks=False, full_eig=False):
    """Time the oracle's solve path on bounded samples of one tile and scale
    each stage to the whole tile (the CPU-baseline leg of bench.py):

      lookup_A   rows [0, rows_A) of the upper triangle, scaled by entry count
      lookup_B   the first rows_B input pixels x all m outputs
      eigen_A    DSYEVR on the leading n_eig x n_eig block of A, scaled by n^3
      build_P    DGEMM Q^T B[:, :cols_P] at full n, scaled by m
      bisect     pix_bisect pixels at full n, scaled by m
      build_T    pix_T DGEMVs at full n, scaled by m

    With full_eig=True, lookup_A and DSYEVR run on the whole tile instead
    (timed, not scaled): DSYEVR's time on a leading or strided block of A does
    not follow n^3 to the whole matrix (measured at n = 16384 on 16 threads:
    blocks of 4096 and 8192 take 4.2 and 21.6 s, the whole matrix 68 s),
    because the block's spectrum is not the whole matrix's.
    L and P (the tile's eigenvalues and projections, any source) feed the
    bisection sample.  With checks=True the scale factors are tested by a
    second sample of twice the size: DSYEVR at 2 n_eig (the fitted exponent of
    n is reported and the larger sample is the one scaled), A rows 2 rows_A
    and T on 2 pix_T pixels (the ratios to the first samples are reported;
    linear scaling gives 2).  Returns (seconds for the whole tile, dict of
    stages, dict of checks).
    """
    import time
    n, m = tile.n, tile.m
    stages = {}
    tab = _f(tile.Alookup)
    ex = np.ascontiguousarray(tile.exp_i, dtype=np.int32)
    x, y = np.ascontiguousarray(tile.x_i), np.ascontiguousarray(tile.y_i)
    chk = {}
    if full_eig:
        t0 = time.perf_counter()
        Ae = lookup_A(tile)
        stages["lookup_A"] = time.perf_counter() - t0
    else:
        rows_A = min(rows_A, n)
        A = np.zeros((n, n), order="F")
        t0 = time.perf_counter()
        lib().or_lookup_A_rows(ctypes.c_int64(n), ctypes.c_int64(0), ctypes.c_int64(rows_A), _p(x),
                               _p(y), ex.ctypes.data_as(_ip), ctypes.c_int(tile.psfconst),
                               ctypes.c_int(tab.shape[0]), ctypes.c_int(tab.shape[1]), _p(tab),
                               ctypes.c_int(tile.npad), ctypes.c_int(tile.npoly),
                               ctypes.c_double(tile.psfxscale), ctypes.c_double(tile.psfyscale), _p(A))
        dt = time.perf_counter() - t0
        done = rows_A * n - rows_A * (rows_A - 1) / 2
        stages["lookup_A"] = dt * (n * (n + 1) / 2) / done
        if checks and 2 * rows_A <= n:
            r2 = 2 * rows_A
            t0 = time.perf_counter()
            lib().or_lookup_A_rows(ctypes.c_int64(n), ctypes.c_int64(0), ctypes.c_int64(r2), _p(x),
                                   _p(y), ex.ctypes.data_as(_ip), ctypes.c_int(tile.psfconst),
                                   ctypes.c_int(tab.shape[0]), ctypes.c_int(tab.shape[1]), _p(tab),
                                   ctypes.c_int(tile.npad), ctypes.c_int(tile.npoly),
                                   ctypes.c_double(tile.psfxscale), ctypes.c_double(tile.psfyscale),
                                   _p(A))
            dt2 = time.perf_counter() - t0
            done2 = r2 * n - r2 * (r2 - 1) / 2
            chk["lookup_A_time_ratio_2x_entries"] = round((dt2 / dt) * (done / done2) * 2, 3)
            stages["lookup_A"] = dt2 * (n * (n + 1) / 2) / done2
        # leading block of A for the eigen sample (computed, not timed)
        ne = min(n_eig, n)
        sub = tile.__class__(**{**tile.__dict__})
        sub.n = ne
        sub.x_i, sub.y_i, sub.exp_i = tile.x_i[:ne], tile.y_i[:ne], tile.exp_i[:ne]
        Ae = lookup_A(sub)
        del A
    rb = min(rows_B, n)
    subB = tile.__class__(**{**tile.__dict__})
    subB.n = rb
    subB.x_i, subB.y_i, subB.exp_i = tile.x_i[:rb], tile.y_i[:rb], tile.exp_i[:rb]
    t0 = time.perf_counter()
    lookup_B(subB)
    stages["lookup_B"] = (time.perf_counter() - t0) * n / rb
    if full_eig:
        t0 = time.perf_counter()
        eigen_A(Ae)
        stages["eigen_A"] = time.perf_counter() - t0
        del Ae
    else:
        t0 = time.perf_counter()
        eigen_A(Ae)
        te = time.perf_counter() - t0
        stages["eigen_A"] = te * (n / ne) ** 3
        if checks and 2 * ne <= n:
            sub2 = tile.__class__(**{**tile.__dict__})
            sub2.n = 2 * ne
            sub2.x_i, sub2.y_i, sub2.exp_i = tile.x_i[:2 * ne], tile.y_i[:2 * ne], tile.exp_i[:2 * ne]
            Ae2 = lookup_A(sub2)
            t0 = time.perf_counter()
            eigen_A(Ae2)
            te2 = time.perf_counter() - t0
            del Ae2
            chk["eigen_A_fitted_exponent"] = round(float(np.log2(te2 / te)), 3)
            chk["eigen_A_sizes"] = [ne, 2 * ne]
            stages["eigen_A"] = te2 * (n / (2 * ne)) ** 3
    rng = np.random.default_rng(0)
    Qr = np.asfortranarray(rng.standard_normal((n, n)))
    cp = min(cols_P, m)
    Bs = np.asfortranarray(rng.standard_normal((n, cp)))
    t0 = time.perf_counter()
    build_P(Qr, Bs)
    stages["build_P"] = (time.perf_counter() - t0) * m / cp
    if L is None or P is None:
        L = np.sort(rng.uniform(1e-17, 0.2, n))[::-1].copy()
        P = rng.standard_normal((n, m)) * 1e-3
    pb = min(pix_bisect, m)
    P2s = np.asfortranarray(np.asarray(P)[:, :pb] ** 2)
    t0 = time.perf_counter()
    K, _, _ = bisect_kappa(L, P2s, tile.C_a, tile.US_max, tile.kappa_min, tile.kappa_max,
                           tile.US_tol, tile.maxNbis, tile.USB, tile.Ndiag)
    stages["bisect"] = (time.perf_counter() - t0) * m / pb
    pt = min(pix_T, m)
    t0 = time.perf_counter()
    T = build_T(Qr, np.asfortranarray(np.asarray(P)[:, :pt]), L, K[:pt] if pt <= pb else np.full(pt, K[0]))
    build_S(T, tile.Ndiag)
    build_image(T, tile.I_i)
    tt = time.perf_counter() - t0
    stages["build_T"] = tt * m / pt
    if checks and 2 * pt <= min(m, pb):
        t0 = time.perf_counter()
        T = build_T(Qr, np.asfortranarray(np.asarray(P)[:, :2 * pt]), L, K[:2 * pt])
        build_S(T, tile.Ndiag)
        build_image(T, tile.I_i)
        tt2 = time.perf_counter() - t0
        chk["build_T_time_ratio_2x_pixels"] = round(tt2 / tt, 3)
        stages["build_T"] = tt2 * m / (2 * pt)
    return float(sum(stages.values())), stages, chk


def solve(tile, Q=None, L=None):
    """The whole reference solve path on one tile (imcom.f90:151-169).
    Returns a dict of every stage's output."""
    A = lookup_A(tile)
    if Q is None:
        Q, L, cond = eigen_A(A)
    else:
        cond = float(L.max() / L.min())
    B = lookup_B(tile)
    P, P2 = build_P(Q, B)
    K, US, ev = bisect_kappa(L, P2, tile.C_a, tile.US_max, tile.kappa_min, tile.kappa_max,
                             tile.US_tol, tile.maxNbis, tile.USB, tile.Ndiag)
    T = build_T(Q, P, L, K)
    if tile.USB:
        U, S = US, build_S(T, tile.Ndiag)
    else:
        S, U = US, build_U(L, P2, tile.C_a, K)
    H = build_image(T, tile.I_i)
    return dict(A=A, B=B, Q=Q, L=L, cond=cond, P=P, P2=P2, K=K, U=U, S=S, T=T, H=H, evals=ev)
