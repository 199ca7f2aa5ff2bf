/*
 * CPU oracle for the IMCOM solve path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference Fortran routines on the solve path,
 * used by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg as
 * the reference the GPU results are compared with.  It is never linked into,
 * or called by, the GPU library.
 *
 * Every array is Fortran column-major with leading dimension n (the
 * imcom_data globals, imcom_data.f90:95-132); exposure ids are 1-based.
 *
 * Restated routines (reference file:line):
 *   or_polint            imcom_proc.f90:351-394   (Neville tableau, NR polint)
 *   or_polin2            imcom_proc.f90:398-429
 *   or_interp_lookup     imcom_matrix.f90:497-542
 *   or_lookup_A          imcom_matrix.f90:206-284
 *   or_lookup_B          imcom_matrix.f90:288-357
 *   or_U_from_kappa      imcom_matrix.f90:904-922  (DDOT restated as an ordered sum)
 *   or_S_from_kappa      imcom_matrix.f90:955-970
 *   or_bisect_by_U       imcom_bisect.f90:174-222
 *   or_bisect_by_S       imcom_bisect.f90:226-267
 *   or_bisect_kappa      imcom_bisect.f90:78-170
 *   or_build_T           imcom_matrix.f90:716-750  (m DGEMVs, then T = -T/2)
 *   or_build_image       imcom_matrix.f90:754-790
 *   or_build_U           imcom_matrix.f90:795-835
 *   or_build_S           imcom_matrix.f90:839-877
 * The eigendecomposition (LAPACK DSYEVR) and P = Q^T B (DGEMM) are taken
 * from the same LAPACK/BLAS routines through scipy/numpy in oracle.py.
 *
 * Pinning: see the header of oracle.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#define OR_EPS 1.1102230246251565e-16 /* DLAMCH('E') = 2^-53 */

/* imcom_proc.f90:351-394.  xa, ya of length n (<= 16). */
static void or_polint(const double *xa, const double *ya, int n, double x, double eps,
                      double *y, double *dy) {
  double c[16], d[16], ho[16], den[16];
  int i, m, ns = 0;
  double best = fabs(x - xa[0]);
  for (i = 0; i < n; ++i) {
    c[i] = ya[i];
    d[i] = ya[i];
    ho[i] = xa[i] - x;
  }
  /* imcom_iminloc: first index of the minimum |x - xa| (1-based ns) */
  for (i = 1; i < n; ++i) {
    double v = fabs(x - xa[i]);
    if (v < best) {
      best = v;
      ns = i;
    }
  }
  *y = ya[ns];
  /* Fortran: ns (1-based) = ns0+1, then ns = ns - 1  ->  nsF = ns0 */
  int nsF = ns;
  *dy = 0.0;
  for (m = 1; m <= n - 1; ++m) {
    for (i = 0; i < n - m; ++i) den[i] = ho[i] - ho[i + m];
    for (i = 0; i < n - m; ++i)
      if (fabs(den[i]) <= eps) abort(); /* identical nodes: reference stops */
    for (i = 0; i < n - m; ++i) den[i] = (c[i + 1] - d[i]) / den[i];
    for (i = 0; i < n - m; ++i) {
      d[i] = ho[i + m] * den[i];
      c[i] = ho[i] * den[i];
    }
    if (2 * nsF < n - m) {
      *dy = c[nsF]; /* Fortran c(ns+1) */
    } else {
      *dy = d[nsF - 1]; /* Fortran d(ns) */
      nsF = nsF - 1;
    }
    *y = *y + *dy;
  }
}

/* imcom_proc.f90:398-429.  ya is (m1 x m2) column-major with leading dim ld. */
static double or_polin2(const double *x1a, int m1, const double *x2a, int m2,
                        const double *ya, int64_t ld, double x1, double x2) {
  double ymtmp[16], yntmp[16], y, dy;
  int j, k;
  for (j = 0; j < m1; ++j) {
    for (k = 0; k < m2; ++k) yntmp[k] = ya[j + (int64_t)k * ld];
    or_polint(x2a, yntmp, m2, x2, OR_EPS, &ymtmp[j], &dy);
  }
  or_polint(x1a, ymtmp, m1, x1, OR_EPS, &y, &dy);
  return y;
}

/* imcom_matrix.f90:497-542 */
double or_interp_lookup(int n1, int n2, const double *lookup, double xarg, double yarg,
                        int npoly) {
  int npm, ia_floor, ja_floor, ia_ceil, ja_ceil, imin, imax, jmin, jmax, ii, jj;
  double icent, jcent, iarg, jarg, x1a[16], x2a[16];
  if (npoly % 2 == 0) npoly -= 1;
  npm = npoly / 2;
  icent = (double)(n1 / 2 + 1);
  jcent = (double)(n2 / 2 + 1);
  iarg = icent + xarg;
  jarg = jcent + yarg;
  ia_floor = (int)floor(iarg);
  ja_floor = (int)floor(jarg);
  ia_ceil = (int)ceil(iarg);
  ja_ceil = (int)ceil(jarg);
  if (ia_floor >= 1 && ja_floor >= 1 && ia_ceil <= n1 && ja_ceil <= n2) {
    imin = ia_floor - npm > 1 ? ia_floor - npm : 1;
    jmin = ja_floor - npm > 1 ? ja_floor - npm : 1;
    imax = ia_ceil + npm < n1 ? ia_ceil + npm : n1;
    jmax = ja_ceil + npm < n2 ? ja_ceil + npm : n2;
    for (ii = imin; ii <= imax; ++ii) x1a[ii - imin] = (double)ii;
    for (jj = jmin; jj <= jmax; ++jj) x2a[jj - jmin] = (double)jj;
    return or_polin2(x1a, imax - imin + 1, x2a, jmax - jmin + 1,
                     lookup + (imin - 1) + (int64_t)(jmin - 1) * n1, n1, iarg, jarg);
  }
  return 0.0;
}

/* imcom_rotate_image (imcom_proc.f90:184-250): rotation about
 * ((n1+1)/2, (n2+1)/2) by Neville interpolation on the clipped stencil;
 * cost, sint as the caller computed them from mod(theta, 360) * pi / 180. */
void or_rotate_image(int n1, int n2, const double *im_in, double cost, double sint, int npoly,
                     int conserve_flux, double *im_out) {
  int npm, iout, jout, is_f, js_f, is_c, js_c, imin, jmin, imax, jmax, ii, jj;
  double x_cent, y_cent, xout, yout, xs, ys, x1a[16], x2a[16], s_in, s_out;
  int64_t k;
  if (npoly % 2 == 0) npoly -= 1;
  npm = npoly / 2;
  x_cent = (double)(n1 + 1) / 2.0;
  y_cent = (double)(n2 + 1) / 2.0;
  for (iout = 1; iout <= n1; ++iout) {
    xout = (double)iout - x_cent;
    for (jout = 1; jout <= n2; ++jout) {
      double *o = im_out + (iout - 1) + (int64_t)(jout - 1) * n1;
      yout = (double)jout - y_cent;
      xs = xout * cost + yout * sint + x_cent;
      ys = -xout * sint + yout * cost + y_cent;
      is_f = (int)floor(xs);
      js_f = (int)floor(ys);
      is_c = (int)ceil(xs);
      js_c = (int)ceil(ys);
      if (is_f >= 1 && js_f >= 1 && is_c <= n1 && js_c <= n2) {
        imin = is_f - npm > 1 ? is_f - npm : 1;
        jmin = js_f - npm > 1 ? js_f - npm : 1;
        imax = is_c + npm < n1 ? is_c + npm : n1;
        jmax = js_c + npm < n2 ? js_c + npm : n2;
        for (ii = imin; ii <= imax; ++ii) x1a[ii - imin] = (double)ii;
        for (jj = jmin; jj <= jmax; ++jj) x2a[jj - jmin] = (double)jj;
        *o = or_polin2(x1a, imax - imin + 1, x2a, jmax - jmin + 1,
                       im_in + (imin - 1) + (int64_t)(jmin - 1) * n1, n1, xs, ys);
      } else {
        *o = 0.0;
      }
    }
  }
  if (conserve_flux != 0) {
    s_in = 0.0;
    s_out = 0.0;
    for (k = 0; k < (int64_t)n1 * n2; ++k) {
      s_in += im_in[k];
      s_out += im_out[k];
    }
    for (k = 0; k < (int64_t)n1 * n2; ++k) im_out[k] = s_in * im_out[k] / s_out;
  }
}

/* imcom_matrix.f90:206-284.  Alookup is (n1al, n2al, nA); A is n x n. */
void or_lookup_A(int64_t n, const double *x_i, const double *y_i, const int32_t *exp_i,
                 int psfconst, int n1al, int n2al, const double *Alookup, int npad, int npoly,
                 double psfxscale, double psfyscale, double *A) {
  int64_t tsz = (int64_t)n1al * n2al;
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = 0; i < n; ++i) {
    for (int64_t j = i; j < n; ++j) {
      int ial = psfconst == 1 ? 1 : exp_i[i] + exp_i[j] * (exp_i[j] - 1) / 2;
      double Delx = (double)npad * (x_i[j] - x_i[i]) / psfxscale;
      double Dely = (double)npad * (y_i[j] - y_i[i]) / psfyscale;
      A[i + j * n] = or_interp_lookup(n1al, n2al, Alookup + (ial - 1) * tsz, Delx, Dely, npoly);
    }
  }
  for (int64_t i = 1; i < n; ++i)
    for (int64_t j = 0; j < i; ++j) A[i + j * n] = A[j + i * n];
}

/* Rows [i0, i1) of imcom_lookup_A's upper triangle (j >= i) into A (n x n):
 * a bounded sample of the A build, for CPU timing. */
void or_lookup_A_rows(int64_t n, int64_t i0, int64_t i1, const double *x_i, const double *y_i,
                      const int32_t *exp_i, int psfconst, int n1al, int n2al,
                      const double *Alookup, int npad, int npoly, double psfxscale,
                      double psfyscale, double *A) {
  int64_t tsz = (int64_t)n1al * n2al;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t i = i0; i < i1; ++i) {
    for (int64_t j = i; j < n; ++j) {
      int ial = psfconst == 1 ? 1 : exp_i[i] + exp_i[j] * (exp_i[j] - 1) / 2;
      double Delx = (double)npad * (x_i[j] - x_i[i]) / psfxscale;
      double Dely = (double)npad * (y_i[j] - y_i[i]) / psfyscale;
      A[i + j * n] = or_interp_lookup(n1al, n2al, Alookup + (ial - 1) * tsz, Delx, Dely, npoly);
    }
  }
}

/* imcom_matrix.f90:288-357.  B is n x m. */
void or_lookup_B(int64_t n, int64_t m, const double *x_i, const double *y_i,
                 const int32_t *exp_i, const double *X_a, const double *Y_a, int psfconst,
                 int n1bl, int n2bl, const double *Blookup, int npad, int npoly,
                 double psfxscale, double psfyscale, double *B) {
  int64_t tsz = (int64_t)n1bl * n2bl;
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = 0; i < n; ++i) {
    const double *tab = Blookup + (psfconst == 1 ? 0 : (int64_t)(exp_i[i] - 1) * tsz);
    for (int64_t a = 0; a < m; ++a) {
      double Delx = (double)npad * (x_i[i] - X_a[a]) / psfxscale;
      double Dely = (double)npad * (y_i[i] - Y_a[a]) / psfyscale;
      B[i + a * n] = or_interp_lookup(n1bl, n2bl, tab, Delx, Dely, npoly);
    }
  }
}

/* imcom_matrix.f90:904-922 */
double or_U_from_kappa(int64_t n, const double *L, const double *P2col, double C_a, double k) {
  double Usum = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    double denom = (L[i] + k) * (L[i] + k);
    Usum += (L[i] + 2.0 * k) * P2col[i] / denom;
  }
  return (C_a - Usum / 4.0) / C_a;
}

/* imcom_matrix.f90:955-970 (Ndiag indexed by eigen index, as the reference) */
double or_S_from_kappa(int64_t n, const double *L, const double *P2col, const double *Ndiag,
                       double k) {
  double Ssum = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    double denom = (L[i] + k) * (L[i] + k);
    Ssum += Ndiag[i] * P2col[i] / denom;
  }
  return Ssum / 4.0;
}

/* imcom_bisect.f90:174-222.  Returns the number of U evaluations. */
int or_bisect_by_U(int64_t n, const double *L, const double *P2col, double C_a, double k_min,
                   double k_max, double U_max, double tol, int maxnits, double *k_root,
                   double *U_out) {
  double lo, hi, root, dx, mid, fmid, f;
  int i, evals = 1;
  *U_out = or_U_from_kappa(n, L, P2col, C_a, k_min);
  f = *U_out - U_max;
  if (f < 0.0) {
    lo = log10(k_min);
    hi = log10(k_max);
  } else {
    *k_root = k_min;
    return evals;
  }
  root = lo;
  dx = hi - lo;
  for (i = 1; i <= maxnits; ++i) {
    dx = 0.5 * dx;
    mid = root + dx;
    fmid = or_U_from_kappa(n, L, P2col, C_a, pow(10.0, mid)) - U_max;
    ++evals;
    if (fmid < 0.0) {
      root = mid;
      *U_out = fmid + U_max;
      if (fabs(fmid) < tol) break;
    }
    if (fabs(fmid) < C_a * OR_EPS) break;
  }
  *k_root = pow(10.0, root);
  return evals;
}

/* imcom_bisect.f90:226-267 */
int or_bisect_by_S(int64_t n, const double *L, const double *P2col, const double *Ndiag,
                   double C_a, double k_min, double k_max, double S_max, double tol,
                   int maxnits, double *k_root, double *S_out) {
  double lo, hi, root, dx, mid, fmid, f;
  int i, evals = 1;
  *S_out = or_S_from_kappa(n, L, P2col, Ndiag, k_max);
  f = *S_out - S_max;
  if (f < 0.0) {
    lo = log10(k_max);
    hi = log10(k_min);
  } else {
    *k_root = k_max;
    return evals;
  }
  root = lo;
  dx = hi - lo;
  for (i = 1; i <= maxnits; ++i) {
    dx = 0.5 * dx;
    mid = root + dx;
    fmid = or_S_from_kappa(n, L, P2col, Ndiag, pow(10.0, mid)) - S_max;
    ++evals;
    if (fmid < 0.0) {
      root = mid;
      *S_out = fmid + S_max;
      if (fabs(fmid) < tol) break;
    }
    if (fabs(fmid) < C_a * OR_EPS) break;
  }
  *k_root = pow(10.0, root);
  return evals;
}

/* imcom_bisect.f90:78-170 (maps and FITS writes are the caller's). */
void or_bisect_kappa(int UnotS, int64_t n, int64_t m, const double *L, const double *P2,
                     const double *Ndiag, double C_a, double US, double k_min, double k_max,
                     double tol, int maxnits, double *K_a, double *US_a, int32_t *evals) {
#pragma omp parallel for schedule(dynamic, 8)
  for (int64_t a = 0; a < m; ++a) {
    int e;
    if (UnotS)
      e = or_bisect_by_U(n, L, P2 + a * n, C_a, k_min, k_max, US, tol, maxnits, &K_a[a],
                         &US_a[a]);
    else
      e = or_bisect_by_S(n, L, P2 + a * n, Ndiag, C_a, k_min, k_max, US, tol, maxnits,
                         &K_a[a], &US_a[a]);
    if (evals) evals[a] = e;
  }
}

/* imcom_matrix.f90:716-750: T(:,a) = Q (P(:,a) / (L + K_a(a))), T = -T/2. */
void or_build_T(int64_t n, int64_t m, const double *Q, const double *P, const double *L,
                const double *K_a, double *T) {
  /* one pixel per chunk: a sample of as many pixels as threads keeps every
     thread busy, as the whole tile does */
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t a = 0; a < m; ++a) {
    double *t = T + a * n;
    const double *p = P + a * n;
    for (int64_t i = 0; i < n; ++i) t[i] = 0.0;
    for (int64_t k = 0; k < n; ++k) {
      double s = p[k] / (L[k] + K_a[a]);
      const double *q = Q + k * n;
      for (int64_t i = 0; i < n; ++i) t[i] += q[i] * s;
    }
    for (int64_t i = 0; i < n; ++i) t[i] = -t[i] / 2.0;
  }
}

/* imcom_matrix.f90:754-790: H = T^T I */
void or_build_image(int64_t n, int64_t m, const double *T, const double *I_i, double *H) {
#pragma omp parallel for
  for (int64_t a = 0; a < m; ++a) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += T[i + a * n] * I_i[i];
    H[a] = s;
  }
}

/* imcom_matrix.f90:795-835 */
void or_build_U(int64_t n, int64_t m, const double *L, const double *P2, double C_a,
                const double *K_a, double *U_a) {
#pragma omp parallel for
  for (int64_t a = 0; a < m; ++a) U_a[a] = or_U_from_kappa(n, L, P2 + a * n, C_a, K_a[a]);
}

/* imcom_matrix.f90:839-877: S_a = sum_i T_ia^2 Ndiag(i) (pixel basis) */
void or_build_S(int64_t n, int64_t m, const double *T, const double *Ndiag, double *S_a) {
#pragma omp parallel for
  for (int64_t a = 0; a < m; ++a) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += T[i + a * n] * (T[i + a * n] * Ndiag[i]);
    S_a[a] = s;
  }
}
