// Kernel (b), part 1: blocked Householder tridiagonalisation of the symmetric
// system matrix A (the SYTRD stage inside DSYEVR, imcom_matrix.f90:617-643),
// and the Householder back-transform of the tridiagonal eigenvectors (the
// ORMTR stage).
//
// Lower-storage one-stage reduction: A = Q T Q^T with Q = H(0) H(1) ...
// H(n-2), H(c) = I - tau_c v_c v_c^T, v_c zero above row c+1, v_c(c+1) = 1.
// Panels of NB columns.  Inside a panel, column c is brought up to date with
// the panel's earlier reflectors (rank-2jj update), its reflector is formed,
// and w_c = tau (A22 v - V (W^T v) - W (V^T v)) - 1/2 tau^2 (w^T v) v is
// accumulated; after the panel the trailing matrix takes one rank-2NB update
// A22 -= V W^T + W V^T on the FP64 matrix cores.  The full symmetric trailing
// matrix is kept, so the matrix-vector product reads whole columns
// (coalesced, one wavefront per column).
//
// Three launches per column:
//   sy_symv      reflector from the column's norm partials, y = A22 v and
//                the 2jj dot products V^T v, W^T v (one wavefront per column)
//   sy_wpre      w_pre = tau (y - V sW - W sV) and partials of w_pre^T v
//   sy_wfin      w = w_pre + alpha v; rank-2(jj+1) update of the next column
//                and its norm partials
// All sums are fixed-order (partials per workgroup, reduced in workgroup
// order), so runs are bitwise reproducible.
//
// Tiles of a batch advance column by column in lockstep; blockIdx.z selects
// the tile and a tile whose columns are exhausted exits.
#include "common.h"
#include "kernels.h"

#include <algorithm>
#include <vector>

namespace imcom {

constexpr int RB = 256;  // rows per workgroup in the row-parallel kernels

__device__ __forceinline__ double dlapy2(double x, double y) {
  const double xa = fabs(x), ya = fabs(y);
  const double w = fmax(xa, ya), z = fmin(xa, ya);
  if (z == 0.0) return w;
  const double q = z / w;
  return w * sqrt(1.0 + q * q);
}

// Column c update for rows [c, n) with the panel's first `jj` reflectors,
// plus per-workgroup partials of the squared norm of rows [c+2, n).
// V aliases A's panel columns (V(i,p) = A(i, k+p)); W is n x NB.
__global__ __launch_bounds__(RB) void sy_colupd(const SyTile* __restrict__ tiles, int k, int c,
                                                int jj) {
  __shared__ double red[RB / 64];
  const SyTile t = tiles[blockIdx.z];
  const int64_t n = t.n;
  if (c >= n - 1) return;
  const int64_t i = c + (int64_t)blockIdx.x * RB + threadIdx.x;
  double a = 0.0;
  if (i < n) {
    a = t.A[i + (int64_t)c * n];
    if (jj > 0) {
      double s = 0.0;
      for (int p = 0; p < jj; ++p) {
        const double* Vp = t.A + (int64_t)(k + p) * n;
        const double* Wp = t.W + (int64_t)p * n;
        s += Vp[i] * Wp[c] + Wp[i] * Vp[c];
      }
      a -= s;
      t.A[i + (int64_t)c * n] = a;
    }
  }
  const double sq = (i >= c + 2 && i < n) ? a * a : 0.0;
  const double tot = block_sum(sq, red);
  if (threadIdx.x == 0) t.part[blockIdx.x] = tot;
}

// Reflector for column c, y = A22 v, and V^T v, W^T v.
// Columns handled (one wavefront each): j in [c+1, n) of A22, then jj columns
// of V (A's panel columns), then jj columns of W.
__global__ __launch_bounds__(256) void sy_symv(const SyTile* __restrict__ tiles, int k, int c,
                                               int jj, int nparts) {
  const SyTile t = tiles[blockIdx.z];
  const int64_t n = t.n;
  if (c >= n - 1) return;
  double* colc = t.A + (int64_t)c * n;
  // reduce the norm partials (every workgroup, same order)
  double xn2 = 0.0;
  for (int q = 0; q < nparts; ++q) xn2 += t.part[q];
  const double alpha = colc[c + 1];
  const double xnorm = sqrt(xn2);
  double tau, beta, scal;
  if (xnorm == 0.0) {
    tau = 0.0;
    beta = alpha;
    scal = 0.0;
  } else {
    beta = -copysign(dlapy2(alpha, xnorm), alpha);
    tau = (beta - alpha) / beta;
    scal = 1.0 / (alpha - beta);
  }
  const int lane = threadIdx.x & 63;
  const int64_t gw = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const int64_t m22 = n - c - 1;  // rows/cols of A22
  const int64_t ncol = m22 + 2 * jj;
  for (int64_t q = gw; q < ncol; q += nw) {
    const double* X;
    if (q < m22)
      X = t.A + (c + 1 + q) * n;
    else if (q < m22 + jj)
      X = t.A + (int64_t)(k + (q - m22)) * n;
    else
      X = t.W + (q - m22 - jj) * n;
    // dot over rows [c+1, n): v_{c+1} = 1, v_i = scal * colc[i]
    double s = 0.0;
    for (int64_t i = c + 2 + lane; i < n; i += 64) s += X[i] * colc[i];
    s = wave_sum(s);
    if (lane == 0) {
      const double dot = X[c + 1] + scal * s;
      if (q < m22)
        t.y[c + 1 + q] = dot;
      else
        t.s[q - m22] = dot;  // s[0:jj) = V^T v, s[jj:2jj) = W^T v
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    t.d[c] = colc[c];
    t.e[c] = beta;
    t.tau[c] = tau;
    t.scal[0] = scal;
  }
}

// Store v into A's column c (unit at c+1) after sy_symv has consumed it, and
// w_pre = tau (y - V sW - W sV) with partials of w_pre^T v.
__global__ __launch_bounds__(RB) void sy_wpre(const SyTile* __restrict__ tiles, int k, int c,
                                              int jj) {
  __shared__ double red[RB / 64];
  const SyTile t = tiles[blockIdx.z];
  const int64_t n = t.n;
  if (c >= n - 1) return;
  const int64_t i = c + 1 + (int64_t)blockIdx.x * RB + threadIdx.x;
  const double tau = t.tau[c];
  const double scal = t.scal[0];
  double contrib = 0.0;
  if (i < n) {
    double* colc = t.A + (int64_t)c * n;
    const double v = (i == c + 1) ? 1.0 : colc[i] * scal;
    double acc = t.y[i];
    for (int p = 0; p < jj; ++p) {
      const double Vip = t.A[i + (int64_t)(k + p) * n];
      const double Wip = t.W[i + (int64_t)p * n];
      acc -= Vip * t.s[jj + p] + Wip * t.s[p];
    }
    const double w = tau * acc;
    t.wpre[i] = w;
    t.vtmp[i] = v;
    contrib = w * v;
  }
  const double tot = block_sum(contrib, red);
  if (threadIdx.x == 0) t.part2[blockIdx.x] = tot;
}

// w = w_pre - 1/2 tau (w_pre^T v) v into W(:, jj); v into A(:, c); then (if
// requested) update column c+1 with reflectors 0..jj and its norm partials.
__global__ __launch_bounds__(RB) void sy_wfin(const SyTile* __restrict__ tiles, int k, int c,
                                              int jj, int nparts2, int in_panel_next) {
  __shared__ double red[RB / 64];
  const SyTile t = tiles[blockIdx.z];
  const int64_t n = t.n;
  if (c >= n - 1) return;
  // column c+1 is updated here when it is in this panel, or when it is the
  // tile's last column (no reflector of its own; its diagonal is d(n-1))
  const int upd_next = in_panel_next || (c + 1 == n - 1);
  double wv = 0.0;
  for (int q = 0; q < nparts2; ++q) wv += t.part2[q];
  const double tau = t.tau[c];
  const double alpha = -0.5 * tau * wv;
  const int64_t c1 = c + 1;
  const int64_t i = c1 + (int64_t)blockIdx.x * RB + threadIdx.x;
  double* W = t.W;
  double a = 0.0;
  if (i < n) {
    const double v = t.vtmp[i];
    const double w = t.wpre[i] + alpha * v;
    W[i + (int64_t)jj * n] = w;
    t.A[i + (int64_t)c * n] = v;
    if (upd_next) {
      // row c+1 of V and W for p <= jj (read from the untouched copies)
      a = t.A[i + c1 * n];
      double s = 0.0;
      for (int p = 0; p < jj; ++p) {
        const double* Vp = t.A + (int64_t)(k + p) * n;
        const double* Wp = W + (int64_t)p * n;
        s += Vp[i] * Wp[c1] + Wp[i] * Vp[c1];
      }
      // p = jj: V(c1, jj) = 1 and W(c1, jj) = wpre[c1] + alpha
      const double Wc1 = t.wpre[c1] + alpha;
      s += v * Wc1 + w * 1.0;
      a -= s;
      t.A[i + c1 * n] = a;
    }
  }
  if (upd_next) {
    const double sq = (i >= c1 + 2 && i < n) ? a * a : 0.0;
    const double tot = block_sum(sq, red);
    if (threadIdx.x == 0) t.part[blockIdx.x] = tot;
    if (c1 == n - 1 && blockIdx.x == 0 && threadIdx.x == 0) t.d[c1] = a;
  }
}

// Copy the panel's V (with explicit zeros above the unit) and W into the
// [V W] and [W V] operands of the trailing update, rows [r0, n).
__global__ void sy_pack(const SyTile* __restrict__ tiles, int k, int nb, int64_t r0) {
  const SyTile t = tiles[blockIdx.z];
  const int64_t n = t.n;
  const int64_t i = r0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (i >= n || k + p >= n - 1) return;
  const double v = (i > k + p) ? t.A[i + (int64_t)(k + p) * n] : 0.0;
  const double w = t.W[i + (int64_t)p * n];
  t.VW[i + (int64_t)p * n] = v;
  t.VW[i + (int64_t)(nb + p) * n] = w;
  t.WV[i + (int64_t)p * n] = w;
  t.WV[i + (int64_t)(nb + p) * n] = v;
}

__global__ void sy_last_diag(const SyTile* __restrict__ tiles) {
  const SyTile t = tiles[blockIdx.z];
  if (t.n == 1) t.d[0] = t.A[0];
}

// ---------------------------------------------------------------------------
// Back-transform helpers.  Reflectors c in [kb, kb+nbb) are gathered into
// Vb (rows [kb+1, n) -> local rows 0.., column p = c - kb; explicit unit and
// zeros), and T (nbb x nbb upper triangular, forward column-wise compact WY,
// as DLARFT) is built from tau and S = Vb^T Vb.
__global__ void bt_gather(const SyTile* __restrict__ tiles, int kb, int nbb) {
  const SyTile t = tiles[blockIdx.z];
  const int64_t n = t.n;
  const int p = blockIdx.y;
  const int64_t c = kb + p;
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // local row
  const int64_t rows = n - kb - 1;
  if (r >= rows) return;
  const int64_t i = kb + 1 + r;
  double v = 0.0;
  if (c < n - 1) v = (i == c + 1) ? 1.0 : (i > c + 1 ? t.A[i + c * n] : 0.0);
  t.Vb[r + (int64_t)p * rows] = v;
}

__global__ __launch_bounds__(256) void bt_build_T(const SyTile* __restrict__ tiles, int kb,
                                                  int nbb) {
  const SyTile t = tiles[blockIdx.z];
  const int64_t n = t.n;
  if (kb >= n - 1) return;
  double* T = t.Tb;  // nbb x nbb, ld nbb
  const double* S = t.Sb;
  const int i = threadIdx.x;
  for (int j = 0; j < nbb; ++j) {
    const int64_t c = kb + j;
    const double tj = (c < n - 1) ? t.tau[c] : 0.0;
    double acc = 0.0;
    if (i < j) {
      for (int l = i; l < j; ++l) acc += T[i + l * nbb] * S[l + j * nbb];
    }
    __syncthreads();
    if (i < nbb) {
      if (i < j)
        T[i + j * nbb] = -tj * acc;
      else if (i == j)
        T[i + j * nbb] = tj;
      else
        T[i + j * nbb] = 0.0;
    }
    __syncthreads();
  }
}

hipError_t sytrd_run(hipStream_t st, const SyTile* d_tiles, int ntiles, int64_t nmax,
                     const std::vector<SyTile>& host, int nb) {
  const int rbl = ceil_div(nmax, RB);
  for (int64_t k = 0; k < nmax - 1; k += nb) {
    const int pb = (int)std::min<int64_t>(nb, nmax - 1 - k);
    // first column of the panel: norm partials only (jj = 0)
    hipLaunchKernelGGL(sy_colupd, dim3(ceil_div(nmax - k, RB), 1, ntiles), dim3(RB), 0, st,
                       d_tiles, (int)k, (int)k, 0);
    for (int jj = 0; jj < pb; ++jj) {
      const int64_t c = k + jj;
      const int nparts = ceil_div(nmax - c, RB);
      const int64_t ncol = (nmax - c - 1) + 2 * jj;
      const int symv_blocks = (int)std::min<int64_t>(std::max<int64_t>(ceil_div(ncol, 4), 1), 2048);
      hipLaunchKernelGGL(sy_symv, dim3(symv_blocks, 1, ntiles), dim3(256), 0, st, d_tiles, (int)k,
                         (int)c, jj, nparts);
      const int nb2 = std::max(ceil_div(nmax - c - 1, RB), 1);
      hipLaunchKernelGGL(sy_wpre, dim3(nb2, 1, ntiles), dim3(RB), 0, st, d_tiles, (int)k, (int)c,
                         jj);
      hipLaunchKernelGGL(sy_wfin, dim3(nb2, 1, ntiles), dim3(RB), 0, st, d_tiles, (int)k, (int)c,
                         jj, nb2, (jj + 1 < pb) ? 1 : 0);
    }
    (void)rbl;
    // trailing update A22 -= [V W] [W V]^T over rows/cols [k+nb, n)
    const int64_t r0 = k + nb;
    if (r0 < nmax) {
      hipLaunchKernelGGL(sy_pack, dim3(ceil_div(nmax - r0, 256), nb, ntiles), dim3(256), 0, st,
                         d_tiles, (int)k, nb, r0);
      for (int b = 0; b < ntiles; ++b) {
        const SyTile& h = host[b];
        if (r0 >= h.n) continue;
        const int64_t mm = h.n - r0;
        IMCOM_HIP_TRY(dgemm(st, false, true, (int)mm, (int)mm, 2 * nb, -1.0, h.VW + r0, h.n,
                            h.WV + r0, h.n, 1.0, h.A + r0 + r0 * h.n, h.n));
      }
    }
  }
  hipLaunchKernelGGL(sy_last_diag, dim3(1, 1, ntiles), dim3(1), 0, st, d_tiles);
  return hipGetLastError();
}

// Q = H(0) ... H(n-2) Z, applied in blocks of NBB reflectors, last block
// first: Z[kb+1:, :] -= Vb (T (Vb^T Z[kb+1:, :])).  Z is n x n (ld n).
hipError_t ormtr_run(hipStream_t st, const SyTile* d_tiles, int ntiles, int64_t nmax,
                     const std::vector<SyTile>& host, int nbb) {
  if (nmax <= 1) return hipSuccess;
  const int64_t nref = nmax - 1;
  const int64_t nblk = (nref + nbb - 1) / nbb;
  for (int64_t bi = nblk - 1; bi >= 0; --bi) {
    const int kb = (int)(bi * nbb);
    hipLaunchKernelGGL(bt_gather, dim3(ceil_div(nmax - kb - 1, 256), nbb, ntiles), dim3(256), 0,
                       st, d_tiles, kb, nbb);
    for (int b = 0; b < ntiles; ++b) {
      const SyTile& h = host[b];
      if (kb >= h.n - 1) continue;
      const int64_t rows = h.n - kb - 1;
      // S = Vb^T Vb
      IMCOM_HIP_TRY(dgemm(st, true, false, nbb, nbb, (int)rows, 1.0, h.Vb, rows, h.Vb, rows, 0.0,
                          h.Sb, nbb));
    }
    hipLaunchKernelGGL(bt_build_T, dim3(1, 1, ntiles), dim3(256), 0, st, d_tiles, kb, nbb);
    for (int b = 0; b < ntiles; ++b) {
      const SyTile& h = host[b];
      if (kb >= h.n - 1) continue;
      const int64_t rows = h.n - kb - 1;
      double* Zr = h.Z + kb + 1;
      // Y = Vb^T Z_r  (nbb x n)
      IMCOM_HIP_TRY(dgemm(st, true, false, nbb, (int)h.n, (int)rows, 1.0, h.Vb, rows, Zr, h.n, 0.0,
                          h.Y, nbb));
      // Y2 = T Y
      IMCOM_HIP_TRY(dgemm(st, false, false, nbb, (int)h.n, nbb, 1.0, h.Tb, nbb, h.Y, nbb, 0.0,
                          h.Y2, nbb));
      // Z_r -= Vb Y2
      IMCOM_HIP_TRY(dgemm(st, false, false, (int)rows, (int)h.n, nbb, -1.0, h.Vb, rows, h.Y2, nbb,
                          1.0, Zr, h.n));
    }
  }
  return hipGetLastError();
}

}  // namespace imcom
