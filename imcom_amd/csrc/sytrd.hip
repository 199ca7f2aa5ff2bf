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
// matrix is kept in memory; the matrix-vector product reads only its lower
// triangle, in 64x64 blocks, each block serving its rows and its columns.
//
// Two launches per column:
//   sy_symv      reflector from the column's norm partials, partial sums of
//                y = A22 v and of v^T A22 v over the lower triangle, the 2jj
//                dot products V^T v, W^T v
//   sy_wupd      w = tau (y - V sW - W sV) + alpha v with
//                alpha = -1/2 tau^2 (v^T A22 v - 2 sV.sW) (the same value as
//                -1/2 tau w_pre^T v, without a second grid-wide reduction);
//                rank-2(jj+1) update of the next column and its norm partials
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

// Householder reflector of column c from its norm partials (LAPACK dlarfg
// conventions): beta = -sign(alpha) |x|, tau = (beta - alpha)/beta, and
// v = [1, x(2:)/(alpha - beta)].
__device__ __forceinline__ void sy_reflector(const SyTile& t, int64_t c, int nparts, double& tau,
                                             double& beta, double& scal) {
  double xn2 = 0.0;
  for (int q = 0; q < nparts; ++q) xn2 += t.part[q];
  const double alpha = t.A[(c + 1) + c * t.n];
  const double xnorm = sqrt(xn2);
  if (xnorm == 0.0) {
    tau = 0.0;
    beta = alpha;
    scal = 0.0;
  } else {
    beta = -copysign(dlapy2(alpha, xnorm), alpha);
    tau = (beta - alpha) / beta;
    scal = 1.0 / (alpha - beta);
  }
}

constexpr int SB = 64;   // symmetric matrix-vector block
constexpr int SEG = 8;   // blocks of one block-column per workgroup

// y = A22 v reading only the lower triangle of A22 (order s = n-c-1, the
// trailing matrix below/right of column c).  Workgroup (seg, J) walks blocks
// I in [J + seg*SEG, J + (seg+1)*SEG) of block-column J: block (I,J) gives the
// rows of block I the product A_IJ v_J (written to ypart[J][rows]) and, for
// I > J, gives the columns of block J the product A_IJ^T v_I (accumulated over
// the segment and written to ycpart[seg][cols]).  sy_wupd sums the partials
// in a fixed order.  Workgroups in row y = nbr compute the 2jj dot products
// V^T v and W^T v, one wavefront each.
__global__ __launch_bounds__(256) void sy_symv(const SyTile* __restrict__ tiles, int k, int c,
                                               int jj, int nparts) {
  __shared__ double vJ[SB];
  __shared__ double red[4][SB + 1];
  const SyTile t = tiles[blockIdx.z];
  const int64_t n = t.n;
  if (c >= n - 1) return;
  const double* __restrict__ colc = t.A + (int64_t)c * n;
  double tau, beta, scal;
  sy_reflector(t, c, nparts, tau, beta, scal);
  const int64_t s = n - c - 1;
  const int nbr = (int)((s + SB - 1) / SB);
  const int J = blockIdx.y;
  const int tid = threadIdx.x, tx = tid & 63, ty = tid >> 6;
  const int64_t qslot = (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
  if (J >= nbr) {
    if (tid == 0) t.qpart[qslot] = 0.0;
    const int q = blockIdx.x * 4 + ty;
    if (J == nbr && q < 2 * jj) {
      const double* X = q < jj ? t.A + (int64_t)(k + q) * n : t.W + (int64_t)(q - jj) * n;
      double acc = 0.0;
      for (int64_t i = c + 2 + tx; i < n; i += 64) acc += X[i] * colc[i];
      acc = wave_sum(acc);
      if (tx == 0) t.s[q] = X[c + 1] + scal * acc;  // s[0:jj) = V^T v, s[jj:2jj) = W^T v
    }
    return;
  }
  const int I0 = J + blockIdx.x * SEG;
  if (I0 >= nbr) {
    if (tid == 0) t.qpart[qslot] = 0.0;
    return;
  }
  const int I1 = min(I0 + SEG, nbr);
  // v in A22-local indexing: v(0) = 1, v(r) = scal * A(c+1+r, c)
  if (tid < SB) {
    const int64_t r = (int64_t)J * SB + tid;
    vJ[tid] = (r < s) ? (r == 0 ? 1.0 : scal * colc[c + 1 + r]) : 0.0;
  }
  __syncthreads();
  const double* __restrict__ A22 = t.A + (c + 1) + (int64_t)(c + 1) * n;
  double colacc[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) colacc[q] = 0.0;
  double qacc = 0.0;  // this workgroup's share of v^T A22 v
  for (int I = I0; I < I1; ++I) {
    const int64_t r = (int64_t)I * SB + tx;
    const bool rv = r < s;
    double a[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int64_t col = (int64_t)J * SB + ty + 4 * q;
      a[q] = (rv && col < s) ? A22[r + col * n] : 0.0;
    }
    double rowp = 0.0;
#pragma unroll
    for (int q = 0; q < 16; ++q) rowp += a[q] * vJ[ty + 4 * q];
    red[ty][tx] = rowp;
    const double vI = rv ? (r == 0 ? 1.0 : scal * colc[c + 1 + r]) : 0.0;
    if (I > J) {
#pragma unroll
      for (int q = 0; q < 16; ++q) colacc[q] += a[q] * vI;
    }
    __syncthreads();
    if (ty == 0 && rv) {
      const double rs = red[0][tx] + red[1][tx] + red[2][tx] + red[3][tx];
      t.ypart[(int64_t)J * n + r] = rs;
      // v_I^T A_IJ v_J, counted twice off the diagonal (the mirrored block)
      qacc += (I == J ? 1.0 : 2.0) * vI * rs;
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const double cs = wave_sum(colacc[q]);
    const int64_t col = (int64_t)J * SB + ty + 4 * q;
    if (tx == 0 && col < s) t.ycpart[(int64_t)blockIdx.x * n + col] = cs;
  }
  {
    __shared__ double qred[4];
    const double qt = block_sum(qacc, qred);
    if (tid == 0) t.qpart[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = qt;
  }
  if (blockIdx.x == 0 && J == 0 && tid == 0) {
    t.d[c] = colc[c];
    t.e[c] = beta;
    t.tau[c] = tau;
    t.scal[0] = scal;
  }
}

// w before its v-correction at row i of the trailing matrix:
//   tau (y_i - sum_p V(i,p) sW_p - W(i,p) sV_p),
// y_i summed from sy_symv's partials in a fixed order.
__device__ __forceinline__ double sy_wrow(const SyTile& t, int k, int64_t c, int jj, double tau,
                                          int64_t i) {
  const int64_t n = t.n;
  const int64_t s = n - c - 1;
  const int nbr = (int)((s + SB - 1) / SB);
  const int64_t r = i - (c + 1);
  const int R = (int)(r / SB);
  double acc = 0.0;
  for (int J = 0; J <= R; ++J) acc += t.ypart[(int64_t)J * n + r];
  const int nseg = (nbr - R + SEG - 1) / SEG;
  for (int sg = 0; sg < nseg; ++sg) acc += t.ycpart[(int64_t)sg * n + r];
  for (int p = 0; p < jj; ++p) {
    const double Vip = t.A[i + (int64_t)(k + p) * n];
    const double Wip = t.W[i + (int64_t)p * n];
    acc -= Vip * t.s[jj + p] + Wip * t.s[p];
  }
  return tau * acc;
}

// Finish column c and start column c+1 in one pass over rows [c+1, n):
//   w = tau (y - V sW - W sV) + alpha v,
//   alpha = -1/2 tau (w_pre^T v) = -1/2 tau^2 (v^T A22 v - 2 sV.sW)
// (v^T A22 v from sy_symv's per-workgroup partials), stored in W(:, jj); v into
// A(:, c); then, when column c+1 is in this panel or is the tile's last column,
// its rank-2(jj+1) update and its norm partials.
__global__ __launch_bounds__(RB) void sy_wupd(const SyTile* __restrict__ tiles, int k, int c,
                                              int jj, int nqparts, int in_panel_next) {
  __shared__ double red[RB / 64];
  const SyTile t = tiles[blockIdx.z];
  const int64_t n = t.n;
  if (c >= n - 1) return;
  const int upd_next = in_panel_next || (c + 1 == n - 1);
  const double tau = t.tau[c];
  const double scal = t.scal[0];
  double qs = 0.0;
  for (int q = threadIdx.x; q < nqparts; q += RB) qs += t.qpart[q];
  const double vAv = block_sum(qs, red);
  double svw = 0.0;
  for (int p = 0; p < jj; ++p) svw += t.s[p] * t.s[jj + p];
  const double alpha = -0.5 * tau * tau * (vAv - 2.0 * svw);
  const int64_t c1 = c + 1;
  const int64_t i = c1 + (int64_t)blockIdx.x * RB + threadIdx.x;
  double* W = t.W;
  double a = 0.0;
  if (i < n) {
    const double* colc = t.A + (int64_t)c * n;
    const double v = (i == c1) ? 1.0 : colc[i] * scal;
    const double w = sy_wrow(t, k, c, jj, tau, i) + alpha * v;
    if (upd_next) {
      // W(c+1, jj): the same row formula at row c+1, where v = 1
      const double Wc1 = sy_wrow(t, k, c, jj, tau, c1) + alpha;
      a = t.A[i + c1 * n];
      double sacc = 0.0;
      for (int p = 0; p < jj; ++p) {
        const double* Vp = t.A + (int64_t)(k + p) * n;
        const double* Wp = W + (int64_t)p * n;
        sacc += Vp[i] * Wp[c1] + Wp[i] * Vp[c1];
      }
      sacc += v * Wc1 + w * 1.0;  // p = jj: V(c+1, jj) = 1
      a -= sacc;
    }
    // each thread reads and writes only its own row of columns c, c+1 and W(:, jj)
    W[i + (int64_t)jj * n] = w;
    t.A[i + (int64_t)c * n] = v;
    if (upd_next) t.A[i + c1 * n] = a;
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
                     const std::vector<SyTile>& host, int nb, Prof* prof) {
  const int rbl = ceil_div(nmax, RB);
  size_t tok = 0;
  for (int64_t k = 0; k < nmax - 1; k += nb) {
    const int pb = (int)std::min<int64_t>(nb, nmax - 1 - k);
    // first column of the panel: norm partials only (jj = 0)
    hipLaunchKernelGGL(sy_colupd, dim3(ceil_div(nmax - k, RB), 1, ntiles), dim3(RB), 0, st,
                       d_tiles, (int)k, (int)k, 0);
    for (int jj = 0; jj < pb; ++jj) {
      const int64_t c = k + jj;
      const int nparts = ceil_div(nmax - c, RB);
      const int nbr = ceil_div(nmax - c - 1, SB);
      const int gx = std::max(ceil_div(nbr, SEG), ceil_div(2 * jj, 4));
      if (prof) IMCOM_HIP_TRY(prof->begin(PS_SYMV, 2, &tok));
      hipLaunchKernelGGL(sy_symv, dim3(std::max(gx, 1), nbr + 1, ntiles), dim3(256), 0, st,
                         d_tiles, (int)k, (int)c, jj, nparts);
      if (prof) IMCOM_HIP_TRY(prof->end(PS_SYMV, tok));
      const int nb2 = std::max(ceil_div(nmax - c - 1, RB), 1);
      const int nq = std::max(gx, 1) * (nbr + 1);  // sy_symv's workgroup count
      hipLaunchKernelGGL(sy_wupd, dim3(nb2, 1, ntiles), dim3(RB), 0, st, d_tiles, (int)k, (int)c,
                         jj, nq, (jj + 1 < pb) ? 1 : 0);
    }
    (void)rbl;
    // trailing update A22 -= [V W] [W V]^T over rows/cols [k+nb, n)
    const int64_t r0 = k + nb;
    if (r0 < nmax) {
      hipLaunchKernelGGL(sy_pack, dim3(ceil_div(nmax - r0, 256), nb, ntiles), dim3(256), 0, st,
                         d_tiles, (int)k, nb, r0);
      if (prof) IMCOM_HIP_TRY(prof->begin(PS_SYR2K, 2, &tok));
      for (int b = 0; b < ntiles; ++b) {
        const SyTile& h = host[b];
        if (r0 >= h.n) continue;
        const int64_t mm = h.n - r0;
        IMCOM_HIP_TRY(dgemm(st, false, true, (int)mm, (int)mm, 2 * nb, -1.0, h.VW + r0, h.n,
                            h.WV + r0, h.n, 1.0, h.A + r0 + r0 * h.n, h.n));
      }
      if (prof) IMCOM_HIP_TRY(prof->end(PS_SYR2K, tok));
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

namespace imcom {

// ---- batched tiles ------------------------------------------------------------
// The per-panel trailing update and the four GEMMs of each back-transform
// block become one grouped launch across the tiles; the descriptors are
// built and uploaded once per batch shape.
void sy_make_plan(const std::vector<SyTile>& host, int64_t nmax, int nb, int nbb, SyPlan& plan) {
  plan.gp = GroupedPlan();
  plan.syr2k.clear();
  plan.bt.clear();
  for (int64_t k = 0; k < nmax - 1; k += nb) {
    const int64_t r0 = k + nb;
    std::vector<GemmDesc> g;
    for (const SyTile& h : host) {
      if (r0 >= h.n) continue;
      const int mm = (int)(h.n - r0);
      g.push_back(GemmDesc{mm, mm, 2 * nb, -1.0, 1.0, h.VW + r0, h.n, h.WV + r0, h.n,
                           h.A + r0 + r0 * h.n, h.n});
    }
    plan.syr2k.push_back(plan.gp.add(false, true, g));
  }
  if (nmax > 1) {
    const int64_t nblk = (nmax - 1 + nbb - 1) / nbb;
    for (int64_t bi = nblk - 1; bi >= 0; --bi) {
      const int kb = (int)(bi * nbb);
      std::vector<GemmDesc> gs, gy, gy2, gz;
      for (const SyTile& h : host) {
        if (kb >= h.n - 1) continue;
        const int rows = (int)(h.n - kb - 1);
        const int n = (int)h.n;
        double* Zr = h.Z + kb + 1;
        gs.push_back(GemmDesc{nbb, nbb, rows, 1.0, 0.0, h.Vb, rows, h.Vb, rows, h.Sb, nbb});
        gy.push_back(GemmDesc{nbb, n, rows, 1.0, 0.0, h.Vb, rows, Zr, h.n, h.Y, nbb});
        gy2.push_back(GemmDesc{nbb, n, nbb, 1.0, 0.0, h.Tb, nbb, h.Y, nbb, h.Y2, nbb});
        gz.push_back(GemmDesc{rows, n, nbb, -1.0, 1.0, h.Vb, rows, h.Y2, nbb, Zr, h.n});
      }
      std::array<int, 4> ids;
      ids[0] = plan.gp.add(true, false, gs);
      ids[1] = plan.gp.add(true, false, gy);
      ids[2] = plan.gp.add(false, false, gy2);
      ids[3] = plan.gp.add(false, false, gz);
      plan.bt.push_back(ids);
    }
  }
}

hipError_t sytrd_run_batched(hipStream_t st, const SyTile* d_tiles, int ntiles, int64_t nmax,
                             int nb, const SyPlan& plan, Prof* prof) {
  size_t tok = 0;
  int panel = 0;
  for (int64_t k = 0; k < nmax - 1; k += nb, ++panel) {
    const int pb = (int)std::min<int64_t>(nb, nmax - 1 - k);
    hipLaunchKernelGGL(sy_colupd, dim3(ceil_div(nmax - k, RB), 1, ntiles), dim3(RB), 0, st,
                       d_tiles, (int)k, (int)k, 0);
    for (int jj = 0; jj < pb; ++jj) {
      const int64_t c = k + jj;
      const int nparts = ceil_div(nmax - c, RB);
      const int nbr = ceil_div(nmax - c - 1, SB);
      const int gx = std::max(ceil_div(nbr, SEG), ceil_div(2 * jj, 4));
      if (prof) IMCOM_HIP_TRY(prof->begin(PS_SYMV, 2, &tok));
      hipLaunchKernelGGL(sy_symv, dim3(std::max(gx, 1), nbr + 1, ntiles), dim3(256), 0, st,
                         d_tiles, (int)k, (int)c, jj, nparts);
      if (prof) IMCOM_HIP_TRY(prof->end(PS_SYMV, tok));
      const int nb2 = std::max(ceil_div(nmax - c - 1, RB), 1);
      const int nq = std::max(gx, 1) * (nbr + 1);  // sy_symv's workgroup count
      hipLaunchKernelGGL(sy_wupd, dim3(nb2, 1, ntiles), dim3(RB), 0, st, d_tiles, (int)k, (int)c,
                         jj, nq, (jj + 1 < pb) ? 1 : 0);
    }
    const int64_t r0 = k + nb;
    if (r0 < nmax) {
      hipLaunchKernelGGL(sy_pack, dim3(ceil_div(nmax - r0, 256), nb, ntiles), dim3(256), 0, st,
                         d_tiles, (int)k, nb, r0);
      if (prof) IMCOM_HIP_TRY(prof->begin(PS_SYR2K, 2, &tok));
      IMCOM_HIP_TRY(plan.gp.run(st, plan.syr2k[panel]));
      if (prof) IMCOM_HIP_TRY(prof->end(PS_SYR2K, tok));
    }
  }
  hipLaunchKernelGGL(sy_last_diag, dim3(1, 1, ntiles), dim3(1), 0, st, d_tiles);
  return hipGetLastError();
}

hipError_t ormtr_run_batched(hipStream_t st, const SyTile* d_tiles, int ntiles, int64_t nmax,
                             int nbb, const SyPlan& plan) {
  if (nmax <= 1) return hipSuccess;
  const int64_t nblk = (nmax - 1 + nbb - 1) / nbb;
  int idx = 0;
  for (int64_t bi = nblk - 1; bi >= 0; --bi, ++idx) {
    const int kb = (int)(bi * nbb);
    const std::array<int, 4>& ids = plan.bt[idx];
    hipLaunchKernelGGL(bt_gather, dim3(ceil_div(nmax - kb - 1, 256), nbb, ntiles), dim3(256), 0,
                       st, d_tiles, kb, nbb);
    IMCOM_HIP_TRY(plan.gp.run(st, ids[0]));
    hipLaunchKernelGGL(bt_build_T, dim3(1, 1, ntiles), dim3(256), 0, st, d_tiles, kb, nbb);
    IMCOM_HIP_TRY(plan.gp.run(st, ids[1]));
    IMCOM_HIP_TRY(plan.gp.run(st, ids[2]));
    IMCOM_HIP_TRY(plan.gp.run(st, ids[3]));
  }
  return hipGetLastError();
}

}  // namespace imcom
