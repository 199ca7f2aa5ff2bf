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
// A22 -= V W^T + W V^T on the FP64 matrix cores, on the tiles on or below
// the diagonal only: nothing reads A above its diagonal.  The matrix-vector
// product reads the lower triangle in 64x64 blocks, each block serving its
// rows and its columns.
//
// Three launches per column (two for one tile, where sy_midwupd does the
// last two):
//   sy_symv      partial sums of u = A22 x for the unscaled column x (so it
//                does not wait for the reflector) and of x^T u over the lower
//                triangle, the 2jj dot products [V W]^T x
//   sy_mid       one workgroup: the reflector from the column's norm
//                partials, v^T A22 v, V^T v, W^T v, alpha and W(c+1, jj)
//   sy_wupd      w = tau (y - V sW - W sV) + alpha v with
//                alpha = -1/2 tau^2 (v^T A22 v - 2 sV.sW) (the same value as
//                -1/2 tau w_pre^T v); rank-2(jj+1) update of the next column
//                and its norm partials
// Grid-wide sums are left as per-workgroup partials and finished by one
// workgroup (sy_mid) rather than with device-scope fences, which on a
// multi-XCD GPU write back the L2 and cost more than the sums themselves.
// All sums are fixed-order (partials per workgroup, reduced in workgroup
// order), so runs are bitwise reproducible.
//
// Tiles of a batch advance column by column in lockstep; blockIdx.z selects
// the tile and a tile whose columns are exhausted exits.
#include <hip/hip_ext.h>

#include "common.h"
#include "kernels.h"

#include <algorithm>
#include <type_traits>
#include <vector>

namespace imcom {

constexpr int RB = 256;  // rows per workgroup in the row-parallel kernels

// Norm partials of column c: two buffers by the column's parity, so that one
// launch can read column c's partials while it writes column c+1's.
__device__ __forceinline__ double* norm_parts(const SyTile& t, int64_t c) {
  return t.part + (c & 1) * kNormParts;
}

__device__ __forceinline__ double dlapy2(double x, double y) {
  const double xa = fabs(x), ya = fabs(y);
  const double w = fmax(xa, ya), z = fmin(xa, ya);
  if (z == 0.0) return w;
  const double q = z / w;
  return w * sqrt(1.0 + q * q);
}

// First column c of a panel (already up to date): per-workgroup partials of
// its squared norm over rows [c+2, n), read by sy_mid.
__global__ __launch_bounds__(RB) void sy_colnorm(const SyTile* __restrict__ tiles, int c) {
  __shared__ double red[RB / 64];
  const SyTile t = tiles[blockIdx.z];
  const int64_t n = t.n;
  if (c >= n - 1) return;  // this tile is reduced (uniform over its workgroups)
  const int64_t i = c + 2 + (int64_t)blockIdx.x * RB + threadIdx.x;
  const double a = (i < n) ? t.A[i + (int64_t)c * n] : 0.0;
  const double tot = block_sum(a * a, red);
  if (threadIdx.x == 0) norm_parts(t, c)[blockIdx.x] = tot;
}

constexpr int SB = 64;  // symmetric matrix-vector block
constexpr int SEG = 4;  // blocks of one block-column per workgroup
constexpr int DCH = 2048;  // rows per dot-product chunk (kDotChunks * DCH >= 32768)
// bytes of the trailing triangle's tail that sy_symv keeps in the 256 MB
// last-level cache from one column to the next (the rest it streams)
constexpr int64_t kSymvKeepBytes = (int64_t)224 << 20;

// The lower-triangle 64x64 blocks of A into the block layout At
// (tri_block_offset), before the first panel.  Later panels' blocks are
// written by the trailing update's GEMM epilogue.
__global__ __launch_bounds__(256) void sy_tile_init(const SyTile* __restrict__ tiles) {
  const SyTile t = tiles[blockIdx.z];
  const int I = blockIdx.x, J = blockIdx.y;
  if (I >= t.nbt || J > I) return;
  const int64_t n = t.n;
  double* __restrict__ dst = t.At + tri_block_offset(I, J, t.nbt);
  for (int e = threadIdx.x; e < 4096; e += 256) {
    const int64_t i = (int64_t)I * 64 + (e & 63), j = (int64_t)J * 64 + (e >> 6);
    dst[e] = (i < n && j < n) ? t.A[i + j * n] : 0.0;
  }
}

// u = A22 x for the unscaled column x = A(c+1:n, c), reading only the lower
// triangle (diagonal included) of the stored trailing matrix A22 = A(c+1:n,
// c+1:n).  The reflector's v is scal * (x - beta e_0), so
// A22 v = scal (u - beta A22 e_0) and
// v^T A22 v = scal^2 (x^T u - 2 beta u_0 + beta^2 A22(0,0)): this kernel
// needs nothing from the reflector and runs before it is formed.
//
// Blocks are the 64x64 tiles of A at absolute positions (rows and columns
// [64I, 64I+64)), read from the block-layout copy At (each block 32 KB
// contiguous: measured 5.5 TB/s against 4.3 for the same blocks read out of
// column-major A).  Within a panel the stored trailing matrix does not
// change.  The trailing matrix lives in At: the trailing update's GEMM reads
// and writes it there, and writes column-major A only for the next panel's
// columns, the only part of the trailing matrix the column kernels read.
// The trailing matrix's first block row and column are partial (entries
// before c+1 are skipped).  With j0 = (c+1)/64, local block-column
// Jl = J - j0.  Workgroup (seg, Jl) covers block-column J over the SEG
// blocks I in [J + seg*SEG, J + (seg+1)*SEG).  Wavefront w owns columns
// 16w..16w+15 of the block-column and reads them two at a time, all SEG
// blocks' rows per step.  Lane l holds rows l, l+64, ... and accumulates
// their row sums over the wavefront's columns (reduced across the 4
// wavefronts in LDS and written to ypart[Jl][row]); each column's sum against
// x over the segment's rows is collected in LDS and the 64 sums go to
// ycpart[seg][cols] in one coalesced store; the workgroup's share of x^T u
// goes to qpart.  Rows and columns are absolute indices.  Diagonal blocks are
// read on and below the diagonal only (the trailing update leaves the upper
// triangle stale): their strictly lower part serves rows and columns, the
// diagonal serves rows, and x^T (L + D + L^T) x = 2 x^T (L + D) x - x^T D x.
// Workgroups in rows y >= nbl form the 2jj unscaled dot products
// sigma = [V W]^T x over rows [c+1, n) in chunks of DCH rows, one chunk per
// wavefront (partials in dpart[q][chunk], summed by sy_mid): a whole column
// per wavefront would leave one long serial wavefront as the launch's tail.
// NT: block-columns J < jkeep are read with nontemporal loads, the tail
// J >= jkeep (a lower triangle of at most kSymvKeepBytes) with ordinary ones,
// so the tail stays in the 256 MB last-level cache for the next column while
// the rest streams past it.  Each workgroup takes one of the two compiled
// loops (a per-load select merges them into ordinary loads).  Measured at
// n = 16384, symv total: all ordinary loads while the triangle exceeds the
// cache 1165 ms, nontemporal above s = 8704 1086 ms, cached tail of 96 / 160
// / 224 / 256 MB 1048 / 1036 / 1027 / 1032 ms.  kkeep: block-columns of
// each tile's cached tail.
// Two columns per step (eight block loads in flight per lane) was measured
// against 4, 8 and 16 at n = 16384 (bitwise the same results): 4 and 8 are
// no faster, and 16 (the whole strip in one step, 183 VGPRs, two waves per
// SIMD) is 10-20% slower at every trailing size above 2000.
template <bool NT>
__global__ __launch_bounds__(256) void sy_symv(const SyTile* __restrict__ tiles, int k, int c,
                                               int jj, int kkeep) {
  __shared__ double xJ[SB];
  __shared__ double csum[SB];
  __shared__ double rred[4][SEG * SB];
  __shared__ double qred[4];
  const SyTile t = tiles[blockIdx.z];
  const int64_t n = t.n;
  if (c >= n - 1) return;  // this tile is reduced (uniform over its workgroups)
  const double* __restrict__ colc = t.A + (int64_t)c * n;
  const int64_t c1 = c + 1;
  const int j0 = (int)(c1 / SB);
  const int nbl = (int)((n + SB - 1) / SB) - j0;  // local block count
  const int jkeep = t.nbt - kkeep;  // first block-column of the cached tail
  const int Jl = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int I0l = Jl + blockIdx.x * SEG;
  double qt = 0.0;  // this workgroup's share of x^T u
  if (Jl >= nbl) {
    const int nch = (int)((n - c1 + DCH - 1) / DCH);
    const int widx = ((Jl - nbl) * gridDim.x + blockIdx.x) * 4 + w;
    const int q = widx / nch, ch = widx % nch;
    if (q < 2 * jj) {
      const double* X = q < jj ? t.A + (int64_t)(k + q) * n : t.W + (int64_t)(q - jj) * n;
      const int64_t i0 = c1 + (int64_t)ch * DCH;
      const int64_t i1 = min(n, i0 + DCH);
      double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
      int64_t i = i0 + lane;
      for (; i + 192 < i1; i += 256) {
        a0 += X[i] * colc[i];
        a1 += X[i + 64] * colc[i + 64];
        a2 += X[i + 128] * colc[i + 128];
        a3 += X[i + 192] * colc[i + 192];
      }
      for (; i < i1; i += 64) a0 += X[i] * colc[i];
      const double acc = wave_sum((a0 + a1) + (a2 + a3));
      if (lane == 0) t.dpart[q * kDotChunks + ch] = acc;
    }
  } else if (I0l < nbl) {
    const int J = j0 + Jl;
    const int I0 = j0 + I0l;
    const int nb = min(SEG, nbl - I0l);  // blocks in this segment
    if (tid < SB) {
      const int64_t j = (int64_t)J * SB + tid;
      xJ[tid] = (j >= c1 && j < n) ? colc[j] : 0.0;
    }
    double xr[SEG], racc[SEG];
#pragma unroll
    for (int u = 0; u < SEG; ++u) {
      const int64_t i = (int64_t)(I0 + u) * SB + lane;
      xr[u] = (u < nb && i >= c1 && i < n) ? colc[i] : 0.0;
      racc[u] = 0.0;
    }
    __syncthreads();  // xJ
    double qacc = 0.0;
    const bool diag = (I0 == J);
    // the block loop, compiled once per load kind; the workgroup's
    // block-column picks one (uniform branch)
    auto blocks = [&](auto ntc) {
    constexpr bool NTL = decltype(ntc)::value;
#pragma unroll 1
    for (int q0 = 0; q0 < 16; q0 += 2) {
      double a[2][SEG];
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        const int cl = 16 * w + q0 + p;
        const int64_t j = (int64_t)J * SB + cl;
        const bool jv = j >= c1 && j < n;
#pragma unroll
        for (int u = 0; u < SEG; ++u) {
          const int64_t i = (int64_t)(I0 + u) * SB + lane;
          const bool ok = jv && (u < nb) && i >= c1 && i < n && !(diag && u == 0 && lane < cl);
          // global address space: global_load with a scalar base (the
          // struct-held pointer would otherwise give flat loads)
          const __attribute__((address_space(1))) double* src =
              (const __attribute__((address_space(1))) double*)t.At +
              tri_block_offset(I0 + u, J, t.nbt) + cl * SB + lane;
          a[p][u] = ok ? (NTL ? __builtin_nontemporal_load(src) : *src) : 0.0;
        }
      }
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        const int cl = 16 * w + q0 + p;
        const double xc = xJ[cl];
        double cs = 0.0;
#pragma unroll
        for (int u = 0; u < SEG; ++u) {
          racc[u] += a[p][u] * xc;
          const bool dg = diag && u == 0 && lane == cl;
          cs += dg ? 0.0 : a[p][u] * xr[u];
          if (dg) qacc -= xr[0] * xr[0] * a[p][0];
        }
        cs = wave_sum(cs);
        if (lane == 0) csum[cl] = cs;
      }
    }
    };
    if (NT && J < jkeep)
      blocks(std::true_type{});
    else
      blocks(std::false_type{});
#pragma unroll
    for (int u = 0; u < SEG; ++u) rred[w][u * SB + lane] = racc[u];
    __syncthreads();
    if (tid < SB) {  // the block-column's 64 column sums, one coalesced store
      const int64_t j = (int64_t)J * SB + tid;
      if (j >= c1 && j < n) t.ycpart[(int64_t)blockIdx.x * n + j] = csum[tid];
    }
    for (int idx = tid; idx < nb * SB; idx += 256) {
      const int u = idx / SB, l = idx % SB;
      const int64_t i = (int64_t)(I0 + u) * SB + l;
      if (i >= c1 && i < n) {
        const double rs = rred[0][idx] + rred[1][idx] + rred[2][idx] + rred[3][idx];
        __builtin_nontemporal_store(rs, t.ypart + (int64_t)Jl * n + i);
        qacc += 2.0 * colc[i] * rs;
      }
    }
    qt = block_sum(qacc, qred);
  }
  if (tid == 0) t.qpart[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = qt;
}

// One workgroup per tile, between sy_symv and sy_wupd of column c: the
// reflector (LAPACK dlarfg conventions: beta = -sign(alpha) |(alpha, x')|,
// tau = (beta - alpha)/beta, v = [1, x'/(alpha - beta)]) from the norm
// partials of x' = x(1:), stored as d[c], e[c], tau[c]; then
//   v^T A22 v = scal^2 (x^T u - 2 beta u_0 + beta^2 A22(0,0)),
//   V^T v = scal (sigma_V - beta V(c+1,:)), W^T v likewise (into s; sigma
//   summed from sy_symv's chunk partials),
//   alpha = -1/2 tau^2 (v^T A22 v - 2 (V^T v).(W^T v)),
//   W(c+1, jj) = tau (y_0 - V(c+1,:) W^T v - W(c+1,:) V^T v) + alpha,
// y_0 = scal (u_0 - beta A22(0,0)).  scal[0..4] = scal, tau, alpha, beta,
// W(c+1, jj).  nnorm / nq: partial counts of the norm and of x^T u.
constexpr int MID_THREADS = 1024;

// Every load is issued up front: the three grid-wide partial sums (column
// norm, x^T u, row c+1 of u) in one pass with one block reduction, the chunk
// sums of the dot products, and V(c+1,:), W(c+1,:); one wavefront then forms
// the jj-term sums.
// The reflector step's results, in LDS: scal[0..4] as the tile's scal, ss as
// its s.
struct MidOut {
  double scal[5];
  double ss[2 * 64];
};

// sy_mid's work for one tile (c < n - 1), run by NTH threads; the results
// go to `out` (read after a barrier) and, with to_global, to the tile's s,
// scal, d, e and tau.  `red`: 3 x NTH/64 doubles of LDS.
template <int NTH>
__device__ __forceinline__ void mid_body(const SyTile& t, int k, int c, int jj, int nnorm, int nq,
                                         double (*red)[NTH / 64], MidOut& out, bool to_global) {
  constexpr int NW = NTH / 64;
  double* ss = out.ss;
  const int64_t n = t.n;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int64_t c1 = c + 1;
  const int nbl = (int)((n + SB - 1) / SB) - (int)(c1 / SB);
  const int nseg0 = (nbl + SEG - 1) / SEG;
  // phase 1: loads
  double pn = 0.0, pu = 0.0;
  const double* __restrict__ part = norm_parts(t, c);
  for (int q = tid; q < nnorm; q += NTH) pn += part[q];
  double qs[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // eight loads in flight
  {
    int q = tid;
    for (; q + 7 * NTH < nq; q += 8 * NTH) {
#pragma unroll
      for (int r = 0; r < 8; ++r) qs[r] += t.qpart[q + r * NTH];
    }
    for (; q < nq; q += NTH) qs[0] += t.qpart[q];
  }
  const double q0 = qs[0] + qs[4], q1 = qs[1] + qs[5], q2 = qs[2] + qs[6], q3 = qs[3] + qs[7];
  for (int sg = tid; sg < nseg0; sg += NTH) pu += t.ycpart[(int64_t)sg * n + c1];
  double sig = 0.0, xc1 = 0.0;
  if (tid < 2 * jj) {
    const int nch = (int)((n - c1 + DCH - 1) / DCH);
    for (int ch = 0; ch < nch; ++ch) sig += t.dpart[tid * kDotChunks + ch];
    const double* X = tid < jj ? t.A + (int64_t)(k + tid) * n : t.W + (int64_t)(tid - jj) * n;
    xc1 = X[c1];
  }
  double vc = 0.0, wc = 0.0;  // V(c+1, p), W(c+1, p) for p = tid < jj
  if (tid < jj) {
    vc = t.A[c1 + (int64_t)(k + tid) * n];
    wc = t.W[c1 + (int64_t)tid * n];
  }
  const double alpha_r = t.A[c1 + c * n];
  const double acc = t.A[c + c * n];
  const double g0 = t.At[tri_block_offset((int)(c1 >> 6), (int)(c1 >> 6), t.nbt) + (c1 & 63) * 65];
  const double y00 = t.ypart[c1];
  // one block reduction for the three sums
  {
    const double a = wave_sum(pn), b = wave_sum((q0 + q1) + (q2 + q3)), d = wave_sum(pu);
    if (lane == 0) {
      red[0][w] = a;
      red[1][w] = b;
      red[2][w] = d;
    }
  }
  __syncthreads();
  double xn2 = 0.0, xu = 0.0, u0 = 0.0;
  for (int v = 0; v < NW; ++v) {
    xn2 += red[0][v];
    xu += red[1][v];
    u0 += red[2][v];
  }
  u0 += y00;
  const double xnorm = sqrt(xn2);
  double tau, beta, scal;
  if (xnorm == 0.0) {
    tau = 0.0;
    beta = alpha_r;
    scal = 0.0;
  } else {
    beta = -copysign(dlapy2(alpha_r, xnorm), alpha_r);
    tau = (beta - alpha_r) / beta;
    scal = 1.0 / (alpha_r - beta);
  }
  if (tid < 2 * jj) ss[tid] = scal * (sig - beta * xc1);
  __syncthreads();
  if (to_global && tid < 2 * jj) t.s[tid] = ss[tid];
  if (w == 0) {  // jj <= 63: one lane per panel column
    const double sv = lane < jj ? ss[lane] * ss[jj + lane] : 0.0;
    const double cr = lane < jj ? vc * ss[jj + lane] + wc * ss[lane] : 0.0;
    const double svw = wave_sum(sv), corr = wave_sum(cr);
    if (lane == 0) {
      const double vAv = scal * scal * (xu - 2.0 * beta * u0 + beta * beta * g0);
      const double alpha = -0.5 * tau * tau * (vAv - 2.0 * svw);
      const double y0 = scal * (u0 - beta * g0);
      const double sc[5] = {scal, tau, alpha, beta, tau * (y0 - corr) + alpha};
#pragma unroll
      for (int q = 0; q < 5; ++q) out.scal[q] = sc[q];
      if (to_global) {
#pragma unroll
        for (int q = 0; q < 5; ++q) t.scal[q] = sc[q];
        t.d[c] = acc;
        t.e[c] = beta;
        t.tau[c] = tau;
      }
    }
  }
}

// The batched launches' reflector step: 256 threads, so that each tile's
// workgroup finds room beside the other streams' matrix-vector workgroups
// (64-tile mosaic as three concurrent batched solves: 1024 threads 785 ms per
// solve, 256 760 ms, 128 762 ms).
constexpr int MID_THREADS_BATCH = 256;

__global__ __launch_bounds__(MID_THREADS_BATCH) void sy_mid(const SyTile* __restrict__ tiles,
                                                            int k, int c, int jj, int nnorm,
                                                            int nq) {
  __shared__ double red[3][MID_THREADS_BATCH / 64];
  __shared__ MidOut mo;
  const SyTile t = tiles[blockIdx.z];
  if (c >= t.n - 1) return;
  mid_body<MID_THREADS_BATCH>(t, k, c, jj, nnorm, nq, red, mo, true);
}

constexpr int WR = 16;  // rows per sy_wupd workgroup
constexpr int WL = 16;  // lanes per row (256 threads)

// Finish column c and start column c+1 over rows [c+1, n), 16 rows per
// row block and 16 lanes per row (lane l sums partials l, l+16, ...):
//   y = scal (u - beta A22(:,0)),  w = tau (y - V sW - W sV) + alpha v,
// stored in W(:, jj); v into A(:, c) below row c+1 (the unit at row c+1 is
// implicit: A(c+1, c) keeps the column's entry before the reflection, which
// sy_mid reads); then, when column c+1 is in this panel or is the tile's last
// column, its rank-2(jj+1) update and its norm partials (for the next
// column's reflector step).
//
// wupd_body: row block b (rows c+1+16b ..), 256 threads (tid), the reflector
// step's results sc[0..4] and sv[0..2jj) (the tile's scal and s, or their
// LDS copies: SV_LDS), LDS scratch red (3 x 4 x WR) and nred (4); row blocks
// b >= nw write no norm partial.
template <bool WIDE, bool SV_LDS>
__device__ __forceinline__ void wupd_body(const SyTile& t, int k, int c, int jj, int in_panel_next,
                                          int b, int tid, int nw, const double* sc,
                                          const double* sv, double (*red)[4][WR], double* nred) {
  const int64_t n = t.n;
  const bool upd_next = in_panel_next || (c + 1 == n - 1);
  const int rl = tid & (WR - 1), lane = tid >> 4, w = tid >> 6;
  const int64_t c1 = c + 1;
  const int j0 = (int)(c1 / SB);
  const int nbl = (int)((n + SB - 1) / SB) - j0;
  const int64_t i = c1 + (int64_t)b * WR + rl;
  auto sval = [&](int q) { return SV_LDS ? sv[q] : ldg(sv + q); };
  // WIDE (one tile): every load is issued before the first sum (straight-line
  // code, one memory round trip): 16 ypart and 4 ycpart partials of u per
  // lane (all of them up to n = 16384; longer rows finish in the loops
  // below), and the panel's V, W entries, up to 4 columns per lane.  This
  // takes 118 VGPRs; the batched launch, where most workgroups exit at once,
  // runs the 36-VGPR form with eight partials in flight per step instead
  // (measured: 8.4 -> 7.0 us per launch for one tile at n = 16384, but
  // 18 -> 22 us for the 64-tile batch).
  const double scal = sc[0], tau = sc[1], alpha = sc[2], beta = sc[3], w1 = sc[4];
  double g = 0.0, xc = 0.0;  // A22(i, 0) as stored, and A(i, c): used after the reduction
  double qu = 0.0, qc = 0.0, q1 = 0.0;  // lane partials: u, V sW + W sV, update
  if (i < n) {
    if (tid < WR) {
      // A22(:, 0) as stored (the last trailing update's value: column c+1 of
      // the block copy; column-major A may not hold it when c+1 starts the
      // next panel)
      g = ldg(t.At + tri_block_offset((int)(i >> 6), (int)(c1 >> 6), t.nbt) + (c1 & 63) * 64 +
              (i & 63));
      xc = ldg(t.A + i + (int64_t)c * n);
    }
    const int R = (int)(i / SB) - j0;  // local block row of row i
    const int nseg = (nbl - R + SEG - 1) / SEG;
    if (WIDE) {
      double u[20];
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int J = lane + q * WL;
        u[q] = (J <= R) ? ldg(t.ypart + (int64_t)J * n + i) : 0.0;
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int sg = lane + q * WL;
        u[16 + q] = (sg < nseg) ? ldg(t.ycpart + (int64_t)sg * n + i) : 0.0;
      }
      double pv[4][6];  // V(i,p), W(i,p), sW(p), sV(p), W(c+1,p), V(c+1,p); p < jj <= 63
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int p = lane + q * WL;
        const bool ok = p < jj, okn = ok && upd_next;
        pv[q][0] = ok ? ldg(t.A + i + (int64_t)(k + p) * n) : 0.0;
        pv[q][1] = ok ? ldg(t.W + i + (int64_t)p * n) : 0.0;
        pv[q][2] = ok ? sval(jj + p) : 0.0;
        pv[q][3] = ok ? sval(p) : 0.0;
        pv[q][4] = okn ? ldg(t.W + c1 + (int64_t)p * n) : 0.0;
        pv[q][5] = okn ? ldg(t.A + c1 + (int64_t)(k + p) * n) : 0.0;
      }
      for (int J = lane + 16 * WL; J <= R; J += WL) u[0] += ldg(t.ypart + (int64_t)J * n + i);
      for (int sg = lane + 4 * WL; sg < nseg; sg += WL) u[16] += ldg(t.ycpart + (int64_t)sg * n + i);
#pragma unroll
      for (int q = 0; q < 10; ++q) u[q] += u[q + 10];
#pragma unroll
      for (int q = 0; q < 5; ++q) u[q] += u[q + 5];
      qu = ((u[0] + u[1]) + (u[2] + u[3])) + u[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        qc += pv[q][0] * pv[q][2] + pv[q][1] * pv[q][3];
        q1 += pv[q][0] * pv[q][4] + pv[q][1] * pv[q][5];
      }
    } else {
      double u[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      int J = lane;
      for (; J + 7 * WL <= R; J += 8 * WL) {
#pragma unroll
        for (int q = 0; q < 8; ++q) u[q] += ldg(t.ypart + (int64_t)(J + q * WL) * n + i);
      }
      for (; J <= R; J += WL) u[0] += ldg(t.ypart + (int64_t)J * n + i);
      int sg = lane;
      for (; sg + 3 * WL < nseg; sg += 4 * WL) {
#pragma unroll
        for (int q = 0; q < 4; ++q) u[4 + q] += ldg(t.ycpart + (int64_t)(sg + q * WL) * n + i);
      }
      for (; sg < nseg; sg += WL) u[1] += ldg(t.ycpart + (int64_t)sg * n + i);
      qu = ((u[0] + u[1]) + (u[2] + u[3])) + ((u[4] + u[5]) + (u[6] + u[7]));
#pragma unroll
      for (int q = 0; q < 4; ++q) {  // p < jj <= 63: at most four per lane
        const int p = lane + q * WL;
        if (p < jj) {
          const double Vip = ldg(t.A + i + (int64_t)(k + p) * n);
          const double Wip = ldg(t.W + i + (int64_t)p * n);
          qc += Vip * sval(jj + p) + Wip * sval(p);
          if (upd_next)
            q1 += Vip * ldg(t.W + c1 + (int64_t)p * n) + Wip * ldg(t.A + c1 + (int64_t)(k + p) * n);
        }
      }
    }
  }
  // lanes of one row: 4 per wavefront (thread bits 4, 5), then across the 4 wavefronts
  qu += __shfl_xor(qu, 16, kWave);
  qc += __shfl_xor(qc, 16, kWave);
  q1 += __shfl_xor(q1, 16, kWave);
  qu += __shfl_xor(qu, 32, kWave);
  qc += __shfl_xor(qc, 32, kWave);
  q1 += __shfl_xor(q1, 32, kWave);
  if ((lane & 3) == 0) {
    red[0][w][rl] = qu;
    red[1][w][rl] = qc;
    red[2][w][rl] = q1;
  }
  __syncthreads();
  double sq = 0.0;
  if (tid < WR && i < n) {
    const double U = red[0][0][rl] + red[0][1][rl] + red[0][2][rl] + red[0][3][rl];
    const double Cr = red[1][0][rl] + red[1][1][rl] + red[1][2][rl] + red[1][3][rl];
    const double Q1 = red[2][0][rl] + red[2][1][rl] + red[2][2][rl] + red[2][3][rl];
    const double v = (i == c1) ? 1.0 : xc * scal;
    const double y = scal * (U - beta * g);
    const double wv = tau * (y - Cr) + alpha * v;
    if (upd_next) {
      // p = jj: V(c+1, jj) = 1, W(c+1, jj) = scal[4]
      const double a = g - (Q1 + v * w1 + wv);
      t.A[i + c1 * n] = a;
      if (i >= c1 + 2) sq = a * a;
      if (c1 == n - 1) t.d[c1] = a;
    }
    // each row's column-c, column-(c+1) and W(:, jj) entries belong to this thread
    t.W[i + (int64_t)jj * n] = wv;
    if (i > c1) t.A[i + (int64_t)c * n] = v;
  }
  if (upd_next && c1 < n - 1) {  // the 4 wavefronts' sums in order, as block_sum
    const double ws = wave_sum(sq);
    __syncthreads();
    if ((tid & 63) == 0) nred[w] = ws;
    __syncthreads();
    double tot = 0.0;
    for (int q = 0; q < 4; ++q) tot += nred[q];
    if (tid == 0 && b < nw) norm_parts(t, c1)[b] = tot;
  }
}

template <bool WIDE>
__global__ __launch_bounds__(256) void sy_wupd(const SyTile* __restrict__ tiles, int k, int c,
                                               int jj, int in_panel_next) {
  __shared__ double red[3][4][WR];
  __shared__ double nred[4];
  const SyTile t = tiles[blockIdx.z];
  if (c >= t.n - 1) return;  // this tile is reduced (uniform over its workgroups)
  wupd_body<WIDE, false>(t, k, c, jj, in_panel_next, blockIdx.x, threadIdx.x, gridDim.x, t.scal,
                         t.s, red, nred);
}

// sy_mid and sy_wupd in one launch, for columns whose partial sums are few:
// every workgroup forms the reflector itself (the same code and summation
// order as sy_mid, so the results are bitwise those of the two launches),
// then its four groups of 256 threads run sy_wupd's row blocks 4b .. 4b+3.
// Workgroup 0 also stores the reflector step's results.  nw: sy_wupd's row
// block count.  One tile only: measured at n = 16384 the reduction went 1401
// -> 1380 ms (the repeated reflector step reads ~17000 partials per
// workgroup at the first columns and still gains), at n = 1024 the step 24.9
// -> 24.2 ms; in the 64-tile batch, where every workgroup of every tile
// repeats it, fusing above trailing size 2000 cost 789 -> 888-903 ms per
// solve, and below it gained nothing.
template <bool WIDE>
__global__ __launch_bounds__(MID_THREADS) void sy_midwupd(const SyTile* __restrict__ tiles, int k,
                                                          int c, int jj, int nnorm, int nq,
                                                          int in_panel_next, int nw) {
  __shared__ double red[3][MID_THREADS / 64];
  __shared__ MidOut mo;
  __shared__ double wred[MID_THREADS / 256][3][4][WR];
  __shared__ double nred[MID_THREADS / 256][4];
  const SyTile t = tiles[blockIdx.z];
  if (c >= t.n - 1) return;
  mid_body<MID_THREADS>(t, k, c, jj, nnorm, nq, red, mo, blockIdx.x == 0);
  __syncthreads();
  const int g = threadIdx.x >> 8;
  wupd_body<WIDE, true>(t, k, c, jj, in_panel_next, blockIdx.x * (MID_THREADS / 256) + g,
                        threadIdx.x & 255, nw, mo.scal, mo.ss, wred[g], nred[g]);
}

// Copy the panel's V (with explicit zeros above the unit) and W into the
// [V W] and [W V] operands of the trailing update, rows [r0, n).
__global__ void sy_pack(const SyTile* __restrict__ tiles, int k, int nb, int64_t r0) {
  const SyTile t = tiles[blockIdx.z];
  const int64_t n = t.n;
  const int64_t i = r0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (i >= n || k + p >= n - 1) return;
  const double v = (i == k + p + 1) ? 1.0 : (i > k + p + 1 ? t.A[i + (int64_t)(k + p) * n] : 0.0);
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

constexpr int kMaxNBB = 128;  // largest diagonal block of T that bt_build_T holds in LDS

// T(j,j) = tau_j, T(0:j, j) = -tau_j T(0:j, 0:j) S(0:j, j), column by column,
// for the diagonal block T(o:o+len, o:o+len), o = blockIdx.x * kMaxNBB,
// len = min(nbb - o, kMaxNBB) (S and T are nbb x nbb, ld nbb).  The block
// lives in LDS (column-major, ld kMaxNBB, 128 KB); thread i owns row i, so a
// column's rows depend only on the same thread's earlier entries and the
// shared S column, which is loaded one column ahead.  Workgroups past the
// first also zero the rows of T below the first diagonal block
// (T(o:o+len, 0:o)); the block above them is formed by sy_t_join.
__global__ __launch_bounds__(256) void bt_build_T(const SyTile* __restrict__ tiles, int kb,
                                                  int nbb) {
  __shared__ double Ts[kMaxNBB * kMaxNBB];
  __shared__ double sS[2][kMaxNBB];
  const SyTile t = tiles[blockIdx.z];
  const int64_t n = t.n;
  if (kb >= n - 1) return;
  const int o = blockIdx.x * kMaxNBB;
  const int len = min(nbb - o, kMaxNBB);
  const double* __restrict__ S = t.Sb + o + (int64_t)o * nbb;  // S(o.., o..), ld nbb
  double* __restrict__ T = t.Tb + o + (int64_t)o * nbb;
  const int i = threadIdx.x;
  double sreg = 0.0;  // S(i, j) for the coming column j
  for (int j = 0; j < len; ++j) {
    if (i < j) sS[j & 1][i] = sreg;
    const double snext = (j + 1 < len && i < j + 1) ? S[i + (int64_t)(j + 1) * nbb] : 0.0;
    __syncthreads();
    const int64_t c = kb + o + j;
    const double tj = (c < n - 1) ? t.tau[c] : 0.0;
    if (i < len) {
      double v;
      if (i < j) {
        const double* sj = sS[j & 1];
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        int l = i;
        for (; l + 3 < j; l += 4) {
          a0 += Ts[l * kMaxNBB + i] * sj[l];
          a1 += Ts[(l + 1) * kMaxNBB + i] * sj[l + 1];
          a2 += Ts[(l + 2) * kMaxNBB + i] * sj[l + 2];
          a3 += Ts[(l + 3) * kMaxNBB + i] * sj[l + 3];
        }
        for (; l < j; ++l) a0 += Ts[l * kMaxNBB + i] * sj[l];
        v = -tj * ((a0 + a1) + (a2 + a3));
      } else {
        v = (i == j) ? tj : 0.0;
      }
      Ts[j * kMaxNBB + i] = v;
    }
    sreg = snext;
  }
  __syncthreads();
  for (int idx = i; idx < len * len; idx += blockDim.x)
    T[idx % len + (int64_t)(idx / len) * nbb] = Ts[(idx / len) * kMaxNBB + idx % len];
  // rows o.., columns 0..o-1: zero (T is upper triangular)
  for (int idx = i; idx < len * o; idx += blockDim.x)
    t.Tb[o + idx % len + (int64_t)(idx / len) * nbb] = 0.0;
}

// Block-columns K at the end of each tile that sy_symv reads with ordinary
// loads while it streams the rest: the tiles' tails (K (K + 1) / 2 blocks of
// 32 KB each) stay in the last-level cache from one column to the next.
// Their total is this context's share of kSymvKeepBytes: reductions of
// `sharers` contexts may run at once.  Measured on 64 tiles (n from 1024 to
// 5766) as three concurrent batched solves: all loads ordinary 852 ms per
// solve, a tail of 0 / 40 / 75 / 110 / 150 / 224 MB per solve 825 / 803 /
// 793 / 802 / 819 / 834 ms.
static int symv_keep_blocks(int ntiles, int sharers) {
  const int64_t keep = kSymvKeepBytes / std::max(1, sharers);
  int K = 0;
  while ((int64_t)ntiles * (K + 1) * (K + 2) / 2 * (SB * SB * 8) <= keep) ++K;
  return K;
}

// The pb columns of the panel starting at column k: the first column's norm
// partials, then three launches per column.
static hipError_t sy_panel_columns(hipStream_t st, const SyTile* d_tiles, int ntiles, int64_t nmax,
                                   int64_t k, int pb, int kkeep, Prof* prof) {
  int nnorm = std::max(ceil_div(nmax - k - 2, RB), 1);
  hipLaunchKernelGGL(sy_colnorm, dim3(nnorm, 1, ntiles), dim3(RB), 0, st, d_tiles, (int)k);
  for (int jj = 0; jj < pb; ++jj) {
    const int64_t c = k + jj;
    const int nbl = ceil_div(nmax, SB) - (int)((c + 1) / SB);  // block rows of the trailing matrix
    const int gx = std::max(ceil_div(nbl, SEG), 1);
    // rows of dot-product workgroups: 2jj dots x chunks, 4 wavefronts each
    const int ndot = ceil_div((int64_t)2 * jj * ceil_div(nmax - c - 1, DCH), 4);
    const int gy = nbl + ceil_div(ndot, gx);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (prof) IMCOM_HIP_TRY(prof->events(PS_SYMV, 2, &e0, &e1));
    // nontemporal reads of the block-columns left of the cached tail while
    // the trailing triangle is larger than that tail
    const dim3 grid(gx, gy, ntiles);
    if (nbl > kkeep)
      hipExtLaunchKernelGGL(sy_symv<true>, grid, dim3(256), 0, st, e0, e1, 0, d_tiles, (int)k,
                            (int)c, jj, kkeep);
    else
      hipExtLaunchKernelGGL(sy_symv<false>, grid, dim3(256), 0, st, e0, e1, 0, d_tiles, (int)k,
                            (int)c, jj, 0);
    const int nw = std::max(ceil_div(nmax - c - 1, WR), 1);
    const int nxt = (jj + 1 < pb) ? 1 : 0;
    if (ntiles == 1) {  // one launch: sy_mid's work repeated by every workgroup
      hipLaunchKernelGGL(sy_midwupd<true>, dim3(ceil_div(nw, MID_THREADS / 256)), dim3(MID_THREADS),
                         0, st, d_tiles, (int)k, (int)c, jj, nnorm, gx * gy, nxt, nw);
    } else {
      hipLaunchKernelGGL(sy_mid, dim3(1, 1, ntiles), dim3(MID_THREADS_BATCH), 0, st, d_tiles,
                         (int)k, (int)c, jj, nnorm, gx * gy);
      hipLaunchKernelGGL(sy_wupd<false>, dim3(nw, 1, ntiles), dim3(256), 0, st, d_tiles, (int)k,
                         (int)c, jj, nxt);
    }
    nnorm = nw;  // sy_wupd leaves the next column's norm partials, one per workgroup
  }
  return hipGetLastError();
}

hipError_t sytrd_run(hipStream_t st, const SyTile* d_tiles, int ntiles, int64_t nmax,
                     const std::vector<SyTile>& host, int nb, int sharers, Prof* prof,
                     int tail_rows, double* tail_xch, int64_t* tail_S) {
  size_t tok = 0;
  const int kkeep = symv_keep_blocks(ntiles, sharers);
  const int nbt = ceil_div(nmax, SB);
  // the first panel boundary from which the persistent kernel takes over
  int64_t k_tail = nmax;
  if (ntiles == 1 && tail_xch && tail_rows >= 2 && nmax >= 2) {
    const int rows = std::min(tail_rows, kTailMaxRows);
    k_tail = 0;
    while (nmax - k_tail > rows) k_tail += nb;
  }
  if (tail_S) *tail_S = 0;
  if (k_tail > 0)  // the block layout the blocked columns read (a whole tail reads A)
    hipLaunchKernelGGL(sy_tile_init, dim3(nbt, nbt, ntiles), dim3(256), 0, st, d_tiles);
  for (int64_t k = 0; k < std::min(nmax - 1, k_tail); k += nb) {
    const int pb = (int)std::min<int64_t>(nb, nmax - 1 - k);
    IMCOM_HIP_TRY(sy_panel_columns(st, d_tiles, ntiles, nmax, k, pb, kkeep, prof));
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
        IMCOM_HIP_TRY(dgemm_lower(st, false, true, (int)mm, 2 * nb, -1.0, h.VW + r0, h.n,
                                  h.WV + r0, h.n, 1.0, h.A + r0 + r0 * h.n, h.n, h.At, r0,
                                  h.nbt, nb));
      }
      if (prof) IMCOM_HIP_TRY(prof->end(PS_SYR2K, tok));
    }
  }
  if (k_tail < nmax - 1) {
    IMCOM_HIP_TRY(sytrd_tail(st, host[0], k_tail, k_tail == 0, tail_xch));
    if (tail_S) *tail_S = nmax - k_tail;
  }
  hipLaunchKernelGGL(sy_last_diag, dim3(1, 1, ntiles), dim3(1), 0, st, d_tiles);
  return hipGetLastError();
}

// Q = H(0) ... H(n-2) Z, applied in blocks of NBB reflectors, last block
// first: Z[kb+1:, :] -= Vb (T (Vb^T Z[kb+1:, :])).  Z is n x n (ld n).  For
// NBB = 256, T's two 128 diagonal blocks come from bt_build_T and the block
// above them is T12 = -T1 (V1^T V2) T2 (the DLARFT recursion).
// The joins that complete T from its kMaxNBB diagonal blocks (bt_build_T):
// for widths w = kMaxNBB, 2 kMaxNBB, .. < nbb and each pair of neighbouring
// finished blocks (i, i + w), T(i:i+w, i+w:i+w+w2) = -T1 S12 T2 with
// T1 = T(i:i+w, i:i+w), T2 = T(i+w:.., i+w:..) and S12 = S(i:i+w, i+w:..);
// X = S12 T2 goes to the scratch `x` (ld w).  Two products per join, in
// this order (a join's T2 is finished by the joins of the width before).
struct TJoin {
  int i, w, w2;
};
static std::vector<TJoin> t_joins(int nbb) {
  std::vector<TJoin> v;
  for (int w = kMaxNBB; w < nbb; w *= 2)
    for (int i = 0; i + w < nbb; i += 2 * w) v.push_back(TJoin{i, w, std::min(w, nbb - i - w)});
  return v;
}

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
      // S = Vb^T Vb (K = rows is long and the output small: split K)
      IMCOM_HIP_TRY(dgemm_splitk(st, true, false, nbb, nbb, (int)rows, 1.0, h.Vb, rows, h.Vb, rows,
                                 h.Sb, nbb, h.Ks));
    }
    hipLaunchKernelGGL(bt_build_T, dim3(ceil_div(nbb, kMaxNBB), 1, ntiles), dim3(256), 0, st,
                       d_tiles, kb, nbb);
    for (int b = 0; b < ntiles; ++b) {
      const SyTile& h = host[b];
      if (kb >= h.n - 1) continue;
      const int64_t rows = h.n - kb - 1;
      double* Zr = h.Z + kb + 1;
      for (const TJoin& j : t_joins(nbb)) {  // X in Y2
        const int64_t c0 = j.i + j.w;  // first column of the block above the diagonal
        IMCOM_HIP_TRY(dgemm(st, false, false, j.w, j.w2, j.w2, 1.0, h.Sb + j.i + c0 * nbb, nbb,
                            h.Tb + c0 + c0 * nbb, nbb, 0.0, h.Y2, j.w));
        IMCOM_HIP_TRY(dgemm(st, false, false, j.w, j.w2, j.w, -1.0, h.Tb + j.i + (int64_t)j.i * nbb,
                            nbb, h.Y2, j.w, 0.0, h.Tb + j.i + c0 * nbb, nbb));
      }
      // Y = Vb^T Z_r  (nbb x n; split K)
      IMCOM_HIP_TRY(dgemm_splitk(st, true, false, nbb, (int)h.n, (int)rows, 1.0, h.Vb, rows, Zr,
                                 h.n, h.Y, nbb, h.Ks));
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
                           h.A + r0 + r0 * h.n, h.n, 1, h.At, r0, h.nbt, nb});
    }
    plan.syr2k.push_back(plan.gp.add(false, true, g));
  }
  if (nmax > 1) {
    const int64_t nblk = (nmax - 1 + nbb - 1) / nbb;
    for (int64_t bi = nblk - 1; bi >= 0; --bi) {
      const int kb = (int)(bi * nbb);
      std::vector<GemmDesc> gs, gy, gy2, gz;
      const std::vector<TJoin> joins = t_joins(nbb);
      std::vector<std::vector<GemmDesc>> gj(2 * joins.size());
      for (const SyTile& h : host) {
        if (kb >= h.n - 1) continue;
        const int rows = (int)(h.n - kb - 1);
        const int n = (int)h.n;
        double* Zr = h.Z + kb + 1;
        gs.push_back(GemmDesc{nbb, nbb, rows, 1.0, 0.0, h.Vb, rows, h.Vb, rows, h.Sb, nbb});
        gy.push_back(GemmDesc{nbb, n, rows, 1.0, 0.0, h.Vb, rows, Zr, h.n, h.Y, nbb});
        gy2.push_back(GemmDesc{nbb, n, nbb, 1.0, 0.0, h.Tb, nbb, h.Y, nbb, h.Y2, nbb});
        gz.push_back(GemmDesc{rows, n, nbb, -1.0, 1.0, h.Vb, rows, h.Y2, nbb, Zr, h.n});
        for (size_t q = 0; q < joins.size(); ++q) {
          const TJoin& j = joins[q];
          const int64_t c0 = j.i + j.w;
          gj[2 * q].push_back(GemmDesc{j.w, j.w2, j.w2, 1.0, 0.0, h.Sb + j.i + c0 * nbb, nbb,
                                       h.Tb + c0 + c0 * nbb, nbb, h.Y2, j.w});
          gj[2 * q + 1].push_back(GemmDesc{j.w, j.w2, j.w, -1.0, 0.0,
                                           h.Tb + j.i + (int64_t)j.i * nbb, nbb, h.Y2, j.w,
                                           h.Tb + j.i + c0 * nbb, nbb});
        }
      }
      std::vector<int> ids;
      ids.push_back(plan.gp.add(true, false, gs));
      ids.push_back(plan.gp.add(true, false, gy));
      ids.push_back(plan.gp.add(false, false, gy2));
      ids.push_back(plan.gp.add(false, false, gz));
      for (auto& g : gj) ids.push_back(plan.gp.add(false, false, g));
      plan.bt.push_back(ids);
    }
  }
}

hipError_t sytrd_run_batched(hipStream_t st, const SyTile* d_tiles, int ntiles, int64_t nmax,
                             int nb, const SyPlan& plan, int sharers, Prof* prof) {
  size_t tok = 0;
  const int kkeep = symv_keep_blocks(ntiles, sharers);
  const int nbt = ceil_div(nmax, SB);
  hipLaunchKernelGGL(sy_tile_init, dim3(nbt, nbt, ntiles), dim3(256), 0, st, d_tiles);
  int panel = 0;
  for (int64_t k = 0; k < nmax - 1; k += nb, ++panel) {
    const int pb = (int)std::min<int64_t>(nb, nmax - 1 - k);
    IMCOM_HIP_TRY(sy_panel_columns(st, d_tiles, ntiles, nmax, k, pb, kkeep, prof));
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
    const std::vector<int>& ids = plan.bt[idx];
    hipLaunchKernelGGL(bt_gather, dim3(ceil_div(nmax - kb - 1, 256), nbb, ntiles), dim3(256), 0,
                       st, d_tiles, kb, nbb);
    IMCOM_HIP_TRY(plan.gp.run(st, ids[0]));
    hipLaunchKernelGGL(bt_build_T, dim3(ceil_div(nbb, kMaxNBB), 1, ntiles), dim3(256), 0, st,
                       d_tiles, kb, nbb);
    for (size_t q = 4; q < ids.size(); ++q) IMCOM_HIP_TRY(plan.gp.run(st, ids[q]));  // T's joins
    IMCOM_HIP_TRY(plan.gp.run(st, ids[1]));
    IMCOM_HIP_TRY(plan.gp.run(st, ids[2]));
    IMCOM_HIP_TRY(plan.gp.run(st, ids[3]));
  }
  return hipGetLastError();
}

}  // namespace imcom
