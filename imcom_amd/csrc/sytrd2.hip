// Kernel (b), two-stage form: the reduction to tridiagonal form inside
// DSYEVR (imcom_matrix.f90:642) as dense -> band -> tridiagonal, with the
// back-transform of the eigenvectors through both stages.
//
// The one-stage reduction streams the whole trailing matrix once per column
// (a matrix-vector product, bound by HBM).  Here the first stage reduces A
// to a band of width kBand = 64 with blocked Householder QR panels whose
// trailing updates are FP64 matrix products (A22 := Q^T A22 Q with
// Q = I - V T V^T); the second stage chases the band down to tridiagonal
// form with reflectors of length kBand that touch only a 64 x 128 window
// next to the diagonal at a time.
//
// Stage 1, per panel of 64 columns k0.. (r0 = k0 + 64 the first row below
// the band):
//   * one persistent launch QR-factors the panel A(r0:n, k0:k0+64): its rows
//     are split among G workgroups (up to 224 rows each, held in LDS), and
//     per column c every workgroup publishes the sums
//     P(c+1:, c)^T P(c+1:, :) over its rows, and the pivot row's owner
//     publishes row c, as entries tagged with the column; once it holds all
//     of them every workgroup forms the same reflector (dlarfg conventions)
//     and w = tau v^T P(c:, :) from these totals and updates its rows.  Workgroup 0 also builds the panel's
//     triangular factor T (LAPACK dlarft, forward, columnwise) from the same
//     sums: V(:, 0:c)^T v_c needs only the pivot row and the published sums.
//   * the panel's R goes to the band (zeros below it), V^T to the upper
//     triangle rows k0..k0+63 (never read as matrix entries again), and V,
//     padded with zeros to 64 columns, to U = [V W V];
//   * X = A22 V, Y = X T, M = T^T (V^T Y), W = Y - 1/2 V M and
//     A22 -= [V W] [W V]^T (both triangles of A22, so the next product can
//     read it as a plain matrix).
//
// Stage 2 (bulge chasing): sweep s annihilates column s below the
// subdiagonal with a reflector of rows s+1..s+64, applied on both sides to
// the 64 x 64 diagonal block; that creates fill in the block below it, whose
// first column the next reflector (rows s+65..) annihilates, and so on down
// the band (Lang's sequential algorithm; the rest of each fill triangle is
// annihilated by the following sweeps).  Task (s, k) touches rows
// r = s+1+64k .. r+63 and columns r-64 .. r+63 only; the task (s, k) may
// start once (s, k-1) and (s-1, k+1) are done (or all of sweep s-1 when it
// has fewer tasks), and any two tasks not ordered by these rules touch
// disjoint entries.  One persistent launch with one workgroup per task
// index; the band is read from column-major A once, and each workgroup's
// part of it stays in its LDS (below).
//
// Back-transform: the eigenvectors Z of the tridiagonal matrix become
// Q1 Q2 Z.  Q2 = H(0,0) H(0,1) ... (creation order) is applied in blocks of
// 32 consecutive sweeps at one task index k: such a block's reflectors are
// staggered by one row (rows s0+1+64k .. s0+95+64k), and applying the blocks
// in the order (sweep group descending, k ascending) with each block as one
// compact-WY factor gives the product in a valid order (reflectors of one
// sweep touch disjoint rows; a later sweep's reflector with a higher task
// index lies wholly below an earlier sweep's with a lower one).  Q1 is the
// product of the panels' block reflectors, applied last panel first.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "common.h"
#include "kernels.h"
#include "wave.h"

namespace imcom {

namespace {

constexpr int B = kBand;       // band width
constexpr int PT = 256;        // threads per workgroup
constexpr int kPanRows = 224;  // panel rows per workgroup at most
constexpr int LDP = B + 1;     // LDS row stride of the panel rows
constexpr unsigned kSpinMax = 1u << 22;

// ---------------------------------------------------------------- stage 1

// A message entry: two agent-scope atomic words, each holding a tag (the
// sweep, or the column) in its upper half and one half of the value in its
// lower half (both
// words single-copy atomic, so an entry whose two tags match is whole).  The
// consumer polls its entries until their tags match: no flag, no wait for
// the producer's stores to complete and no second read.
__device__ __forceinline__ void msg_put(unsigned long long* e, unsigned tag, double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned long long t = (unsigned long long)tag << 32;
  __hip_atomic_store(e, t | (b & 0xffffffffull), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(e + 1, t | (b >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Waits for the entry with the given tag; false after an abort or spin_max
// polls (then raises the abort flag).
__device__ __forceinline__ bool msg_get(const unsigned long long* e, unsigned tag,
                                        unsigned* abort_flag, unsigned spin_max, double* v) {
  for (unsigned it = 0;; ++it) {
    const unsigned long long lo = __hip_atomic_load(e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long hi =
        __hip_atomic_load(e + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((unsigned)(lo >> 32) == tag && (unsigned)(hi >> 32) == tag) {
      *v = __longlong_as_double((long long)((hi << 32) | (lo & 0xffffffffull)));
      return true;
    }
    if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u ||
        it >= spin_max) {
      __hip_atomic_store(abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

// Entries h0 + 4i (i < 16, those below G) of a row of entries B apart,
// loaded together and re-polled until their tags match; x[i] = 0 past G.
__device__ __forceinline__ bool msg_get16(const unsigned long long* src, int h0, int G,
                                          unsigned tag, unsigned* abort_flag, unsigned spin_max,
                                          double* x) {
  unsigned long long lo[16], hi[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const unsigned long long* e = src + (size_t)min(h0 + 4 * i, G - 1) * (2 * B);
    lo[i] = __hip_atomic_load(e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    hi[i] = __hip_atomic_load(e + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  unsigned pend = 0;
#pragma unroll
  for (int i = 0; i < 16; ++i)
    if (h0 + 4 * i < G && ((unsigned)(lo[i] >> 32) != tag || (unsigned)(hi[i] >> 32) != tag))
      pend |= 1u << i;
  for (unsigned it = 0; pend != 0u; ++it) {
    if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u ||
        it >= spin_max) {
      __hip_atomic_store(abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
#pragma unroll
    for (int i = 0; i < 16; ++i)
      if ((pend >> i) & 1u) {
        const unsigned long long* e = src + (size_t)(h0 + 4 * i) * (2 * B);
        lo[i] = __hip_atomic_load(e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        hi[i] = __hip_atomic_load(e + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((unsigned)(lo[i] >> 32) == tag && (unsigned)(hi[i] >> 32) == tag) pend &= ~(1u << i);
      }
  }
#pragma unroll
  for (int i = 0; i < 16; ++i)
    x[i] = h0 + 4 * i < G ? __longlong_as_double((long long)((hi[i] << 32) | (lo[i] & 0xffffffffull)))
                          : 0.0;
  return true;
}

__global__ __launch_bounds__(PT) void band_panel_kernel(double* __restrict__ A, int64_t n,
                                                        int64_t k0, int bc, int nref, int rpw,
                                                        int G, double* __restrict__ U,
                                                        int64_t ldu, double* __restrict__ T1,
                                                        unsigned long long* __restrict__ xch,
                                                        unsigned* __restrict__ sync,
                                                        unsigned base, unsigned spin_max) {
  extern __shared__ double P[];  // owned rows, row-major with stride LDP
  __shared__ double part[4][B];
  __shared__ double sv[B], pr[B], wv[B], yv[B];
  __shared__ double vl[kPanRows];
  __shared__ double Tl[B][B + 1];  // the triangular factor (workgroup 0)
  __shared__ int ok_sh;
  const int tid = threadIdx.x, g = blockIdx.x;
  const int jj = tid & 63, q = tid >> 6;
  const int64_t r0 = k0 + B;
  const int m = (int)(n - r0);
  const int i0 = g * rpw;
  const int nr = min(rpw, m - i0);
  // tagged entries (two words each), each written once per band reduction
  unsigned long long* part_x = xch;                             // [nref][G][B]
  unsigned long long* piv_x = xch + (size_t)nref * G * B * 2;  // [nref][B]

  for (int j = 0; j < B; ++j)
    for (int i = tid; i < nr; i += PT)
      P[i * LDP + j] = j < bc ? A[(r0 + i0 + i) + (k0 + j) * n] : 0.0;
  if (tid == 0) ok_sh = 1;
  __syncthreads();

  bool ok = true;
  for (int c = 0; c < nref; ++c) {
    // (a) this workgroup's share of P(c+1:, c)^T P(c+1:, j), every j: rows
    // i = q mod 4 from the first below the pivot, four independent chains
    {
      int i = max(c + 1 - i0, 0);
      i += (q - i) & 3;
      double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
      for (; i + 12 < nr; i += 16) {
        s0 += P[i * LDP + c] * P[i * LDP + jj];
        s1 += P[(i + 4) * LDP + c] * P[(i + 4) * LDP + jj];
        s2 += P[(i + 8) * LDP + c] * P[(i + 8) * LDP + jj];
        s3 += P[(i + 12) * LDP + c] * P[(i + 12) * LDP + jj];
      }
      for (; i < nr; i += 4) s0 += P[i * LDP + c] * P[i * LDP + jj];
      part[q][jj] = (s0 + s1) + (s2 + s3);
    }
    __syncthreads();
    if (G == 1) {  // one workgroup: the totals are its own
      if (tid < B) {
        sv[tid] = (part[0][tid] + part[1][tid]) + (part[2][tid] + part[3][tid]);
        pr[tid] = P[c * LDP + tid];
      }
      __syncthreads();
    } else {
      double tp = 0.0;
      if (tid < B) tp = (part[0][tid] + part[1][tid]) + (part[2][tid] + part[3][tid]);
      __syncthreads();  // part is rewritten below
      const unsigned tag = (unsigned)(k0 + c + 1);  // this column's, unique in the reduction
      if (tid < B) {
        msg_put(&part_x[(((size_t)c * G + g) * B + tid) * 2], tag, tp);
        if (c >= i0 && c < i0 + nr) msg_put(&piv_x[((size_t)c * B + tid) * 2], tag, P[(c - i0) * LDP + tid]);
      }
      // (b) totals in a fixed order (the same in every workgroup), each
      // workgroup's entry polled until its tag is this column's, sixteen in
      // flight at a time (h = q, q + 4, ...)
      bool okt = true;
      {
        const unsigned long long* src = part_x + ((size_t)c * G * B + jj) * 2;
        double t0 = 0.0, t1 = 0.0;
        for (int h0 = q; h0 < G && okt; h0 += 64) {
          double x[16];
          okt = msg_get16(src, h0, G, tag, &sync[1], spin_max, x);
#pragma unroll
          for (int i = 0; i < 16; i += 2) {
            t0 += x[i];
            t1 += x[i + 1];
          }
        }
        part[q][jj] = t0 + t1;
      }
      if (tid < B && okt) {
        double v;
        okt = msg_get(&piv_x[((size_t)c * B + tid) * 2], tag, &sync[1], spin_max, &v);
        pr[tid] = v;
      }
      if (!okt) ok_sh = 0;
      __syncthreads();
      if (!ok_sh) {
        ok = false;
        break;
      }
      if (tid < B) sv[tid] = (part[0][tid] + part[1][tid]) + (part[2][tid] + part[3][tid]);
      __syncthreads();
    }
    const double alpha = pr[c];
    const double xnorm = sqrt(sv[c]);
    double tau, beta, sc;
    if (xnorm == 0.0) {
      tau = 0.0;
      beta = alpha;
      sc = 0.0;
    } else {
      beta = -copysign(dlapy2_t(alpha, xnorm), alpha);
      tau = (beta - alpha) / beta;
      sc = 1.0 / (alpha - beta);
    }
    if (tid < B) {
      // w_j = tau v^T P(c:, j); V(:, j)^T v for the columns before c
      wv[tid] = (tid > c && tid < bc) ? tau * (pr[tid] + sc * sv[tid]) : 0.0;
      yv[tid] = tid < c ? pr[tid] + sc * sv[tid] : 0.0;
    }
    for (int i = tid; i < nr; i += PT) vl[i] = P[i * LDP + c] * sc;
    __syncthreads();
    if (g == 0 && tid < B) {  // T(0:c, c) = -tau T(0:c, 0:c) V(:, 0:c)^T v, T(c, c) = tau
      double t = 0.0;
      if (tid < c) {
        for (int l = tid; l < c; ++l) t += Tl[tid][l] * yv[l];
        t *= -tau;
      } else if (tid == c) {
        t = tau;
      }
      Tl[tid][c] = t;
    }
    // (c) the owned rows: P(i, j) -= v_i w_j below the pivot, the pivot row
    // less w, column c becomes v (beta on the pivot row)
    if (jj >= c) {
      const double wj = wv[jj];
      int i = max(c + 1 - i0, 0);
      i += (q - i) & 3;
      if (jj > c) {  // four rows at a time, their reads issued together
        for (; i + 12 < nr; i += 16) {
          const double p0 = P[i * LDP + jj], p1 = P[(i + 4) * LDP + jj];
          const double p2 = P[(i + 8) * LDP + jj], p3 = P[(i + 12) * LDP + jj];
          const double v0 = vl[i], v1 = vl[i + 4], v2 = vl[i + 8], v3 = vl[i + 12];
          P[i * LDP + jj] = p0 - v0 * wj;
          P[(i + 4) * LDP + jj] = p1 - v1 * wj;
          P[(i + 8) * LDP + jj] = p2 - v2 * wj;
          P[(i + 12) * LDP + jj] = p3 - v3 * wj;
        }
        for (; i < nr; i += 4) P[i * LDP + jj] -= vl[i] * wj;
      } else {
        for (; i < nr; i += 4) P[i * LDP + jj] = vl[i];
      }
      const int ip = c - i0;  // the pivot row
      if (ip >= 0 && ip < nr && (ip & 3) == q) P[ip * LDP + jj] = jj == c ? beta : P[ip * LDP + jj] - wj;
    }
    __syncthreads();
  }
  if (!ok) return;

  // R into the band (zeros below it)
  for (int j = 0; j < bc; ++j)
    for (int i = tid; i < nr; i += PT) {
      const int ig = i0 + i;
      A[(r0 + ig) + (k0 + j) * n] = ig <= j ? P[i * LDP + j] : 0.0;
    }
  // V (unit diagonal, zeros above, zero columns past nref) into U's first
  // and third blocks of columns
  for (int j = 0; j < B; ++j)
    for (int i = tid; i < nr; i += PT) {
      const int ig = i0 + i;
      double v = 0.0;
      if (j < nref) v = ig == j ? 1.0 : (ig > j ? P[i * LDP + j] : 0.0);
      U[ig + j * ldu] = v;
      U[ig + (2 * B + j) * ldu] = v;
    }
  // V^T into the upper triangle: row k0 + j, column r0 + ig
  if (jj < bc)
    for (int i = q; i < nr; i += 4) {
      const int ig = i0 + i;
      double v = 0.0;
      if (jj < nref) v = ig == jj ? 1.0 : (ig > jj ? P[i * LDP + jj] : 0.0);
      A[(k0 + jj) + (r0 + ig) * n] = v;
    }
  if (g == 0)
    for (int idx = tid; idx < B * B; idx += PT) {
      const int j = idx / B, i = idx % B;
      T1[idx] = (i <= j && j < nref) ? Tl[i][j] : 0.0;
    }
}

__global__ void msg_zero_kernel(unsigned* w, size_t nw) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nw) w[i] = 0u;
}

__global__ void band_zero_kernel(unsigned* w, int nw) {
  for (int i = threadIdx.x; i < nw; i += blockDim.x)
    __hip_atomic_store(&w[i], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// workgroups and rows per workgroup for m panel rows
void panel_shape(int m, int* G, int* rpw) {
  const int g = (m + kPanRows - 1) / kPanRows;
  *G = g;
  *rpw = (m + g - 1) / g;
}

// ---------------------------------------------------------------- stage 2

// Workgroup k owns task index k of every sweep and keeps its window (rows
// r..r+63, columns r-64..r+63, r = s+1+64k) in LDS as a circular array,
// both triangles of the diagonal block: sliding from sweep s to s+1 drops
// row r (sent to workgroup k-1, whose window it enters as the bottom row)
// and takes row r+64 from workgroup k+1 (its top row after task (s, k+1));
// the entries of the new row left of column r are zero (annihilated
// columns).  Task (s, k) needs only the reflector of (s, k-1), sent by
// workgroup k-1.  Messages go through rings of 4 slots per workgroup (a
// producer is at most one sweep ahead of its consumers); each entry of a
// message carries the sweep's tag, and its consumer polls it until the tag
// matches.  Workgroup 0's top rows are the
// tridiagonal matrix: d and e come out as the window slides.
constexpr int WLD = 65;  // LDS stride of a window column slot (64 row slots + 1)
constexpr int BT = 512;  // threads per bulge-chasing workgroup: row ri = tid / 8, eighth tid % 8

// Sum over each group of eight consecutive lanes (every lane of the group
// holds it; partner sums commutative, so the order is fixed).
__device__ __forceinline__ double oct_sum_dpp(double v) {
  v += dpp_d<0xb1>(v);   // quad_perm [1, 0, 3, 2]
  v += dpp_d<0x4e>(v);   // quad_perm [2, 3, 0, 1]
  v += dpp_d<0x141>(v);  // row_half_mirror: the other quad of the eight
  return v;
}

__global__ __launch_bounds__(BT, 1) void bulge_kernel(const double* __restrict__ A, int64_t n,
                                                      double* __restrict__ V2,
                                                      double* __restrict__ d,
                                                      double* __restrict__ e,
                                                      unsigned long long* __restrict__ msg,
                                                      unsigned* __restrict__ flags,
                                                      unsigned spin_max) {
  __shared__ double Wl[2 * B * WLD];
  __shared__ double vp[B], v[B], y[B], z[B], w[B], sc[2];
  __shared__ int ok_sh;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int ri = tid >> 3, q8 = tid & 7;
  const int k = blockIdx.x, K = gridDim.x;
  const int64_t s_last = min<int64_t>(n - 3, n - 2 - (int64_t)B * k);
  unsigned* abort_flag = flags + 2 * K;
  constexpr int ME = 2 * (B + 1);                  // words per message: B + 1 entries
  unsigned long long* vmsg = msg;                  // [K][4][ME]: v, tau
  unsigned long long* rmsg = msg + (size_t)K * 4 * ME;  // [K][4][ME]: row r, columns r-64..r
  auto W = [&](int64_t i, int64_t j) -> double& {
    return Wl[(int)(j & (2 * B - 1)) * WLD + (int)(i & (B - 1))];
  };
  // the window of sweep 0 (entries outside the matrix are zero; the upper
  // triangle of the diagonal block mirrored from the lower)
  {
    const int64_t r = 1 + (int64_t)B * k;
    for (int idx = tid; idx < 2 * B * B; idx += BT) {
      const int jj = idx >> 6, ii = idx & 63;
      const int64_t i = r + ii, j = r - B + jj;
      double val = 0.0;
      if (i < n && j >= 0 && j < n) val = j <= i ? A[i + j * n] : A[j + i * n];
      W(i, j) = val;
    }
    if (k == 0 && tid == 0) d[0] = A[0];
    if (tid == 0) ok_sh = 1;
  }
  __syncthreads();
  for (int64_t s = 0; s <= s_last; ++s) {
    const int64_t r = s + 1 + (int64_t)B * k;
    const int L = (int)min<int64_t>(B, n - r);
    const int64_t rn = r + B - 1;  // the window's bottom row
    if (s > 0) {  // slide: row r-1 and column r-65 become row rn and column rn
      if (tid < 2 * B) W(rn, r - B + tid) = 0.0;
      if (tid < B) W(r + tid, rn) = 0.0;
      __syncthreads();
    }
    const bool need_row = s > 0 && rn < n;
    if (need_row || k > 0) {
      // the row of (s-1, k+1) and the reflector of (s, k-1), each entry
      // polled by the thread that stores it
      if (need_row && tid <= B) {  // row rn, columns rn-64 .. rn (mirrored into column rn)
        double val;
        if (msg_get(&rmsg[((size_t)(k + 1) * 4 + ((s - 1) & 3)) * ME + 2 * tid], (unsigned)s,
                    abort_flag, spin_max, &val)) {
          const int64_t c = rn - B + tid;
          W(rn, c) = val;
          if (c >= r && c < rn) W(c, rn) = val;
        } else {
          ok_sh = 0;
        }
      }
      if (k > 0 && tid >= 2 * B && tid <= 3 * B) {
        double val;
        if (msg_get(&vmsg[((size_t)(k - 1) * 4 + (s & 3)) * ME + 2 * (tid - 2 * B)],
                    (unsigned)(s + 1), abort_flag, spin_max, &val)) {
          if (tid < 3 * B)
            vp[tid - 2 * B] = val;
          else
            sc[0] = val;  // tau of (s, k-1)
        } else {
          ok_sh = 0;
        }
      }
      __syncthreads();
      if (!ok_sh) return;
    }
    const bool send_v = k + 1 < K;
    unsigned long long* vslot = vmsg + ((size_t)k * 4 + (s & 3)) * ME;
    if (k == 0) {  // the reflector of column s = r-1; published at once
      if (wv == 0) {
        const double x = lane < L ? W(r + lane, r - 1) : 0.0;
        const double s2 = wave_sum_dpp(lane >= 1 ? x * x : 0.0);
        const double alpha = readlane_d(x, 0);
        double t = 0.0, bt = alpha, scal = 0.0;
        if (L >= 2 && s2 != 0.0) {
          bt = -copysign(dlapy2_t(alpha, sqrt(s2)), alpha);
          t = (bt - alpha) / bt;
          scal = 1.0 / (alpha - bt);
        }
        const double vl = lane == 0 ? 1.0 : (lane < L ? x * scal : 0.0);
        v[lane] = vl;
        if (lane < L) W(r + lane, r - 1) = lane == 0 ? bt : 0.0;
        if (lane == 0) {
          sc[0] = t;
          sc[1] = bt;
        }
        if (send_v) {
          msg_put(&vslot[2 * lane], (unsigned)(s + 1), vl);
          if (lane == 0) msg_put(&vslot[2 * B], (unsigned)(s + 1), t);
        }
      }
      __syncthreads();
    } else {
      {  // E := E (I - taup vp vp^T)
        const double taup = sc[0];
        double t = 0.0;
#pragma unroll
        for (int j = 8 * q8; j < 8 * q8 + 8; ++j) t += W(r + ri, r - B + j) * vp[j];
        t = oct_sum_dpp(t) * taup;
#pragma unroll
        for (int j = 8 * q8; j < 8 * q8 + 8; ++j) W(r + ri, r - B + j) -= t * vp[j];
      }
      __syncthreads();
      if (wv == 0) {  // the reflector of E(:, 0) (rows past L are zero); published at once
        const double x = W(r + lane, r - B);
        const double s2 = wave_sum_dpp(lane >= 1 ? x * x : 0.0);
        const double alpha = readlane_d(x, 0);
        double t = 0.0, bt = alpha, scal = 0.0;
        if (L >= 2 && s2 != 0.0) {
          bt = -copysign(dlapy2_t(alpha, sqrt(s2)), alpha);
          t = (bt - alpha) / bt;
          scal = 1.0 / (alpha - bt);
        }
        const double vl = lane == 0 ? 1.0 : (lane < L ? x * scal : 0.0);
        v[lane] = vl;
        if (lane == 0) {
          sc[0] = t;
          sc[1] = bt;
        }
        if (send_v) {
          msg_put(&vslot[2 * lane], (unsigned)(s + 1), vl);
          if (lane == 0) msg_put(&vslot[2 * B], (unsigned)(s + 1), t);
        }
      }
      __syncthreads();
    }
    const double tau = sc[0], beta = sc[1];
    // column sums of E against v (k > 0) and p = tau D v
    if (k > 0) {
      double t = 0.0;
#pragma unroll
      for (int i = 8 * q8; i < 8 * q8 + 8; ++i) t += v[i] * W(r + i, r - B + ri);
      t = oct_sum_dpp(t);
      if (q8 == 0) z[ri] = t;
    }
    {
      double t = 0.0;
#pragma unroll
      for (int j = 8 * q8; j < 8 * q8 + 8; ++j) t += W(r + ri, r + j) * v[j];
      t = oct_sum_dpp(t) * tau;
      if (q8 == 0) y[ri] = t;
    }
    __syncthreads();
    if (k > 0) {  // E(:, 1:) := (I - tau v v^T) E(:, 1:); column 0 becomes beta e1
      const double tv = tau * v[ri];
#pragma unroll
      for (int j = 8 * q8; j < 8 * q8 + 8; ++j) {
        double& x = W(r + ri, r - B + j);
        x = j == 0 ? (ri == 0 ? beta : 0.0) : x - tv * z[j];
      }
    }
    if (wv == 0) {  // w = p - 1/2 tau (p^T v) v
      const double pv = wave_sum_dpp(y[lane] * v[lane]);
      w[lane] = y[lane] - 0.5 * tau * pv * v[lane];
    }
    __syncthreads();
    // row r is final once D(0, 0) is: send it to workgroup k-1 before the
    // rest of the D update
    if (k > 0 && wv == 0) {
      unsigned long long* rslot = rmsg + ((size_t)k * 4 + (s & 3)) * ME;
      msg_put(&rslot[2 * lane], (unsigned)(s + 1), W(r, r - B + lane));
      if (lane == 0) msg_put(&rslot[2 * B], (unsigned)(s + 1), W(r, r) - (v[0] * w[0] + w[0] * v[0]));
    }
    {  // D -= v w^T + w v^T (both triangles; v, w vanish past L)
      const double vi = v[ri], wi = w[ri];
#pragma unroll
      for (int j = 8 * q8; j < 8 * q8 + 8; ++j) W(r + ri, r + j) -= vi * w[j] + wi * v[j];
    }
    __syncthreads();
    // outputs: the reflector and (workgroup 0) d, e
    if (tid < L) V2[s * n + (int64_t)k * B + tid] = tid == 0 ? tau : v[tid];
    if (k == 0 && tid == 0) {
      e[s] = W(r, r - 1);
      d[s + 1] = W(r, r);
    }
  }
  if (k == 0 && tid == 0) {
    e[n - 2] = W(n - 1, n - 2);
    d[n - 1] = W(n - 1, n - 1);
  }
}

// ------------------------------------------------- bulge back-transform

constexpr int QL = kQ2Group;       // reflectors per block (consecutive sweeps)
constexpr int QR = 96;             // block rows (B + QL - 1 = 95, padded)
constexpr int LV = QL + 1;         // LDS row stride of V
// A block's image, the apply kernel's LDS copy of it: V (QR x QL, unit
// diagonal, zero outside the reflectors) in rows of stride LV, then
// Y = -V T (QR x QL: (I - V T V^T) Z = Z + Y (V^T Z)) in rows of stride
// LV, zero padding.
constexpr int QIMG = kQ2Image;
constexpr int QVT = QR * LV;       // offset of Y in an image
constexpr int QPL = QIMG / 512;    // image loads of two doubles per thread of the apply kernel
static_assert(QIMG % 512 == 0 && QVT + QR * LV <= QIMG, "Q2 block image layout");
typedef double v2d __attribute__((ext_vector_type(2)));

typedef double v4d __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int q2_ntask(int64_t n, int64_t s) { return (int)((n - 2 - s) / B) + 1; }

// V of block (S, k) into Vs (QR x QL, zero outside the reflectors) and
// tau into tau_s: column j is reflector (s0 + j, k), rows j .. j + L - 1.
__device__ __forceinline__ void q2_load_V(const double* __restrict__ V2, int64_t n, int64_t s0,
                                          int k, double* Vs, double* tau_s) {
  const int64_t nsw = n - 2;
  for (int idx = threadIdx.x; idx < QR * QL; idx += blockDim.x) Vs[(idx / QL) * LV + idx % QL] = 0.0;
  __syncthreads();
  const int64_t rb = s0 + 1 + (int64_t)k * B;
  for (int idx = threadIdx.x; idx < QL * B; idx += blockDim.x) {
    const int j = idx / B, i = idx % B;
    const int64_t s = s0 + j;
    double v = 0.0;
    bool have = s < nsw && k < q2_ntask(n, s);
    int L = 0;
    if (have) {
      L = (int)min<int64_t>(B, n - (rb + j));
      if (i < L) v = V2[s * n + (int64_t)k * B + i];
    }
    if (i == 0) {
      tau_s[j] = have ? v : 0.0;
      v = have ? 1.0 : 0.0;
    }
    if (have && i < L) Vs[(j + i) * LV + j] = v;
  }
  __syncthreads();
}

// T of every block (LAPACK dlarft, forward, columnwise): one workgroup per
// block, blocks numbered group by group (blk_off[S] + k).
__global__ __launch_bounds__(256) void q2_tfactor_kernel(const double* __restrict__ V2, int64_t n,
                                                         const int* __restrict__ blk_off, int nS,
                                                         double* __restrict__ img) {
  __shared__ double Vs[QR * LV];
  __shared__ double tau_s[QL];
  __shared__ double Gm[QL][QL + 1];
  __shared__ double Ts[QL][QL + 1];
  const int S = blockIdx.y, k = blockIdx.x;
  if (S >= nS || k >= blk_off[S + 1] - blk_off[S]) return;
  const int64_t s0 = (int64_t)S * QL;
  q2_load_V(V2, n, s0, k, Vs, tau_s);
  const int tid = threadIdx.x;
  // Gram matrix of V's columns (upper triangle)
  for (int idx = tid; idx < QL * QL; idx += blockDim.x) {
    const int i = idx / QL, j = idx % QL;
    double g = 0.0;
    if (i < j)
      for (int r = j; r < min(QR, i + B); ++r) g += Vs[r * LV + i] * Vs[r * LV + j];
    Gm[i][j] = g;
  }
  __syncthreads();
  for (int j = 0; j < QL; ++j) {
    if (tid < QL) {
      double t = 0.0;
      if (tid < j) {
        for (int l = tid; l < j; ++l) t += Ts[tid][l] * Gm[l][j];
        t *= -tau_s[j];
      } else if (tid == j) {
        t = tau_s[j];
      }
      Ts[tid][j] = t;
    }
    __syncthreads();
  }
  // the block's image: V as the apply kernel's LDS holds it, then
  // Y = -V T (T upper triangular: Y(r, j) sums over l <= j)
  double* out = img + ((size_t)blk_off[S] + k) * QIMG;
  for (int idx = tid; idx < QIMG; idx += blockDim.x) {
    double v = 0.0;
    if (idx < QVT) {
      if (idx % LV < QL) v = Vs[idx];
    } else if (idx < QVT + QR * LV) {
      const int r = (idx - QVT) / LV, c = (idx - QVT) % LV;
      if (c < QL) {
        double y = 0.0;
        for (int l = 0; l <= c; ++l) y += Vs[r * LV + l] * Ts[l][c];
        v = -y;
      }
    }
    out[idx] = v;
  }
}

// Z := Q2 Z on Z stored by rows (Zt = Z^T column-major: Z(i, c) at
// Zt[c + i n]), 64 columns of Z per workgroup: blocks (S descending, k
// ascending), each Z(rows, :) += Y (V^T Z(rows, :)) with Y = -V T on FP64
// MFMA.  Wavefront w owns the 16-column stripe 16w.. and keeps its 96-row
// window of Z in registers in the 16x16x4 MFMA's operand layout: lane l
// holds rows 4s + l/16 (s = 0..23) of column l%16, which is both the B
// operand of K step s and the accumulator layout of row tile s/4.  So
// W = V^T Z reads Z from registers, Z + Y W reads W from the accumulators
// that produced it, and the update of Z accumulates into Z itself; only
// the block images (V and Y, 52 KB, double-buffered) go through the LDS,
// with one barrier per block, and no part of Z moves between wavefronts.
// The last QL - 1 rows of a block's window are the first of the next
// (k + 1): the window shifts by B rows in each lane's registers and only the
// new rows are read.  The next block's image and rows are loaded at the
// start of a block, every load unconditional, and consumed at its end.
// With Z stored by rows, one load or store of a wavefront covers four rows
// of 16 consecutive entries: four whole 128-byte lines (stored by columns
// it would touch 16 lines, and the vector cache's line rate would bound the
// kernel).
// Z is read and written through a buffer resource based at the block's
// first row and the workgroup's first column, whose range ends with the
// matrix: rows past the matrix read 0.  A load or store that must not
// happen (columns past n - 1, rows not stored) is given an offset from
// kQ2Drop on (past the range, whose size is capped at it), so every store
// is issued without a branch: a store under a lane mask is not counted as
// issued by the compiler's wait bookkeeping, which then waits for all
// memory operations, stores included, before using any later load.
constexpr int QS = 16;       // columns per wavefront
constexpr int QCW = 4 * QS;  // columns per workgroup
constexpr unsigned kQ2Drop = 0x80000000u;  // window offsets stay below it for n < 2 million
typedef unsigned int v2u __attribute__((ext_vector_type(2)));

__device__ __forceinline__ double q2_ldz(__amdgpu_buffer_rsrc_t r, unsigned off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}
__device__ __forceinline__ void q2_stz(__amdgpu_buffer_rsrc_t r, unsigned off, double v) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2u, v), r, (int)off, 0, 0);
}
// Buffer resource for Zt from row rb, column cw on to the end of the
// matrix; base and size from scalar registers (a resource the compiler
// cannot show uniform is used one lane value at a time).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t q2_zrsrc(double* Zt, int64_t n, int64_t rb,
                                                          int64_t cw) {
  const uint64_t bytes = ((uint64_t)(n - rb) * (uint64_t)n - (uint64_t)cw) * 8u;
  const uint64_t p = (uint64_t)(Zt + rb * n + cw);
  const uint64_t pu = ((uint64_t)__builtin_amdgcn_readfirstlane((unsigned)(p >> 32)) << 32) |
                      (unsigned)__builtin_amdgcn_readfirstlane((unsigned)p);
  const unsigned nb = bytes < kQ2Drop ? (unsigned)bytes : kQ2Drop;
  return __builtin_amdgcn_make_buffer_rsrc((void*)pu, 0, (int)__builtin_amdgcn_readfirstlane(nb),
                                           0x00020000);
}

__global__ __launch_bounds__(256, 1) void q2_apply_kernel(const double* __restrict__ img, int64_t n,
                                                          const int* __restrict__ blk_off, int nS,
                                                          double* __restrict__ Zt) {
  __shared__ double Ls[2][QIMG];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int64_t cw = (int64_t)blockIdx.x * QCW;
  const bool colok = cw + QS * wv + lr < n;
  if (nS <= 0) return;
  // byte offsets: the lane's column and window row lk from the resource's
  // base (kQ2Drop on for a column past the matrix), four rows per step s
  const unsigned zl = colok ? (unsigned)(((int64_t)lk * n + QS * wv + lr) * 8) : kQ2Drop;
  const unsigned row4 = (unsigned)(32 * n);
  v4d zt[QR / 16];  // window row 4s + lk in zt[s / 4][s % 4]
  // the window of block (S, 0)
  auto load_window = [&](int S_) {
    const __amdgpu_buffer_rsrc_t zr = q2_zrsrc(Zt, n, (int64_t)S_ * QL + 1, cw);
#pragma unroll
    for (int s = 0; s < QR / 4; ++s) zt[s / 4][s % 4] = q2_ldz(zr, zl + s * row4);
  };
  // thread tid moves image entries 2 tid + 512 u, +1 (16 bytes)
  const v2d* ib = reinterpret_cast<const v2d*>(img) + tid;
  int off = blk_off[nS - 1];
  {
    v2d ri[QPL];
    const v2d* p = ib + (size_t)off * (QIMG / 2);
#pragma unroll
    for (int u = 0; u < QPL; ++u) ri[u] = p[256 * u];
    load_window(nS - 1);
    v2d* L0 = reinterpret_cast<v2d*>(Ls[0]) + tid;
#pragma unroll
    for (int u = 0; u < QPL; ++u) L0[256 * u] = ri[u];
  }
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0);  // no load left in flight into the loop
  __syncthreads();
  int S = nS - 1, k = 0, buf = 0;
  while (true) {
    const int64_t rb = (int64_t)S * QL + 1 + (int64_t)k * B;
    const int R = (int)min<int64_t>(QR - 1, n - rb);
    const int nk = q2_ntask(n, (int64_t)S * QL);
    const bool last = k + 1 == nk;
    const int nS2 = last ? S - 1 : S, nk2 = last ? 0 : k + 1;
    const int off2 = last ? off - (S > 0 ? q2_ntask(n, (int64_t)(S - 1) * QL) : 0) : off;
    const __amdgpu_buffer_rsrc_t zr = q2_zrsrc(Zt, n, rb, cw);
    // prefetch: the next block's image (this block's again after the last)
    // and the next window's new rows 31..95, B rows further down (4s + lk
    // for s = 7..23; rows 28..30 of s = 7 are carried; read but not used
    // after the last block of a group)
    v2d ri[QPL];
    double zn[QR / 4 - 7];
    {
      const v2d* p = ib + (size_t)(nS2 >= 0 ? off2 + nk2 : off + k) * (QIMG / 2);
#pragma unroll
      for (int u = 0; u < QPL; ++u) ri[u] = p[256 * u];
#pragma unroll
      for (int s = 7; s < QR / 4; ++s) zn[s - 7] = q2_ldz(zr, zl + (s + B / 4) * row4);
    }
    const double* Lv = Ls[buf];
    const double* Ly = Ls[buf] + QVT;
    // W = V^T Z: row tile 0 over window rows 0..79, tile 1 over rows 16..95
    // (V's columns 16 c .. 16 c + 15 have entries in rows 16 c .. 16 c + 78)
    v4d w0 = (v4d){0.0, 0.0, 0.0, 0.0}, w1 = w0;
#pragma unroll
    for (int s = 0; s < QR / 4; ++s) {
      const double bz = zt[s / 4][s % 4];
      if (s < 20)
        w0 = __builtin_amdgcn_mfma_f64_16x16x4f64(Lv[(4 * s + lk) * LV + lr], bz, w0, 0, 0, 0);
      if (s >= 4)
        w1 = __builtin_amdgcn_mfma_f64_16x16x4f64(Lv[(4 * s + lk) * LV + 16 + lr], bz, w1, 0, 0, 0);
    }
    // the prefetch loads spread over the products of W, one after each
    // (issued together they queue in the vector memory unit while the
    // wavefront waits to issue them)
#pragma unroll
    for (int i = 0; i < QPL + QR / 4 - 7; ++i) {
      __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // one vector memory read
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // one MFMA
    }
    // Z += Y W (Y's rows 80-95 have no entries before column 17), tile by
    // tile; rows 0..63 are final after their tile (no later block of the
    // group touches them) and go back to Z at once, under the next tiles'
    // products (rows past the block's rows, nout on, are not stored)
    const int nout = (last ? R : min(B, R)) - lk;
#pragma unroll
    for (int t = 0; t < QR / 16; ++t) {
#pragma unroll
      for (int q = 0; q < QL / 4; ++q) {
        if (t == QR / 16 - 1 && q < 4) continue;
        const double bw = q < 4 ? w0[q] : w1[q - 4];
        zt[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(Ly[(16 * t + lr) * LV + 4 * q + lk], bw, zt[t],
                                                     0, 0, 0);
      }
      if (t < B / 16) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          q2_stz(zr, (4 * (4 * t + r) < nout ? zl : kQ2Drop) + (4 * t + r) * row4, zt[t][r]);
      }
    }
    // within a group the window shifts (rows 64.. to the top, the
    // prefetched rows below them); after a group's last block its other
    // rows go back to Z and the next group's window is read
    v2d* Ln = reinterpret_cast<v2d*>(Ls[buf ^ 1]) + tid;
    if (!last) {
      zt[0] = zt[4];
      zt[1] = zt[5];
      if (lk == 3) zt[1][3] = zn[0];
#pragma unroll
      for (int s = 8; s < QR / 4; ++s) zt[s / 4][s % 4] = zn[s - 7];
#pragma unroll
      for (int u = 0; u < QPL; ++u) Ln[256 * u] = ri[u];
      // the new window in registers within this block (left to the next,
      // its moves would wait there for this block's stores too)
#pragma unroll
      for (int t = 0; t < QR / 16; ++t) asm volatile("" : "+v"(zt[t]));
    } else {
#pragma unroll
      for (int s = B / 4; s < QR / 4; ++s)
        q2_stz(zr, (4 * s < nout ? zl : kQ2Drop) + s * row4, zt[s / 4][s % 4]);
      if (nS2 < 0) break;
      // the next group's window holds rows this wavefront just stored
      asm volatile("" ::: "memory");
      __builtin_amdgcn_s_waitcnt(0);
      load_window(nS2);
#pragma unroll
      for (int u = 0; u < QPL; ++u) Ln[256 * u] = ri[u];
      asm volatile("" ::: "memory");
      __builtin_amdgcn_s_waitcnt(0);  // no load left in flight across the loop's back edge
    }
    lds_sync();  // double buffering: the other image was last read before the previous barrier
    buf ^= 1;
    S = nS2;
    k = nk2;
    off = off2;
  }
}

// dst = src^T, n x n column-major, 64 x 64 tiles through the LDS
__global__ __launch_bounds__(256) void transpose_kernel(int64_t n, const double* __restrict__ src,
                                                        double* __restrict__ dst) {
  __shared__ double t[64][65];
  const int64_t i0 = (int64_t)blockIdx.x * 64, j0 = (int64_t)blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int jj = ty + 4 * r;
    if (i0 + tx < n && j0 + jj < n) t[jj][tx] = src[(i0 + tx) + (j0 + jj) * n];
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int ii = ty + 4 * r;
    if (j0 + tx < n && i0 + ii < n) dst[(j0 + tx) + (i0 + ii) * n] = t[tx][ii];
  }
}

}  // namespace

size_t band_xch_doubles(int64_t n) {
  int G, rpw;
  panel_shape((int)std::max<int64_t>(n - B, 1), &G, &rpw);
  return ((size_t)B * G * B + (size_t)B * B) * 2;  // tagged entries: two words each
}

size_t band_kwork_doubles(int64_t n) {
  size_t mx = 8;
  for (int64_t k0 = 0; n - (k0 + B) >= 2; k0 += B) {
    const int m = (int)(n - k0 - B);
    mx = std::max(mx, dgemm_splitk_work(m, B, m));
    mx = std::max(mx, dgemm_splitk_work(B, B, m));
  }
  // the band back-transform, in groups of 4 panels (fewer in the last)
  const int np = band_panels(n);
  for (int p0 = 0; p0 < np; p0 += 4) {
    const int nbg = std::min(4, np - p0) * B;
    const int m = (int)(n - (int64_t)p0 * B - B);
    mx = std::max(mx, dgemm_splitk_work(nbg, nbg, m));
    mx = std::max(mx, dgemm_splitk_work(nbg, (int)n, m));
  }
  return mx / 8 + 1;
}

int band_panels(int64_t n) {
  int np = 0;
  for (int64_t k0 = 0; n - (k0 + B) >= 2; k0 += B) ++np;
  return np;
}

hipError_t band_reduce(hipStream_t st, double* A, int64_t n, const TwoStageBufs& w) {
  hipLaunchKernelGGL(band_zero_kernel, dim3(1), dim3(64), 0, st, w.sync, 2);
  IMCOM_HIP_TRY(hipGetLastError());
  // tags from an earlier reduction must not match: clear the exchange
  const size_t xw = band_xch_doubles(n) * 2;
  hipLaunchKernelGGL(msg_zero_kernel, dim3((unsigned)((xw + 255) / 256)), dim3(256), 0, st,
                     reinterpret_cast<unsigned*>(w.xch), xw);
  IMCOM_HIP_TRY(hipGetLastError());
  static bool attr_set = false;
  if (!attr_set) {
    IMCOM_HIP_TRY(hipFuncSetAttribute((const void*)band_panel_kernel,
                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                      kPanRows * LDP * 8));
    attr_set = true;
  }
  unsigned base = 0;
  const int64_t ldu = n;
  int p = 0;
  for (int64_t k0 = 0;; k0 += B, ++p) {
    const int64_t r0 = k0 + B;
    const int m = (int)(n - r0);
    if (m < 2) break;
    const int bc = (int)std::min<int64_t>(B, n - k0);
    const int nref = std::min(bc, m - 1);
    int G, rpw;
    panel_shape(m, &G, &rpw);
    double* T1p = w.T1 + (size_t)p * B * B;
    hipLaunchKernelGGL(band_panel_kernel, dim3(G), dim3(PT), (size_t)rpw * LDP * 8, st, A, n, k0,
                       bc, nref, rpw, G, w.U, ldu, T1p,
                       reinterpret_cast<unsigned long long*>(w.xch), w.sync, base, kSpinMax);
    IMCOM_HIP_TRY(hipGetLastError());
    base += (unsigned)nref * G;
    double* A22 = A + r0 + r0 * n;
    double* V = w.U;
    double* Wb = w.U + B * ldu;
    // X = A22 V (split K: m / 64 tiles alone leave the CUs short of work)
    IMCOM_HIP_TRY(dgemm_splitk(st, false, false, m, B, m, 1.0, A22, n, V, ldu, w.X, m, w.kwork));
    // Y = X T (into W's place)
    IMCOM_HIP_TRY(dgemm(st, false, false, m, B, B, 1.0, w.X, m, T1p, B, 0.0, Wb, ldu));
    // Mt = V^T Y, M = T^T Mt
    IMCOM_HIP_TRY(dgemm_splitk(st, true, false, B, B, m, 1.0, V, ldu, Wb, ldu, w.Mt, B, w.kwork));
    IMCOM_HIP_TRY(dgemm(st, true, false, B, B, B, 1.0, T1p, B, w.Mt, B, 0.0, w.Mm, B));
    // W = Y - 1/2 V M
    IMCOM_HIP_TRY(dgemm(st, false, false, m, B, B, -0.5, V, ldu, w.Mm, B, 1.0, Wb, ldu));
    // A22 -= [V W] [W V]^T
    IMCOM_HIP_TRY(dgemm(st, false, true, m, m, 2 * B, -1.0, w.U, ldu, Wb, ldu, 1.0, A22, n));
  }
  return hipSuccess;
}

size_t bulge_msg_doubles(int64_t n) {
  const int64_t K = n >= 3 ? (n - 2) / B + 1 : 1;
  return (size_t)2 * K * 4 * 2 * (B + 1);  // 64-bit words: two per entry
}

size_t bulge_flag_words(int64_t n) {
  const int64_t K = n >= 3 ? (n - 2) / B + 1 : 1;
  return (size_t)2 * K + 1;
}

hipError_t bulge_chase(hipStream_t st, double* A, int64_t n, double* V2, double* d, double* e,
                       double* msg, unsigned* flags, Prof* prof) {
  if (n < 3) return hipErrorInvalidValue;
  const int K = (int)((n - 2) / B + 1);
  int dev = 0, ncu = 0;
  IMCOM_HIP_TRY(hipGetDevice(&dev));
  IMCOM_HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  // 70 KB of LDS and eight wavefronts: two workgroups per CU, so K <= 2 x CUs
  // workgroups are all resident on an otherwise idle GPU
  if (K > 2 * ncu) return hipErrorInvalidValue;
  const int nw = (int)bulge_flag_words(n);
  hipLaunchKernelGGL(band_zero_kernel, dim3(1), dim3(256), 0, st, flags, nw);
  IMCOM_HIP_TRY(hipGetLastError());
  // message tags from an earlier chase must not match: clear the rings
  const size_t mw = bulge_msg_doubles(n) * 2;
  hipLaunchKernelGGL(msg_zero_kernel, dim3((unsigned)((mw + 255) / 256)), dim3(256), 0, st,
                     reinterpret_cast<unsigned*>(msg), mw);
  IMCOM_HIP_TRY(hipGetLastError());
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (prof) IMCOM_HIP_TRY(prof->events(PS_BULGE, 2, &e0, &e1));
  hipExtLaunchKernelGGL(bulge_kernel, dim3(K), dim3(BT), 0, st, e0, e1, 0, A, n, V2, d, e,
                        reinterpret_cast<unsigned long long*>(msg), flags, kSpinMax);
  return hipGetLastError();
}

// Abort flags of the panel barriers and of the bulge chase (0: none).
hipError_t twostage_status(hipStream_t st, const TwoStageBufs& w, int64_t n, unsigned* flags) {
  IMCOM_HIP_TRY(
      hipMemcpyAsync(&flags[0], w.sync + 1, 4, hipMemcpyDeviceToHost, st));
  flags[1] = 0;
  if (n >= 3)
    IMCOM_HIP_TRY(hipMemcpyAsync(&flags[1], w.prog + (bulge_flag_words(n) - 1), 4,
                                 hipMemcpyDeviceToHost, st));
  return hipStreamSynchronize(st);
}

namespace {

// Q1 in groups of kQ1Group panels: V^T of a group (nb_g x m, zeros above
// each panel's first row) gathered from A's upper triangle, and the group's
// T built from the panels' T by the block recurrence of compact WY
// (T(0:i, i) = -T(0:i, 0:i) G(0:i, i) T_i, G = V^T V).
constexpr int kQ1Group = 4;
constexpr int QG = kQ1Group * B;  // reflectors per group at most

__global__ void q1_gather_kernel(const double* __restrict__ A, int64_t n, int64_t p0, int nbg,
                                 int64_t r0, int64_t m, double* __restrict__ Vt) {
  const int64_t total = (int64_t)nbg * m;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(idx % nbg);
    const int64_t x = idx / nbg;
    const int off = (c / B) * B;  // the panel's first row within the group
    Vt[idx] = x >= off ? A[(p0 * B + c) + (r0 + x) * n] : 0.0;
  }
}

__global__ void q1_tinit_kernel(const double* __restrict__ T1, int64_t p0, int nbg,
                                double* __restrict__ Tg) {
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < nbg * nbg;
       idx += gridDim.x * blockDim.x) {
    const int i = idx % nbg, j = idx / nbg;
    double v = 0.0;
    if (i / B == j / B) v = T1[(size_t)(p0 + i / B) * B * B + (i % B) + (j % B) * B];
    Tg[i + (size_t)j * nbg] = v;
  }
}

}  // namespace

size_t q1_group_work_doubles(int64_t n) {
  // Vt (QG x n), T and G (QG x QG), X (QG x B)
  return (size_t)QG * n + 2 * (size_t)QG * QG + (size_t)QG * B;
}

// Q := Q1 Q with Q1 the product of the panels' block reflectors (V^T in the
// upper triangle of A, T per panel), last group of panels first.
hipError_t apply_q1(hipStream_t st, const double* A, int64_t n, const TwoStageBufs& w, double* Q,
                    int64_t ncols) {
  const int np = band_panels(n);
  const int ng = (np + kQ1Group - 1) / kQ1Group;
  double* Vt = w.q1work;
  double* Tg = Vt + (size_t)QG * n;
  double* G = Tg + (size_t)QG * QG;
  double* X = G + (size_t)QG * QG;
  for (int g = ng - 1; g >= 0; --g) {
    const int p0 = g * kQ1Group, npan = std::min(kQ1Group, np - p0);
    const int nbg = npan * B;
    const int64_t r0 = (int64_t)p0 * B + B;
    const int m = (int)(n - r0);
    const int64_t total = (int64_t)nbg * m;
    hipLaunchKernelGGL(q1_gather_kernel, dim3((unsigned)std::min<int64_t>((total + 255) / 256, 4096)),
                       dim3(256), 0, st, A, n, (int64_t)p0, nbg, r0, (int64_t)m, Vt);
    IMCOM_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(q1_tinit_kernel, dim3((nbg * nbg + 255) / 256), dim3(256), 0, st, w.T1,
                       (int64_t)p0, nbg, Tg);
    IMCOM_HIP_TRY(hipGetLastError());
    if (npan > 1) {
      // G = V^T V (the blocks above the diagonal are used)
      IMCOM_HIP_TRY(dgemm_splitk(st, false, true, nbg, nbg, m, 1.0, Vt, nbg, Vt, nbg, G, nbg,
                                 w.kwork));
      for (int i = 1; i < npan; ++i) {
        const int c0 = i * B;
        IMCOM_HIP_TRY(dgemm(st, false, false, c0, B, c0, 1.0, Tg, nbg, G + (size_t)c0 * nbg, nbg,
                            0.0, X, c0));
        IMCOM_HIP_TRY(dgemm(st, false, false, c0, B, B, -1.0, X, c0,
                            Tg + c0 + (size_t)c0 * nbg, nbg, 0.0, Tg + (size_t)c0 * nbg, nbg));
      }
    }
    IMCOM_HIP_TRY(
        dgemm_splitk(st, false, false, nbg, (int)ncols, m, 1.0, Vt, nbg, Q + r0, n, w.Y1, nbg,
                     w.kwork));
    IMCOM_HIP_TRY(
        dgemm(st, false, false, nbg, (int)ncols, nbg, 1.0, Tg, nbg, w.Y1, nbg, 0.0, w.Y2, nbg));
    IMCOM_HIP_TRY(
        dgemm(st, true, false, m, (int)ncols, nbg, -1.0, Vt, nbg, w.Y2, nbg, 1.0, Q + r0, n));
  }
  return hipSuccess;
}

}  // namespace imcom

namespace imcom {

int q2_groups(int64_t n) { return (int)((std::max<int64_t>(n - 2, 0) + QL - 1) / QL); }

// Blocks of group S: the task count of its first sweep.  blk_off has
// q2_groups(n) + 1 entries.
void q2_block_offsets(int64_t n, std::vector<int>& blk_off) {
  const int nS = q2_groups(n);
  blk_off.assign(nS + 1, 0);
  for (int S = 0; S < nS; ++S)
    blk_off[S + 1] = blk_off[S] + (int)((n - 2 - (int64_t)S * QL) / B) + 1;
}

hipError_t apply_q2(hipStream_t st, double* V2, int64_t n, const int* d_blk_off, int nS,
                    int maxk, double* img, double* Z, Prof* prof) {
  if (nS <= 0) return hipSuccess;
  if (n >= (int64_t)1 << 21) return hipErrorInvalidValue;  // window offsets below kQ2Drop
  hipLaunchKernelGGL(q2_tfactor_kernel, dim3((unsigned)maxk, (unsigned)nS), dim3(256), 0, st, V2,
                     n, d_blk_off, nS, img);
  IMCOM_HIP_TRY(hipGetLastError());
  // V2 is consumed: its storage holds Z by rows for the apply
  const dim3 tg((unsigned)((n + 63) / 64), (unsigned)((n + 63) / 64));
  hipLaunchKernelGGL(transpose_kernel, tg, dim3(256), 0, st, n, Z, V2);
  IMCOM_HIP_TRY(hipGetLastError());
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (prof) IMCOM_HIP_TRY(prof->events(PS_Q2, 2, &e0, &e1));
  hipExtLaunchKernelGGL(q2_apply_kernel, dim3((unsigned)((n + QCW - 1) / QCW)), dim3(256), 0, st,
                        e0, e1, 0, img, n, d_blk_off, nS, V2);
  IMCOM_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(transpose_kernel, tg, dim3(256), 0, st, n, V2, Z);
  return hipGetLastError();
}

}  // namespace imcom
