// Kernel (b), part 1b: the last columns of the tridiagonal reduction in one
// persistent launch (the SYTD2 stage inside DSYEVR, imcom_matrix.f90:642).
//
// Once the trailing matrix is small (S <= kTailMaxRows columns), each column
// of the blocked reduction is a few microseconds of work behind two or three
// dependent launches.  Here the whole trailing matrix lives in the LDS of G
// workgroups instead (column j of it on workgroup j % G, slot j / G, all S
// rows), and the columns are reduced one by one (unblocked, LAPACK dsytd2's
// lower form) inside one launch, with one grid-wide barrier per column:
//
//   (B) every workgroup forms p_j = tau A22(:, j)^T v for its columns j > c
//       (each column whole on one workgroup, so no cross-workgroup sum) and
//       its share of p^T v; the owner of column c+1 also publishes that
//       column as it stands;
//   barrier
//   (D) every workgroup reads p, the partial sums and column c+1, forms
//       w = p - 1/2 tau (p^T v) v, brings column c+1 up to date and forms its
//       reflector (dlarfg conventions, as the blocked kernels) itself: the
//       same inputs in the same order give every workgroup the same v, tau;
//   (F) every workgroup updates its columns, A22 -= v w^T + w v^T.
//
// Values exchanged between workgroups are written and read with agent-scope
// atomic stores and loads (coherent across the XCDs' L2 caches); a thread
// waits for its stores to complete before the workgroup arrives at the
// barrier (an atomic counter), so the data is in place when the count is
// reached.  Every wait is bounded: a workgroup that waits too long raises an
// abort flag that ends every workgroup's wait, and the host reports the
// flag as an error.  Each workgroup takes 150 KB of LDS, so a CU holds one,
// and G never exceeds the CU count: on an otherwise idle GPU all G are
// resident at once, and work of other streams only delays them.
//
// Outputs are the blocked reduction's: d, e, tau of columns k0.., and each
// reflector's v(2:) in column-major A below the subdiagonal.
#include <hip/hip_runtime.h>

#include "common.h"
#include "kernels.h"
#include "wave.h"

namespace imcom {

namespace {

constexpr int TT = 512;                  // threads per workgroup
constexpr int kWaves = TT / 64;
constexpr int kRPT = (kTailMaxRows + TT - 1) / TT;  // rows per thread (register arrays)
constexpr int kMaxCPW = 16;              // columns per workgroup at most
// ... and at least: the LDS holds kTailLdsDoubles / kTailMaxRows columns of
// the largest trailing matrix
constexpr int kMinCPW = 9;
// LDS for the owned columns: cpw * S doubles
constexpr int kTailLdsDoubles = (152 * 1024) / 8;
static_assert(kRPT * TT >= kTailMaxRows, "rows per thread");
static_assert(kMinCPW * kTailMaxRows <= kTailLdsDoubles, "LDS holds the largest tail");

// Element (i, j) of the symmetric trailing matrix from the lower-triangle
// block layout (the upper triangle of a diagonal block is stale).
__device__ __forceinline__ double at_elem(const double* __restrict__ At, int nbt, int64_t i,
                                          int64_t j) {
  if (i < j) {
    const int64_t s = i;
    i = j;
    j = s;
  }
  return At[tri_block_offset((int)(i >> 6), (int)(j >> 6), nbt) + (j & 63) * 64 + (i & 63)];
}

// Fixed-order workgroup sum (wavefront sums, then the wavefronts in
// order); every thread gets the result.
__device__ __forceinline__ double tail_sum(double v, double* red) {
  v = wave_sum_dpp(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) t += red[w];
  return t;
}

// Grid barrier: arrival count sync[0] (monotonic within a launch), abort
// flag sync[1].  Thread 0 arrives with an agent-scope release (the L2's dirty
// lines are written back first, so the workgroup's stores are visible to every
// XCD) and leaves with an agent-scope acquire (stale L2 lines are dropped);
// the workgroup barriers around it order the other threads' accesses.
__device__ __forceinline__ bool tail_barrier(unsigned* sync, unsigned target, unsigned spin_max,
                                             int* ok_sh) {
  __builtin_amdgcn_s_waitcnt(0);  // every wavefront's stores have completed
  __syncthreads();
  if (threadIdx.x == 0) {
    int ok = 1;
    __hip_atomic_fetch_add(&sync[0], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    unsigned it = 0;
    while (__hip_atomic_load(&sync[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (__hip_atomic_load(&sync[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u ||
          ++it > spin_max) {
        __hip_atomic_store(&sync[1], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    *ok_sh = ok;
  }
  __syncthreads();
  return *ok_sh != 0;
}

// The reflector of relative column cc from its entries x (registers, rows
// il = tid + TT r): d = x(cc), alpha = x(cc+1), x' = x(cc+2..).  Returns v
// (v(cc+1) = 1, v(cc+2..) = x' / (alpha - beta), zero above) in vn, and
// tau, beta and d in rf[0..2].
__device__ __forceinline__ void tail_reflector(int S, int cc, const double (&x)[kRPT],
                                               double (&vn)[kRPT], double* rf, double* red,
                                               double* bc) {
  const int tid = threadIdx.x;
  double s2 = 0.0;
#pragma unroll
  for (int r = 0; r < kRPT; ++r) {
    const int il = tid + TT * r;
    if (il >= cc + 2 && il < S) s2 += x[r] * x[r];
    if (il == cc) bc[0] = x[r];
    if (il == cc + 1) bc[1] = x[r];
  }
  const double xn2 = tail_sum(s2, red);  // its barriers also publish bc
  const double dcc = bc[0], alpha_r = bc[1];
  const double xnorm = sqrt(xn2);
  double tau, beta, scal;
  if (xnorm == 0.0) {
    tau = 0.0;
    beta = alpha_r;
    scal = 0.0;
  } else {
    beta = -copysign(dlapy2_t(alpha_r, xnorm), alpha_r);
    tau = (beta - alpha_r) / beta;
    scal = 1.0 / (alpha_r - beta);
  }
#pragma unroll
  for (int r = 0; r < kRPT; ++r) {
    const int il = tid + TT * r;
    double v = 0.0;
    if (il == cc + 1) v = 1.0;
    else if (il >= cc + 2 && il < S) v = x[r] * scal;
    vn[r] = v;
  }
  rf[0] = tau;
  rf[1] = beta;
  rf[2] = dcc;
}

// The blocked reduction's outputs for relative column cc: d, e, tau and
// v(2:) below the subdiagonal of column-major A.
__device__ __forceinline__ void tail_store(const SyTile& t, int64_t k0, int S, int cc,
                                           const double (&vn)[kRPT], const double* rf) {
  const int64_t C = k0 + cc;
#pragma unroll
  for (int r = 0; r < kRPT; ++r) {
    const int il = threadIdx.x + TT * r;
    if (il >= cc + 2 && il < S) t.A[(k0 + il) + C * t.n] = vn[r];
  }
  if (threadIdx.x == 0) {
    t.d[C] = rf[2];
    t.e[C] = rf[1];
    t.tau[C] = rf[0];
  }
}

}  // namespace

// xch: the exchange slots of tail_xch_doubles; sync: two words, zero at
// launch.  from_A: read the trailing matrix from column-major A (k0 = 0: A
// is whole) instead of the block layout At.  Thread tid holds the rows
// il = tid + TT r (r < kRPT) of every per-row vector for the whole launch, so
// v, w and the next column's v stay in its registers; the LDS holds the
// owned columns only.
template <int CPW>
__global__ __launch_bounds__(TT) void sy_tail_kernel(SyTile t, int64_t k0, int S, int G,
                                                     int from_A, double* __restrict__ xch,
                                                     unsigned* __restrict__ sync,
                                                     unsigned spin_max) {
  constexpr int cpw = CPW;
  extern __shared__ double L[];                 // cpw x S, column-major by slot
  __shared__ double red[kWaves];
  __shared__ double bc[2];
  __shared__ double redp[CPW][kWaves];          // wavefront partials of the owned p_j
  __shared__ double vjs[CPW], wjs[CPW];         // v_j, w_j at the owned columns
  __shared__ double vw1[2];                     // v and w at row c+1
  __shared__ int ok_sh;
  const int g = blockIdx.x, tid = threadIdx.x;
  const int64_t n = t.n;
  // write-once exchange slots, one per column (see tail_xch_doubles)
  double* __restrict__ pbuf = xch;                              // [S][S]: p of column c
  double* __restrict__ cbuf = xch + (size_t)S * S;              // [S][S]: column c+1
  double* __restrict__ qbuf = xch + (size_t)2 * S * S;          // [S][G]: shares of p^T v

  // A count of G or more before the first barrier is a count that was not
  // reset: flag it (2) and stop rather than pass barriers early.
  if (tid == 0) {
    const unsigned c0 = __hip_atomic_load(&sync[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    ok_sh = c0 < (unsigned)G;
    if (!ok_sh) __hip_atomic_store(&sync[1], 2u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  if (!ok_sh) return;
  // the owned columns, and column 0 for the first reflector
  auto src = [&](int il, int jl) -> double {
    const int64_t i = k0 + il, j = k0 + jl;
    if (!from_A) return at_elem(t.At, t.nbt, i, j);
    return i >= j ? t.A[i + j * n] : t.A[j + i * n];  // the lower triangle, as the blocked path
  };
  for (int s = 0; s < cpw; ++s) {  // slots past the last column hold zeros
    const int jl = g + G * s;
    for (int il = tid; il < S; il += TT) L[s * S + il] = jl < S ? src(il, jl) : 0.0;
  }
  double x[kRPT], v[kRPT], w[kRPT], vn[kRPT], pr[kRPT];
#pragma unroll
  for (int r = 0; r < kRPT; ++r) {
    const int il = tid + TT * r;
    x[r] = il < S ? src(il, 0) : 0.0;
  }
  // column 0's reflector; workgroup 0 stores it after the first barrier, when
  // every workgroup has read its columns (from A, whose column 0 it
  // overwrites)
  double rf[3];
  tail_reflector(S, 0, x, v, rf, red, bc);
  double tau = rf[0];
  const double rf0[3] = {rf[0], rf[1], rf[2]};
  // slot of the owned column at each of this thread's rows (-1: not owned),
  // fixed for the launch; owned_rows copies a per-row vector's values at the
  // owned columns into LDS
  int own_s[kRPT];
#pragma unroll
  for (int r = 0; r < kRPT; ++r) {
    const int il = tid + TT * r;
    own_s[r] = (il < S && il >= g && (il - g) % G == 0) ? (il - g) / G : -1;
  }
  auto owned_rows = [&](const double (&a)[kRPT], double* dst) {
#pragma unroll
    for (int r = 0; r < kRPT; ++r)
      if (own_s[r] >= 0) dst[own_s[r]] = a[r];
  };
  owned_rows(v, vjs);
  double part[CPW];
#pragma unroll
  for (int s = 0; s < CPW; ++s) part[s] = 0.0;
#pragma unroll
  for (int r = 0; r < kRPT; ++r) {
    const int il = tid + TT * r;
    if (il > 0 && il < S) {
#pragma unroll
      for (int s = 0; s < CPW; ++s) part[s] += L[s * S + il] * v[r];
    }
  }

  for (int c = 0; c < S - 1; ++c) {
    double* __restrict__ pc = pbuf + (size_t)c * S;
    double* __restrict__ cc1 = cbuf + (size_t)c * S;
    double* __restrict__ qc = qbuf + (size_t)c * G;
    // (B) p_j = tau A22(:, j)^T v for the owned columns j > c (one barrier
    // for all their sums), their share of p^T v, and column c+1
#pragma unroll
    for (int s = 0; s < CPW; ++s) {
      const double ws = wave_sum_dpp(part[s]);
      if ((tid & 63) == 0) redp[s][tid >> 6] = ws;
    }
    __syncthreads();
    // p_j by lanes 0..cpw-1 of wavefront 0, and its share of p^T v by that
    // wavefront's sum (the other lanes add zeros)
    if (tid < 64) {
      double q = 0.0;
      if (tid < cpw) {
        const int jl = g + G * tid;
        if (jl > c && jl < S) {
          double tot = 0.0;
#pragma unroll
          for (int w = 0; w < kWaves; ++w) tot += redp[tid][w];
          const double p = tau * tot;
          st_coh(&pc[jl], p);
          q = p * vjs[tid];
        }
      }
      const double qs = wave_sum_dpp(q);
      if (tid == 0) st_coh(&qc[g], qs);
    }
    if ((c + 1) % G == g) {  // the owner of column c+1 publishes it
      const int s1 = (c + 1) / G;
#pragma unroll
      for (int r = 0; r < kRPT; ++r) {
        const int il = tid + TT * r;
        if (il > c && il < S) st_coh(&cc1[il], L[s1 * S + il]);
      }
    }
    if (!tail_barrier(sync, (unsigned)(c + 1) * (unsigned)G, spin_max, &ok_sh)) return;
    if (c == 0 && g == 0) tail_store(t, k0, S, 0, v, rf0);

    // (D) w = p - 1/2 tau (p^T v) v and column c+1; every load issued before
    // the first is used (one round trip)
    double qv = 0.0;
    for (int qi = tid; qi < G; qi += TT) qv += ld_coh(&qc[qi]);
#pragma unroll
    for (int r = 0; r < kRPT; ++r) {
      const int il = tid + TT * r;
      const bool on = il > c && il < S;
      pr[r] = on ? ld_coh(&pc[il]) : 0.0;
      x[r] = on ? ld_coh(&cc1[il]) : 0.0;
    }
    const double ptv = tail_sum(qv, red);
    const double alpha = -0.5 * tau * ptv;
#pragma unroll
    for (int r = 0; r < kRPT; ++r) {
      const int il = tid + TT * r;
      w[r] = (il > c && il < S) ? pr[r] + alpha * v[r] : 0.0;
      if (il == c + 1) {
        vw1[0] = v[r];
        vw1[1] = w[r];
      }
    }
    owned_rows(w, wjs);
    __syncthreads();
    // (E) column c+1 up to date, then its reflector (or the last diagonal)
    {
      const double v1 = vw1[0], w1 = vw1[1];
#pragma unroll
      for (int r = 0; r < kRPT; ++r) {
        const int il = tid + TT * r;
        if (il > c && il < S) x[r] -= v[r] * w1 + w[r] * v1;
      }
    }
    double taun = 0.0;
    const bool writer = ((c + 1) % G) == g;
    if (c + 1 < S - 1) {
      tail_reflector(S, c + 1, x, vn, rf, red, bc);
      taun = rf[0];
      if (writer) tail_store(t, k0, S, c + 1, vn, rf);
    } else {
#pragma unroll
      for (int r = 0; r < kRPT; ++r) {
        vn[r] = 0.0;
        if (writer && tid + TT * r == c + 1) t.d[k0 + c + 1] = x[r];
      }
    }
    // (F) the owned columns j > c+1: A22 -= v w^T + w v^T over rows > c, and
    // in the same pass their partial sums against the next v (columns
    // j <= c+1 take v_j = w_j = 0: their entries stay unchanged)
    double vj[CPW], wj[CPW];
#pragma unroll
    for (int s = 0; s < CPW; ++s) {
      const int jl = g + G * s;
      const bool on = jl > c + 1 && jl < S;
      vj[s] = on ? vjs[s] : 0.0;
      wj[s] = on ? wjs[s] : 0.0;
      part[s] = 0.0;
    }
    // a row's reads are all issued before its first update (the LDS
    // latency is paid once per row, not once per column)
#pragma unroll
    for (int r = 0; r < kRPT; ++r) {
      const int il = tid + TT * r;
      if (il > c && il < S) {
        double l[CPW];
#pragma unroll
        for (int s = 0; s < CPW; ++s) l[s] = L[s * S + il];
#pragma unroll
        for (int s = 0; s < CPW; ++s) {
          l[s] -= v[r] * wj[s] + w[r] * vj[s];
          part[s] += l[s] * vn[r];
          L[s * S + il] = l[s];
        }
      }
    }
#pragma unroll
    for (int r = 0; r < kRPT; ++r) v[r] = vn[r];
    tau = taun;
    __syncthreads();  // vjs was read above
    owned_rows(v, vjs);
  }
}

// The kernel instantiated for cpw (kMinCPW..kMaxCPW columns per workgroup).
template <int C>
static hipError_t tail_launch(int cpw, dim3 grid, size_t lds, hipStream_t st, const SyTile& t,
                              int64_t k0, int S, int G, int from_A, double* xch, unsigned* sync,
                              unsigned spin) {
  if constexpr (C <= kMaxCPW) {
    if (cpw == C) {
      IMCOM_HIP_TRY(hipFuncSetAttribute((const void*)sy_tail_kernel<C>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      hipLaunchKernelGGL(sy_tail_kernel<C>, grid, dim3(TT), lds, st, t, k0, S, G, from_A, xch,
                         sync, spin);
      return hipGetLastError();
    }
    return tail_launch<C + 1>(cpw, grid, lds, st, t, k0, S, G, from_A, xch, sync, spin);
  } else {
    return hipErrorInvalidValue;
  }
}

// The barrier's words are reset with the agent-scope atomic stores the
// barrier itself uses.
__global__ void sy_tail_zero(unsigned* sync) {
  if (threadIdx.x < 2)
    __hip_atomic_store(&sync[threadIdx.x], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Workgroups and columns per workgroup for a trailing size S: the fewest
// workgroups (cheapest barrier) whose LDS holds cpw columns each.
void tail_shape(int S, int* G, int* cpw) {
  int c = kTailLdsDoubles / S;
  c = c < kMaxCPW ? c : kMaxCPW;
  c = c > kMinCPW ? c : kMinCPW;  // small trailing sizes: more slots than columns
  *cpw = c;
  *G = (S + c - 1) / c;
}

// Exchange space: p, column c+1 and the shares of p^T v of every column in
// slots of their own.  A slot is written once (write-through) before the
// barrier that publishes it and read after it, so no XCD's L2 can hold an
// older copy of the line (slots reused every other column were measured to
// return stale values now and then).  Then the two barrier words.
size_t tail_xch_doubles(int S) {
  int G, cpw;
  tail_shape(S, &G, &cpw);
  return (size_t)2 * S * S + (size_t)S * G + 2;
}

hipError_t sytrd_tail(hipStream_t st, const SyTile& t, int64_t k0, bool from_A, double* xch) {
  const int S = (int)(t.n - k0);
  if (S < 2 || S > kTailMaxRows) return hipErrorInvalidValue;
  int G, cpw;
  tail_shape(S, &G, &cpw);
  const size_t lds = (size_t)cpw * S * sizeof(double);
  unsigned* sync = reinterpret_cast<unsigned*>(xch + tail_xch_doubles(S) - 2);

  hipLaunchKernelGGL(sy_tail_zero, dim3(1), dim3(64), 0, st, sync);
  IMCOM_HIP_TRY(hipGetLastError());
  int dev = 0, ncu = 0;
  IMCOM_HIP_TRY(hipGetDevice(&dev));
  IMCOM_HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  if (G > ncu) return hipErrorInvalidConfiguration;  // one workgroup per CU at most
  const unsigned spin = 1u << 22;  // about 4 s of polling: a bound, never reached in a healthy run
  return tail_launch<kMinCPW>(cpw, dim3(G), lds, st, t, k0, S, G, from_A ? 1 : 0, xch, sync,
                              spin);
}

// The abort flag of the last tail launch (1: a wait ran out, 2: the barrier
// count was not reset).
hipError_t sytrd_tail_status(hipStream_t st, int64_t S, const double* xch, unsigned* flag) {
  int G, cpw;
  tail_shape((int)S, &G, &cpw);
  const unsigned* sync = reinterpret_cast<const unsigned*>(xch + tail_xch_doubles((int)S) - 2);
  IMCOM_HIP_TRY(hipMemcpyAsync(flag, sync + 1, sizeof(unsigned), hipMemcpyDeviceToHost, st));
  return hipStreamSynchronize(st);
}

}  // namespace imcom
