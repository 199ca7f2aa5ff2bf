// Kernel (d): per-output-pixel bisection for the Lagrange multiplier kappa,
// and the small per-pixel epilogues around T.
//
// Reference: imcom_bisect_kappa / imcom_bisect_by_U / imcom_bisect_by_S
// (imcom_bisect.f90:78-267) over imcom_U_from_kappa / imcom_S_from_kappa
// (imcom_matrix.f90:904-970).  The decision logic (start at kappa_min for U,
// kappa_max for S; halve dx in log10 kappa; accept mid when f(mid) < 0; the
// two exit tests) is kept step for step.
//
// One workgroup per output pixel a.  Each lane keeps its strided share of
// lambda_i and P_ia (and N_i for S) in registers, so every evaluation of
// U(kappa) or S(kappa) is a register sweep plus one fixed-order block sum;
// P is read from HBM once per pixel.  The block size is chosen so that the
// per-lane share is <= PER.
//
// Optional learned bracket (lo, hi in log10 kappa per pixel): the bracket is
// validated with one evaluation at each end; if it does not hold the root the
// pixel falls back to the full [kappa_min, kappa_max] search.
#include "common.h"
#include "kernels.h"

#include <algorithm>

namespace imcom {

template <int PER>
__global__ void bisect_kernel(int64_t n, int64_t m, const double* __restrict__ L,
                              const double* __restrict__ P, int64_t ldp,
                              const double* __restrict__ Nd, double C_a, double US_max,
                              double tol, double kmin, double kmax, int maxnits, int UnotS,
                              const float* __restrict__ bracket, double* __restrict__ K_out,
                              double* __restrict__ US_out, int32_t* __restrict__ evals_out,
                              int32_t* __restrict__ hits_out) {
  __shared__ double red[16];
  const int64_t a = blockIdx.x;
  if (a >= m) return;
  const int bs = blockDim.x;
  double lam[PER], p2[PER], nd[PER];
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int64_t i = threadIdx.x + (int64_t)r * bs;
    if (i < n) {
      lam[r] = L[i];
      const double p = P[i + a * ldp];
      p2[r] = p * p;
      nd[r] = UnotS ? 0.0 : Nd[i];
    } else {
      lam[r] = 1.0;
      p2[r] = 0.0;
      nd[r] = 0.0;
    }
  }
  const double eps = 1.1102230246251565e-16;  // DLAMCH('E')

  // U(k) = (C - sum (L+2k) P2 / (L+k)^2 / 4) / C ;  S(k) = sum N P2 / (L+k)^2 / 4
  auto feval = [&](double k) -> double {
    double s = 0.0;
    if (UnotS) {
#pragma unroll
      for (int r = 0; r < PER; ++r) {
        const double den = (lam[r] + k) * (lam[r] + k);
        s += (lam[r] + 2.0 * k) * p2[r] / den;
      }
      s = block_sum(s, red);
      return (C_a - s / 4.0) / C_a;
    } else {
#pragma unroll
      for (int r = 0; r < PER; ++r) {
        const double den = (lam[r] + k) * (lam[r] + k);
        s += nd[r] * p2[r] / den;
      }
      s = block_sum(s, red);
      return s / 4.0;
    }
  };

  int evals = 0;
  double k_root, out;
  bool use_full = true;
  // U: start from kappa_min and search upward; S: from kappa_max downward
  const double kstart = UnotS ? kmin : kmax;
  const double kend = UnotS ? kmax : kmin;
  out = feval(kstart);
  ++evals;
  double f = out - US_max;
  if (!(f < 0.0)) {
    k_root = kstart;
  } else {
    double lo = log10(kstart), hi = log10(kend);
    if (bracket != nullptr) {
      // learned bracket [blo, bhi] (log10 kappa), oriented like (lo, hi)
      double blo = (double)bracket[2 * a], bhi = (double)bracket[2 * a + 1];
      if (!UnotS) {
        const double t = blo;
        blo = bhi;
        bhi = t;
      }
      const bool inside = UnotS ? (blo > lo && bhi < hi && blo < bhi)
                                : (blo < lo && bhi > hi && blo > bhi);
      if (inside) {
        const double flo = feval(exp10(blo)) - US_max;
        ++evals;
        if (flo < 0.0) {
          const double fhi = feval(exp10(bhi)) - US_max;
          ++evals;
          if (!(fhi < 0.0)) {
            lo = blo;
            hi = bhi;
            out = flo + US_max;
            use_full = false;
          }
        }
      }
    }
    double root = lo, dx = hi - lo;
    for (int it = 1; it <= maxnits; ++it) {
      dx = 0.5 * dx;
      const double mid = root + dx;
      const double fmid = feval(exp10(mid)) - US_max;
      ++evals;
      if (fmid < 0.0) {
        root = mid;
        out = fmid + US_max;
        if (fabs(fmid) < tol) break;
      }
      if (fabs(fmid) < C_a * eps) break;
    }
    k_root = exp10(root);
  }
  if (threadIdx.x == 0) {
    K_out[a] = k_root;
    US_out[a] = out;
    if (evals_out) evals_out[a] = evals;
    if (hits_out) hits_out[a] = use_full ? 0 : 1;
  }
}

hipError_t launch_bisect(hipStream_t st, int64_t n, int64_t m, const double* L, const double* P,
                         int64_t ldp, const double* Nd, double C_a, double US_max, double tol,
                         double kmin, double kmax, int maxnits, int UnotS, const float* bracket,
                         double* K, double* US, int32_t* evals, int32_t* hits) {
  int bs = 64;
  while (bs < 1024 && (n + bs - 1) / bs > 16) bs *= 2;
  const int64_t per = (n + bs - 1) / bs;
  dim3 grid((unsigned)m), block(bs);
#define IMCOM_BISECT(PV)                                                                     \
  hipLaunchKernelGGL(bisect_kernel<PV>, grid, block, 0, st, n, m, L, P, ldp, Nd, C_a, US_max, \
                     tol, kmin, kmax, maxnits, UnotS, bracket, K, US, evals, hits)
  if (per <= 4)
    IMCOM_BISECT(4);
  else if (per <= 8)
    IMCOM_BISECT(8);
  else if (per <= 16)
    IMCOM_BISECT(16);
  else if (per <= 32)
    IMCOM_BISECT(32);
  else
    return hipErrorInvalidValue;
#undef IMCOM_BISECT
  return hipGetLastError();
}

// W(i,a) = P(i,a) / (L(i) + K(a))  (the right-hand vector of each DGEMV in
// imcom_build_T, imcom_matrix.f90:734)
__global__ void scale_W_kernel(int64_t n, int64_t m, const double* __restrict__ P, int64_t ldp,
                               const double* __restrict__ L, const double* __restrict__ K,
                               double* __restrict__ W, int64_t ldw) {
  const int64_t a = blockIdx.y;
  const double k = K[a];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    W[i + a * ldw] = P[i + a * ldp] / (L[i] + k);
}

hipError_t launch_scale_W(hipStream_t st, int64_t n, int64_t m, const double* P, int64_t ldp,
                          const double* L, const double* K, double* W, int64_t ldw) {
  dim3 grid((unsigned)((n + 255) / 256 < 64 ? (n + 255) / 256 : 64), (unsigned)m);
  hipLaunchKernelGGL(scale_W_kernel, grid, dim3(256), 0, st, n, m, P, ldp, L, K, W, ldw);
  return hipGetLastError();
}

// Per-output-pixel reductions over T's column a, one workgroup per pixel:
//   H_a = sum_i T_ia I_i          (imcom_build_image, imcom_matrix.f90:771)
//   S_a = sum_i T_ia (T_ia N_i)   (imcom_build_S, imcom_matrix.f90:847-862)
__global__ __launch_bounds__(256) void maps_kernel(int64_t n, int64_t m,
                                                   const double* __restrict__ T, int64_t ldt,
                                                   const double* __restrict__ I,
                                                   const double* __restrict__ Nd,
                                                   double* __restrict__ H, double* __restrict__ S) {
  __shared__ double red[4];
  const int64_t a = blockIdx.x;
  double h = 0.0, s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    const double t = T[i + a * ldt];
    if (H) h += t * I[i];
    if (S) s += t * (t * Nd[i]);
  }
  if (H) {
    h = block_sum(h, red);
    if (threadIdx.x == 0) H[a] = h;
  }
  if (S) {
    s = block_sum(s, red);
    if (threadIdx.x == 0) S[a] = s;
  }
}

hipError_t launch_maps(hipStream_t st, int64_t n, int64_t m, const double* T, int64_t ldt,
                       const double* I, const double* Nd, double* H, double* S) {
  hipLaunchKernelGGL(maps_kernel, dim3((unsigned)m), dim3(256), 0, st, n, m, T, ldt, I, Nd, H, S);
  return hipGetLastError();
}

// U_a = U(K_a) for the S-mode run (imcom_build_U, imcom_matrix.f90:795-835)
__global__ __launch_bounds__(256) void U_from_K_kernel(int64_t n, const double* __restrict__ L,
                                                       const double* __restrict__ P, int64_t ldp,
                                                       double C_a, const double* __restrict__ K,
                                                       double* __restrict__ U) {
  __shared__ double red[4];
  const int64_t a = blockIdx.x;
  const double k = K[a];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    const double p = P[i + a * ldp];
    const double den = (L[i] + k) * (L[i] + k);
    s += (L[i] + 2.0 * k) * (p * p) / den;
  }
  s = block_sum(s, red);
  if (threadIdx.x == 0) U[a] = (C_a - s / 4.0) / C_a;
}

hipError_t launch_U_from_K(hipStream_t st, int64_t n, int64_t m, const double* L, const double* P,
                           int64_t ldp, double C_a, const double* K, double* U) {
  hipLaunchKernelGGL(U_from_K_kernel, dim3((unsigned)m), dim3(256), 0, st, n, L, P, ldp, C_a, K, U);
  return hipGetLastError();
}

}  // namespace imcom
