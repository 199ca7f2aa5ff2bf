// Kernel (a): build A (n x n) and B (n x m) by 2-D polynomial interpolation
// of the PSF-correlation lookup tables.
//
// Reference: imcom_lookup_A (imcom_matrix.f90:206-284), imcom_lookup_B
// (:288-357), imcom_interp_lookup (:497-542) and the Neville routines
// imcom_polin2/imcom_polint (imcom_proc.f90:351-429).
//
// The reference evaluates the interpolating polynomial through the clipped
// stencil [floor(arg)-npm, ceil(arg)+npm] with Neville's tableau.  Here the
// same polynomial (same nodes, same argument) is evaluated in Lagrange form:
// per axis, the <=8 weights come from prefix/suffix products of (arg - node)
// and a table of exact reciprocal node-difference products, then the value
// is the separable 2-D weighted sum of the <=8x8 window.  Identical in exact
// arithmetic; rounding differs at the 1e-16 relative level.
//
// Layout: A and B column-major with leading dimension n; the tables are
// column-major (n1, n2) slices, `tstride` doubles apart.
#include "common.h"
#include "kernels.h"

namespace imcom {

// c_invden[cnt-1][k] = 1 / prod_{l != k, l < cnt} (k - l)
__constant__ double c_invden[8][8];

static bool g_invden_ready[64] = {false};

static hipError_t ensure_invden(int dev) {
  if (dev >= 0 && dev < 64 && g_invden_ready[dev]) return hipSuccess;
  double h[8][8];
  for (int cnt = 1; cnt <= 8; ++cnt)
    for (int k = 0; k < 8; ++k) {
      double p = 1.0;
      if (k < cnt) {
        for (int l = 0; l < cnt; ++l)
          if (l != k) p *= (double)(k - l);
        h[cnt - 1][k] = 1.0 / p;
      } else {
        h[cnt - 1][k] = 0.0;
      }
    }
  IMCOM_HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_invden), h, sizeof(h)));
  if (dev >= 0 && dev < 64) g_invden_ready[dev] = true;
  return hipSuccess;
}

// Loads the weight table on `dev` (idempotent; call before any graph capture).
hipError_t interp_init(int dev) { return ensure_invden(dev); }

// Lagrange weights at 1-based table coordinate `arg` for nodes lo..lo+cnt-1.
__device__ __forceinline__ void lagrange8(double arg, int lo, int cnt, double w[8]) {
  double d[8];
#pragma unroll
  for (int l = 0; l < 8; ++l) d[l] = (l < cnt) ? (arg - (double)(lo + l)) : 1.0;
  double pre[8], suf[8];
  pre[0] = 1.0;
#pragma unroll
  for (int l = 1; l < 8; ++l) pre[l] = pre[l - 1] * d[l - 1];
  suf[7] = 1.0;
#pragma unroll
  for (int l = 6; l >= 0; --l) suf[l] = suf[l + 1] * d[l + 1];
#pragma unroll
  for (int k = 0; k < 8; ++k) w[k] = pre[k] * suf[k] * c_invden[cnt - 1][k];
}

// imcom_interp_lookup: table value at offset (xarg, yarg) table pixels from
// the centre pixel (n1/2+1, n2/2+1); 0 outside the table.
__device__ __forceinline__ double interp_lookup(const double* __restrict__ tab, int n1, int n2,
                                                int npm, double xarg, double yarg) {
  const double iarg = (double)(n1 / 2 + 1) + xarg;
  const double jarg = (double)(n2 / 2 + 1) + yarg;
  const double fi = floor(iarg), fj = floor(jarg);
  const double ci = ceil(iarg), cj = ceil(jarg);
  if (!(fi >= 1.0 && fj >= 1.0 && ci <= (double)n1 && cj <= (double)n2)) return 0.0;
  const int imin = max(1, (int)fi - npm), jmin = max(1, (int)fj - npm);
  const int imax = min(n1, (int)ci + npm), jmax = min(n2, (int)cj + npm);
  const int cx = imax - imin + 1, cy = jmax - jmin + 1;
  double wx[8], wy[8];
  lagrange8(iarg, imin, cx, wx);
  lagrange8(jarg, jmin, cy, wy);
  const double* base = tab + (imin - 1) + (int64_t)(jmin - 1) * n1;
  double acc = 0.0;
#pragma unroll
  for (int l = 0; l < 8; ++l) {
    const double* col = base + (int64_t)min(l, cy - 1) * n1;
    double r = 0.0;
#pragma unroll
    for (int k = 0; k < 8; ++k) r += wx[k] * col[min(k, cx - 1)];
    acc += wy[l] * r;
  }
  return acc;
}

// Upper-triangular 64x64 tiles of A; each tile is also written mirrored.
__global__ __launch_bounds__(256) void lookup_A_kernel(
    int64_t n, const double* __restrict__ x, const double* __restrict__ y,
    const int32_t* __restrict__ ex, int psfconst, const double* __restrict__ tabs,
    int64_t tstride, int n1, int n2, int npm, double npad, double sx, double sy,
    double* __restrict__ A) {
  __shared__ double tileS[64][65];
  // blockIdx.x -> (bi, bj) with bi <= bj, row-major over the upper triangle
  const int64_t t = blockIdx.x;
  int64_t bj = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while ((bj + 1) * (bj + 2) / 2 <= t) ++bj;
  while (bj * (bj + 1) / 2 > t) --bj;
  const int64_t bi = t - bj * (bj + 1) / 2;
  const int64_t i0 = bi * 64, j0 = bj * 64;
  const int li = threadIdx.x & 63, lq = threadIdx.x >> 6;
  const int64_t i = i0 + li;
  double xi = 0.0, yi = 0.0;
  int ei = 1;
  if (i < n) {
    xi = x[i];
    yi = y[i];
    ei = ex[i];
  }
#pragma unroll 2
  for (int s = 0; s < 16; ++s) {
    const int lj = lq + 4 * s;
    const int64_t j = j0 + lj;
    double v = 0.0;
    if (i < n && j < n && j >= i) {
      const int ej = ex[j];
      const int ial = psfconst == 1 ? 1 : ei + ej * (ej - 1) / 2;
      const double Delx = npad * (x[j] - xi) / sx;
      const double Dely = npad * (y[j] - yi) / sy;
      v = interp_lookup(tabs + (int64_t)(ial - 1) * tstride, n1, n2, npm, Delx, Dely);
      A[i + j * n] = v;
    }
    tileS[lj][li] = v;
  }
  __syncthreads();
  // mirror: A(j, i) = A(i, j) for j > i, consecutive threads -> consecutive j
#pragma unroll 4
  for (int s = 0; s < 16; ++s) {
    const int lj = threadIdx.x & 63;
    const int li2 = (threadIdx.x >> 6) + 4 * s;
    const int64_t jj = j0 + lj, ii = i0 + li2;
    if (jj < n && ii < n && jj > ii) A[jj + ii * n] = tileS[lj][li2];
  }
}

__global__ __launch_bounds__(256) void lookup_B_kernel(
    int64_t n, int64_t m, const double* __restrict__ x, const double* __restrict__ y,
    const int32_t* __restrict__ ex, const double* __restrict__ Xa, const double* __restrict__ Ya,
    int psfconst, const double* __restrict__ tabs, int64_t tstride, int n1, int n2, int npm,
    double npad, double sx, double sy, double* __restrict__ B) {
  const int64_t i = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
  const int64_t a0 = (int64_t)blockIdx.y * 64;
  const int lq = threadIdx.x >> 6;
  if (i >= n) return;
  const double xi = x[i], yi = y[i];
  const double* tab = tabs + (psfconst == 1 ? 0 : (int64_t)(ex[i] - 1) * tstride);
#pragma unroll 2
  for (int s = 0; s < 16; ++s) {
    const int64_t a = a0 + lq + 4 * s;
    if (a >= m) break;
    const double Delx = npad * (xi - Xa[a]) / sx;
    const double Dely = npad * (yi - Ya[a]) / sy;
    B[i + a * n] = interp_lookup(tab, n1, n2, npm, Delx, Dely);
  }
}

hipError_t launch_lookup_A(hipStream_t st, int dev, int64_t n, const double* x, const double* y,
                           const int32_t* ex, int psfconst, const double* tabs, int64_t tstride,
                           int n1, int n2, int npm, int npad, double sx, double sy, double* A) {
  IMCOM_HIP_TRY(ensure_invden(dev));
  const int64_t T = (n + 63) / 64;
  const int64_t nblk = T * (T + 1) / 2;
  hipLaunchKernelGGL(lookup_A_kernel, dim3((unsigned)nblk), dim3(256), 0, st, n, x, y, ex, psfconst,
                     tabs, tstride, n1, n2, npm, (double)npad, sx, sy, A);
  return hipGetLastError();
}

hipError_t launch_lookup_B(hipStream_t st, int dev, int64_t n, int64_t m, const double* x,
                           const double* y, const int32_t* ex, const double* Xa, const double* Ya,
                           int psfconst, const double* tabs, int64_t tstride, int n1, int n2,
                           int npm, int npad, double sx, double sy, double* B) {
  IMCOM_HIP_TRY(ensure_invden(dev));
  dim3 grid((unsigned)((n + 63) / 64), (unsigned)((m + 63) / 64));
  hipLaunchKernelGGL(lookup_B_kernel, grid, dim3(256), 0, st, n, m, x, y, ex, Xa, Ya, psfconst, tabs,
                     tstride, n1, n2, npm, (double)npad, sx, sy, B);
  return hipGetLastError();
}

// ---- regular grids (USERXY=0): one interpolation per distinct separation ----
//
// With x_i = kx * gimxscale(e) + dither(e, 1) (imcom_build_xy,
// imcom_proc.f90:56-89) and X_a = ax * outxscale + outpos1 (imcom_build_XaYa,
// imcom_proc.f90:93-119), the arguments of imcom_lookup_A and imcom_lookup_B
// (imcom_matrix.f90:230-231, 310-311) take few distinct values: for A, the
// separation of two pixels depends only on their exposures and grid offsets,
// (2 n1gim - 1) (2 n2gim - 1) values per exposure pair when the scales are
// equal; for B, on the exposure and (P kx - Q ax) when the input and output
// pitches are in the ratio P/Q.  Each such value is interpolated once into a
// lattice (lattice_eval_kernel), and A and B are written from the lattices
// with coalesced stores: the matrices' cost becomes their write bandwidth.
// The lattice argument (u g + doff) equals the reference's (x_j - x_i) in
// exact arithmetic; rounded, it differs by a few units in the last place of
// the coordinates.

__global__ __launch_bounds__(256) void grid_check_kernel(
    int64_t n, int64_t m, const double* __restrict__ x, const double* __restrict__ y,
    const int32_t* __restrict__ ex, const int32_t* __restrict__ mask_subs, int n1gim, int n2gim,
    int nexp, const double* __restrict__ gsx, const double* __restrict__ gsy,
    const double* __restrict__ dith, int n1out, double osx, double osy, double op1, double op2,
    const double* __restrict__ Xa, const double* __restrict__ Ya, int32_t* __restrict__ kx,
    int32_t* __restrict__ ky, int32_t* __restrict__ bad) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  bool ok = true;
  auto near = [](double a, double b) { return fabs(a - b) <= 1e-12 * (1.0 + fabs(b)); };
  if (t < n) {
    const int64_t sub = (int64_t)mask_subs[t] - 1;
    const int64_t plane = (int64_t)n1gim * n2gim;
    if (sub < 0 || sub >= plane * nexp) {
      ok = false;
      kx[t] = ky[t] = 0;
    } else {
      const int ix = (int)(sub % n1gim), iy = (int)((sub / n1gim) % n2gim);
      const int e = (int)(sub / plane);  // 0-based
      kx[t] = ix;
      ky[t] = iy;
      ok = (ex[t] == e + 1) && near((double)ix * gsx[e] + dith[e], x[t]) &&
           near((double)iy * gsy[e] + dith[nexp + e], y[t]);
    }
  } else if (t < n + m) {
    const int64_t a = t - n;
    const int ax = (int)(a % n1out), ay = (int)(a / n1out);
    ok = near((double)ax * osx + op1, Xa[a]) && near((double)ay * osy + op2, Ya[a]);
  }
  if (!ok) atomicAdd(bad, 1);
}

__global__ __launch_bounds__(256) void lattice_eval_kernel(
    const LatDesc* __restrict__ desc, int npairs, int64_t total, int psfconst,
    const double* __restrict__ tabs, int64_t tstride, int n1, int n2, int npm, double npad,
    double sx, double sy, double* __restrict__ vals) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= total) return;
  int p = 0;
  while (p + 1 < npairs && desc[p + 1].off <= t) ++p;
  const LatDesc d = desc[p];
  const int64_t r = t - d.off;
  const int ux = (int)(r % d.x.L), uy = (int)(r / d.x.L);
  const double Dx = (double)(ux - d.x.P * (d.x.Nr - 1)) * d.x.g + d.x.doff;
  const double Dy = (double)(uy - d.y.P * (d.y.Nr - 1)) * d.y.g + d.y.doff;
  const double Delx = npad * Dx / sx;
  const double Dely = npad * Dy / sy;
  const int tab = psfconst == 1 ? 0 : d.table;
  vals[t] = interp_lookup(tabs + (int64_t)tab * tstride, n1, n2, npm, Delx, Dely);
}

constexpr int kLatCols = 16;  // matrix columns per workgroup in the expansions

// Pair (e_lo, e_hi) of A is descriptor (e_lo - 1) + e_hi (e_hi - 1) / 2, the
// reference's table index ial less one.
__global__ __launch_bounds__(256) void lattice_A_kernel(int64_t n, const int32_t* __restrict__ kx,
                                                        const int32_t* __restrict__ ky,
                                                        const int32_t* __restrict__ ex,
                                                        const LatDesc* __restrict__ desc,
                                                        const double* __restrict__ vals,
                                                        double* __restrict__ A) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t j0 = (int64_t)blockIdx.y * kLatCols;
  if (i >= n) return;
  const int kxi = kx[i], kyi = ky[i], ei = ex[i];
  const int64_t jn = min(n, j0 + kLatCols);
  for (int64_t j = j0; j < jn; ++j) {
    const int kxj = kx[j], kyj = ky[j], ej = ex[j];
    const bool up = i <= j;  // pixel i is the subtracted one
    const int elo = up ? ei : ej, ehi = up ? ej : ei;
    const int xlo = up ? kxi : kxj, xhi = up ? kxj : kxi;
    const int ylo = up ? kyi : kyj, yhi = up ? kyj : kyi;
    const LatDesc* d = desc + (elo - 1) + ehi * (ehi - 1) / 2;
    const int ux = d->x.Q * xhi - d->x.P * xlo + d->x.P * (d->x.Nr - 1);
    const int uy = d->y.Q * yhi - d->y.P * ylo + d->y.P * (d->y.Nr - 1);
    A[i + j * n] = vals[d->off + ux + (int64_t)uy * d->x.L];
  }
}

// B(i, a): the input pixel's grid is the added one, the output grid the
// subtracted one (x_i - X_a); descriptor e - 1.
__global__ __launch_bounds__(256) void lattice_B_kernel(int64_t n, int64_t m, int n1out,
                                                        const int32_t* __restrict__ kx,
                                                        const int32_t* __restrict__ ky,
                                                        const int32_t* __restrict__ ex,
                                                        const LatDesc* __restrict__ desc,
                                                        const double* __restrict__ vals,
                                                        double* __restrict__ B) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t a0 = (int64_t)blockIdx.y * kLatCols;
  if (i >= n) return;
  const LatDesc* d = desc + (ex[i] - 1);
  const int bx = d->x.Q * kx[i] + d->x.P * (d->x.Nr - 1);
  const int by = d->y.Q * ky[i] + d->y.P * (d->y.Nr - 1);
  const double* v = vals + d->off;
  const int Lx = d->x.L, Px = d->x.P, Py = d->y.P;
  const int64_t an = min(m, a0 + kLatCols);
  for (int64_t a = a0; a < an; ++a) {
    const int ax = (int)(a % n1out), ay = (int)(a / n1out);
    B[i + a * n] = v[(bx - Px * ax) + (int64_t)(by - Py * ay) * Lx];
  }
}

hipError_t launch_grid_check(hipStream_t st, int64_t n, int64_t m, const double* x, const double* y,
                             const int32_t* ex, const int32_t* mask_subs, int n1gim, int n2gim,
                             int nexp, const double* gsx, const double* gsy, const double* dith,
                             int n1out, double osx, double osy, double op1, double op2,
                             const double* Xa, const double* Ya, int32_t* kx, int32_t* ky,
                             int32_t* bad) {
  const unsigned nb = (unsigned)((n + m + 255) / 256);
  hipLaunchKernelGGL(grid_check_kernel, dim3(nb), dim3(256), 0, st, n, m, x, y, ex, mask_subs,
                     n1gim, n2gim, nexp, gsx, gsy, dith, n1out, osx, osy, op1, op2, Xa, Ya, kx,
                     ky, bad);
  return hipGetLastError();
}

hipError_t launch_lattice_eval(hipStream_t st, int dev, const LatDesc* desc, int npairs,
                               int64_t total, int psfconst, const double* tabs, int64_t tstride,
                               int n1, int n2, int npm, int npad, double sx, double sy,
                               double* vals) {
  IMCOM_HIP_TRY(ensure_invden(dev));
  hipLaunchKernelGGL(lattice_eval_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st,
                     desc, npairs, total, psfconst, tabs, tstride, n1, n2, npm, (double)npad, sx,
                     sy, vals);
  return hipGetLastError();
}

hipError_t launch_lattice_A(hipStream_t st, int64_t n, const int32_t* kx, const int32_t* ky,
                            const int32_t* ex, const LatDesc* desc, const double* vals, double* A) {
  dim3 grid((unsigned)((n + 255) / 256), (unsigned)((n + kLatCols - 1) / kLatCols));
  hipLaunchKernelGGL(lattice_A_kernel, grid, dim3(256), 0, st, n, kx, ky, ex, desc, vals, A);
  return hipGetLastError();
}

hipError_t launch_lattice_B(hipStream_t st, int64_t n, int64_t m, int n1out, const int32_t* kx,
                            const int32_t* ky, const int32_t* ex, const LatDesc* desc,
                            const double* vals, double* B) {
  dim3 grid((unsigned)((n + 255) / 256), (unsigned)((m + kLatCols - 1) / kLatCols));
  hipLaunchKernelGGL(lattice_B_kernel, grid, dim3(256), 0, st, n, m, n1out, kx, ky, ex, desc, vals,
                     B);
  return hipGetLastError();
}

// imcom_rotate_image (imcom_proc.f90:184-250): output pixel (iout, jout),
// 1-based, takes the polynomial through the clipped stencil around the
// rotated position (xs, ys), or 0 when the position's floor/ceiling leave the
// image.  cost/sint come from the host; the position is formed with the
// reference's operation order and no fused multiply-adds.
__global__ __launch_bounds__(256) void rotate_kernel(const double* __restrict__ in, int n1, int n2,
                                                     double cost, double sint, int npm,
                                                     double* __restrict__ out) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n1 * n2) return;
  const int iout = (int)(idx % n1) + 1, jout = (int)(idx / n1) + 1;
  const double xc = (double)(n1 + 1) / 2.0, yc = (double)(n2 + 1) / 2.0;
  double xs, ys;
  {
#pragma clang fp contract(off)  // the rotated position rounded as the reference forms it
    const double xo = (double)iout - xc, yo = (double)jout - yc;
    xs = xo * cost + yo * sint + xc;
    ys = -xo * sint + yo * cost + yc;
  }
  const double fi = floor(xs), fj = floor(ys), ci = ceil(xs), cj = ceil(ys);
  double v = 0.0;
  if (fi >= 1.0 && fj >= 1.0 && ci <= (double)n1 && cj <= (double)n2) {
    const int imin = max(1, (int)fi - npm), jmin = max(1, (int)fj - npm);
    const int imax = min(n1, (int)ci + npm), jmax = min(n2, (int)cj + npm);
    const int cx = imax - imin + 1, cy = jmax - jmin + 1;
    double wx[8], wy[8];
    lagrange8(xs, imin, cx, wx);
    lagrange8(ys, jmin, cy, wy);
    const double* base = in + (imin - 1) + (int64_t)(jmin - 1) * n1;
#pragma unroll
    for (int l = 0; l < 8; ++l) {
      const double* col = base + (int64_t)min(l, cy - 1) * n1;
      double r = 0.0;
#pragma unroll
      for (int k = 0; k < 8; ++k) r += wx[k] * col[min(k, cx - 1)];
      v += wy[l] * r;
    }
  }
  out[idx] = v;
}

// sums[0] = sum(in), sums[1] = sum(out): one workgroup, fixed order
__global__ __launch_bounds__(1024) void rotate_sums_kernel(const double* __restrict__ in,
                                                           const double* __restrict__ out,
                                                           int64_t cnt, double* __restrict__ sums) {
  __shared__ double red[16];
  double a = 0.0, b = 0.0;
  for (int64_t i = threadIdx.x; i < cnt; i += 1024) {
    a += in[i];
    b += out[i];
  }
  const double sa = block_sum(a, red);
  const double sb = block_sum(b, red);
  if (threadIdx.x == 0) {
    sums[0] = sa;
    sums[1] = sb;
  }
}

// im_out = sum(im_in) * im_out / sum(im_out)
__global__ __launch_bounds__(256) void rotate_flux_kernel(double* __restrict__ out, int64_t cnt,
                                                          const double* __restrict__ sums) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cnt) out[i] = sums[0] * out[i] / sums[1];
}

hipError_t launch_rotate(hipStream_t st, int dev, const double* in, int n1, int n2, double cost,
                         double sint, int npm, int conserve_flux, double* out, double* sums) {
  IMCOM_HIP_TRY(ensure_invden(dev));
  const int64_t cnt = (int64_t)n1 * n2;
  const unsigned nb = (unsigned)((cnt + 255) / 256);
  hipLaunchKernelGGL(rotate_kernel, dim3(nb), dim3(256), 0, st, in, n1, n2, cost, sint, npm, out);
  IMCOM_HIP_TRY(hipGetLastError());
  if (conserve_flux) {
    hipLaunchKernelGGL(rotate_sums_kernel, dim3(1), dim3(1024), 0, st, in, out, cnt, sums);
    IMCOM_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(rotate_flux_kernel, dim3(nb), dim3(256), 0, st, out, cnt, sums);
  }
  return hipGetLastError();
}

}  // namespace imcom
