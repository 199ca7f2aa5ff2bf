// Tile set-up on the device: the masked input vectors (imcom_alloc_gims,
// imcom_io.f90:84-123, and imcom_build_xy, imcom_proc.f90:56-89), the PSF
// rotation (imcom_rotate_image, imcom_proc.f90:184-250; the kernel is in
// interp.hip), the PSF spectra
// (imcom_ft_r2c, imcom_proc.f90:433-475), the A and B interpolation tables
// (imcom_build_Alookup / imcom_build_Blookup, imcom_matrix.f90:361-493) and
// C (imcom_calc_C, imcom_matrix.f90:546-552).
//
// The reference zero-pads each n x n product spectrum (conj(G_i) G_j, or
// conj(Gamma) G_i) to N = npad*n with two pairs of circular shifts, runs an
// unnormalised inverse c2c FFT of N x N and re-centres the real part.  With
// f in [-n/2, n/2) the signed frequency of a spectrum entry (FFT order
// f mod n), that composition is, exactly,
//
//   table(x1, x2) = s * Re sum_{f1, f2} (-1)^(f1+f2) S[f1 mod n, f2 mod n]
//                          * exp(+2 pi i (f1 x1 + f2 x2) / N)
//
// (the output shift by N/2 is the (-1)^f factor; s = 1/n^2 for A, -2/n^2 for
// B).  Only n of the N frequencies per axis are non-zero, so the transform is
// two dense products with the N x n matrix E(x, a) = exp(2 pi i (a - n/2) x / N)
// of inner size 2n (real and imaginary parts stacked), which run on the FP64
// matrix cores:
//
//   Xs = [Re X; Im X] = [Sr | Si] [Er | -Ei ; Ei | Er]^T     (2n x N)
//   table = s [Er | -Ei] Xs                                  (N x N)
//
// E is held as the N x 4n matrix [Er | -Ei | Ei | Er]; the forward DFT of the
// PSF images uses the same matrix with N = n and no centring.  The DFT
// matrices come from sincospi of the exactly reduced phase (f x mod N).
#include "ctx.h"

#include <algorithm>
#include <cmath>

namespace {

// E (N x 4n, column-major): [C | -S | S | C] with
// C(x, a) = cos(2 pi (a - f0) x / N), S(x, a) = sin(2 pi (a - f0) x / N).
__global__ void dft_matrix_kernel(double* __restrict__ E, int N, int n, int f0) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)N * n) return;
  const int64_t x = idx % N, a = idx / N;
  int64_t r = ((a - f0) * x) % N;
  if (r < 0) r += N;
  double s, c;
  sincospi(2.0 * (double)r / (double)N, &s, &c);
  const int64_t ld = N, nn = n;
  E[x + a * ld] = c;
  E[x + (nn + a) * ld] = -s;
  E[x + (2 * nn + a) * ld] = s;
  E[x + (3 * nn + a) * ld] = c;
}

// Sp (n x 2n) = [Re S' | Im S'], S'(a1, a2) = (-1)^(f1+f2) conj(P) Q at
// (f1 mod n, f2 mod n), f = a - n/2.  P, Q: interleaved complex n x n.
__global__ void table_spectrum_kernel(const double* __restrict__ P, const double* __restrict__ Q,
                                      int n, double* __restrict__ Sp) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  const int a1 = (int)(idx % n), a2 = (int)(idx / n);
  const int f1 = a1 - n / 2, f2 = a2 - n / 2;
  const int i1 = f1 < 0 ? f1 + n : f1, i2 = f2 < 0 ? f2 + n : f2;
  const int64_t s = 2 * ((int64_t)i1 + (int64_t)i2 * n);
  const double pr = P[s], pi = P[s + 1], qr = Q[s], qi = Q[s + 1];
  // conj(p) * q, as Fortran's complex product of dconjg(p) and q
  const double re = pr * qr - (-pi) * qi;
  const double im = pr * qi + (-pi) * qr;
  const double sg = ((f1 + f2) & 1) ? -1.0 : 1.0;
  Sp[idx] = sg * re;
  Sp[idx + (int64_t)n * n] = sg * im;
}

// Full spectrum from the computed real/imaginary parts, with the upper half
// of the first axis replaced by the conjugates of the lower half, as the
// Hermitian completion of imcom_ft_r2c writes it (imcom_proc.f90:459-472).
__global__ void r2c_finish_kernel(const double* __restrict__ Sr, const double* __restrict__ Si,
                                  int n1, int n2, double* __restrict__ out) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n1 * n2) return;
  const int r = (int)(idx % n1), j = (int)(idx / n1);
  double re, im;
  if (r <= n1 / 2) {
    re = Sr[idx];
    im = Si[idx];
  } else {
    const int64_t s = (int64_t)(n1 - r) + (int64_t)((n2 - j) % n2) * n1;
    re = Sr[s];
    im = -Si[s];
  }
  out[2 * idx] = re;
  out[2 * idx + 1] = im;
}

// C = sum |Gamma~|^2 / (n1 n2), one workgroup, fixed summation order.
__global__ __launch_bounds__(1024) void calc_C_kernel(const double* __restrict__ G, int64_t cnt,
                                                      double scale, double* __restrict__ out) {
  __shared__ double red[16];
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < cnt; i += 1024) {
    const double re = G[2 * i], im = G[2 * i + 1];
    acc += re * re - im * (-im);
  }
  const double tot = block_sum(acc, red);
  if (threadIdx.x == 0) out[0] = tot * scale;
}

// Masked input vectors in one workgroup: pixels (i1, i2, e) in column-major
// order with I < saturation (strict, as the reference's where), compacted in
// order by a block-wide exclusive scan per 1024-pixel chunk.
//   x = (i1 - 1) * gimxscale(e) + dither(e, 1), y likewise with i2 and
//   dither(e, 2); exp = e (1-based); I = Im_unmasked.
// dither is Fortran (nexp, 2).  count[0] receives n.
__global__ __launch_bounds__(1024) void assemble_kernel(int n1, int n2, int nexp,
                                                        const double* __restrict__ Im, double sat,
                                                        const double* __restrict__ sx,
                                                        const double* __restrict__ sy,
                                                        const double* __restrict__ dither,
                                                        double* __restrict__ x, double* __restrict__ y,
                                                        int32_t* __restrict__ ex,
                                                        double* __restrict__ I,
                                                        int64_t* __restrict__ count) {
#pragma clang fp contract(off)  // x = i * scale + dither rounded twice, as the reference
  __shared__ int wsum[16];
  __shared__ int64_t base;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int64_t npx = (int64_t)n1 * n2, total = npx * nexp;
  if (tid == 0) base = 0;
  __syncthreads();
  for (int64_t c0 = 0; c0 < total; c0 += 1024) {
    const int64_t idx = c0 + tid;
    const bool in = idx < total;
    const double v = in ? Im[idx] : 0.0;
    const int keep = (in && v < sat) ? 1 : 0;
    // exclusive scan: within the wavefront by ballot, then over wavefronts
    const unsigned long long bal = __ballot(keep);
    const int before = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[w] = __popcll(bal);
    __syncthreads();
    int woff = 0, tot = 0;
    for (int q = 0; q < 16; ++q) {
      woff += q < w ? wsum[q] : 0;
      tot += wsum[q];
    }
    if (keep) {
      const int64_t o = base + woff + before;
      const int e = (int)(idx / npx);
      const int64_t r = idx % npx;
      const int i1 = (int)(r % n1), i2 = (int)(r / n1);
      x[o] = (double)i1 * sx[e] + dither[e];
      y[o] = (double)i2 * sy[e] + dither[nexp + e];
      ex[o] = e + 1;
      I[o] = v;
    }
    __syncthreads();
    if (tid == 0) base += tot;
    __syncthreads();
  }
  if (tid == 0) count[0] = base;
}

inline unsigned blocks_for(int64_t cnt, int bs) { return (unsigned)((cnt + bs - 1) / bs); }

// The DFT matrix for (N, n, f0), cached in the context: slot 0 holds the
// forward PSF transform's, slot 1 the padded table transform's.
hipError_t dft_matrix(imcom_gpu_ctx* c, int slot, int N, int n, int f0, const double** out) {
  TableWork& w = c->tw;
  DevBuf& E = slot == 0 ? w.E_f : w.E_t;
  int& cN = slot == 0 ? w.Ef_N : w.Et_N;
  int& cn = slot == 0 ? w.Ef_n : w.Et_n;
  int& cf = slot == 0 ? w.Ef_f0 : w.Et_f0;
  if (!(cN == N && cn == n && cf == f0)) {
    IMCOM_HIP_TRY(E.reserve((size_t)N * 4 * n * 8));
    const int64_t cnt = (int64_t)N * n;
    hipLaunchKernelGGL(dft_matrix_kernel, dim3(blocks_for(cnt, 256)), dim3(256), 0, c->stream,
                       E.as<double>(), N, n, f0);
    IMCOM_HIP_TRY(hipGetLastError());
    cN = N;
    cn = n;
    cf = f0;
  }
  *out = E.as<double>();
  return hipSuccess;
}

// One table: dst (N x N) = scale * Re(E S' E^T) for the spectra P, Q (device,
// interleaved complex n x n).
hipError_t one_table(imcom_gpu_ctx* c, int n, int N, const double* P, const double* Q,
                     double scale, double* dst) {
  TableWork& w = c->tw;
  const double* E = nullptr;
  IMCOM_HIP_TRY(dft_matrix(c, 1, N, n, n / 2, &E));
  IMCOM_HIP_TRY(w.Sp.reserve((size_t)n * 2 * n * 8));
  IMCOM_HIP_TRY(w.X.reserve((size_t)2 * n * N * 8));
  double* Sp = w.Sp.as<double>();
  double* X = w.X.as<double>();
  hipLaunchKernelGGL(table_spectrum_kernel, dim3(blocks_for((int64_t)n * n, 256)), dim3(256), 0,
                     c->stream, P, Q, n, Sp);
  IMCOM_HIP_TRY(hipGetLastError());
  // Re X = Sp [Er | -Ei]^T, Im X = Sp [Ei | Er]^T, stacked as the rows of X
  IMCOM_HIP_TRY(dgemm(c->stream, false, true, n, N, 2 * n, 1.0, Sp, n, E, N, 0.0, X, 2 * n));
  IMCOM_HIP_TRY(dgemm(c->stream, false, true, n, N, 2 * n, 1.0, Sp, n, E + (int64_t)2 * n * N, N,
                      0.0, X + n, 2 * n));
  // table = scale [Er | -Ei] X
  IMCOM_HIP_TRY(dgemm(c->stream, false, false, N, N, 2 * n, scale, E, N, X, 2 * n, 0.0, dst, N));
  return hipSuccess;
}

int check_psf_grid(imcom_gpu_ctx* ctx, int psfconst, int nexppsf, int n1psf, int n2psf, int npad) {
  NEED(psfconst == 0 || psfconst == 1, IMCOM_EARG, "build tables: PSFCONST must be 1 or 0");
  NEED(nexppsf >= 1 && (psfconst == 0 || nexppsf == 1), IMCOM_EARG,
       "build tables: nexppsf must be >= 1 (and 1 when psfconst = 1)");
  // the reference forces a square padded PSF grid (imcom_io.f90:281-282) of
  // even side (2 * nint(...)), which the centred frequency range needs
  NEED(n1psf == n2psf && n1psf >= 2 && n1psf % 2 == 0, IMCOM_EARG,
       "build tables: the PSF grid must be square with an even side");
  NEED(npad >= 1 && (int64_t)npad * n1psf <= 65536, IMCOM_EARG, "build tables: bad npad");
  return IMCOM_OK;
}

}  // namespace

extern "C" {

int imcom_gpu_assemble_inputs(imcom_gpu_ctx* ctx, int n1gim, int n2gim, int nexp,
                              const double* Im_unmasked, double saturation,
                              const double* gimxscale, const double* gimyscale,
                              const double* dither, int64_t* n_out, double* x_out, double* y_out,
                              int32_t* exp_out, double* I_out) {
  if (!ctx) return IMCOM_EARG;
  NEED(n1gim >= 1 && n2gim >= 1 && nexp >= 1, IMCOM_EARG, "assemble_inputs: bad sizes");
  NEED(Im_unmasked && gimxscale && gimyscale && dither && n_out && x_out && y_out && exp_out &&
           I_out,
       IMCOM_EARG, "assemble_inputs: NULL array");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  TableWork& w = ctx->tw;
  const size_t total = (size_t)n1gim * n2gim * nexp;
  // inputs: image, then the scales and dithers; outputs x, y, I, exp, count
  CK(w.img.reserve(total * 8 + (size_t)4 * nexp * 8), "alloc assemble inputs");
  CK(w.spec.reserve(total * (3 * 8 + 4) + 64), "alloc assemble outputs");
  double* dIm = w.img.as<double>();
  double* dsc = dIm + total;
  CK(hipMemcpyAsync(dIm, Im_unmasked, total * 8, hipMemcpyHostToDevice, ctx->stream), "upload");
  CK(hipMemcpyAsync(dsc, gimxscale, nexp * 8, hipMemcpyHostToDevice, ctx->stream), "upload");
  CK(hipMemcpyAsync(dsc + nexp, gimyscale, nexp * 8, hipMemcpyHostToDevice, ctx->stream), "upload");
  CK(hipMemcpyAsync(dsc + 2 * nexp, dither, 2 * nexp * 8, hipMemcpyHostToDevice, ctx->stream),
     "upload");
  double* dx = w.spec.as<double>();
  double* dy = dx + total;
  double* dI = dy + total;
  int64_t* dn = (int64_t*)(dI + total);
  int32_t* de = (int32_t*)(dn + 8);
  hipLaunchKernelGGL(assemble_kernel, dim3(1), dim3(1024), 0, ctx->stream, n1gim, n2gim, nexp, dIm,
                     saturation, dsc, dsc + nexp, dsc + 2 * nexp, dx, dy, de, dI, dn);
  CK(hipGetLastError(), "assemble_inputs");
  int64_t n = 0;
  CK(hipMemcpyAsync(&n, dn, 8, hipMemcpyDeviceToHost, ctx->stream), "download n");
  CK(hipStreamSynchronize(ctx->stream), "assemble sync");
  *n_out = n;
  if (n > 0) {
    CK(hipMemcpyAsync(x_out, dx, n * 8, hipMemcpyDeviceToHost, ctx->stream), "download x");
    CK(hipMemcpyAsync(y_out, dy, n * 8, hipMemcpyDeviceToHost, ctx->stream), "download y");
    CK(hipMemcpyAsync(I_out, dI, n * 8, hipMemcpyDeviceToHost, ctx->stream), "download I");
    CK(hipMemcpyAsync(exp_out, de, n * 4, hipMemcpyDeviceToHost, ctx->stream), "download exp");
    CK(hipStreamSynchronize(ctx->stream), "assemble sync");
  }
  return IMCOM_OK;
}

int imcom_gpu_rotate_image(imcom_gpu_ctx* ctx, int n1, int n2, const double* im_in,
                           double theta_deg, int npoly, int conserve_flux, double* im_out) {
  if (!ctx) return IMCOM_EARG;
  NEED(n1 >= 1 && n2 >= 1 && im_in && im_out, IMCOM_EARG, "rotate_image: bad sizes or NULL arrays");
  NEED(npoly >= 1 && npoly <= 8, IMCOM_EARG, "rotate_image: npoly must be in 1..7 (8 is taken as 7)");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  TableWork& w = ctx->tw;
  const size_t px = (size_t)n1 * n2;
  CK(w.img.reserve(px * 2 * 8), "alloc rotate images");
  CK(w.Cv.reserve(16), "alloc rotate sums");
  CK(hipMemcpyAsync(w.img.p, im_in, px * 8, hipMemcpyHostToDevice, ctx->stream), "upload image");
  // the reference's angle and its trigonometry in double on the host
  const double pi = 3.141592653589793238462643383279502884;
  const double th = std::fmod(theta_deg, 360.0) * pi / 180.0;
  const int np = npoly % 2 == 0 ? npoly - 1 : npoly;
  double* out = w.img.as<double>() + px;
  CK(launch_rotate(ctx->stream, ctx->device, w.img.as<double>(), n1, n2, std::cos(th),
                   std::sin(th), np / 2, conserve_flux, out, w.Cv.as<double>()),
     "rotate_image");
  CK(hipMemcpyAsync(im_out, out, px * 8, hipMemcpyDeviceToHost, ctx->stream), "download image");
  CK(hipStreamSynchronize(ctx->stream), "rotate_image sync");
  return IMCOM_OK;
}

int imcom_gpu_ft_r2c(imcom_gpu_ctx* ctx, int n1, int n2, int count, const double* img,
                     double* spec_out) {
  if (!ctx) return IMCOM_EARG;
  NEED(n1 >= 1 && n2 >= 1 && count >= 1 && img && spec_out, IMCOM_EARG,
       "ft_r2c: bad sizes or NULL arrays");
  // PSF grids are square (imcom_io.f90:281-282): one DFT matrix serves both axes
  NEED(n1 == n2, IMCOM_EARG, "ft_r2c: the image must be square");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  TableWork& w = ctx->tw;
  const int n = n1;
  const size_t px = (size_t)n * n;
  CK(w.img.reserve(px * count * 8), "alloc ft_r2c images");
  CK(w.spec.reserve(px * count * 16), "alloc ft_r2c spectra");
  CK(w.Y.reserve(px * 2 * 8), "alloc ft_r2c workspace");
  CK(w.Sr.reserve(px * 2 * 8), "alloc ft_r2c workspace");
  const double* E = nullptr;
  CK(dft_matrix(ctx, 0, n, n, 0, &E), "DFT matrix");
  CK(upload(ctx, w.img, img, px * count * 8), "upload images");
  double* Y = w.Y.as<double>();
  double* Sr = w.Sr.as<double>();
  double* Si = Sr + px;
  for (int k = 0; k < count; ++k) {
    const double* G = w.img.as<double>() + (int64_t)k * px;
    // Y = [G C ; G S] (2n x n), the second-axis transform (C, S symmetric)
    CK(dgemm(ctx->stream, false, false, n, n, n, 1.0, G, n, E, n, 0.0, Y, 2 * n), "ft_r2c GEMM");
    CK(dgemm(ctx->stream, false, false, n, n, n, 1.0, G, n, E + (int64_t)2 * n * n, n, 0.0,
             Y + n, 2 * n),
       "ft_r2c GEMM");
    // first axis: Re F = [C | -S] Y,  Im F = -[S | C] Y
    CK(dgemm(ctx->stream, false, false, n, n, 2 * n, 1.0, E, n, Y, 2 * n, 0.0, Sr, n),
       "ft_r2c GEMM");
    CK(dgemm(ctx->stream, false, false, n, n, 2 * n, -1.0, E + (int64_t)2 * n * n, n, Y, 2 * n,
             0.0, Si, n),
       "ft_r2c GEMM");
    hipLaunchKernelGGL(r2c_finish_kernel, dim3(blocks_for((int64_t)px, 256)), dim3(256), 0,
                       ctx->stream, Sr, Si, n, n, w.spec.as<double>() + (int64_t)k * 2 * px);
    CK(hipGetLastError(), "ft_r2c finish");
  }
  CK(download(ctx, spec_out, w.spec, px * count * 16), "download spectra");
  return IMCOM_OK;
}

int imcom_gpu_build_Alookup(imcom_gpu_ctx* ctx, int psfconst, int nexppsf, int n1psf, int n2psf,
                            int npad, const double* Gt_rot, double* Alookup_out) {
  if (!ctx) return IMCOM_EARG;
  int rc = check_psf_grid(ctx, psfconst, nexppsf, n1psf, n2psf, npad);
  if (rc) return rc;
  NEED(Gt_rot, IMCOM_EARG, "build_Alookup: NULL Gt_rot");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  TableWork& w = ctx->tw;
  const int n = n1psf, N = npad * n1psf;
  const size_t sp = (size_t)n * n * 16, ts = (size_t)N * N * 8;
  const int nA = nexppsf * (nexppsf + 1) / 2;
  CK(w.Gt.reserve(sp * nexppsf), "alloc spectra");
  CK(upload(ctx, w.Gt, Gt_rot, sp * nexppsf), "upload Gt_rot");
  CK(ctx->Atab.reserve(ts * nA), "alloc Alookup");
  StageTimer tm(ctx, PS_TABLES);
  const double scale = 1.0 / (double)n1psf / (double)n2psf;
  for (int jexp = 1; jexp <= nexppsf; ++jexp)
    for (int iexp = 1; iexp <= jexp; ++iexp) {
      const int ial = iexp + jexp * (jexp - 1) / 2;  // packed upper triangle
      const double* Gi = w.Gt.as<double>() + (int64_t)(iexp - 1) * n * n * 2;
      const double* Gj = w.Gt.as<double>() + (int64_t)(jexp - 1) * n * n * 2;
      CK(one_table(ctx, n, N, Gi, Gj, scale, ctx->Atab.as<double>() + (int64_t)(ial - 1) * N * N),
         "build_Alookup");
    }
  if (ctx->have_Btab && (ctx->n1 != N || ctx->psfconst != psfconst || ctx->nexppsf != nexppsf))
    ctx->have_Btab = false;
  ctx->psfconst = psfconst;
  ctx->nexppsf = nexppsf;
  ctx->n1 = ctx->n2 = N;
  ctx->have_Atab = true;
  ctx->have_A = false;
  CK(download(ctx, Alookup_out, ctx->Atab, ts * nA), "download Alookup");
  return IMCOM_OK;
}

int imcom_gpu_build_Blookup(imcom_gpu_ctx* ctx, int psfconst, int nexppsf, int n1psf, int n2psf,
                            int npad, const double* Gammat, const double* Gt_rot,
                            double* Blookup_out) {
  if (!ctx) return IMCOM_EARG;
  int rc = check_psf_grid(ctx, psfconst, nexppsf, n1psf, n2psf, npad);
  if (rc) return rc;
  NEED(Gammat && Gt_rot, IMCOM_EARG, "build_Blookup: NULL Gammat or Gt_rot");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  TableWork& w = ctx->tw;
  const int n = n1psf, N = npad * n1psf;
  const size_t sp = (size_t)n * n * 16, ts = (size_t)N * N * 8;
  CK(w.Gt.reserve(sp * nexppsf), "alloc spectra");
  CK(w.Gam.reserve(sp), "alloc Gammat");
  CK(upload(ctx, w.Gt, Gt_rot, sp * nexppsf), "upload Gt_rot");
  CK(upload(ctx, w.Gam, Gammat, sp), "upload Gammat");
  CK(ctx->Btab.reserve(ts * nexppsf), "alloc Blookup");
  StageTimer tm(ctx, PS_TABLES);
  const double scale = -2.0 / (double)n1psf / (double)n2psf;
  for (int iexp = 1; iexp <= nexppsf; ++iexp) {
    const double* Gi = w.Gt.as<double>() + (int64_t)(iexp - 1) * n * n * 2;
    CK(one_table(ctx, n, N, w.Gam.as<double>(), Gi, scale,
                 ctx->Btab.as<double>() + (int64_t)(iexp - 1) * N * N),
       "build_Blookup");
  }
  if (ctx->have_Atab && (ctx->n1 != N || ctx->psfconst != psfconst || ctx->nexppsf != nexppsf))
    ctx->have_Atab = false;
  ctx->psfconst = psfconst;
  ctx->nexppsf = nexppsf;
  ctx->n1 = ctx->n2 = N;
  ctx->have_Btab = true;
  ctx->have_B = false;
  CK(download(ctx, Blookup_out, ctx->Btab, ts * nexppsf), "download Blookup");
  return IMCOM_OK;
}

int imcom_gpu_calc_C(imcom_gpu_ctx* ctx, int n1psf, int n2psf, const double* Gammat,
                     double* C_out) {
  if (!ctx) return IMCOM_EARG;
  NEED(n1psf >= 1 && n2psf >= 1 && Gammat && C_out, IMCOM_EARG, "calc_C: bad arguments");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  TableWork& w = ctx->tw;
  const int64_t cnt = (int64_t)n1psf * n2psf;
  CK(w.Gam.reserve((size_t)cnt * 16), "alloc Gammat");
  CK(w.Cv.reserve(8), "alloc C");
  CK(upload(ctx, w.Gam, Gammat, (size_t)cnt * 16), "upload Gammat");
  hipLaunchKernelGGL(calc_C_kernel, dim3(1), dim3(1024), 0, ctx->stream, w.Gam.as<double>(), cnt,
                     1.0 / (double)n1psf / (double)n2psf, w.Cv.as<double>());
  CK(hipGetLastError(), "calc_C");
  CK(download(ctx, C_out, w.Cv, 8), "download C");
  return IMCOM_OK;
}

}  // extern "C"
