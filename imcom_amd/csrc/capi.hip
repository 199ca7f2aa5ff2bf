// C ABI of the solve path (declared in include/imcom_gpu.h).  Each entry
// point stands in for one procedure of the reference's imcom_matrix /
// imcom_bisect modules; the context owns the device copies of the tile's
// inputs and intermediates.
#include "../../include/imcom_gpu.h"
#include "common.h"
#include "kernels.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace imcom;

#include "ctx.h"

// Q(:, j) = Z(:, n-1-j), L(j) = lam(n-1-j): ascending -> descending
__global__ void reverse_cols_kernel(int64_t n, const double* __restrict__ Z, double* __restrict__ Q,
                                    const double* __restrict__ lam, double* __restrict__ L) {
  const int64_t j = blockIdx.y;
  const int64_t src = n - 1 - j;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    Q[i + j * n] = Z[i + src * n];
  if (blockIdx.x == 0 && threadIdx.x == 0) L[j] = lam[src];
}

// contexts open per device: they share its last-level cache
static std::atomic<int> g_live[64];
int live_contexts(int device) {
  return (device >= 0 && device < 64) ? std::max(1, g_live[device].load()) : 1;
}

namespace imcom {
hipError_t launch_reverse_cols(hipStream_t st, int64_t n, const double* Zin, double* Qout,
                               const double* lam_in, double* L_out) {
  dim3 grid((unsigned)std::min<int64_t>((n + 255) / 256, 64), (unsigned)n);
  hipLaunchKernelGGL(reverse_cols_kernel, grid, dim3(256), 0, st, n, Zin, Qout, lam_in, L_out);
  return hipGetLastError();
}
}  // namespace imcom

extern "C" {

int imcom_gpu_init(imcom_gpu_ctx** out, int device) {
  if (!out) return IMCOM_EARG;
  *out = nullptr;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0) return IMCOM_EHIP;
  if (device < 0 || device >= ndev || device >= 64) return IMCOM_EARG;
  e = hipSetDevice(device);
  if (e != hipSuccess) return IMCOM_EHIP;
  e = noalloc_scratch_setup(device);
  if (e != hipSuccess) return IMCOM_EHIP;
  imcom_gpu_ctx* ctx = new imcom_gpu_ctx();
  ctx->device = device;
  e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete ctx;
    return IMCOM_EHIP;
  }
  ctx->dc = dc_create();
  ctx->prof.st = ctx->stream;
  g_live[device].fetch_add(1);
  *out = ctx;
  return IMCOM_OK;
}

void imcom_gpu_finalize(imcom_gpu_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  (void)fetch_wait_all(ctx);
  if (ctx->copy_stream) (void)hipStreamSynchronize(ctx->copy_stream);
  DevBuf* bufs[] = {&ctx->x,     &ctx->y,    &ctx->ex,     &ctx->Xa,    &ctx->Ya,   &ctx->Nd,
                    &ctx->Atab,  &ctx->Btab, &ctx->I,      &ctx->A,     &ctx->B,    &ctx->Q,
                    &ctx->L,     &ctx->P,    &ctx->K,      &ctx->US,    &ctx->T,    &ctx->W,
                    &ctx->H,     &ctx->S,    &ctx->U,      &ctx->evals, &ctx->bracket,
                    &ctx->split_work, &ctx->mlp_params, &ctx->mlp_feat, &ctx->mlp_bracket,
                    &ctx->mlp_bins, &ctx->mlp_hits, &ctx->mlp_X, &ctx->mlp_Y, &ctx->mlp_perm,
                    &ctx->mlp_loss, &ctx->mlp_grad, &ctx->mlp_m, &ctx->mlp_v, &ctx->mlp_out,
                    &ctx->ms_Xn, &ctx->ms_H1, &ctx->ms_dZ1, &ctx->ms_H2, &ctx->ms_dZ2,
                    &ctx->ms_dY,    &ctx->gkx,  &ctx->gky,   &ctx->latA_desc, &ctx->latB_desc,
                    &ctx->lat_vals};
  for (DevBuf* b : bufs) b->release();
  ctx->ew.release();
  ctx->tw.release();
  ctx->tsw.release();
  imcom_batch_free(ctx);
  dc_destroy(ctx->dc);
  ctx->prof.destroy();
  for (hipEvent_t& e : ctx->fetch_ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->fetch_ready) (void)hipEventDestroy(ctx->fetch_ready);
  if (ctx->copy_stream) (void)hipStreamDestroy(ctx->copy_stream);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  g_live[ctx->device].fetch_sub(1);
  delete ctx;
}

const char* imcom_gpu_last_error(const imcom_gpu_ctx* ctx) {
  return ctx ? ctx->err.c_str() : "no context";
}

int imcom_gpu_synchronize(imcom_gpu_ctx* ctx) {
  if (!ctx) return IMCOM_EARG;
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  CK(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  return IMCOM_OK;
}

int imcom_gpu_set_tile(imcom_gpu_ctx* ctx, int64_t n, int64_t m, const double* x_i,
                       const double* y_i, const int32_t* exp_i, const double* X_a,
                       const double* Y_a, const double* Ndiag, double C_a) {
  if (!ctx) return IMCOM_EARG;
  NEED(n >= 1 && m >= 1, IMCOM_EARG, "set_tile: n and m must be >= 1");
  NEED(n <= 32768, IMCOM_EARG, "set_tile: n > 32768 is not supported");
  NEED(x_i && y_i && exp_i && X_a && Y_a && Ndiag, IMCOM_EARG, "set_tile: NULL input");
  NEED(C_a > 0.0, IMCOM_EARG, "set_tile: C_a must be positive");
  for (int64_t i = 0; i < n; ++i) {
    NEED(exp_i[i] >= 1, IMCOM_EARG, "set_tile: exposure ids are 1-based");
    NEED(i == 0 || exp_i[i] >= exp_i[i - 1], IMCOM_EARG,
         "set_tile: pixels must be exposure-major (exp_i non-decreasing)");
  }
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  CK(fetch_wait_all(ctx), "set_tile: wait for fetches");
  ctx->n = n;
  ctx->m = m;
  ctx->C_a = C_a;
  ctx->h_exp.assign(exp_i, exp_i + n);
  CK(upload(ctx, ctx->x, x_i, n * 8), "upload x_i");
  CK(upload(ctx, ctx->y, y_i, n * 8), "upload y_i");
  CK(upload(ctx, ctx->ex, exp_i, n * 4), "upload exp_i");
  CK(upload(ctx, ctx->Xa, X_a, m * 8), "upload X_a");
  CK(upload(ctx, ctx->Ya, Y_a, m * 8), "upload Y_a");
  CK(upload(ctx, ctx->Nd, Ndiag, n * 8), "upload Ndiag");
  CK(hipStreamSynchronize(ctx->stream), "set_tile sync");
  ctx->have_tile = true;
  ctx->have_A = ctx->have_B = ctx->have_QL = ctx->have_P = ctx->have_K = ctx->have_T = false;
  ctx->have_grid = ctx->lat_A = ctx->lat_B = false;  // a new tile's grids are declared anew
  return IMCOM_OK;
}

// sr / sc = P / Q with coprime P, Q <= 64 (smallest P + Q first), to 1e-13
// relative: the lattice of two grid pitches.
static bool pitch_ratio(double sr, double sc, int& P, int& Q) {
  if (!(sr > 0.0 && sc > 0.0) || !std::isfinite(sr) || !std::isfinite(sc)) return false;
  for (int tot = 2; tot <= 128; ++tot)
    for (int p = std::max(1, tot - 64); p <= std::min(64, tot - 1); ++p) {
      const int q = tot - p;
      if (std::gcd(p, q) != 1) continue;
      if (std::fabs(sr * q - sc * p) <= 1e-13 * sr * q) {
        P = p;
        Q = q;
        return true;
      }
    }
  return false;
}

// The lattice of the subtracted grid (pitch sr, offset dr, Nr points) and the
// added grid (sc, dc, Nc); false when the pitches have no small ratio.
static bool lat_axis(double sr, double dr, int Nr, double sc, double dc, int Nc, LatAxis& a) {
  if (!pitch_ratio(sr, sc, a.P, a.Q)) return false;
  a.Nr = Nr;
  a.Nc = Nc;
  a.L = a.P * (Nr - 1) + a.Q * (Nc - 1) + 1;
  a.g = sr / a.P;
  a.doff = dc - dr;
  return true;
}

// Lattices beyond this many values (for all pairs of one matrix) are not
// formed; the per-entry kernels serve such grids.
static const int64_t kLatMaxValues = (int64_t)1 << 25;

int imcom_gpu_set_grid(imcom_gpu_ctx* ctx, int n1gim, int n2gim, int nexp, const double* gimxscale,
                       const double* gimyscale, const double* dither, const int32_t* mask_subs_i,
                       int n1out, int n2out, double outxscale, double outyscale, double outpos1,
                       double outpos2) {
  if (!ctx) return IMCOM_EARG;
  NEED(ctx->have_tile, IMCOM_ESTATE, "set_grid: tile inputs not set (imcom_gpu_set_tile)");
  ctx->have_grid = ctx->lat_A = ctx->lat_B = false;
  if (n1gim == 0) return IMCOM_OK;  // grids cleared: per-entry lookups
  NEED(n1gim >= 1 && n2gim >= 1 && nexp >= 1 && n1out >= 1 && n2out >= 1, IMCOM_EARG,
       "set_grid: grid sizes must be >= 1");
  NEED(gimxscale && gimyscale && dither && mask_subs_i, IMCOM_EARG, "set_grid: NULL input");
  NEED((int64_t)n1out * n2out == ctx->m, IMCOM_EARG, "set_grid: n1out * n2out must equal m");
  NEED((int64_t)n1gim * n2gim * nexp >= ctx->n, IMCOM_EARG,
       "set_grid: fewer grid pixels than input pixels");
  const int64_t n = ctx->n, m = ctx->m;
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  // per-exposure pitches and dithers, mask_subs, the check counter
  std::vector<double> geo(4 * (size_t)nexp);
  for (int e = 0; e < nexp; ++e) {
    geo[e] = gimxscale[e];
    geo[nexp + e] = gimyscale[e];
    geo[2 * nexp + e] = dither[e];         // dither(e, 1)
    geo[3 * nexp + e] = dither[nexp + e];  // dither(e, 2)
  }
  DevBuf dgeo, dsubs, dbad;
  CK(dgeo.reserve(geo.size() * 8), "alloc grid");
  CK(dsubs.reserve((size_t)n * 4), "alloc mask_subs");
  CK(dbad.reserve(4), "alloc grid check");
  CK(ctx->gkx.reserve((size_t)n * 4), "alloc grid x");
  CK(ctx->gky.reserve((size_t)n * 4), "alloc grid y");
  int32_t bad = 0;
  hipError_t e = hipMemcpyAsync(dgeo.p, geo.data(), geo.size() * 8, hipMemcpyHostToDevice,
                                ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dsubs.p, mask_subs_i, (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dbad.p, &bad, 4, hipMemcpyHostToDevice, ctx->stream);
  const double* g = dgeo.as<double>();
  if (e == hipSuccess)
    e = launch_grid_check(ctx->stream, n, m, ctx->x.as<double>(), ctx->y.as<double>(),
                          ctx->ex.as<int32_t>(), dsubs.as<int32_t>(), n1gim, n2gim, nexp, g,
                          g + nexp, g + 2 * nexp, n1out, outxscale, outyscale, outpos1, outpos2,
                          ctx->Xa.as<double>(), ctx->Ya.as<double>(), ctx->gkx.as<int32_t>(),
                          ctx->gky.as<int32_t>(), dbad.as<int32_t>());
  if (e == hipSuccess) e = hipMemcpyAsync(&bad, dbad.p, 4, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  dgeo.release();
  dsubs.release();
  dbad.release();
  CK(e, "set_grid: grid check");
  NEED(bad == 0, IMCOM_EARG,
       "set_grid: " + std::to_string(bad) +
           " pixels do not lie on the declared grids (x_i, y_i, exp_i, X_a, Y_a)");
  // A: exposure pairs e <= f (descriptor e + f (f + 1) / 2, 0-based), grid e subtracted
  std::vector<LatDesc> dA, dB;
  bool okA = true, okB = true;
  int64_t tot = 0;
  for (int f = 0; f < nexp && okA; ++f)
    for (int e2 = 0; e2 <= f && okA; ++e2) {
      LatDesc d{};
      okA = lat_axis(gimxscale[e2], dither[e2], n1gim, gimxscale[f], dither[f], n1gim, d.x) &&
            lat_axis(gimyscale[e2], dither[nexp + e2], n2gim, gimyscale[f], dither[nexp + f],
                     n2gim, d.y);
      d.table = e2 + f * (f + 1) / 2;
      d.off = tot;
      tot += (int64_t)d.x.L * d.y.L;
      dA.push_back(d);
    }
  okA = okA && tot <= kLatMaxValues;
  ctx->lat_totA = tot;
  // B: exposure e against the output grid (subtracted)
  tot = 0;
  for (int e2 = 0; e2 < nexp && okB; ++e2) {
    LatDesc d{};
    okB = lat_axis(outxscale, outpos1, n1out, gimxscale[e2], dither[e2], n1gim, d.x) &&
          lat_axis(outyscale, outpos2, n2out, gimyscale[e2], dither[nexp + e2], n2gim, d.y);
    d.table = e2;
    d.off = tot;
    tot += (int64_t)d.x.L * d.y.L;
    dB.push_back(d);
  }
  okB = okB && tot <= kLatMaxValues;
  ctx->lat_totB = tot;
  if (okA) {
    CK(upload(ctx, ctx->latA_desc, dA.data(), dA.size() * sizeof(LatDesc)), "upload lattice A");
    ctx->lat_npA = (int)dA.size();
  }
  if (okB) {
    CK(upload(ctx, ctx->latB_desc, dB.data(), dB.size() * sizeof(LatDesc)), "upload lattice B");
    ctx->lat_npB = (int)dB.size();
  }
  CK(hipStreamSynchronize(ctx->stream), "set_grid sync");
  ctx->grid_n1out = n1out;
  ctx->have_grid = true;
  ctx->lat_A = okA;
  ctx->lat_B = okB;
  return IMCOM_OK;
}

int imcom_gpu_lattice_state(imcom_gpu_ctx* ctx, int* A_lattice, int* B_lattice) {
  if (!ctx) return IMCOM_EARG;
  if (A_lattice) *A_lattice = ctx->lat_A ? 1 : 0;
  if (B_lattice) *B_lattice = ctx->lat_B ? 1 : 0;
  return IMCOM_OK;
}

int imcom_gpu_set_tables(imcom_gpu_ctx* ctx, int psfconst, int nexppsf, int n1, int n2,
                         const double* Alookup, const double* Blookup) {
  if (!ctx) return IMCOM_EARG;
  NEED(psfconst == 0 || psfconst == 1, IMCOM_EARG, "set_tables: PSFCONST must be 1 or 0");
  NEED(nexppsf >= 1 && n1 >= 1 && n2 >= 1, IMCOM_EARG, "set_tables: bad sizes");
  NEED(psfconst == 0 || nexppsf == 1, IMCOM_EARG, "set_tables: psfconst=1 needs nexppsf=1");
  NEED(Alookup || Blookup, IMCOM_EARG, "set_tables: NULL tables");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  const size_t ts = (size_t)n1 * n2 * 8;
  const int nA = nexppsf * (nexppsf + 1) / 2;
  if (Alookup) CK(upload(ctx, ctx->Atab, Alookup, ts * nA), "upload Alookup");
  if (Blookup) CK(upload(ctx, ctx->Btab, Blookup, ts * nexppsf), "upload Blookup");
  CK(hipStreamSynchronize(ctx->stream), "set_tables sync");
  ctx->psfconst = psfconst;
  ctx->nexppsf = nexppsf;
  ctx->n1 = n1;
  ctx->n2 = n2;
  ctx->have_Atab = Alookup != nullptr;
  ctx->have_Btab = Blookup != nullptr;
  return IMCOM_OK;
}

static int check_interp(imcom_gpu_ctx* ctx, bool table, const char* which, int npoly, int npad,
                        double sx, double sy) {
  NEED(ctx->have_tile, IMCOM_ESTATE, "tile inputs not set (imcom_gpu_set_tile)");
  NEED(table, IMCOM_ESTATE,
       std::string(which) +
           " lookup table not set (imcom_gpu_set_tables or imcom_gpu_build_" + which + "lookup)");
  NEED(npoly >= 1 && npoly <= 8, IMCOM_EARG, "npoly must be in 1..7 (8 is taken as 7)");
  NEED(npad >= 1, IMCOM_EARG, "npad must be >= 1");
  NEED(sx > 0.0 && sy > 0.0, IMCOM_EARG, "PSF pixel scales must be positive");
  const int ne = ctx->psfconst == 1 ? 1 : ctx->nexppsf;
  NEED(ctx->h_exp.back() <= ne || ctx->psfconst == 1, IMCOM_EARG,
       "exposure id exceeds the number of PSF tables");
  return IMCOM_OK;
}

int imcom_gpu_lookup_A(imcom_gpu_ctx* ctx, int npoly, int npad, double psfxscale,
                       double psfyscale, double* A_out) {
  if (!ctx) return IMCOM_EARG;
  int rc = check_interp(ctx, ctx->have_Atab, "A", npoly, npad, psfxscale, psfyscale);
  if (rc) return rc;
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  CK(fetch_guard(ctx, 1u << IMCOM_FETCH_A), "lookup_A: wait for fetches");
  const int64_t n = ctx->n;
  CK(ctx->A.reserve((size_t)n * n * 8), "alloc A");
  int np = npoly % 2 == 0 ? npoly - 1 : npoly;
  if (ctx->lat_A) CK(ctx->lat_vals.reserve((size_t)ctx->lat_totA * 8), "alloc lattice");
  {
    StageTimer tm(ctx, PS_LOOKUP_A);
    if (ctx->lat_A) {  // regular grids: one interpolation per separation, then A
      CK(launch_lattice_eval(ctx->stream, ctx->device, ctx->latA_desc.as<LatDesc>(), ctx->lat_npA,
                             ctx->lat_totA, ctx->psfconst, ctx->Atab.as<double>(),
                             (int64_t)ctx->n1 * ctx->n2, ctx->n1, ctx->n2, np / 2, npad, psfxscale,
                             psfyscale, ctx->lat_vals.as<double>()),
         "lookup_A lattice");
      CK(launch_lattice_A(ctx->stream, n, ctx->gkx.as<int32_t>(), ctx->gky.as<int32_t>(),
                          ctx->ex.as<int32_t>(), ctx->latA_desc.as<LatDesc>(),
                          ctx->lat_vals.as<double>(), ctx->A.as<double>()),
         "lookup_A");
    } else {
      CK(launch_lookup_A(ctx->stream, ctx->device, n, ctx->x.as<double>(), ctx->y.as<double>(),
                         ctx->ex.as<int32_t>(), ctx->psfconst, ctx->Atab.as<double>(),
                         (int64_t)ctx->n1 * ctx->n2, ctx->n1, ctx->n2, np / 2, npad, psfxscale,
                         psfyscale, ctx->A.as<double>()),
         "lookup_A");
    }
  }
  ctx->have_A = true;
  CK(download(ctx, A_out, ctx->A, (size_t)n * n * 8), "download A");
  return IMCOM_OK;
}

int imcom_gpu_set_A(imcom_gpu_ctx* ctx, const double* A) {
  if (!ctx) return IMCOM_EARG;
  NEED(ctx->have_tile && A, IMCOM_EARG, "set_A: tile not set or NULL A");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  CK(fetch_guard(ctx, 1u << IMCOM_FETCH_A), "set_A: wait for fetches");
  CK(upload(ctx, ctx->A, A, (size_t)ctx->n * ctx->n * 8), "upload A");
  CK(hipStreamSynchronize(ctx->stream), "set_A sync");
  ctx->have_A = true;
  return IMCOM_OK;
}

int imcom_gpu_lookup_B(imcom_gpu_ctx* ctx, int npoly, int npad, double psfxscale,
                       double psfyscale, double* B_out) {
  if (!ctx) return IMCOM_EARG;
  int rc = check_interp(ctx, ctx->have_Btab, "B", npoly, npad, psfxscale, psfyscale);
  if (rc) return rc;
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  CK(fetch_guard(ctx, 1u << IMCOM_FETCH_B), "lookup_B: wait for fetches");
  const int64_t n = ctx->n, m = ctx->m;
  CK(ctx->B.reserve((size_t)n * m * 8), "alloc B");
  int np = npoly % 2 == 0 ? npoly - 1 : npoly;
  if (ctx->lat_B) CK(ctx->lat_vals.reserve((size_t)ctx->lat_totB * 8), "alloc lattice");
  {
    StageTimer tm(ctx, PS_LOOKUP_B);
    if (ctx->lat_B) {
      CK(launch_lattice_eval(ctx->stream, ctx->device, ctx->latB_desc.as<LatDesc>(), ctx->lat_npB,
                             ctx->lat_totB, ctx->psfconst, ctx->Btab.as<double>(),
                             (int64_t)ctx->n1 * ctx->n2, ctx->n1, ctx->n2, np / 2, npad, psfxscale,
                             psfyscale, ctx->lat_vals.as<double>()),
         "lookup_B lattice");
      CK(launch_lattice_B(ctx->stream, n, m, ctx->grid_n1out, ctx->gkx.as<int32_t>(),
                          ctx->gky.as<int32_t>(), ctx->ex.as<int32_t>(),
                          ctx->latB_desc.as<LatDesc>(), ctx->lat_vals.as<double>(),
                          ctx->B.as<double>()),
         "lookup_B");
    } else {
      CK(launch_lookup_B(ctx->stream, ctx->device, n, m, ctx->x.as<double>(), ctx->y.as<double>(),
                         ctx->ex.as<int32_t>(), ctx->Xa.as<double>(), ctx->Ya.as<double>(),
                         ctx->psfconst, ctx->Btab.as<double>(), (int64_t)ctx->n1 * ctx->n2,
                         ctx->n1, ctx->n2, np / 2, npad, psfxscale, psfyscale,
                         ctx->B.as<double>()),
         "lookup_B");
    }
  }
  ctx->have_B = true;
  CK(download(ctx, B_out, ctx->B, (size_t)n * m * 8), "download B");
  return IMCOM_OK;
}

int imcom_gpu_set_B(imcom_gpu_ctx* ctx, const double* B) {
  if (!ctx) return IMCOM_EARG;
  NEED(ctx->have_tile && B, IMCOM_EARG, "set_B: tile not set or NULL B");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  CK(fetch_guard(ctx, 1u << IMCOM_FETCH_B), "set_B: wait for fetches");
  CK(upload(ctx, ctx->B, B, (size_t)ctx->n * ctx->m * 8), "upload B");
  CK(hipStreamSynchronize(ctx->stream), "set_B sync");
  ctx->have_B = true;
  return IMCOM_OK;
}

// A persistent reduction tail whose grid barrier ran out of time (its
// workgroups could not all be resident) leaves an abort flag: the reduction
// is then incomplete and reported as an error.
static int check_tail(imcom_gpu_ctx* ctx, int64_t tail_S) {
  if (tail_S < 2) return IMCOM_OK;
  unsigned flag = 0;
  CK(sytrd_tail_status(ctx->stream, tail_S, ctx->ew.stail.as<double>(), &flag),
     "reduction tail status");
  NEED(flag != 1, IMCOM_EHIP,
       "tridiagonal reduction: the persistent tail kernel's grid barrier timed out "
       "(its workgroups were not all resident); set_reduction_tail(0) avoids it");
  NEED(flag == 0, IMCOM_EHIP,
       "tridiagonal reduction: the persistent tail kernel found its barrier count not reset; "
       "set_reduction_tail(0) avoids it");
  return IMCOM_OK;
}

static hipError_t reserve_eigen(imcom_gpu_ctx* ctx, int64_t n, SyTile& t) {
  if (ctx->tail_rows >= 2) {
    const int S = (int)std::min<int64_t>(n, std::min(ctx->tail_rows, kTailMaxRows));
    if (S >= 2) IMCOM_HIP_TRY(ctx->ew.stail.reserve(tail_xch_doubles(S) * 8));
  }
  IMCOM_HIP_TRY(ctx->Q.reserve((size_t)n * n * 8));
  IMCOM_HIP_TRY(ctx->L.reserve((size_t)n * 8));
  return ctx->ew.reserve(n, ctx->nb, ctx->nbb, ctx->A.as<double>(), ctx->Q.as<double>(), t);
}

static bool use_twostage(const imcom_gpu_ctx* ctx, int64_t n) {
  return ctx->stages == 2 && n >= ctx->twostage_min && n >= 3;
}

// Two-stage reduction of A (overwritten): d, e into the eigensolver
// workspace, the bulge-chasing reflectors into Q's buffer, the panels' V^T
// in A's upper triangle and T in the two-stage workspace.
static int twostage_reduce(imcom_gpu_ctx* ctx, int64_t n) {
  TwoStageBufs w;
  CK(ctx->tsw.reserve(n, w), "alloc two-stage workspace");
  CK(ctx->ew.sd.reserve((size_t)n * 8), "alloc d");
  CK(ctx->ew.se.reserve((size_t)n * 8), "alloc e");
  CK(ctx->Q.reserve((size_t)n * n * 8), "alloc Q");
  {
    StageTimer tm(ctx, PS_SYTRD);
    CK(band_reduce(ctx->stream, ctx->A.as<double>(), n, w), "band reduction");
    CK(bulge_chase(ctx->stream, ctx->A.as<double>(), n, ctx->Q.as<double>(),
                   ctx->ew.sd.as<double>(), ctx->ew.se.as<double>(), w.msg, w.prog, &ctx->prof),
       "bulge chasing");
  }
  unsigned flags[2] = {0, 0};
  CK(twostage_status(ctx->stream, w, n, flags), "two-stage status");
  NEED(flags[0] == 0, IMCOM_EHIP,
       "band reduction: a panel's grid barrier timed out (its workgroups were not all resident)");
  NEED(flags[1] == 0, IMCOM_EHIP,
       "bulge chasing: a wait for the previous sweep timed out (its workgroups were not all "
       "resident)");
  return IMCOM_OK;
}

// Eigendecomposition through the two-stage reduction: Q = Q1 Q2 Z with Z
// the tridiagonal matrix's eigenvectors, descending order as eigen_A's.
static int twostage_eigen(imcom_gpu_ctx* ctx, int64_t n) {
  {
    int rc = twostage_reduce(ctx, n);
    if (rc) return rc;
  }
  TwoStageBufs w;
  CK(ctx->tsw.reserve(n, w), "alloc two-stage workspace");
  CK(ctx->tsw.reserve_q2(n, ctx->stream), "alloc bulge back-transform workspace");
  CK(ctx->ew.slam.reserve((size_t)n * 8), "alloc lam");
  CK(ctx->ew.sZ.reserve((size_t)n * n * 8), "alloc Z");
  CK(ctx->L.reserve((size_t)n * 8), "alloc L");
  {
    StageTimer tm(ctx, PS_STEDC);
    CK(dc_solve(ctx->stream, ctx->dc, n, ctx->ew.sd.as<double>(), ctx->ew.se.as<double>(),
                ctx->ew.slam.as<double>(), ctx->ew.sZ.as<double>()),
       "divide-and-conquer");
  }
  {
    StageTimer tm(ctx, PS_ORMTR);
    const int nS = q2_groups(n);
    const int maxk = nS > 0 ? ctx->tsw.h_blk_off[1] - ctx->tsw.h_blk_off[0] : 0;
    CK(apply_q2(ctx->stream, ctx->Q.as<double>(), n, ctx->tsw.blk_off.as<int>(), nS, maxk,
                ctx->tsw.Tq.as<double>(), ctx->ew.sZ.as<double>(), &ctx->prof),
       "bulge back-transform");
    CK(launch_reverse_cols(ctx->stream, n, ctx->ew.sZ.as<double>(), ctx->Q.as<double>(),
                           ctx->ew.slam.as<double>(), ctx->L.as<double>()),
       "reverse");
    CK(apply_q1(ctx->stream, ctx->A.as<double>(), n, w, ctx->Q.as<double>(), n),
       "band back-transform");
  }
  return IMCOM_OK;
}

static int eigen_finish(imcom_gpu_ctx* ctx, double* Q_out, double* L_out, double* condition);

int imcom_gpu_set_reduction_stages(imcom_gpu_ctx* ctx, int stages, int64_t min_n) {
  if (!ctx) return IMCOM_EARG;
  NEED((stages == 1 || stages == 2) && min_n >= 3, IMCOM_EARG,
       "set_reduction_stages: stages must be 1 or 2 and min_n >= 3");
  ctx->stages = stages;
  ctx->twostage_min = min_n;
  return IMCOM_OK;
}

int imcom_gpu_band_reduce(imcom_gpu_ctx* ctx, double* A_out) {
  if (!ctx) return IMCOM_EARG;
  NEED(ctx->have_A, IMCOM_ESTATE, "band_reduce: A not built (lookup_A or set_A first)");
  NEED(ctx->n >= 3, IMCOM_EARG, "band_reduce: n < 3");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  CK(fetch_guard(ctx, 1u << IMCOM_FETCH_A), "band_reduce: wait for fetches");
  const int64_t n = ctx->n;
  TwoStageBufs w;
  CK(ctx->tsw.reserve(n, w), "alloc two-stage workspace");
  CK(band_reduce(ctx->stream, ctx->A.as<double>(), n, w), "band reduction");
  unsigned flags[2] = {0, 0};
  CK(twostage_status(ctx->stream, w, 2, flags), "two-stage status");
  NEED(flags[0] == 0, IMCOM_EHIP,
       "band reduction: a panel's grid barrier timed out (its workgroups were not all resident)");
  ctx->have_A = false;
  CK(download(ctx, A_out, ctx->A, (size_t)n * n * 8), "download A");
  CK(hipStreamSynchronize(ctx->stream), "sync");
  return IMCOM_OK;
}

int imcom_gpu_eigen_A(imcom_gpu_ctx* ctx, double* Q_out, double* L_out, double* condition) {
  if (!ctx) return IMCOM_EARG;
  NEED(ctx->have_A, IMCOM_ESTATE, "eigen_A: A not built (lookup_A or set_A first)");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  CK(fetch_guard(ctx, (1u << IMCOM_FETCH_A) | (1u << IMCOM_FETCH_Q) | (1u << IMCOM_FETCH_L)),
     "eigen_A: wait for fetches");
  const int64_t n = ctx->n;
  if (use_twostage(ctx, n)) {
    int rc = twostage_eigen(ctx, n);
    ctx->have_A = false;
    if (rc) return rc;
    return eigen_finish(ctx, Q_out, L_out, condition);
  }
  SyTile t;
  CK(reserve_eigen(ctx, n, t), "alloc eigensolver workspace");
  CK(hipMemcpyAsync(ctx->ew.stile.p, &t, sizeof(SyTile), hipMemcpyHostToDevice, ctx->stream),
     "upload tile descriptor");
  std::vector<SyTile> host(1, t);
  const SyTile* dt = ctx->ew.stile.as<SyTile>();
  int64_t tail_S = 0;
  // 1. A = H T H^T (reflectors left in A's lower part)
  {
    StageTimer tm(ctx, PS_SYTRD);
    CK(sytrd_run(ctx->stream, dt, 1, n, host, ctx->nb, live_contexts(ctx->device), &ctx->prof,
                 ctx->tail_rows, ctx->ew.stail.as<double>(), &tail_S),
       "tridiagonal reduction");
  }
  // 2. T = Z diag(lam) Z^T, ascending
  {
    StageTimer tm(ctx, PS_STEDC);
    CK(dc_solve(ctx->stream, ctx->dc, n, ctx->ew.sd.as<double>(), ctx->ew.se.as<double>(),
                ctx->ew.slam.as<double>(), ctx->ew.sZ.as<double>()),
       "divide-and-conquer");
  }
  {
    StageTimer tm(ctx, PS_ORMTR);
    // 3. descending order into Q, L (imcom_matrix.f90:656-663)
    CK(launch_reverse_cols(ctx->stream, n, ctx->ew.sZ.as<double>(), ctx->Q.as<double>(),
                           ctx->ew.slam.as<double>(), ctx->L.as<double>()),
       "reverse");
    // 4. Q = H Q
    CK(ormtr_run(ctx->stream, dt, 1, n, host, ctx->nbb), "back-transform");
  }
  ctx->have_A = false;  // A was overwritten by the reduction
  {
    int rc = check_tail(ctx, tail_S);
    if (rc) return rc;
  }
  return eigen_finish(ctx, Q_out, L_out, condition);
}

// Outputs of eigen_A once Q and L are formed.
static int eigen_finish(imcom_gpu_ctx* ctx, double* Q_out, double* L_out, double* condition) {
  const int64_t n = ctx->n;
  ctx->have_QL = true;
  ctx->have_P = ctx->have_K = ctx->have_T = false;
  CK(download(ctx, Q_out, ctx->Q, (size_t)n * n * 8), "download Q");
  CK(download(ctx, L_out, ctx->L, (size_t)n * 8), "download L");
  if (condition) {
    double ends[2];
    CK(hipMemcpyAsync(&ends[0], ctx->L.as<double>(), 8, hipMemcpyDeviceToHost, ctx->stream),
       "download L0");
    CK(hipMemcpyAsync(&ends[1], ctx->L.as<double>() + (n - 1), 8, hipMemcpyDeviceToHost,
                      ctx->stream),
       "download Ln");
    CK(hipStreamSynchronize(ctx->stream), "sync");
    *condition = ends[0] / ends[1];
  }
  return IMCOM_OK;
}

int imcom_gpu_tridiagonalize(imcom_gpu_ctx* ctx, double* d_out, double* e_out) {
  if (!ctx) return IMCOM_EARG;
  NEED(ctx->have_A, IMCOM_ESTATE, "tridiagonalize: A not built (lookup_A or set_A first)");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  CK(fetch_guard(ctx, 1u << IMCOM_FETCH_A), "tridiagonalize: wait for fetches");
  const int64_t n = ctx->n;
  if (use_twostage(ctx, n)) {
    int rc = twostage_reduce(ctx, n);
    ctx->have_A = false;
    if (rc) return rc;
    CK(download(ctx, d_out, ctx->ew.sd, (size_t)n * 8), "download d");
    CK(download(ctx, e_out, ctx->ew.se, (size_t)(n - 1) * 8), "download e");
    CK(hipStreamSynchronize(ctx->stream), "sync");
    return IMCOM_OK;
  }
  SyTile t;
  CK(reserve_eigen(ctx, n, t), "alloc eigensolver workspace");
  CK(hipMemcpyAsync(ctx->ew.stile.p, &t, sizeof(SyTile), hipMemcpyHostToDevice, ctx->stream),
     "upload tile descriptor");
  std::vector<SyTile> host(1, t);
  int64_t tail_S = 0;
  CK(sytrd_run(ctx->stream, ctx->ew.stile.as<SyTile>(), 1, n, host, ctx->nb,
               live_contexts(ctx->device), &ctx->prof, ctx->tail_rows,
               ctx->ew.stail.as<double>(), &tail_S),
     "tridiagonal reduction");
  ctx->have_A = false;
  {
    int rc = check_tail(ctx, tail_S);
    if (rc) return rc;
  }
  CK(download(ctx, d_out, ctx->ew.sd, (size_t)n * 8), "download d");
  if (n > 1) CK(download(ctx, e_out, ctx->ew.se, (size_t)(n - 1) * 8), "download e");
  CK(hipStreamSynchronize(ctx->stream), "sync");
  return IMCOM_OK;
}

int imcom_gpu_tridiag_eig(imcom_gpu_ctx* ctx, int64_t n, const double* d, const double* e,
                          double* lam_out, double* Z_out) {
  if (!ctx) return IMCOM_EARG;
  NEED(n >= 1 && n <= 32768 && d && (e || n == 1), IMCOM_EARG, "tridiag_eig: bad arguments");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  CK(fetch_guard(ctx, (1u << IMCOM_FETCH_Q) | (1u << IMCOM_FETCH_L)),
     "tridiag_eig: wait for fetches");
  CK(upload(ctx, ctx->ew.sd, d, (size_t)n * 8), "upload d");
  CK(ctx->ew.se.reserve((size_t)n * 8), "alloc e");
  if (n > 1) CK(hipMemcpyAsync(ctx->ew.se.p, e, (size_t)(n - 1) * 8, hipMemcpyHostToDevice, ctx->stream), "upload e");
  CK(ctx->ew.slam.reserve((size_t)n * 8), "alloc lam");
  CK(ctx->ew.sZ.reserve((size_t)n * n * 8), "alloc Z");
  CK(dc_solve(ctx->stream, ctx->dc, n, ctx->ew.sd.as<double>(), ctx->ew.se.as<double>(),
              ctx->ew.slam.as<double>(), ctx->ew.sZ.as<double>()),
     "divide-and-conquer");
  CK(download(ctx, lam_out, ctx->ew.slam, (size_t)n * 8), "download lam");
  CK(download(ctx, Z_out, ctx->ew.sZ, (size_t)n * n * 8), "download Z");
  CK(hipStreamSynchronize(ctx->stream), "sync");
  return IMCOM_OK;
}

int imcom_gpu_set_QL(imcom_gpu_ctx* ctx, const double* Q, const double* L) {
  if (!ctx) return IMCOM_EARG;
  NEED(ctx->have_tile && Q && L, IMCOM_EARG, "set_QL: tile not set or NULL input");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  CK(fetch_guard(ctx, (1u << IMCOM_FETCH_Q) | (1u << IMCOM_FETCH_L)), "set_QL: wait for fetches");
  const int64_t n = ctx->n;
  CK(upload(ctx, ctx->Q, Q, (size_t)n * n * 8), "upload Q");
  CK(upload(ctx, ctx->L, L, (size_t)n * 8), "upload L");
  CK(hipStreamSynchronize(ctx->stream), "set_QL sync");
  ctx->have_QL = true;
  ctx->have_P = ctx->have_K = ctx->have_T = false;
  return IMCOM_OK;
}

static int gemm_prec(imcom_gpu_ctx* ctx, int precision, bool ta, int M, int N, int K,
                     double alpha, const double* A, int64_t lda, const double* B, int64_t ldb,
                     double* C, int64_t ldc, const char* what) {
  if (precision == IMCOM_PREC_FP64) {
    CK(dgemm(ctx->stream, ta, false, M, N, K, alpha, A, lda, B, ldb, 0.0, C, ldc), what);
    return IMCOM_OK;
  }
  NEED(precision == IMCOM_PREC_FP32X3 || precision == IMCOM_PREC_BF16X6, IMCOM_EARG,
       "precision must be IMCOM_PREC_FP64, _FP32X3 or _BF16X6");
  const size_t wb = gemm_split_work_bytes(precision, M, N, K);
  CK(ctx->split_work.reserve(wb), "alloc split workspace");
  CK(gemm_split(ctx->stream, precision, ta, M, N, K, alpha, A, lda, B, ldb, C, ldc,
                ctx->split_work.p, wb),
     what);
  return IMCOM_OK;
}

int imcom_gpu_build_P(imcom_gpu_ctx* ctx, int precision, double* P_out) {
  if (!ctx) return IMCOM_EARG;
  NEED(ctx->have_QL, IMCOM_ESTATE, "build_P: Q, L not available (eigen_A or set_QL first)");
  NEED(ctx->have_B, IMCOM_ESTATE, "build_P: B not built (lookup_B or set_B first)");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  CK(fetch_guard(ctx, 1u << IMCOM_FETCH_P), "build_P: wait for fetches");
  const int64_t n = ctx->n, m = ctx->m;
  CK(ctx->P.reserve((size_t)n * m * 8), "alloc P");
  {
    StageTimer tm(ctx, PS_BUILD_P);
    int rc = gemm_prec(ctx, precision, true, (int)n, (int)m, (int)n, 1.0, ctx->Q.as<double>(), n,
                       ctx->B.as<double>(), n, ctx->P.as<double>(), n, "P = Q^T B");
    if (rc) return rc;
  }
  ctx->have_P = true;
  ctx->have_K = ctx->have_T = false;
  CK(download(ctx, P_out, ctx->P, (size_t)n * m * 8), "download P");
  return IMCOM_OK;
}

int imcom_gpu_set_P(imcom_gpu_ctx* ctx, const double* P) {
  if (!ctx) return IMCOM_EARG;
  NEED(ctx->have_tile && P, IMCOM_EARG, "set_P: tile not set or NULL P");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  CK(fetch_guard(ctx, 1u << IMCOM_FETCH_P), "set_P: wait for fetches");
  CK(upload(ctx, ctx->P, P, (size_t)ctx->n * ctx->m * 8), "upload P");
  CK(hipStreamSynchronize(ctx->stream), "set_P sync");
  ctx->have_P = true;
  ctx->have_K = ctx->have_T = false;
  return IMCOM_OK;
}

int imcom_gpu_bisect_kappa(imcom_gpu_ctx* ctx, int UnotS, double US_max, double k_min,
                           double k_max, double tol, int maxnits, const float* bracket,
                           double* K_out, double* US_out, int32_t* evals_out) {
  if (!ctx) return IMCOM_EARG;
  NEED(ctx->have_P && ctx->have_QL, IMCOM_ESTATE, "bisect_kappa: needs L and P");
  NEED(k_min > 0.0 && k_max > k_min, IMCOM_EARG, "bisect_kappa: need 0 < k_min < k_max");
  NEED(maxnits >= 0, IMCOM_EARG, "bisect_kappa: maxnits must be >= 0");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  const int64_t n = ctx->n, m = ctx->m;
  CK(ctx->K.reserve(m * 8), "alloc K");
  CK(ctx->US.reserve(m * 8), "alloc US");
  CK(ctx->evals.reserve(m * 4), "alloc evals");
  const float* dbr = nullptr;
  if (bracket) {
    CK(upload(ctx, ctx->bracket, bracket, (size_t)m * 2 * sizeof(float)), "upload bracket");
    dbr = ctx->bracket.as<float>();
  }
  {
    StageTimer tm(ctx, PS_BISECT);
    CK(launch_bisect(ctx->stream, n, m, ctx->L.as<double>(), ctx->P.as<double>(), n,
                     ctx->Nd.as<double>(), ctx->C_a, US_max, tol, k_min, k_max, maxnits,
                     UnotS ? 1 : 0, dbr, ctx->K.as<double>(), ctx->US.as<double>(),
                     ctx->evals.as<int32_t>(), nullptr),
       "bisect");
  }
  ctx->have_K = true;
  ctx->have_T = false;
  CK(download(ctx, K_out, ctx->K, m * 8), "download K");
  CK(download(ctx, US_out, ctx->US, m * 8), "download U/S");
  CK(download(ctx, evals_out, ctx->evals, m * 4), "download evals");
  return IMCOM_OK;
}

int imcom_gpu_build_T(imcom_gpu_ctx* ctx, int precision, double* T_out) {
  if (!ctx) return IMCOM_EARG;
  NEED(ctx->have_K && ctx->have_P && ctx->have_QL, IMCOM_ESTATE,
       "build_T: needs Q, L, P and K (bisect_kappa first)");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  CK(fetch_guard(ctx, 1u << IMCOM_FETCH_T), "build_T: wait for fetches");
  const int64_t n = ctx->n, m = ctx->m;
  CK(ctx->W.reserve((size_t)n * m * 8), "alloc W");
  CK(ctx->T.reserve((size_t)n * m * 8), "alloc T");
  {
    StageTimer tm(ctx, PS_BUILD_T);
    CK(launch_scale_W(ctx->stream, n, m, ctx->P.as<double>(), n, ctx->L.as<double>(),
                      ctx->K.as<double>(), ctx->W.as<double>(), n),
       "W = P / (L + K)");
    int rc = gemm_prec(ctx, precision, false, (int)n, (int)m, (int)n, -0.5, ctx->Q.as<double>(),
                       n, ctx->W.as<double>(), n, ctx->T.as<double>(), n, "T = -Q W / 2");
    if (rc) return rc;
  }
  ctx->have_T = true;
  CK(download(ctx, T_out, ctx->T, (size_t)n * m * 8), "download T");
  return IMCOM_OK;
}

int imcom_gpu_set_T(imcom_gpu_ctx* ctx, const double* T) {
  if (!ctx) return IMCOM_EARG;
  NEED(ctx->have_tile && T, IMCOM_EARG, "set_T: tile not set or NULL T");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  CK(fetch_guard(ctx, 1u << IMCOM_FETCH_T), "set_T: wait for fetches");
  CK(upload(ctx, ctx->T, T, (size_t)ctx->n * ctx->m * 8), "upload T");
  CK(hipStreamSynchronize(ctx->stream), "set_T sync");
  ctx->have_T = true;
  return IMCOM_OK;
}

int imcom_gpu_build_S(imcom_gpu_ctx* ctx, double* S_out) {
  if (!ctx) return IMCOM_EARG;
  NEED(ctx->have_T, IMCOM_ESTATE, "build_S: T not built");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  CK(ctx->S.reserve(ctx->m * 8), "alloc S");
  StageTimer tm(ctx, PS_MAPS);
  CK(launch_maps(ctx->stream, ctx->n, ctx->m, ctx->T.as<double>(), ctx->n, nullptr,
                 ctx->Nd.as<double>(), nullptr, ctx->S.as<double>()),
     "build_S");
  CK(download(ctx, S_out, ctx->S, ctx->m * 8), "download S");
  return IMCOM_OK;
}

int imcom_gpu_build_U(imcom_gpu_ctx* ctx, double* U_out) {
  if (!ctx) return IMCOM_EARG;
  NEED(ctx->have_K && ctx->have_P && ctx->have_QL, IMCOM_ESTATE, "build_U: needs L, P and K");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  CK(ctx->U.reserve(ctx->m * 8), "alloc U");
  StageTimer tm(ctx, PS_MAPS);
  CK(launch_U_from_K(ctx->stream, ctx->n, ctx->m, ctx->L.as<double>(), ctx->P.as<double>(), ctx->n,
                     ctx->C_a, ctx->K.as<double>(), ctx->U.as<double>()),
     "build_U");
  CK(download(ctx, U_out, ctx->U, ctx->m * 8), "download U");
  return IMCOM_OK;
}

int imcom_gpu_build_image(imcom_gpu_ctx* ctx, const double* I_i, double* H_out) {
  if (!ctx) return IMCOM_EARG;
  NEED(ctx->have_T, IMCOM_ESTATE, "build_image: T not built");
  NEED(I_i != nullptr, IMCOM_EARG, "build_image: NULL I_i");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  CK(upload(ctx, ctx->I, I_i, ctx->n * 8), "upload I_i");
  CK(ctx->H.reserve(ctx->m * 8), "alloc H");
  StageTimer tm(ctx, PS_MAPS);
  CK(launch_maps(ctx->stream, ctx->n, ctx->m, ctx->T.as<double>(), ctx->n, ctx->I.as<double>(),
                 nullptr, ctx->H.as<double>(), nullptr),
     "build_image");
  CK(download(ctx, H_out, ctx->H, ctx->m * 8), "download H");
  return IMCOM_OK;
}

int imcom_gpu_set_profiling(imcom_gpu_ctx* ctx, int level) {
  if (!ctx) return IMCOM_EARG;
  NEED(level >= 0 && level <= 2, IMCOM_EARG, "profiling level must be 0, 1 or 2");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  CK(ctx->prof.collect(), "collect stage times");
  ctx->prof.level = level;
  ctx->prof.reset();
  return IMCOM_OK;
}

int imcom_gpu_stage_times(imcom_gpu_ctx* ctx, double* ms_out, int64_t* counts_out, int nslots) {
  if (!ctx) return IMCOM_EARG;
  NEED(ms_out && nslots > 0, IMCOM_EARG, "stage_times: NULL output");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  CK(ctx->prof.collect(), "collect stage times");
  for (int i = 0; i < nslots; ++i) {
    ms_out[i] = i < PS_NSLOTS ? ctx->prof.ms[i] : 0.0;
    if (counts_out) counts_out[i] = i < PS_NSLOTS ? ctx->prof.count[i] : 0;
  }
  return IMCOM_OK;
}

}  // extern "C"
