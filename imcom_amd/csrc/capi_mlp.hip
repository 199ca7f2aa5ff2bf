// C ABI of the learned kappa bracket (declared in include/imcom_gpu.h).
// No reference counterpart: the reference always bisects the full
// [kappa_min, kappa_max] range (imcom_bisect.f90:174-267).  The network only
// narrows where that same bisection starts; a bracket that does not hold the
// root falls back to the full range.
#include "ctx.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>

extern "C" {

int imcom_gpu_mlp_nparams(void) { return mlp_nparams(); }

int imcom_gpu_mlp_set_params(imcom_gpu_ctx* ctx, const float* params) {
  if (!ctx) return IMCOM_EARG;
  NEED(params != nullptr, IMCOM_EARG, "mlp_set_params: NULL params");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  for (int i = 0; i < 32; ++i)
    NEED(params[mlp_ntrain() + 32 + i] > 0.0f, IMCOM_EARG, "mlp_set_params: feature std <= 0");
  NEED(params[mlp_nparams() - 1] > 0.0f, IMCOM_EARG, "mlp_set_params: target std <= 0");
  CK(upload(ctx, ctx->mlp_params, params, (size_t)mlp_nparams() * 4), "upload params");
  CK(hipStreamSynchronize(ctx->stream), "sync");
  ctx->have_mlp = true;
  ctx->mlp_step = 0;
  return IMCOM_OK;
}

int imcom_gpu_mlp_get_params(imcom_gpu_ctx* ctx, float* params) {
  if (!ctx) return IMCOM_EARG;
  NEED(ctx->have_mlp && params, IMCOM_ESTATE, "mlp_get_params: no parameters set");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  CK(download(ctx, params, ctx->mlp_params, (size_t)mlp_nparams() * 4), "download params");
  return IMCOM_OK;
}

int imcom_gpu_kappa_features(imcom_gpu_ctx* ctx, double US_target, float* X_out) {
  if (!ctx) return IMCOM_EARG;
  NEED(ctx->have_P && ctx->have_QL, IMCOM_ESTATE, "kappa_features: needs L and P");
  NEED(US_target > 0.0, IMCOM_EARG, "kappa_features: target must be positive");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  const int64_t n = ctx->n, m = ctx->m;
  CK(ctx->mlp_bins.reserve(n), "alloc bins");
  CK(ctx->mlp_feat.reserve((size_t)m * 32 * 4), "alloc features");
  CK(launch_features(ctx->stream, n, m, ctx->L.as<double>(), ctx->mlp_bins.as<int8_t>(),
                     ctx->P.as<double>(), n, ctx->C_a, US_target, ctx->mlp_feat.as<float>()),
     "features");
  CK(download(ctx, X_out, ctx->mlp_feat, (size_t)m * 32 * 4), "download features");
  return IMCOM_OK;
}

int imcom_gpu_mlp_infer(imcom_gpu_ctx* ctx, const float* X, int64_t N, float* out) {
  if (!ctx) return IMCOM_EARG;
  NEED(ctx->have_mlp, IMCOM_ESTATE, "mlp_infer: no parameters set");
  NEED(X && out && N >= 0, IMCOM_EARG, "mlp_infer: bad arguments");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  CK(upload(ctx, ctx->mlp_X, X, (size_t)N * 32 * 4), "upload X");
  CK(ctx->mlp_out.reserve((size_t)N * 2 * 4), "alloc out");
  CK(launch_mlp_infer(ctx->stream, ctx->mlp_params.as<float>(), ctx->mlp_X.as<float>(), N, 0.0f,
                      ctx->mlp_out.as<float>(), 0),
     "mlp_infer");
  CK(download(ctx, out, ctx->mlp_out, (size_t)N * 2 * 4), "download out");
  return IMCOM_OK;
}

static hipError_t mlp_reserve(imcom_gpu_ctx* ctx, int64_t B) {
  const size_t f = 4;
  IMCOM_HIP_TRY(ctx->ms_Xn.reserve(B * 32 * f));
  IMCOM_HIP_TRY(ctx->ms_H1.reserve(B * 64 * f));
  IMCOM_HIP_TRY(ctx->ms_dZ1.reserve(B * 64 * f));
  IMCOM_HIP_TRY(ctx->ms_H2.reserve(B * 64 * f));
  IMCOM_HIP_TRY(ctx->ms_dZ2.reserve(B * 64 * f));
  IMCOM_HIP_TRY(ctx->ms_dY.reserve(B * 2 * f));
  IMCOM_HIP_TRY(ctx->mlp_grad.reserve((size_t)mlp_ntrain() * f));
  return hipSuccess;
}

static MlpScratch scratch(imcom_gpu_ctx* ctx) {
  return MlpScratch{ctx->ms_Xn.as<float>(), ctx->ms_H1.as<float>(), ctx->ms_dZ1.as<float>(),
                    ctx->ms_H2.as<float>(), ctx->ms_dZ2.as<float>(), ctx->ms_dY.as<float>()};
}

int imcom_gpu_mlp_grad(imcom_gpu_ctx* ctx, const float* X, const float* y, int64_t N,
                       float* grad_out, float* loss_out) {
  if (!ctx) return IMCOM_EARG;
  NEED(ctx->have_mlp, IMCOM_ESTATE, "mlp_grad: no parameters set");
  NEED(X && y && N > 0, IMCOM_EARG, "mlp_grad: bad arguments");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  CK(upload(ctx, ctx->mlp_X, X, (size_t)N * 32 * 4), "upload X");
  CK(upload(ctx, ctx->mlp_Y, y, (size_t)N * 4), "upload y");
  CK(mlp_reserve(ctx, N), "alloc mlp scratch");
  CK(ctx->mlp_loss.reserve((size_t)N * 4), "alloc loss");
  MlpScratch s = scratch(ctx);
  CK(mlp_gradient(ctx->stream, ctx->mlp_params.as<float>(), ctx->mlp_X.as<float>(),
                  ctx->mlp_Y.as<float>(), nullptr, 0, N, s, ctx->mlp_grad.as<float>(),
                  ctx->mlp_loss.as<float>()),
     "mlp_gradient");
  CK(download(ctx, grad_out, ctx->mlp_grad, (size_t)mlp_ntrain() * 4), "download grad");
  if (loss_out) {
    std::vector<float> l(N);
    CK(download(ctx, l.data(), ctx->mlp_loss, (size_t)N * 4), "download loss");
    double acc = 0.0;
    for (float v : l) acc += v;
    *loss_out = (float)(acc / (double)N);
  }
  return IMCOM_OK;
}

int imcom_gpu_mlp_train(imcom_gpu_ctx* ctx, const float* X, const float* y, int64_t N, int epochs,
                        int batch, float lr, uint64_t seed, float* loss_out) {
  if (!ctx) return IMCOM_EARG;
  NEED(ctx->have_mlp, IMCOM_ESTATE, "mlp_train: set initial parameters first");
  NEED(X && y && N > 0 && epochs >= 0 && batch > 0 && lr > 0.0f, IMCOM_EARG,
       "mlp_train: bad arguments");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  const int np = mlp_ntrain();
  CK(upload(ctx, ctx->mlp_X, X, (size_t)N * 32 * 4), "upload X");
  CK(upload(ctx, ctx->mlp_Y, y, (size_t)N * 4), "upload y");
  const int64_t B = std::min<int64_t>(batch, N);
  CK(mlp_reserve(ctx, B), "alloc mlp scratch");
  CK(ctx->mlp_loss.reserve((size_t)N * 4), "alloc loss");
  CK(ctx->mlp_perm.reserve((size_t)N * 4), "alloc perm");
  if (ctx->mlp_step == 0) {
    CK(ctx->mlp_m.reserve((size_t)np * 4), "alloc adam m");
    CK(ctx->mlp_v.reserve((size_t)np * 4), "alloc adam v");
    CK(hipMemsetAsync(ctx->mlp_m.p, 0, (size_t)np * 4, ctx->stream), "zero adam m");
    CK(hipMemsetAsync(ctx->mlp_v.p, 0, (size_t)np * 4, ctx->stream), "zero adam v");
  }
  std::mt19937_64 rng(seed);
  std::vector<int> perm(N);
  std::vector<float> losses(N);
  MlpScratch s = scratch(ctx);
  for (int ep = 0; ep < epochs; ++ep) {
    std::iota(perm.begin(), perm.end(), 0);
    std::shuffle(perm.begin(), perm.end(), rng);
    CK(hipMemcpyAsync(ctx->mlp_perm.p, perm.data(), (size_t)N * 4, hipMemcpyHostToDevice,
                      ctx->stream),
       "upload perm");
    for (int64_t off = 0; off + B <= N; off += B) {
      CK(mlp_gradient(ctx->stream, ctx->mlp_params.as<float>(), ctx->mlp_X.as<float>(),
                      ctx->mlp_Y.as<float>(), ctx->mlp_perm.as<int>(), off, B, s,
                      ctx->mlp_grad.as<float>(), ctx->mlp_loss.as<float>() + off),
         "mlp_gradient");
      ctx->mlp_step += 1;
      CK(launch_adam(ctx->stream, ctx->mlp_params.as<float>(), ctx->mlp_grad.as<float>(),
                     ctx->mlp_m.as<float>(), ctx->mlp_v.as<float>(), np, lr, ctx->mlp_step),
         "adam");
    }
    const int64_t used = (N / B) * B;
    CK(download(ctx, losses.data(), ctx->mlp_loss, (size_t)used * 4), "download loss");
    double acc = 0.0;
    for (int64_t i = 0; i < used; ++i) acc += losses[i];
    if (loss_out) loss_out[ep] = (float)(acc / (double)std::max<int64_t>(used, 1));
  }
  return IMCOM_OK;
}

int imcom_gpu_bisect_kappa_learned(imcom_gpu_ctx* ctx, int UnotS, double US_max, double k_min,
                                   double k_max, double tol, int maxnits, float zsig,
                                   double* K_out, double* US_out, int32_t* evals_out,
                                   int32_t* hits_out) {
  if (!ctx) return IMCOM_EARG;
  NEED(ctx->have_mlp, IMCOM_ESTATE, "bisect_kappa_learned: no network parameters set");
  NEED(ctx->have_P && ctx->have_QL, IMCOM_ESTATE, "bisect_kappa_learned: needs L and P");
  NEED(k_min > 0.0 && k_max > k_min && US_max > 0.0 && zsig > 0.0f, IMCOM_EARG,
       "bisect_kappa_learned: bad arguments");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  const int64_t n = ctx->n, m = ctx->m;
  CK(ctx->mlp_bins.reserve(n), "alloc bins");
  CK(ctx->mlp_feat.reserve((size_t)m * 32 * 4), "alloc features");
  CK(ctx->mlp_bracket.reserve((size_t)m * 2 * 4), "alloc bracket");
  CK(ctx->mlp_hits.reserve((size_t)m * 4), "alloc hits");
  CK(ctx->K.reserve(m * 8), "alloc K");
  CK(ctx->US.reserve(m * 8), "alloc US");
  CK(ctx->evals.reserve(m * 4), "alloc evals");
  {
    StageTimer tm(ctx, PS_BISECT);
    CK(launch_features(ctx->stream, n, m, ctx->L.as<double>(), ctx->mlp_bins.as<int8_t>(),
                       ctx->P.as<double>(), n, ctx->C_a, US_max, ctx->mlp_feat.as<float>()),
       "features");
    CK(launch_mlp_infer(ctx->stream, ctx->mlp_params.as<float>(), ctx->mlp_feat.as<float>(), m, zsig,
                        ctx->mlp_bracket.as<float>(), 1),
       "mlp_infer");
    CK(launch_bisect(ctx->stream, n, m, ctx->L.as<double>(), ctx->P.as<double>(), n,
                     ctx->Nd.as<double>(), ctx->C_a, US_max, tol, k_min, k_max, maxnits,
                     UnotS ? 1 : 0, ctx->mlp_bracket.as<float>(), ctx->K.as<double>(),
                     ctx->US.as<double>(), ctx->evals.as<int32_t>(), ctx->mlp_hits.as<int32_t>()),
       "bisect");
  }
  ctx->have_K = true;
  ctx->have_T = false;
  CK(download(ctx, K_out, ctx->K, m * 8), "download K");
  CK(download(ctx, US_out, ctx->US, m * 8), "download U/S");
  CK(download(ctx, evals_out, ctx->evals, m * 4), "download evals");
  CK(download(ctx, hits_out, ctx->mlp_hits, m * 4), "download hits");
  return IMCOM_OK;
}

}  // extern "C"
