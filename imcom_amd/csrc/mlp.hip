// Learned kappa bracket: features, a 32-64-64-2 GELU MLP (FP32 weights) with
// hand-written inference, forward/backward and Adam kernels.
//
// The network reads, per output pixel a, 32 features of the tile's spectrum
// and of the pixel's projection column, and predicts the log10 kappa root of
// the bisection (imcom_bisect.f90:174-267) as a mean mu and a log-spread s:
// the bracket handed to the bisection kernel is [mu - z sigma, mu + z sigma]
// in log10 kappa, sigma = exp(s).  The bisection validates the bracket and
// falls back to the full range on a miss, so the network only ever shortens
// the search.
//
// Features (all FP32, from FP64 inputs):
//   0..11   log10(|lambda_q| / C) at 12 evenly spaced quantiles of the
//           descending eigenvalues (shared by the tile's pixels)
//   12..30  log10(sum of P_ia^2 / C over eigenvalues in decade b below
//           lambda_max, b = 0..18; b = 18 also takes lambda <= 0)
//   31      log10(U or S target)
// Parameter block (floats): W1[32][64] b1[64] W2[64][64] b2[64] W3[64][2] b3[2]
// then the feature standardisation xmean[32] xstd[32] and ymean, ystd.
#include "common.h"
#include "kernels.h"

namespace imcom {

constexpr int NF = 32, NH = 64, NO = 2;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NF * NH, OFF_W2 = OFF_B1 + NH,
              OFF_B2 = OFF_W2 + NH * NH, OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NH * NO,
              N_TRAIN = OFF_B3 + NO, OFF_XM = N_TRAIN, OFF_XS = OFF_XM + NF, OFF_YM = OFF_XS + NF,
              OFF_YS = OFF_YM + 1, N_PARAMS = OFF_YS + 1;
static_assert(N_TRAIN == 6402, "parameter layout");

int mlp_nparams() { return N_PARAMS; }
int mlp_ntrain() { return N_TRAIN; }

__device__ __forceinline__ float gelu(float x) {
  return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f));
}
__device__ __forceinline__ float gelu_grad(float x) {
  const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
  return cdf + x * pdf;
}

// ---- features -------------------------------------------------------------
// bin of each eigenvalue: decade below lambda_max (0..18)
__global__ void eig_bins_kernel(int64_t n, const double* __restrict__ L, int8_t* __restrict__ bin) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double lmax = L[0];
  const double li = L[i];
  int b = 18;
  if (li > 0.0 && lmax > 0.0) {
    const double d = -log10(li / lmax);
    b = d < 0.0 ? 0 : (d >= 18.0 ? 18 : (int)d);
  }
  bin[i] = (int8_t)b;
}

__global__ __launch_bounds__(256) void features_kernel(int64_t n, int64_t m,
                                                       const double* __restrict__ L,
                                                       const int8_t* __restrict__ bin,
                                                       const double* __restrict__ P, int64_t ldp,
                                                       double C_a, double target,
                                                       float* __restrict__ X) {
  __shared__ double red[4][19];
  const int64_t a = blockIdx.x;
  double acc[19];
#pragma unroll
  for (int b = 0; b < 19; ++b) acc[b] = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    const double p = P[i + a * ldp];
    const int b = bin[i];
#pragma unroll
    for (int q = 0; q < 19; ++q) acc[q] += (q == b) ? p * p : 0.0;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    const double s = wave_sum(acc[q]);
    if (lane == 0) red[wid][q] = s;
  }
  __syncthreads();
  float* x = X + a * NF;
  if (threadIdx.x < 19) {
    const int q = threadIdx.x;
    const double s = red[0][q] + red[1][q] + red[2][q] + red[3][q];
    x[12 + q] = (float)log10(s / C_a + 1e-30);
  } else if (threadIdx.x >= 32 && threadIdx.x < 44) {
    const int q = threadIdx.x - 32;
    const int64_t idx = (int64_t)llround((double)q * (double)(n - 1) / 11.0);
    x[q] = (float)log10(fabs(L[idx]) / C_a + 1e-300);
  } else if (threadIdx.x == 44) {
    x[31] = (float)log10(target);
  }
}

// ---- inference ---------------------------------------------------------------
// The bracket's sigma is never below kSigmaFloor decades: the network's own
// sigma is calibrated on the tiles it was trained on, and a tile it has not
// seen can sit further from its prediction than that.
constexpr float kSigmaFloor = 0.1f;

// One thread per sample; parameters staged in LDS (every lane reads the same
// weight: LDS broadcast).  Writes raw outputs (mu, s) in standardised units, or
// the bracket [mu - z sigma', mu + z sigma'] in log10 kappa, sigma' =
// sqrt(sigma^2 + kSigmaFloor^2).
__global__ __launch_bounds__(256) void mlp_infer_kernel(const float* __restrict__ params,
                                                        const float* __restrict__ X, int64_t N,
                                                        float zsig, float* __restrict__ out,
                                                        int as_bracket) {
  __shared__ float w[N_PARAMS];
  for (int i = threadIdx.x; i < N_PARAMS; i += blockDim.x) w[i] = params[i];
  __syncthreads();
  const int64_t sidx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (sidx >= N) return;
  float x[NF];
#pragma unroll
  for (int i = 0; i < NF; ++i) x[i] = (X[sidx * NF + i] - w[OFF_XM + i]) / w[OFF_XS + i];
  float h1[NH];
#pragma unroll
  for (int o = 0; o < NH; ++o) {
    float z = w[OFF_B1 + o];
#pragma unroll
    for (int i = 0; i < NF; ++i) z += x[i] * w[OFF_W1 + i * NH + o];
    h1[o] = gelu(z);
  }
  float y0 = w[OFF_B3], y1 = w[OFF_B3 + 1];
  for (int o = 0; o < NH; ++o) {
    float z = w[OFF_B2 + o];
#pragma unroll
    for (int i = 0; i < NH; ++i) z += h1[i] * w[OFF_W2 + i * NH + o];
    const float h = gelu(z);
    y0 += h * w[OFF_W3 + o * NO];
    y1 += h * w[OFF_W3 + o * NO + 1];
  }
  if (as_bracket) {
    const float ym = w[OFF_YM], ys = w[OFF_YS];
    const float mu = ym + ys * y0;
    const float sn = ys * expf(fminf(fmaxf(y1, -10.f), 5.f));
    const float sg = sqrtf(sn * sn + kSigmaFloor * kSigmaFloor);
    out[2 * sidx] = mu - zsig * sg;
    out[2 * sidx + 1] = mu + zsig * sg;
  } else {
    out[2 * sidx] = y0;
    out[2 * sidx + 1] = y1;
  }
}

// ---- training ------------------------------------------------------------------
// Per sample: forward, Gaussian negative log-likelihood of the standardised
// target t = (y - ymean)/ystd under N(mu, exp(s)^2),
//   loss = 0.5 ((t - mu) / sigma)^2 + s,
// and backward to the pre-activation gradients; the activations and
// gradients go to global memory for the weight-gradient reductions.
// Sample indices come through `perm` (minibatch order).
__global__ __launch_bounds__(256) void mlp_fwdbwd_kernel(
    const float* __restrict__ params, const float* __restrict__ X, const float* __restrict__ Y,
    const int* __restrict__ perm, int64_t off, int64_t B, float* __restrict__ Xn,
    float* __restrict__ H1, float* __restrict__ dZ1, float* __restrict__ H2,
    float* __restrict__ dZ2, float* __restrict__ dY, float* __restrict__ loss) {
  __shared__ float w[N_PARAMS];
  for (int i = threadIdx.x; i < N_PARAMS; i += blockDim.x) w[i] = params[i];
  __syncthreads();
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const int64_t sidx = perm ? perm[off + b] : off + b;
  float x[NF];
#pragma unroll
  for (int i = 0; i < NF; ++i) {
    x[i] = (X[sidx * NF + i] - w[OFF_XM + i]) / w[OFF_XS + i];
    Xn[b * NF + i] = x[i];
  }
  float z1[NH], h1[NH];
#pragma unroll
  for (int o = 0; o < NH; ++o) {
    float z = w[OFF_B1 + o];
#pragma unroll
    for (int i = 0; i < NF; ++i) z += x[i] * w[OFF_W1 + i * NH + o];
    z1[o] = z;
    h1[o] = gelu(z);
    H1[b * NH + o] = h1[o];
  }
  float z2[NH];
  float y0 = w[OFF_B3], y1 = w[OFF_B3 + 1];
  for (int o = 0; o < NH; ++o) {
    float z = w[OFF_B2 + o];
#pragma unroll
    for (int i = 0; i < NH; ++i) z += h1[i] * w[OFF_W2 + i * NH + o];
    z2[o] = z;
    const float h = gelu(z);
    H2[b * NH + o] = h;
    y0 += h * w[OFF_W3 + o * NO];
    y1 += h * w[OFF_W3 + o * NO + 1];
  }
  const float t = (Y[sidx] - w[OFF_YM]) / w[OFF_YS];
  const float sc = fminf(fmaxf(y1, -10.f), 5.f);
  const float inv = expf(-sc);
  const float r = (t - y0) * inv;
  loss[b] = 0.5f * r * r + sc;
  // d loss / d mu = -(t - mu)/sigma^2 ; d loss / d s = 1 - r^2 (zero when clamped)
  const float g0 = -r * inv / (float)B;
  const float g1 = (y1 > -10.f && y1 < 5.f) ? (1.0f - r * r) / (float)B : 0.0f;
  dY[b * NO] = g0;
  dY[b * NO + 1] = g1;
  // dH2 = dY W3^T ; dZ2 = dH2 * gelu'(z2)
  float dz2[NH];
  for (int o = 0; o < NH; ++o) {
    const float dh = g0 * w[OFF_W3 + o * NO] + g1 * w[OFF_W3 + o * NO + 1];
    dz2[o] = dh * gelu_grad(z2[o]);
    dZ2[b * NH + o] = dz2[o];
  }
  // dH1 = dZ2 W2^T ; dZ1 = dH1 * gelu'(z1)
  for (int i = 0; i < NH; ++i) {
    float dh = 0.0f;
#pragma unroll
    for (int o = 0; o < NH; ++o) dh += dz2[o] * w[OFF_W2 + i * NH + o];
    dZ1[b * NH + i] = dh * gelu_grad(z1[i]);
  }
}

// grad[W(i,o)] = sum_b A[b][i] * G[b][o], grad[bias(o)] = sum_b G[b][o]
// (fixed summation order over b).  One thread per weight entry.
__global__ void mlp_wgrad_kernel(const float* __restrict__ A, int na, const float* __restrict__ G,
                                 int ng, int64_t B, float* __restrict__ gW, float* __restrict__ gb) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < na * ng) {
    const int i = e / ng, o = e % ng;
    float s = 0.0f;
    for (int64_t b = 0; b < B; ++b) s += A[b * na + i] * G[b * ng + o];
    gW[e] = s;
  } else if (e < na * ng + ng) {
    const int o = e - na * ng;
    float s = 0.0f;
    for (int64_t b = 0; b < B; ++b) s += G[b * ng + o];
    gb[o] = s;
  }
}

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ mo,
                            float* __restrict__ ve, int n, float lr, float b1, float b2, float eps,
                            float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float m = b1 * mo[i] + (1.0f - b1) * gi;
  const float v = b2 * ve[i] + (1.0f - b2) * gi * gi;
  mo[i] = m;
  ve[i] = v;
  p[i] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

// ---- launchers -------------------------------------------------------------------
hipError_t launch_features(hipStream_t st, int64_t n, int64_t m, const double* L, int8_t* bins,
                           const double* P, int64_t ldp, double C_a, double target, float* X) {
  hipLaunchKernelGGL(eig_bins_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, L, bins);
  hipLaunchKernelGGL(features_kernel, dim3((unsigned)m), dim3(256), 0, st, n, m, L, bins, P, ldp,
                     C_a, target, X);
  return hipGetLastError();
}

hipError_t launch_mlp_infer(hipStream_t st, const float* params, const float* X, int64_t N,
                            float zsig, float* out, int as_bracket) {
  if (N <= 0) return hipSuccess;
  hipLaunchKernelGGL(mlp_infer_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, params,
                     X, N, zsig, out, as_bracket);
  return hipGetLastError();
}

// Gradient of the mean loss over samples perm[off .. off+B) into grad
// (N_TRAIN floats, parameter layout); per-sample losses into loss[0..B).
hipError_t mlp_gradient(hipStream_t st, const float* params, const float* X, const float* Y,
                        const int* perm, int64_t off, int64_t B, MlpScratch& s, float* grad,
                        float* loss) {
  hipLaunchKernelGGL(mlp_fwdbwd_kernel, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, st, params,
                     X, Y, perm, off, B, s.Xn, s.H1, s.dZ1, s.H2, s.dZ2, s.dY, loss);
  auto wg = [&](const float* A, int na, const float* G, int ng, int offW, int offb) {
    const int tot = na * ng + ng;
    hipLaunchKernelGGL(mlp_wgrad_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, A, na,
                       G, ng, B, grad + offW, grad + offb);
  };
  wg(s.Xn, NF, s.dZ1, NH, OFF_W1, OFF_B1);
  wg(s.H1, NH, s.dZ2, NH, OFF_W2, OFF_B2);
  wg(s.H2, NH, s.dY, NO, OFF_W3, OFF_B3);
  return hipGetLastError();
}

hipError_t launch_adam(hipStream_t st, float* p, const float* g, float* m, float* v, int n,
                       float lr, int step) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, g, m, v, n,
                     lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

}  // namespace imcom
