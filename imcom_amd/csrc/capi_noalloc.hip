// imcom_gpu_bisect_kappa_learned_noalloc (declared in include/imcom_gpu.h):
// the learned-bracket bisection for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops).  No reference
// counterpart; the arithmetic is imcom_gpu_bisect_kappa_learned's (the same
// three kernels), but the caller passes device pointers for every input and
// output and the kernels' scratch -- per-pixel features, predicted brackets
// and the eigenvalue decade of each input -- lives in one file-scope device
// array instead of context buffers.
//
// The array is one per GPU and is shared by every call on that GPU: calls
// that overlap on one GPU (two streams, two threads) race on it.  Each call
// overwrites all of the scratch it reads (features of its m pixels, their
// brackets, the bins of its n eigenvalues), so nothing needs clearing.
#include "ctx.h"

#include <mutex>

namespace {

constexpr int64_t kNoallocMaxN = 32768;  // input pixels
constexpr int64_t kNoallocMaxM = 16384;  // output pixels
constexpr size_t kFeatBytes = (size_t)kNoallocMaxM * 32 * sizeof(float);
constexpr size_t kBracketBytes = (size_t)kNoallocMaxM * 2 * sizeof(float);
constexpr size_t kBinBytes = (size_t)kNoallocMaxN;
constexpr size_t kScratchBytes = kFeatBytes + kBracketBytes + kBinBytes;  // 2.16 MiB

// Host code must not take this array's address directly: the name is a
// host-side stand-in.  Its device address comes from hipGetSymbolAddress,
// per device, at imcom_gpu_init.
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

}  // namespace

// Called by imcom_gpu_init with `device` current: records this GPU's address
// of the scratch array (once per device).
hipError_t noalloc_scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (!p) return hipErrorInvalidSymbol;
  g_scratch[device] = p;
  return hipSuccess;
}

static void* scratch_of(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  return g_scratch[device];
}

extern "C" int imcom_gpu_bisect_kappa_learned_noalloc(
    imcom_gpu_ctx* ctx, void* stream, int64_t n, int64_t m, const double* L, const double* P,
    int64_t ldp, const double* Ndiag, double C_a, const float* params, int UnotS, double US_max,
    double k_min, double k_max, double tol, int maxnits, float zsig, double* K, double* US,
    int32_t* evals, int32_t* hits) {
  if (!ctx) return IMCOM_EARG;
  NEED(n >= 1 && m >= 1, IMCOM_EARG, "bisect_kappa_learned_noalloc: n and m must be >= 1");
  NEED(n <= kNoallocMaxN && m <= kNoallocMaxM, IMCOM_EARG,
       "bisect_kappa_learned_noalloc: n <= 32768 and m <= 16384 (the size of its fixed "
       "scratch); use imcom_gpu_bisect_kappa_learned for larger tiles");
  NEED(ldp >= n, IMCOM_EARG, "bisect_kappa_learned_noalloc: ldp must be >= n");
  NEED(L && P && params && K && US, IMCOM_EARG,
       "bisect_kappa_learned_noalloc: NULL L, P, params, K or US");
  NEED(UnotS || Ndiag, IMCOM_EARG, "bisect_kappa_learned_noalloc: S mode needs Ndiag");
  NEED(C_a > 0.0 && k_min > 0.0 && k_max > k_min && US_max > 0.0 && zsig > 0.0f && maxnits >= 0,
       IMCOM_EARG, "bisect_kappa_learned_noalloc: bad arguments");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  unsigned char* s = static_cast<unsigned char*>(scratch_of(ctx->device));
  NEED(s != nullptr, IMCOM_ESTATE,
       "bisect_kappa_learned_noalloc: scratch address not set up for this device");
  float* feat = reinterpret_cast<float*>(s);
  float* bracket = reinterpret_cast<float*>(s + kFeatBytes);
  int8_t* bins = reinterpret_cast<int8_t*>(s + kFeatBytes + kBracketBytes);
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  CK(launch_features(st, n, m, L, bins, P, ldp, C_a, US_max, feat), "features");
  CK(launch_mlp_infer(st, params, feat, m, zsig, bracket, 1), "mlp_infer");
  CK(launch_bisect(st, n, m, L, P, ldp, Ndiag, C_a, US_max, tol, k_min, k_max, maxnits,
                   UnotS ? 1 : 0, bracket, K, US, evals, hits),
     "bisect");
  return IMCOM_OK;
}
