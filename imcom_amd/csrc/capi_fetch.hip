// Asynchronous copies of stage outputs to host memory, so that the stage-file
// writes of the driver (A, Q, L, B, P, T: imcom_matrix.f90:275-278, 348-351,
// 672-677, 700-703, 748-749) run beside the later stages instead of between
// them.  The copies run on a second stream of the context, ordered after the
// work issued so far on the compute stream; a stage that overwrites a buffer
// with a copy in flight waits for it on the device (fetch_guard).
#include <hip/hip_runtime.h>

#include "ctx.h"

using namespace imcom;

namespace {

// The device buffer and byte count of stage output `what`, or an error
// message when that stage has not run.
const char* fetch_source(imcom_gpu_ctx* ctx, int what, const DevBuf** buf, size_t* bytes) {
  const size_t n = (size_t)ctx->n, m = (size_t)ctx->m;
  switch (what) {
    case IMCOM_FETCH_A:
      *buf = &ctx->A;
      *bytes = n * n * 8;
      return ctx->have_A ? nullptr : "fetch_async: A not built (or overwritten by eigen_A)";
    case IMCOM_FETCH_Q:
      *buf = &ctx->Q;
      *bytes = n * n * 8;
      return ctx->have_QL ? nullptr : "fetch_async: Q not built";
    case IMCOM_FETCH_L:
      *buf = &ctx->L;
      *bytes = n * 8;
      return ctx->have_QL ? nullptr : "fetch_async: L not built";
    case IMCOM_FETCH_B:
      *buf = &ctx->B;
      *bytes = n * m * 8;
      return ctx->have_B ? nullptr : "fetch_async: B not built";
    case IMCOM_FETCH_P:
      *buf = &ctx->P;
      *bytes = n * m * 8;
      return ctx->have_P ? nullptr : "fetch_async: P not built";
    case IMCOM_FETCH_T:
      *buf = &ctx->T;
      *bytes = n * m * 8;
      return ctx->have_T ? nullptr : "fetch_async: T not built";
    default:
      return "fetch_async: unknown output";
  }
}

}  // namespace

extern "C" {

int imcom_gpu_fetch_async(imcom_gpu_ctx* ctx, int what, double* host) {
  if (!ctx) return IMCOM_EARG;
  NEED(what >= 0 && what < IMCOM_FETCH_N, IMCOM_EARG, "fetch_async: unknown output");
  NEED(host != nullptr, IMCOM_EARG, "fetch_async: NULL host buffer");
  const DevBuf* buf = nullptr;
  size_t bytes = 0;
  if (const char* msg = fetch_source(ctx, what, &buf, &bytes)) return set_err(ctx, IMCOM_ESTATE, msg);
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  if (!ctx->copy_stream)
    CK(hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking), "copy stream");
  if (!ctx->fetch_ready)
    CK(hipEventCreateWithFlags(&ctx->fetch_ready, hipEventDisableTiming), "fetch event");
  if (!ctx->fetch_ev[what])
    CK(hipEventCreateWithFlags(&ctx->fetch_ev[what], hipEventDisableTiming), "fetch event");
  CK(hipEventRecord(ctx->fetch_ready, ctx->stream), "fetch_async record");
  CK(hipStreamWaitEvent(ctx->copy_stream, ctx->fetch_ready, 0), "fetch_async wait");
  CK(hipMemcpyAsync(host, buf->p, bytes, hipMemcpyDeviceToHost, ctx->copy_stream),
     "fetch_async copy");
  CK(hipEventRecord(ctx->fetch_ev[what], ctx->copy_stream), "fetch_async record");
  ctx->fetch_pending[what] = true;
  return IMCOM_OK;
}

int imcom_gpu_fetch_wait(imcom_gpu_ctx* ctx, int what) {
  if (!ctx) return IMCOM_EARG;
  NEED(what >= -1 && what < IMCOM_FETCH_N, IMCOM_EARG, "fetch_wait: unknown output");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  if (what < 0) {
    CK(fetch_wait_all(ctx), "fetch_wait");
    return IMCOM_OK;
  }
  if (ctx->fetch_pending[what]) {
    CK(hipEventSynchronize(ctx->fetch_ev[what]), "fetch_wait");
    ctx->fetch_pending[what] = false;
  }
  return IMCOM_OK;
}

int imcom_gpu_host_alloc(size_t bytes, void** ptr) {
  if (!ptr) return IMCOM_EARG;
  *ptr = nullptr;
  if (bytes == 0) return IMCOM_OK;
  const hipError_t e = hipHostMalloc(ptr, bytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    *ptr = nullptr;
    return e == hipErrorOutOfMemory ? IMCOM_ENOMEM : IMCOM_EHIP;
  }
  return IMCOM_OK;
}

int imcom_gpu_host_free(void* ptr) {
  if (!ptr) return IMCOM_OK;
  return hipHostFree(ptr) == hipSuccess ? IMCOM_OK : IMCOM_EHIP;
}

}  // extern "C"
