// Context of the C ABI (shared by the capi*.hip translation units).
#pragma once
#include "../../include/imcom_gpu.h"
#include "common.h"
#include "kernels.h"

#include <algorithm>
#include <string>
#include <vector>

using namespace imcom;

namespace imcom_ctx_detail {

constexpr int kNB = 32;    // tridiagonal-reduction panel width
constexpr int kNBB = 128;  // back-transform block

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  hipError_t reserve(size_t bytes) {
    if (bytes <= cap && p) return hipSuccess;
    if (p) {
      hipError_t e = hipFree(p);
      p = nullptr;
      cap = 0;
      if (e != hipSuccess) return e;
    }
    if (bytes == 0) bytes = 8;
    hipError_t e = hipMalloc(&p, bytes);
    if (e == hipSuccess) cap = bytes;
    return e;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

}  // namespace imcom_ctx_detail
using imcom_ctx_detail::DevBuf;
using imcom_ctx_detail::kNB;
using imcom_ctx_detail::kNBB;

// Device workspace of one tile's eigendecomposition (tridiagonal reduction,
// divide-and-conquer output, back-transform), sized for n, panel width nb and
// back-transform block nbb.
struct EigWork {
  DevBuf sW, sVW, sWV, sy, sycp, swpre, svtmp, sd, se, stau, spart, spart2, ss, sscal, sVb, sSb,
      sTb, sY, sY2, slam, sZ, stile, sdp, sAt, sKs;
  // exchange space of the persistent reduction tail (sytrd_tail; single
  // tiles only, reserved by eigen_A / tridiagonalize)
  DevBuf stail;
  // splitk: workspace for the single-tile back-transform's split-K products
  // (the batched back-transform uses grouped GEMMs instead)
  hipError_t reserve(int64_t n, int nb, int nbb, double* A, double* Z, SyTile& t,
                     bool splitk = true) {
    const size_t d = 8;
    IMCOM_HIP_TRY(sW.reserve(n * nb * d));
    IMCOM_HIP_TRY(sVW.reserve(n * 2 * nb * d));
    IMCOM_HIP_TRY(sWV.reserve(n * 2 * nb * d));
    const int64_t nbr = (n + 63) / 64;
    IMCOM_HIP_TRY(sy.reserve((size_t)nbr * n * d));
    IMCOM_HIP_TRY(sycp.reserve((size_t)((nbr + 3) / 4) * n * d));
    IMCOM_HIP_TRY(swpre.reserve(n * d));
    IMCOM_HIP_TRY(svtmp.reserve(n * d));
    IMCOM_HIP_TRY(sd.reserve(n * d));
    IMCOM_HIP_TRY(se.reserve(n * d));
    IMCOM_HIP_TRY(stau.reserve(n * d));
    IMCOM_HIP_TRY(spart.reserve(2 * kNormParts * d));  // two buffers by column parity
    IMCOM_HIP_TRY(spart2.reserve(kQParts * d));
    IMCOM_HIP_TRY(ss.reserve(2 * nb * d));
    IMCOM_HIP_TRY(sdp.reserve((size_t)2 * nb * kDotChunks * d));
    IMCOM_HIP_TRY(sscal.reserve(8 * d));
    IMCOM_HIP_TRY(sVb.reserve((size_t)(n > 1 ? n - 1 : 1) * nbb * d));
    IMCOM_HIP_TRY(sSb.reserve((size_t)nbb * nbb * d));
    IMCOM_HIP_TRY(sTb.reserve((size_t)nbb * nbb * d));
    IMCOM_HIP_TRY(sY.reserve((size_t)nbb * n * d));
    IMCOM_HIP_TRY(sY2.reserve((size_t)nbb * n * d));
    IMCOM_HIP_TRY(slam.reserve(n * d));
    IMCOM_HIP_TRY(sZ.reserve((size_t)n * n * d));
    IMCOM_HIP_TRY(stile.reserve(sizeof(SyTile)));
    const int nbt = (int)((n + 63) / 64);
    IMCOM_HIP_TRY(sAt.reserve((size_t)nbt * (nbt + 1) / 2 * 4096 * d));
    if (splitk)
      IMCOM_HIP_TRY(sKs.reserve(std::max(dgemm_splitk_work(nbb, nbb, (int)n),
                                         dgemm_splitk_work(nbb, (int)n, (int)n))));
    t.n = n;
    t.A = A;
    t.At = sAt.as<double>();
    t.nbt = nbt;
    t.W = sW.as<double>();
    t.VW = sVW.as<double>();
    t.WV = sWV.as<double>();
    t.ypart = sy.as<double>();
    t.ycpart = sycp.as<double>();
    t.wpre = swpre.as<double>();
    t.vtmp = svtmp.as<double>();
    t.d = sd.as<double>();
    t.e = se.as<double>();
    t.tau = stau.as<double>();
    t.part = spart.as<double>();
    t.part2 = spart2.as<double>();
    t.qpart = spart2.as<double>();
    t.s = ss.as<double>();
    t.dpart = sdp.as<double>();
    t.scal = sscal.as<double>();
    t.Vb = sVb.as<double>();
    t.Sb = sSb.as<double>();
    t.Tb = sTb.as<double>();
    t.Y = sY.as<double>();
    t.Y2 = sY2.as<double>();
    t.Z = Z;
    t.Ks = sKs.as<double>();
    return hipSuccess;
  }
  void release() {
    DevBuf* b[] = {&sW, &sVW, &sWV, &sy, &sycp, &swpre, &svtmp, &sd, &se, &stau, &spart,
                   &spart2, &ss, &sscal, &sVb, &sSb, &sTb, &sY, &sY2, &slam, &sZ, &stile, &sdp,
                   &sAt, &sKs, &stail};
    for (DevBuf* p : b) p->release();
  }
};

// Workspace of the two-stage reduction (sytrd2.hip).
struct TwoStageWork {
  DevBuf U, X, Mt, Mm, T1, xch, kwork, Y1, Y2, q1w, sync, prog, msg;
  // the bulge chase's back-transform: block offsets per sweep group (host
  // and device, for n_blk) and each block's triangular factor
  DevBuf blk_off, Tq;
  std::vector<int> h_blk_off;
  int64_t n_blk = -1;
  hipError_t reserve(int64_t n, imcom::TwoStageBufs& w) {
    const size_t d = 8, b = imcom::kBand;
    IMCOM_HIP_TRY(U.reserve((size_t)n * 3 * b * d));
    IMCOM_HIP_TRY(X.reserve((size_t)n * b * d));
    IMCOM_HIP_TRY(Mt.reserve(b * b * d));
    IMCOM_HIP_TRY(Mm.reserve(b * b * d));
    IMCOM_HIP_TRY(T1.reserve((size_t)std::max(1, imcom::band_panels(n)) * b * b * d));
    IMCOM_HIP_TRY(xch.reserve(imcom::band_xch_doubles(n) * d));
    IMCOM_HIP_TRY(kwork.reserve(imcom::band_kwork_doubles(n) * d));
    IMCOM_HIP_TRY(Y1.reserve((size_t)n * 4 * b * d));
    IMCOM_HIP_TRY(Y2.reserve((size_t)n * 4 * b * d));
    IMCOM_HIP_TRY(q1w.reserve(imcom::q1_group_work_doubles(n) * d));
    IMCOM_HIP_TRY(sync.reserve(2 * 4));
    IMCOM_HIP_TRY(prog.reserve(imcom::bulge_flag_words(n) * 4));
    IMCOM_HIP_TRY(msg.reserve(imcom::bulge_msg_doubles(n) * d));
    w.U = U.as<double>();
    w.X = X.as<double>();
    w.Mt = Mt.as<double>();
    w.Mm = Mm.as<double>();
    w.T1 = T1.as<double>();
    w.xch = xch.as<double>();
    w.kwork = kwork.as<double>();
    w.Y1 = Y1.as<double>();
    w.Y2 = Y2.as<double>();
    w.q1work = q1w.as<double>();
    w.sync = sync.as<unsigned>();
    w.prog = prog.as<unsigned>();
    w.msg = msg.as<double>();
    return hipSuccess;
  }
  // block offsets and T factors of the bulge back-transform (uploaded on
  // the stream; the host copy stays alive)
  hipError_t reserve_q2(int64_t n, hipStream_t st) {
    if (n_blk != n) {
      imcom::q2_block_offsets(n, h_blk_off);
      IMCOM_HIP_TRY(blk_off.reserve(h_blk_off.size() * sizeof(int)));
      IMCOM_HIP_TRY(hipMemcpyAsync(blk_off.p, h_blk_off.data(), h_blk_off.size() * sizeof(int),
                                   hipMemcpyHostToDevice, st));
      IMCOM_HIP_TRY(hipStreamSynchronize(st));
      n_blk = n;
    }
    const size_t nblocks = h_blk_off.empty() ? 0 : (size_t)h_blk_off.back();
    return Tq.reserve(std::max<size_t>(nblocks, 1) * imcom::kQ2Image * 8);
  }
  void release() {
    DevBuf* b[] = {&U, &X, &Mt, &Mm, &T1, &xch, &kwork, &Y1, &Y2, &q1w, &sync, &prog, &msg,
                   &blk_off, &Tq};
    for (DevBuf* p : b) p->release();
    h_blk_off.clear();
    n_blk = -1;
  }
};

// Workspace of the device table set-up (tables.hip): the cached DFT
// matrices (E_f for the PSF spectra, E_t for the padded table transform,
// with the (N, n, f0) each was built for), the product spectrum, the
// half-transformed stack and the uploaded spectra.
struct TableWork {
  DevBuf E_f, E_t, Sp, X, img, spec, Y, Sr, Gt, Gam, Cv;
  int Ef_N = 0, Ef_n = 0, Ef_f0 = -1, Et_N = 0, Et_n = 0, Et_f0 = -1;
  void release() {
    DevBuf* b[] = {&E_f, &E_t, &Sp, &X, &img, &spec, &Y, &Sr, &Gt, &Gam, &Cv};
    for (DevBuf* p : b) p->release();
    Ef_N = Ef_n = Et_N = Et_n = 0;
    Ef_f0 = Et_f0 = -1;
  }
};

struct imcom_gpu_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int64_t n = 0, m = 0;
  double C_a = 0.0;
  int nb = kNB, nbb = kNBB;  // eigensolver tuning (imcom_gpu_set_tuning)
  // largest trailing size the persistent kernel reduces (imcom_gpu_set_reduction_tail;
  // 0: off)
  int tail_rows = imcom::kTailMaxRows;
  // 1: one-stage reduction; 2: two-stage (dense -> band -> tridiagonal) for
  // n >= twostage_min (imcom_gpu_set_reduction_stages).  By default the
  // two-stage reduction serves n >= 7168, where it was measured faster
  // (n = 16384: 1.52 against 1.63 s; n = 12288: 0.83 against 0.80 s).
  int stages = 2;
  int64_t twostage_min = 7168;
  TwoStageWork tsw;
  std::vector<int32_t> h_exp;
  // tables
  int psfconst = 1, nexppsf = 1, n1 = 0, n2 = 0;
  bool have_Atab = false, have_Btab = false;  // set_tables or build_Alookup / build_Blookup
  TableWork tw;
  // stage flags
  bool have_tile = false, have_A = false, have_B = false, have_QL = false, have_P = false,
       have_K = false, have_T = false;
  // inputs
  DevBuf x, y, ex, Xa, Ya, Nd, Atab, Btab, I;
  // regular input and output grids (imcom_gpu_set_grid): each pixel's grid
  // position, and the separation lattices of A's exposure pairs and of B's
  // exposures (lat_A / lat_B: the grids' pitches are in small integer ratios,
  // so lookup_A / lookup_B interpolate each separation once)
  bool have_grid = false, lat_A = false, lat_B = false;
  int grid_n1out = 0, lat_npA = 0, lat_npB = 0;
  int64_t lat_totA = 0, lat_totB = 0;
  DevBuf gkx, gky, latA_desc, latB_desc, lat_vals;
  // stages
  DevBuf A, B, Q, L, P, K, US, T, W, H, S, U, evals, bracket;
  // eigensolver workspace
  EigWork ew;
  DcWork* dc = nullptr;
  DevBuf split_work;
  Prof prof;
  // learned kappa bracket
  DevBuf mlp_params, mlp_feat, mlp_bracket, mlp_bins, mlp_hits, mlp_X, mlp_Y, mlp_perm, mlp_loss,
      mlp_grad, mlp_m, mlp_v, mlp_out;
  DevBuf ms_Xn, ms_H1, ms_dZ1, ms_H2, ms_dZ2, ms_dY;
  bool have_mlp = false;
  int mlp_step = 0;
  // batched tiles
  struct BatchState* batch = nullptr;
  // asynchronous fetches of stage outputs (capi_fetch.hip): the copy stream,
  // the event the copies start after, and per output the event after which
  // its device buffer may be overwritten
  hipStream_t copy_stream = nullptr;
  hipEvent_t fetch_ready = nullptr;
  hipEvent_t fetch_ev[IMCOM_FETCH_N] = {};
  bool fetch_pending[IMCOM_FETCH_N] = {};
};

// A stage about to write outputs with a fetch in flight (bits of
// 1 << IMCOM_FETCH_*) makes the compute stream wait for those copies; the
// host does not block.
inline hipError_t fetch_guard(imcom_gpu_ctx* ctx, unsigned mask) {
  for (int w = 0; w < IMCOM_FETCH_N; ++w) {
    if (!((mask >> w) & 1u) || !ctx->fetch_pending[w]) continue;
    const hipError_t e = hipStreamWaitEvent(ctx->stream, ctx->fetch_ev[w], 0);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Host-side wait for every fetch in flight (before buffers are reallocated
// or released).
inline hipError_t fetch_wait_all(imcom_gpu_ctx* ctx) {
  for (int w = 0; w < IMCOM_FETCH_N; ++w) {
    if (!ctx->fetch_pending[w]) continue;
    const hipError_t e = hipEventSynchronize(ctx->fetch_ev[w]);
    if (e != hipSuccess) return e;
    ctx->fetch_pending[w] = false;
  }
  return hipSuccess;
}

void imcom_batch_free(imcom_gpu_ctx* ctx);

// capi_noalloc.hip: records this device's address of the no-allocation entry
// point's scratch array (called by imcom_gpu_init, device current)
hipError_t noalloc_scratch_setup(int device);

// contexts open on the device (at least 1): the tridiagonal reduction gives
// each an equal share of the last-level cache
int live_contexts(int device);

// stage timing helper: events recorded only when profiling is on
struct StageTimer {
  imcom_gpu_ctx* c;
  int slot;
  size_t tok = 0;
  StageTimer(imcom_gpu_ctx* ctx_, int slot_) : c(ctx_), slot(slot_) {
    (void)c->prof.begin(slot, 1, &tok);
  }
  ~StageTimer() { (void)c->prof.end(slot, tok); }
};

inline int set_err(imcom_gpu_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

inline int hip_err(imcom_gpu_ctx* c, hipError_t e, const char* where) {
  if (e == hipSuccess) return IMCOM_OK;
  const int code = (e == hipErrorOutOfMemory) ? IMCOM_ENOMEM : IMCOM_EHIP;
  return set_err(c, code, std::string(where) + ": " + hipGetErrorString(e));
}

#define CK(expr, where)                                   \
  do {                                                    \
    hipError_t _e = (expr);                               \
    if (_e != hipSuccess) return hip_err(ctx, _e, where); \
  } while (0)

#define NEED(cond, code, msg) \
  do {                        \
    if (!(cond)) return set_err(ctx, code, msg); \
  } while (0)

inline hipError_t upload(imcom_gpu_ctx* ctx, DevBuf& b, const void* h, size_t bytes) {
  hipError_t e = b.reserve(bytes);
  if (e != hipSuccess) return e;
  if (bytes == 0) return hipSuccess;
  return hipMemcpyAsync(b.p, h, bytes, hipMemcpyHostToDevice, ctx->stream);
}

inline hipError_t download(imcom_gpu_ctx* ctx, void* h, const DevBuf& b, size_t bytes) {
  if (!h || bytes == 0) return hipSuccess;
  hipError_t e = hipMemcpyAsync(h, b.p, bytes, hipMemcpyDeviceToHost, ctx->stream);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(ctx->stream);
}

