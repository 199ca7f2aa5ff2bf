// C ABI of batched tiles (declared in include/imcom_gpu.h): many independent
// tiles, each with its own n, m, tables and targets, through the whole solve
// path together.  No reference counterpart -- the reference solves one tile
// per process -- but every tile's result is what the single-tile entry points
// give for it.
//
// Execution: lookup_A of every tile, then the tridiagonal reduction of all
// tiles in lockstep (blockIdx.z = tile, one grouped GEMM per panel for the
// trailing updates), the divide-and-conquer of all tiles level by level (a
// launch plan fixed at batch_set; deflation counts stay on the device), the
// back-transform (grouped GEMMs per block), and per-tile lookup_B, P, bisection
// and T.  With use_graph the whole sequence is captured once per batch (and
// per precision / bracket setting) into one HIP graph and replayed: one
// hipGraphLaunch per solve.
#include "ctx.h"

#include <algorithm>

struct BatchTile {
  imcom_gpu_tile in{};
  DevBuf x, y, ex, Xa, Ya, Nd, I, Atab, Btab, A, Q, L, B, P, W, T, K, US, SU, evals, hits, H, feat,
      bracket, bins;
  EigWork ew;
  SyTile sy{};
};

struct BatchState {
  std::vector<BatchTile*> tiles;
  DevBuf d_sy;
  SyPlan plan;
  GroupedPlan pt;  // P and T GEMMs
  int launch_P = -1, launch_T = -1;
  int64_t nmax = 0;
  int nb = 32, nbb = 128;
  hipGraphExec_t graph = nullptr;
  int graph_key = -1;
  // split-precision workspace address baked into the graph's products at
  // capture; the single-tile entry points may reallocate that buffer
  const void* graph_split = nullptr;
  // the batch's own divide-and-conquer workspace and launch plan (the
  // single-tile path's plan changes with its tile size; a captured graph
  // keeps pointing at this one)
  DcWork* dc = nullptr;
  ~BatchState() {
    clear();
    dc_destroy(dc);
  }
  void drop_graphs() {
    if (graph) (void)hipGraphExecDestroy(graph);
    graph = nullptr;
    graph_key = -1;
    graph_split = nullptr;
  }
  void clear() {
    drop_graphs();
    for (BatchTile* t : tiles) {
      DevBuf* b[] = {&t->x, &t->y, &t->ex, &t->Xa, &t->Ya, &t->Nd, &t->I, &t->Atab, &t->Btab,
                     &t->A, &t->Q, &t->L, &t->B, &t->P, &t->W, &t->T, &t->K, &t->US, &t->SU,
                     &t->evals, &t->hits, &t->H, &t->feat, &t->bracket, &t->bins};
      for (DevBuf* p : b) p->release();
      t->ew.release();
      delete t;
    }
    tiles.clear();
    d_sy.release();
    plan.gp.release();
    pt.release();
  }
};

void imcom_batch_free(imcom_gpu_ctx* ctx) {
  delete ctx->batch;
  ctx->batch = nullptr;
}

static int check_tile(imcom_gpu_ctx* ctx, const imcom_gpu_tile& t, int idx) {
  const std::string w = "batch tile " + std::to_string(idx) + ": ";
  NEED(t.n >= 1 && t.m >= 1 && t.n <= 32768, IMCOM_EARG, w + "bad n or m");
  NEED(t.x_i && t.y_i && t.exp_i && t.X_a && t.Y_a && t.Ndiag && t.I_i, IMCOM_EARG, w + "NULL input");
  NEED(t.Alookup && t.Blookup, IMCOM_EARG, w + "NULL lookup table");
  NEED(t.C_a > 0.0 && t.kappa_min > 0.0 && t.kappa_max > t.kappa_min && t.US_max > 0.0,
       IMCOM_EARG, w + "bad C_a, kappa range or target");
  NEED(t.psfconst == 0 || t.psfconst == 1, IMCOM_EARG, w + "PSFCONST must be 1 or 0");
  NEED(t.npoly >= 1 && t.npoly <= 8 && t.npad >= 1 && t.n1 >= 1 && t.n2 >= 1, IMCOM_EARG,
       w + "bad interpolation parameters");
  const int ne = t.psfconst == 1 ? 1 : t.nexppsf;
  for (int64_t i = 0; i < t.n; ++i) {
    NEED(t.exp_i[i] >= 1 && (i == 0 || t.exp_i[i] >= t.exp_i[i - 1]), IMCOM_EARG,
         w + "exp_i must be 1-based and non-decreasing");
    NEED(t.psfconst == 1 || t.exp_i[i] <= ne, IMCOM_EARG, w + "exposure id beyond the tables");
  }
  return IMCOM_OK;
}

extern "C" {

int imcom_gpu_batch_set(imcom_gpu_ctx* ctx, int ntiles, const imcom_gpu_tile* tiles) {
  if (!ctx) return IMCOM_EARG;
  NEED(ntiles >= 1 && tiles, IMCOM_EARG, "batch_set: no tiles");
  for (int i = 0; i < ntiles; ++i) {
    int rc = check_tile(ctx, tiles[i], i);
    if (rc) return rc;
  }
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  CK(interp_init(ctx->device), "interpolation weights");
  if (!ctx->batch) ctx->batch = new BatchState();
  BatchState* B = ctx->batch;
  B->clear();
  B->nb = ctx->nb;
  B->nbb = ctx->nbb;
  B->nmax = 0;
  std::vector<SyTile> host;
  for (int i = 0; i < ntiles; ++i) {
    const imcom_gpu_tile& s = tiles[i];
    BatchTile* t = new BatchTile();
    B->tiles.push_back(t);
    t->in = s;
    const int64_t n = s.n, m = s.m;
    const size_t ts = (size_t)s.n1 * s.n2 * 8;
    const int nA = s.psfconst == 1 ? 1 : s.nexppsf * (s.nexppsf + 1) / 2;
    const int nB = s.psfconst == 1 ? 1 : s.nexppsf;
    CK(upload(ctx, t->x, s.x_i, n * 8), "upload x_i");
    CK(upload(ctx, t->y, s.y_i, n * 8), "upload y_i");
    CK(upload(ctx, t->ex, s.exp_i, n * 4), "upload exp_i");
    CK(upload(ctx, t->Xa, s.X_a, m * 8), "upload X_a");
    CK(upload(ctx, t->Ya, s.Y_a, m * 8), "upload Y_a");
    CK(upload(ctx, t->Nd, s.Ndiag, n * 8), "upload Ndiag");
    CK(upload(ctx, t->I, s.I_i, n * 8), "upload I_i");
    CK(upload(ctx, t->Atab, s.Alookup, ts * nA), "upload Alookup");
    CK(upload(ctx, t->Btab, s.Blookup, ts * nB), "upload Blookup");
    CK(t->A.reserve((size_t)n * n * 8), "alloc A");
    CK(t->Q.reserve((size_t)n * n * 8), "alloc Q");
    CK(t->L.reserve(n * 8), "alloc L");
    CK(t->B.reserve((size_t)n * m * 8), "alloc B");
    CK(t->P.reserve((size_t)n * m * 8), "alloc P");
    CK(t->W.reserve((size_t)n * m * 8), "alloc W");
    CK(t->T.reserve((size_t)n * m * 8), "alloc T");
    DevBuf* small[] = {&t->K, &t->US, &t->SU, &t->H};
    for (DevBuf* b : small) CK(b->reserve(m * 8), "alloc per-pixel outputs");
    CK(t->evals.reserve(m * 4), "alloc evals");
    CK(t->hits.reserve(m * 4), "alloc hits");
    CK(t->feat.reserve((size_t)m * 32 * 4), "alloc features");
    CK(t->bracket.reserve((size_t)m * 2 * 4), "alloc bracket");
    CK(t->bins.reserve(n), "alloc bins");
    CK(t->ew.reserve(n, B->nb, B->nbb, t->A.as<double>(), t->Q.as<double>(), t->sy, false),
       "alloc eig");
    host.push_back(t->sy);
    B->nmax = std::max(B->nmax, n);
  }
  CK(upload(ctx, B->d_sy, host.data(), host.size() * sizeof(SyTile)), "upload tile descriptors");
  {
    if (!B->dc) B->dc = dc_create();
    std::vector<int64_t> ns;
    std::vector<double*> zp;
    for (BatchTile* t : B->tiles) {
      ns.push_back(t->in.n);
      zp.push_back(t->ew.sZ.as<double>());
    }
    CK(dc_prepare(ctx->stream, B->dc, (int)ns.size(), ns.data(), zp.data()),
       "divide-and-conquer plan");
  }
  sy_make_plan(host, B->nmax, B->nb, B->nbb, B->plan);
  CK(B->plan.gp.upload(ctx->stream), "upload reduction plan");
  std::vector<GemmDesc> gP, gT;
  for (BatchTile* t : B->tiles) {
    const int n = (int)t->in.n, m = (int)t->in.m;
    gP.push_back(GemmDesc{n, m, n, 1.0, 0.0, t->Q.as<double>(), n, t->B.as<double>(), n,
                          t->P.as<double>(), n});
    gT.push_back(GemmDesc{n, m, n, -0.5, 0.0, t->Q.as<double>(), n, t->W.as<double>(), n,
                          t->T.as<double>(), n});
  }
  B->launch_P = B->pt.add(true, false, gP);
  B->launch_T = B->pt.add(false, false, gT);
  CK(B->pt.upload(ctx->stream), "upload P/T plan");
  CK(hipStreamSynchronize(ctx->stream), "batch_set sync");
  return IMCOM_OK;
}

// lookup_A of every tile, then the reduction to tridiagonal form
static hipError_t batch_seq1(imcom_gpu_ctx* ctx) {
  BatchState* B = ctx->batch;
  hipStream_t st = ctx->stream;
  for (BatchTile* t : B->tiles) {
    const imcom_gpu_tile& s = t->in;
    const int np = s.npoly % 2 == 0 ? s.npoly - 1 : s.npoly;
    IMCOM_HIP_TRY(launch_lookup_A(st, ctx->device, s.n, t->x.as<double>(), t->y.as<double>(),
                                  t->ex.as<int32_t>(), s.psfconst, t->Atab.as<double>(),
                                  (int64_t)s.n1 * s.n2, s.n1, s.n2, np / 2, s.npad, s.psfxscale,
                                  s.psfyscale, t->A.as<double>()));
  }
  return sytrd_run_batched(st, B->d_sy.as<SyTile>(), (int)B->tiles.size(), B->nmax, B->nb, B->plan,
                           live_contexts(ctx->device), nullptr);
}

// descending reorder, back-transform, lookup_B, P, bisection, T and maps
static hipError_t batch_gemm(imcom_gpu_ctx* ctx, int precision, bool isP) {
  BatchState* B = ctx->batch;
  if (precision == IMCOM_PREC_FP64) return B->pt.run(ctx->stream, isP ? B->launch_P : B->launch_T);
  for (BatchTile* t : B->tiles) {
    const int n = (int)t->in.n, m = (int)t->in.m;
    if (isP)
      IMCOM_HIP_TRY(gemm_split(ctx->stream, precision, true, n, m, n, 1.0, t->Q.as<double>(), n,
                               t->B.as<double>(), n, t->P.as<double>(), n, ctx->split_work.p,
                               ctx->split_work.cap));
    else
      IMCOM_HIP_TRY(gemm_split(ctx->stream, precision, false, n, m, n, -0.5, t->Q.as<double>(), n,
                               t->W.as<double>(), n, t->T.as<double>(), n, ctx->split_work.p,
                               ctx->split_work.cap));
  }
  return hipSuccess;
}

static hipError_t batch_seq2(imcom_gpu_ctx* ctx, int prec_P, int prec_T, int learned,
                             float zsig) {
  BatchState* B = ctx->batch;
  hipStream_t st = ctx->stream;
  for (BatchTile* t : B->tiles)
    IMCOM_HIP_TRY(launch_reverse_cols(st, t->in.n, t->ew.sZ.as<double>(), t->Q.as<double>(),
                                      t->ew.slam.as<double>(), t->L.as<double>()));
  IMCOM_HIP_TRY(ormtr_run_batched(st, B->d_sy.as<SyTile>(), (int)B->tiles.size(), B->nmax, B->nbb,
                                  B->plan));
  for (BatchTile* t : B->tiles) {
    const imcom_gpu_tile& s = t->in;
    const int np = s.npoly % 2 == 0 ? s.npoly - 1 : s.npoly;
    IMCOM_HIP_TRY(launch_lookup_B(st, ctx->device, s.n, s.m, t->x.as<double>(), t->y.as<double>(),
                                  t->ex.as<int32_t>(), t->Xa.as<double>(), t->Ya.as<double>(),
                                  s.psfconst, t->Btab.as<double>(), (int64_t)s.n1 * s.n2, s.n1,
                                  s.n2, np / 2, s.npad, s.psfxscale, s.psfyscale,
                                  t->B.as<double>()));
  }
  IMCOM_HIP_TRY(batch_gemm(ctx, prec_P, true));
  for (BatchTile* t : B->tiles) {
    const imcom_gpu_tile& s = t->in;
    const float* br = nullptr;
    if (learned) {
      IMCOM_HIP_TRY(launch_features(st, s.n, s.m, t->L.as<double>(), t->bins.as<int8_t>(),
                                    t->P.as<double>(), s.n, s.C_a, s.US_max, t->feat.as<float>()));
      IMCOM_HIP_TRY(launch_mlp_infer(st, ctx->mlp_params.as<float>(), t->feat.as<float>(), s.m, zsig,
                                     t->bracket.as<float>(), 1));
      br = t->bracket.as<float>();
    }
    IMCOM_HIP_TRY(launch_bisect(st, s.n, s.m, t->L.as<double>(), t->P.as<double>(), s.n,
                                t->Nd.as<double>(), s.C_a, s.US_max, s.US_tol, s.kappa_min,
                                s.kappa_max, s.maxnits, s.UnotS ? 1 : 0, br, t->K.as<double>(),
                                t->US.as<double>(), t->evals.as<int32_t>(), t->hits.as<int32_t>()));
    IMCOM_HIP_TRY(launch_scale_W(st, s.n, s.m, t->P.as<double>(), s.n, t->L.as<double>(),
                                 t->K.as<double>(), t->W.as<double>(), s.n));
  }
  IMCOM_HIP_TRY(batch_gemm(ctx, prec_T, false));
  for (BatchTile* t : B->tiles) {
    const imcom_gpu_tile& s = t->in;
    if (s.UnotS)
      IMCOM_HIP_TRY(launch_maps(st, s.n, s.m, t->T.as<double>(), s.n, t->I.as<double>(),
                                t->Nd.as<double>(), t->H.as<double>(), t->SU.as<double>()));
    else {
      IMCOM_HIP_TRY(launch_maps(st, s.n, s.m, t->T.as<double>(), s.n, t->I.as<double>(), nullptr,
                                t->H.as<double>(), nullptr));
      IMCOM_HIP_TRY(launch_U_from_K(st, s.n, s.m, t->L.as<double>(), t->P.as<double>(), s.n, s.C_a,
                                    t->K.as<double>(), t->SU.as<double>()));
    }
  }
  return hipGetLastError();
}

// the divide-and-conquer of every tile's tridiagonal, level by level
static hipError_t batch_dc(imcom_gpu_ctx* ctx) {
  BatchState* B = ctx->batch;
  std::vector<int64_t> ns;
  std::vector<const double*> dp, ep;
  std::vector<double*> lp, zp;
  for (BatchTile* t : B->tiles) {
    ns.push_back(t->in.n);
    dp.push_back(t->ew.sd.as<double>());
    ep.push_back(t->ew.se.as<double>());
    lp.push_back(t->ew.slam.as<double>());
    zp.push_back(t->ew.sZ.as<double>());
  }
  return dc_solve_batched(ctx->stream, B->dc, (int)ns.size(), ns.data(), dp.data(), ep.data(),
                          lp.data(), zp.data());
}

// the whole solve: reduction, divide-and-conquer, back-transform to the maps
static hipError_t batch_all(imcom_gpu_ctx* ctx, int prec_P, int prec_T, int learned, float zsig) {
  IMCOM_HIP_TRY(batch_seq1(ctx));
  IMCOM_HIP_TRY(batch_dc(ctx));
  return batch_seq2(ctx, prec_P, prec_T, learned, zsig);
}

static hipError_t run_graph(imcom_gpu_ctx* ctx, hipGraphExec_t* exec, int prec_P, int prec_T,
                            int learned, float zsig) {
  hipStream_t st = ctx->stream;
  if (!*exec) {
    hipGraph_t g = nullptr;
    IMCOM_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    hipError_t e = batch_all(ctx, prec_P, prec_T, learned, zsig);
    hipError_t e2 = hipStreamEndCapture(st, &g);
    if (e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return e;
    }
    IMCOM_HIP_TRY(e2);
    e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    IMCOM_HIP_TRY(e);
  }
  return hipGraphLaunch(*exec, st);
}

// The split-precision workspace a batched solve with these precisions needs.
static int batch_reserve_split(imcom_gpu_ctx* ctx, int precision_P, int precision_T) {
  if (precision_P == IMCOM_PREC_FP64 && precision_T == IMCOM_PREC_FP64) return IMCOM_OK;
  size_t wb = 0;
  for (BatchTile* t : ctx->batch->tiles) {
    const int n = (int)t->in.n, m = (int)t->in.m;
    wb = std::max(wb, gemm_split_work_bytes(std::max(precision_P, precision_T), n, m, n));
    if (precision_P) wb = std::max(wb, gemm_split_work_bytes(precision_P, n, m, n));
    if (precision_T) wb = std::max(wb, gemm_split_work_bytes(precision_T, n, m, n));
  }
  if (ctx->split_work.cap < wb) CK(ctx->split_work.reserve(wb), "alloc split workspace");
  return IMCOM_OK;
}

int imcom_gpu_batch_reserve(imcom_gpu_ctx* ctx, int precision_P, int precision_T) {
  if (!ctx) return IMCOM_EARG;
  NEED(precision_P >= 0 && precision_P <= 2 && precision_T >= 0 && precision_T <= 2, IMCOM_EARG,
       "batch_reserve: precision must be IMCOM_PREC_FP64, _FP32X3 or _BF16X6");
  NEED(ctx->batch && !ctx->batch->tiles.empty(), IMCOM_ESTATE, "batch_reserve: no batch set");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  return batch_reserve_split(ctx, precision_P, precision_T);
}

int imcom_gpu_batch_solve(imcom_gpu_ctx* ctx, int precision_P, int precision_T, int learned,
                          float zsig, int use_graph) {
  if (!ctx) return IMCOM_EARG;
  NEED(precision_P >= 0 && precision_P <= 2 && precision_T >= 0 && precision_T <= 2, IMCOM_EARG,
       "batch_solve: precision must be IMCOM_PREC_FP64, _FP32X3 or _BF16X6");
  NEED(ctx->batch && !ctx->batch->tiles.empty(), IMCOM_ESTATE, "batch_solve: no batch set");
  NEED(!learned || ctx->have_mlp, IMCOM_ESTATE, "batch_solve: learned bracket needs parameters");
  NEED(!learned || zsig > 0.0f, IMCOM_EARG, "batch_solve: zsig must be positive");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  BatchState* B = ctx->batch;
  {
    int rc = batch_reserve_split(ctx, precision_P, precision_T);
    if (rc) return rc;
  }
  const int key = (learned ? 1 + (int)(zsig * 1000.0f) : 0) * 16 + precision_P * 4 + precision_T;
  // re-capture when the controls change or when the split workspace moved
  // (grown here, or reallocated by imcom_gpu_build_P / _build_T since)
  if (B->graph && (B->graph_key != key || B->graph_split != ctx->split_work.p)) {
    (void)hipGraphExecDestroy(B->graph);
    B->graph = nullptr;
  }
  B->graph_key = key;
  B->graph_split = ctx->split_work.p;
  if (use_graph) {
    // one graph: the whole solve lands in the reduction's time slot
    StageTimer tm(ctx, PS_SYTRD);
    CK(run_graph(ctx, &B->graph, precision_P, precision_T, learned, zsig), "batch solve (graph)");
  } else {
    {
      StageTimer tm(ctx, PS_SYTRD);
      CK(batch_seq1(ctx), "batch reduction");
    }
    {
      StageTimer tm(ctx, PS_STEDC);
      CK(batch_dc(ctx), "batch divide-and-conquer");
    }
    {
      StageTimer tm(ctx, PS_ORMTR);
      CK(batch_seq2(ctx, precision_P, precision_T, learned, zsig), "batch back-transform to T");
    }
  }
  CK(hipStreamSynchronize(ctx->stream), "batch_solve sync");
  return IMCOM_OK;
}

int imcom_gpu_batch_result(imcom_gpu_ctx* ctx, int tile, double* K_out, double* US_out,
                           double* SU_out, double* H_out, int32_t* evals_out, int32_t* hits_out,
                           double* T_out) {
  if (!ctx) return IMCOM_EARG;
  NEED(ctx->batch && tile >= 0 && tile < (int)ctx->batch->tiles.size(), IMCOM_EARG,
       "batch_result: no such tile");
  CK(hipSetDevice(ctx->device), "hipSetDevice");
  BatchTile* t = ctx->batch->tiles[tile];
  const int64_t n = t->in.n, m = t->in.m;
  CK(download(ctx, K_out, t->K, m * 8), "download K");
  CK(download(ctx, US_out, t->US, m * 8), "download U/S");
  CK(download(ctx, SU_out, t->SU, m * 8), "download S/U");
  CK(download(ctx, H_out, t->H, m * 8), "download H");
  CK(download(ctx, evals_out, t->evals, m * 4), "download evals");
  CK(download(ctx, hits_out, t->hits, m * 4), "download hits");
  CK(download(ctx, T_out, t->T, (size_t)n * m * 8), "download T");
  return IMCOM_OK;
}

}  // extern "C"

extern "C" int imcom_gpu_set_reduction_tail(imcom_gpu_ctx* ctx, int max_rows) {
  if (!ctx) return IMCOM_EARG;
  NEED(max_rows == 0 || (max_rows >= 2 && max_rows <= kTailMaxRows), IMCOM_EARG,
       "set_reduction_tail: max_rows must be 0 or in 2.." + std::to_string(kTailMaxRows));
  ctx->tail_rows = max_rows;
  return IMCOM_OK;
}

extern "C" int imcom_gpu_set_tuning(imcom_gpu_ctx* ctx, int nb, int nbb) {
  if (!ctx) return IMCOM_EARG;
  NEED(nb == 16 || nb == 32 || nb == 64, IMCOM_EARG, "set_tuning: nb must be 16, 32 or 64");
  NEED(nbb == 64 || nbb == 128 || nbb == 256 || nbb == 512, IMCOM_EARG,
       "set_tuning: nbb must be 64, 128, 256 or 512");
  ctx->nb = nb;
  ctx->nbb = nbb;
  return IMCOM_OK;
}
