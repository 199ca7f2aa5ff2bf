/*
 * imcom_gpu.h -- C ABI of the MI355X solve path of IMCOM.
 *
 * One context per GPU tile stream.  The context owns the device copies of the
 * tile's inputs and of every intermediate of the solve path (A, B, Q, L, P,
 * K, T); each entry point below replaces one procedure of the reference's
 * imcom_matrix / imcom_bisect modules and keeps its argument meaning.  All
 * matrices are Fortran column-major with leading dimension n; exposure ids
 * are 1-based; all host buffers are plain pointers.  A NULL output pointer
 * leaves that stage's result on the device only (the reference writes it to
 * a FITS cache only when the file name is not 'None').
 *
 * Every function returns IMCOM_OK (0) or an error code; the message is in
 * imcom_gpu_last_error().  Calls on one context must come from one host
 * thread at a time; the context's work runs in order on its own HIP stream
 * and every call that returns host data synchronises.
 */
#ifndef IMCOM_GPU_H
#define IMCOM_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct imcom_gpu_ctx imcom_gpu_ctx;

enum {
  IMCOM_OK = 0,
  IMCOM_EARG = 1,   /* bad argument (sizes, ids, NULL where required) */
  IMCOM_EHIP = 2,   /* HIP runtime error */
  IMCOM_ESTATE = 3, /* a stage was called before the stages it needs */
  IMCOM_ENOMEM = 4  /* device allocation failed */
};

/* arithmetic of the Q^T B and Q W products */
enum {
  IMCOM_PREC_FP64 = 0,   /* FP64 matrix cores */
  IMCOM_PREC_FP32X3 = 1, /* split FP32 x3 on FP32 MFMA, FP64 recombination */
  IMCOM_PREC_BF16X6 = 2  /* split BF16 x6 on BF16 MFMA, FP64 recombination */
};

/* stage outputs for imcom_gpu_fetch_async */
enum {
  IMCOM_FETCH_A = 0, /* A_aij (n x n) */
  IMCOM_FETCH_Q = 1, /* Q_ij (n x n) */
  IMCOM_FETCH_L = 2, /* L_i (n) */
  IMCOM_FETCH_B = 3, /* B_ia (n x m) */
  IMCOM_FETCH_P = 4, /* P_ia (n x m) */
  IMCOM_FETCH_T = 5, /* T_ia (n x m) */
  IMCOM_FETCH_N = 6
};

/* Context lifetime.  No reference counterpart (the reference keeps module
 * globals, imcom_data.f90:24-132). */
int imcom_gpu_init(imcom_gpu_ctx** ctx, int device);
void imcom_gpu_finalize(imcom_gpu_ctx* ctx);
const char* imcom_gpu_last_error(const imcom_gpu_ctx* ctx);
int imcom_gpu_synchronize(imcom_gpu_ctx* ctx);

/* Tile inputs: the imcom_data globals the get_* procedures read
 * (x_i, y_i, exp_i, X_a, Y_a, Ndiag, C_a; imcom_data.f90:95-132).  exp_i must
 * be exposure-major (non-decreasing), as imcom_alloc_gims builds it
 * (imcom_io.f90:84-123). */
int imcom_gpu_set_tile(imcom_gpu_ctx* ctx, int64_t n, int64_t m, const double* x_i,
                       const double* y_i, const int32_t* exp_i, const double* X_a,
                       const double* Y_a, const double* Ndiag, double C_a);

/* The regular grids of a USERXY=0 tile, as the imcom_data globals hold them:
 * input pixel (k1, k2) of exposure e (0-based k) at
 * x = k1 * gimxscale(e) + dither(e, 1), y = k2 * gimyscale(e) + dither(e, 2)
 * (imcom_build_xy, imcom_proc.f90:56-89; dither is (nexp, 2) column-major),
 * mask_subs_i (n entries, 1-based into the (n1gim, n2gim, nexp) order) the
 * unmasked pixels, and output pixel (a1, a2) at a1 * outxscale + outpos1,
 * a2 * outyscale + outpos2 (imcom_build_XaYa, imcom_proc.f90:93-119).  Call
 * after imcom_gpu_set_tile; every pixel's coordinates are checked against the
 * grids (IMCOM_EARG if one differs by more than 1e-12 relative).  With grids
 * whose pitches are in small integer ratios, imcom_gpu_lookup_A / _B
 * interpolate each distinct pixel separation once and write the matrices from
 * those values (the same as the per-entry lookups to the rounding of the
 * separations).  n1gim = 0 clears the grids; imcom_gpu_set_tile clears them
 * too.  No reference counterpart: imcom_lookup_A/B (imcom_matrix.f90:206-357)
 * read only x_i, y_i, X_a, Y_a. */
int imcom_gpu_set_grid(imcom_gpu_ctx* ctx, int n1gim, int n2gim, int nexp, const double* gimxscale,
                       const double* gimyscale, const double* dither, const int32_t* mask_subs_i,
                       int n1out, int n2out, double outxscale, double outyscale, double outpos1,
                       double outpos2);
/* Whether imcom_gpu_lookup_A / _B take the grid lattices (1) or evaluate
 * every entry (0). */
int imcom_gpu_lattice_state(imcom_gpu_ctx* ctx, int* A_lattice, int* B_lattice);

/* Lookup tables built on the host by imcom_build_Alookup / imcom_build_Blookup
 * (imcom_matrix.f90:361-493): Alookup is (n1, n2, nexppsf*(nexppsf+1)/2),
 * Blookup is (n1, n2, nexppsf), psfconst as imcom_test_psfconst returns it. */
int imcom_gpu_set_tables(imcom_gpu_ctx* ctx, int psfconst, int nexppsf, int n1, int n2,
                         const double* Alookup, const double* Blookup);

/* Device set-up of the tile: input vectors, PSF rotation, spectra and lookup
 * tables (the CPU steps of imcom.f90:55-116 and imcom_build_Alookup /
 * imcom_build_Blookup).  Spectra are Fortran complex(8)
 * arrays (interleaved re, im), n1psf x n2psf each, stacked over exposures;
 * the PSF grid is square with an even side, as the reference pads it
 * (imcom_io.f90:279-286).  The tables stay on the device for
 * imcom_gpu_lookup_A / _B (in place of imcom_gpu_set_tables); the host copies
 * are optional. */

/* Masked input vectors for USERXY=0 regular grids: the masking of
 * imcom_alloc_gims (imcom_io.f90:84-123) and imcom_build_xy
 * (imcom_proc.f90:56-89).  Im_unmasked is (n1gim, n2gim, nexp); pixels with
 * I < saturation, in that column-major order, give x_i, y_i, exp_i (1-based)
 * and I_i; dither is (nexp, 2).  *n_out receives n; each output must hold
 * n1gim * n2gim * nexp entries. */
int imcom_gpu_assemble_inputs(imcom_gpu_ctx* ctx, int n1gim, int n2gim, int nexp,
                              const double* Im_unmasked, double saturation,
                              const double* gimxscale, const double* gimyscale,
                              const double* dither, int64_t* n_out, double* x_out, double* y_out,
                              int32_t* exp_out, double* I_out);
/* imcom_rotate_image(im_in, theta_deg, npoly_in, conserve_flux, im_out)
 * (imcom_proc.f90:184-250): im_in, im_out n1 x n2; npoly <= 7 (8 is taken
 * as 7, as the reference does with a warning). */
int imcom_gpu_rotate_image(imcom_gpu_ctx* ctx, int n1, int n2, const double* im_in,
                           double theta_deg, int npoly, int conserve_flux, double* im_out);
/* imcom_ft_r2c(n1, n2, plan, F, Ft)  (imcom_proc.f90:433-475): the full
 * complex DFT of count real n1 x n2 images (count images stacked), with the
 * reference's Hermitian completion of the upper half of the first axis. */
int imcom_gpu_ft_r2c(imcom_gpu_ctx* ctx, int n1, int n2, int count, const double* img,
                     double* spec_out);
/* imcom_build_Alookup(npad)  (imcom_matrix.f90:361-430): Gt_rot holds
 * nexppsf spectra; Alookup_out (npad n1psf, npad n2psf,
 * nexppsf (nexppsf + 1) / 2) may be NULL. */
int imcom_gpu_build_Alookup(imcom_gpu_ctx* ctx, int psfconst, int nexppsf, int n1psf, int n2psf,
                            int npad, const double* Gt_rot, double* Alookup_out);
/* imcom_build_Blookup(npad)  (imcom_matrix.f90:434-493): Blookup_out
 * (npad n1psf, npad n2psf, nexppsf) may be NULL. */
int imcom_gpu_build_Blookup(imcom_gpu_ctx* ctx, int psfconst, int nexppsf, int n1psf, int n2psf,
                            int npad, const double* Gammat, const double* Gt_rot,
                            double* Blookup_out);
/* imcom_calc_C  (imcom_matrix.f90:546-552): C = sum |Gammat|^2 / (n1psf n2psf). */
int imcom_gpu_calc_C(imcom_gpu_ctx* ctx, int n1psf, int n2psf, const double* Gammat,
                     double* C_out);

/* imcom_lookup_A(npoly, npad, saveA)  (imcom_matrix.f90:206-284).
 * A_out (n*n) may be NULL.  npoly <= 7 (stencil of at most 8 points). */
int imcom_gpu_lookup_A(imcom_gpu_ctx* ctx, int npoly, int npad, double psfxscale,
                       double psfyscale, double* A_out);
/* imcom_get_A cache hit: A read from AFILE (imcom_matrix.f90:49-68). */
int imcom_gpu_set_A(imcom_gpu_ctx* ctx, const double* A);

/* imcom_eigen_A(saveQL)  (imcom_matrix.f90:586-678): Q (n*n) and L (n) in
 * descending eigenvalue order, condition = max(L)/min(L).  Consumes the
 * device copy of A.  Q_out, L_out, condition may be NULL. */
int imcom_gpu_eigen_A(imcom_gpu_ctx* ctx, double* Q_out, double* L_out, double* condition);
/* The two stages inside imcom_eigen_A's DSYEVR call (imcom_matrix.f90:642),
 * exposed for staged checks: the Householder reduction of A to tridiagonal
 * form (d: n, e: n-1; consumes the device copy of A), and the tridiagonal
 * eigensolver on host (d, e) returning ascending eigenvalues lam (n) and the
 * eigenvectors Z (n*n) of the tridiagonal matrix. */
int imcom_gpu_tridiagonalize(imcom_gpu_ctx* ctx, double* d_out, double* e_out);
int imcom_gpu_tridiag_eig(imcom_gpu_ctx* ctx, int64_t n, const double* d, const double* e,
                          double* lam_out, double* Z_out);
/* imcom_get_QL cache hit: Q, L read from QFILE/LFILE (imcom_matrix.f90:164-201). */
int imcom_gpu_set_QL(imcom_gpu_ctx* ctx, const double* Q, const double* L);

/* imcom_lookup_B(npoly, npad, saveB)  (imcom_matrix.f90:288-357); B_out (n*m) may be NULL. */
int imcom_gpu_lookup_B(imcom_gpu_ctx* ctx, int npoly, int npad, double psfxscale,
                       double psfyscale, double* B_out);
/* imcom_get_B cache hit (imcom_matrix.f90:85-108). */
int imcom_gpu_set_B(imcom_gpu_ctx* ctx, const double* B);

/* imcom_build_P(saveP)  (imcom_matrix.f90:682-712): P = Q^T B; P_out (n*m)
 * may be NULL.  P2 = P*P is formed on the fly where it is used. */
int imcom_gpu_build_P(imcom_gpu_ctx* ctx, int precision, double* P_out);
/* imcom_get_P cache hit (imcom_matrix.f90:122-147). */
int imcom_gpu_set_P(imcom_gpu_ctx* ctx, const double* P);

/* imcom_bisect_kappa(UnotS, US, k_min, k_max, tol, maxnits)
 * (imcom_bisect.f90:78-170): K_out (m) and US_out (m, U or S at the root).
 * evals_out (m, may be NULL) counts U/S evaluations per pixel.  bracket
 * (2*m floats, log10 kappa low/high per pixel, may be NULL) starts each
 * pixel's bisection from a predicted bracket; a bracket that does not hold
 * the root falls back to the full range. */
int imcom_gpu_bisect_kappa(imcom_gpu_ctx* ctx, int UnotS, double US_max, double k_min,
                           double k_max, double tol, int maxnits, const float* bracket,
                           double* K_out, double* US_out, int32_t* evals_out);

/* imcom_build_T  (imcom_matrix.f90:716-750): T = -1/2 Q (P / (L + K_a));
 * T_out (n*m) may be NULL. */
int imcom_gpu_build_T(imcom_gpu_ctx* ctx, int precision, double* T_out);
/* imcom_get_T cache hit: T read from TFILE instead of re-solved when forceT
 * is 0 (imcom_bisect.f90:47-62); build_image and build_S then use it. */
int imcom_gpu_set_T(imcom_gpu_ctx* ctx, const double* T);
/* imcom_build_S  (imcom_matrix.f90:839-877): S_a = sum_i T_ia^2 Ndiag(i). */
int imcom_gpu_build_S(imcom_gpu_ctx* ctx, double* S_out);
/* imcom_build_U  (imcom_matrix.f90:795-835): U_a = U(K_a). */
int imcom_gpu_build_U(imcom_gpu_ctx* ctx, double* U_out);
/* imcom_build_image  (imcom_matrix.f90:754-790): H = T^T I. */
int imcom_gpu_build_image(imcom_gpu_ctx* ctx, const double* I_i, double* H_out);

/* Learned kappa bracket.  No reference counterpart: the reference always
 * bisects the full [kappa_min, kappa_max] range (imcom_bisect.f90:174-267).
 * A 32-64-64-2 GELU MLP (FP32) reads 32 features per output pixel (12
 * log eigenvalue quantiles, 19 log-binned energies of the pixel's column of
 * P, log of the U or S target) and predicts mu and log sigma of log10 kappa.
 * Parameter block: imcom_gpu_mlp_nparams() floats (weights, then the
 * feature and target standardisation).  imcom_gpu_mlp_train runs Adam on the
 * mean Gaussian negative log-likelihood over minibatches of (X: N*32,
 * y: N log10 kappa); loss_out gets one mean loss per epoch.
 * imcom_gpu_bisect_kappa_learned is imcom_gpu_bisect_kappa started from the
 * bracket mu +- zsig*sigma; hits_out (m, may be NULL) is 1 where the bracket
 * held the root and 0 where the pixel fell back to the full range. */
int imcom_gpu_mlp_nparams(void);
int imcom_gpu_mlp_set_params(imcom_gpu_ctx* ctx, const float* params);
int imcom_gpu_mlp_get_params(imcom_gpu_ctx* ctx, float* params);
int imcom_gpu_kappa_features(imcom_gpu_ctx* ctx, double US_target, float* X_out);
int imcom_gpu_mlp_infer(imcom_gpu_ctx* ctx, const float* X, int64_t N, float* out);
int imcom_gpu_mlp_grad(imcom_gpu_ctx* ctx, const float* X, const float* y, int64_t N,
                       float* grad_out, float* loss_out);
int imcom_gpu_mlp_train(imcom_gpu_ctx* ctx, const float* X, const float* y, int64_t N, int epochs,
                        int batch, float lr, uint64_t seed, float* loss_out);
int imcom_gpu_bisect_kappa_learned(imcom_gpu_ctx* ctx, int UnotS, double US_max, double k_min,
                                   double k_max, double tol, int maxnits, float zsig,
                                   double* K_out, double* US_out, int32_t* evals_out,
                                   int32_t* hits_out);

/* imcom_gpu_bisect_kappa_learned for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Every pointer is a
 * device pointer: L (n, descending), P (n x m, leading dimension ldp >= n),
 * Ndiag (n; only read in S mode, may be NULL in U mode), params
 * (imcom_gpu_mlp_nparams() floats, as imcom_gpu_mlp_set_params takes them,
 * not validated here), outputs K and US (m), evals and hits (m, may be NULL).
 * stream: a hipStream_t, or NULL for the context's stream; the call only
 * enqueues work and returns.  The kernels' scratch (per-pixel features and
 * brackets, eigenvalue bins) is a fixed device array of 2.16 MiB per GPU,
 * held until the process ends, which limits n <= 32768 and m <= 16384
 * (larger sizes return IMCOM_EARG).  Calls that overlap on one GPU share that
 * array and race.  Nothing is allocated: no hipMalloc, hipFree or allocating
 * library routine. */
int imcom_gpu_bisect_kappa_learned_noalloc(imcom_gpu_ctx* ctx, void* stream, int64_t n, int64_t m,
                                           const double* L, const double* P, int64_t ldp,
                                           const double* Ndiag, double C_a, const float* params,
                                           int UnotS, double US_max, double k_min, double k_max,
                                           double tol, int maxnits, float zsig, double* K,
                                           double* US, int32_t* evals, int32_t* hits);

/* Batched tiles.  No reference counterpart (the reference solves one tile per
 * process); each tile's result is what the single-tile entry points give for
 * it.  A tile carries its own inputs (the fields of imcom_gpu_set_tile and
 * imcom_gpu_set_tables, the image I_i, the target and bisection controls of
 * imcom_gpu_bisect_kappa, and the interpolation controls of
 * imcom_gpu_lookup_A/B).  imcom_gpu_batch_set uploads them;
 * imcom_gpu_batch_solve runs get_A, get_QL, get_B, get_P, get_T and
 * build_image for every tile (use_graph: the whole solve is captured into
 * one HIP graph on the first call with these precision and bracket settings
 * and replayed, one graph launch per solve); imcom_gpu_batch_result returns one
 * tile's K, U or S at the root (US_out), the other map (SU_out: S in U mode,
 * U in S mode), H, evaluation counts, bracket hits and T (any may be NULL). */
typedef struct imcom_gpu_tile {
  int64_t n, m;
  const double *x_i, *y_i;
  const int32_t* exp_i;
  const double *X_a, *Y_a, *Ndiag, *I_i;
  double C_a, kappa_min, kappa_max, US_max, US_tol;
  int UnotS, maxnits;
  int psfconst, nexppsf, n1, n2;
  const double *Alookup, *Blookup;
  int npoly, npad;
  double psfxscale, psfyscale;
} imcom_gpu_tile;

/* Eigensolver tuning picked by the autotuner (imcom_amd/autotune.py): the
 * reduction panel width nb (16, 32, 64) and the back-transform block nbb
 * (64, 128 or 256); applies to later eigen_A calls and batch_set. */
int imcom_gpu_set_tuning(imcom_gpu_ctx* ctx, int nb, int nbb);
/* Single-tile eigen_A and tridiagonalize reduce the last columns, from the
 * first panel boundary where at most max_rows remain (2..2048; default 2048),
 * in one persistent launch that keeps the trailing matrix in on-chip memory
 * and synchronises its workgroups per column; 0 reduces every column with
 * the blocked kernels.  Part of the DSYTRD stage of DSYEVR
 * (imcom_matrix.f90:642); no reference counterpart of its own. */
int imcom_gpu_set_reduction_tail(imcom_gpu_ctx* ctx, int max_rows);

/* Selects the reduction to tridiagonal form inside eigen_A / tridiagonalize
 * (the DSYTRD stage of DSYEVR, imcom_matrix.f90:642): stages = 1, the
 * one-stage blocked reduction; stages = 2, for n >= min_n, the
 * two-stage reduction (A to a band of width 64 by blocked QR panels with
 * matrix-product updates, then bulge chasing to tridiagonal form), with the
 * eigenvectors back-transformed through both stages.  Same outputs and
 * conventions either way; the eigenvalues agree to rounding.  Default:
 * stages = 2 with min_n = 7168 (the two-stage reduction where it was
 * measured faster, the one-stage one below). */
int imcom_gpu_set_reduction_stages(imcom_gpu_ctx* ctx, int stages, int64_t min_n);
/* Diagnostic: the first stage of the two-stage reduction alone.  A (n >= 3)
 * is reduced in place and downloaded (n x n, column-major): its lower
 * triangle holds the symmetric band matrix (bandwidth 64, zeros below the
 * band), orthogonally similar to the input; the upper triangle holds the
 * panels' Householder vectors.  A is consumed, as by eigen_A.  No reference
 * counterpart (a stage inside DSYEVR, imcom_matrix.f90:642). */
int imcom_gpu_band_reduce(imcom_gpu_ctx* ctx, double* A_out);

int imcom_gpu_batch_set(imcom_gpu_ctx* ctx, int ntiles, const imcom_gpu_tile* tiles);
int imcom_gpu_batch_solve(imcom_gpu_ctx* ctx, int precision_P, int precision_T, int learned,
                          float zsig, int use_graph);
/* Allocates the workspace imcom_gpu_batch_solve needs for these precisions
 * (batch_solve allocates it itself otherwise): a caller that runs batched
 * solves of several contexts from several threads calls it first, so that
 * no device allocation runs while another thread captures a solve graph. */
int imcom_gpu_batch_reserve(imcom_gpu_ctx* ctx, int precision_P, int precision_T);
int imcom_gpu_batch_result(imcom_gpu_ctx* ctx, int tile, double* K_out, double* US_out,
                           double* SU_out, double* H_out, int32_t* evals_out, int32_t* hits_out,
                           double* T_out);

/* Device-time accounting (no reference counterpart; the reference prints
 * progress lines only).  level 0: off; 1: per-stage HIP-event times; 2: also
 * every launch of the one-stage reduction's matrix-vector kernel and its
 * trailing-update GEMMs, and the two-stage reduction's bulge-chasing and
 * bulge back-transform launches.  Setting the level resets the counters.
 * imcom_gpu_stage_times fills ms_out[i] / counts_out[i] for the slots
 *   0 lookup_A, 1 tridiagonal reduction, 2 divide-and-conquer,
 *   3 back-transform, 4 lookup_B, 5 build_P, 6 bisect_kappa, 7 build_T,
 *   8 maps (S, U, H), 9 matrix-vector launches, 10 trailing updates,
 *   11 device lookup-table set-up, 12 bulge chasing, 13 bulge back-transform
 *   (nslots at most 14). */
int imcom_gpu_set_profiling(imcom_gpu_ctx* ctx, int level);

int imcom_gpu_stage_times(imcom_gpu_ctx* ctx, double* ms_out, int64_t* counts_out, int nslots);

/* Stage files written while the device works on.  The reference writes A, Q,
 * L, B, P and T synchronously between stages when their file names are not
 * None (imcom_matrix.f90:275-278, 348-351, 672-677, 700-703, 748-749); the
 * synchronous form here is the stage's own output pointer.  These three
 * calls have no reference counterpart:
 * imcom_gpu_fetch_async copies the stage output `what` (IMCOM_FETCH_*), as
 * it stands after the work issued so far, to `host` on a second stream and
 * returns at once.  The output must exist (the stage has run, or its set_*
 * call).  Stages that overwrite that device buffer (lookup_A and eigen_A for
 * A, eigen_A for Q and L, lookup_B for B, build_P for P, build_T for T, and
 * the set_* calls) wait for the copy on the device.  `host` must stay valid
 * until imcom_gpu_fetch_wait(ctx, what) (what = -1: every output) returns;
 * set_tile and finalize wait for every copy.  Page-locked memory from
 * imcom_gpu_host_alloc lets the copies run beside the later stages. */
int imcom_gpu_fetch_async(imcom_gpu_ctx* ctx, int what, double* host);
int imcom_gpu_fetch_wait(imcom_gpu_ctx* ctx, int what);
int imcom_gpu_host_alloc(size_t bytes, void** ptr);
int imcom_gpu_host_free(void* ptr);

#ifdef __cplusplus
}
#endif

#endif /* IMCOM_GPU_H */
