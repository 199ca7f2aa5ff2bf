// Internal launcher interfaces of the IMCOM solve-path kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <array>
#include <vector>

namespace imcom {

// ---- profiling --------------------------------------------------------------
// Device-time accounting with HIP events on the context's stream.  Slots
// name the stages of the solve path; per-launch timing of the tridiagonal
// reduction's matrix-vector kernel is kept in its own slot (level 2), with
// the events recorded by the kernel dispatch itself.
enum ProfSlot {
  PS_LOOKUP_A = 0,
  PS_SYTRD,
  PS_STEDC,
  PS_ORMTR,
  PS_LOOKUP_B,
  PS_BUILD_P,
  PS_BISECT,
  PS_BUILD_T,
  PS_MAPS,
  PS_SYMV,      // sy_symv launches only (level 2)
  PS_SYR2K,     // trailing-update GEMMs only (level 2)
  PS_TABLES,    // device lookup-table set-up (build_Alookup / build_Blookup)
  PS_BULGE,     // the bulge-chasing launch of the two-stage reduction (level 2)
  PS_Q2,        // the bulge back-transform's apply launch (level 2)
  PS_NSLOTS
};

struct Prof {
  int level = 0;
  hipStream_t st = nullptr;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  std::vector<std::pair<int, size_t>> pending;  // (slot, index of the first event)
  double ms[PS_NSLOTS] = {0};
  int64_t count[PS_NSLOTS] = {0};
  hipError_t begin(int slot, int need_level, size_t* token);
  hipError_t end(int slot, size_t token);
  // a start/stop event pair for one kernel launch (hipExtLaunchKernelGGL records
  // them at the kernel's own start and end); both null below need_level
  hipError_t events(int slot, int need_level, hipEvent_t* e0, hipEvent_t* e1);
  hipError_t collect();
  void reset();
  void destroy();
};

// ---- GEMM -----------------------------------------------------------------
struct GemmDesc {
  int M, N, K;
  double alpha, beta;
  const double* A;
  int64_t lda;
  const double* B;
  int64_t ldb;
  double* C;
  int64_t ldc;
  int lower = 0;  // 1: only the tiles of C on or below its diagonal (M == N)
  // Optional block-layout C (see tri_block_offset): element (i, j) of C is
  // element (ct_off + i, ct_off + j) of a symmetric matrix whose lower
  // triangle is stored in 64x64 blocks at Ct.  When set, C's old value is
  // read from and the result written to Ct (blocks on or below the block
  // diagonal only), and the column-major C receives only columns
  // j < cm_cols.
  double* Ct = nullptr;
  int64_t ct_off = 0;
  int ct_nbt = 0;
  int cm_cols = 0;
  // Split K (non-grouped launches only): blockIdx.y = s takes k in
  // [s*kchunk, min(K, (s+1)*kchunk)) and writes C + s*c_split_stride.
  int kchunk = 0;
  int64_t c_split_stride = 0;
};

// Offset (in doubles) of block (I, J), I >= J, in the lower-triangle block
// layout of an nbt x nbt grid of 64x64 blocks: block-columns one after
// another, each block column-major (element (i, j) of the block at
// j * 64 + i).
__host__ __device__ __forceinline__ int64_t tri_block_offset(int I, int J, int nbt) {
  return ((int64_t)J * nbt - (int64_t)J * (J + 1) / 2 + I) * 4096;
}

hipError_t dgemm(hipStream_t st, bool ta, bool tb, int M, int N, int K, double alpha,
                 const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                 double* C, int64_t ldc);
// C's tiles on or below the diagonal only (M == N): the symmetric trailing
// update, whose upper triangle is never read.
hipError_t dgemm_lower(hipStream_t st, bool ta, bool tb, int M, int K, double alpha,
                       const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                       double* C, int64_t ldc, double* Ct = nullptr, int64_t ct_off = 0,
                       int ct_nbt = 0, int cm_cols = 0);
int gemm_tiles64(int M, int N);
// C (M x N, ldc) = alpha * op(A) op(B) with K split over workgroups; the
// partial products go to `work` (ksplit * M * N doubles, from
// dgemm_splitk_work) and are summed in a fixed order.
size_t dgemm_splitk_work(int M, int N, int K);
hipError_t dgemm_splitk(hipStream_t st, bool ta, bool tb, int M, int N, int K, double alpha,
                        const double* A, int64_t lda, const double* B, int64_t ldb, double* C,
                        int64_t ldc, double* work);
int gemm_tiles64(const GemmDesc& g);
// tile: 64 or 128; total_blocks and d_prefix count tiles of that size
hipError_t dgemm_grouped(hipStream_t st, bool ta, bool tb, const GemmDesc* d_descs,
                         const int* d_prefix, int ngroups, int total_blocks, int tile = 64);

// A sequence of grouped GEMM launches whose descriptors are uploaded once
// (before any graph capture) and replayed by index.
struct GroupedPlan {
  struct Launch {
    int desc_off, prefix_off, ngroups, blocks;
    bool ta, tb;
    int tile;  // 64 or 128 (square output tiles)
  };
  std::vector<GemmDesc> descs;
  std::vector<int> prefix;
  std::vector<Launch> launches;
  GemmDesc* d_descs = nullptr;
  int* d_prefix = nullptr;
  // append one launch made of `groups` (empty groups are dropped); returns its index
  int add(bool ta, bool tb, const std::vector<GemmDesc>& groups);
  hipError_t upload(hipStream_t st);  // (re)allocates and copies; synchronises
  hipError_t run(hipStream_t st, int launch) const;
  void release();
};

// Split-precision GEMM (FP32 x3 or BF16 x6 products on MFMA with FP32
// accumulation and FP64 recombination); scheme 1 = FP32x3, 2 = BF16x6.
hipError_t gemm_split(hipStream_t st, int scheme, bool ta, int M, int N, int K, double alpha,
                      const double* A, int64_t lda, const double* B, int64_t ldb, double* C,
                      int64_t ldc, void* work, size_t work_bytes);
size_t gemm_split_work_bytes(int scheme, int M, int N, int K);

// ---- (a) interpolation ----------------------------------------------------
hipError_t interp_init(int dev);
// imcom_rotate_image on the device; sums: 2 doubles of device scratch
hipError_t launch_rotate(hipStream_t st, int dev, const double* in, int n1, int n2, double cost,
                         double sint, int npm, int conserve_flux, double* out, double* sums);
hipError_t launch_lookup_A(hipStream_t st, int dev, int64_t n, const double* x, const double* y,
                           const int32_t* ex, int psfconst, const double* tabs, int64_t tstride,
                           int n1, int n2, int npm, int npad, double sx, double sy, double* A);
hipError_t launch_lookup_B(hipStream_t st, int dev, int64_t n, int64_t m, const double* x,
                           const double* y, const int32_t* ex, const double* Xa, const double* Ya,
                           int psfconst, const double* tabs, int64_t tstride, int n1, int n2,
                           int npm, int npad, double sx, double sy, double* B);

// Separations between two regular 1-D grids, r(k) = k sr + dr (k < Nr, the
// subtracted grid) and c(l) = l sc + dc (l < Nc), with sr = P g, sc = Q g for
// coprime P, Q: c(l) - r(k) = (Q l - P k) g + doff, doff = dc - dr, so the
// separations are the lattice index u = Q l - P k + P (Nr - 1) in [0, L),
// L = P (Nr - 1) + Q (Nc - 1) + 1.
struct LatAxis {
  int P, Q, Nr, Nc, L;
  double g, doff;
};
// One pair of grids (an exposure pair for A, an exposure and the output grid
// for B): its x and y axes, the lookup table it reads unless all exposures
// share table 0 (psfconst = 1), and where its Lx * Ly interpolated values
// start in the value array (x index fastest).
struct LatDesc {
  LatAxis x, y;
  int table;
  int64_t off;
};
// The 0-based grid position of every pixel of a regular (USERXY=0) tile,
// checked against the coordinates: kx, ky (n each) from the 1-based
// mask_subs into the (n1gim, n2gim, nexp) order; bad counts the pixels whose
// exposure or position do not match (and the output pixels off their grid).
hipError_t launch_grid_check(hipStream_t st, int64_t n, int64_t m, const double* x, const double* y,
                             const int32_t* ex, const int32_t* mask_subs, int n1gim, int n2gim,
                             int nexp, const double* gsx, const double* gsy, const double* dith,
                             int n1out, double osx, double osy, double op1, double op2,
                             const double* Xa, const double* Ya, int32_t* kx, int32_t* ky,
                             int32_t* bad);
// Every pair's lattice values (total entries, descriptors in desc), then A
// (n x n, both triangles: pixel lo = min(i, j) is the subtracted one, so
// A(i, j) = A(j, i) exactly) or B (n x m) expanded from them.
hipError_t launch_lattice_eval(hipStream_t st, int dev, const LatDesc* desc, int npairs,
                               int64_t total, int psfconst, const double* tabs, int64_t tstride,
                               int n1, int n2, int npm, int npad, double sx, double sy,
                               double* vals);
hipError_t launch_lattice_A(hipStream_t st, int64_t n, const int32_t* kx, const int32_t* ky,
                            const int32_t* ex, const LatDesc* desc, const double* vals, double* A);
hipError_t launch_lattice_B(hipStream_t st, int64_t n, int64_t m, int n1out, const int32_t* kx,
                            const int32_t* ky, const int32_t* ex, const LatDesc* desc,
                            const double* vals, double* B);

// ---- (d) bisection and per-pixel epilogues --------------------------------
hipError_t launch_bisect(hipStream_t st, int64_t n, int64_t m, const double* L, const double* P,
                         int64_t ldp, const double* Nd, double C_a, double US_max, double tol,
                         double kmin, double kmax, int maxnits, int UnotS, const float* bracket,
                         double* K, double* US, int32_t* evals, int32_t* hits);
hipError_t launch_scale_W(hipStream_t st, int64_t n, int64_t m, const double* P, int64_t ldp,
                          const double* L, const double* K, double* W, int64_t ldw);
hipError_t launch_maps(hipStream_t st, int64_t n, int64_t m, const double* T, int64_t ldt,
                       const double* I, const double* Nd, double* H, double* S);
hipError_t launch_U_from_K(hipStream_t st, int64_t n, int64_t m, const double* L, const double* P,
                           int64_t ldp, double C_a, const double* K, double* U);

// ---- (b) eigensolver ------------------------------------------------------
// One tile of the tridiagonal reduction / back-transform.  All pointers are
// device pointers sized for this tile's n (NB = panel width, NBB = back-
// transform block).
struct SyTile {
  int64_t n;
  double* A;      // n x n, overwritten by the reflectors (lower part)
  double* At;     // lower-triangle 64x64 blocks of A (tri_block_offset), read by sy_symv
  int nbt;        // ceil(n / 64)
  double* W;      // n x NB
  double* VW;     // n x 2NB  [V W]
  double* WV;     // n x 2NB  [W V]
  double* ypart;  // ceil(n/64) x n   row partials of A22 x (absolute row index)
  double* ycpart; // ceil(n/256) x n  column partials of A22 x (absolute column index)
  double* qpart;  // kQParts            per-workgroup partials of x^T A22 x
  double* wpre;   // n
  double* vtmp;   // n
  double* d;      // n     tridiagonal diagonal
  double* e;      // n-1   off-diagonal
  double* tau;    // n-1
  double* part;   // 2 x kNormParts  per-workgroup partials of a column's squared norm
  double* part2;  // kQParts
  double* s;      // 2 NB dot products
  double* dpart;  // 2 NB x kDotChunks  chunk partials of the dot products
  double* scal;   // 5: reflector scale 1/(alpha-beta), tau, alpha, beta, W(c+1, jj)
  double* Vb;     // (n-1) x NBB
  double* Sb;     // NBB x NBB
  double* Tb;     // NBB x NBB
  double* Y;      // NBB x n
  double* Y2;     // NBB x n
  double* Z;      // n x n  eigenvectors (in: of the tridiagonal, out: of A)
  double* Ks;     // split-K partials of the back-transform's V^T V and V^T Z
};

// Partial-sum slots of the reduction kernels, sized for the largest tile
// (n <= 32768) so that tiles of a batch can share launch grids.
constexpr int kQParts = 70000;
constexpr int kNormParts = 2048;
constexpr int kDotChunks = 16;

// `sharers`: contexts whose reductions may run at once on the device; each
// keeps 1/sharers of the matrix-vector product's cached share of the
// last-level cache.
// Single tile (ntiles == 1) with tail_rows >= 2 and tail_xch set: the
// columns from the first panel boundary k0 with n - k0 <= tail_rows on are
// reduced by the persistent kernel (sytrd_tail); *tail_S receives n - k0
// (0 when it did not run).
hipError_t sytrd_run(hipStream_t st, const SyTile* d_tiles, int ntiles, int64_t nmax,
                     const std::vector<SyTile>& host, int nb, int sharers, Prof* prof,
                     int tail_rows = 0, double* tail_xch = nullptr, int64_t* tail_S = nullptr);

// The persistent tail of the reduction (sytrd_tail.hip): trailing sizes up to
// kTailMaxRows, whose columns, spread over at most 256 workgroups, fit in
// their LDS.  xch: tail_xch_doubles(S) doubles of device exchange space.
constexpr int kTailMaxRows = 2048;
size_t tail_xch_doubles(int S);
void tail_shape(int S, int* G, int* cpw);
hipError_t sytrd_tail(hipStream_t st, const SyTile& t, int64_t k0, bool from_A, double* xch);
// The abort flag the last tail launch left in xch (synchronises the stream):
// 0 completed, 1 a barrier wait ran out, 2 the barrier count was not reset.
hipError_t sytrd_tail_status(hipStream_t st, int64_t S, const double* xch, unsigned* flag);

// Two-stage tridiagonal reduction (sytrd2.hip): A (column-major, both
// triangles) to a band of width kBand by blocked QR panels, the band to
// tridiagonal form by bulge chasing, and the back-transforms.
constexpr int kBand = 64;
constexpr int kQ2Group = 32;  // reflectors per block of the bulge-chasing back-transform
// doubles per block image of that back-transform: V and -V T (96 rows
// each), rows of kQ2Group + 1, padded to a multiple of 512 (two doubles per
// thread per load of a 256-thread workgroup)
constexpr int kQ2Image = 6656;
struct TwoStageBufs {
  double* U = nullptr;      // n x 3 kBand: [V W V] of the current panel
  double* X = nullptr;      // n x kBand
  double* Mt = nullptr;     // kBand^2
  double* Mm = nullptr;     // kBand^2
  double* T1 = nullptr;     // band_panels(n) x kBand^2: the panels' triangular factors
  double* xch = nullptr;    // band_xch_doubles(n): the panel kernel's exchange slots
  double* kwork = nullptr;  // band_kwork_doubles(n): split-K partial products
  double* Y1 = nullptr;     // 4 kBand x n
  double* Y2 = nullptr;     // 4 kBand x n
  double* q1work = nullptr; // q1_group_work_doubles(n): a panel group's V^T, T, G
  unsigned* sync = nullptr; // 2 words: panel barrier count, abort flag
  unsigned* prog = nullptr; // bulge_flag_words(n): bulge-chasing message counters, abort flag
  double* msg = nullptr;    // bulge_msg_doubles(n): bulge-chasing message rings
};
size_t band_xch_doubles(int64_t n);
size_t band_kwork_doubles(int64_t n);
int band_panels(int64_t n);
// Stage 1: the band in A's lower triangle (zeros below it), each panel's V^T
// in the upper triangle rows of its columns, T per panel in w.T1.
hipError_t band_reduce(hipStream_t st, double* A, int64_t n, const TwoStageBufs& w);
// Stage 2 (n >= 3): d, e of the tridiagonal matrix; reflector (s, k) in
// column s of V2 (n x n), rows 64k.. (tau in place of the unit first entry).
size_t bulge_msg_doubles(int64_t n);
size_t bulge_flag_words(int64_t n);
hipError_t bulge_chase(hipStream_t st, double* A, int64_t n, double* V2, double* d, double* e,
                       double* msg, unsigned* flags, Prof* prof = nullptr);
// flags[0]: a panel barrier timed out; flags[1]: a bulge-chasing wait timed out.
hipError_t twostage_status(hipStream_t st, const TwoStageBufs& w, int64_t n, unsigned* flags);
// Q2 blocks: kQ2Group consecutive sweeps at one task index, numbered group
// by group (blk_off: q2_groups(n) + 1 prefix counts).
int q2_groups(int64_t n);
void q2_block_offsets(int64_t n, std::vector<int>& blk_off);
// Z := Q2 Z (n x n, leading dimension n), Q2 the product of the bulge
// chase's reflectors (V2 as bulge_chase leaves it, n x n; overwritten: once
// the reflectors are in the block images its storage holds Z^T for the
// apply); img holds blk_off[nS] x kQ2Image doubles, maxk = the first
// group's block count.
hipError_t apply_q2(hipStream_t st, double* V2, int64_t n, const int* d_blk_off, int nS,
                    int maxk, double* img, double* Z, Prof* prof = nullptr);
size_t q1_group_work_doubles(int64_t n);
// Q := Q1 Q (ncols columns of Q, leading dimension n).
hipError_t apply_q1(hipStream_t st, const double* A, int64_t n, const TwoStageBufs& w, double* Q,
                    int64_t ncols);
hipError_t ormtr_run(hipStream_t st, const SyTile* d_tiles, int ntiles, int64_t nmax,
                     const std::vector<SyTile>& host, int nbb);

struct SyPlan {
  GroupedPlan gp;
  std::vector<int> syr2k;            // grouped launch per panel
  std::vector<std::vector<int>> bt;  // grouped launches per back-transform block: S, Y,
                                     // Y2, Z update, then T's joins (two each)
};
void sy_make_plan(const std::vector<SyTile>& host, int64_t nmax, int nb, int nbb, SyPlan& plan);
hipError_t sytrd_run_batched(hipStream_t st, const SyTile* d_tiles, int ntiles, int64_t nmax,
                             int nb, const SyPlan& plan, int sharers, Prof* prof);
hipError_t ormtr_run_batched(hipStream_t st, const SyTile* d_tiles, int ntiles, int64_t nmax,
                             int nbb, const SyPlan& plan);

// Divide-and-conquer symmetric tridiagonal eigensolver (ascending
// eigenvalues, eigenvectors in Z, ld n).  `ws` is the workspace object.
struct DcWork;
DcWork* dc_create();
void dc_destroy(DcWork* w);
hipError_t dc_solve(hipStream_t st, DcWork* w, int64_t n, const double* d, const double* e,
                    double* lam, double* Z);
// Builds the launch plan of dc_solve_batched for these tile sizes and
// eigenvector buffers (host work and uploads; synchronises st): call before
// capturing a dc_solve_batched into a HIP graph.
hipError_t dc_prepare(hipStream_t st, DcWork* w, int ntiles, const int64_t* ns, double* const* Z_t);
hipError_t dc_solve_batched(hipStream_t st, DcWork* w, int ntiles, const int64_t* ns,
                            const double* const* d_t, const double* const* e_t,
                            double* const* lam_t, double* const* Z_t);

// ---- learned kappa bracket ---------------------------------------------------
struct MlpScratch {
  float *Xn, *H1, *dZ1, *H2, *dZ2, *dY;
};
int mlp_nparams();
int mlp_ntrain();
hipError_t launch_features(hipStream_t st, int64_t n, int64_t m, const double* L, int8_t* bins,
                           const double* P, int64_t ldp, double C_a, double target, float* X);
hipError_t launch_mlp_infer(hipStream_t st, const float* params, const float* X, int64_t N,
                            float zsig, float* out, int as_bracket);
hipError_t mlp_gradient(hipStream_t st, const float* params, const float* X, const float* Y,
                        const int* perm, int64_t off, int64_t B, MlpScratch& s, float* grad,
                        float* loss);
hipError_t launch_adam(hipStream_t st, float* p, const float* g, float* m, float* v, int n,
                       float lr, int step);

// misc
hipError_t launch_reverse_cols(hipStream_t st, int64_t n, const double* Zin, double* Qout,
                               const double* lam_in, double* L_out);
hipError_t launch_minmax_ratio(hipStream_t st, int64_t n, const double* L, double* out);

}  // namespace imcom
