// Internal launcher interfaces of the IMCOM solve-path kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace imcom {

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
};

hipError_t dgemm(hipStream_t st, bool ta, bool tb, int M, int N, int K, double alpha,
                 const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                 double* C, int64_t ldc);
int gemm_tiles64(int M, int N);
hipError_t dgemm_grouped(hipStream_t st, bool ta, bool tb, const GemmDesc* d_descs,
                         const int* d_prefix, int ngroups, int total_blocks);

// Split-precision GEMM (FP32 x3 or BF16 x6 products on MFMA with FP32
// accumulation and FP64 recombination); scheme 1 = FP32x3, 2 = BF16x6.
hipError_t gemm_split(hipStream_t st, int scheme, bool ta, int M, int N, int K, double alpha,
                      const double* A, int64_t lda, const double* B, int64_t ldb, double* C,
                      int64_t ldc, void* work, size_t work_bytes);
size_t gemm_split_work_bytes(int scheme, int M, int N, int K);

// ---- (a) interpolation ----------------------------------------------------
hipError_t launch_lookup_A(hipStream_t st, int dev, int64_t n, const double* x, const double* y,
                           const int32_t* ex, int psfconst, const double* tabs, int64_t tstride,
                           int n1, int n2, int npm, int npad, double sx, double sy, double* A);
hipError_t launch_lookup_B(hipStream_t st, int dev, int64_t n, int64_t m, const double* x,
                           const double* y, const int32_t* ex, const double* Xa, const double* Ya,
                           int psfconst, const double* tabs, int64_t tstride, int n1, int n2,
                           int npm, int npad, double sx, double sy, double* B);

// ---- (d) bisection and per-pixel epilogues --------------------------------
hipError_t launch_bisect(hipStream_t st, int64_t n, int64_t m, const double* L, const double* P,
                         int64_t ldp, const double* Nd, double C_a, double US_max, double tol,
                         double kmin, double kmax, int maxnits, int UnotS, const float* bracket,
                         double* K, double* US, int32_t* evals);
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
  double* W;      // n x NB
  double* VW;     // n x 2NB  [V W]
  double* WV;     // n x 2NB  [W V]
  double* y;      // n
  double* wpre;   // n
  double* vtmp;   // n
  double* d;      // n     tridiagonal diagonal
  double* e;      // n-1   off-diagonal
  double* tau;    // n-1
  double* part;   // >= ceil(n/256) partial sums
  double* part2;  // >= ceil(n/256)
  double* s;      // 2 NB dot products
  double* scal;   // 1
  double* Vb;     // (n-1) x NBB
  double* Sb;     // NBB x NBB
  double* Tb;     // NBB x NBB
  double* Y;      // NBB x n
  double* Y2;     // NBB x n
  double* Z;      // n x n  eigenvectors (in: of the tridiagonal, out: of A)
};

hipError_t sytrd_run(hipStream_t st, const SyTile* d_tiles, int ntiles, int64_t nmax,
                     const std::vector<SyTile>& host, int nb);
hipError_t ormtr_run(hipStream_t st, const SyTile* d_tiles, int ntiles, int64_t nmax,
                     const std::vector<SyTile>& host, int nbb);

// Divide-and-conquer symmetric tridiagonal eigensolver (ascending
// eigenvalues, eigenvectors in Z, ld n).  `ws` is the workspace object.
struct DcWork;
DcWork* dc_create();
void dc_destroy(DcWork* w);
hipError_t dc_solve(hipStream_t st, DcWork* w, int64_t n, const double* d, const double* e,
                    double* lam, double* Z);

// misc
hipError_t launch_reverse_cols(hipStream_t st, int64_t n, const double* Zin, double* Qout,
                               const double* lam_in, double* L_out);
hipError_t launch_minmax_ratio(hipStream_t st, int64_t n, const double* L, double* out);

}  // namespace imcom
