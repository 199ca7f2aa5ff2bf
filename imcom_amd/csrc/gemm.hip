// FP64 GEMM on CDNA4 matrix cores (v_mfma_f64_16x16x4_f64).
//
//   C (M x N) = alpha * op(A) * op(B) + beta * C      (column-major)
//
// Used for P = Q^T B (imcom_matrix.f90:699, DGEMM('T','N')), for
// T = -1/2 Q W with W = P / (L + K_a) (imcom_matrix.f90:731-745: the m DGEMVs
// done as one GEMM), and inside the eigensolver (trailing updates of the
// tridiagonal reduction, divide-and-conquer eigenvector products, Householder
// back-transform).  A grouped form runs many independent GEMMs of different
// shapes in one launch (one descriptor per group, blocks mapped to groups by
// a prefix array).
//
// Tiling: a 256-thread workgroup (4 wavefronts as 2x2) owns a BM x BN tile of
// C; each wavefront owns (BM/2) x (BN/2) as 16x16 MFMA sub-tiles.  K advances
// in BK=16 slabs staged through LDS with register prefetch of the next slab
// (two LDS buffers, one barrier per slab).  The MFMA takes the B slab as its
// first operand, which puts 16 consecutive rows of one C column in the 16
// lanes of a lane group: stores leave as 128-byte column segments.
// (f64 MFMA lane layout, measured on gfx950: A operand lane l = A[l%16][l/16],
// B operand lane l = B[l/16][l%16], result register r of lane l =
// D[4r + l/16][l%16].)
#include "common.h"
#include "kernels.h"

#include <algorithm>

namespace imcom {

constexpr int BK = 16;

// 128x128 tiles are held to at most 256 registers per lane (accumulators
// included) so that two wavefronts share each SIMD: with one, every barrier
// and LDS wait idles the matrix core (measured: 45 -> 55 TF/s on 16384^2
// operands).
template <bool TA, bool TB, int BM, int BN, bool GROUPED>
__global__ __launch_bounds__(256, (BM >= 128 ? 2 : 1)) void dgemm_kernel(GemmDesc one, const GemmDesc* __restrict__ descs,
                                                    const int* __restrict__ prefix, int ngroups) {
  constexpr int PA = BM + 1;  // padded LDS row lengths
  constexpr int PB = BN + 1;
  __shared__ double As[2][BK * PA];
  __shared__ double Bs[2][BK * PB];
  constexpr int LA = BM * BK / 256;  // doubles per thread per slab
  constexpr int LB = BN * BK / 256;
  constexpr int WM = BM / 2, WN = BN / 2;  // per-wave tile
  constexpr int TM = WM / 16, TN = WN / 16;

  int b = blockIdx.x;
  GemmDesc g;
  if (GROUPED) {
    int lo = 0, hi = ngroups - 1;  // last group with prefix[grp] <= b
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (prefix[mid] <= b)
        lo = mid;
      else
        hi = mid - 1;
    }
    g = descs[lo];
    b -= prefix[lo];
  } else {
    g = one;
  }
  const int M = g.M, N = g.N;
  int K = g.K;
  const double* Ag = g.A;
  const double* Bg = g.B;
  double* Cg = g.C;
  if (!GROUPED && g.kchunk > 0) {  // split K: this block's k range and partial
    const int64_t k0 = (int64_t)blockIdx.y * g.kchunk;
    K = (int)min((int64_t)g.kchunk, (int64_t)g.K - k0);
    Ag += TA ? k0 : k0 * g.lda;
    Bg += TB ? k0 * g.ldb : k0;
    Cg += (int64_t)blockIdx.y * g.c_split_stride;
  }
  const int tiles_m = (M + BM - 1) / BM, tiles_n = (N + BN - 1) / BN;
  int tm, tn;
  if (g.lower) {
    // row-major walk of the tiles with tm >= tn
    if (b >= tiles_m * (tiles_m + 1) / 2) return;
    int r = (int)((sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
    while ((r + 1) * (r + 2) / 2 <= b) ++r;
    while (r * (r + 1) / 2 > b) --r;
    tm = r;
    tn = b - r * (r + 1) / 2;
  } else {
    const int nt = tiles_m * tiles_n;
    if (b >= nt) return;
    // blocks are handed round-robin to the 8 XCDs: give each XCD a contiguous
    // run of tiles so neighbouring tiles share A rows / B columns in its L2
    if (!GROUPED && (nt & 7) == 0) b = (b & 7) * (nt >> 3) + (b >> 3);
    // column-major walk in groups of 8 tile rows
    const int group = 8 * tiles_n;
    const int gi0 = b / group;
    const int first_m = gi0 * 8;
    const int gm = min(tiles_m - first_m, 8);
    tm = first_m + (b % group) % gm;
    tn = (b % group) / gm;
  }
  const int m0 = tm * BM, n0 = tn * BN;
  const double* __restrict__ A = Ag;
  const double* __restrict__ B = Bg;
  const int64_t lda = g.lda, ldb = g.ldb;

  const int tid = threadIdx.x;
  const int lane = tid & 63, wid = tid >> 6;
  const int wm0 = (wid & 1) * WM, wn0 = (wid >> 1) * WN;

  v4d acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = (v4d){0.0, 0.0, 0.0, 0.0};

  double ra[LA], rb[LB];
  auto load_slab = [&](int k0) {
#pragma unroll
    for (int q = 0; q < LA; ++q) {
      const int idx = tid + q * 256;
      const int i = TA ? idx / BK : idx % BM;
      const int k = TA ? idx % BK : idx / BM;
      const int gi = m0 + i, gk = k0 + k;
      double v = 0.0;
      if (gi < M && gk < K) v = TA ? A[gk + (int64_t)gi * lda] : A[gi + (int64_t)gk * lda];
      ra[q] = v;
    }
#pragma unroll
    for (int q = 0; q < LB; ++q) {
      const int idx = tid + q * 256;
      const int j = TB ? idx % BN : idx / BK;
      const int k = TB ? idx / BN : idx % BK;
      const int gj = n0 + j, gk = k0 + k;
      double v = 0.0;
      if (gj < N && gk < K) v = TB ? B[gj + (int64_t)gk * ldb] : B[gk + (int64_t)gj * ldb];
      rb[q] = v;
    }
  };
  auto store_slab = [&](int buf) {
#pragma unroll
    for (int q = 0; q < LA; ++q) {
      const int idx = tid + q * 256;
      const int i = TA ? idx / BK : idx % BM;
      const int k = TA ? idx % BK : idx / BM;
      As[buf][k * PA + i] = ra[q];
    }
#pragma unroll
    for (int q = 0; q < LB; ++q) {
      const int idx = tid + q * 256;
      const int j = TB ? idx % BN : idx / BK;
      const int k = TB ? idx / BN : idx % BK;
      Bs[buf][k * PB + j] = rb[q];
    }
  };

  const int nslab = (K + BK - 1) / BK;
  if (nslab > 0) {
    load_slab(0);
    store_slab(0);
  }
  __syncthreads();
  const int lr = lane & 15, lk = lane >> 4;
  for (int s = 0; s < nslab; ++s) {
    const int buf = s & 1;
    if (s + 1 < nslab) load_slab((s + 1) * BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      double av[TM], bv[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) av[i] = As[buf][(kk + lk) * PA + wm0 + i * 16 + lr];
#pragma unroll
      for (int j = 0; j < TN; ++j) bv[j] = Bs[buf][(kk + lk) * PB + wn0 + j * 16 + lr];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[j], av[i], acc[i][j], 0, 0, 0);
    }
    if (s + 1 < nslab) store_slab(buf ^ 1);
    __syncthreads();
  }

  // D'[j][i] = sum_k B[k][j] A[i][k] = C[i][j].  The f64 16x16x4 MFMA leaves
  // D'[4*r + l/16][l%16] in lane l, register r: C row (l%16), column
  // 4*r + l/16 of the sub-tile.  The epilogue runs per 16-row group: all of
  // the group's old C values are loaded before any store, so the loads are in
  // flight together (a load-then-store per element serialises on the memory
  // latency, which dominates short-K products such as the rank-2nb updates).
  const double alpha = g.alpha, beta = g.beta;
  double* __restrict__ C = Cg;
  const int64_t ldc = g.ldc;
  double* __restrict__ Ct = g.Ct;
  const bool readc = beta != 0.0;
  if (Ct) {
    // Ct offset of C(gi, gj) = row part (gi only) + column part (gj only);
    // both are recomputed for the stores rather than held in registers
    const int ct_nbt = g.ct_nbt, cm_cols = g.cm_cols;
    const int64_t ct_off = g.ct_off;
    auto col_part = [&](int gj) -> int64_t {
      const int64_t aj = ct_off + gj;
      const int64_t bj = aj >> 6;
      return (bj * ct_nbt - bj * (bj + 1) / 2) * 4096 + (aj & 63) * 64;
    };
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      const int gi = m0 + wm0 + i * 16 + lr;
      const int64_t ai = ct_off + gi;
      const int64_t bi = ai >> 6;
      double* __restrict__ ctrow = Ct + bi * 4096 + (ai & 63);
      double v[TN][4];
      unsigned mask = 0;
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int gj = n0 + wn0 + j * 16 + 4 * r + lk;
          const bool in = gi < M && gj < N && bi >= ((ct_off + gj) >> 6);
          mask |= in ? (1u << (j * 4 + r)) : 0u;
          v[j][r] = (in && readc) ? ctrow[col_part(gj)] : 0.0;
        }
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          v[j][r] = alpha * acc[i][j][r] + beta * v[j][r];
          // keep the update here, ahead of the conditional stores: sunk into
          // them, each would wait on every memory operation before it
          asm volatile("" ::"v"(v[j][r]));
        }
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          if (!(mask & (1u << (j * 4 + r)))) continue;
          const int gj = n0 + wn0 + j * 16 + 4 * r + lk;
          ctrow[col_part(gj)] = v[j][r];
          if (gj < cm_cols) C[gi + (int64_t)gj * ldc] = v[j][r];
        }
    }
  } else {
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      const int gi = m0 + wm0 + i * 16 + lr;
      double* __restrict__ crow = C + gi;
      double v[TN][4];
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int gj = n0 + wn0 + j * 16 + 4 * r + lk;
          v[j][r] = (readc && gi < M && gj < N) ? crow[(int64_t)gj * ldc] : 0.0;
        }
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          v[j][r] = alpha * acc[i][j][r] + beta * v[j][r];
          // keep the update here, ahead of the conditional stores: sunk into
          // them, each would wait on every memory operation before it
          asm volatile("" ::"v"(v[j][r]));
        }
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int gj = n0 + wn0 + j * 16 + 4 * r + lk;
          if (gi < M && gj < N) crow[(int64_t)gj * ldc] = v[j][r];
        }
    }
  }
}

template <int BM, int BN, bool GROUPED>
static hipError_t launch_sz(hipStream_t st, bool ta, bool tb, dim3 grid, const GemmDesc& one,
                            const GemmDesc* descs, const int* prefix, int ngroups) {
#define IMCOM_DG(TA_, TB_)                                                                   \
  hipLaunchKernelGGL((dgemm_kernel<TA_, TB_, BM, BN, GROUPED>), grid, dim3(256), 0, st, one, \
                     descs, prefix, ngroups)
  if (!ta && !tb)
    IMCOM_DG(false, false);
  else if (ta && !tb)
    IMCOM_DG(true, false);
  else if (!ta && tb)
    IMCOM_DG(false, true);
  else
    IMCOM_DG(true, true);
#undef IMCOM_DG
  return hipGetLastError();
}

hipError_t dgemm(hipStream_t st, bool ta, bool tb, int M, int N, int K, double alpha,
                 const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                 double* C, int64_t ldc) {
  if (M <= 0 || N <= 0) return hipSuccess;
  GemmDesc g{M, N, K, alpha, beta, A, lda, B, ldb, C, ldc};
  // 128x128 tiles when that still gives >= 2 tiles per CU, else 64x64
  const int64_t t128 = (int64_t)((M + 127) / 128) * ((N + 127) / 128);
  if (t128 >= 512) {
    dim3 grid((unsigned)t128);
    return launch_sz<128, 128, false>(st, ta, tb, grid, g, nullptr, nullptr, 0);
  }
  const int64_t t64 = (int64_t)((M + 63) / 64) * ((N + 63) / 64);
  dim3 grid((unsigned)t64);
  return launch_sz<64, 64, false>(st, ta, tb, grid, g, nullptr, nullptr, 0);
}

hipError_t dgemm_lower(hipStream_t st, bool ta, bool tb, int M, int K, double alpha,
                       const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                       double* C, int64_t ldc, double* Ct, int64_t ct_off, int ct_nbt,
                       int cm_cols) {
  if (M <= 0) return hipSuccess;
  GemmDesc g{M, M, K, alpha, beta, A, lda, B, ldb, C, ldc, 1, Ct, ct_off, ct_nbt, cm_cols};
  const int64_t T128 = (M + 127) / 128;
  if (T128 * (T128 + 1) / 2 >= 512) {
    dim3 grid((unsigned)(T128 * (T128 + 1) / 2));
    return launch_sz<128, 128, false>(st, ta, tb, grid, g, nullptr, nullptr, 0);
  }
  const int64_t T64 = (M + 63) / 64;
  dim3 grid((unsigned)(T64 * (T64 + 1) / 2));
  return launch_sz<64, 64, false>(st, ta, tb, grid, g, nullptr, nullptr, 0);
}

// C = alpha * sum_s W_s in a fixed order (W_s: M x N, leading dimension M)
__global__ __launch_bounds__(256) void splitk_reduce_kernel(int M, int N, int ksplit,
                                                            const double* __restrict__ W,
                                                            double alpha, double* __restrict__ C,
                                                            int64_t ldc) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t mn = (int64_t)M * N;
  if (idx >= mn) return;
  double s = 0.0;
  for (int k = 0; k < ksplit; ++k) s += W[idx + k * mn];
  const int64_t i = idx % M, j = idx / M;
  C[i + j * ldc] = alpha * s;
}

// K chunks of at least 512, about 1024 blocks in all; 128x128 tiles unless
// that leaves fewer than 256 blocks
static void splitk_shape(int M, int N, int K, int* bm, int* bn, int* ksplit) {
  const int maxs = std::max(1, K / 512);
  const int64_t t128 = (int64_t)((M + 127) / 128) * ((N + 127) / 128);
  int s = (int)std::min<int64_t>(maxs, std::max<int64_t>(1, (1024 + t128 - 1) / t128));
  // a 128-wide tile over a dimension of at most 64 would compute half zeros
  if (t128 * s >= 256 && M > 64 && N > 64) {
    *bm = *bn = 128;
    *ksplit = s;
    return;
  }
  // one column of 64: 128 x 64 tiles (a wavefront's 4 x 2 MFMA tiles take
  // eight products per six operand reads, against four per four on 64 x 64)
  if (N <= 64 && M > 64) {
    const int64_t t = (M + 127) / 128;
    *bm = 128;
    *bn = 64;
    *ksplit = (int)std::min<int64_t>(maxs, std::max<int64_t>(1, (1024 + t - 1) / t));
    return;
  }
  const int64_t t64 = (int64_t)((M + 63) / 64) * ((N + 63) / 64);
  *bm = *bn = 64;
  *ksplit = (int)std::min<int64_t>(maxs, std::max<int64_t>(1, (1024 + t64 - 1) / t64));
}

size_t dgemm_splitk_work(int M, int N, int K) {
  int bm, bn, ks;
  splitk_shape(M, N, K, &bm, &bn, &ks);
  return (size_t)ks * M * N * sizeof(double);
}

hipError_t dgemm_splitk(hipStream_t st, bool ta, bool tb, int M, int N, int K, double alpha,
                        const double* A, int64_t lda, const double* B, int64_t ldb, double* C,
                        int64_t ldc, double* work) {
  if (M <= 0 || N <= 0) return hipSuccess;
  int bm, bn, ks;
  splitk_shape(M, N, K, &bm, &bn, &ks);
  const int kchunk = (K + ks - 1) / ks;
  ks = (K + kchunk - 1) / kchunk;  // no empty chunk
  GemmDesc g{M, N, K, 1.0, 0.0, A, lda, B, ldb, work, M};
  g.kchunk = kchunk;
  g.c_split_stride = (int64_t)M * N;
  if (bm == 128 && bn == 128) {
    dim3 grid((unsigned)(((M + 127) / 128) * ((N + 127) / 128)), (unsigned)ks);
    IMCOM_HIP_TRY((launch_sz<128, 128, false>(st, ta, tb, grid, g, nullptr, nullptr, 0)));
  } else if (bm == 128) {
    dim3 grid((unsigned)(((M + 127) / 128) * ((N + 63) / 64)), (unsigned)ks);
    IMCOM_HIP_TRY((launch_sz<128, 64, false>(st, ta, tb, grid, g, nullptr, nullptr, 0)));
  } else {
    dim3 grid((unsigned)(((M + 63) / 64) * ((N + 63) / 64)), (unsigned)ks);
    IMCOM_HIP_TRY((launch_sz<64, 64, false>(st, ta, tb, grid, g, nullptr, nullptr, 0)));
  }
  const int64_t mn = (int64_t)M * N;
  hipLaunchKernelGGL(splitk_reduce_kernel, dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, st, M,
                     N, ks, work, alpha, C, ldc);
  return hipGetLastError();
}

static int gemm_tiles(int M, int N, int tile) {
  return ((M + tile - 1) / tile) * ((N + tile - 1) / tile);
}

static int gemm_tiles(const GemmDesc& g, int tile) {
  if (g.lower) {
    const int T = (g.M + tile - 1) / tile;
    return T * (T + 1) / 2;
  }
  return gemm_tiles(g.M, g.N, tile);
}

int gemm_tiles64(int M, int N) { return gemm_tiles(M, N, 64); }

int gemm_tiles64(const GemmDesc& g) { return gemm_tiles(g, 64); }

hipError_t dgemm_grouped(hipStream_t st, bool ta, bool tb, const GemmDesc* d_descs,
                         const int* d_prefix, int ngroups, int total_blocks, int tile) {
  if (ngroups <= 0 || total_blocks <= 0) return hipSuccess;
  GemmDesc dummy{};
  dim3 grid((unsigned)total_blocks);
  if (tile == 128)
    return launch_sz<128, 128, true>(st, ta, tb, grid, dummy, d_descs, d_prefix, ngroups);
  return launch_sz<64, 64, true>(st, ta, tb, grid, dummy, d_descs, d_prefix, ngroups);
}

}  // namespace imcom

namespace imcom {

// 128x128 tiles once they give at least two workgroups per CU over the
// launch; smaller launches keep 64x64 tiles to fill the GPU.
constexpr int kGrouped128Min = 512;

int GroupedPlan::add(bool ta, bool tb, const std::vector<GemmDesc>& groups) {
  Launch L{(int)descs.size(), (int)prefix.size(), 0, 0, ta, tb, 64};
  int blocks128 = 0;
  for (const GemmDesc& g : groups)
    if (g.M > 0 && g.N > 0) blocks128 += gemm_tiles(g, 128);
  if (blocks128 >= kGrouped128Min) L.tile = 128;
  prefix.push_back(0);
  for (const GemmDesc& g : groups) {
    if (g.M <= 0 || g.N <= 0) continue;
    descs.push_back(g);
    L.blocks += gemm_tiles(g, L.tile);
    prefix.push_back(L.blocks);
    L.ngroups += 1;
  }
  launches.push_back(L);
  return (int)launches.size() - 1;
}

hipError_t GroupedPlan::upload(hipStream_t st) {
  release();
  if (!descs.empty()) {
    IMCOM_HIP_TRY(hipMalloc(&d_descs, descs.size() * sizeof(GemmDesc)));
    IMCOM_HIP_TRY(hipMemcpyAsync(d_descs, descs.data(), descs.size() * sizeof(GemmDesc),
                                 hipMemcpyHostToDevice, st));
  }
  if (!prefix.empty()) {
    IMCOM_HIP_TRY(hipMalloc(&d_prefix, prefix.size() * sizeof(int)));
    IMCOM_HIP_TRY(hipMemcpyAsync(d_prefix, prefix.data(), prefix.size() * sizeof(int),
                                 hipMemcpyHostToDevice, st));
  }
  return hipStreamSynchronize(st);
}

hipError_t GroupedPlan::run(hipStream_t st, int i) const {
  const Launch& L = launches[i];
  if (L.ngroups == 0) return hipSuccess;
  return dgemm_grouped(st, L.ta, L.tb, d_descs + L.desc_off, d_prefix + L.prefix_off, L.ngroups,
                       L.blocks, L.tile);
}

void GroupedPlan::release() {
  if (d_descs) (void)hipFree(d_descs);
  if (d_prefix) (void)hipFree(d_prefix);
  d_descs = nullptr;
  d_prefix = nullptr;
}

}  // namespace imcom
