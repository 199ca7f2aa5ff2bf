// Split-precision GEMM for the Q^T B and Q W products (imcom_matrix.f90:699
// and :731-745) on the low-precision matrix cores.
//
//   scheme 2, BF16 x6:  a = a1 + a2 + a3 (bf16 pieces, each the rounded
//     residual of the previous), likewise b; the six products with
//     i + j <= 4 (1-based) run on v_mfma_f32_16x16x32_bf16 with FP32
//     accumulation in three magnitude groups (a1b1 | a1b2+a2b1 |
//     a1b3+a2b2+a3b1), recombined in FP64 every 256 k; the last two groups
//     share one FP32 accumulator in the 128 x 128-tile kernel.
//   scheme 1, FP32 x3:  a = a1 + a2 (fp32 pieces), products a1b1 | a1b2+a2b1
//     on v_mfma_f32_16x16x4_f32, the same FP64 recombination.
//
// Operands are first split into K-contiguous piece arrays (rows of op(A),
// columns of op(B)), zero-padded to the tile grid (the bf16 pieces in 32-k
// slabs), so every MFMA operand is one 16-byte (bf16) or 4-byte (fp32) read
// from LDS.  The FP64 result is
// C = alpha * sum; the result tile layout and store path follow the FP64
// kernel (lane l, register r -> C row l%16, column 4*(l/16)+r for the f32
// MFMA result layout D[4*(l/16)+r][l%16] with op(B) as the first operand).
#include "common.h"
#include "kernels.h"

#include <algorithm>

namespace imcom {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

constexpr int ST = 64;     // C tile (ST x ST) per workgroup
constexpr int FLUSH = 8;   // slabs between FP64 recombinations

static inline int64_t pad_to(int64_t v, int64_t q) { return (v + q - 1) / q * q; }

// ---- splitting ---------------------------------------------------------------
// Piece p of X(r, k), where X(r, k) = kcontig ? X[k + r*ldx] : X[r + k*ldx],
// zero for r >= R or k >= K (padding), goes to out_p[k + r * ldo] (srows = 0)
// or, in the 32-k slab layout of the bf16 pieces (srows = padded row count),
// to out_p[((k / 32) * srows + r) * 32 + k % 32].  64x64 tiles via LDS so both
// the read and the write are coalesced.
template <int SCHEME>
__global__ __launch_bounds__(256) void split_kernel(const double* __restrict__ X, int64_t ldx,
                                                    int kcontig, int R, int K, void* out0,
                                                    void* out1, void* out2, int64_t ldo,
                                                    int64_t srows) {
  __shared__ double tile[64][65];
  const int r0 = blockIdx.y * 64, k0 = blockIdx.x * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  // load tile[rr][kk] = X(r0+rr, k0+kk)
  for (int q = ty; q < 64; q += 4) {
    if (kcontig) {
      const int r = r0 + q, k = k0 + tx;
      tile[q][tx] = (r < R && k < K) ? X[k + (int64_t)r * ldx] : 0.0;
    } else {
      const int r = r0 + tx, k = k0 + q;
      tile[tx][q] = (r < R && k < K) ? X[r + (int64_t)k * ldx] : 0.0;
    }
  }
  __syncthreads();
  for (int q = ty; q < 64; q += 4) {
    const double v = tile[q][tx];
    const int64_t r = r0 + q, k = k0 + tx;
    const int64_t o = srows > 0 ? ((k >> 5) * srows + r) * 32 + (k & 31) : r * ldo + k;
    if (SCHEME == 2) {
      const __bf16 p1 = (__bf16)(float)v;
      const double r1 = v - (double)(float)p1;
      const __bf16 p2 = (__bf16)(float)r1;
      const double r2 = r1 - (double)(float)p2;
      const __bf16 p3 = (__bf16)(float)r2;
      ((__bf16*)out0)[o] = p1;
      ((__bf16*)out1)[o] = p2;
      ((__bf16*)out2)[o] = p3;
    } else {
      const float p1 = (float)v;
      const float p2 = (float)(v - (double)p1);
      ((float*)out0)[o] = p1;
      ((float*)out1)[o] = p2;
    }
  }
}

// ---- BF16 x6 GEMM, 128 x 128 tiles ------------------------------------------
// Pieces in the 32-k slab layout: one tile's slab of a piece (128 rows x 32 k)
// is 8 KB contiguous, two 16-byte loads per thread.  Four wavefronts (2 x 2)
// own 64 x 64 each as 4 x 4 MFMA sub-tiles; per slab and sub-tile the six
// products go to two FP32 accumulators (a1 b1 | the five smaller ones: their
// sum sits 2^-8 below a1 b1, so one FP32 accumulator keeps them to well below
// the a1 b1 accumulator's own rounding), recombined in FP64 every FLUSH
// slabs.  The products are issued product-major over the 16 sub-tiles, so no
// MFMA waits on the one before it.  Double-buffered LDS (120 KB), one barrier
// per slab; one workgroup per CU (up to 512 registers per lane).
constexpr int TB6 = 128;
constexpr int KS6 = 32;
constexpr int LD6 = KS6 + 8;  // padded LDS row (bf16): 80 B, conflict-free 16-byte reads

__global__ __launch_bounds__(256, 1) void gemm_bf16x6_t128_kernel(
    int M, int N, int Kp, double alpha, const __bf16* __restrict__ A1, const __bf16* __restrict__ A2,
    const __bf16* __restrict__ A3, const __bf16* __restrict__ B1, const __bf16* __restrict__ B2,
    const __bf16* __restrict__ B3, int64_t Mp, int64_t Np, double* __restrict__ C, int64_t ldc,
    int tiles_m, int tiles_n) {
  __shared__ __attribute__((aligned(16))) __bf16 As[2][3][TB6 * LD6];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[2][3][TB6 * LD6];
  int b = blockIdx.x;
  const int nt = tiles_m * tiles_n;
  // blocks go round-robin to the 8 XCDs: give each XCD a contiguous run of
  // tiles, walked column-major in groups of 8 tile rows, so neighbouring
  // tiles share A and B slabs in its L2
  if ((nt & 7) == 0) b = (b & 7) * (nt >> 3) + (b >> 3);
  const int group = 8 * tiles_n;
  const int first_m = (b / group) * 8;
  const int gm = min(tiles_m - first_m, 8);
  const int tm = first_m + (b % group) % gm, tn = (b % group) / gm;
  const int64_t m0 = (int64_t)tm * TB6, n0 = (int64_t)tn * TB6;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int wm0 = (wid & 1) * 64, wn0 = (wid >> 1) * 64;
  // slab sl of each piece for this tile: 512 16-byte chunks, two per thread
  const int64_t offA = m0 * KS6, offB = n0 * KS6, strA = Mp * KS6, strB = Np * KS6;
  uint4 a10, a11, a20, a21, a30, a31, b10, b11, b20, b21, b30, b31;
  auto load = [&](int sl) {
    const uint4* p;
    p = (const uint4*)(A1 + offA + sl * strA); a10 = p[tid]; a11 = p[tid + 256];
    p = (const uint4*)(A2 + offA + sl * strA); a20 = p[tid]; a21 = p[tid + 256];
    p = (const uint4*)(A3 + offA + sl * strA); a30 = p[tid]; a31 = p[tid + 256];
    p = (const uint4*)(B1 + offB + sl * strB); b10 = p[tid]; b11 = p[tid + 256];
    p = (const uint4*)(B2 + offB + sl * strB); b20 = p[tid]; b21 = p[tid + 256];
    p = (const uint4*)(B3 + offB + sl * strB); b30 = p[tid]; b31 = p[tid + 256];
  };
  // chunk q = tid (+256): row q / 4, 8-element chunk q % 4
  const int so0 = (tid >> 2) * LD6 + (tid & 3) * 8, so1 = so0 + 64 * LD6;
  auto store = [&](int buf) {
    *(uint4*)(&As[buf][0][so0]) = a10; *(uint4*)(&As[buf][0][so1]) = a11;
    *(uint4*)(&As[buf][1][so0]) = a20; *(uint4*)(&As[buf][1][so1]) = a21;
    *(uint4*)(&As[buf][2][so0]) = a30; *(uint4*)(&As[buf][2][so1]) = a31;
    *(uint4*)(&Bs[buf][0][so0]) = b10; *(uint4*)(&Bs[buf][0][so1]) = b11;
    *(uint4*)(&Bs[buf][1][so0]) = b20; *(uint4*)(&Bs[buf][1][so1]) = b21;
    *(uint4*)(&Bs[buf][2][so0]) = b30; *(uint4*)(&Bs[buf][2][so1]) = b31;
  };
  v4f hi[4][4], lo[4][4];
  v4d acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      hi[i][j] = lo[i][j] = (v4f){0.f, 0.f, 0.f, 0.f};
      acc[i][j] = (v4d){0.0, 0.0, 0.0, 0.0};
    }
  const int nslab = Kp / KS6;
  load(0);
  store(0);
  __syncthreads();
  const int lr = lane & 15, lk = (lane >> 4) * 8;
  for (int sl = 0; sl < nslab; ++sl) {
    const int buf = sl & 1;
    if (sl + 1 < nslab) load(sl + 1);
    v8bf a[4][3], bb[4][3];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int p = 0; p < 3; ++p) {
        a[i][p] = *(const v8bf*)(&As[buf][p][(wm0 + i * 16 + lr) * LD6 + lk]);
        bb[i][p] = *(const v8bf*)(&Bs[buf][p][(wn0 + i * 16 + lr) * LD6 + lk]);
      }
#define IMCOM_P6(ACC, PB, PA)                                                                   \
  _Pragma("unroll") for (int i = 0; i < 4; ++i) _Pragma("unroll") for (int j = 0; j < 4; ++j)    \
      ACC[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bb[j][PB], a[i][PA], ACC[i][j], 0, 0, 0);
    IMCOM_P6(hi, 0, 0)
    IMCOM_P6(lo, 1, 0)
    IMCOM_P6(lo, 0, 1)
    IMCOM_P6(lo, 2, 0)
    IMCOM_P6(lo, 1, 1)
    IMCOM_P6(lo, 0, 2)
#undef IMCOM_P6
    if ((sl + 1) % FLUSH == 0 || sl + 1 == nslab) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
#pragma unroll
          for (int r = 0; r < 4; ++r) acc[i][j][r] += (double)hi[i][j][r] + (double)lo[i][j][r];
          hi[i][j] = lo[i][j] = (v4f){0.f, 0.f, 0.f, 0.f};
        }
    }
    if (sl + 1 < nslab) store(buf ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gi = m0 + wm0 + i * 16 + lr;
        const int64_t gj = n0 + wn0 + j * 16 + 4 * (lane >> 4) + r;
        if (gi < M && gj < N) C[gi + gj * ldc] = alpha * acc[i][j][r];
      }
}

// ---- FP32 x3 GEMM -----------------------------------------------------------
constexpr int BKF = 16;         // k per slab (four 16x16x4 steps)
constexpr int LDF = BKF + 1;    // padded LDS row (floats)

__global__ __launch_bounds__(256) void gemm_f32x3_kernel(
    int M, int N, int Kp, double alpha, const float* __restrict__ A1, const float* __restrict__ A2,
    const float* __restrict__ B1, const float* __restrict__ B2, int64_t ldp, double* __restrict__ C,
    int64_t ldc, int tiles_m) {
  __shared__ float As[2][2][ST * LDF];
  __shared__ float Bs[2][2][ST * LDF];
  const int tm = blockIdx.x % tiles_m, tn = blockIdx.x / tiles_m;
  const int m0 = tm * ST, n0 = tn * ST;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int wm0 = (wid & 1) * 32, wn0 = (wid >> 1) * 32;
  const float* Ap[2] = {A1, A2};
  const float* Bp[2] = {B1, B2};
  // 64 rows x 16 k = 1024 floats per piece: 4 per thread (row tid/4, k 4*(tid%4)..)
  const int lrow = tid >> 2, lk4 = (tid & 3) * 4;
  float4 ra[2], rb[2];
  auto load = [&](int k0) {
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      ra[p] = *(const float4*)(Ap[p] + (int64_t)(m0 + lrow) * ldp + k0 + lk4);
      rb[p] = *(const float4*)(Bp[p] + (int64_t)(n0 + lrow) * ldp + k0 + lk4);
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      float* a = &As[buf][p][lrow * LDF + lk4];
      float* b = &Bs[buf][p][lrow * LDF + lk4];
      a[0] = ra[p].x; a[1] = ra[p].y; a[2] = ra[p].z; a[3] = ra[p].w;
      b[0] = rb[p].x; b[1] = rb[p].y; b[2] = rb[p].z; b[3] = rb[p].w;
    }
  };
  v4f hi[2][2], mid[2][2];
  double acc[2][2][4];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      hi[i][j] = mid[i][j] = (v4f){0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[i][j][r] = 0.0;
    }
  const int nslab = Kp / BKF;
  load(0);
  store(0);
  __syncthreads();
  const int lr = lane & 15, lk = lane >> 4;
  for (int s = 0; s < nslab; ++s) {
    const int buf = s & 1;
    if (s + 1 < nslab) load((s + 1) * BKF);
#pragma unroll
    for (int kk = 0; kk < BKF; kk += 4) {
      float a[2][2], b[2][2];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int p = 0; p < 2; ++p) {
          a[i][p] = As[buf][p][(wm0 + i * 16 + lr) * LDF + kk + lk];
          b[i][p] = Bs[buf][p][(wn0 + i * 16 + lr) * LDF + kk + lk];
        }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          hi[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(b[j][0], a[i][0], hi[i][j], 0, 0, 0);
          mid[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(b[j][1], a[i][0], mid[i][j], 0, 0, 0);
          mid[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(b[j][0], a[i][1], mid[i][j], 0, 0, 0);
        }
    }
    if ((s + 1) % (2 * FLUSH) == 0 || s + 1 == nslab) {
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
#pragma unroll
          for (int r = 0; r < 4; ++r) acc[i][j][r] += (double)hi[i][j][r] + (double)mid[i][j][r];
          hi[i][j] = mid[i][j] = (v4f){0.f, 0.f, 0.f, 0.f};
        }
    }
    if (s + 1 < nslab) store(buf ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gi = m0 + wm0 + i * 16 + lr;
        const int gj = n0 + wn0 + j * 16 + 4 * (lane >> 4) + r;
        if (gi < M && gj < N) C[gi + (int64_t)gj * ldc] = alpha * acc[i][j][r];
      }
}

// ---- host side -----------------------------------------------------------------
struct SplitLayout {
  int64_t Mp, Np, Kp, esz, npieces;
  size_t a_bytes() const { return (size_t)(Mp * Kp * esz); }
  size_t b_bytes() const { return (size_t)(Np * Kp * esz); }
  size_t total() const { return npieces * (a_bytes() + b_bytes()); }
};

static SplitLayout layout(int scheme, int M, int N, int K) {
  SplitLayout L;
  L.Mp = pad_to(M, scheme == 2 ? TB6 : 64);
  L.Np = pad_to(N, scheme == 2 ? TB6 : 64);
  L.Kp = pad_to(K, 64);  // multiple of both slab depths and of the split tile
  L.esz = scheme == 2 ? 2 : 4;
  L.npieces = scheme == 2 ? 3 : 2;
  return L;
}

size_t gemm_split_work_bytes(int scheme, int M, int N, int K) {
  return layout(scheme, M, N, K).total() + 256;
}

hipError_t gemm_split(hipStream_t st, int scheme, bool ta, int M, int N, int K, double alpha,
                      const double* A, int64_t lda, const double* B, int64_t ldb, double* C,
                      int64_t ldc, void* work, size_t work_bytes) {
  if (scheme != 1 && scheme != 2) return hipErrorInvalidValue;
  if (M <= 0 || N <= 0) return hipSuccess;
  const SplitLayout L = layout(scheme, M, N, K);
  if (work_bytes < L.total()) return hipErrorInvalidValue;
  char* w = (char*)work;
  void* ap[3];
  void* bp[3];
  for (int p = 0; p < L.npieces; ++p) {
    ap[p] = w;
    w += L.a_bytes();
  }
  for (int p = 0; p < L.npieces; ++p) {
    bp[p] = w;
    w += L.b_bytes();
  }
  if (L.npieces == 2) ap[2] = bp[2] = nullptr;
  // rows of op(A): op(A)(i,k) = ta ? A[k + i*lda] (k-contiguous) : A[i + k*lda]
  // columns of op(B) (B not transposed): op(B)(k,j) = B[k + j*ldb] (k-contiguous)
  dim3 ga((unsigned)(L.Kp / 64), (unsigned)(L.Mp / 64)), gb((unsigned)(L.Kp / 64), (unsigned)(L.Np / 64));
  if (scheme == 2) {
    hipLaunchKernelGGL(split_kernel<2>, ga, dim3(256), 0, st, A, lda, ta ? 1 : 0, M, K, ap[0], ap[1],
                       ap[2], L.Kp, L.Mp);
    hipLaunchKernelGGL(split_kernel<2>, gb, dim3(256), 0, st, B, ldb, 1, N, K, bp[0], bp[1], bp[2],
                       L.Kp, L.Np);
    const int tiles_m = (int)(L.Mp / TB6), tiles_n = (int)(L.Np / TB6);
    hipLaunchKernelGGL(gemm_bf16x6_t128_kernel, dim3(tiles_m * tiles_n), dim3(256), 0, st, M, N,
                       (int)L.Kp, alpha, (const __bf16*)ap[0], (const __bf16*)ap[1],
                       (const __bf16*)ap[2], (const __bf16*)bp[0], (const __bf16*)bp[1],
                       (const __bf16*)bp[2], L.Mp, L.Np, C, ldc, tiles_m, tiles_n);
  } else {
    hipLaunchKernelGGL(split_kernel<1>, ga, dim3(256), 0, st, A, lda, ta ? 1 : 0, M, K, ap[0], ap[1],
                       nullptr, L.Kp, (int64_t)0);
    hipLaunchKernelGGL(split_kernel<1>, gb, dim3(256), 0, st, B, ldb, 1, N, K, bp[0], bp[1], nullptr,
                       L.Kp, (int64_t)0);
    const int tiles_m = (int)(L.Mp / ST), tiles_n = (int)(L.Np / ST);
    hipLaunchKernelGGL(gemm_f32x3_kernel, dim3(tiles_m * tiles_n), dim3(256), 0, st, M, N,
                       (int)L.Kp, alpha, (const float*)ap[0], (const float*)ap[1],
                       (const float*)bp[0], (const float*)bp[1], L.Kp, C, ldc, tiles_m);
  }
  return hipGetLastError();
}

}  // namespace imcom
