// Split-precision GEMM for the Q^T B and Q W products (imcom_matrix.f90:699
// and :731-745) on the low-precision matrix cores.
//
//   scheme 2, BF16 x6:  a = a1 + a2 + a3 (bf16 pieces, each the rounded
//     residual of the previous), likewise b; the six products with
//     i + j <= 4 (1-based) run on v_mfma_f32_16x16x32_bf16 with FP32
//     accumulation in three magnitude groups (a1b1 | a1b2+a2b1 |
//     a1b3+a2b2+a3b1), recombined in FP64 every 256 k.
//   scheme 1, FP32 x3:  a = a1 + a2 (fp32 pieces), products a1b1 | a1b2+a2b1
//     on v_mfma_f32_16x16x4_f32, the same FP64 recombination.
//
// Operands are first split into K-contiguous piece arrays (rows of op(A),
// columns of op(B)), zero-padded to the tile grid, so every MFMA operand is
// one 16-byte (bf16) or 4-byte (fp32) read from LDS.  The FP64 result is
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
// out_p[k + r * ldo] = piece p of X(r, k) where X(r, k) = kcontig ? X[k + r*ldx]
// : X[r + k*ldx]; zero for r >= R or k >= K (padding).  64x64 tiles via LDS so
// both the read and the write are coalesced.
template <int SCHEME>
__global__ __launch_bounds__(256) void split_kernel(const double* __restrict__ X, int64_t ldx,
                                                    int kcontig, int R, int K, void* out0,
                                                    void* out1, void* out2, int64_t ldo) {
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
    const int64_t o = (int64_t)(r0 + q) * ldo + (k0 + tx);
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

// ---- BF16 x6 GEMM ---------------------------------------------------------
// A pieces: [Mp][Kp] bf16 (k contiguous); B pieces: [Np][Kp].
constexpr int BKB = 32;          // k per slab (one 16x16x32 MFMA step)
constexpr int LDB16 = BKB + 8;   // padded LDS row (bf16 elements)

__global__ __launch_bounds__(256) void gemm_bf16x6_kernel(
    int M, int N, int Kp, double alpha, const __bf16* __restrict__ A1, const __bf16* __restrict__ A2,
    const __bf16* __restrict__ A3, const __bf16* __restrict__ B1, const __bf16* __restrict__ B2,
    const __bf16* __restrict__ B3, int64_t ldp, double* __restrict__ C, int64_t ldc, int tiles_m) {
  __shared__ __attribute__((aligned(16))) __bf16 As[2][3][ST * LDB16];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[2][3][ST * LDB16];
  const int tm = blockIdx.x % tiles_m, tn = blockIdx.x / tiles_m;
  const int m0 = tm * ST, n0 = tn * ST;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int wm0 = (wid & 1) * 32, wn0 = (wid >> 1) * 32;
  const __bf16* Ap[3] = {A1, A2, A3};
  const __bf16* Bp[3] = {B1, B2, B3};
  // each thread moves one 16-byte chunk per piece per operand: row tid/4, chunk tid%4
  const int lrow = tid >> 2, lch = tid & 3;
  uint4 ra[3], rb[3];
  auto load = [&](int k0) {
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      ra[p] = *(const uint4*)(Ap[p] + (int64_t)(m0 + lrow) * ldp + k0 + lch * 8);
      rb[p] = *(const uint4*)(Bp[p] + (int64_t)(n0 + lrow) * ldp + k0 + lch * 8);
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      *(uint4*)(&As[buf][p][lrow * LDB16 + lch * 8]) = ra[p];
      *(uint4*)(&Bs[buf][p][lrow * LDB16 + lch * 8]) = rb[p];
    }
  };
  v4f hi[2][2], mid[2][2], lo[2][2];
  double acc[2][2][4];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      hi[i][j] = mid[i][j] = lo[i][j] = (v4f){0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[i][j][r] = 0.0;
    }
  const int nslab = Kp / BKB;
  load(0);
  store(0);
  __syncthreads();
  const int lr = lane & 15, lk = (lane >> 4) * 8;
  for (int s = 0; s < nslab; ++s) {
    const int buf = s & 1;
    if (s + 1 < nslab) load((s + 1) * BKB);
    v8bf a[2][3], b[2][3];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int p = 0; p < 3; ++p) {
        a[i][p] = *(const v8bf*)(&As[buf][p][(wm0 + i * 16 + lr) * LDB16 + lk]);
        b[i][p] = *(const v8bf*)(&Bs[buf][p][(wn0 + i * 16 + lr) * LDB16 + lk]);
      }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        hi[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[j][0], a[i][0], hi[i][j], 0, 0, 0);
        mid[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[j][1], a[i][0], mid[i][j], 0, 0, 0);
        mid[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[j][0], a[i][1], mid[i][j], 0, 0, 0);
        lo[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[j][2], a[i][0], lo[i][j], 0, 0, 0);
        lo[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[j][1], a[i][1], lo[i][j], 0, 0, 0);
        lo[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[j][0], a[i][2], lo[i][j], 0, 0, 0);
      }
    if ((s + 1) % FLUSH == 0 || s + 1 == nslab) {
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
#pragma unroll
          for (int r = 0; r < 4; ++r)
            acc[i][j][r] += ((double)hi[i][j][r] + (double)mid[i][j][r]) + (double)lo[i][j][r];
          hi[i][j] = mid[i][j] = lo[i][j] = (v4f){0.f, 0.f, 0.f, 0.f};
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
  L.Mp = pad_to(M, 64);
  L.Np = pad_to(N, 64);
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
                       ap[2], L.Kp);
    hipLaunchKernelGGL(split_kernel<2>, gb, dim3(256), 0, st, B, ldb, 1, N, K, bp[0], bp[1], bp[2],
                       L.Kp);
    const int tiles_m = (int)(L.Mp / ST), tiles_n = (int)(L.Np / ST);
    hipLaunchKernelGGL(gemm_bf16x6_kernel, dim3(tiles_m * tiles_n), dim3(256), 0, st, M, N,
                       (int)L.Kp, alpha, (const __bf16*)ap[0], (const __bf16*)ap[1],
                       (const __bf16*)ap[2], (const __bf16*)bp[0], (const __bf16*)bp[1],
                       (const __bf16*)bp[2], L.Kp, C, ldc, tiles_m);
  } else {
    hipLaunchKernelGGL(split_kernel<1>, ga, dim3(256), 0, st, A, lda, ta ? 1 : 0, M, K, ap[0], ap[1],
                       nullptr, L.Kp);
    hipLaunchKernelGGL(split_kernel<1>, gb, dim3(256), 0, st, B, ldb, 1, N, K, bp[0], bp[1], nullptr,
                       L.Kp);
    const int tiles_m = (int)(L.Mp / ST), tiles_n = (int)(L.Np / ST);
    hipLaunchKernelGGL(gemm_f32x3_kernel, dim3(tiles_m * tiles_n), dim3(256), 0, st, M, N,
                       (int)L.Kp, alpha, (const float*)ap[0], (const float*)ap[1],
                       (const float*)bp[0], (const float*)bp[1], L.Kp, C, ldc, tiles_m);
  }
  return hipGetLastError();
}

}  // namespace imcom
