// Shared definitions for the IMCOM solve-path kernels (gfx950 / CDNA4).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace imcom {

// 64-wide wavefronts on CDNA.
constexpr int kWave = 64;

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// Error carrying: launchers return hipError_t; the C-ABI layer turns it into
// a status code plus a message.
struct Status {
  int code = 0;
  std::string msg;
  bool ok() const { return code == 0; }
};

#define IMCOM_HIP_TRY(expr)                                   \
  do {                                                        \
    hipError_t _e = (expr);                                   \
    if (_e != hipSuccess) return _e;                          \
  } while (0)

// A load through the global address space (global_load with a scalar or
// vector base); a pointer read from a struct would otherwise give a flat load.
__device__ __forceinline__ double ldg(const double* p) {
  return *(const __attribute__((address_space(1))) double*)p;
}

__device__ __forceinline__ double wave_sum(double v) {
  // butterfly over the 64 lanes of a wavefront
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Block-wide sum with a fixed reduction order (deterministic): each wave
// reduces, then wave 0 sums the per-wave partials in wave order.  `red` must
// hold blockDim.x/64 doubles.  Every thread gets the result.
__device__ __forceinline__ double block_sum(double v, double* red) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < nw; ++w) s += red[w];
  return s;
}

inline int ceil_div(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

}  // namespace imcom
