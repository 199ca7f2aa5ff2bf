// Device helpers shared by the persistent kernels: agent-scope coherent
// loads and stores, DPP wavefront sums, dlapy2.
#pragma once
#include <hip/hip_runtime.h>

namespace imcom {

// A double stored and loaded with agent-scope atomics (coherent across the
// XCDs' L2 caches).
__device__ __forceinline__ void st_coh(double* p, double v) {
  __hip_atomic_store(reinterpret_cast<unsigned long long*>(p),
                     static_cast<unsigned long long>(__double_as_longlong(v)), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ld_coh(const double* p) {
  return __longlong_as_double(static_cast<long long>(__hip_atomic_load(
      reinterpret_cast<const unsigned long long*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)));
}

__device__ __forceinline__ double dlapy2_t(double x, double y) {
  const double xa = fabs(x), ya = fabs(y);
  const double w = fmax(xa, ya), z = fmin(xa, ya);
  if (z == 0.0) return w;
  const double q = z / w;
  return w * sqrt(1.0 + q * q);
}

// A double moved between lanes by a DPP pattern (both 32-bit halves).
template <int CTRL>
__device__ __forceinline__ double dpp_d(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double readlane_d(double x, int lane) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
// Workgroup barrier for data exchanged through the LDS only: waits for the
// wavefront's LDS operations (lgkmcnt 0; vmcnt and expcnt left at their
// maxima) and meets the other wavefronts, without the vmcnt(0) wait that
// __syncthreads' workgroup-scope fences imply, so global loads issued before
// it stay in flight.  The empty asm statements keep the compiler from moving
// memory accesses across it.
__device__ __forceinline__ void lds_sync() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0xc07f);
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// Sum over the 64 lanes of a wavefront with DPP moves (a few cycles each,
// against a round trip through the LDS unit for each step of a shuffle
// butterfly): pairs, quads, half-rows and rows of 16 by the quad_perm and
// mirror patterns, whose partner sums are commutative, so that every lane of
// a row holds the same row sum; then the four rows in order.  Fixed order,
// the same value in every lane.
__device__ __forceinline__ double wave_sum_dpp(double v) {
  v += dpp_d<0xb1>(v);   // quad_perm [1, 0, 3, 2]
  v += dpp_d<0x4e>(v);   // quad_perm [2, 3, 0, 1]
  v += dpp_d<0x141>(v);  // row_half_mirror
  v += dpp_d<0x140>(v);  // row_mirror
  return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
}
// Sum over each group of four consecutive lanes (every lane of the quad
// holds it; the two partner sums are commutative, so the order is fixed).
__device__ __forceinline__ double quad_sum_dpp(double v) {
  v += dpp_d<0xb1>(v);  // quad_perm [1, 0, 3, 2]
  v += dpp_d<0x4e>(v);  // quad_perm [2, 3, 0, 1]
  return v;
}

}  // namespace imcom
