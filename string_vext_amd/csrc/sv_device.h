// Device helpers shared by the gfx950 kernels of libsv_vext.so.
#pragma once
#include "sv_internal.h"

namespace sv {

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Number of set bits of m in lanes below the calling lane.
__device__ __forceinline__ int lanes_below(unsigned long long m) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
// Wave-uniform values loaded through pointers (vector loads) or computed by
// VALU land in vector registers; readfirstlane moves them to scalar registers.
__device__ __forceinline__ float uniform_f(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ double uniform_d(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)b);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// p[i] read through the constant address space, for data no kernel writes
// while it runs (plan and pose tables): with a wave-uniform address the
// compiler issues a scalar load through the scalar cache instead of a vector
// load and a readfirstlane.
template <class T>
__device__ __forceinline__ T ld_scalar(const T* p, long long i) {
  return ((const __attribute__((address_space(4))) T*)p)[i];
}

// Blocks are dispatched round-robin over the 8 XCDs (block b on XCD b % 8).
// Map hardware block b to a logical block so that runs of kXcdRun consecutive
// logical blocks (same tile, same candidate list) execute on one XCD and share
// its L2, while the runs rotate over the XCDs: every XCD gets an even share of
// each part of the work range, so batches sorted by cost stay balanced.  The
// last incomplete round maps to itself.
constexpr long long kXcdRun = 64;
__device__ __forceinline__ long long xcd_block(long long b, long long nb) {
  const long long round = 8 * kXcdRun;
  if (b >= nb / round * round) return b;
  const long long x = b % 8, i = b / 8;
  return ((i / kXcdRun) * 8 + x) * kXcdRun + i % kXcdRun;
}

}  // namespace sv
