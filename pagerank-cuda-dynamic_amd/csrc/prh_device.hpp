// prh_device.hpp -- device-side building blocks for the gfx950 (CDNA4, wave64) PageRank kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "prh_internal.hpp"

namespace prh {

// ---------------------------------------------------------------------------------------------------------
// Cross-lane moves. DPP (data-parallel primitives) moves within a 16-lane row are free operand modifiers on
// VALU moves; ds_swizzle crosses the two rows of a 32-lane half; readlane crosses the halves. A 64-bit value
// moves as its two 32-bit halves. All lanes of the wave must be active when these are called.
// ---------------------------------------------------------------------------------------------------------
enum : int {
  kDppQuadSwap1 = 0xB1,       // quad_perm [1,0,3,2]: lane i <- lane i^1
  kDppQuadSwap2 = 0x4E,       // quad_perm [2,3,0,1]: lane i <- lane i^2
  kDppRowHalfMirror = 0x141,  // lane i <- lane 7-i within each 8-lane half-row
  kDppRowMirror = 0x140,      // lane i <- lane 15-i within each 16-lane row
};

template <int CTRL>
__device__ __forceinline__ double dpp_move(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, int(b), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, int(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// ds_swizzle in bit-mask mode with xor 16 (offset = and 0x1F | xor 0x10 << 10): lane i <- lane i^16 within 32.
__device__ __forceinline__ double swizzle_xor16(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_ds_swizzle(int(b), 0x401F);
  const int hi = __builtin_amdgcn_ds_swizzle(int(b >> 32), 0x401F);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

__device__ __forceinline__ double read_lane(double x, int lane) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane(int(b), lane);
  const int hi = __builtin_amdgcn_readlane(int(b >> 32), lane);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Sum over aligned groups of W lanes (W = 1, 2, 4, ..., 64). Each step pairs lanes symmetrically, so every lane
// of a group ends with the same value (for W = 64 the result is wave-uniform).
template <int W>
__device__ __forceinline__ double group_sum(double x) {
  if constexpr (W >= 2) x += dpp_move<kDppQuadSwap1>(x);
  if constexpr (W >= 4) x += dpp_move<kDppQuadSwap2>(x);
  if constexpr (W >= 8) x += dpp_move<kDppRowHalfMirror>(x);
  if constexpr (W >= 16) x += dpp_move<kDppRowMirror>(x);
  if constexpr (W >= 32) x += swizzle_xor16(x);
  if constexpr (W >= 64) x = read_lane(x, 0) + read_lane(x, 32);
  return x;
}

template <int W>
__device__ __forceinline__ double group_max(double x) {
  if constexpr (W >= 2) x = fmax(x, dpp_move<kDppQuadSwap1>(x));
  if constexpr (W >= 4) x = fmax(x, dpp_move<kDppQuadSwap2>(x));
  if constexpr (W >= 8) x = fmax(x, dpp_move<kDppRowHalfMirror>(x));
  if constexpr (W >= 16) x = fmax(x, dpp_move<kDppRowMirror>(x));
  if constexpr (W >= 32) x = fmax(x, swizzle_xor16(x));
  if constexpr (W >= 64) x = fmax(read_lane(x, 0), read_lane(x, 32));
  return x;
}

__device__ __forceinline__ int lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0)); }

// Number of set bits of a 64-bit wave mask below this lane.
__device__ __forceinline__ unsigned lanes_below(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi(unsigned(mask >> 32), __builtin_amdgcn_mbcnt_lo(unsigned(mask), 0));
}

// Workgroup-wide max of a non-negative double, result valid in thread 0. Every thread must call it.
__device__ __forceinline__ double block_max(double x, double *lds4) {
  x = group_max<kWave>(x);
  const int w = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) lds4[w] = x;
  __syncthreads();
  double m = 0.0;
  if (threadIdx.x == 0) m = fmax(fmax(lds4[0], lds4[1]), fmax(lds4[2], lds4[3]));
  return m;
}

// Non-negative doubles order like their bit patterns, so an unsigned 64-bit max is an exact L-inf reduction.
// The slot only grows, so any value read from it is a lower bound of its final value: a workgroup whose maximum
// does not exceed what it reads has nothing to add and skips the device-scope atomic (tens of thousands of
// workgroups per launch would otherwise queue on one address).
__device__ __forceinline__ void atomic_max_nonneg(unsigned long long *slot, double x) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  if (x > 0.0 && b > *(const unsigned long long *)slot) atomicMax(slot, b);
}

// ---------------------------------------------------------------------------------------------------------
// Iteration control on the device. err[l] collects iteration l's L-inf change as bits; *done records the first
// iteration index l at which err[l-1] < tolerance was seen, after which every kernel of later iterations exits
// at once. This keeps "stop at the first iteration whose change is below tolerance" without a host round trip.
// ---------------------------------------------------------------------------------------------------------
struct LoopCtl {
  unsigned long long *err;  // [max iterations] L-inf bits, zeroed before the loop
  int *done;                // 0 while running, else the number of iterations performed
  int l;                    // this launch's iteration index, or -1: read it from done[1] (replayed graphs, where
                            // the same launch serves many iterations and a device counter says which)
  double tol;
};

__device__ __forceinline__ int iter_index(const LoopCtl &c) { return c.l >= 0 ? c.l : c.done[1]; }

// Entry test for the first kernel of iteration l; all threads of a block agree.
// done and err[l-1] were last written by earlier launches, whose completion made them visible: plain loads (both
// issued together) suffice. A volatile read here would be a system-scope load at the start of every wave.
__device__ __forceinline__ bool iteration_begin(const LoopCtl &c) {
  const int done = *c.done;
  const int l = iter_index(c);
  const double e = l > 0 ? __longlong_as_double((long long)c.err[l - 1]) : c.tol;
  if (done) return false;
  if (e < c.tol) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *(volatile int *)c.done = l;
    return false;
  }
  return true;
}

// Entry test for kernels that follow the rank update within iteration l (frontier extension).
__device__ __forceinline__ bool iteration_continues(const LoopCtl &c) {
  const int done = *c.done;
  const double e = __longlong_as_double((long long)c.err[iter_index(c)]);
  return !done && e >= c.tol;
}

}  // namespace prh
