// gw_device.hpp -- device helpers shared by the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "gw_kernels.hpp"

namespace gw {

#define GW_DEV __device__ __forceinline__

// p + byte_off for a wave-uniform p and a 32-bit per-lane byte offset (lets
// the compiler use the scalar-base + 32-bit vector-offset addressing form).
template <class T>
GW_DEV T* lane_ptr(T* p, unsigned byte_off) {
  return reinterpret_cast<T*>(reinterpret_cast<char*>(p) + byte_off);
}
template <class T>
GW_DEV const T* lane_ptr(const T* p, unsigned byte_off) {
  return reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + byte_off);
}

// Workgroup barrier that waits for this wave's LDS operations only: its
// global loads and stores stay in flight (__syncthreads' fence waits for
// them all).  For LDS buffers reused between passes whose global accesses
// never alias.
GW_DEV void lds_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0); vmcnt and expcnt at their maxima (no wait)
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// A wave-uniform element offset pinned in SGPRs: base + offset + lane byte
// offset then uses the scalar-base + 32-bit vector-offset addressing form.
GW_DEV int64_t soff(int64_t x) {
  asm volatile("" : "+s"(x));
  return x;
}

// Marks a wave-uniform 64-bit value as such (it then lives in SGPRs).
GW_DEV int64_t uniform64(int64_t x) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)(uint64_t)x);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)x >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Value of lane-1 (wave_shr:1) / lane+1 (wave_shl:1) across the whole wave;
// lane 0 (resp. 63) reads 0.
GW_DEV uint32_t dpp_shr1(uint32_t x) { return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x138, 0xf, 0xf, true); }
GW_DEV uint32_t dpp_shl1(uint32_t x) { return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x130, 0xf, 0xf, true); }
GW_DEV double wave_shr1(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint64_t r = (uint64_t)dpp_shr1((uint32_t)b) | ((uint64_t)dpp_shr1((uint32_t)(b >> 32)) << 32);
  return __builtin_bit_cast(double, r);
}
GW_DEV double wave_shl1(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint64_t r = (uint64_t)dpp_shl1((uint32_t)b) | ((uint64_t)dpp_shl1((uint32_t)(b >> 32)) << 32);
  return __builtin_bit_cast(double, r);
}
GW_DEV float wave_shr1(float v) { return __builtin_bit_cast(float, dpp_shr1(__builtin_bit_cast(uint32_t, v))); }
GW_DEV float wave_shl1(float v) { return __builtin_bit_cast(float, dpp_shl1(__builtin_bit_cast(uint32_t, v))); }

// Sum of x over the block with a fixed-order tree (deterministic for a given
// block size NT); red holds NT doubles.
template <int NT>
GW_DEV double block_sum(double x, double* red) {
  red[threadIdx.x] = x;
  __syncthreads();
#pragma unroll
  for (int h = NT / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// Level-sum layout of the block that combines the chunk partials: kFinLP
// levels per pass, kFinS chunk slices.
constexpr int kFinLP = 128, kFinS = 8, kFinThreads = kFinLP * kFinS;

// Per-level sum of the chunk partials ld(0) .. ld(n-1) by a kFinThreads
// block: thread (l, s) = (t % kFinLP, t / kFinLP) adds chunks s, s + kFinS, ...
// in order, 4 loads in flight; the slices are then added in slice order.  The
// result is valid in the threads with s == 0; every thread must call it.
template <class F>
GW_DEV double level_sum(int n, F ld, double* red) {
  const int t = threadIdx.x, l = t % kFinLP, s = t / kFinLP;
  double acc = 0.0;
  int q = s;
  for (; q + 3 * kFinS < n; q += 4 * kFinS) {
    const double x0 = ld(q), x1 = ld(q + kFinS), x2 = ld(q + 2 * kFinS), x3 = ld(q + 3 * kFinS);
    acc += x0;
    acc += x1;
    acc += x2;
    acc += x3;
  }
  for (; q < n; q += kFinS) acc += ld(q);
  red[t] = acc;
  __syncthreads();
  double r = 0.0;
  if (s == 0)
#pragma unroll
    for (int ss = 0; ss < kFinS; ++ss) r += red[ss * kFinLP + l];
  __syncthreads();
  return r;
}

// Scenario s with prefix[s] <= idx < prefix[s+1].
GW_DEV int find_scen(const int64_t* prefix, int n, int64_t idx) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= idx) lo = mid; else hi = mid - 1;
  }
  return lo;
}

}  // namespace gw
