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
