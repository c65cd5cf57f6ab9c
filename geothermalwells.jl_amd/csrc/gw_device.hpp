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

// kLineClassMax[line_class(n)] on the device: the longest line of n's
// line-length class (the padded length of its coefficient sets).
GW_DEV int line_class_max(int n) {
  int m = kLineClassMax[kLineClasses - 1];
#pragma unroll
  for (int q = kLineClasses - 1; q >= 0; --q)
    if (n <= kLineClassMax[q]) m = kLineClassMax[q];
  return m;
}

// LDS DMA (global_load_lds_dwordx4): each active lane copies the 16 bytes at
// g straight into LDS, lane i to lds_wave + 16 i (lds_wave wave-uniform), no
// registers involved.  It completes under vmcnt; a __syncthreads() after it
// waits for it (the barrier's fence waits for vmcnt(0) while LDS DMA is
// pending).
GW_DEV void lds_dma16(const void* g, void* lds_wave) {
  __builtin_amdgcn_global_load_lds(g, (__attribute__((address_space(3))) void*)lds_wave, 16, 0, 0);
}

// The same with 4 bytes per lane (global_load_lds_dword): lane i to
// lds_wave + 4 i.
GW_DEV void lds_dma4(const void* g, void* lds_wave) {
  __builtin_amdgcn_global_load_lds(g, (__attribute__((address_space(3))) void*)lds_wave, 4, 0, 0);
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

// ---------------------------------------------------------------------------
// Stochastic rounding FP64 -> FP32 (GW_F32_STATE_SR).  A double's low 29
// mantissa bits are what FP32 drops; adding a uniform 29-bit integer to the
// bit pattern and clearing those bits rounds the magnitude up with
// probability (dropped bits) / 2^29, so E[result] = x (the carry into bit 29
// moves to the next float, across a binade too).  r is a 24-bit uniform
// number, placed at the top of the 29 bits (the probability is exact to
// 2^-24).  The result is an FP32 value, so the final conversion is exact;
// values FP32 holds exactly are returned unchanged.
// ---------------------------------------------------------------------------
GW_DEV float sr_f32(double x, uint32_t r24) {
  const uint64_t b = (__builtin_bit_cast(uint64_t, x) + ((uint64_t)r24 << 5)) & ~(uint64_t)0x1FFFFFFFu;
  return (float)__builtin_bit_cast(double, b);
}

// 32-bit integer hash (a bijection; "lowbias32", C. Wellons' hash prospector).
__host__ __device__ constexpr uint32_t lowbias32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// The random numbers of a line's elements: the states of the 24-bit LCG
// s <- (kLcgA s + kLcgC) mod 2^24 (full period: a = 1 mod 4, c odd; one
// 24-bit multiply-add per number), element l taking the l-th state.
// lcg_jump(n) is the map of n steps, s -> a s + c (mod 2^24); kLcgAInv steps
// back.
constexpr uint32_t kLcgA = 16598013u, kLcgC = 12820163u, kLcgMask = 0xFFFFFFu;
struct LcgJump {
  uint32_t a, c;
};
__host__ __device__ constexpr LcgJump lcg_jump(uint32_t n) {
  uint32_t A = 1u, C = 0u, ma = kLcgA, mc = kLcgC;
  while (n) {
    if (n & 1u) { A *= ma; C = C * ma + mc; }
    mc = mc * (ma + 1u);
    ma *= ma;
    n >>= 1;
  }
  return LcgJump{A & kLcgMask, C & kLcgMask};
}
__host__ __device__ constexpr uint32_t lcg_inverse_a() {
  uint32_t x = kLcgA;  // correct to 3 bits (a odd), each Newton step doubles them
  for (int i = 0; i < 5; ++i) x *= 2u - kLcgA * x;
  return x & kLcgMask;
}
constexpr uint32_t kLcgAInv = lcg_inverse_a();
static_assert(((kLcgA * kLcgAInv) & kLcgMask) == 1u, "LCG multiplier inverse");
static_assert(((lcg_jump(3).a * 5u + lcg_jump(3).c) & kLcgMask) ==
                  ((kLcgA * (kLcgA * (kLcgA * 5u + kLcgC) + kLcgC) + kLcgC) & kLcgMask), "LCG jump");
GW_DEV uint32_t lcg_next(uint32_t s) { return (__umul24(s, kLcgA) + kLcgC) & kLcgMask; }
GW_DEV uint32_t lcg_prev(uint32_t s) { return __umul24((s - kLcgC) & kLcgMask, kLcgAInv) & kLcgMask; }
GW_DEV uint32_t lcg_apply(LcgJump j, uint32_t s) { return (__umul24(s, j.a) + j.c) & kLcgMask; }

// Random-number stream of one rounding point (p) of step n for the element
// with flat index idx (a line's first element: the line's stream start).
GW_DEV uint32_t sr_stream(uint32_t idx, uint32_t n, uint32_t p) {
  return lowbias32(idx ^ lowbias32(n * 8u + p)) & kLcgMask;
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
