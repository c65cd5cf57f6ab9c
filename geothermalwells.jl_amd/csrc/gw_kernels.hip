// gw_kernels.hip -- gfx950 kernels of the per-timestep hot path.
//
//   k_zrhs          dphi = A_z phi                      (rhs_diffusion_z!, src/solver.jl:23-58)
//   k_zstep         phi = R(dt A_z) phi, R the minimum-degree ROCK2 polynomial
//   k_line          x/y lines: explicit (I + cA), implicit (I - cA)^-1, or the
//                   implicit solve followed by the explicit half-step of the same
//                   direction (thomas_I_minus_A! then diffusion_1D!,
//                   src/solver.jl:450,455), fused
//   k_pipe_reduce   outlet mean and turnaround level means (src/solver.jl:253-257,
//                   267-282, 321-335) with a fixed-order (deterministic) reduction
//   k_pipe_advect   semi-Lagrangian update of the pipe columns (src/solver.jl:285-396)
//
// Every kernel is bandwidth-bound (no MFMA).  Lanes always run along k, the
// contiguous axis, so each wave touches 64 consecutive elements (512 B in FP64).
#include <hip/hip_runtime.h>

#include <type_traits>

#include "gw_kernels.hpp"

namespace gw {

#define GW_DEV __device__ __forceinline__
#ifndef GW_SCHED_FENCE
#define GW_SCHED_FENCE() ((void)0)
#endif

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

// ---------------------------------------------------------------------------
// dphi = A_z phi  (one thread per cell)
// ---------------------------------------------------------------------------
template <class TS>
__global__ __launch_bounds__(256) void k_zrhs(DevView v, TS* __restrict__ dphi,
                                              const TS* __restrict__ phi) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= v.cells_total) return;
  const int s = find_scen(v.cell_prefix, v.nscen, g);
  const ScenDev S = v.sc[s];
  const int64_t q = g - v.cell_prefix[s];
  const int nz = S.nz;
  const int64_t col = q / nz;
  const int k = (int)(q - col * nz);
  const int c = v.cls[S.o_cls + col];
  const double* zc = v.zcoef + 2 * (S.o_zcoef + (int64_t)c * nz + k);
  const TS* p = phi + S.off + col * nz;
  const TS u = p[k];
  const TS up = (k + 1 < nz) ? p[k + 1] : u;
  const TS um = (k > 0) ? p[k - 1] : u;
  dphi[S.off + q] = (TS)zc[0] * (up - u) - (TS)zc[1] * (u - um);
}

// ---------------------------------------------------------------------------
// phi = R(dt A_z) phi, Horner form  w = u + a1 A u;  w = u + a2 A w;  w = u + a3 A w
// with a1 = 2 c3 dt, a2 = dt/2, a3 = dt.  One 256-thread block per (i, j)
// column; the column lives in registers (R values per thread), the stage
// vector in LDS.
// ---------------------------------------------------------------------------
template <class TS, int R>
__global__ __launch_bounds__(256) void k_zstep(DevView v, TS* __restrict__ phi, TS a1, TS a2, TS a3) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  TS* w = reinterpret_cast<TS*>(smem_raw);
  const int64_t g = blockIdx.x;
  const int s = find_scen(v.col_prefix, v.nscen, g);
  const ScenDev S = v.sc[s];
  const int64_t col = g - v.col_prefix[s];
  const int nz = S.nz;
  const int c = v.cls[S.o_cls + col];
  const double* zc = v.zcoef + 2 * (S.o_zcoef + (int64_t)c * nz);
  TS* p = phi + S.off + col * nz;
  const int t = threadIdx.x;
  TS u[R], ea[R], eb[R], nw[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int k = t + 256 * r;
    if (k < nz) {
      u[r] = p[k];
      const double2 e = *reinterpret_cast<const double2*>(zc + 2 * k);
      ea[r] = (TS)e.x;
      eb[r] = (TS)e.y;
      w[k] = u[r];
    } else {
      u[r] = 0; ea[r] = 0; eb[r] = 0;
    }
  }
  __syncthreads();
  const TS alpha[3] = {a1, a2, a3};
#pragma unroll
  for (int st = 0; st < 3; ++st) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int k = t + 256 * r;
      if (k < nz) {
        const TS w0 = w[k];
        const TS wm = (k > 0) ? w[k - 1] : w0;
        const TS wp = (k + 1 < nz) ? w[k + 1] : w0;
        nw[r] = u[r] + alpha[st] * (ea[r] * (wp - w0) - eb[r] * (w0 - wm));
      }
    }
    if (st < 2) {
      __syncthreads();
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int k = t + 256 * r;
        if (k < nz) w[k] = nw[r];
      }
      __syncthreads();
    }
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int k = t + 256 * r;
    if (k < nz) p[k] = nw[r];
  }
}

// ---------------------------------------------------------------------------
// Line kernel.  A tile is a window of 64 adjacent k-lines (k = kw .. kw+63,
// one per lane) of one x-line (fixed j) or y-line (fixed i); its outputs are
// the lanes with ko0 <= k < ko1.  The block has NW waves; wave w holds line
// elements [w*S, w*S + S) of its lane's line in registers (NW*S >= N).
// Loads and stores of one element index are 64 consecutive k values:
// coalesced.  The Thomas recurrences run segment-locally from zero carry and
// are then corrected by the carried value times the running product of the
// multipliers; the carries are an NW-step scan in LDS.  The lanes of a window
// use at most three coefficient sets (one per material zone met in z), staged
// in LDS and padded with identity rows so the element loops have no branches.
//
// Modes (the reference's operators on the same direction, src/solver.jl):
//   EXPLICIT            u <- (I + cA) u                               (:34-58)
//   IMPLICIT            u <- (I - cA)^-1 u                            (:111-232)
//   IMPLICIT_EXPLICIT   u <- (I + cA)(I - cA)^-1 u   x-lines, :450 + :455
//   ZY                  u <- (I + cA_y) R(dt A_z) u   next step's z-step and
//                       explicit y (:23-31 under ROCK2, then :444)
//   YZY                 u <- (I + cA_y) R(dt A_z) (I - cA_y)^-1 u      (:461, then
//                       the next step's z-step and explicit y)
// ZY / YZY run on y-lines with a 3-lane halo in k: the z-stencil is applied
// three times across lanes (shuffles), so only lanes 3..60 of a window are
// outputs; the halo lanes re-solve their own lines.
// ---------------------------------------------------------------------------
template <int NW_, int S_, int LW_>
struct LineCfg {
  static constexpr int NW = NW_, S = S_;
  static constexpr int LW = LW_;             // k-window width: lanes per line element
  static constexpr int H = 64 / LW_;         // line segments per wave (one per half-wave when LW = 32)
  static constexpr int G = NW_ * H;          // line segments per tile
  static constexpr int NMAX = G * S_;        // longest line
  static constexpr int THREADS = NW_ * 64;
  static_assert(LW_ == 64 || LW_ == 32, "k-window of a full or half wave");
};

constexpr int kSlots = 3;      // coefficient sets per window
constexpr int kZcMax = 16;     // column classes staged for the fused z-step

template <int MODE> struct ModeTraits {
  static constexpr bool halo = MODE == LINE_ZY || MODE == LINE_YZY;
  static constexpr bool impl = MODE == LINE_IMPLICIT || MODE == LINE_IMPLICIT_EXPLICIT || MODE == LINE_YZY;
  static constexpr bool expl = MODE != LINE_IMPLICIT;
  // coefficient arrays staged per set slot: bd, fw, bu (implicit) then ea, eb
  // (explicit), each NMAX long
  static constexpr int NA = (impl ? 3 : 0) + (expl ? 2 : 0);
  static constexpr int A_BD = 0, A_FW = 1, A_BU = 2, A_EA = impl ? 3 : 0, A_EB = A_EA + 1;
};

template <class Cfg, int MODE>
constexpr size_t line_lds_doubles() {
  using MT = ModeTraits<MODE>;
  return 2 * Cfg::G * Cfg::LW + kSlots * MT::NA * Cfg::NMAX +
         (MT::halo ? kZcMax * Cfg::LW * 2 + (Cfg::NMAX + 1) / 2 : 0);
}

GW_DEV int64_t xcd_remap(int64_t b, int64_t n) {
  // Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
  // contiguous range of tiles so neighbouring k-windows of a line share an L2.
  const int64_t per = n / 8, rem = n % 8;
  const int64_t x = b % 8, i = b / 8;
  return x * per + (x < rem ? x : rem) + i;
}

template <class Cfg, class TS, int MODE>
__global__ __launch_bounds__(Cfg::THREADS, 2) void k_line(DevView v, int dir, TS* out, const TS* in,
                                                       double c, double a1, double a2, double a3) {
  using MT = ModeTraits<MODE>;
  // The line is held in registers in the storage type; every recurrence runs
  // in FP64 and rounds to the storage type only when it writes an element.
  using RT = TS;
  constexpr int S = Cfg::S, NM = Cfg::NMAX, LW = Cfg::LW, H = Cfg::H, G = Cfg::G;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* shA = reinterpret_cast<double*>(smem_raw);  // [G][LW] carries / edges
  double* shB = shA + G * LW;                          // [G][LW]
  double* shC = shB + G * LW;                          // [kSlots][MT::NA][NM] coefficient arrays
  double* shZ = shC + kSlots * MT::NA * NM;            // [kZcMax][LW][2]: z coefficients
  int* shK = reinterpret_cast<int*>(shZ + kZcMax * LW * 2);  // [NM] column class per line element

  constexpr int FAM = LW == 32 ? 1 : 0;
  const LineTile* tiles = MT::halo ? v.htiles[FAM] : (dir == 0 ? v.xtiles[FAM] : v.ytiles[FAM]);
  const int64_t ntiles =
      MT::halo ? v.htiles_total[FAM] : (dir == 0 ? v.xtiles_total[FAM] : v.ytiles_total[FAM]);
  const LineTile T = tiles[xcd_remap(blockIdx.x, ntiles)];
  const ScenDev S_ = v.sc[T.scen];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int kk = lane % LW;    // k offset in the window
  const int hh = lane / LW;    // half-wave (0 when LW = 64)
  const int seg = wv * H + hh;
  const int N = dir == 0 ? S_.nx : S_.ny;
  const int nz = S_.nz;
  const int k = T.kw + kk;
  const int kc = k < 0 ? 0 : (k >= nz ? nz - 1 : k);
  const int kmin = T.kw < 0 ? 0 : T.kw;
  // lane byte offset from the wave's first segment: the k offset plus, for
  // the upper half-wave, S line elements (below 2^32: nx, ny <= 640,
  // nz <= 4096, S <= 52)
  const unsigned kb = (unsigned)(kc - kmin) * (unsigned)sizeof(TS) +
                      (unsigned)(hh * S) * (unsigned)(dir == 0 ? nz * S_.ny : nz) * (unsigned)sizeof(TS);
  const int64_t stride = dir == 0 ? (int64_t)nz * S_.ny : (int64_t)nz;
  const int64_t base = S_.off + (dir == 0 ? (int64_t)nz * T.line : (int64_t)nz * S_.ny * T.line) + kmin;

  // coefficient sets of the window's lanes, identity rows beyond N
  const int32_t* sids = v.setid + (dir == 0 ? S_.o_xset : S_.o_yset) + (int64_t)T.line * S_.nzone;
  const int sid = sids[v.zone[S_.o_zone + kc]];
  const int slot = sid == T.set[1] ? 1 : (sid == T.set[2] ? 2 : 0);
  for (int q = 0; q < T.nset; ++q) {
    const int sq = q == 0 ? T.set[0] : (q == 1 ? T.set[1] : T.set[2]);
    const double* g = v.coef + v.set_off[sq];
    double* dst = shC + q * MT::NA * NM;
    for (int l = tid; l < NM; l += Cfg::THREADS) {
      const bool ok = l < N;
      if (MT::impl) {
        dst[MT::A_BD * NM + l] = ok ? g[SET_BD * N + l] : 1.0;
        dst[MT::A_FW * NM + l] = ok ? g[SET_FW * N + l] : 0.0;
        dst[MT::A_BU * NM + l] = ok ? g[SET_BU * N + l] : 0.0;
      }
      if (MT::expl) {
        dst[MT::A_EA * NM + l] = ok ? g[SET_EA * N + l] : 0.0;
        dst[MT::A_EB * NM + l] = ok ? g[SET_EB * N + l] : 0.0;
      }
    }
  }
  if (MT::halo) {
    const double* zc = v.zcoef + 2 * S_.o_zcoef;
    const int ncls = S_.ncls < kZcMax ? S_.ncls : kZcMax;
    for (int q = wv; q < ncls; q += Cfg::NW) {
      shZ[(q * LW + kk) * 2 + hh] = zc[2 * ((int64_t)q * nz + kc) + hh];
      if (H == 1) shZ[(q * LW + kk) * 2 + 1] = zc[2 * ((int64_t)q * nz + kc) + 1];
    }
    const uint16_t* clsl = v.cls + S_.o_cls + (int64_t)S_.ny * T.line;
    for (int l = tid; l < NM; l += Cfg::THREADS) shK[l] = clsl[l < N ? l : N - 1];
  }

  RT d[S];
  const int l0 = seg * S;
  const int64_t ustride = uniform64(stride);
  const int64_t useg = uniform64(base + (int64_t)(wv * H * S) * stride);
  const TS* pin = in + useg;
  if (l0 + S <= N) {
#pragma unroll
    for (int r = 0; r < S; ++r) d[r] = (RT)*lane_ptr(pin + (int64_t)r * ustride, kb);
  } else {
#pragma unroll
    for (int r = 0; r < S; ++r) d[r] = (l0 + r < N) ? (RT)*lane_ptr(pin + (int64_t)r * ustride, kb) : (RT)0;
  }
  __syncthreads();

  const double* cs = shC + slot * MT::NA * NM;
  const double* bd = cs + MT::A_BD * NM;
  const double* fw = cs + MT::A_FW * NM;
  const double* bu = cs + MT::A_BU * NM;
  const double* ea = cs + MT::A_EA * NM;
  const double* eb = cs + MT::A_EB * NM;

  // Every element loop below runs in chunks of KC elements: the LDS operands
  // of the next chunk are fetched while the current one computes, and a
  // scheduling barrier keeps the compiler from hoisting further (which would
  // spill the register-resident line).
  constexpr int KC = 4;
  static_assert(S % KC == 0, "segment length must be a multiple of the chunk");
  if (MT::impl) {
    // forward: y_l = b_l - fw_l y_{l-1}
    auto wmul = [&](int l) { return fw[l]; };
    double y = 0.0, P = 1.0, wn[KC];
#pragma unroll
    for (int q = 0; q < KC; ++q) wn[q] = wmul(l0 + q);
#pragma unroll
    for (int r0 = 0; r0 < S; r0 += KC) {
      double wc[KC];
#pragma unroll
      for (int q = 0; q < KC; ++q) wc[q] = wn[q];
      if (r0 + KC < S) {
#pragma unroll
        for (int q = 0; q < KC; ++q) wn[q] = wmul(l0 + r0 + KC + q);
      }
#pragma unroll
      for (int q = 0; q < KC; ++q) {
        y = (double)d[r0 + q] - wc[q] * y;
        P = -wc[q] * P;
        d[r0 + q] = y;
      }
      GW_SCHED_FENCE();
    }
    shA[seg * LW + kk] = y;
    shB[seg * LW + kk] = P;
    __syncthreads();
    double C = 0.0;
    for (int g = 0; g < seg; ++g) C = shA[g * LW + kk] + shB[g * LW + kk] * C;
    __syncthreads();
    P = 1.0;
#pragma unroll
    for (int q = 0; q < KC; ++q) wn[q] = wmul(l0 + q);
#pragma unroll
    for (int r0 = 0; r0 < S; r0 += KC) {
      double wc[KC];
#pragma unroll
      for (int q = 0; q < KC; ++q) wc[q] = wn[q];
      if (r0 + KC < S) {
#pragma unroll
        for (int q = 0; q < KC; ++q) wn[q] = wmul(l0 + r0 + KC + q);
      }
#pragma unroll
      for (int q = 0; q < KC; ++q) {
        P = -wc[q] * P;
        d[r0 + q] = (double)d[r0 + q] + P * C;
      }
      GW_SCHED_FENCE();
    }
    // backward: x_l = bd_l y_l - bu_l x_{l+1}
    double x = 0.0, Q = 1.0, bn[KC], un[KC];
#pragma unroll
    for (int q = 0; q < KC; ++q) {
      const int l = l0 + S - 1 - q;
      bn[q] = bd[l];
      un[q] = bu[l];
    }
#pragma unroll
    for (int r0 = 0; r0 < S; r0 += KC) {
      double bc[KC], uc[KC];
#pragma unroll
      for (int q = 0; q < KC; ++q) { bc[q] = bn[q]; uc[q] = un[q]; }
      if (r0 + KC < S) {
#pragma unroll
        for (int q = 0; q < KC; ++q) {
          const int l = l0 + S - 1 - (r0 + KC + q);
          bn[q] = bd[l];
          un[q] = bu[l];
        }
      }
#pragma unroll
      for (int q = 0; q < KC; ++q) {
        const int r = S - 1 - (r0 + q);
        x = bc[q] * (double)d[r] - uc[q] * x;
        Q = -uc[q] * Q;
        d[r] = x;
      }
      GW_SCHED_FENCE();
    }
    shA[seg * LW + kk] = x;
    shB[seg * LW + kk] = Q;
    __syncthreads();
    double X = 0.0;
    for (int g = G - 1; g > seg; --g) X = shA[g * LW + kk] + shB[g * LW + kk] * X;
    __syncthreads();
    Q = 1.0;
#pragma unroll
    for (int q = 0; q < KC; ++q) un[q] = bu[l0 + S - 1 - q];
#pragma unroll
    for (int r0 = 0; r0 < S; r0 += KC) {
      double uc[KC];
#pragma unroll
      for (int q = 0; q < KC; ++q) uc[q] = un[q];
      if (r0 + KC < S) {
#pragma unroll
        for (int q = 0; q < KC; ++q) un[q] = bu[l0 + S - 1 - (r0 + KC + q)];
      }
#pragma unroll
      for (int q = 0; q < KC; ++q) {
        const int r = S - 1 - (r0 + q);
        Q = -uc[q] * Q;
        d[r] = (double)d[r] + Q * X;
      }
      GW_SCHED_FENCE();
    }
  }

  if (MT::halo) {
    // z-step R(dt A_z) across lanes (lane = k), column class per line element;
    // neighbours k-1 / k+1 come from lanes -1 / +1 by DPP wave shifts (the
    // window edge lanes keep their own value: they are halo, never output)
    TS ean[KC], ebn[KC];
#pragma unroll
    for (int q = 0; q < KC; ++q) {
      const int cl = shK[l0 + q];
      ean[q] = (TS)shZ[(cl * LW + kk) * 2];
      ebn[q] = (TS)shZ[(cl * LW + kk) * 2 + 1];
    }
#pragma unroll
    for (int r0 = 0; r0 < S; r0 += KC) {
      TS eac[KC], ebc[KC];
#pragma unroll
      for (int q = 0; q < KC; ++q) { eac[q] = ean[q]; ebc[q] = ebn[q]; }
      if (r0 + KC < S) {
#pragma unroll
        for (int q = 0; q < KC; ++q) {
          const int cl = shK[l0 + r0 + KC + q];
          ean[q] = (TS)shZ[(cl * LW + kk) * 2];
          ebn[q] = (TS)shZ[(cl * LW + kk) * 2 + 1];
        }
      }
      // the chunk's KC elements advance through the three stages together
      // (KC independent chains per stage); the stage loop stays a real loop:
      // unrolled, the compiler interleaves chunks and the line spills
      TS u0[KC], w[KC];
#pragma unroll
      for (int q = 0; q < KC; ++q) { u0[q] = (TS)d[r0 + q]; w[q] = u0[q]; }
#pragma unroll 1
      for (int st = 0; st < 3; ++st) {
        const TS al = st == 0 ? (TS)a1 : (st == 1 ? (TS)a2 : (TS)a3);
#pragma unroll
        for (int q = 0; q < KC; ++q) {
          const TS wm = wave_shr1(w[q]);
          const TS wp = wave_shl1(w[q]);
          w[q] = u0[q] + al * (eac[q] * (wp - w[q]) - ebc[q] * (w[q] - wm));
        }
      }
#pragma unroll
      for (int q = 0; q < KC; ++q) d[r0 + q] = (RT)w[q];
      GW_SCHED_FENCE();
    }
  }

  if (MT::expl) {
    // explicit half-step of the line direction: u + c (ea (u+ - u) - eb (u - u-))
    shA[seg * LW + kk] = (double)d[0];
    shB[seg * LW + kk] = (double)d[S - 1];
    __syncthreads();
    const TS cc = (TS)c;
    TS prev = (seg > 0) ? (TS)shB[(seg - 1) * LW + kk] : (TS)d[0];
    const TS edge = (l0 + S < N) ? (TS)shA[(seg + 1) * LW + kk] : (TS)d[S - 1];
    TS an[KC], bn2[KC];
#pragma unroll
    for (int q = 0; q < KC; ++q) { an[q] = (TS)ea[l0 + q]; bn2[q] = (TS)eb[l0 + q]; }
#pragma unroll
    for (int r0 = 0; r0 < S; r0 += KC) {
      TS ac[KC], bc2[KC];
#pragma unroll
      for (int q = 0; q < KC; ++q) { ac[q] = an[q]; bc2[q] = bn2[q]; }
      if (r0 + KC < S) {
#pragma unroll
        for (int q = 0; q < KC; ++q) { an[q] = (TS)ea[l0 + r0 + KC + q]; bn2[q] = (TS)eb[l0 + r0 + KC + q]; }
      }
#pragma unroll
      for (int q = 0; q < KC; ++q) {
        const int r = r0 + q;
        const TS x = (TS)d[r];
        const TS xp = (r < S - 1) ? (TS)d[r + 1] : edge;
        const TS o = x + cc * (ac[q] * (xp - x) - bc2[q] * (x - prev));
        prev = x;
        d[r] = (RT)o;
      }
      GW_SCHED_FENCE();
    }
  }

  if (k >= T.ko0 && k < T.ko1) {
    TS* pout = out + useg;
    if (l0 + S <= N) {
#pragma unroll
      for (int r = 0; r < S; ++r) *lane_ptr(pout + (int64_t)r * ustride, kb) = (TS)d[r];
    } else {
#pragma unroll
      for (int r = 0; r < S; ++r)
        if (l0 + r < N) *lane_ptr(pout + (int64_t)r * ustride, kb) = (TS)d[r];
    }
  }
}

// ---------------------------------------------------------------------------
// Pipe kernels.
// ---------------------------------------------------------------------------
template <int NT>
GW_DEV double block_sum(double x, double* red) {
  // fixed-order tree: deterministic for a given block size
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

// Pass 1: one block per reduction chunk (up to 64 pipe columns of one
// borehole); thread t owns level t (coalesced within each column) and sums
// the chunk's columns in order.
template <class TS>
__global__ __launch_bounds__(256) void k_pipe_partial(DevView v, const TS* __restrict__ phi) {
  const RedChunk R = v.rchunk[blockIdx.x];
  const BhDev B = v.bh[R.bh];
  const ScenDev S = v.sc[B.scen];
  const int k0 = R.part == 0 ? 0 : B.klo;
  const int nl = R.part == 0 ? B.n_outlev : B.nlev;
  const TS* base = phi + S.off + k0;
  for (int lev = threadIdx.x; lev < nl; lev += blockDim.x) {
    double acc = 0.0;
    for (int c = 0; c < R.ncol; ++c) {
      const PipeCol pc = v.pcol[R.col0 + c];
      acc += (double)base[((int64_t)pc.j + (int64_t)S.ny * pc.i) * S.nz + lev];
    }
    v.partial[R.poff + lev] = acc;
  }
}

// Pass 2: one block per borehole; chunk partials are combined in chunk order
// and the outlet levels with a fixed tree, so the result is deterministic.
__global__ __launch_bounds__(256) void k_pipe_final(DevView v) {
  __shared__ double red[256];
  const int b = blockIdx.x;
  const BhDev B = v.bh[b];
  const int t = threadIdx.x;
  double acc = 0.0;
  for (int lev = t; lev < B.n_outlev; lev += 256) {
    double s = 0.0;
    for (int q = 0; q < B.n_red_in; ++q) s += v.partial[v.rchunk[B.red_in0 + q].poff + lev];
    acc += s;
  }
  const double tot = block_sum<256>(acc, red);
  const double tout = tot / ((double)B.n_in * (double)B.n_outlev);
  for (int lev = t; lev < B.nlev; lev += 256) {
    double s = 0.0;
    for (int q = 0; q < B.n_red_an; ++q) s += v.partial[v.rchunk[B.red_an0 + q].poff + lev];
    v.means[B.mean_off + lev] = s / (double)B.n_an;
  }
  if (t == 0) {
    v.tout[b] = tout;
    double in;
    if (B.inlet_kind == 0) in = B.inlet_value;
    else if (B.inlet_kind == 1) in = tout - B.inlet_value;
    else in = v.inlet_tab[b];  // per-well table, or host-written values
    v.inlet[b] = in;
  }
}

// One block per pipe column; the column's advected levels are staged in LDS
// so the update can run in place.
template <class TS>
__global__ __launch_bounds__(256) void k_pipe_advect(DevView v, TS* __restrict__ phi) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  TS* colv = reinterpret_cast<TS*>(smem_raw);
  const PipeCol pc = v.pcol[blockIdx.x];
  const BhDev B = v.bh[pc.bh];
  const ScenDev S = v.sc[B.scen];
  TS* p = phi + S.off + ((int64_t)pc.j + (int64_t)S.ny * pc.i) * S.nz;
  const int nlev = S.countz;
  for (int k = threadIdx.x; k < nlev; k += blockDim.x) colv[k] = p[k];
  __syncthreads();
  const Dep* dp = v.dep + B.dep_off + (pc.inner ? 0 : 1);
  for (int k = threadIdx.x; k < nlev; k += blockDim.x) {
    const Dep e = dp[2 * k];
    if (e.type == DEP_KEEP) continue;
    TS o;
    if (e.type == DEP_OWN) {
      const TS a = (TS)e.a;
      o = ((TS)1 - a) * colv[e.kl] + a * colv[e.kr];
    } else if (e.type == DEP_TURN) {
      const double a = e.a;
      o = (TS)((1.0 - a) * v.means[B.mean_off + e.kl - B.klo] + a * v.means[B.mean_off + e.kr - B.klo]);
    } else {
      o = (TS)v.inlet[pc.bh];
    }
    p[k] = o;
  }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
hipError_t launch_zrhs(const DevView& v, int prec, void* dphi, const void* phi, hipStream_t st) {
  const int64_t nb = (v.cells_total + 255) / 256;
  if (prec == 0)
    hipLaunchKernelGGL((k_zrhs<double>), dim3((unsigned)nb), dim3(256), 0, st, v, (double*)dphi, (const double*)phi);
  else
    hipLaunchKernelGGL((k_zrhs<float>), dim3((unsigned)nb), dim3(256), 0, st, v, (float*)dphi, (const float*)phi);
  return hipGetLastError();
}

template <class TS>
static hipError_t zstep_t(const DevView& v, TS* phi, double dt, double c3, hipStream_t st) {
  const dim3 grid((unsigned)v.cols_total), block(256);
  const size_t sh = (size_t)v.nz_max * sizeof(TS);
  const TS a1 = (TS)(2.0 * c3 * dt), a2 = (TS)(0.5 * dt), a3 = (TS)dt;
  const int R = (v.nz_max + 255) / 256;
  if (R <= 1) hipLaunchKernelGGL((k_zstep<TS, 1>), grid, block, sh, st, v, phi, a1, a2, a3);
  else if (R <= 2) hipLaunchKernelGGL((k_zstep<TS, 2>), grid, block, sh, st, v, phi, a1, a2, a3);
  else if (R <= 4) hipLaunchKernelGGL((k_zstep<TS, 4>), grid, block, sh, st, v, phi, a1, a2, a3);
  else if (R <= 8) hipLaunchKernelGGL((k_zstep<TS, 8>), grid, block, sh, st, v, phi, a1, a2, a3);
  else if (R <= 16) hipLaunchKernelGGL((k_zstep<TS, 16>), grid, block, sh, st, v, phi, a1, a2, a3);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_zstep(const DevView& v, int prec, void* phi, double dt, double c3, hipStream_t st) {
  if (prec == 0) return zstep_t<double>(v, (double*)phi, dt, c3, st);
  return zstep_t<float>(v, (float*)phi, dt, c3, st);
}

template <class Cfg, class TS, int MODE>
static void line_launch(const DevView& v, int dir, TS* out, const TS* in, double c, double a1, double a2,
                        double a3, hipStream_t st) {
  using MT = ModeTraits<MODE>;
  constexpr int FAM = Cfg::LW == 32 ? 1 : 0;
  const int64_t ntiles =
      MT::halo ? v.htiles_total[FAM] : (dir == 0 ? v.xtiles_total[FAM] : v.ytiles_total[FAM]);
  if (ntiles == 0) return;
  const size_t sh = line_lds_doubles<Cfg, MODE>() * sizeof(double);
  hipLaunchKernelGGL((k_line<Cfg, TS, MODE>), dim3((unsigned)ntiles), dim3(Cfg::THREADS), sh, st, v, dir, out, in,
                     c, a1, a2, a3);
}

template <class Cfg, class TS>
static hipError_t line_cfg(const DevView& v, int mode, int dir, TS* out, const TS* in, double c, double a1,
                           double a2, double a3, hipStream_t st) {
  switch (mode) {
    case LINE_EXPLICIT: line_launch<Cfg, TS, LINE_EXPLICIT>(v, dir, out, in, c, a1, a2, a3, st); break;
    case LINE_IMPLICIT: line_launch<Cfg, TS, LINE_IMPLICIT>(v, dir, out, in, c, a1, a2, a3, st); break;
    case LINE_IMPLICIT_EXPLICIT:
      line_launch<Cfg, TS, LINE_IMPLICIT_EXPLICIT>(v, dir, out, in, c, a1, a2, a3, st);
      break;
    case LINE_ZY: line_launch<Cfg, TS, LINE_ZY>(v, dir, out, in, c, a1, a2, a3, st); break;
    case LINE_YZY: line_launch<Cfg, TS, LINE_YZY>(v, dir, out, in, c, a1, a2, a3, st); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Tile shapes: waves per tile NW, line elements per lane S, k-window LW.
// LW = 32 puts two line segments in a wave (one per half-wave): a tile needs
// half the waves of an LW = 64 tile of the same line, so two tiles share a CU
// and one computes while the other loads or stores.
using W64a = LineCfg<1, 48, 64>;
using W64b = LineCfg<2, 48, 64>;
using W64c = LineCfg<4, 48, 64>;
using W64d = LineCfg<8, 52, 64>;
using W64e = LineCfg<8, 80, 64>;
using W32a = LineCfg<1, 24, 32>;
using W32b = LineCfg<1, 48, 32>;
using W32c = LineCfg<2, 48, 32>;
using W32d = LineCfg<4, 52, 32>;
using W32e = LineCfg<8, 40, 32>;

bool line_window_ok(int nmax, int lw) {
  return nmax >= 1 && ((lw == 64 && nmax <= W64e::NMAX) || (lw == 32 && nmax <= W32e::NMAX));
}

int line_window_default(int nmax) { return nmax <= W64d::NMAX ? 64 : 32; }

template <class TS>
static hipError_t line_t(const DevView& v, int mode, int dir, int lw, TS* out, const TS* in, double c, double a1,
                         double a2, double a3, int nmax, hipStream_t st) {
  if (lw == 64) {
    if (nmax <= W64a::NMAX) return line_cfg<W64a>(v, mode, dir, out, in, c, a1, a2, a3, st);
    if (nmax <= W64b::NMAX) return line_cfg<W64b>(v, mode, dir, out, in, c, a1, a2, a3, st);
    if (nmax <= W64c::NMAX) return line_cfg<W64c>(v, mode, dir, out, in, c, a1, a2, a3, st);
    if (nmax <= W64d::NMAX) return line_cfg<W64d>(v, mode, dir, out, in, c, a1, a2, a3, st);
    if (nmax <= W64e::NMAX) return line_cfg<W64e>(v, mode, dir, out, in, c, a1, a2, a3, st);
  } else if (lw == 32) {
    if (nmax <= W32a::NMAX) return line_cfg<W32a>(v, mode, dir, out, in, c, a1, a2, a3, st);
    if (nmax <= W32b::NMAX) return line_cfg<W32b>(v, mode, dir, out, in, c, a1, a2, a3, st);
    if (nmax <= W32c::NMAX) return line_cfg<W32c>(v, mode, dir, out, in, c, a1, a2, a3, st);
    if (nmax <= W32d::NMAX) return line_cfg<W32d>(v, mode, dir, out, in, c, a1, a2, a3, st);
    if (nmax <= W32e::NMAX) return line_cfg<W32e>(v, mode, dir, out, in, c, a1, a2, a3, st);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_line(const DevView& v, int prec, int mode, int dir, int lw, void* out, const void* in,
                       double c, double dt_z, double c3, hipStream_t st) {
  const int nmax = dir == 0 ? v.nx_max : v.ny_max;
  const double a1 = 2.0 * c3 * dt_z, a2 = 0.5 * dt_z, a3 = dt_z;
  if (prec == 0)
    return line_t<double>(v, mode, dir, lw, (double*)out, (const double*)in, c, a1, a2, a3, nmax, st);
  return line_t<float>(v, mode, dir, lw, (float*)out, (const float*)in, c, a1, a2, a3, nmax, st);
}

hipError_t launch_pipe_reduce(const DevView& v, int prec, const void* phi, hipStream_t st) {
  if (v.nbh == 0) return hipSuccess;
  if (prec == 0)
    hipLaunchKernelGGL((k_pipe_partial<double>), dim3(v.nrchunk), dim3(256), 0, st, v, (const double*)phi);
  else
    hipLaunchKernelGGL((k_pipe_partial<float>), dim3(v.nrchunk), dim3(256), 0, st, v, (const float*)phi);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_pipe_final, dim3(v.nbh), dim3(256), 0, st, v);
  return hipGetLastError();
}

hipError_t launch_pipe_advect(const DevView& v, int prec, void* phi, hipStream_t st) {
  if (v.npcol == 0) return hipSuccess;
  const size_t sh = (size_t)v.countz_max * (prec == 0 ? 8 : 4);
  if (prec == 0)
    hipLaunchKernelGGL((k_pipe_advect<double>), dim3(v.npcol), dim3(256), sh, st, v, (double*)phi);
  else
    hipLaunchKernelGGL((k_pipe_advect<float>), dim3(v.npcol), dim3(256), sh, st, v, (float*)phi);
  return hipGetLastError();
}

}  // namespace gw
