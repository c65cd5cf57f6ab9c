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
// Line kernel.  A tile is up to L adjacent k-lines (k0 .. k0+nk-1) of one
// x-line (fixed j) or y-line (fixed i) that share one coefficient set (the
// host splits 64-wide k ranges at material zone boundaries).  The block has
// NW waves; a wave's 64 lanes cover L lines x (64/L) segments, so the block
// covers G = NW*64/L segments of S consecutive line elements (G*S >= N).
// Each lane keeps its segment in registers.  The Thomas recurrences run
// segment-locally with zero carry and are then corrected by the carried value
// times the running product of the multipliers; the carries are a short scan
// over the segments in LDS.  Loads and stores of one element index cover L
// consecutive k values: coalesced.  The coefficient set is staged in LDS,
// padded to G*S with identity rows so the element loops carry no branches.
// ---------------------------------------------------------------------------
template <int L_, int NW_, int S_>
struct LineCfg {
  static constexpr int L = L_, NW = NW_, S = S_;
  static constexpr int SPW = 64 / L_;     // segments per wave
  static constexpr int G = NW_ * SPW;     // segments per tile
  static constexpr int NMAX = G * S_;     // longest line
  static constexpr int THREADS = NW_ * 64;
};

template <class Cfg, class TS, int MODE, class RT>
GW_DEV void line_tile_body(const double* cf, RT (&d)[Cfg::S], double* sh_carry, int N, int seg,
                           int ll, double c) {
  constexpr int S = Cfg::S, G = Cfg::G, L = Cfg::L, NM = Cfg::NMAX;
  const double* fw = cf;
  const double* bd = cf + NM;
  const double* bu = cf + 2 * NM;
  const double* ea = cf + 3 * NM;
  const double* eb = cf + 4 * NM;
  const int l0 = seg * S;
  double* fy = sh_carry;              // [G][L] forward local end value
  double* fP = sh_carry + G * L;      // [G][L] forward multiplier product
  double* bx = sh_carry + 2 * G * L;  // [G][L] backward local start value
  double* bQ = sh_carry + 3 * G * L;  // [G][L] backward multiplier product
  double* xf = sh_carry + 4 * G * L;  // [G][L] first element of the segment
  double* xl = sh_carry + 5 * G * L;  // [G][L] last element of the segment

  if (MODE != LINE_EXPLICIT) {
    // forward elimination y_l = b_l - fw_l y_{l-1}, segment-local
    double y = 0.0, P = 1.0;
#pragma unroll
    for (int r = 0; r < S; ++r) {
      const double w = fw[l0 + r];
      y = (double)d[r] - w * y;
      P = -w * P;
      d[r] = y;
    }
    fy[seg * L + ll] = y;
    fP[seg * L + ll] = P;
    __syncthreads();
    double C = 0.0;
    for (int g = 0; g < seg; ++g) C = fy[g * L + ll] + fP[g * L + ll] * C;
    P = 1.0;
#pragma unroll
    for (int r = 0; r < S; ++r) {
      P = -fw[l0 + r] * P;
      d[r] = (double)d[r] + P * C;
    }
    // back substitution x_l = bd_l y_l - bu_l x_{l+1}, segment-local
    double x = 0.0, Q = 1.0;
#pragma unroll
    for (int r = S - 1; r >= 0; --r) {
      const double u = bu[l0 + r];
      x = bd[l0 + r] * (double)d[r] - u * x;
      Q = -u * Q;
      d[r] = x;
    }
    bx[seg * L + ll] = x;
    bQ[seg * L + ll] = Q;
    __syncthreads();
    double X = 0.0;
    for (int g = G - 1; g > seg; --g) X = bx[g * L + ll] + bQ[g * L + ll] * X;
    Q = 1.0;
#pragma unroll
    for (int r = S - 1; r >= 0; --r) {
      Q = -bu[l0 + r] * Q;
      d[r] = (double)d[r] + Q * X;
    }
  }

  if (MODE != LINE_IMPLICIT) {
    // explicit half-step of the same direction: u + c (ea (u+ - u) - eb (u - u-))
    xf[seg * L + ll] = (double)d[0];
    xl[seg * L + ll] = (double)d[S - 1];
    __syncthreads();
    const TS cc = (TS)c;
    TS prev = (seg > 0) ? (TS)xl[(seg - 1) * L + ll] : (TS)d[0];
    const TS edge = (l0 + S < N) ? (TS)xf[(seg + 1) * L + ll] : (TS)d[S - 1];
#pragma unroll
    for (int r = 0; r < S; ++r) {
      const TS x = (TS)d[r];
      const TS xp = (r < S - 1) ? (TS)d[r + 1] : edge;
      const TS o = x + cc * ((TS)ea[l0 + r] * (xp - x) - (TS)eb[l0 + r] * (x - prev));
      prev = x;
      d[r] = (RT)o;
    }
  }
}

template <class Cfg, class TS, int MODE>
__global__ __launch_bounds__(Cfg::THREADS) void k_line(DevView v, int dir, TS* out, const TS* in,
                                                       double c) {
  using RT = typename std::conditional<MODE == LINE_EXPLICIT, TS, double>::type;
  constexpr int S = Cfg::S, L = Cfg::L, G = Cfg::G, NM = Cfg::NMAX;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* sh_carry = reinterpret_cast<double*>(smem_raw);  // 6*G*L doubles
  double* sh_coef = sh_carry + 6 * G * L;                  // 5*NMAX doubles

  const LineTile T = (dir == 0 ? v.xtiles : v.ytiles)[blockIdx.x];
  const ScenDev S_ = v.sc[T.scen];
  const int tid = threadIdx.x;
  const int ll = tid % L;
  const int seg = (L == 64) ? __builtin_amdgcn_readfirstlane(tid / 64) : tid / L;
  const int N = dir == 0 ? S_.nx : S_.ny;
  const int64_t stride = dir == 0 ? (int64_t)S_.nz * S_.ny : (int64_t)S_.nz;
  // idle lanes shadow the tile's last line; the lane offset is a 32-bit
  // unsigned value over a wave-uniform base so loads/stores use the
  // scalar-base addressing form (no 64-bit address per element in VGPRs)
  const unsigned kb = (unsigned)(ll < T.nk ? ll : T.nk - 1) * (unsigned)sizeof(TS);
  const int64_t base = S_.off + (dir == 0 ? (int64_t)S_.nz * T.line : (int64_t)S_.nz * S_.ny * T.line) + T.k0;

  // stage the coefficient set, identity rows beyond N
  const double* g = v.coef + v.set_off[T.sid];
  for (int q = tid; q < NM; q += Cfg::THREADS) {
    const bool ok = q < N;
    sh_coef[q] = ok ? g[q] : 0.0;
    sh_coef[NM + q] = ok ? g[N + q] : 1.0;
    sh_coef[2 * NM + q] = ok ? g[2 * N + q] : 0.0;
    sh_coef[3 * NM + q] = ok ? g[3 * N + q] : 0.0;
    sh_coef[4 * NM + q] = ok ? g[4 * N + q] : 0.0;
  }

  RT d[S];
  const int l0 = seg * S;
  const int64_t ustride = uniform64(stride);
  const int64_t useg = uniform64(base + (int64_t)l0 * stride);
  const TS* pin = in + useg;
  if (l0 + S <= N) {
#pragma unroll
    for (int r = 0; r < S; ++r) {
      const TS* pr = pin + (int64_t)r * ustride;
      d[r] = (RT)*lane_ptr(pr, kb);
    }
  } else {
#pragma unroll
    for (int r = 0; r < S; ++r) {
      const TS* pr = pin + (int64_t)r * ustride;
      d[r] = (l0 + r < N) ? (RT)*lane_ptr(pr, kb) : (RT)0;
    }
  }
  __syncthreads();

  line_tile_body<Cfg, TS, MODE>(sh_coef, d, sh_carry, N, seg, ll, c);

  if (ll < T.nk) {
    TS* pout = out + useg;
    if (l0 + S <= N) {
#pragma unroll
      for (int r = 0; r < S; ++r) {
        TS* pr = pout + (int64_t)r * ustride;
        *lane_ptr(pr, kb) = (TS)d[r];
      }
    } else {
#pragma unroll
      for (int r = 0; r < S; ++r) {
        TS* pr = pout + (int64_t)r * ustride;
        if (l0 + r < N) *lane_ptr(pr, kb) = (TS)d[r];
      }
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

template <class Cfg, class TS>
static hipError_t line_cfg(const DevView& v, int mode, int dir, TS* out, const TS* in, double c,
                           int nmax, hipStream_t st) {
  const int64_t ntiles = dir == 0 ? v.xtiles_total : v.ytiles_total;
  (void)nmax;
  const size_t sh = (size_t)(6 * Cfg::G * Cfg::L + 5 * Cfg::NMAX) * sizeof(double);
  const dim3 grid((unsigned)ntiles), block(Cfg::THREADS);
  if (mode == LINE_EXPLICIT)
    hipLaunchKernelGGL((k_line<Cfg, TS, LINE_EXPLICIT>), grid, block, sh, st, v, dir, out, in, c);
  else if (mode == LINE_IMPLICIT)
    hipLaunchKernelGGL((k_line<Cfg, TS, LINE_IMPLICIT>), grid, block, sh, st, v, dir, out, in, c);
  else
    hipLaunchKernelGGL((k_line<Cfg, TS, LINE_IMPLICIT_EXPLICIT>), grid, block, sh, st, v, dir, out, in, c);
  return hipGetLastError();
}

// Tile shapes: lines per tile 64, waves per tile NW, elements per lane S.
using Cfg48 = LineCfg<64, 1, 48>;
using Cfg96 = LineCfg<64, 2, 48>;
using Cfg192 = LineCfg<64, 4, 48>;
using Cfg416 = LineCfg<64, 8, 52>;
using Cfg640 = LineCfg<64, 8, 80>;

template <class TS>
static hipError_t line_t(const DevView& v, int mode, int dir, TS* out, const TS* in, double c,
                         int nmax, hipStream_t st) {
  if (nmax <= Cfg48::NMAX) return line_cfg<Cfg48>(v, mode, dir, out, in, c, nmax, st);
  if (nmax <= Cfg96::NMAX) return line_cfg<Cfg96>(v, mode, dir, out, in, c, nmax, st);
  if (nmax <= Cfg192::NMAX) return line_cfg<Cfg192>(v, mode, dir, out, in, c, nmax, st);
  if (nmax <= Cfg416::NMAX) return line_cfg<Cfg416>(v, mode, dir, out, in, c, nmax, st);
  if (nmax <= Cfg640::NMAX) return line_cfg<Cfg640>(v, mode, dir, out, in, c, nmax, st);
  return hipErrorInvalidValue;
}

hipError_t launch_line(const DevView& v, int prec, int mode, int dir, void* out, const void* in,
                       double c, hipStream_t st) {
  const int nmax = dir == 0 ? v.nx_max : v.ny_max;
  if (prec == 0) return line_t<double>(v, mode, dir, (double*)out, (const double*)in, c, nmax, st);
  return line_t<float>(v, mode, dir, (float*)out, (const float*)in, c, nmax, st);
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
