// gw_kernels.hip -- gfx950 kernels of the per-timestep hot path.
//
//   k_zrhs          dphi = A_z phi                      (rhs_diffusion_z!, src/solver.jl:23-58)
//   k_zstep         phi = R(dt A_z) phi, R the minimum-degree ROCK2 polynomial
//   (k_line, the x/y line kernel, is in gw_line.hpp)
//   k_pipe_partial  outlet mean and turnaround level means (src/solver.jl:253-257,
//                   267-282, 321-335) with a fixed-order (deterministic) reduction
//   k_pipe_advect   semi-Lagrangian update of the pipe columns (src/solver.jl:285-396)
//
// Every kernel is bandwidth-bound (no MFMA).  Lanes always run along k, the
// contiguous axis, so each wave touches 64 consecutive elements (512 B in FP64).
#include <type_traits>

#include "gw_device.hpp"

namespace gw {

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
// SR (FP32 state, GW_F32_STATE_SR): the stages run in FP64 and the result is
// rounded stochastically, element k of column g drawing
// lowbias32(flat index ^ key(step, point 6)).
template <class TS, int R, bool SR>
__global__ __launch_bounds__(256) void k_zstep(DevView v, TS* __restrict__ phi, double a1_, double a2_, double a3_) {
  using AT = typename std::conditional<SR, double, TS>::type;  // arithmetic type
  const AT a1 = (AT)a1_, a2 = (AT)a2_, a3 = (AT)a3_;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  AT* w = reinterpret_cast<AT*>(smem_raw);
  const int64_t g = blockIdx.x;
  const int s = find_scen(v.col_prefix, v.nscen, g);
  const ScenDev S = v.sc[s];
  const int64_t col = g - v.col_prefix[s];
  const int nz = S.nz;
  const int c = v.cls[S.o_cls + col];
  const double* zc = v.zcoef + 2 * (S.o_zcoef + (int64_t)c * nz);
  TS* p = phi + S.off + col * nz;
  const int t = threadIdx.x;
  AT u[R], ea[R], eb[R], nw[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int k = t + 256 * r;
    if (k < nz) {
      u[r] = (AT)p[k];
      const double2 e = *reinterpret_cast<const double2*>(zc + 2 * k);
      ea[r] = (AT)e.x;
      eb[r] = (AT)e.y;
      w[k] = u[r];
    } else {
      u[r] = 0; ea[r] = 0; eb[r] = 0;
    }
  }
  __syncthreads();
  const AT alpha[3] = {a1, a2, a3};
#pragma unroll
  for (int st = 0; st < 3; ++st) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int k = t + 256 * r;
      if (k < nz) {
        const AT w0 = w[k];
        const AT wm = (k > 0) ? w[k - 1] : w0;
        const AT wp = (k + 1 < nz) ? w[k + 1] : w0;
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
  uint32_t key = 0u;
  if constexpr (SR) key = lowbias32(*v.step_ctr * 8u + 6u);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int k = t + 256 * r;
    if (k < nz) {
      if constexpr (SR)
        p[k] = sr_f32(nw[r], lowbias32((uint32_t)(S.off + col * nz + k) ^ key) >> 8);
      else
        p[k] = (TS)nw[r];
    }
  }
}

// ---------------------------------------------------------------------------
// Pipe kernels.
// ---------------------------------------------------------------------------
// Sum of ld(0) .. ld(n-1) added in index order, NB loads in flight at a time
// (the padding of the last batch adds +0.0, which leaves the sum unchanged).
template <int NB, class F>
GW_DEV double ordered_sum(int n, F ld) {
  double acc = 0.0;
  for (int c = 0; c < n; c += NB) {
    double x[NB];
#pragma unroll
    for (int q = 0; q < NB; ++q) x[q] = c + q < n ? ld(c + q) : 0.0;
#pragma unroll
    for (int q = 0; q < NB; ++q) acc += x[q];
  }
  return acc;
}

// Outlet mean, turnaround level means and inlet value of borehole b from the
// chunk partials (level_sum: a fixed order, so the result is deterministic).
GW_DEV void pipe_final(const DevView& v, int b, double* red) {
  const BhDev B = v.bh[b];
  const int t = threadIdx.x, l = t % kFinLP, s = t / kFinLP;
  // a borehole's chunk partials are contiguous: chunk q of a part at base + q * levels
  const double* pin = v.partial + (B.n_red_in > 0 ? v.rchunk[B.red_in0].poff : 0);
  const double* pan = v.partial + (B.n_red_an > 0 ? v.rchunk[B.red_an0].poff : 0);
  double acc = 0.0;
  for (int l0 = 0; l0 < B.n_outlev; l0 += kFinLP) {
    const int lev = l0 + l;
    const double x = level_sum(lev < B.n_outlev ? B.n_red_in : 0,
                               [&](int q) { return pin[(int64_t)q * B.n_outlev + lev]; }, red);
    if (s == 0) acc += x;
  }
  const double tot = block_sum<kFinThreads>(acc, red);
  const double tout = tot / ((double)B.n_in * (double)B.n_outlev);
  for (int l0 = 0; l0 < B.nlev; l0 += kFinLP) {
    const int lev = l0 + l;
    const double x = level_sum(lev < B.nlev ? B.n_red_an : 0,
                               [&](int q) { return pan[(int64_t)q * B.nlev + lev]; }, red);
    if (s == 0 && lev < B.nlev) v.means[B.mean_off + lev] = x / (double)B.n_an;
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

// Pass 1: each block sums the levels of its chunks (up to kPartChunks
// consecutive chunks of at most kPartItems levels in all, one thread per
// chunk level; a longer chunk alone, thread t owning levels t, t + 256, ...):
// a level's sum adds the chunk's columns in order, all loads in flight at
// once (coalesced within each column).  Small chunks share a block so a
// batch of many shallow reductions is not a launch of mostly idle blocks.
// (A single pass with a last-block combine would need device-scope fences,
// which on this multi-XCD part write back and invalidate the whole L2 per
// block.)
template <class TS>
__global__ __launch_bounds__(256) void k_pipe_partial(DevView v, const TS* __restrict__ phi) {
  const int c0 = v.pblock[blockIdx.x], nc = v.pblock[blockIdx.x + 1] - c0;
  const int t = threadIdx.x;
  // the chunks' column offsets and headers through LDS (in the level loop
  // each column would cost a table load before its value load)
  __shared__ int64_t coff[kPartChunks * kRedChunk];
  __shared__ int32_t item0[kPartChunks], nl[kPartChunks], ncol[kPartChunks], poff[kPartChunks];
  if (t < nc * kRedChunk) {
    const PartChunk P = v.pchunk[c0 + t / kRedChunk];
    const int c = t % kRedChunk;
    if (c < P.ncol) coff[t] = P.base + v.pcol_off[P.col0 + c];
    if (c == 0) {
      const int q = t / kRedChunk;
      item0[q] = P.item0; nl[q] = P.nl; ncol[q] = P.ncol; poff[q] = P.poff;
    }
  }
  __syncthreads();
  if (nc == 1) {
    for (int lev = t; lev < nl[0]; lev += blockDim.x)
      v.partial[poff[0] + lev] = ordered_sum<kRedChunk>(ncol[0], [&](int c) { return (double)phi[coff[c] + lev]; });
  } else {
    int q = 0;
    for (int i = 1; i < nc; ++i) q = item0[i] <= t ? i : q;
    const int lev = t - item0[q];
    if (lev < nl[q])
      v.partial[poff[q] + lev] =
          ordered_sum<kRedChunk>(ncol[q], [&](int c) { return (double)phi[coff[q * kRedChunk + c] + lev]; });
  }
}

// Pass 2: one block per borehole; jacc (may be NULL) accumulates the outlet
// of every launch per borehole (the block owns its slot).
__global__ __launch_bounds__(kFinThreads) void k_pipe_final(DevView v, double* jacc) {
  __shared__ double red[kFinThreads];
  pipe_final(v, blockIdx.x, red);
  if (jacc && threadIdx.x == 0) jacc[blockIdx.x] += v.tout[blockIdx.x];
}

// One block per advection group (up to kAdvGroupMax pipe columns of one
// borehole and one kind, inner or annulus): the departure entries of the
// thread's levels (t + 256 r) and the columns' offsets are loaded once for the
// group; the columns are staged in LDS two at a time (the next pair's loads in
// flight meanwhile) so the update runs in place.
template <class TS, int R>
__global__ __launch_bounds__(256) void k_pipe_advect(DevView v, TS* __restrict__ phi) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  TS* colv = reinterpret_cast<TS*>(smem_raw);  // [2][256 R]
  const RedChunk Gp = v.agroup[blockIdx.x];
  const BhDev B = v.bh[Gp.bh];
  const ScenDev S = v.sc[B.scen];
  const int nlev = S.countz;
  if (nlev <= 0) return;
  const int t = threadIdx.x;
  const Dep* dp = v.dep + B.dep_off + Gp.part;  // entries (inner, annulus) per level
  const int moff = B.mean_off - B.klo;  // means index of turnaround level k: moff + k
  // Slot r of the thread is level t + 256 r, clamped into the column: a slot
  // past the column's end repeats its last level (same departure, same
  // value, same address), so every load, LDS write and store runs without a
  // branch and the compiler can count the stores in flight.  Per slot: the
  // departure kind, the interpolation of an own departure and the value of a
  // turnaround or inlet departure (the same for every column of the group).
  // Load order: departure entries, then (one barrier) the column offsets,
  // then the first column, then the turnaround means the entries point at,
  // so that no wait holds back a load that does not depend on it.
  int kc[R];
  Dep e[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    kc[r] = t + 256 * r < nlev ? t + 256 * r : nlev - 1;
    e[r] = dp[2 * kc[r]];
  }
  // column offsets through LDS: a vector load of the column table inside the
  // loop would wait for every load and store in flight (the table could alias
  // phi as far as the compiler knows)
  __shared__ int64_t coff[kAdvGroupMax];
  if (t < Gp.ncol) {
    const PipeCol pc = v.pcol[Gp.col0 + t];
    coff[t] = S.off + ((int64_t)pc.j + (int64_t)S.ny * pc.i) * S.nz;
  }
  __syncthreads();
  auto colp = [&](int c) { return phi + uniform64(coff[c]); };
  const int ncol = Gp.ncol;
  TS ca[R], cb[R];
  {
    const TS* p0 = colp(0);
    const TS* p1 = colp(ncol >= 2 ? 1 : 0);
#pragma unroll
    for (int r = 0; r < R; ++r) ca[r] = p0[kc[r]];
    if (ncol >= 2) {
#pragma unroll
      for (int r = 0; r < R; ++r) cb[r] = p1[kc[r]];
    }
  }
  int typ[R], kl[R], kr[R];
  TS w[R], cval[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    typ[r] = e[r].type; kl[r] = 0; kr[r] = 0; w[r] = (TS)0; cval[r] = (TS)0;
    if (e[r].type == DEP_OWN) {
      kl[r] = e[r].kl; kr[r] = e[r].kr; w[r] = (TS)e[r].a;
    } else if (e[r].type == DEP_TURN) {
      cval[r] = (TS)((1.0 - e[r].a) * v.means[moff + e[r].kl] + e[r].a * v.means[moff + e[r].kr]);
    } else if (e[r].type == DEP_INLET) {
      cval[r] = (TS)v.inlet[Gp.bh];
    }
  }
  // Columns go in pairs: both of a pair are staged in the two LDS buffers
  // while the next pair's loads are in flight (a load index past the group
  // repeats its last column; such values are never used).
  auto load_col = [&](TS (&dst)[R], int c) {
    const TS* pn = colp(c < ncol ? c : ncol - 1);
#pragma unroll
    for (int r = 0; r < R; ++r) dst[r] = pn[kc[r]];
  };
  auto stage = [&](TS* buf, const TS (&cur)[R]) {
#pragma unroll
    for (int r = 0; r < R; ++r) buf[t + 256 * r] = cur[r];
  };
  auto update = [&](int c, const TS* buf) {
    TS* p = colp(c);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const TS own = ((TS)1 - w[r]) * buf[kl[r]] + w[r] * buf[kr[r]];
      const TS o = typ[r] == DEP_OWN ? own : (typ[r] == DEP_KEEP ? buf[t + 256 * r] : cval[r]);
      p[kc[r]] = o;  // a kept level is written back unchanged
    }
  };
  TS* buf0 = colv;
  TS* buf1 = colv + 256 * R;
  // the columns of a group are distinct, so only the LDS buffers need the
  // barriers: the next pair's loads and this pair's stores stay in flight
  auto pair = [&](int c) {
    stage(buf0, ca);
    stage(buf1, cb);
    load_col(ca, c + 2);
    load_col(cb, c + 3);
    lds_barrier();
    update(c, buf0);
    update(c + 1, buf1);
    lds_barrier();
  };
  if (ncol >= 2) {
    // the first pair outside the loop: every loop entry then has the same
    // accesses in flight and the loop waits for the staged pair's loads only
    pair(0);
    for (int c = 2; c + 1 < ncol; c += 2) pair(c);
  }
  if (ncol & 1) {
    stage(buf0, ca);  // the last column, loaded as the first of a pair
    lds_barrier();
    update(ncol - 1, buf0);
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

template <class TS, bool SR>
static hipError_t zstep_t(const DevView& v, TS* phi, double dt, double c3, hipStream_t st) {
  const dim3 grid((unsigned)v.cols_total), block(256);
  const size_t sh = (size_t)v.nz_max * (SR ? sizeof(double) : sizeof(TS));
  // Horner coefficients rounded to the arithmetic type
  const double a1 = 2.0 * c3 * dt, a2 = 0.5 * dt, a3 = dt;
  const int R = (v.nz_max + 255) / 256;
  if (R <= 1) hipLaunchKernelGGL((k_zstep<TS, 1, SR>), grid, block, sh, st, v, phi, a1, a2, a3);
  else if (R <= 2) hipLaunchKernelGGL((k_zstep<TS, 2, SR>), grid, block, sh, st, v, phi, a1, a2, a3);
  else if (R <= 4) hipLaunchKernelGGL((k_zstep<TS, 4, SR>), grid, block, sh, st, v, phi, a1, a2, a3);
  else if (R <= 8) hipLaunchKernelGGL((k_zstep<TS, 8, SR>), grid, block, sh, st, v, phi, a1, a2, a3);
  else if (R <= 16) hipLaunchKernelGGL((k_zstep<TS, 16, SR>), grid, block, sh, st, v, phi, a1, a2, a3);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_zstep(const DevView& v, int prec, void* phi, double dt, double c3, hipStream_t st) {
  if (prec == 0) return zstep_t<double, false>(v, (double*)phi, dt, c3, st);
  if (prec == 2) return zstep_t<float, true>(v, (float*)phi, dt, c3, st);
  return zstep_t<float, false>(v, (float*)phi, dt, c3, st);
}

__global__ void k_step_ctr(uint32_t* ctr, int set, uint32_t value) {
  if (threadIdx.x == 0) ctr[0] = set ? value : ctr[0] + 1u;
}

hipError_t launch_step_ctr(uint32_t* ctr, bool set, uint32_t value, hipStream_t st) {
  hipLaunchKernelGGL(k_step_ctr, dim3(1), dim3(64), 0, st, ctr, set ? 1 : 0, value);
  return hipGetLastError();
}

int line_window_default(int nmax) { return nmax <= kLineClassMax[8] ? 64 : 32; }

hipError_t launch_line(const DevView& v, int prec, int mode, int dir, int lw, void* out, const void* in,
                       double c, double dt_z, double c3, hipStream_t st, const LineStreams* ls) {
  const double a1 = 2.0 * c3 * dt_z, a2 = 0.5 * dt_z, a3 = dt_z;
  DevView v2 = v;
  if (lw & kLwTwoTiles) v2.lds_floor = kLdsTwoTiles;
  if (lw & kLwHaloWhole) v2.halo_whole = 1;
  const LineStreams* side = (lw & kLwConcurrent) ? ls : nullptr;
  const int w = lw & ~kLwFlags;
  if (w == 64)
    return prec == 0 ? line_w64_f64(v2, mode, dir, out, in, c, a1, a2, a3, st, side)
         : prec == 2 ? line_w64_f32sr(v2, mode, dir, out, in, c, a1, a2, a3, st, side)
                     : line_w64_f32(v2, mode, dir, out, in, c, a1, a2, a3, st, side);
  if (w == 32)
    return prec == 0 ? line_w32_f64(v2, mode, dir, out, in, c, a1, a2, a3, st, side)
         : prec == 2 ? line_w32_f32sr(v2, mode, dir, out, in, c, a1, a2, a3, st, side)
                     : line_w32_f32(v2, mode, dir, out, in, c, a1, a2, a3, st, side);
  return hipErrorInvalidValue;
}

hipError_t launch_pipe_reduce(const DevView& v, int prec, const void* phi, hipStream_t st, double* jacc) {
  if (v.nbh == 0) return hipSuccess;
  if (v.npblock > 0) {
    if (prec == 0)
      hipLaunchKernelGGL((k_pipe_partial<double>), dim3(v.npblock), dim3(256), 0, st, v, (const double*)phi);
    else
      hipLaunchKernelGGL((k_pipe_partial<float>), dim3(v.npblock), dim3(256), 0, st, v, (const float*)phi);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_pipe_final, dim3(v.nbh), dim3(kFinThreads), 0, st, v, jacc);
  return hipGetLastError();
}

template <class TS>
static hipError_t advect_t(const DevView& v, TS* phi, hipStream_t st) {
  const dim3 grid((unsigned)v.nagroup), block(256);
  const int R = (v.countz_max + 255) / 256;
  // two column buffers of 256 R levels (every thread's slots)
  auto sh = [](int r) { return 2 * (size_t)256 * r * sizeof(TS); };
  if (R <= 1) hipLaunchKernelGGL((k_pipe_advect<TS, 1>), grid, block, sh(1), st, v, phi);
  else if (R <= 2) hipLaunchKernelGGL((k_pipe_advect<TS, 2>), grid, block, sh(2), st, v, phi);
  else if (R <= 3) hipLaunchKernelGGL((k_pipe_advect<TS, 3>), grid, block, sh(3), st, v, phi);
  else if (R <= 4) hipLaunchKernelGGL((k_pipe_advect<TS, 4>), grid, block, sh(4), st, v, phi);
  else if (R <= 6) hipLaunchKernelGGL((k_pipe_advect<TS, 6>), grid, block, sh(6), st, v, phi);
  else if (R <= 8) hipLaunchKernelGGL((k_pipe_advect<TS, 8>), grid, block, sh(8), st, v, phi);
  else if (R <= 16) hipLaunchKernelGGL((k_pipe_advect<TS, 16>), grid, block, sh(16), st, v, phi);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_pipe_advect(const DevView& v, int prec, void* phi, hipStream_t st) {
  if (v.nagroup == 0) return hipSuccess;
  if (prec == 0) return advect_t<double>(v, (double*)phi, st);
  return advect_t<float>(v, (float*)phi, st);
}

}  // namespace gw
