// gw_adjoint.hip -- adjoint kernels that hold no line in registers (FP64, one
// scenario): the conductivity gradient of an operator along x or y, the
// transposed z-step with its gradient, the transposed advection sub-step and
// the lambda = W mu conversion.  mu = lambda / W, W = hx hy hz rho_c (the
// transposed diffusion operators are W op W^-1, see AdjGeo).
#include "gw_device.hpp"

namespace gw {

constexpr int kStreamWaves = 4;   // waves per streaming block; each walks a quarter of the line
constexpr int kZcMaxStream = 16;  // column classes staged by k_adj_z (the adjoint requires at most this many)

GW_DEV double cell_w(const AdjGeo& ag, const DevView& v, const ScenDev& S, int i, int j, int k) {
  const int cl = v.cls[S.o_cls + (int64_t)i * S.ny + j];
  return ag.hx[i] * ag.hy[j] * ag.hz[k] * ag.rho[(int64_t)cl * S.nz + k];
}

// Sums the waves' per-lane values in order and adds them to the line's slot of
// part for the tile's output levels.
GW_DEV void add_line_part(double acc, double* part, const LineTile& T, int nz, double* sh) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  sh[w * 64 + lane] = acc;
  __syncthreads();
  const int k = T.kw + lane;
  if (w == 0 && k >= T.ko0 && k < T.ko1) {
    double tot = 0.0;
    for (int q = 0; q < kStreamWaves; ++q) tot += sh[q * 64 + lane];
    part[(int64_t)T.line * nz + k] += tot;
  }
}

// grad += scale * G_dir(mu1 + mu2, x) on the 64-lane tiles of the direction:
// per face (l-1, l) along the line, -(A_f / D_f)(dmu)(dx) with A_f / D_f =
// h_other(line) hz(k) / (g[l] - g[l-1]), half to each rock cell.  Wave w
// walks elements [l_a, l_b) of the line, starting from element l_a - 1.
__global__ __launch_bounds__(64 * kStreamWaves) void k_grad_stream(DevView v, AdjGeo ag, int dir, const double* mu1,
                                                                  const double* mu2, const double* x, double scale) {
  __shared__ double sh[64 * kStreamWaves];
  const LineTile T = (dir == 0 ? v.xtiles[0] : v.ytiles[0])[blockIdx.x];
  const ScenDev S = v.sc[T.scen];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int N = dir == 0 ? S.nx : S.ny, nz = S.nz;
  const int k = T.kw + lane;
  const int kc = k < 0 ? 0 : (k >= nz ? nz - 1 : k);
  const bool kin = k >= 0 && k < nz;
  const int per = (N + kStreamWaves - 1) / kStreamWaves;
  const int la = w * per, lb = min(N, la + per);
  auto cell = [&](int l) {
    const int64_t col = dir == 0 ? (int64_t)l * S.ny + T.line : (int64_t)T.line * S.ny + l;
    return S.off + col * nz + kc;
  };
  auto rock_at = [&](int l) {
    const int64_t col = dir == 0 ? (int64_t)l * S.ny + T.line : (int64_t)T.line * S.ny + l;
    return ag.rock[(int64_t)v.cls[S.o_cls + col] * nz + kc];
  };
  const double* gi = dir == 0 ? ag.ix : ag.iy;
  // faces (l-1, l) for l in [max(la, 1), lb): the wave starts from element
  // la - 1 (or 0), so the face at a wave boundary is counted exactly once
  double g = 0.0;
  if (la < lb) {
    int l = la > 0 ? la - 1 : 0;
    const int64_t c0 = cell(l);
    double mp = mu1[c0] + (mu2 ? mu2[c0] : 0.0), xp = x[c0], rp = rock_at(l);
    for (l = l + 1; l < lb; ++l) {
      const int64_t c1 = cell(l);
      const double m = mu1[c1] + (mu2 ? mu2[c1] : 0.0), xv = x[c1], r = rock_at(l);
      g += gi[l - 1] * (m - mp) * (xv - xp) * (rp + r);
      mp = m; xp = xv; rp = r;
    }
  }
  const double hline = dir == 0 ? ag.hy[T.line] : ag.hx[T.line];
  const double acc = kin ? -0.5 * scale * hline * ag.hz[kc] * g : 0.0;
  add_line_part(acc, dir == 0 ? ag.gx_part : ag.gy_part, T, nz, sh);
}

// Transposed z-step R(dt A_z) in mu on the 64-lane halo tiles of the y-lines
// (3-level halo, outputs 3..60), one line element (column (i, j)) at a time,
// with the z-face gradient of the three Horner stages:
//   forward  w1 = u + a1 A u,  w2 = u + a2 A w1         (recomputed from u)
//   adjoint  m2 = a3 A m,  m1 = a2 A m2,  out = m + m2 + m1 + a1 A m1
//   grad     sum_s a_s G_z(m_s, w_{s-1}) over (m, w2), (m2, w1), (m1, u)
__global__ __launch_bounds__(64 * kStreamWaves) void k_adj_z(DevView v, AdjGeo ag, double* out, const double* mu,
                                                            const double* u, double a1, double a2, double a3) {
  __shared__ double shZ[kZcMaxStream * 64 * 2];
  __shared__ double shR[kZcMaxStream * 64];
  __shared__ double sh[64 * kStreamWaves];
  const LineTile T = v.htiles[0][blockIdx.x];
  const ScenDev S = v.sc[T.scen];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int N = S.ny, nz = S.nz, i = T.line;
  const int k = T.kw + lane;
  const int kc = k < 0 ? 0 : (k >= nz ? nz - 1 : k);
  const int ncls = S.ncls < kZcMaxStream ? S.ncls : kZcMaxStream;
  for (int q = w; q < ncls; q += kStreamWaves) {
    const double* zc = v.zcoef + 2 * (S.o_zcoef + (int64_t)q * nz + kc);
    shZ[(q * 64 + lane) * 2] = zc[0];
    shZ[(q * 64 + lane) * 2 + 1] = zc[1];
    shR[q * 64 + lane] = ag.rock[(int64_t)q * nz + kc];
  }
  __syncthreads();
  const double izk = (k >= 0 && k < nz - 1) ? ag.iz[k] : 0.0;
  const double hxi = ag.hx[i];
  const bool outl = k >= T.ko0 && k < T.ko1;
  const int per = (N + kStreamWaves - 1) / kStreamWaves;
  const int la = w * per, lb = min(N, la + per);
  double acc = 0.0;
  for (int j = la; j < lb; ++j) {
    const int cl = __builtin_amdgcn_readfirstlane(v.cls[S.o_cls + (int64_t)i * S.ny + j]);
    const double ea = shZ[(cl * 64 + lane) * 2], eb = shZ[(cl * 64 + lane) * 2 + 1];
    auto Aop = [&](double q) { return ea * (wave_shl1(q) - q) - eb * (q - wave_shr1(q)); };
    auto face = [&](double mm, double xx) { return (wave_shl1(mm) - mm) * (wave_shl1(xx) - xx); };
    const int64_t cidx = S.off + ((int64_t)i * S.ny + j) * nz + kc;
    const double uu = u[cidx], m = mu[cidx];
    const double w1 = uu + a1 * Aop(uu);
    const double w2 = uu + a2 * Aop(w1);
    const double m2 = a3 * Aop(m);
    const double m1 = a2 * Aop(m2);
    const double o = m + m2 + m1 + a1 * Aop(m1);
    const double fs = a3 * face(m, w2) + a2 * face(m2, w1) + a1 * face(m1, uu);
    const double vup = -hxi * ag.hy[j] * izk * fs;
    const double vdn = wave_shr1(vup);
    acc += 0.5 * shR[cl * 64 + lane] * (vup + vdn);
    if (outl) out[cidx] = o;
  }
  add_line_part(acc, ag.gy_part, T, nz, sh);
}

// ---------------------------------------------------------------------------
// Transposed advection sub-step.
// ---------------------------------------------------------------------------
template <int NB, class F>
GW_DEV double ordered_sum_adj(int n, F ld) {
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

// Pass 1: per reduction chunk, the column sums of lambda = W mu at the levels
// the transposed turnaround (inner chunks) or the inlet gradient (annulus
// chunks) needs.
__global__ __launch_bounds__(256) void k_adj_partial(DevView v, AdjGeo ag, AdjPipe ap, const double* mu) {
  const RedChunk R = ap.chunk[blockIdx.x];
  const BhDev B = v.bh[R.bh];
  const AdjBh A = ap.bh[R.bh];
  const ScenDev S = v.sc[B.scen];
  const int k0 = R.part == 0 ? A.t0 : A.i0;
  const int nl = R.part == 0 ? A.nt : A.ni;
  for (int t = threadIdx.x; t < nl; t += blockDim.x) {
    const int lev = k0 + t;
    ap.partial[R.poff + t] = ordered_sum_adj<16>(R.ncol, [&](int c) {
      const PipeCol pc = v.pcol[R.col0 + c];
      const double m = mu[S.off + ((int64_t)pc.i * S.ny + pc.j) * S.nz + lev];
      return cell_w(ag, v, S, pc.i, pc.j, lev) * m;
    });
  }
}

// Pass 2: per borehole.  Lam_k = sum over the inner chunks (turnaround levels);
// s_j = (1/n_an) sum over the turnaround entries of level j of w * Lam_k (the
// adjoint of the annulus level means); the inlet gradient adds the annulus
// inlet levels' sums.
__global__ __launch_bounds__(256) void k_adj_final(DevView v, AdjPipe ap) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* lam = reinterpret_cast<double*>(smem_raw);  // [nt]
  __shared__ double red[256];
  const int b = blockIdx.x;
  const BhDev B = v.bh[b];
  const AdjBh A = ap.bh[b];
  const int t = threadIdx.x;
  for (int q = t; q < A.nt; q += 256) {
    const int lev = A.t0 + q;
    const bool turn = v.dep[B.dep_off + 2 * lev].type == DEP_TURN;
    lam[q] = turn ? ordered_sum_adj<16>(A.nred_in, [&](int c) {
      return ap.partial[ap.chunk[A.red0 + c].poff + q];
    }) : 0.0;
  }
  __syncthreads();
  for (int q = t; q < B.nlev; q += 256) {
    double s = 0.0;
    for (int e = ap.toff[A.turn_off + q]; e < ap.toff[A.turn_off + q + 1]; ++e)
      s += ap.tent[e].w * lam[ap.tent[e].k - A.t0];
    ap.sbuf[A.soff + q] = s / (double)B.n_an;
  }
  double g = 0.0;
  for (int q = t; q < A.ni; q += 256) {
    const int lev = A.i0 + q;
    if (v.dep[B.dep_off + 2 * lev + 1].type == DEP_INLET)
      g += ordered_sum_adj<16>(A.nred_an, [&](int c) {
        return ap.partial[ap.chunk[A.red0 + A.nred_in + c].poff + q];
      });
  }
  red[t] = g;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) red[t] += red[t + h];
    __syncthreads();
  }
  if (t == 0) ap.gtin[b] += red[0];
}

// Pass 3: per advection group (up to 8 columns of one borehole and kind), for
// each column: lambda_out = W mu staged in LDS, then for every level k' below
// countz lambda_in = sum over the gather list of k' (KEEP and OWN departures)
// + the annulus turnaround adjoint + the objective seed (inner outlet levels),
// written back as mu = lambda_in / W.
template <int R>
__global__ __launch_bounds__(256) void k_adj_apply(DevView v, AdjGeo ag, AdjPipe ap, double* mu, const double* seed) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* lout = reinterpret_cast<double*>(smem_raw);  // [countz_max]
  const RedChunk Gp = v.agroup[blockIdx.x];
  const BhDev B = v.bh[Gp.bh];
  const AdjBh A = ap.bh[Gp.bh];
  const ScenDev S = v.sc[B.scen];
  const int nlev = S.countz;
  const int t = threadIdx.x;
  const int* off = ap.off + (Gp.part == 0 ? A.csr_in : A.csr_an);
  const double sd = (Gp.part == 0 && seed) ? seed[Gp.bh] : 0.0;
  for (int c = 0; c < Gp.ncol; ++c) {
    const PipeCol pc = v.pcol[Gp.col0 + c];
    double* p = mu + S.off + ((int64_t)pc.i * S.ny + pc.j) * S.nz;
    double wk[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int k = t + 256 * r;
      wk[r] = k < nlev ? cell_w(ag, v, S, pc.i, pc.j, k) : 1.0;
      if (k < nlev) lout[k] = wk[r] * p[k];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int k = t + 256 * r;
      if (k < nlev) {
        double lin = 0.0;
        for (int e = off[k]; e < off[k + 1]; ++e) lin += ap.ent[e].w * lout[ap.ent[e].k];
        if (Gp.part == 1 && k >= B.klo && k < B.klo + B.nlev) lin += ap.sbuf[A.soff + k - B.klo];
        if (Gp.part == 0 && k < B.n_outlev) lin += sd;
        p[k] = lin / wk[r];
      }
    }
    // outlet levels below the borehole bottom (v_inner dt/2 > h): cells the
    // advection leaves alone but the outlet mean still reads
    if (sd != 0.0)
      for (int k = nlev + t; k < B.n_outlev; k += 256) p[k] += sd / cell_w(ag, v, S, pc.i, pc.j, k);
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_mul_w(DevView v, AdjGeo ag, double* lam, const double* mu) {
  const ScenDev S = v.sc[0];
  const int64_t n = (int64_t)S.nx * S.ny * S.nz;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(g % S.nz);
    const int64_t col = g / S.nz;
    const int i = (int)(col / S.ny), j = (int)(col % S.ny);
    lam[S.off + g] = cell_w(ag, v, S, i, j, k) * mu[S.off + g];
  }
}

hipError_t launch_grad_stream(const DevView& v, const AdjGeo& ag, int dir, const double* mu1, const double* mu2,
                              const double* x, double scale, hipStream_t st) {
  const int64_t n = dir == 0 ? v.xcls[0][kLineClasses] : v.ycls[0][kLineClasses];
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_grad_stream, dim3((unsigned)n), dim3(64 * kStreamWaves), 0, st, v, ag, dir, mu1, mu2, x,
                     scale);
  return hipGetLastError();
}

hipError_t launch_adj_z(const DevView& v, const AdjGeo& ag, double* out, const double* mu, const double* u,
                        double dt, double c3, hipStream_t st) {
  const int64_t n = v.hcls[0][kLineClasses];
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_adj_z, dim3((unsigned)n), dim3(64 * kStreamWaves), 0, st, v, ag, out, mu, u, 2.0 * c3 * dt,
                     0.5 * dt, dt);
  return hipGetLastError();
}

hipError_t launch_adj_pipe(const DevView& v, const AdjGeo& ag, const AdjPipe& ap, double* mu, const double* seed,
                           hipStream_t st) {
  if (v.nbh == 0) return hipSuccess;
  if (ap.nchunk > 0) {
    hipLaunchKernelGGL(k_adj_partial, dim3((unsigned)ap.nchunk), dim3(256), 0, st, v, ag, ap, mu);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_adj_final, dim3((unsigned)v.nbh), dim3(256), (size_t)v.countz_max * sizeof(double), st, v,
                     ap);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (v.nagroup == 0) return hipSuccess;
  const size_t sh = (size_t)v.countz_max * sizeof(double);
  const int R = (v.countz_max + 255) / 256;
  const dim3 grid((unsigned)v.nagroup), block(256);
  if (R <= 1) hipLaunchKernelGGL((k_adj_apply<1>), grid, block, sh, st, v, ag, ap, mu, seed);
  else if (R <= 2) hipLaunchKernelGGL((k_adj_apply<2>), grid, block, sh, st, v, ag, ap, mu, seed);
  else if (R <= 4) hipLaunchKernelGGL((k_adj_apply<4>), grid, block, sh, st, v, ag, ap, mu, seed);
  else if (R <= 8) hipLaunchKernelGGL((k_adj_apply<8>), grid, block, sh, st, v, ag, ap, mu, seed);
  else if (R <= 16) hipLaunchKernelGGL((k_adj_apply<16>), grid, block, sh, st, v, ag, ap, mu, seed);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_mul_w(const DevView& v, const AdjGeo& ag, double* lam, const double* mu, hipStream_t st) {
  hipLaunchKernelGGL(k_mul_w, dim3(2048), dim3(256), 0, st, v, ag, lam, mu);
  return hipGetLastError();
}

}  // namespace gw
