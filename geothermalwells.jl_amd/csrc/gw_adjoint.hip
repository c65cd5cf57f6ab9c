// gw_adjoint.hip -- adjoint kernels that hold no line in registers (FP64, one
// scenario): the transposed explicit y and z-step with their conductivity
// gradients (streaming along the y-lines), the transposed advection sub-step
// and the lambda = W mu conversion.  mu = lambda / W, W = hx hy hz rho_c (the
// transposed diffusion operators are W op W^-1, see AdjGeo).
#include "gw_device.hpp"

namespace gw {

constexpr int kStreamWaves = 4;   // waves per streaming block
constexpr int kWalk = 16;         // elements per wave of k_adj_yz (fully unrolled: no loop back-edge)
constexpr int kZcMaxStream = 16;  // column classes staged by k_adj_yz (the adjoint requires at most this many)

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

// Transposed (explicit y then z-step) in mu on the 64-lane halo tiles of the
// y-lines (3-level halo, outputs 3..60), with both gradients.  The line is
// split over ag.yparts blocks of kStreamWaves waves; each wave walks kWalk
// consecutive elements (columns (i, j)), loads two to three elements ahead,
// and keeps only a few values per lane.  Per element:
//   EY:       m = mu + c (ea (mu_{j+1} - mu) - eb (mu - mu_{j-1}))   (explicit y in mu)
//             grad_y += face (j-1, j): (mu_j - mu_{j-1})(Z_j - Z_{j-1}), Z = R(dt A_z) u
//   forward   w1 = u + a1 A u,  w2 = u + a2 A w1,  Z = u + a3 A w2   (from u, in the wave)
//   adjoint   m2 = a3 A m,  m1 = a2 A m2,  out = m + m2 + m1 + a1 A m1
//   grad_z    sum_s a_s G_z(m_s, w_{s-1}) over (m, w2), (m2, w1), (m1, u)
// Without EY, m = mu (the transposed z-step alone).  Gradients go to slot
// (part, line) of gy_part.
template <bool EY>
// Eight waves per SIMD: the walk is latency-bound, so occupancy buys more than
// deeper per-wave prefetch (64 registers fit without spills).
__global__ __launch_bounds__(64 * kStreamWaves, 8) void k_adj_yz(DevView v, AdjGeo ag, double* out, const double* mu,
                                                             const double* u, double c, double a1, double a2,
                                                             double a3) {
  // z coefficients [ncls][64][2] and rock terms [ncls][64] of the scenario's
  // column classes: launched with LDS for its class count only
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int ncls_l = ag.ncls < kZcMaxStream ? ag.ncls : kZcMaxStream;
  double* shZ = reinterpret_cast<double*>(smem_raw);
  double* shR = shZ + ncls_l * 64 * 2;
  __shared__ double sh[64 * kStreamWaves];
  // per-element values that are the same for every lane (hy_j, iy_{j-1} and
  // the column class), for the block's elements and their neighbours
  constexpr int kJs = kStreamWaves * kWalk + 4;
  __shared__ double shHy[kJs], shIy[kJs];
  __shared__ int shCl[kJs];
  // explicit-y coefficients ea, eb of the tile's coefficient sets (at most
  // four: the sets of its lanes) for the same elements
  __shared__ double shEa[4 * kJs], shEb[4 * kJs];
  const int parts = ag.yparts;
  const int64_t tb = blockIdx.x / parts;
  const int p = (int)(blockIdx.x - tb * parts);
  const LineTile T = v.htiles[0][tb];
  const ScenDev S = v.sc[T.scen];
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int N = S.ny, nz = S.nz, i = T.line;
  const int k = T.kw + lane;
  const int kc = k < 0 ? 0 : (k >= nz ? nz - 1 : k);
  const int ncls = ncls_l;
  // the lane's explicit-y set: its slot among the tile's sets (from the tile)
  const int eslot = EY ? (int)((T.slot_b0 >> lane) & 1u) | ((int)((T.slot_b1 >> lane) & 1u) << 1) : 0;
  // block's elements j (clamped) from jlo: la - 1 .. lb + 2 of its waves
  const int jlo = max(0, p * kStreamWaves * kWalk - 1);
  for (int q = threadIdx.x; q < kJs; q += 64 * kStreamWaves) {
    const int jj = min(jlo + q, N - 1);
    shHy[q] = ag.hy[jj];
    shIy[q] = ag.iy[jj > 0 ? jj - 1 : 0];
    shCl[q] = v.cls[S.o_cls + (int64_t)i * S.ny + jj];
  }
  if (EY) {
    // from the sets' padded copies (the line kernels' layout: ea, eb are
    // arrays 3 and 4 of the set's NM-element arrays), addressed from the tile
    const int NM = line_class_max(N);
    for (int q = threadIdx.x; q < T.nset * kJs; q += 64 * kStreamWaves) {
      const int s = q / kJs, jj = min(jlo + q - s * kJs, N - 1);
      const double* g = v.pcoef[0] + (s == 0 ? T.poff[0] : (s == 1 ? T.poff[1] : (s == 2 ? T.poff[2] : T.poff[3])));
      shEa[q] = g[3 * NM + jj];
      shEb[q] = g[4 * NM + jj];
    }
  }
  // z coefficients and rock terms of the column classes: all of a wave's loads
  // before its LDS writes (one round trip)
  {
    constexpr int QW = kZcMaxStream / kStreamWaves;
    double z0[QW], z1[QW], zr[QW];
#pragma unroll
    for (int s = 0; s < QW; ++s) {
      const int q = w + s * kStreamWaves;
      const bool ok = q < ncls;
      const int qc = ok ? q : 0;
      const double* zc = v.zcoef + 2 * (S.o_zcoef + (int64_t)qc * nz + kc);
      z0[s] = ok ? zc[0] : 0.0;
      z1[s] = ok ? zc[1] : 0.0;
      zr[s] = ok ? ag.rock[(int64_t)qc * nz + kc] : 0.0;
    }
#pragma unroll
    for (int s = 0; s < QW; ++s) {
      const int q = w + s * kStreamWaves;
      if (q < ncls) {
        shZ[(q * 64 + lane) * 2] = z0[s];
        shZ[(q * 64 + lane) * 2 + 1] = z1[s];
        shR[q * 64 + lane] = zr[s];
      }
    }
  }
  __syncthreads();
  const double izk = (k >= 0 && k < nz - 1) ? ag.iz[k] : 0.0;
  const double hxi = ag.hx[i];
  const bool kin = k >= 0 && k < nz;
  const bool outl = k >= T.ko0 && k < T.ko1;
  // each wave walks kWalk elements (fewer at the end of the line)
  const int la = min(N, (p * kStreamWaves + w) * kWalk), lb = min(N, la + kWalk);
  // element j of this lane: wave-uniform base of column (i, j) plus the
  // lane's byte offset (scalar base + 32-bit vector offset addressing)
  const int64_t cbase = uniform64(S.off + (int64_t)i * S.ny * nz);
  const unsigned kcb = (unsigned)kc * (unsigned)sizeof(double);
  auto at = [&](const double* f, int j) { return *lane_ptr(f + cbase + (int64_t)j * nz, kcb); };
  auto cls_at = [&](int j) { return __builtin_amdgcn_readfirstlane(shCl[j - jlo]); };
  // a value with its lane neighbours (levels k+1, k-1): each shifted once and
  // shared by the z operator and the face products (the kernel is bound by
  // instruction issue, and a 64-bit lane shift is two DPP moves)
  struct Nb {
    double c, p, m;
  };
  auto nb = [&](double q) { return Nb{q, wave_shl1(q), wave_shr1(q)}; };
  auto face = [&](const Nb& a, const Nb& b) { return (a.p - a.c) * (b.p - b.c); };
  double gy = 0.0, gz = 0.0;
  if (la < lb) {
    // element la - 1 (left neighbour of the first element and of the first face)
    const int j0 = la > 0 ? la - 1 : la;
    double m_prev = at(mu, j0);
    double z_prev = 0.0, r_prev = 0.0;
    if (EY && la > 0) {
      const int cl = cls_at(j0);
      const double ez = shZ[(cl * 64 + lane) * 2], fz = shZ[(cl * 64 + lane) * 2 + 1];
      auto Az = [&](const Nb& q) { return ez * (q.p - q.c) - fz * (q.c - q.m); };
      const double up = at(u, j0);
      const Nb nu = nb(up);
      const Nb n1 = nb(up + a1 * Az(nu));
      const Nb n2 = nb(up + a2 * Az(n1));
      z_prev = up + a3 * Az(n2);
      r_prev = shR[cl * 64 + lane];
    }
    // operands in flight in 4-slot register rings (slot of element j: (j - la)
    // % 4, the loop unrolled 4 times so every slot index is a constant and no
    // value moves between registers while its load is in flight): mu three
    // elements ahead (the explicit step needs mu_{j+1}), u two ahead (the
    // per-element coefficients come from LDS)
    auto jc = [&](int j) { return j < N ? j : N - 1; };
    double rm[4], ru[4];
    auto load_mu = [&](int slot, int j) { rm[slot & 3] = at(mu, jc(j)); };
    auto load_rest = [&](int slot, int j) { ru[slot & 3] = at(u, jc(j)); };
#pragma unroll
    for (int q = 0; q < 3; ++q) load_mu(q, la + q);
#pragma unroll
    for (int q = 0; q < 2; ++q) load_rest(q, la + q);
    auto step = [&](int s, int j, bool act) __attribute__((always_inline)) {
      load_mu(s + 3, j + 3);
      load_rest(s + 2, j + 2);
      const double m_cur = rm[s & 3], m_nx = rm[(s + 1) & 3], u_cur = ru[s & 3];
      const int jj = jc(j);
      const int cl = cls_at(jj);
      const double ez = shZ[(cl * 64 + lane) * 2], fz = shZ[(cl * 64 + lane) * 2 + 1];
      const double rk = shR[cl * 64 + lane];
      auto Az = [&](const Nb& q) { return ez * (q.p - q.c) - fz * (q.c - q.m); };
      const Nb nu = nb(u_cur);
      const Nb n1 = nb(u_cur + a1 * Az(nu));
      const Nb n2 = nb(u_cur + a2 * Az(n1));
      double m = m_cur;
      if (EY) {
        const double z = u_cur + a3 * Az(n2);
        const double fy = shIy[jj - jlo] * (m_cur - m_prev) * (z - z_prev) * (r_prev + rk);  // used for j > 0 only
        gy += (act && j > 0) ? fy : 0.0;
        const double ea = shEa[eslot * kJs + jj - jlo], eb = shEb[eslot * kJs + jj - jlo];
        m = m_cur + c * (ea * (m_nx - m_cur) - eb * (m_cur - m_prev));
        z_prev = z;
        r_prev = rk;
      }
      const Nb nm = nb(m);
      const Nb nm2 = nb(a3 * Az(nm));
      const Nb nm1 = nb(a2 * Az(nm2));
      const double o = m + nm2.c + nm1.c + a1 * Az(nm1);
      const double fs = a3 * face(nm, n2) + a2 * face(nm2, n1) + a1 * face(nm1, nu);
      const double vup = -hxi * shHy[jj - jlo] * izk * fs;
      const double fz2 = 0.5 * rk * (vup + wave_shr1(vup));
      gz += act ? fz2 : 0.0;
      // the sums are completed here: otherwise the compiler defers every
      // step's terms past the conditional stores and holds them all live
      asm volatile("" : "+v"(gy), "+v"(gz));
      if (outl && act) *lane_ptr(out + cbase + (int64_t)j * nz, kcb) = o;
      m_prev = m_cur;
      // keeps the scheduler from interleaving steps (which would hold several
      // elements' values at once and spill)
      __builtin_amdgcn_sched_barrier(0);
    };
    // straight-line code over the wave's kWalk elements: a loop back-edge
    // would have to move the rings' registers and so wait for every load in
    // flight; steps past the wave's range read clamped elements and neither
    // store nor add to the gradient
#pragma unroll
    for (int q = 0; q < kWalk; ++q) step(q & 3, la + q, la + q < lb);
  }
  const double acc = gz + (EY && kin ? -0.5 * c * hxi * ag.hz[kc] * gy : 0.0);
  add_line_part(acc, ag.gy_part + (int64_t)p * S.nx * nz, T, nz, sh);
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
// the transposed turnaround (inner chunks: kept per level for pass 2) or the
// inlet gradient (annulus chunks: summed over the inlet levels, added to the
// chunk's own accumulator, so the order is fixed) needs.
__global__ __launch_bounds__(256) void k_adj_partial(DevView v, AdjGeo ag, AdjPipe ap, const double* mu) {
  const RedChunk R = ap.chunk[blockIdx.x];
  const BhDev B = v.bh[R.bh];
  const AdjBh A = ap.bh[R.bh];
  const ScenDev S = v.sc[B.scen];
  const int k0 = R.part == 0 ? A.t0 : A.i0;
  const int nl = R.part == 0 ? A.nt : A.ni;
  __shared__ AdjCol scol[kRedChunk];  // the chunk's column entries, loaded once
  __shared__ double red[256];
  if ((int)threadIdx.x < R.ncol) scol[threadIdx.x] = ap.col[R.col0 + threadIdx.x];
  __syncthreads();
  double g = 0.0;
  for (int t = threadIdx.x; t < nl; t += blockDim.x) {
    const int lev = k0 + t;
    const double x = ordered_sum_adj<kRedChunk>(R.ncol, [&](int c) {
      const AdjCol ac = scol[c];
      return ac.hxy * ag.hzrho[(int64_t)ac.cls * S.nz + lev] * mu[ac.off + lev];
    });
    if (R.part == 0) ap.partial[R.poff + t] = x;
    else if (v.dep[B.dep_off + 2 * lev + 1].type == DEP_INLET) g += x;
  }
  if (R.part == 1) {
    g = block_sum<256>(g, red);
    if (threadIdx.x == 0) ap.gacc[blockIdx.x] += g;
  }
}

// Pass 3: per advection group (up to 8 columns of one borehole and kind), for
// each column: lambda_out = W mu staged in LDS, then for every level k' below
// countz lambda_in = sum over the gather list of k' (KEEP and OWN departures)
// + the annulus turnaround adjoint + the objective seed (inner outlet levels),
// written back as mu = lambda_in / W.
template <int R>
__global__ __launch_bounds__(256, R <= 3 ? 6 : 1) void k_adj_apply(DevView v, AdjGeo ag, AdjPipe ap, double* mu, const double* seed) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* lout = reinterpret_cast<double*>(smem_raw);  // [countz_max]
  const RedChunk Gp = v.agroup[blockIdx.x];
  const BhDev B = v.bh[Gp.bh];
  const AdjBh A = ap.bh[Gp.bh];
  const ScenDev S = v.sc[B.scen];
  const int nlev = S.countz;
  const int t = threadIdx.x;
  // a column's mu and cell weights W = hxy hz rho at the thread's levels;
  // the first column's go out before everything else (the gather lists and
  // turnaround partials below overlap them), each next column's while the
  // current one gathers
  double pv[R], wv[R];
  auto load_col = [&](const AdjCol& a) {
    const double* hr = ag.hzrho + (int64_t)a.cls * S.nz;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int k = t + 256 * r;
      pv[r] = k < nlev ? mu[a.off + k] : 0.0;
      wv[r] = k < nlev ? hr[k] : 1.0;
    }
  };
  AdjCol ac = ap.col[Gp.col0];
  load_col(ac);
  const int* off = ap.off + (Gp.part == 0 ? A.csr_in : A.csr_an);
  const double sd = (Gp.part == 0 && seed) ? seed[Gp.bh] : 0.0;
  // the gather lists do not depend on the column: the first two entries of
  // each level, staged in LDS once per block (a level of a semi-Lagrangian
  // shift gathers from at most a few departures); the rest of a longer list
  // is read from the table
  int* gk = reinterpret_cast<int*>(lout + 2 * (size_t)v.countz_max);   // [2][countz_max]
  double* gw = lout + 3 * (size_t)v.countz_max;                         // [2][countz_max]
  int e_beg[R], e_end[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int k = t + 256 * r;
    e_beg[r] = k < nlev ? off[k] : 0;
    e_end[r] = k < nlev ? off[k + 1] : 0;
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int k = t + 256 * r;
    if (k < nlev) {
      const AdjEnt a = e_end[r] > e_beg[r] ? ap.ent[e_beg[r]] : AdjEnt{0, 0, 0.0};
      const AdjEnt b = e_end[r] > e_beg[r] + 1 ? ap.ent[e_beg[r] + 1] : AdjEnt{0, 0, 0.0};
      gk[k] = a.k;
      gk[v.countz_max + k] = b.k;
      gw[k] = a.w;
      gw[v.countz_max + k] = b.w;
    }
  }
  // the group's column entries through LDS (a vector load of the table in
  // the loop would wait for every access in flight)
  __shared__ AdjCol scol[kAdvGroupMax];
  if (t < Gp.ncol) scol[t] = ap.col[Gp.col0 + t];
  // annulus groups: the adjoint of the turnaround means at the thread's
  // levels, s_k = (1/n_an) sum over the turnaround entries of level k of
  // w Lam_k', with Lam_k' the inner chunks' partials of turnaround level k'
  // added in chunk order (every block of the borehole gets the same values)
  double sk[R];
#pragma unroll
  for (int r = 0; r < R; ++r) sk[r] = 0.0;
  const bool turn = Gp.part == 1 && A.nt > 0 && A.nred_in > 0;
  if (turn) {
    double* lam = lout + v.countz_max;  // [nt]
    const double* pin = ap.partial + ap.chunk[A.red0].poff;
    for (int q = t; q < A.nt; q += 256)
      lam[q] = ordered_sum_adj<8>(A.nred_in, [&](int c) { return pin[(int64_t)c * A.nt + q]; });
  }
  __syncthreads();
  if (turn) {
    const double* lam = lout + v.countz_max;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int q = t + 256 * r - B.klo;
      if (t + 256 * r < nlev && q >= 0 && q < B.nlev) {
        double sm = 0.0;
        for (int e = ap.toff[A.turn_off + q]; e < ap.toff[A.turn_off + q + 1]; ++e)
          sm += ap.tent[e].w * lam[ap.tent[e].k - A.t0];
        sk[r] = sm / (double)B.n_an;
      }
    }
  }
  for (int c = 0; c < Gp.ncol; ++c) {
    double* p = mu + ac.off;
    double wk[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int k = t + 256 * r;
      wk[r] = k < nlev ? ac.hxy * wv[r] : 1.0;
      if (k < nlev) lout[k] = wk[r] * pv[r];
    }
    const AdjCol an = c + 1 < Gp.ncol ? scol[c + 1] : ac;
    if (c + 1 < Gp.ncol) load_col(an);  // distinct columns: no overlap with this column's stores
    lds_barrier();  // distinct columns: only lout needs the barrier
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int k = t + 256 * r;
      if (k < nlev) {
        double lin = gw[k] * lout[gk[k]] + gw[v.countz_max + k] * lout[gk[v.countz_max + k]];
        for (int e = e_beg[r] + 2; e < e_end[r]; ++e) lin += ap.ent[e].w * lout[ap.ent[e].k];
        lin += sk[r];
        if (Gp.part == 0 && k < B.n_outlev) lin += sd;
        p[k] = lin / wk[r];
      }
    }
    // outlet levels below the borehole bottom (v_inner dt/2 > h): cells the
    // advection leaves alone but the outlet mean still reads
    if (sd != 0.0) {
      const double* hr = ag.hzrho + (int64_t)ac.cls * S.nz;
      for (int k = nlev + t; k < B.n_outlev; k += 256) p[k] += sd / (ac.hxy * hr[k]);
    }
    lds_barrier();
    ac = an;
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

hipError_t launch_adj_yz(const DevView& v, const AdjGeo& ag, bool ey, double* out, const double* mu,
                         const double* u, double c, double dt, double c3, hipStream_t st) {
  const int64_t n = v.hcls[0][kLineClasses];
  if (n == 0) return hipSuccess;
  const double a1 = 2.0 * c3 * dt, a2 = 0.5 * dt, a3 = dt;
  const dim3 grid((unsigned)(n * ag.yparts)), block(64 * kStreamWaves);
  const size_t sh = (size_t)std::min(ag.ncls, kZcMaxStream) * 64 * 3 * sizeof(double);
  if (ey) hipLaunchKernelGGL(k_adj_yz<true>, grid, block, sh, st, v, ag, out, mu, u, c, a1, a2, a3);
  else hipLaunchKernelGGL(k_adj_yz<false>, grid, block, sh, st, v, ag, out, mu, u, c, a1, a2, a3);
  return hipGetLastError();
}

int adj_yparts(int64_t htiles, int ny) {
  (void)htiles;
  return std::max(1, (ny + kStreamWaves * kWalk - 1) / (kStreamWaves * kWalk));
}

hipError_t launch_adj_pipe(const DevView& v, const AdjGeo& ag, const AdjPipe& ap, double* mu, const double* seed,
                           hipStream_t st) {
  if (v.nbh == 0) return hipSuccess;
  if (ap.nchunk > 0) {
    hipLaunchKernelGGL(k_adj_partial, dim3((unsigned)ap.nchunk), dim3(256), 0, st, v, ag, ap, mu);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (v.nagroup == 0) return hipSuccess;
  // lout, lam, the gather lists' first two source levels (int) and weights:
  // 40 bytes per level (24 KB at 602 levels, six blocks per CU)
  const size_t sh = 2 * (size_t)v.countz_max * sizeof(double) + 2 * (size_t)v.countz_max * sizeof(int) +
                    2 * (size_t)v.countz_max * sizeof(double);
  const int R = (v.countz_max + 255) / 256;
  const dim3 grid((unsigned)v.nagroup), block(256);
  if (R <= 1) hipLaunchKernelGGL((k_adj_apply<1>), grid, block, sh, st, v, ag, ap, mu, seed);
  else if (R <= 2) hipLaunchKernelGGL((k_adj_apply<2>), grid, block, sh, st, v, ag, ap, mu, seed);
  else if (R <= 3) hipLaunchKernelGGL((k_adj_apply<3>), grid, block, sh, st, v, ag, ap, mu, seed);
  else if (R <= 4) hipLaunchKernelGGL((k_adj_apply<4>), grid, block, sh, st, v, ag, ap, mu, seed);
  else if (R <= 6) hipLaunchKernelGGL((k_adj_apply<6>), grid, block, sh, st, v, ag, ap, mu, seed);
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
