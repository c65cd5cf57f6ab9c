// gw_line_adj.hpp -- adjoint line kernels (FP64), on the tiles of the line
// kernel (gw_line.hpp).  The state adjoint runs in mu = lambda / W, where the
// transposed diffusion operators are the forward ones (see AdjGeo), so the
// solve and the explicit half-step below repeat the forward kernel's
// arithmetic on mu.  Each mode also accumulates the gradient of the
// objective with respect to the rock conductivity: per face of the operator's
// direction the term -(A_f / D_f) (mu2 - mu1)(x2 - x1), x the forward field
// the operator acted on, half to each rock cell of the face, summed per
// (line, level) over the line (one lane per level) and added to the line's
// slot of gx_part / gy_part by one lane (no atomics; fixed order).
#pragma once

#include "gw_line.hpp"

namespace gw {

template <int MODE> struct AdjTraits {
  static constexpr bool halo = false;
  static constexpr bool impl = true;
  static constexpr bool expl = MODE == ADJ_X;
  static constexpr int NA = (impl ? 3 : 0) + (expl ? 2 : 0);
  static constexpr int A_BD = 0, A_FW = 1, A_BU = 2, A_EA = impl ? 3 : 0, A_EB = A_EA + 1;
};

template <class Cfg, int MODE>
size_t line_adj_lds_doubles(int zcls) {
  using AT = AdjTraits<MODE>;
  // carries/edges, coefficient sets, z coefficients (halo) and rock flags of
  // the zcls column classes, inverse face spacing and h along the line, class
  // per element
  return 2 * Cfg::G * Cfg::LW + kSlots * AT::NA * Cfg::NMAX + (size_t)zcls * Cfg::LW * (AT::halo ? 3 : 1) +
         2 * Cfg::NMAX + (Cfg::NMAX + 1) / 2;
}

template <class Cfg, int MODE>
__global__ __launch_bounds__(Cfg::THREADS) void k_line_adj(DevView v, AdjGeo ag, int dir, double* out,
                                                           const double* in, const double* x1, const double* x2,
                                                           double c, double a1, double a2, double a3,
                                                           int64_t tile0, int64_t ntiles) {
  using AT = AdjTraits<MODE>;
  constexpr int S = Cfg::S, NM = Cfg::NMAX, LW = Cfg::LW, H = Cfg::H, G = Cfg::G, KC = S % 4 == 0 ? 4 : 2;
  static_assert(S % KC == 0, "segment length must be a multiple of the chunk");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* shA = reinterpret_cast<double*>(smem_raw);  // [G][LW]
  double* shB = shA + G * LW;                          // [G][LW]
  double* shC = shB + G * LW;                          // [kSlots][NA][NM]
  const int zcls = v.zc_cls;                           // column classes staged
  double* shZ = shC + kSlots * AT::NA * NM;            // [zcls][LW][2] (halo modes)
  double* shR = shZ + (AT::halo ? zcls * LW * 2 : 0);  // [zcls][LW] rock flags
  double* shI = shR + zcls * LW;                       // [NM] 1 / face spacing along the line
  double* shH = shI + NM;                              // [NM] control length along the line
  int* shK = reinterpret_cast<int*>(shH + NM);         // [NM] column class per element

  constexpr int FAM = LW == 32 ? 1 : 0;
  const LineTile* tiles = AT::halo ? v.htiles[FAM] : (dir == 0 ? v.xtiles[FAM] : v.ytiles[FAM]);
  const LineTile T = tiles[tile0 + xcd_remap(blockIdx.x, ntiles)];
  const ScenDev S_ = v.sc[T.scen];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int kk = lane % LW;
  const int hh = lane / LW;
  const int seg = wv * H + hh;
  const int N = T.N;
  const int nz = T.nz;
  const int k = T.kw + kk;
  const int kc = k < 0 ? 0 : (k >= nz ? nz - 1 : k);
  const int kmin = T.kw < 0 ? 0 : T.kw;
  const unsigned kb = (unsigned)(kc - kmin) * (unsigned)sizeof(double) +
                      (unsigned)(hh * S) * (unsigned)T.stride * (unsigned)sizeof(double);
  const int64_t stride = T.stride;
  const int64_t base = T.base;

  // As the forward kernel: the coefficient sets (its padded copies: bd, fw,
  // bu, ea, eb) and the rock flags go into LDS by LDS DMA, the per-element
  // values through registers written after the line's loads are issued.
  const int slot = (int)((T.slot_b0 >> kk) & 1u) | ((int)((T.slot_b1 >> kk) & 1u) << 1);
  {
    constexpr int NCH = AT::NA * NM / 2;  // 16-byte chunks of a slot
    for (int q = 0; q < T.nset; ++q) {
      const double* pg = v.pcoef[FAM] + (q == 0 ? T.poff[0] : (q == 1 ? T.poff[1] : T.poff[2]));
      double* dst = shC + q * AT::NA * NM;
#pragma unroll
      for (int i = 0; i < (NCH + Cfg::THREADS - 1) / Cfg::THREADS; ++i) {
        const int cw = i * Cfg::THREADS + wv * 64;  // the wave's first chunk
        if (cw < NCH) {
          const int ch = cw + lane;
          if (ch < NCH) lds_dma16(pg + 2 * ch, dst + 2 * cw);
        }
      }
    }
  }
  // rock flags of the window's levels per column class: 4-byte DMA loads,
  // lane i copying word i & 1 of level kw + i / 2 (+ 32 in the second load)
  const int ncls = S_.ncls < kZcMax ? S_.ncls : kZcMax;
  for (int q = wv; q < ncls; q += Cfg::NW)
#pragma unroll
    for (int half = 0; half < LW / 32; ++half) {
      const int kq = T.kw + half * 32 + (lane >> 1);
      const int kqc = kq < 0 ? 0 : (kq >= nz ? nz - 1 : kq);
      lds_dma4(reinterpret_cast<const uint32_t*>(ag.rock + (int64_t)q * nz + kqc) + (lane & 1),
               shR + q * LW + half * 32);
    }
  // classes, inverse face spacings and control lengths along the line
  constexpr int NL = (NM + Cfg::THREADS - 1) / Cfg::THREADS;
  int kv[NL];
  double iv[NL], hv[NL];
  {
    const double* gi = dir == 0 ? ag.ix : ag.iy;
    const double* gh = dir == 0 ? ag.hx : ag.hy;
#pragma unroll
    for (int i = 0; i < NL; ++i) {
      const int l = tid + i * Cfg::THREADS;
      const int lc = l < N ? l : N - 1;
      const int64_t col = dir == 0 ? (int64_t)lc * S_.ny + T.line : (int64_t)T.line * S_.ny + lc;
      kv[i] = l < NM ? v.cls[S_.o_cls + col] : 0;
      iv[i] = l + 1 < N ? gi[l] : 0.0;
      hv[i] = l < NM ? gh[lc] : 0.0;
    }
  }

  double d[S];
  const int l0 = seg * S;
  const int64_t ustride = uniform64(stride);
  const int64_t useg = uniform64(base + (int64_t)(wv * H * S) * stride);
  {
#pragma unroll
    for (int r = 0; r < S; ++r) d[r] = (l0 + r < N) ? *lane_ptr(in + soff(useg + (int64_t)r * ustride), kb) : 0.0;
  }
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    const int l = tid + i * Cfg::THREADS;
    if (l < NM) {
      shK[l] = kv[i];
      shI[l] = iv[i];
      shH[l] = hv[i];
    }
  }
  __syncthreads();

  const double* cs = shC + slot * AT::NA * NM;
  const double* bd = cs + AT::A_BD * NM;
  const double* fw = cs + AT::A_FW * NM;
  const double* bu = cs + AT::A_BU * NM;
  const double* ea = cs + AT::A_EA * NM;
  const double* eb = cs + AT::A_EB * NM;
  const bool kin = k >= 0 && k < nz;
  // face factor of this lane's level: the other two control lengths
  const double hline = dir == 0 ? ag.hy[T.line] : ag.hx[T.line];
  const double phi_line = kin ? hline * ag.hz[kc] : 0.0;
  double acc = 0.0;

  // sum over this segment's faces along the line of (mu2-mu1)(x2-x1) times
  // the half rock weights of the two cells, times -c * phi_line (mu: dd)
  auto grad_line = [&](const double(&dd)[S], const double* x) __attribute__((always_inline)) {
    shA[seg * LW + kk] = dd[0];
    __syncthreads();
    const double mu_edge = (l0 + S < N) ? shA[(seg + 1) * LW + kk] : 0.0;
    __syncthreads();
    // x in chunks of KG elements, the next chunk's loads in flight while the
    // current one computes; the barrier keeps the compiler from hoisting every
    // load of the segment (which would spill the line)
    constexpr int KG = S % 8 == 0 ? 8 : (S % 4 == 0 ? 4 : 2);
    const double xafter = (l0 + S < N) ? *lane_ptr(x + soff(useg + (int64_t)S * ustride), kb) : 0.0;
    double xn[KG];
#pragma unroll
    for (int q = 0; q < KG; ++q) xn[q] = (l0 + q < N) ? *lane_ptr(x + soff(useg + (int64_t)q * ustride), kb) : 0.0;
    double gsum = 0.0;
#pragma unroll
    for (int r0 = 0; r0 < S; r0 += KG) {
      double xc[KG];
#pragma unroll
      for (int q = 0; q < KG; ++q) xc[q] = xn[q];
      if (r0 + KG < S) {
#pragma unroll
        for (int q = 0; q < KG; ++q)
          xn[q] = (l0 + r0 + KG + q < N) ? *lane_ptr(x + soff(useg + (int64_t)(r0 + KG + q) * ustride), kb) : 0.0;
      }
#pragma unroll
      for (int q = 0; q < KG; ++q) {
        const int r = r0 + q, l = l0 + r;
        const double xnb = q + 1 < KG ? xc[q + 1] : (r0 + KG < S ? xn[0] : xafter);
        const double mn = (r + 1 < S) ? dd[r + 1] : mu_edge;
        if (l + 1 < N) {
          const double wgt = shR[shK[l] * LW + kk] + shR[shK[l + 1] * LW + kk];
          gsum += shI[l] * (mn - dd[r]) * (xnb - xc[q]) * wgt;
        }
      }
      asm volatile("" ::: "memory");
    }
    acc += -0.5 * c * phi_line * gsum;
  };

  // the solve and the explicit step keep the forward kernel's chunked loops
  // (the next chunk's LDS operands fetched while the current one computes)
  auto thomas = [&]() __attribute__((always_inline)) {
    double y = 0.0, P = 1.0, wn[KC];
#pragma unroll
    for (int q = 0; q < KC; ++q) wn[q] = fw[l0 + q];
#pragma unroll
    for (int r0 = 0; r0 < S; r0 += KC) {
      double wc[KC];
#pragma unroll
      for (int q = 0; q < KC; ++q) wc[q] = wn[q];
      if (r0 + KC < S) {
#pragma unroll
        for (int q = 0; q < KC; ++q) wn[q] = fw[l0 + r0 + KC + q];
      }
#pragma unroll
      for (int q = 0; q < KC; ++q) {
        y = d[r0 + q] - wc[q] * y;
        P = -wc[q] * P;
        d[r0 + q] = y;
      }
    }
    shA[seg * LW + kk] = y;
    shB[seg * LW + kk] = P;
    __syncthreads();
    double C = 0.0;
    for (int g = 0; g < seg; ++g) C = shA[g * LW + kk] + shB[g * LW + kk] * C;
    __syncthreads();
    P = 1.0;
#pragma unroll
    for (int q = 0; q < KC; ++q) wn[q] = fw[l0 + q];
#pragma unroll
    for (int r0 = 0; r0 < S; r0 += KC) {
      double wc[KC];
#pragma unroll
      for (int q = 0; q < KC; ++q) wc[q] = wn[q];
      if (r0 + KC < S) {
#pragma unroll
        for (int q = 0; q < KC; ++q) wn[q] = fw[l0 + r0 + KC + q];
      }
#pragma unroll
      for (int q = 0; q < KC; ++q) {
        P = -wc[q] * P;
        d[r0 + q] = d[r0 + q] + P * C;
      }
    }
    double x = 0.0, Q = 1.0, bn[KC], un[KC];
#pragma unroll
    for (int q = 0; q < KC; ++q) {
      bn[q] = bd[l0 + S - 1 - q];
      un[q] = bu[l0 + S - 1 - q];
    }
#pragma unroll
    for (int r0 = 0; r0 < S; r0 += KC) {
      double bc[KC], uc[KC];
#pragma unroll
      for (int q = 0; q < KC; ++q) { bc[q] = bn[q]; uc[q] = un[q]; }
      if (r0 + KC < S) {
#pragma unroll
        for (int q = 0; q < KC; ++q) {
          bn[q] = bd[l0 + S - 1 - (r0 + KC + q)];
          un[q] = bu[l0 + S - 1 - (r0 + KC + q)];
        }
      }
#pragma unroll
      for (int q = 0; q < KC; ++q) {
        const int r = S - 1 - (r0 + q);
        x = bc[q] * d[r] - uc[q] * x;
        Q = -uc[q] * Q;
        d[r] = x;
      }
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
        d[r] = d[r] + Q * X;
      }
    }
  };

  auto explicit_step = [&]() __attribute__((always_inline)) {
    shA[seg * LW + kk] = d[0];
    shB[seg * LW + kk] = d[S - 1];
    __syncthreads();
    double prev = (seg > 0) ? shB[(seg - 1) * LW + kk] : d[0];
    const double edge = (l0 + S < N) ? shA[(seg + 1) * LW + kk] : d[S - 1];
    __syncthreads();
    double an[KC], bn2[KC];
#pragma unroll
    for (int q = 0; q < KC; ++q) { an[q] = ea[l0 + q]; bn2[q] = eb[l0 + q]; }
#pragma unroll
    for (int r0 = 0; r0 < S; r0 += KC) {
      double ac[KC], bc2[KC];
#pragma unroll
      for (int q = 0; q < KC; ++q) { ac[q] = an[q]; bc2[q] = bn2[q]; }
      if (r0 + KC < S) {
#pragma unroll
        for (int q = 0; q < KC; ++q) { an[q] = ea[l0 + r0 + KC + q]; bn2[q] = eb[l0 + r0 + KC + q]; }
      }
#pragma unroll
      for (int q = 0; q < KC; ++q) {
        const int r = r0 + q;
        const double xv = d[r];
        const double xp = (r < S - 1) ? d[r + 1] : edge;
        d[r] = xv + c * (ac[q] * (xp - xv) - bc2[q] * (xv - prev));
        prev = xv;
      }
    }
  };

  // the line is complete after each solve: without the register barrier the
  // compiler interleaves the solve's last pass with the next phase and holds
  // both sets of values
  auto settle = [&](double(&dd)[S]) __attribute__((always_inline)) {
#pragma unroll
    for (int r = 0; r < S; ++r) asm volatile("" : "+v"(dd[r]));
  };
  if (MODE == ADJ_Y) {
    thomas();
    settle(d);
    grad_line(d, x1);
  } else if (S <= 32) {
    // G(mu_in, x) + G(mu_out, x) = G(mu_in + mu_out, x): one pass over x,
    // with the input line kept in registers
    double d0[S];
#pragma unroll
    for (int r = 0; r < S; ++r) d0[r] = d[r];
    explicit_step();
    thomas();
    settle(d);
#pragma unroll
    for (int r = 0; r < S; ++r) d0[r] += d[r];
    grad_line(d0, x1);
  } else {
    grad_line(d, x1);
    explicit_step();
    thomas();
    settle(d);
    grad_line(d, x1);
  }

  // per-(line, level) gradient: segments summed in order by segment 0
  shA[seg * LW + kk] = acc;
  __syncthreads();
  if (seg == 0 && k >= T.ko0 && k < T.ko1) {
    double tot = 0.0;
    for (int g = 0; g < G; ++g) tot += shA[g * LW + kk];
    double* part = dir == 0 ? ag.gx_part : ag.gy_part;
    part[(int64_t)T.line * nz + k] += tot;
  }

  if (k >= T.ko0 && k < T.ko1) {
#pragma unroll
    for (int r = 0; r < S; ++r)
      if (l0 + r < N) *lane_ptr(out + soff(useg + (int64_t)r * ustride), kb) = d[r];
  }
}

template <class Cfg, int MODE>
static void line_adj_launch(const DevView& v, const AdjGeo& ag, int dir, double* out, const double* in,
                            const double* x1, const double* x2, double c, double a1, double a2, double a3,
                            int64_t tile0, int64_t ntiles, hipStream_t st) {
  const size_t sh = line_adj_lds_doubles<Cfg, MODE>(v.zc_cls) * sizeof(double);
  hipLaunchKernelGGL((k_line_adj<Cfg, MODE>), dim3((unsigned)ntiles), dim3(Cfg::THREADS), sh, st, v, ag, dir, out,
                     in, x1, x2, c, a1, a2, a3, tile0, ntiles);
}

template <class Cfg>
static hipError_t line_adj_cfg(const DevView& v, const AdjGeo& ag, int mode, double* out, const double* in,
                               const double* x1, const double* x2, double c, double a1, double a2, double a3,
                               int64_t t0, int64_t n, hipStream_t st) {
  switch (mode) {
    case ADJ_Y: line_adj_launch<Cfg, ADJ_Y>(v, ag, 1, out, in, x1, x2, c, a1, a2, a3, t0, n, st); break;
    case ADJ_X: line_adj_launch<Cfg, ADJ_X>(v, ag, 0, out, in, x1, x2, c, a1, a2, a3, t0, n, st); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Tile shapes of the adjoint line kernels.  The forward kernels' tiles, with
// the segments of the 33-320-element classes cut to 16-24 elements on twice
// the waves: ADJ_X keeps its input line beside the solved one (one pass over
// the forward field by linearity), which at 32 elements per lane held it to
// two waves per SIMD.
template <int LW> struct LineFamilyAdj;
template <> struct LineFamilyAdj<64> {
  using C0 = LineCfg<1, 48, 64>;
  using C1 = LineCfg<4, 16, 64>;
  using C2 = LineCfg<4, 24, 64>;
  using C3 = LineCfg<8, 16, 64>;
  using C4 = LineCfg<8, 20, 64>;
  using C5 = LineCfg<8, 24, 64>;
  using C6 = LineCfg<16, 16, 64>;
  using C7 = LineCfg<16, 20, 64>;
  using C8 = LineCfg<16, 26, 64>;
  using C9 = LineCfg<8, 80, 64>;
};
template <> struct LineFamilyAdj<32> {
  using C0 = LineCfg<1, 24, 32>;
  using C1 = LineCfg<2, 16, 32>;
  using C2 = LineCfg<2, 24, 32>;
  using C3 = LineCfg<4, 16, 32>;
  using C4 = LineCfg<4, 20, 32>;
  using C5 = LineCfg<4, 24, 32>;
  using C6 = LineCfg<8, 16, 32>;
  using C7 = LineCfg<8, 20, 32>;
  using C8 = LineCfg<8, 26, 32>;
  using C9 = LineCfg<8, 40, 32>;
};
template <int LW>
constexpr bool adj_family_matches_classes() {
  using F = LineFamilyAdj<LW>;
  return F::C0::NMAX == kLineClassMax[0] && F::C1::NMAX == kLineClassMax[1] && F::C2::NMAX == kLineClassMax[2] &&
         F::C3::NMAX == kLineClassMax[3] && F::C4::NMAX == kLineClassMax[4] && F::C5::NMAX == kLineClassMax[5] &&
         F::C6::NMAX == kLineClassMax[6] && F::C7::NMAX == kLineClassMax[7] && F::C8::NMAX == kLineClassMax[8] &&
         F::C9::NMAX == kLineClassMax[9];
}
static_assert(adj_family_matches_classes<64>() && adj_family_matches_classes<32>(), "tile shapes follow the classes");

// One launch per non-empty line-length class of the mode's tiles.
template <int LW>
hipError_t line_adj_family(const DevView& v, const AdjGeo& ag, int mode, double* out, const double* in,
                           const double* x1, const double* x2, double c, double a1, double a2, double a3,
                           hipStream_t st) {
  using F = LineFamilyAdj<LW>;
  const int f = LW == 32 ? 1 : 0;
  const int64_t* cls = mode == ADJ_X ? v.xcls[f] : v.ycls[f];
  for (int q = 0; q < kLineClasses; ++q) {
    const int64_t t0 = cls[q], n = cls[q + 1] - cls[q];
    if (n == 0) continue;
    hipError_t e;
    switch (q) {
      case 0: e = line_adj_cfg<typename F::C0>(v, ag, mode, out, in, x1, x2, c, a1, a2, a3, t0, n, st); break;
      case 1: e = line_adj_cfg<typename F::C1>(v, ag, mode, out, in, x1, x2, c, a1, a2, a3, t0, n, st); break;
      case 2: e = line_adj_cfg<typename F::C2>(v, ag, mode, out, in, x1, x2, c, a1, a2, a3, t0, n, st); break;
      case 3: e = line_adj_cfg<typename F::C3>(v, ag, mode, out, in, x1, x2, c, a1, a2, a3, t0, n, st); break;
      case 4: e = line_adj_cfg<typename F::C4>(v, ag, mode, out, in, x1, x2, c, a1, a2, a3, t0, n, st); break;
      case 5: e = line_adj_cfg<typename F::C5>(v, ag, mode, out, in, x1, x2, c, a1, a2, a3, t0, n, st); break;
      case 6: e = line_adj_cfg<typename F::C6>(v, ag, mode, out, in, x1, x2, c, a1, a2, a3, t0, n, st); break;
      case 7: e = line_adj_cfg<typename F::C7>(v, ag, mode, out, in, x1, x2, c, a1, a2, a3, t0, n, st); break;
      case 8: e = line_adj_cfg<typename F::C8>(v, ag, mode, out, in, x1, x2, c, a1, a2, a3, t0, n, st); break;
      default: e = line_adj_cfg<typename F::C9>(v, ag, mode, out, in, x1, x2, c, a1, a2, a3, t0, n, st); break;
    }
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace gw
