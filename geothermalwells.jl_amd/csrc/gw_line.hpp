// gw_line.hpp -- the line kernel (x/y lines, register-resident, with the
// fused z-step modes) and its per-family launcher.  Each k-window family and
// state precision is instantiated in its own translation unit
// (gw_line_w<LW>_<prec>.hip).
#pragma once

#include <algorithm>
#include <type_traits>

#include "gw_device.hpp"

namespace gw {

// ---------------------------------------------------------------------------
// Line kernel.  A tile is a window of LW adjacent k-lines (k = kw .. kw+LW-1)
// of one x-line (fixed j) or y-line (fixed i); its outputs are the lanes with
// ko0 <= k < ko1.  The line is cut into G segments of S elements (G*S >= N),
// each held in registers by LW lanes (one per k): a whole wave when LW = 64,
// a half-wave when LW = 32 (the wave's two halves hold consecutive segments).
// Loads and stores of one element index are LW consecutive k values:
// coalesced.  The Thomas recurrences run segment-locally from zero carry and
// are then corrected by the carried value times the running product of the
// multipliers; the carries are a G-step scan in LDS.  The lanes of a window
// use at most three coefficient sets (one per material zone met in z; four in
// the halo windows of ZY / YZY), staged in LDS and padded with identity rows
// so the element loops have no branches.
//
// Modes (the reference's operators on the same direction, src/solver.jl):
//   EXPLICIT            u <- (I + cA) u                               (:34-58)
//   IMPLICIT            u <- (I - cA)^-1 u                            (:111-232)
//   IMPLICIT_EXPLICIT   u <- (I + cA)(I - cA)^-1 u   x-lines, :450 + :455
//   ZY                  u <- (I + cA_y) R(dt A_z) u   next step's z-step and
//                       explicit y (:23-31 under ROCK2, then :444)
//   YZY                 u <- (I + cA_y) R(dt A_z) (I - cA_y)^-1 u      (:461, then
//                       the next step's z-step and explicit y)
//   IMPLICIT_EXPLICIT_KEEP   IMPLICIT_EXPLICIT that also writes the implicit
//                       result to DevView::aux (the adjoint's recompute)
// ZY / YZY run on y-lines with a 3-lane halo in k: the z-stencil is applied
// three times across lanes (wave-wide DPP shifts), so only lanes 3..LW-4 of a
// window are outputs; the halo lanes re-solve their own lines.  A shift that
// crosses from one half-wave into the other (LW = 32) reaches a halo lane
// only, and in three stages its value travels at most three lanes: the
// outputs never see it.
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

constexpr int kSlots = 3;      // coefficient sets per plain window (LDS slots)
constexpr int kSlotsHalo = 4;  // most sets per z-halo window

template <int MODE> struct ModeTraits {
  static constexpr bool halo = MODE == LINE_ZY || MODE == LINE_YZY || MODE == LINE_YZY_KEEP;
  static constexpr bool impl = MODE == LINE_IMPLICIT || MODE == LINE_IMPLICIT_EXPLICIT || MODE == LINE_YZY ||
                               MODE == LINE_IMPLICIT_EXPLICIT_KEEP || MODE == LINE_YZY_KEEP;
  static constexpr bool expl = MODE != LINE_IMPLICIT;
  static constexpr bool keep = MODE == LINE_IMPLICIT_EXPLICIT_KEEP || MODE == LINE_YZY_KEEP;
  // coefficient arrays staged per set slot: bd, fw, bu (implicit) then ea, eb
  // (explicit), each NMAX long
  static constexpr int NA = (impl ? 3 : 0) + (expl ? 2 : 0);
  static constexpr int A_BD = 0, A_FW = 1, A_BU = 2, A_EA = impl ? 3 : 0, A_EB = A_EA + 1;
};

// Type of the explicit-step and z-step coefficients a tile keeps in LDS: the
// state type in the FP32 64-lane kernels (rounded once at staging, not at
// every use: the same values), FP64 elsewhere (the FP32 32-lane kernels'
// register allocation does worse with it).
template <class TS, int LW>
using LineCT = typename std::conditional<sizeof(TS) == 4 && LW == 64, TS, double>::type;

// LDS of a tile, in doubles: carries and edges, the z coefficients of the
// zcls column classes and the column class per element (halo modes), then
// `slots` coefficient-set slots (last, so the slot count only changes the
// launch's LDS size).
template <class Cfg, class TS, int MODE>
__host__ __device__ inline size_t line_lds_fixed(int zcls) {
  using MT = ModeTraits<MODE>;
  constexpr size_t zc_per_cls = Cfg::LW * 2 * sizeof(LineCT<TS, Cfg::LW>) / sizeof(double);
  return 2 * Cfg::G * Cfg::LW + (MT::halo ? (size_t)zcls * zc_per_cls + (Cfg::NMAX + 1) / 2 : 0);
}
template <class Cfg, class TS, int MODE>
size_t line_lds_doubles(int slots, int zcls) {
  return line_lds_fixed<Cfg, TS, MODE>(zcls) + (size_t)slots * ModeTraits<MODE>::NA * Cfg::NMAX;
}

GW_DEV int64_t xcd_remap(int64_t b, int64_t n) {
  // Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
  // contiguous range of tiles so neighbouring k-windows of a line share an L2.
  const int64_t per = n / 8, rem = n % 8;
  const int64_t x = b % 8, i = b / 8;
  return x * per + (x < rem ? x : rem) + i;
}

// Waves per SIMD the register budget aims at: a tile of NW waves puts NW/4 on
// each SIMD, so four-wave tiles run three to a CU when the line fits in 168
// registers (S <= 52), eight-wave tiles two to a CU when it fits in 128
// (sixteen-wave tiles, 1024 threads, one).  An FP32 line takes half the
// registers: four-wave FP32 tiles aim at five per CU (96 registers) up to
// S = 32 and four (128) up to S = 48.
template <class Cfg, class TS, int MODE, bool SR = false>
constexpr int line_min_waves() {
  if (sizeof(TS) == 4 && Cfg::NW <= 4 && Cfg::S > 32 && Cfg::S <= 40) return 4;
  // the stochastic-rounding 8-wave 40-element tiles: four per SIMD (128
  // registers, a few spilled) run faster than three without spills
  if (SR && Cfg::NW == 8 && Cfg::S == 40) return 4;
  if (MODE == LINE_IMPLICIT_EXPLICIT || MODE == LINE_ZY) return 2;
  return Cfg::NW <= 4 ? (Cfg::S <= 52 ? 3 : 2) : (Cfg::S <= 32 ? 4 : 2);
}

// Carries of the segmented recurrences.  Segment g leaves the map
// v -> m.x + m.y v (its end value from zero carry, and the product of its
// multipliers) in sh[g][kk]; a lane's carry composes the maps of the
// segments before it (forward) or after it (backward), reading four of them
// per LDS round trip.  The loop bounds are the wave's (its last lane's
// segment forward, its first backward), the lane's own segment masks the rest.
template <int G, int LW, int H, int B>
GW_DEV double carry_fwd(const double2* sh, int kk, int seg, int wv) {
  double C = 0.0;
  const int gmax = wv * H + H - 1;  // most segments any lane of the wave composes
#pragma unroll
  for (int g0 = 0; g0 < G; g0 += B) {
    if (g0 >= gmax) break;
    double2 m[B];
#pragma unroll
    for (int i = 0; i < B; ++i)
      if (g0 + i < G) m[i] = sh[(g0 + i) * LW + kk];
#pragma unroll
    for (int i = 0; i < B; ++i)
      if (g0 + i < G && g0 + i < seg) C = m[i].x + m[i].y * C;
  }
  return C;
}
template <int G, int LW, int H, int B>
GW_DEV double carry_bwd(const double2* sh, int kk, int seg, int wv) {
  double X = 0.0;
  const int gmin = wv * H;  // the wave's first segment: every lane composes the segments after its own
#pragma unroll
  for (int g0 = G - 1; g0 >= 0; g0 -= B) {
    if (g0 <= gmin) break;
    double2 m[B];
#pragma unroll
    for (int i = 0; i < B; ++i)
      if (g0 - i >= 0) m[i] = sh[(g0 - i) * LW + kk];
#pragma unroll
    for (int i = 0; i < B; ++i)
      if (g0 - i >= 0 && g0 - i > seg) X = m[i].x + m[i].y * X;
  }
  return X;
}

// Column class of each line element (the last one's past N), loaded into
// registers and then written to LDS.
template <class Cfg>
GW_DEV void load_line_classes(const uint16_t* clsl, int N, int tid, uint16_t* kv) {
  constexpr int NM = Cfg::NMAX, NL = (NM + Cfg::THREADS - 1) / Cfg::THREADS;
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    const int l = tid + i * Cfg::THREADS;
    kv[i] = l < NM ? clsl[l < N ? l : N - 1] : 0;
  }
}
template <class Cfg>
GW_DEV void store_line_classes(const uint16_t* kv, int* shK, int tid) {
  constexpr int NM = Cfg::NMAX, NL = (NM + Cfg::THREADS - 1) / Cfg::THREADS;
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    const int l = tid + i * Cfg::THREADS;
    if (l < NM) shK[l] = kv[i];
  }
}

// SR (FP32 state only): every rounding of an updated value to FP32 is
// stochastic (sr_f32): the implicit solve's two stores, the fused z-step's
// result and the explicit step's result; round to nearest otherwise.
template <class Cfg, class TS, int MODE, bool SR>
__global__ __launch_bounds__(Cfg::THREADS, (line_min_waves<Cfg, TS, MODE, SR>())) void k_line(DevView v, int dir, TS* out, const TS* in,
                                                          double c, double a1, double a2, double a3,
                                                          int64_t tile0, int64_t ntiles) {
  using MT = ModeTraits<MODE>;
  static_assert(!SR || sizeof(TS) == 4, "stochastic rounding of an FP32 state");
  // The line is held in registers in the storage type; every recurrence runs
  // in FP64.  In FP32-state kernels the values kept between the passes (the
  // line, z_l = bd_l y_l, the solution) are FP32: each element is rounded
  // twice in an implicit solve.
  using RT = TS;
  constexpr int S = Cfg::S, NM = Cfg::NMAX, LW = Cfg::LW, H = Cfg::H, G = Cfg::G;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* shA = reinterpret_cast<double*>(smem_raw);  // [G][LW] carries / edges
  double* shB = shA + G * LW;                          // [G][LW]
  double2* shM = reinterpret_cast<double2*>(shA);      // [G][LW] segment maps of the carry scans (over shA, shB)
  double* shZ = shB + G * LW;                          // [zcls][LW][2]: z coefficients (halo modes)
  // The explicit step and the z-step run in the state type, their LDS
  // coefficients are of type CT (LineCT).
  using CT = LineCT<TS, LW>;
  CT* shZt = reinterpret_cast<CT*>(shZ);
  const int zcls = v.zc_cls;
  int* shK = reinterpret_cast<int*>(shZt + zcls * LW * 2);  // [NM] column class per line element
  double* shC = shA + line_lds_fixed<Cfg, TS, MODE>(zcls);  // [nset][MT::NA][NM] coefficient arrays

  constexpr int FAM = LW == 32 ? 1 : 0;
  const LineTile* tiles = MT::halo ? v.htiles[FAM] : (dir == 0 ? v.xtiles[FAM] : v.ytiles[FAM]);
  const LineTile T = tiles[tile0 + xcd_remap(blockIdx.x, ntiles)];
  const ScenDev S_ = v.sc[T.scen];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int kk = lane % LW;    // k offset in the window
  const int hh = lane / LW;    // half-wave (0 when LW = 64)
  const int seg = wv * H + hh;
  // the line's geometry from the tile descriptor (the scenario's fields are
  // only needed by the z-step staging and the stochastic rounding)
  const int N = T.N;
  const int nz = T.nz;
  const int k = T.kw + kk;
  const int kc = k < 0 ? 0 : (k >= nz ? nz - 1 : k);
  const int kmin = T.kw < 0 ? 0 : T.kw;
  // lane byte offset from the wave's first segment: the k offset plus, for
  // the upper half-wave, S line elements (below 2^32: nx, ny <= 640,
  // nz <= 4096, S <= 52)
  const unsigned kb = (unsigned)(kc - kmin) * (unsigned)sizeof(TS) +
                      (unsigned)(hh * S) * (unsigned)T.stride * (unsigned)sizeof(TS);
  const int64_t stride = T.stride;
  const int64_t base = T.base;

  // the coefficient slot of this lane's level (set per tile on the host)
  const int slot = (int)((T.slot_b0 >> kk) & 1u) | ((int)((T.slot_b1 >> kk) & 1u) << 1);
  constexpr int NL = (NM + Cfg::THREADS - 1) / Cfg::THREADS;  // staged elements per thread and array
  // A tile's loads go out right after its descriptor's scalar round trip:
  // the coefficient sets and z coefficients are copied into LDS by LDS DMA
  // (no registers, no wait: the barrier after the line's loads waits for
  // them), from copies laid out as the LDS holds them; the column classes
  // go through registers, written to LDS after the line's loads are issued.
  {
    // the sets' padded copies (bd, fw, bu, explicit coefficients, NM each:
    // the LDS slot's layout), 16 bytes per lane and DMA load
    constexpr int NCH = MT::NA * NM / 2;  // 16-byte chunks of a slot
    const double* pc = v.pcoef[FAM] + (MT::impl ? 0 : 3 * NM);
    for (int q = 0; q < T.nset; ++q) {
      const double* pg = pc + (q == 0 ? T.poff[0] : (q == 1 ? T.poff[1] : (q == 2 ? T.poff[2] : T.poff[3])));
      double* dst = shC + q * MT::NA * NM;
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
  if (MT::halo) {
    const int ncls = S_.ncls < kZcMax ? S_.ncls : kZcMax;
    if constexpr (std::is_same<CT, double>::value) {
      // {ea, eb} of level kc: one class per half-wave (LW = 32) or wave
      // (LW = 64), lane (hh, kk) copies class q0 + hh
      const double* zc = v.zcoef + 2 * S_.o_zcoef;
      for (int q0 = wv * H; q0 < ncls; q0 += Cfg::NW * H) {
        const int q = q0 + hh;
        if (q < ncls) lds_dma16(zc + 2 * ((int64_t)q * nz + kc), shZ + q0 * LW * 2);
      }
    } else {
      // float pairs (64-lane FP32 tiles): a class is 64 lanes x 8 bytes, two
      // 4-byte DMA loads per class, lane i copying component i & 1 of the
      // pair of level kw + i / 2 (+ 32 in the second load)
      const float* zc = v.zcoef32 + 2 * S_.o_zcoef;
      for (int q = wv; q < ncls; q += Cfg::NW)
#pragma unroll
        for (int half = 0; half < 2; ++half) {
          const int kq = T.kw + half * 32 + (lane >> 1);
          const int kqc = kq < 0 ? 0 : (kq >= nz ? nz - 1 : kq);
          lds_dma4(zc + 2 * ((int64_t)q * nz + kqc) + (lane & 1), shZt + (q * LW + half * 32) * 2);
        }
    }
  }
  uint16_t kv[NL];
  if (MT::halo) load_line_classes<Cfg>(v.cls + S_.o_cls + (int64_t)S_.ny * T.line, N, tid, kv);

  RT d[S];
  const int l0 = seg * S;
  // Stochastic rounding: rounding point p (0, 1: the implicit solve's
  // forward and backward stores, 2: the z-step, 3: the explicit step; + 4 on
  // y-lines) of step n draws, for element l of this line at level kc, the
  // l-th state of the stream sr_stream(first element's flat index, n, p).
  // Every tile that holds the element (as output or as halo) therefore draws
  // the same number whatever its tile shape, and a run split into several
  // calls rounds exactly like one call.
  uint32_t sr_idx = 0u, sr_n = 0u;
  LcgJump sr_j{1u, 0u};
  if constexpr (SR) {
    sr_n = *v.step_ctr;
    sr_idx = (uint32_t)(base - kmin + kc);  // the first element's flat index
    constexpr LcgJump js = lcg_jump(S);  // to element l0 = seg * S
    for (int g = 0; g < seg; ++g) sr_j = LcgJump{(sr_j.a * js.a) & kLcgMask, (sr_j.c * js.a + js.c) & kLcgMask};
  }
  auto sr_first = [&](uint32_t n, int p) { return lcg_apply(sr_j, sr_stream(sr_idx, n, (uint32_t)(dir * 4 + p))); };
  // the z-step and explicit y of YZY belong to the next step
  const uint32_t sr_n2 = MODE == LINE_YZY ? sr_n + 1u : sr_n;
  const int64_t ustride = uniform64(stride);
  const int64_t useg = uniform64(base + (int64_t)(wv * H * S) * stride);
  // element r of the lane's segment: field base + an element offset kept in
  // SGPRs (soff; left to itself the compiler forms 64-bit vector addresses,
  // two to three VALU instructions per access) + the lane's byte offset
  if (l0 + S <= N) {
#pragma unroll
    for (int r = 0; r < S; ++r) d[r] = (RT)*lane_ptr(in + soff(useg + (int64_t)r * ustride), kb);
  } else {
#pragma unroll
    for (int r = 0; r < S; ++r) d[r] = (l0 + r < N) ? (RT)*lane_ptr(in + soff(useg + (int64_t)r * ustride), kb) : (RT)0;
  }
  if (MT::halo) store_line_classes<Cfg>(kv, shK, tid);
  __syncthreads();

  const double* cs = shC + slot * MT::NA * NM;
  const double* bd = cs + MT::A_BD * NM;
  const double* fw = cs + MT::A_FW * NM;
  const double* bu = cs + MT::A_BU * NM;
  const CT* ea = reinterpret_cast<const CT*>(cs + MT::A_EA * NM);
  const CT* eb = reinterpret_cast<const CT*>(cs + MT::A_EB * NM);

  // Every element loop below runs in chunks of KC elements: the LDS operands
  // of the next chunk are fetched while the current one computes.  The long
  // FP64 segments of the fused z-step modes use chunks of two (their
  // registers are otherwise short at three waves per SIMD), and so do the
  // stochastic-rounding fused z-step 8-wave 40-element tiles (at four waves
  // per SIMD), and segments whose length is not a multiple of four.
  constexpr int KC = ((MT::halo && S > 40 && sizeof(TS) == 8) || (sizeof(TS) == 4 && Cfg::NW <= 4 && S > 32 && S <= 40) ||
                      (SR && MT::halo && Cfg::NW == 8 && S == 40) || S % 4 != 0) ? 2 : 4;
  static_assert(S % KC == 0, "segment length must be a multiple of the chunk");
  // The carry scans of the FP64 sixteen-segment fused z-step tiles read four
  // segment maps per LDS round trip (interleaved {value, product} pairs:
  // 1.5% off the 401-point lines' fused kernel).  The other shapes read one
  // map per step from two arrays
  // (the batched form's extra registers cost several of them a wave per
  // SIMD, and it made the 401-point implicit-y tiles 4.6% slower).
  constexpr bool kScan4 = sizeof(TS) == 8 && G >= 16 && MT::halo;
  if constexpr (MT::impl && sizeof(TS) == 4) {
    // Forward elimination y_l = b_l - fw_l y_{l-1}, then back substitution
    // x_l = bd_l y_l - bu_l x_{l+1}, each in two passes over the segment:
    // first the segment's map from its incoming to its outgoing value
    // (v_out = V + M v_in, from zero carry, nothing stored), the carries by a
    // G-step scan of those maps in LDS, then the exact recurrence from the
    // incoming carry.  The forward pass stores z_l = bd_l y_l, so the back
    // substitution is one FMA per element.
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
        y = (double)d[r0 + q] - wc[q] * y;
        P = -wc[q] * P;
      }
    }
    shA[seg * LW + kk] = y;
    shB[seg * LW + kk] = P;
    __syncthreads();
    y = 0.0;
    for (int g = 0; g < seg; ++g) y = shA[g * LW + kk] + shB[g * LW + kk] * y;
    __syncthreads();
    double bn[KC];
    uint32_t rs = 0u;
    if constexpr (SR) rs = sr_first(sr_n, 0);
#pragma unroll
    for (int q = 0; q < KC; ++q) { wn[q] = fw[l0 + q]; bn[q] = bd[l0 + q]; }
#pragma unroll
    for (int r0 = 0; r0 < S; r0 += KC) {
      double wc[KC], bc[KC];
#pragma unroll
      for (int q = 0; q < KC; ++q) { wc[q] = wn[q]; bc[q] = bn[q]; }
      if (r0 + KC < S) {
#pragma unroll
        for (int q = 0; q < KC; ++q) { wn[q] = fw[l0 + r0 + KC + q]; bn[q] = bd[l0 + r0 + KC + q]; }
      }
#pragma unroll
      for (int q = 0; q < KC; ++q) {
        y = (double)d[r0 + q] - wc[q] * y;
        if constexpr (SR) {
          d[r0 + q] = sr_f32(bc[q] * y, rs);
          rs = lcg_next(rs);
        } else {
          d[r0 + q] = bc[q] * y;
        }
      }
    }
    double x = 0.0, Q = 1.0, un[KC];
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
        x = (double)d[S - 1 - (r0 + q)] - uc[q] * x;
        Q = -uc[q] * Q;
      }
    }
    shA[seg * LW + kk] = x;
    shB[seg * LW + kk] = Q;
    __syncthreads();
    x = 0.0;
    for (int g = G - 1; g > seg; --g) x = shA[g * LW + kk] + shB[g * LW + kk] * x;
    __syncthreads();
    if constexpr (SR) {
      constexpr LcgJump jb = lcg_jump(S - 1);  // the segment's last element, stepped back from there
      rs = lcg_apply(jb, sr_first(sr_n, 1));
    }
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
        x = (double)d[r] - uc[q] * x;
        if constexpr (SR) {
          d[r] = sr_f32(x, rs);
          rs = lcg_prev(rs);
        } else {
          d[r] = x;
        }
      }
    }
  }

  if constexpr (MT::impl && sizeof(TS) == 8) {
    // FP64 lines: each pass runs segment-locally from zero carry and is then
    // corrected by the carried value times the running product of the
    // multipliers (four passes; measured faster than the carry-first scheme
    // for the FP64 register budgets).
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
    }
    double C = 0.0;
    if constexpr (kScan4) {
      shM[seg * LW + kk] = make_double2(y, P);
      __syncthreads();
      C = carry_fwd<G, LW, H, 4>(shM, kk, seg, wv);
    } else {
      shA[seg * LW + kk] = y;
      shB[seg * LW + kk] = P;
      __syncthreads();
      for (int g = 0; g < seg; ++g) C = shA[g * LW + kk] + shB[g * LW + kk] * C;
    }
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
    }
    double X = 0.0;
    if constexpr (kScan4) {
      shM[seg * LW + kk] = make_double2(x, Q);
      __syncthreads();
      X = carry_bwd<G, LW, H, 4>(shM, kk, seg, wv);
    } else {
      shA[seg * LW + kk] = x;
      shB[seg * LW + kk] = Q;
      __syncthreads();
      for (int g = G - 1; g > seg; --g) X = shA[g * LW + kk] + shB[g * LW + kk] * X;
    }
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
    }
  }

  // the solved line is complete here: without this the compiler interleaves
  // the last fix-up pass with the next phase and holds both sets of values
  if (MT::impl) {
#pragma unroll
    for (int r = 0; r < S; ++r) asm volatile("" : "+v"(d[r]));
  }

  if (MT::keep && k >= T.ko0 && k < T.ko1) {
#pragma unroll
    for (int r = 0; r < S; ++r)
      if (l0 + r < N) *lane_ptr(static_cast<TS*>(v.aux) + soff(useg + (int64_t)r * ustride), kb) = (TS)d[r];
  }

  if (MT::halo) {
    // z-step R(dt A_z) across lanes (lane = k), column class per line element;
    // neighbours k-1 / k+1 come from lanes -1 / +1 by DPP wave shifts (the
    // window edge lanes keep their own value: they are halo, never output)
    constexpr int KZ = KC;
    uint32_t rz = 0u;
    if constexpr (SR) rz = sr_first(sr_n2, 2);
    TS ean[KZ], ebn[KZ];
#pragma unroll
    for (int q = 0; q < KZ; ++q) {
      const int cl = shK[l0 + q];
      ean[q] = (TS)shZt[(cl * LW + kk) * 2];
      ebn[q] = (TS)shZt[(cl * LW + kk) * 2 + 1];
    }
#pragma unroll
    for (int r0 = 0; r0 < S; r0 += KZ) {
      TS eac[KZ], ebc[KZ];
#pragma unroll
      for (int q = 0; q < KZ; ++q) { eac[q] = ean[q]; ebc[q] = ebn[q]; }
      if (r0 + KZ < S) {
#pragma unroll
        for (int q = 0; q < KZ; ++q) {
          const int cl = shK[l0 + r0 + KZ + q];
          ean[q] = (TS)shZt[(cl * LW + kk) * 2];
          ebn[q] = (TS)shZt[(cl * LW + kk) * 2 + 1];
        }
      }
      // the chunk's KZ elements advance through the three stages together
      // (KZ independent chains per stage); the stage loop stays a real loop:
      // unrolled, the compiler interleaves chunks and the line spills
      if constexpr (SR) {
        // stages in increments: w = u0 + e, e <- al (ea (du+ + e+ - e) -
        // eb (du- + e - e-)) with the differences du of the loaded values
        // (exact), so no stage rounds u0 + e; one stochastic rounding at the end
        TS u0[KZ], dp[KZ], dm[KZ], e[KZ];
#pragma unroll
        for (int q = 0; q < KZ; ++q) {
          u0[q] = (TS)d[r0 + q];
          dp[q] = wave_shl1(u0[q]) - u0[q];
          dm[q] = u0[q] - wave_shr1(u0[q]);
          e[q] = (TS)0;
        }
#pragma unroll 1
        for (int st = 0; st < 3; ++st) {
          const TS al = st == 0 ? (TS)a1 : (st == 1 ? (TS)a2 : (TS)a3);
#pragma unroll
          for (int q = 0; q < KZ; ++q) {
            const TS em = wave_shr1(e[q]);
            const TS ep = wave_shl1(e[q]);
            e[q] = al * (eac[q] * (dp[q] + (ep - e[q])) - ebc[q] * (dm[q] + (e[q] - em)));
          }
        }
#pragma unroll
        for (int q = 0; q < KZ; ++q) {
          d[r0 + q] = sr_f32((double)u0[q] + (double)e[q], rz);
          rz = lcg_next(rz);
        }
      } else {
        TS u0[KZ], w[KZ];
#pragma unroll
        for (int q = 0; q < KZ; ++q) { u0[q] = (TS)d[r0 + q]; w[q] = u0[q]; }
#pragma unroll 1
        for (int st = 0; st < 3; ++st) {
          const TS al = st == 0 ? (TS)a1 : (st == 1 ? (TS)a2 : (TS)a3);
#pragma unroll
          for (int q = 0; q < KZ; ++q) {
            const TS wm = wave_shr1(w[q]);
            const TS wp = wave_shl1(w[q]);
            w[q] = u0[q] + al * (eac[q] * (wp - w[q]) - ebc[q] * (w[q] - wm));
          }
        }
#pragma unroll
        for (int q = 0; q < KZ; ++q) d[r0 + q] = (RT)w[q];
      }
    }
  }

  if (MT::expl) {
    // explicit half-step of the line direction: u + c (ea (u+ - u) - eb (u - u-)),
    // FP32 lines with the products c ea, c eb staged in LDS
    const TS cc = (TS)c;
    uint32_t rx = 0u;
    if constexpr (SR) rx = sr_first(sr_n2, 3);
    shA[seg * LW + kk] = (double)d[0];
    shB[seg * LW + kk] = (double)d[S - 1];
    __syncthreads();
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
        TS o;
        if constexpr (SR) {
          o = sr_f32((double)x + (double)(ac[q] * (xp - x) - bc2[q] * (x - prev)), rx);
          rx = lcg_next(rx);
        } else if constexpr (sizeof(TS) == 4) {
          o = x + ac[q] * (xp - x) - bc2[q] * (x - prev);
        } else {
          o = x + cc * (ac[q] * (xp - x) - bc2[q] * (x - prev));
        }
        prev = x;
        d[r] = (RT)o;
      }
    }
  }

  if (k >= T.ko0 && k < T.ko1) {
    if (l0 + S <= N) {
#pragma unroll
      for (int r = 0; r < S; ++r) *lane_ptr(out + soff(useg + (int64_t)r * ustride), kb) = (TS)d[r];
    } else {
#pragma unroll
      for (int r = 0; r < S; ++r)
        if (l0 + r < N) *lane_ptr(out + soff(useg + (int64_t)r * ustride), kb) = (TS)d[r];
    }
  }
}


template <class Cfg, class TS, int MODE, bool SR>
static void line_launch(const DevView& v, int dir, TS* out, const TS* in, double c, double a1, double a2,
                        double a3, int64_t tile0, int64_t ntiles, int slots, hipStream_t st) {
  const size_t sh = std::max(line_lds_doubles<Cfg, TS, MODE>(slots, v.zc_cls) * sizeof(double), (size_t)v.lds_floor);
  hipLaunchKernelGGL((k_line<Cfg, TS, MODE, SR>), dim3((unsigned)ntiles), dim3(Cfg::THREADS), sh, st, v, dir, out,
                     in, c, a1, a2, a3, tile0, ntiles);
}

template <class Cfg, class TS, bool SR>
static hipError_t line_cfg(const DevView& v, int mode, int dir, TS* out, const TS* in, double c, double a1,
                           double a2, double a3, int64_t tile0, int64_t ntiles, int hslots, hipStream_t st) {
  switch (mode) {
    case LINE_EXPLICIT:
      line_launch<Cfg, TS, LINE_EXPLICIT, SR>(v, dir, out, in, c, a1, a2, a3, tile0, ntiles, kSlots, st);
      break;
    case LINE_IMPLICIT:
      line_launch<Cfg, TS, LINE_IMPLICIT, SR>(v, dir, out, in, c, a1, a2, a3, tile0, ntiles, kSlots, st);
      break;
    case LINE_IMPLICIT_EXPLICIT:
      line_launch<Cfg, TS, LINE_IMPLICIT_EXPLICIT, SR>(v, dir, out, in, c, a1, a2, a3, tile0, ntiles, kSlots, st);
      break;
    case LINE_ZY: line_launch<Cfg, TS, LINE_ZY, SR>(v, dir, out, in, c, a1, a2, a3, tile0, ntiles, hslots, st); break;
    case LINE_YZY: line_launch<Cfg, TS, LINE_YZY, SR>(v, dir, out, in, c, a1, a2, a3, tile0, ntiles, hslots, st); break;
    case LINE_IMPLICIT_EXPLICIT_KEEP:
      if constexpr (std::is_same<TS, double>::value) {
        line_launch<Cfg, TS, LINE_IMPLICIT_EXPLICIT_KEEP, SR>(v, dir, out, in, c, a1, a2, a3, tile0, ntiles, kSlots, st);
        break;
      }
      return hipErrorInvalidValue;
    case LINE_YZY_KEEP:
      if constexpr (std::is_same<TS, double>::value) {
        line_launch<Cfg, TS, LINE_YZY_KEEP, SR>(v, dir, out, in, c, a1, a2, a3, tile0, ntiles, hslots, st);
        break;
      }
      return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Tile shapes: waves per tile NW, line elements per lane S, k-window LW.
// LW = 32 puts two line segments in a wave (one per half-wave): a tile needs
// half the waves of an LW = 64 tile of the same line, so two tiles share a CU
// and one computes while the other loads or stores.

// Tile shapes of one k-window family, one per line-length class
// (kLineClassMax): waves per tile NW, line elements per lane S.  LW = 32 puts
// two line segments in a wave (one per half-wave), so a tile needs half the
// waves of an LW = 64 tile of the same line and two tiles share a CU: one
// computes while the other loads or stores.
template <int LW> struct LineFamily;
template <> struct LineFamily<64> {
  using C0 = LineCfg<1, 48, 64>;
  using C1 = LineCfg<2, 32, 64>;
  using C2 = LineCfg<2, 48, 64>;
  using C3 = LineCfg<4, 32, 64>;
  using C4 = LineCfg<4, 40, 64>;
  using C5 = LineCfg<4, 48, 64>;
  using C6 = LineCfg<8, 32, 64>;
  using C7 = LineCfg<8, 40, 64>;
  using C8 = LineCfg<16, 26, 64>;
  using C9 = LineCfg<8, 80, 64>;
};
template <> struct LineFamily<32> {
  using C0 = LineCfg<1, 24, 32>;
  using C1 = LineCfg<1, 32, 32>;
  using C2 = LineCfg<1, 48, 32>;
  using C3 = LineCfg<2, 32, 32>;
  using C4 = LineCfg<2, 40, 32>;
  using C5 = LineCfg<2, 48, 32>;
  using C6 = LineCfg<4, 32, 32>;
  using C7 = LineCfg<4, 40, 32>;
  using C8 = LineCfg<8, 26, 32>;
  using C9 = LineCfg<8, 40, 32>;
};

// FP32-state tiles of the 64-lane family: half the segment length on twice
// the waves of the FP64 shapes up to the 416-element class (an FP32 line
// takes half the registers, so the register budget no longer asks for long
// segments; the FP64 recurrences' dependent chains halve).  The 640-element
// class keeps its FP64 shape (sixteen-wave 40-element tiles measured slower).
struct LineFamily64F32 {
  using C0 = LineCfg<2, 24, 64>;
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
// The stochastic-rounding FP32 kernels' 64-lane family: the 640-element class
// on sixteen waves of 40 elements (the eight-wave 80-element tiles need all
// 256 registers with the rounding's FP64 sums: one tile per CU, half the
// waves).  5x5 array, 601-point lines: implicit + explicit x 1.83 -> 1.76-1.79
// ms, step +1.2%; round to nearest keeps the eight-wave shape (0.5% faster).
struct LineFamily64F32SR : LineFamily64F32 {
  using C9 = LineCfg<16, 40, 64>;
};
// Tile shapes of a k-window family, state type and rounding.
template <int LW, class TS, bool SR = false> struct LineShapes { using F = LineFamily<LW>; };
template <> struct LineShapes<64, float, false> { using F = LineFamily64F32; };
template <> struct LineShapes<64, float, true> { using F = LineFamily64F32SR; };

template <class F>
constexpr bool family_matches_classes() {
  return F::C0::NMAX == kLineClassMax[0] && F::C1::NMAX == kLineClassMax[1] && F::C2::NMAX == kLineClassMax[2] &&
         F::C3::NMAX == kLineClassMax[3] && F::C4::NMAX == kLineClassMax[4] && F::C5::NMAX == kLineClassMax[5] &&
         F::C6::NMAX == kLineClassMax[6] && F::C7::NMAX == kLineClassMax[7] && F::C8::NMAX == kLineClassMax[8] &&
         F::C9::NMAX == kLineClassMax[9];
}
static_assert(kLineClasses == 10, "one tile shape per line-length class");
static_assert(family_matches_classes<LineShapes<64, double>::F>() && family_matches_classes<LineShapes<32, double>::F>() &&
                  family_matches_classes<LineShapes<64, float>::F>() && family_matches_classes<LineShapes<32, float>::F>() &&
                  family_matches_classes<LineShapes<64, float, true>::F>(),
              "tile shapes follow the classes");

// One launch per non-empty line-length class, each with its own tile shape;
// with side streams (ls) the largest launch stays on st and the others run
// concurrently on the side streams between a fork and join events.
template <int LW, class TS, bool SR = false>
hipError_t line_family(const DevView& v, int mode, int dir, void* outp, const void* inp, double c, double a1,
                       double a2, double a3, hipStream_t st, const LineStreams* ls) {
  using F = typename LineShapes<LW, TS, SR>::F;
  TS* out = static_cast<TS*>(outp);
  const TS* in = static_cast<const TS*>(inp);
  const bool halo = mode == LINE_ZY || mode == LINE_YZY || mode == LINE_YZY_KEEP;
  const int f = LW == 32 ? 1 : 0;
  const int64_t* cls = halo ? v.hcls[f] : (dir == 0 ? v.xcls[f] : v.ycls[f]);
  // halo tiles: those with one coefficient set first, launched with one LDS
  // slot (most windows meet a single material zone), then the others; or
  // (halo_whole) each class as one range with its largest set count
  const bool split = halo && !v.halo_whole;
  struct Launch { int q; int64_t t0, n; int hs; };
  Launch L[2 * kLineClasses];
  int nl = 0, big = 0;
  for (int q = 0; q < kLineClasses; ++q)
    for (int r = 0; r < (split ? 2 : 1); ++r) {
      const int64_t t0 = (split && r == 1) ? v.hone[f][q] : cls[q];
      const int64_t t1 = (split && r == 0) ? v.hone[f][q] : cls[q + 1];
      if (t1 == t0) continue;
      L[nl] = Launch{q, t0, t1 - t0, !halo ? kSlots : (split && r == 0 ? 1 : v.hslots[f][q])};
      if (L[nl].n > L[big].n) big = nl;
      ++nl;
    }
  const bool conc = ls && nl > 1;
  hipError_t e = hipSuccess;
  if (conc) {
    if ((e = hipEventRecord(ls->fork, st)) != hipSuccess) return e;
    for (int s = 0; s < kSideStreams && s < nl - 1; ++s)
      if ((e = hipStreamWaitEvent(ls->side[s], ls->fork, 0)) != hipSuccess) return e;
  }
  for (int i = 0, k = 0; i < nl; ++i) {
    const Launch& x = L[i];
    const hipStream_t s = (!conc || i == big) ? st : ls->side[(k++) % kSideStreams];
    switch (x.q) {
      case 0: e = line_cfg<typename F::C0, TS, SR>(v, mode, dir, out, in, c, a1, a2, a3, x.t0, x.n, x.hs, s); break;
      case 1: e = line_cfg<typename F::C1, TS, SR>(v, mode, dir, out, in, c, a1, a2, a3, x.t0, x.n, x.hs, s); break;
      case 2: e = line_cfg<typename F::C2, TS, SR>(v, mode, dir, out, in, c, a1, a2, a3, x.t0, x.n, x.hs, s); break;
      case 3: e = line_cfg<typename F::C3, TS, SR>(v, mode, dir, out, in, c, a1, a2, a3, x.t0, x.n, x.hs, s); break;
      case 4: e = line_cfg<typename F::C4, TS, SR>(v, mode, dir, out, in, c, a1, a2, a3, x.t0, x.n, x.hs, s); break;
      case 5: e = line_cfg<typename F::C5, TS, SR>(v, mode, dir, out, in, c, a1, a2, a3, x.t0, x.n, x.hs, s); break;
      case 6: e = line_cfg<typename F::C6, TS, SR>(v, mode, dir, out, in, c, a1, a2, a3, x.t0, x.n, x.hs, s); break;
      case 7: e = line_cfg<typename F::C7, TS, SR>(v, mode, dir, out, in, c, a1, a2, a3, x.t0, x.n, x.hs, s); break;
      case 8: e = line_cfg<typename F::C8, TS, SR>(v, mode, dir, out, in, c, a1, a2, a3, x.t0, x.n, x.hs, s); break;
      default: e = line_cfg<typename F::C9, TS, SR>(v, mode, dir, out, in, c, a1, a2, a3, x.t0, x.n, x.hs, s); break;
    }
    if (e != hipSuccess) return e;
  }
  if (conc)
    for (int s = 0; s < kSideStreams && s < nl - 1; ++s) {
      if ((e = hipEventRecord(ls->join[s], ls->side[s])) != hipSuccess) return e;
      if ((e = hipStreamWaitEvent(st, ls->join[s], 0)) != hipSuccess) return e;
    }
  return hipSuccess;
}

}  // namespace gw
