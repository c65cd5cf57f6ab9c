// gw_layout.hpp -- device-side problem layout shared by the host set-up code
// (gw_host.cpp) and the gfx950 kernels (gw_kernels.hip).
//
// HBM layout of one scenario's temperature field: the reference's
// column-major phi[k, j, i] of size (Nz, Ny, Nx) (src/physics.jl:336), i.e.
// element offset k + Nz*(j + Ny*i); z (k) is contiguous.  A batch packs its
// scenarios back to back; ScenDev::off is the first element of each.
//
// Materials are not looked up point-wise.  Every (i, j) column gets a class id
// (columns with the same k(z), rho*c(z) profile share it) and every level k a
// zone id (levels whose material values agree in every class share it).  The
// 1-D operators along x and y then depend only on (line, zone); their
// coefficients are stored once per distinct (line, zone) as a "line set".
#pragma once
#include <stdint.h>

namespace gw {

struct ScenDev {
  int32_t nx, ny, nz, countz;
  int32_t nzone, ncls, nbh, bh0;   // bh0: first global borehole index
  int64_t off;                     // first state element of this scenario
  int64_t o_gz;                    // gridz offset in the double pool
  int64_t o_cls;                   // class map (uint16, index j + ny*i)
  int64_t o_zone;                  // zone map (int32, index k)
  int64_t o_zcoef;                 // z coefficients (double2 {ea, eb}, index cls*nz + k)
  int64_t o_xset;                  // x-line set ids (int32, index j*nzone + zone)
  int64_t o_yset;                  // y-line set ids (int32, index i*nzone + zone)
};

// Per borehole (global index over the batch).
struct BhDev {
  int32_t scen;
  int32_t col_in0, n_in;    // inner-pipe columns [col_in0, col_in0 + n_in) in the pipe list
  int32_t col_an0, n_an;    // annulus columns
  int32_t n_outlev;         // outlet levels: k < n_outlev  <=>  gridz[k] <= v_inner*c
  int32_t klo, nlev;        // annulus level means needed for the turnaround: [klo, klo+nlev)
  int32_t mean_off;         // offset of those means in the mean buffer
  int32_t dep_off;          // departure table offset (entries), 2 entries per level
  int32_t inlet_kind;       // GW_INLET_*
  int32_t red_in0, n_red_in;  // reduction chunks of the inner columns
  int32_t red_an0, n_red_an;  // reduction chunks of the annulus columns
  int32_t pad;
  double inlet_value;       // ConstantInlet value or HeatExchangerInlet dT
};

// A reduction chunk: pipe columns [col0, col0+ncol) of borehole `bh`; part 0
// sums inner-pipe columns over the outlet levels k < n_outlev, part 1 annulus
// columns over the turnaround levels [klo, klo+nlev).  Per-level partial sums
// go to partial[poff + level].
struct RedChunk {
  int32_t bh, part, col0, ncol, poff, pad;
};

// A reduction chunk as the partial-sum kernel reads it: the state offset of
// its first summed level (scenario offset + the part's first level), its
// columns [col0, col0+ncol) of the pipe list, its nl levels, where its sums
// go (partial[poff + level]) and its first item in its block (a block packs
// consecutive chunks, one thread per chunk level).
struct PartChunk {
  int64_t base;
  int32_t col0, ncol, nl, poff, item0, pad;
};

// One pipe column: (i, j) of a borehole; inner-pipe columns of a borehole come
// first, then its annulus columns (src/cache.jl:17-30 order).
struct PipeCol {
  int32_t i, j, bh, inner;
};

// Semi-Lagrangian departure of one (borehole, level, pipe) entry, computed on
// the host with the reference's exact branch arithmetic (src/solver.jl:299-357).
enum DepType : int32_t { DEP_KEEP = 0, DEP_OWN = 1, DEP_TURN = 2, DEP_INLET = 3 };
struct Dep {
  int32_t type, kl, kr, pad;
  double a;
};

// A line tile: the 64 lanes k = kw .. kw+63 (clamped to the grid) of x-line
// j (or y-line i) = `line` in scenario `scen`; lanes with ko0 <= k < ko1 are
// written.  Its lanes use the nset coefficient sets in `set` (at most 3 in
// plain tiles, 4 in the z-halo tiles of the fused z-step; unused entries -1).
// Lane kk's set is set[slot], slot = bit kk of slot_b0 + 2 * bit kk of
// slot_b1 (computed on the host from the zone and set-id maps).  N, stride
// and nz are the line's length, its element stride in the state and the
// scenario's nz; base is the state offset of (element 0, level max(kw, 0));
// poff[q] the offset of set[q]'s padded copy in DevView::pcoef: a tile
// addresses its line and coefficients from its descriptor alone.
struct LineTile {
  int32_t scen, line, kw, ko0, ko1, nset;
  int32_t set[4];
  uint64_t slot_b0, slot_b1;
  int32_t N, stride, nz, pad;
  int64_t base;
  uint32_t poff[4];
};

// Coefficient set of one 1-D line (length N) for one half-step length c,
// stored as 5 consecutive arrays:
//   bd[N]  1 / pivot of (I - cA) = LU (Thomas without pivoting)
//   ea[N]  explicit coefficient of (u_{l+1} - u_l)
//   eb[N]  explicit coefficient of (u_l - u_{l-1})
//   fw[N]  forward multiplier  (-c eb_l) bd_{l-1}   (y_l = b_l - fw_l y_{l-1})
//   bu[N]  backward multiplier (-c ea_l) bd_l       (x_l = bd_l y_l - bu_l x_{l+1})
// with (A u)_l = ea_l (u_{l+1} - u_l) - eb_l (u_l - u_{l-1}); fw_0 uses bd_0.
enum { SET_BD = 0, SET_EA = 1, SET_EB = 2, SET_FW = 3, SET_BU = 4, SET_NARR = 5 };

}  // namespace gw
