// gw_kernels.hpp -- host-callable launchers of the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gw_layout.hpp"

namespace gw {

// Device view of one context's tables, passed by value to every launcher.
struct DevView {
  const ScenDev* sc;
  int32_t nscen;
  const int64_t* cell_prefix;   // [nscen+1] cells
  const int64_t* col_prefix;    // [nscen+1] (i, j) columns
  // line tiles of each k-window family (index 0: 64-lane windows, 1: 32-lane)
  const LineTile* xtiles[2];    // x-line tiles
  const LineTile* ytiles[2];    // y-line tiles
  const LineTile* htiles[2];    // y-line windows with a 3-lane k halo (fused z-step)
  const uint16_t* cls;
  const int32_t* zone;
  const double* zcoef;          // pairs {ea, eb}
  const int32_t* setid;
  // dt-dependent
  const int64_t* set_off;
  const double* coef;
  const BhDev* bh;
  int32_t nbh;
  const PipeCol* pcol;
  int32_t npcol;
  const Dep* dep;
  const RedChunk* rchunk;
  int32_t nrchunk;
  double* partial;
  double* means;
  double* tout;                 // [nbh] outlet of the current sub-step
  double* inlet;                // [nbh] inlet value of the current sub-step
  const double* inlet_tab;      // [nbh] per-well table
  int32_t nx_max, ny_max, nz_max, countz_max;
  int64_t cells_total, cols_total;
  int64_t xtiles_total[2], ytiles_total[2], htiles_total[2];
};

enum LineMode { LINE_EXPLICIT = 0, LINE_IMPLICIT = 1, LINE_IMPLICIT_EXPLICIT = 2, LINE_ZY = 3, LINE_YZY = 4 };

// Whether the line kernels of k-window width lw (64 or 32 lanes) take lines
// of nmax elements, and the default width for that length.
bool line_window_ok(int nmax, int lw);
int line_window_default(int nmax);

hipError_t launch_zrhs(const DevView& v, int prec, void* dphi, const void* phi, hipStream_t st);
hipError_t launch_zstep(const DevView& v, int prec, void* phi, double dt, double c3, hipStream_t st);
// c: sub-step of the line operators; dt_z, c3: the fused z-step (modes ZY, YZY);
// lw: k-window family (64 or 32).
hipError_t launch_line(const DevView& v, int prec, int mode, int dir, int lw, void* out, const void* in,
                       double c, double dt_z, double c3, hipStream_t st);
// Outlet sums, annulus level means and inlet values of one sub-step.
hipError_t launch_pipe_reduce(const DevView& v, int prec, const void* phi, hipStream_t st);
// Semi-Lagrangian update of every pipe column, in place.
hipError_t launch_pipe_advect(const DevView& v, int prec, void* phi, hipStream_t st);

}  // namespace gw
