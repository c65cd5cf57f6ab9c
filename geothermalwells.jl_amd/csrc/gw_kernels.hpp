// gw_kernels.hpp -- host-callable launchers of the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gw_layout.hpp"

namespace gw {

// Device view of one context's tables, passed by value to every launcher.
// Line-length classes of the line kernels: class q takes lines of at most
// kLineClassMax[q] elements (the same ladder for both k-window widths).
constexpr int kLineClasses = 10;
constexpr int kLineClassMax[kLineClasses] = {48, 64, 96, 128, 160, 192, 256, 320, 416, 640};
inline int line_class(int n) {
  for (int q = 0; q < kLineClasses; ++q)
    if (n <= kLineClassMax[q]) return q;
  return -1;
}

struct DevView {
  const ScenDev* sc;
  int32_t nscen;
  const int64_t* cell_prefix;   // [nscen+1] cells
  const int64_t* col_prefix;    // [nscen+1] (i, j) columns
  // line tiles of each k-window family (index 0: 64-lane windows, 1: 32-lane),
  // ordered by line-length class (kLineClassMax); *cls[f][q] .. *cls[f][q+1]
  // are the tiles of class q
  const LineTile* xtiles[2];    // x-line tiles
  const LineTile* ytiles[2];    // y-line tiles
  const LineTile* htiles[2];    // y-line windows with a 3-lane k halo (fused z-step)
  int64_t xcls[2][kLineClasses + 1], ycls[2][kLineClasses + 1], hcls[2][kLineClasses + 1];
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
  const RedChunk* agroup;       // advection groups (col0, ncol, bh, part; poff unused)
  int32_t nagroup;
  double* partial;
  double* means;
  double* tout;                 // [nbh] outlet of the current sub-step
  double* inlet;                // [nbh] inlet value of the current sub-step
  const double* inlet_tab;      // [nbh] per-well table
  int32_t nx_max, ny_max, nz_max, countz_max;
  int64_t cells_total, cols_total;
};

enum LineMode { LINE_EXPLICIT = 0, LINE_IMPLICIT = 1, LINE_IMPLICIT_EXPLICIT = 2, LINE_ZY = 3, LINE_YZY = 4 };

// Default k-window width (64 or 32 lanes) for lines of up to nmax elements.
int line_window_default(int nmax);
// The line kernels of one k-window family and state precision (a1..a3: the
// z-step's Horner coefficients), one launch per non-empty line-length class.
hipError_t line_w64_f64(const DevView& v, int mode, int dir, void* out, const void* in, double c, double a1,
                        double a2, double a3, hipStream_t st);
hipError_t line_w64_f32(const DevView& v, int mode, int dir, void* out, const void* in, double c, double a1,
                        double a2, double a3, hipStream_t st);
hipError_t line_w32_f64(const DevView& v, int mode, int dir, void* out, const void* in, double c, double a1,
                        double a2, double a3, hipStream_t st);
hipError_t line_w32_f32(const DevView& v, int mode, int dir, void* out, const void* in, double c, double a1,
                        double a2, double a3, hipStream_t st);

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
