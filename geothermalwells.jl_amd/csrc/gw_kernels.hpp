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

// Pipe columns per reduction chunk (one partial sum per level).  Fixed per
// borehole, so the summation order does not depend on the batch.
constexpr int kRedChunk = 16;
// Partial-sum blocks: consecutive chunks with at most kPartItems levels in
// all and at most kPartChunks chunks share a block (a chunk of more levels
// has a block of its own and loops over them).
constexpr int kPartItems = 256, kPartChunks = 16;
constexpr int kAdvGroupMax = 8;  // most pipe columns of an advection group
constexpr int kZcMax = 16;       // most column classes a scenario may have for the fused z-step kernels

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
  int32_t hslots[2][kLineClasses];  // most coefficient sets of a halo tile of the class (LDS slots)
  int64_t hone[2][kLineClasses];    // halo tiles of class q with one set: [hcls[q], hone[q]) (they come first)
  const uint16_t* cls;
  const int32_t* zone;
  const double* zcoef;          // pairs {ea, eb}
  const float* zcoef32;         // the same rounded to float (FP32 states only)
  const int32_t* setid;
  // dt-dependent
  const int64_t* set_off;
  const double* coef;
  // the sets padded in the line kernels' LDS slot layout, per k-window family
  // (index 0: 64-lane, 1: 32-lane): bd, fw, bu, then the explicit
  // coefficients (FP64 states ea, eb; FP32 states c ea, c eb, as float
  // arrays in the 64-lane family), each kLineClassMax[line_class(N)] long
  // with identity rows past N, at the offsets LineTile::poff gives
  const double* pcoef[2];
  const BhDev* bh;
  int32_t nbh;
  const PipeCol* pcol;
  int32_t npcol;
  const Dep* dep;
  const RedChunk* rchunk;
  int32_t nrchunk;
  // the chunks again for the partial-sum kernel, and its blocks: block b
  // sums chunks pblock[b] .. pblock[b+1]-1
  const PartChunk* pchunk;
  const int32_t* pblock;
  int32_t npblock;
  const int64_t* pcol_off;      // per pipe column: its level-0 offset within its scenario's state
  const RedChunk* agroup;       // advection groups (col0, ncol, bh, part; poff unused)
  int32_t nagroup;
  double* partial;
  double* means;
  double* tout;                 // [nbh] outlet of the current sub-step
  double* inlet;                // [nbh] inlet value of the current sub-step
  const double* inlet_tab;      // [nbh] per-well table
  int32_t nx_max, ny_max, nz_max, countz_max;
  int64_t cells_total, cols_total;
  void* aux;                    // LINE_IMPLICIT_EXPLICIT_KEEP: receives the implicit result
  int32_t lds_floor;            // bytes: each line tile requests at least this much LDS
  int32_t halo_whole;           // 1: the fused z-step modes launch each class as one range
  int32_t zc_cls;               // column classes whose z coefficients a halo tile stages (most of any scenario)
  // GW_F32_STATE_SR: the step index that keys the random numbers of the
  // stochastic rounding (set by the host at a call's first step, advanced on
  // the device after each step)
  uint32_t* step_ctr;
};

// The k-window argument of launch_line is 64 or 32, optionally with
// kLwTwoTiles set: every tile then requests kLdsTwoTiles bytes of LDS, so at
// most two share a CU (some operators run faster with fewer concurrent
// tiles); and kLwHaloWhole: the fused z-step modes launch each class's tiles
// as one range with LDS for the class's largest set count, instead of two
// ranges split by set count (one launch less, fewer tiles per CU).  The
// autotune decides per operator.
constexpr int kLwTwoTiles = 1;
constexpr int kLwHaloWhole = 2;
// kLwConcurrent: an operator that launches several line-length classes (a
// ragged batch) puts its largest launch on the context's stream and the others
// on side streams, concurrently, between a fork event and join events (the
// pattern graph capture records as parallel branches), so the short launches
// of small classes overlap the long one instead of queueing behind it.
constexpr int kLwConcurrent = 4;
constexpr int kLwFlags = kLwTwoTiles | kLwHaloWhole | kLwConcurrent;
constexpr int kSideStreams = 3;
struct LineStreams {
  hipStream_t side[kSideStreams];
  hipEvent_t fork;
  hipEvent_t join[kSideStreams];
};
constexpr int kLdsTwoTiles = 80 * 1024;

// LINE_IMPLICIT_EXPLICIT_KEEP / LINE_YZY_KEEP are LINE_IMPLICIT_EXPLICIT /
// LINE_YZY that also write the implicit result to DevView::aux (FP64 only;
// the adjoint's recompute keeps it).  They share the k-window width chosen
// for LINE_IMPLICIT_EXPLICIT / LINE_YZY.
enum LineMode {
  LINE_EXPLICIT = 0,
  LINE_IMPLICIT = 1,
  LINE_IMPLICIT_EXPLICIT = 2,
  LINE_ZY = 3,
  LINE_YZY = 4,
  LINE_IMPLICIT_EXPLICIT_KEEP = 5,
  LINE_YZY_KEEP = 6
};

// ---------------------------------------------------------------------------
// Adjoint (one scenario, FP64).  The 1-D operators are A = D^-1 K with K
// symmetric and D = h * rho_c (h the control length of the direction), so
// with the cell weight W = hx hy hz rho_c every transposed diffusion operator
// is W op W^-1: the adjoint runs the forward line operators on mu = lambda / W
// (reverse order), and only the advection needs a transposed kernel.
// ---------------------------------------------------------------------------
struct AdjGeo {
  const double* hx;      // [nx] control lengths (h_l = (dp + dm) / 2, mirrored ends)
  const double* hy;      // [ny]
  const double* hz;      // [nz]
  const double* ix;      // [nx] 1 / (g[l+1] - g[l]); 0 at the last point
  const double* iy;      // [ny]
  const double* iz;      // [nz]
  const double* rho;     // [ncls][nz] volumetric heat capacity per column class
  const double* hzrho;   // [ncls][nz] hz[k] rho[cls][k] (a pipe cell's weight W over its hx hy)
  const double* rock;    // [ncls][nz] 1 where the conductivity is the rock layer's, else 0
  double* gx_part;       // [ny][nz] conductivity-gradient sums of the x-line kernels (per cell level)
  double* gy_part;       // [yparts][nx][nz] ... of the y-line kernels (y faces and z faces); the
                         // line kernels use part 0, the streaming kernel one part per block of a line
  int32_t yparts;        // blocks per y-line tile of the streaming kernel (adj_yparts)
  int32_t ncls;          // column classes of the scenario
};

// Adjoint line modes (mu in, mu out; gradients added to the part buffers):
//   ADJ_Y    mu <- S_y mu,                 grad c G_y(mu_out, x1)
//   ADJ_X    mu <- S_x E_x mu,             grad c G_x(mu_in, x1) + c G_x(mu_out, x1)
// with G_d(mu, x) = sum over faces of direction d of -(A_f / D_f) (mu2 - mu1)(x2 - x1), half to each
// rock cell of the face.
enum AdjLineMode { ADJ_Y = 0, ADJ_X = 1 };

hipError_t line_adj_w64(const DevView& v, const AdjGeo& ag, int mode, double* out, const double* in,
                        const double* x1, const double* x2, double c, double a1, double a2, double a3,
                        hipStream_t st);
hipError_t line_adj_w32(const DevView& v, const AdjGeo& ag, int mode, double* out, const double* in,
                        const double* x1, const double* x2, double c, double a1, double a2, double a3,
                        hipStream_t st);

// Streaming adjoint kernel (no line held in registers) on the halo tiles of
// the y-lines: out = R(dt A_z) E_y mu (ey) or R(dt A_z) mu, the transposes in
// mu of the forward z-step (from u) followed by the explicit y half-step, with
// the z-face gradient of the three Horner stages and (ey) the explicit-y
// gradient c G_y(mu, R(dt A_z) u); out != mu.  adj_yparts: blocks per tile.
hipError_t launch_adj_yz(const DevView& v, const AdjGeo& ag, bool ey, double* out, const double* mu,
                         const double* u, double c, double dt, double c3, hipStream_t st);
int adj_yparts(int64_t htiles, int ny);

// Transposed advection of one sub-step, in mu: per borehole, gather lists of
// the departure table (CSR over the destination level), the turnaround
// transpose (CSR over annulus levels), chunked reductions over the pipe
// columns.  seed (may be NULL): per-borehole objective weight added to the
// inner-pipe outlet levels (the sub-step whose outlet enters the objective).
struct AdjEnt {
  int32_t k, pad;
  double w;
};
struct AdjBh {
  int32_t csr_in, csr_an;   // offsets into adj_off of the inner / annulus level CSR (countz + 1 entries)
  int32_t turn_off;         // offset into turn_off_tab (nlev + 1 entries) of the turnaround transpose
  int32_t t0, nt;           // inner levels reduced for the turnaround (t0 .. t0+nt-1)
  int32_t i0, ni;           // annulus levels reduced for the inlet (i0 .. i0+ni-1)
  int32_t red0, nred_in, nred_an, pad;  // chunks: red0 .. red0+nred_in-1 inner, then nred_an annulus
};
// One pipe column for the adjoint kernels: its first cell, column class and
// hx * hy (W at level k = hxy hz[k] rho[cls][k]).
struct AdjCol {
  int64_t off;
  double hxy;
  int32_t cls, pad;
};
struct AdjPipe {
  const AdjBh* bh;
  const AdjCol* col;        // per pipe column (the pipe column list's order)
  const int32_t* off;       // CSR offsets
  const AdjEnt* ent;        // CSR entries (source level, weight)
  const int32_t* toff;      // turnaround CSR offsets (per annulus level)
  const AdjEnt* tent;       // turnaround entries (inner level, weight)
  const RedChunk* chunk;    // reduction chunks (part 0: inner/turn levels, 1: annulus/inlet levels)
  int32_t nchunk;
  double* partial;          // chunk partial sums (inner chunks, per turnaround level)
  double* gacc;             // [nchunk] annulus chunks: accumulated d objective / d inlet temperature
};
hipError_t launch_adj_pipe(const DevView& v, const AdjGeo& ag, const AdjPipe& ap, double* mu, const double* seed,
                           hipStream_t st);
// lam = W mu over the whole field (the state adjoint from mu)
hipError_t launch_mul_w(const DevView& v, const AdjGeo& ag, double* lam, const double* mu, hipStream_t st);

// Default k-window width (64 or 32 lanes) for lines of up to nmax elements.
int line_window_default(int nmax);
// The line kernels of one k-window family and state precision (a1..a3: the
// z-step's Horner coefficients), one launch per non-empty line-length class.
hipError_t line_w64_f64(const DevView& v, int mode, int dir, void* out, const void* in, double c, double a1,
                        double a2, double a3, hipStream_t st, const LineStreams* ls);
hipError_t line_w64_f32(const DevView& v, int mode, int dir, void* out, const void* in, double c, double a1,
                        double a2, double a3, hipStream_t st, const LineStreams* ls);
hipError_t line_w32_f64(const DevView& v, int mode, int dir, void* out, const void* in, double c, double a1,
                        double a2, double a3, hipStream_t st, const LineStreams* ls);
hipError_t line_w32_f32(const DevView& v, int mode, int dir, void* out, const void* in, double c, double a1,
                        double a2, double a3, hipStream_t st, const LineStreams* ls);
// FP32 state with stochastic rounding of every state update (GW_F32_STATE_SR)
hipError_t line_w64_f32sr(const DevView& v, int mode, int dir, void* out, const void* in, double c, double a1,
                          double a2, double a3, hipStream_t st, const LineStreams* ls);
hipError_t line_w32_f32sr(const DevView& v, int mode, int dir, void* out, const void* in, double c, double a1,
                          double a2, double a3, hipStream_t st, const LineStreams* ls);
// *ctr = value (set) or *ctr += 1, by a one-thread kernel (graph-capturable)
hipError_t launch_step_ctr(uint32_t* ctr, bool set, uint32_t value, hipStream_t st);

hipError_t launch_zrhs(const DevView& v, int prec, void* dphi, const void* phi, hipStream_t st);
hipError_t launch_zstep(const DevView& v, int prec, void* phi, double dt, double c3, hipStream_t st);
// c: sub-step of the line operators; dt_z, c3: the fused z-step (modes ZY, YZY);
// lw: k-window family (64 or 32).
hipError_t launch_line(const DevView& v, int prec, int mode, int dir, int lw, void* out, const void* in,
                       double c, double dt_z, double c3, hipStream_t st, const LineStreams* ls = nullptr);
// Outlet sums, annulus level means and inlet values of one sub-step.
// jacc (may be NULL): per-borehole outlet accumulator, += T_outlet of this sub-step.
hipError_t launch_pipe_reduce(const DevView& v, int prec, const void* phi, hipStream_t st, double* jacc = nullptr);
// Semi-Lagrangian update of every pipe column, in place.
hipError_t launch_pipe_advect(const DevView& v, int prec, void* phi, hipStream_t st);

// Post-processing (src/util.jl:274-356).  Per-level means of phi over two
// lists of column offsets (off[0 .. n_in) then off[n_in .. n_in + n_an)), for
// levels 0 .. nlev-1: means[0 .. nlev) and means[nlev .. 2 nlev).
hipError_t launch_level_means(int prec, const void* phi, const int64_t* off, int n_in, int n_an, int nlev,
                              double* means, hipStream_t st);
// out[i] = phi(k_lo, j, i) (k_lo == k_hi) or w_lo phi(k_lo, j, i) + w_hi phi(k_hi, j, i), i < nx.
hipError_t launch_x_profile(int prec, const void* phi, int nx, int ny, int nz, int j, int k_lo, int k_hi,
                            double w_lo, double w_hi, double* out, hipStream_t st);

}  // namespace gw
