/* geowells.h -- C ABI of libgeowells_hip, the gfx950 (MI355X) implementation of
 * GeothermalWells.jl's per-timestep hot path.
 *
 * The Julia package keeps its API (create_cache, rhs_diffusion_z!, the
 * ODEProblem/ROCK2 driver, get_simulation_callback); each entry point below
 * replaces the reference function named in its comment (paths relative to the
 * GeothermalWells.jl v0.1.3 source tree).  Plain pointers and sizes only.
 *
 * Conventions
 *  - Return value: GW_OK (0) or a negative gw_status; gw_last_error() gives a
 *    message for the calling thread.  Nothing aborts or throws across the ABI.
 *  - State buffers are device pointers owned by the caller, valid only during
 *    the call.  A field is the reference's column-major phi[k, j, i] of size
 *    (Nz, Ny, Nx), element type double (GW_F64) or float (GW_F32_STATE*).  A
 *    batch packs its scenarios back to back (gw_scenario_offset).
 *  - A context is bound to one HIP stream and is not re-entrant.  Calls are
 *    stream-ordered and asynchronous except gw_outlet, gw_create and a
 *    sub-step whose inlet is evaluated on the host (GW_INLET_HOST).
 */
#ifndef GEOWELLS_H
#define GEOWELLS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum gw_status {
  GW_OK = 0,
  GW_EINVAL = -1,  /* invalid argument or a configuration the reference turns into NaN */
  GW_EHIP = -2,    /* HIP runtime error */
  GW_ENOMEM = -3,  /* device allocation failed */
  GW_ESTATE = -4   /* call not valid in the context's current state */
} gw_status;

/* State precision.  GW_F32_STATE: FP32 field, stencils and advection in FP32,
 * the Thomas recurrences and the outlet reductions in FP64.  An implicit solve
 * keeps its line in FP32 registers between its two passes: each element is
 * rounded to nearest twice (the forward pass's result bd_l y_l, then the
 * solution); the recurrences themselves, their carries and the coefficients
 * are FP64.  GW_F32_STATE_SR: the same FP32 field, with every
 * rounding of an updated value (the implicit solves' stores, the z-step's and
 * the explicit steps' results, computed as FP32 increments added in FP64)
 * stochastic, so sub-ulp changes are kept in expectation instead of lost: no
 * counterpart in the reference (which is FP64); for runs long enough that
 * round-to-nearest stagnation biases the FP32 field.  Its random numbers are
 * keyed by the step index t / dt, the element and the rounding point: results
 * are reproducible, and a run split into gw_steps calls equals one call. */
enum { GW_F64 = 0, GW_F32_STATE = 1, GW_F32_STATE_SR = 2 };

/* gw_problem.flags.  The line kernels come in two k-window widths (64 and 32
 * lanes) and launch variants; by default the first gw_steps call times them
 * for each line operator of the step and keeps the fastest.  GW_FLAG_LINE_W64
 * / _W32 force one width; with one of them, GW_FLAG_LINE_TWO_TILES pads each
 * tile's LDS so at most two share a CU, and GW_FLAG_HALO_WHOLE launches the
 * fused z-step tiles of a line class as one range (default: split by their
 * coefficient-set count); GW_FLAG_LINE_CONCURRENT runs the launches of the
 * line-length classes of a batch concurrently on side streams.  gw_steps runs
 * the steps of a call by replaying one captured hipGraph per step or by
 * launching the same kernels directly; the first call of 9 or more steps times
 * both and keeps the faster, GW_FLAG_STEPS_GRAPH / _DIRECT force one.  In a
 * ragged batch, a line-length class holding under 5% of the batch's line
 * cells (in x and in y) runs with the next larger class's tile shape, one
 * launch less per operator; GW_FLAG_KEEP_LINE_CLASSES gives every class its
 * own launch.  Results do not depend on these choices beyond rounding: a tile
 * shape with other segment lengths splits the segmented recurrences of a line
 * elsewhere. */
enum { GW_FLAG_LINE_W64 = 1, GW_FLAG_LINE_W32 = 2, GW_FLAG_LINE_TWO_TILES = 4, GW_FLAG_HALO_WHOLE = 8,
       GW_FLAG_LINE_CONCURRENT = 16, GW_FLAG_STEPS_GRAPH = 32, GW_FLAG_STEPS_DIRECT = 64,
       GW_FLAG_KEEP_LINE_CLASSES = 128 };

/* Inlet models (src/physics.jl:283-318). */
enum {
  GW_INLET_CONSTANT = 0,       /* ConstantInlet(T):        T                       */
  GW_INLET_HEAT_EXCHANGER = 1, /* HeatExchangerInlet(dT):  T_outlet[bh] - dT       */
  GW_INLET_PER_WELL = 2,       /* CustomInlet((bh,_,_) -> T_in[bh]) as a table   */
  GW_INLET_HOST = 3            /* CustomInlet(f): f evaluated on the host per sub-step */
};

/* Borehole{T}(xc, yc, h, r_inner, t_inner, r_outer, t_outer, r_backfill, mdot,
 * insulation_depth) constructor arguments (src/physics.jl:48). */
typedef struct gw_borehole {
  double xc, yc, h, r_inner, t_inner, r_outer, t_outer, r_backfill, mdot, insulation_depth;
} gw_borehole;

/* HomogenousMaterialProperties (src/physics.jl:159-170) when n_layers == 0,
 * else StratifiedMaterialProperties{n_layers} (src/physics.jl:81-94). */
typedef struct gw_materials {
  int32_t n_layers;
  const double *k_rock_layers, *rho_c_rock_layers, *layer_depths;
  double k_rock, rho_c_rock;
  double k_water, rho_c_water, k_steel, rho_c_steel;
  double k_insulating, rho_c_insulating, k_backfill, rho_c_backfill;
} gw_materials;

/* Host inlet callback: (1-based borehole index, outlet temperatures of this
 * scenario's boreholes, their count, time) -> inlet temperature. */
typedef double (*gw_inlet_fn)(int32_t bh, const double *T_outlet, int32_t n_bh, double t,
                              void *user);

typedef struct gw_inlet {
  int32_t kind;            /* GW_INLET_* */
  double value;            /* CONSTANT: T; HEAT_EXCHANGER: dT */
  const double *per_well;  /* PER_WELL: n_bh values */
  gw_inlet_fn fn;          /* HOST */
  void *user;
} gw_inlet;

/* One scenario: the arguments of create_cache (src/cache.jl:55). */
typedef struct gw_scenario {
  int32_t nx, ny, nz;
  const double *gridx, *gridy, *gridz;
  int32_t n_bh;
  const gw_borehole *boreholes;
  gw_materials materials;
  gw_inlet inlet;
} gw_scenario;

typedef struct gw_problem {
  int32_t n_scen;
  const gw_scenario *scen;
  int32_t precision;   /* GW_F64, GW_F32_STATE or GW_F32_STATE_SR */
  void *stream;        /* hipStream_t; NULL: the context creates its own */
  double rock2_c3;     /* cubic coefficient of the minimum-degree ROCK2 polynomial; <= 0: default 1/16 */
  int32_t flags;       /* GW_FLAG_*, or 0 */
} gw_problem;

typedef struct gw_ctx gw_ctx;

/* create_cache (src/cache.jl:55-116) and create_advection_index_lists
 * (src/cache.jl:9-44): classification, index lists, coefficient tables and
 * eigen_estimator_pre_calculation (src/solver.jl:474-488).  Synchronous. */
int gw_create(const gw_problem *p, gw_ctx **out);
int gw_destroy(gw_ctx *ctx);
const char *gw_last_error(void);

/* rhs_diffusion_z!(dphi, phi, cache, t) (src/solver.jl:23-31): dphi = A_z phi. */
int gw_rhs_z(gw_ctx *ctx, void *dphi, const void *phi, double t);

/* ADI_and_ADV_callback!(integrator) (src/solver.jl:432-464) with
 * t = integrator.t (the step's end) and dt = t - integrator.tprev.  Runs in
 * place on phi; `scratch` stands for integrator.uprev, which the reference
 * overwrites: it may be NULL and is not written. */
int gw_adi_adv(gw_ctx *ctx, void *phi, void *scratch, double t, double dt);

/* gw_adi_adv without any device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same kernels and results
 * as gw_adi_adv; the pipe reductions' partial sums and turnaround level means
 * live in a fixed 8 MiB device array per GPU (held until the process ends)
 * instead of the context's tables.  Needs the context's tables for this dt,
 * built by one earlier allocating call with the same dt (gw_adi_adv, gw_step,
 * gw_steps): GW_ESTATE otherwise.  GW_EINVAL when the reductions need more
 * than 2^20 doubles of scratch, or for a GW_INLET_HOST inlet (its host round
 * trip synchronises).  Every call on one GPU shares that array: calls that
 * overlap on one GPU (two streams, two threads) race. */
int gw_adi_adv_noalloc(gw_ctx *ctx, void *phi, double t, double dt);

/* The ROCK2 z-step of one step at its minimum degree, fused:
 * phi = R(dt*A_z) phi with R(z) = 1 + z + z^2/2 + c3 z^3 (the polynomial the
 * examples' ROCK2(eigen_est=eigen_estimator) applies, src/solver.jl:23-31,
 * 495).  GW_EINVAL when dt times the Gershgorin bound of A_z leaves the
 * polynomial's real stability interval. */
int gw_zstep(gw_ctx *ctx, void *phi, double dt);

/* One full step from t to t+dt: gw_zstep then gw_adi_adv(t+dt). */
int gw_step(gw_ctx *ctx, void *phi, double t, double dt);

/* nsteps full steps from t as the fused step pipeline (two full-field sweeps
 * per step: implicit+explicit x, and implicit y fused with the next step's
 * z-step and explicit y), with the arithmetic of gw_zstep + gw_adi_adv.
 * Steps 2 .. nsteps-1 replay one captured hipGraph per step or launch the
 * same kernels directly (graphs need inlet models other than GW_INLET_HOST);
 * see GW_FLAG_STEPS_GRAPH.  The first call times the line kernels' two
 * k-window widths (unless gw_problem.flags forces one); phi is only read by
 * that.  The result is in phi. */
int gw_steps(gw_ctx *ctx, void *phi, double t, double dt, int32_t nsteps);

/* cache.T_outlet after the last advection sub-step (src/solver.jl:253-257):
 * all boreholes of all scenarios in order.  Synchronises the stream. */
int gw_outlet(gw_ctx *ctx, double *T_outlet_host);

/* Replace the per-well inlet table (GW_INLET_PER_WELL scenarios only),
 * values for every borehole of the batch in order. */
int gw_set_inlet(gw_ctx *ctx, const double *T_in);

/* Single operators, for callers that sequence the step themselves:
 * explicit (I + c A_dir) (diffusion_1D! with plus_I, src/solver.jl:34-58),
 * implicit (I - c A_dir)^-1 (thomas_I_minus_A!, src/solver.jl:111-232),
 * dir 0 = x, 1 = y; in == out is allowed.  gw_advect is advection!
 * (src/solver.jl:250-264) of one sub-step of length c at time t, in place. */
int gw_explicit(gw_ctx *ctx, void *out, const void *in, double c, int32_t dir);
int gw_implicit(gw_ctx *ctx, void *out, const void *in, double c, int32_t dir);
int gw_advect(gw_ctx *ctx, void *phi, double c, double t);

/* One line operator of the fused step pipeline, for callers that sequence
 * steps themselves and for testing: mode 0 explicit (I + cA_dir), 1 implicit
 * (I - cA_dir)^-1, 2 implicit then explicit, 3 z-step R(dt_z A_z) then
 * explicit y, 4 implicit y then z-step then explicit y (the reference's
 * src/solver.jl:461, then :23-31 under ROCK2 and :444 of the next step).
 * Modes 3 and 4 are y-direction only and need out != in. */
int gw_line_op(gw_ctx *ctx, void *out, const void *in, double c, double dt_z, int32_t dir, int32_t mode);

/* Gradient of the outlet objective
 *   J = (1/nsteps) sum_n sum_b w_b T_out_b(n),
 * T_out_b(n) borehole b's outlet in step n's second advection sub-step
 * (src/solver.jl:253-257, 458), with respect to the rock conductivity of every
 * layer (k_rock when homogeneous: StratifiedMaterialProperties /
 * HomogenousMaterialProperties, src/physics.jl:81-94, 159-170) and every
 * borehole's inlet temperature, from phi0 (read only) over nsteps steps of dt
 * from t0 -- the derivative a caller of the reference obtains by
 * differentiating solve(...) through its parameters.  Reverse sweeps over
 * segments of seg steps (one checkpoint each, recomputed).  One scenario,
 * GW_F64, constant or per-well inlets.  w_bh NULL: w_b = 1/n_bh.  dJ_dk has
 * n_dk = max(1, n_layers) entries; dJ_dTin has n_bh (for a constant inlet
 * dJ/dT is their sum).  dJ_du0 (device, may be NULL) receives dJ/dphi0.
 * Allocates its working fields; synchronous. */
int gw_adjoint(gw_ctx *ctx, const void *phi0, double t0, double dt, int32_t nsteps, int32_t seg,
               const double *w_bh, double *J, double *dJ_dk, int32_t n_dk, double *dJ_dTin, void *dJ_du0);

/* The pieces of gw_adjoint's reverse step, for callers that compose their
 * own adjoint (and for testing), in mu = lambda / W with W = hx hy hz rho_c
 * per cell (h the control lengths of the grid): op 0 transposed advection
 * sub-step of length dt/2; 1 transposed implicit y (x: the forward result,
 * for the conductivity gradient, discarded here); 2 transposed implicit +
 * explicit x (x: the forward implicit-x result); 3 transposed explicit y;
 * 4 transposed z-step R(dt A_z) (x: the forward field it acted on);
 * 5 lambda = W mu; 6 transposed z-step followed by explicit y (the pair
 * gw_line_op mode 3 applies; x: the field the z-step acted on).  Ops 4 and 6
 * need out distinct from in and x.  One FP64 scenario; synchronous; no
 * counterpart in the reference (it has no adjoint). */
int gw_adjoint_op(gw_ctx *ctx, int32_t op, void *out, const void *in, const void *x, double dt);

/* Grid generators (src/grids.jl), host only: no context, no GPU needed.
 * The two grid functions write at most `cap` values to `out` (which may be
 * NULL) and return the grid's length, or a negative gw_status; Julia's float
 * ranges are reproduced exactly (decimal endpoints as exact rationals, each
 * element rounded once).
 * compute_domain (src/grids.jl:8-18): out[6] = xmin, xmax, ymin, ymax, zmin, zmax. */
int gw_compute_domain(const gw_borehole *bh, int32_t n_bh, double buffer_x, double buffer_y, double buffer_z,
                      double *out);
/* create_uniform_gridz_with_borehole_depths (src/grids.jl:29-48): zmin:dz:zmax
 * plus every borehole depth, sorted, duplicates removed. */
int64_t gw_uniform_gridz(double zmin, double zmax, double dz, const gw_borehole *bh, int32_t n_bh, double *out,
                         int64_t cap);
/* create_adaptive_grid_1d (src/grids.jl:80-147), direction 0 = x, 1 = y: fine
 * uniform spacing across each borehole's backfill diameter, geometric growth
 * (capped at dx_max) towards the domain ends and between boreholes. */
int64_t gw_adaptive_grid_1d(double xmin, double xmax, double dx_fine, double growth_factor, double dx_max,
                            const gw_borehole *bh, int32_t n_bh, int32_t direction, double *out, int64_t cap);

/* get_temperatures_along_z_single_well(phi, cache) (src/util.jl:330-356) on
 * the resident field: for the levels with gridz < h of the scenario's first
 * borehole (their count in *nlev), the mean over the Idx_list_Inner columns
 * (T_inner) and over the Idx_list_Outer columns (T_outer), FP64, host arrays
 * of at least that count (cap).  T_inner = T_outer = NULL only reports *nlev.
 * Synchronous. */
int gw_pipe_level_means(gw_ctx *ctx, const void *phi, int32_t scen, double *T_inner, double *T_outer, int32_t cap,
                        int32_t *nlev);
/* extract_x_profile(T, gridx, gridy, gridz, z_target, xc, yc, full_profile=true)
 * (src/util.jl:274-318) of a device field T (nx, ny, nz as everywhere here):
 * the x-line at the y index closest to yc, linearly interpolated between the
 * levels bracketing z_target (the first or last level outside the grid),
 * into the host array profile[nx].  No context; runs on `stream` (NULL: the
 * null stream), synchronous. */
int gw_extract_x_profile(const void *T, int32_t precision, int32_t nx, int32_t ny, int32_t nz, const double *gridy,
                         const double *gridz, double z_target, double yc, double *profile, void *stream);

/* Introspection (cache fields the reference exposes). */
int64_t gw_state_elems(gw_ctx *ctx);
int64_t gw_scenario_offset(gw_ctx *ctx, int32_t scen);
int gw_counts(gw_ctx *ctx, int32_t scen, int32_t *countz, int32_t *n_inner, int32_t *n_outer);
/* Idx_list_Inner / Idx_list_Outer as 0-based (i, j, bh) triples. */
int gw_index_lists(gw_ctx *ctx, int32_t scen, int32_t *inner, int32_t *outer);
double gw_eigen_estimate(gw_ctx *ctx, int32_t scen);
int gw_sync(gw_ctx *ctx);

/* The hipStream_t the context launches on (for callers that record events). */
void *gw_stream(gw_ctx *ctx);

/* How gw_steps runs a call's steps: 0 not decided yet (graph replays until a
 * call of 9 or more steps has timed both), 1 graph replays, 2 direct launches;
 * < 0 on error. */
int gw_steps_mode(gw_ctx *ctx);

/* The launch form the context uses for line operator `op` (0: z-step +
 * explicit y, 1: implicit + explicit x, 2: implicit y + z-step + explicit y,
 * 3: implicit y): the k-window width (64 or 32) plus the
 * GW_FLAG_LINE_TWO_TILES, GW_FLAG_HALO_WHOLE and GW_FLAG_LINE_CONCURRENT bits,
 * as forced or as the first gw_steps call chose them; < 0 on error.  No
 * counterpart in the reference (which has one launch form). */
int gw_line_launch(gw_ctx *ctx, int32_t op);

/* Measurement: nsteps full steps (no graph) with HIP events between the
 * kernels of the fused step pipeline, on the context's stream.  ms[] adds up:
 * [0] z-step + explicit y of the first step (once), [1] advection sub-step 1,
 * [2] implicit x + explicit x, [3] advection sub-step 2, [4] implicit y fused
 * with the next step's z-step and explicit y (nsteps-1 launches), [5] the
 * last step's implicit y (once).  Synchronises. */
int gw_profile_steps(gw_ctx *ctx, void *phi, double t, double dt, int32_t nsteps, double *ms);

#ifdef __cplusplus
}
#endif
#endif /* GEOWELLS_H */
