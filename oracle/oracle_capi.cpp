// oracle_capi.cpp -- TEST INFRASTRUCTURE ONLY: flat C entry points over the
// CPU restatement in gw_oracle.hpp, loaded by the test-suite (ctypes) and by
// bench.py's CPU-baseline leg.  Nothing in the product links this library.
#include <omp.h>

#include <chrono>
#include <cstring>
#include <memory>
#include <string>

#include "gw_oracle.hpp"

using gwo::Borehole;
using gwo::Problem;

extern "C" {

// Raw constructor arguments, mirroring the reference's constructors:
// bh[10*b + 0..9] = Borehole{T}(xc, yc, h, r_inner, t_inner, r_outer, t_outer,
// r_backfill, mdot, insulation_depth) (src/physics.jl:48);
// mat[0..9] = HomogenousMaterialProperties field order (src/physics.jl:159-170);
// when n_layers > 0, mat[0..1] are ignored and the layer arrays are used
// (StratifiedMaterialProperties, src/physics.jl:81-94).
typedef struct or_scenario {
  int nx, ny, nz;
  const double *gx, *gy, *gz;
  int nbh;
  const double* bh;
  int n_layers;
  const double *k_layers, *rc_layers, *depths;
  const double* mat;
  int inlet_kind;  // 0 constant, 1 heat exchanger, 2 per-well table
  double inlet_value;
  const double* inlet_per_well;
} or_scenario;

static thread_local std::string g_err;
const char* or_last_error(void) { return g_err.c_str(); }

void or_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int or_max_threads(void) { return omp_get_max_threads(); }

static std::vector<Borehole> make_bhs(int nbh, const double* a) {
  std::vector<Borehole> v;
  for (int b = 0; b < nbh; ++b) {
    const double* q = a + 10 * b;
    v.push_back(gwo::make_borehole(q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], q[8], q[9]));
  }
  return v;
}

}  // extern "C"

template <class T>
static void fill_problem(Problem<T>* p, const or_scenario* s) {
  p->gx.assign(s->gx, s->gx + s->nx);
  p->gy.assign(s->gy, s->gy + s->ny);
  p->gz.assign(s->gz, s->gz + s->nz);
  p->bhs = make_bhs(s->nbh, s->bh);
  auto& m = p->mat;
  if (s->n_layers > 0) {
    m.k_rock_layers.assign(s->k_layers, s->k_layers + s->n_layers);
    m.rho_c_rock_layers.assign(s->rc_layers, s->rc_layers + s->n_layers);
    m.layer_depths.assign(s->depths, s->depths + s->n_layers);
  }
  m.k_rock = s->mat[0]; m.rho_c_rock = s->mat[1];
  m.k_water = s->mat[2]; m.rho_c_water = s->mat[3];
  m.k_steel = s->mat[4]; m.rho_c_steel = s->mat[5];
  m.k_insulating = s->mat[6]; m.rho_c_insulating = s->mat[7];
  m.k_backfill = s->mat[8]; m.rho_c_backfill = s->mat[9];
  p->inlet.kind = s->inlet_kind;
  p->inlet.value = s->inlet_value;
  if (s->inlet_kind == 2) p->inlet.per_well.assign(s->inlet_per_well, s->inlet_per_well + s->nbh);
  gwo::build_cache(*p);
}

extern "C" {

void* or_create(const or_scenario* s) {
  try {
    auto* p = new Problem<double>();
    fill_problem(p, s);
    return p;
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

// Forward-mode derivative (dual numbers, the role ForwardDiff plays for the
// reference) of the outlet objective
//   J = (1/nsteps) sum_n sum_b w_b T_out_b(n),
// T_out_b(n) the outlet of step n's second advection sub-step, from phi0 at t0.
// Parameters, in order: the rock conductivity of every layer (k_rock when the
// materials are homogeneous), then the inlet: its constant temperature
// (constant inlet) or one temperature per well (per-well table); a
// heat-exchanger inlet contributes no parameter.  n_params receives their
// count, dJ their derivatives (at most 32).
int or_objective_dual(const or_scenario* s, const double* phi0, double t0, double dt, int nsteps, double c3,
                      const double* w_bh, double* J, double* dJ, int* n_params) {
  constexpr int N = 32;
  using D = gwo::Dual<N>;
  try {
    const int nk = s->n_layers > 0 ? s->n_layers : 1;
    const int nt = s->inlet_kind == 0 ? 1 : (s->inlet_kind == 2 ? s->nbh : 0);
    if (nk + nt > N) { g_err = "more than 32 parameters"; return -1; }
    Problem<D> p;
    fill_problem(&p, s);
    if (s->n_layers > 0)
      for (int l = 0; l < nk; ++l) p.mat.k_rock_layers[l].d[l] = 1.0;
    else
      p.mat.k_rock.d[0] = 1.0;
    if (s->inlet_kind == 0) p.inlet.value.d[nk] = 1.0;
    if (s->inlet_kind == 2)
      for (int b = 0; b < s->nbh; ++b) p.inlet.per_well[b].d[nk + b] = 1.0;
    const gwo::Rock2Coef coef = gwo::rock2_coef_from_c3(c3);
    std::vector<D> phi(phi0, phi0 + p.cells()), temp(p.cells()), s0, s1, s2;
    D acc(0.0);
    for (int n = 0; n < nsteps; ++n) {
      gwo::full_step(p, phi.data(), temp.data(), t0 + (n + 1) * dt, dt, coef, s0, s1, s2);
      for (int b = 0; b < s->nbh; ++b) acc += D(w_bh[b]) * p.T_outlet[b];
    }
    acc = acc / D((double)nsteps);
    *J = acc.v;
    for (int q = 0; q < nk + nt; ++q) dJ[q] = acc.d[q];
    *n_params = nk + nt;
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

void or_destroy(void* h) { delete static_cast<Problem<double>*>(h); }

// 1 when ROCK2's degree rule (src/solver.jl:474-488 estimate) stays at its
// minimum for this dt.
int or_rock2_minimum_degree(void* h, double dt) {
  return gwo::rock2_minimum_degree(dt, static_cast<Problem<double>*>(h)->eigen_estimate) ? 1 : 0;
}

// info[0..5] = nx, ny, nz, countz, n_inner, n_outer
int or_info(void* h, long long* info, double* eigen) {
  auto* p = static_cast<Problem<double>*>(h);
  info[0] = p->nx; info[1] = p->ny; info[2] = p->nz; info[3] = p->countz;
  info[4] = (long long)p->idx_inner.size(); info[5] = (long long)p->idx_outer.size();
  if (eigen) *eigen = p->eigen_estimate;
  return 0;
}

// (i, j, bh) triples, 0-based, reference order.
int or_index_lists(void* h, int* inner, int* outer) {
  auto* p = static_cast<Problem<double>*>(h);
  for (size_t c = 0; c < p->idx_inner.size(); ++c)
    for (int q = 0; q < 3; ++q) inner[3 * c + q] = p->idx_inner[c][q];
  for (size_t c = 0; c < p->idx_outer.size(); ++c)
    for (int q = 0; q < 3; ++q) outer[3 * c + q] = p->idx_outer[c][q];
  return 0;
}

// Point-wise k and rho*c at every cell, layout (k fastest, then j, then i).
int or_material_maps(void* h, double* kmap, double* rcmap) {
  auto* p = static_cast<Problem<double>*>(h);
#pragma omp parallel for collapse(2)
  for (int i = 0; i < p->nx; ++i)
    for (int j = 0; j < p->ny; ++j)
      for (int k = 0; k < p->nz; ++k) {
        const size_t q = p->at(k, j, i);
        kmap[q] = gwo::thermal_conductivity(p->gx[i], p->gy[j], p->gz[k], p->bhs, p->mat);
        rcmap[q] = gwo::heat_capacity(p->gx[i], p->gy[j], p->gz[k], p->bhs, p->mat);
      }
  return 0;
}

int or_rhs_z(void* h, double* dphi, const double* phi) {
  gwo::rhs_diffusion_z(*static_cast<Problem<double>*>(h), dphi, phi);
  return 0;
}

int or_diffusion(void* h, double* out, const double* in, double dt, int dir, int plus_I) {
  gwo::diffusion_1d(*static_cast<Problem<double>*>(h), out, in, dt, dir, plus_I != 0);
  return 0;
}

int or_thomas(void* h, double* out, const double* in, double dt, int dir) {
  gwo::thomas(*static_cast<Problem<double>*>(h), out, in, dt, dir);
  return 0;
}

int or_advection(void* h, double* phi, double dt, double t, double* T_out) {
  auto* p = static_cast<Problem<double>*>(h);
  gwo::advection(*p, phi, dt, t);
  if (T_out) std::memcpy(T_out, p->T_outlet.data(), sizeof(double) * p->T_outlet.size());
  return 0;
}

int or_adi_adv(void* h, double* phi, double* temp, double t, double dt, double* T_out) {
  auto* p = static_cast<Problem<double>*>(h);
  gwo::adi_and_adv(*p, phi, temp, t, dt);
  if (T_out) std::memcpy(T_out, p->T_outlet.data(), sizeof(double) * p->T_outlet.size());
  return 0;
}

void or_rock2_coef(double c3, double* mu_sigma_tau) {
  gwo::Rock2Coef c = gwo::rock2_coef_from_c3(c3);
  mu_sigma_tau[0] = c.mu; mu_sigma_tau[1] = c.sigma; mu_sigma_tau[2] = c.tau;
}

int or_rock2_step(void* h, double* phi, double dt, double c3) {
  auto* p = static_cast<Problem<double>*>(h);
  std::vector<double> a, b, c;
  gwo::rock2_min_step(*p, phi, dt, gwo::rock2_coef_from_c3(c3), a, b, c);
  return 0;
}

// Runs `nsteps` full steps from t0.  hist (may be NULL) receives, per step and
// borehole, the outlet temperature of the step's second advection sub-step.
// seconds (may be NULL) receives the wall time of the stepping loop.
int or_run(void* h, double* phi, double t0, double dt, int nsteps, double c3, double* hist,
           double* seconds) {
  auto* p = static_cast<Problem<double>*>(h);
  const gwo::Rock2Coef coef = gwo::rock2_coef_from_c3(c3);
  std::vector<double> temp(p->cells()), s0, s1, s2;
  const size_t nb = p->bhs.size();
  auto t_start = std::chrono::steady_clock::now();
  for (int n = 0; n < nsteps; ++n) {
    const double t_end = t0 + (n + 1) * dt;
    gwo::full_step(*p, phi, temp.data(), t_end, dt, coef, s0, s1, s2);
    if (hist) std::memcpy(hist + nb * n, p->T_outlet.data(), sizeof(double) * nb);
  }
  auto t_stop = std::chrono::steady_clock::now();
  if (seconds) *seconds = std::chrono::duration<double>(t_stop - t_start).count();
  return 0;
}

int or_initial_condition(void* h, double* phi, double T_surface, double gradient) {
  gwo::initial_condition(*static_cast<Problem<double>*>(h), phi, T_surface, gradient);
  return 0;
}

// Grid generators (src/grids.jl).  Return the number of points written, or -1
// when `cap` is too small.
int or_adaptive_grid(double xmin, double xmax, double dx_fine, double growth, double dx_max,
                     int nbh, const double* bh, int dir_x, double* out, int cap) {
  auto g = gwo::adaptive_grid_1d(xmin, xmax, dx_fine, growth, dx_max, make_bhs(nbh, bh), dir_x != 0);
  if ((int)g.size() > cap) return -1;
  std::memcpy(out, g.data(), sizeof(double) * g.size());
  return (int)g.size();
}

int or_uniform_gridz(double zmin, double zmax, double dz, int nbh, const double* bh, double* out,
                     int cap) {
  auto g = gwo::uniform_gridz(zmin, zmax, dz, make_bhs(nbh, bh));
  if ((int)g.size() > cap) return -1;
  std::memcpy(out, g.data(), sizeof(double) * g.size());
  return (int)g.size();
}

// Borehole derived fields, 16 values in the reference's field order.
void or_borehole_fields(const double* args, double* out16) {
  Borehole b = make_bhs(1, args)[0];
  const double f[16] = {b.xc, b.yc, b.h, b.r_inner, b.t_inner, b.r_and_t_inner_sq, b.r_outer,
                        b.r_outer_sq, b.t_outer, b.r_and_t_outer_sq, b.r_backfill, b.r_backfill_sq,
                        b.mdot, b.v_inner, b.v_outer, b.insulation_depth};
  std::memcpy(out16, f, sizeof f);
}

// Rock lookups and d_max for known-answer tests.
double or_rock_k(void* h, double z) { return gwo::rock_conductivity_by_depth(z, static_cast<Problem<double>*>(h)->mat); }
double or_rock_rc(void* h, double z) { return gwo::rock_heat_capacity_by_depth(z, static_cast<Problem<double>*>(h)->mat); }
double or_dmax(void* h) { return gwo::dmax(static_cast<Problem<double>*>(h)->mat); }

}  // extern "C"
