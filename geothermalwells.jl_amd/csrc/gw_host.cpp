// gw_host.cpp -- set-up and C ABI of libgeowells_hip.
//
// Host work done once per context (create_cache, src/cache.jl:55-116):
//   * Borehole derived fields exactly as the reference constructor forms them
//     (src/physics.jl:48-60);
//   * material classification of every (i, j) column with the reference's
//     point-wise rules (src/physics.jl:199-274), evaluated in FP64 without
//     contraction so every branch decision matches; columns with identical
//     k(z), rho*c(z) profiles share a class, levels with identical values in
//     all classes share a zone;
//   * coefficients of the 1-D operators (src/solver.jl:34-109, 143-205) per
//     class and level (z) and per distinct (line, zone) (x and y);
//   * the pipe index lists in the reference's order (src/cache.jl:9-44).
// Work done once per sub-step length c = dt/2 (the reference's dt is fixed):
//   * LU factors of (I - cA) for every line set (Thomas without pivoting,
//     src/solver.jl:207-218);
//   * semi-Lagrangian departure tables per (borehole, level), with the
//     reference's branch arithmetic (src/solver.jl:299-376), and the level
//     ranges of the turnaround means and the outlet.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "geowells.h"
#include "gw_kernels.hpp"

using namespace gw;

// Scratch memory of gw_adi_adv_noalloc: the pipe reductions' chunk partials
// and turnaround level means of a sub-step, in a fixed device array (one copy
// per GPU, kept until the process ends) instead of the context's allocated
// tables.  Host code never passes this name to a kernel: gw_create resolves
// the current device's copy once (resolve_scratch) and the kernels receive
// that address.
constexpr size_t kNoAllocScratchDoubles = size_t(1) << 20;  // 8 MiB
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

struct HipFail {
  hipError_t e;
};
inline void hip_check(hipError_t e) {
  if (e != hipSuccess) throw HipFail{e};
}

struct HostBh {
  double xc, yc, h, r_inner, t_inner, r_and_t_inner_sq, r_outer, r_outer_sq, t_outer,
      r_and_t_outer_sq, r_backfill, r_backfill_sq, mdot, v_inner, v_outer, insulation_depth;
};

// Borehole{T}(...)  (src/physics.jl:48-60)
HostBh make_bh(const gw_borehole& a) {
  const double rho_water = 998.2;
  const double pi = 3.14159265358979323846;
  HostBh b;
  b.xc = a.xc; b.yc = a.yc; b.h = a.h; b.r_inner = a.r_inner; b.t_inner = a.t_inner;
  b.r_and_t_inner_sq = (a.r_inner + a.t_inner) * (a.r_inner + a.t_inner);
  b.r_outer = a.r_outer;
  b.r_outer_sq = a.r_outer * a.r_outer;
  b.t_outer = a.t_outer;
  b.r_and_t_outer_sq = (a.r_outer + a.t_outer) * (a.r_outer + a.t_outer);
  b.r_backfill = a.r_backfill;
  b.r_backfill_sq = a.r_backfill * a.r_backfill;
  b.mdot = a.mdot;
  const double A_inner = pi * (a.r_inner * a.r_inner);
  const double A_outer = pi * b.r_outer_sq - pi * b.r_and_t_inner_sq;
  b.v_inner = a.mdot / (A_inner * rho_water);
  b.v_outer = a.mdot / (A_outer * rho_water);
  b.insulation_depth = a.insulation_depth;
  return b;
}

struct Mat {
  std::vector<double> kl, rl, depth;  // layers (empty: homogeneous)
  double k_rock, rc_rock, k_water, rc_water, k_steel, rc_steel, k_ins, rc_ins, k_bf, rc_bf;
  double rock_k(double z) const {
    if (depth.empty()) return k_rock;
    for (size_t l = 0; l < depth.size(); ++l) if (z <= depth[l]) return kl[l];
    return kl.back();
  }
  double rock_rc(double z) const {
    if (depth.empty()) return rc_rock;
    for (size_t l = 0; l < depth.size(); ++l) if (z <= depth[l]) return rl[l];
    return rl.back();
  }
  // eigen_estimator_get_dmax (src/physics.jl:134-149, 177-185)
  double dmax() const {
    double d = 0;
    if (depth.empty()) d = k_rock / rc_rock;
    else for (size_t l = 0; l < kl.size(); ++l) d = std::max(d, kl[l] / rl[l]);
    d = std::max(d, k_water / rc_water);
    d = std::max(d, k_steel / rc_steel);
    d = std::max(d, k_ins / rc_ins);
    d = std::max(d, k_bf / rc_bf);
    return d;
  }
};

// get_thermal_conductivity / get_volumetric_heat_capacity at one point,
// restricted to the boreholes `near` (all others are skipped by the
// reference's radius test anyway).
void point_material(double x, double y, double z, const std::vector<HostBh>& bhs,
                    const std::vector<int>& near, const Mat& m, double& k, double& rc, bool* rock = nullptr) {
  k = m.rock_k(z);
  rc = m.rock_rc(z);
  if (rock) *rock = true;
  for (int bi : near) {
    const HostBh& bh = bhs[bi];
    const double dx = x - bh.xc, dy = y - bh.yc;
    const double r_sq = dx * dx + dy * dy;
    if (!(r_sq > bh.r_backfill_sq)) {  // src/physics.jl:207-228
      if (rock && z <= bh.h + 3) *rock = false;
      if (z <= bh.h) {
        if (r_sq < bh.r_inner * bh.r_inner) k = m.k_water;
        else if (r_sq < bh.r_and_t_inner_sq) k = m.k_ins;
        else if (r_sq < bh.r_outer_sq) k = m.k_water;
        else if (r_sq < bh.r_and_t_outer_sq) k = (z <= bh.insulation_depth) ? m.k_ins : m.k_steel;
        else k = m.k_bf;
      } else if (z <= bh.h + 3) {
        k = m.k_steel;
      }
    }
    if (!(r_sq > bh.r_and_t_outer_sq)) {  // src/physics.jl:251-270
      if (z <= bh.h) {
        if (r_sq < bh.r_inner * bh.r_inner) rc = m.rc_water;
        else if (r_sq < bh.r_and_t_inner_sq) rc = m.rc_ins;
        else if (r_sq < bh.r_outer_sq) rc = m.rc_water;
        else rc = (z <= bh.insulation_depth) ? m.rc_ins : m.rc_steel;
      } else if (z <= bh.h + 3) {
        rc = m.rc_steel;
      }
    }
  }
}

// Explicit operator coefficients along one line (src/solver.jl:34-109):
// (A u)_l = ea_l (u_{l+1} - u_l) - eb_l (u_l - u_{l-1}), zero-flux ends.
void line_coefficients(const double* g, int N, const double* kc, const double* rc, double* ea,
                       double* eb) {
  for (int l = 0; l < N; ++l) {
    const double dp = (l == N - 1) ? (g[l] - g[l - 1]) : (g[l + 1] - g[l]);
    const double dm = (l == 0) ? dp : (g[l] - g[l - 1]);
    const double kp = (l == N - 1) ? kc[l] : kc[l + 1];
    const double km = (l == 0) ? kc[l] : kc[l - 1];
    const double den = (dp + dm) * 0.5;
    ea[l] = (l == N - 1) ? 0.0 : ((kp + kc[l]) * 0.5 / dp) / den / rc[l];
    eb[l] = (l == 0) ? 0.0 : ((kc[l] + km) * 0.5 / dm) / den / rc[l];
  }
}

struct Scen {
  int nx, ny, nz;
  std::vector<double> gx, gy, gz;
  std::vector<HostBh> bh;
  Mat mat;
  gw_inlet inlet;
  std::vector<double> per_well;
  // classification
  std::vector<uint16_t> cls;      // nx*ny, index j + ny*i
  int ncls = 0;
  std::vector<double> kprof, rprof;  // ncls*nz
  std::vector<uint8_t> rockprof;      // ncls*nz: 1 where the conductivity is the rock layer's
  std::vector<int32_t> zone;      // nz
  int nzone = 0;
  std::vector<int> zone_rep;      // representative level of each zone
  std::vector<double> zcoef;      // ncls*nz*2
  std::vector<int32_t> xset, yset;  // ny*nzone, nx*nzone (global set ids)
  // pipes (reference order)
  std::vector<std::array<int, 3>> idx_inner, idx_outer;
  int countz = 0;
  double eigen = 0;
  double zrho = 0;                // Gershgorin bound of the spectral radius of A_z
  int64_t off = 0;
};

struct LineSet {
  int N;
  std::vector<double> ea, eb;
};

struct CTab {
  double c = 0;
  std::vector<BhDev> h_bh;
  int64_t* d_set_off = nullptr;
  double* d_coef = nullptr;
  double* d_pcoef[2] = {};          // the sets again, padded in the line kernels' LDS layout (DevView::pcoef)
  BhDev* d_bh = nullptr;
  Dep* d_dep = nullptr;
  double* d_means = nullptr;
  RedChunk* d_rchunk = nullptr;
  PartChunk* d_pchunk = nullptr;
  int32_t* d_pblock = nullptr;
  RedChunk* d_agroup = nullptr;
  double* d_partial = nullptr;
  std::vector<Dep> h_dep;           // host copy of the departure table (adjoint)
  std::string adv_err;              // why an advection sub-step of this length would yield NaN (none: empty)
  int nmeans = 0;
  int npartial = 0;                 // doubles of d_partial
  int nrchunk = 0;
  int npblock = 0;
  int nagroup = 0;
};

template <class T>
T* dev_upload(const std::vector<T>& h) {
  T* d = nullptr;
  const size_t n = std::max<size_t>(h.size(), 1) * sizeof(T);
  hip_check(hipMalloc(&d, n));
  if (!h.empty()) hip_check(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return d;
}

}  // namespace

// gw_steps launch modes: not yet timed, graph replays, direct launches.
enum { kStepsUndecided = 0, kStepsGraph = 1, kStepsEager = 2 };

struct gw_ctx {
  int prec = GW_F64;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  double c3 = 1.0 / 16.0;
  double zstab = 0.0;  // stability interval of the z-step polynomial
  std::vector<Scen> sc;
  std::vector<LineSet> sets;
  int nbh_total = 0;
  std::vector<PipeCol> pcol;
  std::vector<int> bh_col_in0, bh_n_in, bh_col_an0, bh_n_an, bh_scen;
  // c-independent device data
  std::vector<void*> allocs;
  ScenDev* d_sc = nullptr;
  int64_t *d_cell_prefix = nullptr, *d_col_prefix = nullptr;
  LineTile *d_xtiles[2] = {}, *d_ytiles[2] = {}, *d_htiles[2] = {};  // per k-window family
  uint16_t* d_cls = nullptr;
  int32_t* d_zone = nullptr;
  double* d_zcoef = nullptr;
  float* d_zcoef32 = nullptr;       // the same rounded to float (FP32 states)
  int32_t* d_setid = nullptr;
  PipeCol* d_pcol = nullptr;
  int64_t* d_pcol_off = nullptr;    // per pipe column: (j + ny i) nz, its level-0 offset in its scenario
  double *d_tout = nullptr, *d_inlet = nullptr, *d_inlet_tab = nullptr;
  uint32_t* d_step_ctr = nullptr;  // GW_F32_STATE_SR: step index keying the stochastic rounding
  int nx_max = 0, ny_max = 0, nz_max = 0, countz_max = 0;
  int64_t cells_total = 0, cols_total = 0;
  // tile offsets of the line-length classes, per k-window family
  int64_t xcls[2][kLineClasses + 1] = {}, ycls[2][kLineClasses + 1] = {}, hcls[2][kLineClasses + 1] = {};
  int32_t hslots[2][kLineClasses] = {};
  int64_t hone[2][kLineClasses] = {};
  // k-window width (64 or 32) of each line operator (mode, direction): the
  // length default until the first gw_steps times both widths (autotune), or
  // the width the problem flags force
  int lw_sel[5][2] = {};
  int force_lw = 0;
  // line-length class whose tile shape runs each class's lines (identity
  // unless small classes of a ragged batch are merged into larger ones)
  int cls_map[kLineClasses] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9};
  int line_cls(int n) const { return cls_map[line_class(n)]; }
  bool tuned = false;
  int steps_mode = kStepsUndecided;     // gw_steps: graph replays or direct launches (kSteps*)
  int adj_lw[2] = {0, 0};  // k-window width of the adjoint line modes ADJ_Y, ADJ_X (0: not yet timed)
  int lw(int mode, int dir) const { return lw_sel[mode][dir]; }
  bool fused_ok = false;  // fused z-step line kernels usable (few enough column classes)
  std::map<double, CTab> ctabs;
  bool host_inlet = false;
  // captured step graph
  hipGraphExec_t graph_exec = nullptr;   // step body, state phi -> buf2
  hipGraphExec_t graph_exec2 = nullptr;  // step body, state buf2 -> phi
  void* graph_phi = nullptr;
  double graph_dt = 0;
  // second field buffer of the fused pipeline (its halo kernels write out of place)
  void* buf2 = nullptr;
  // device the context was created on, and that device's copy of the
  // file-scope scratch array (gw_adi_adv_noalloc)
  // post-processing: per scenario, the device column offsets of the inner
  // then the annulus index list (built on first use), and an output buffer
  std::map<int, int64_t*> post_off;
  double* post_buf = nullptr;
  // side streams of the concurrent class launches (batched contexts)
  LineStreams ls{};
  bool has_ls = false;
  const LineStreams* lsp() const { return has_ls ? &ls : nullptr; }
  int device = 0;
  double* scratch_dev = nullptr;
  std::string scratch_err;  // why scratch_dev is NULL

  template <class T>
  T* up(const std::vector<T>& h) {
    T* d = dev_upload(h);
    allocs.push_back(d);
    return d;
  }
  ~gw_ctx() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph_exec2) (void)hipGraphExecDestroy(graph_exec2);
    if (buf2) (void)hipFree(buf2);
    for (auto& kv : post_off) (void)hipFree(kv.second);
    if (post_buf) (void)hipFree(post_buf);
    for (auto& kv : ctabs) {
      (void)hipFree(kv.second.d_set_off); (void)hipFree(kv.second.d_coef); (void)hipFree(kv.second.d_bh);
      (void)hipFree(kv.second.d_pcoef[0]); (void)hipFree(kv.second.d_pcoef[1]);
      (void)hipFree(kv.second.d_dep); (void)hipFree(kv.second.d_means);
      (void)hipFree(kv.second.d_rchunk); (void)hipFree(kv.second.d_partial);
      (void)hipFree(kv.second.d_pchunk); (void)hipFree(kv.second.d_pblock);
      (void)hipFree(kv.second.d_agroup);
    }
    for (void* p : allocs) (void)hipFree(p);
    if (has_ls) {
      for (int q = 0; q < kSideStreams; ++q) {
        (void)hipStreamDestroy(ls.side[q]);
        (void)hipEventDestroy(ls.join[q]);
      }
      (void)hipEventDestroy(ls.fork);
    }
    if (own_stream && stream) (void)hipStreamDestroy(stream);
  }
  DevView view(const CTab* t) const {
    DevView v{};
    v.sc = d_sc; v.nscen = (int)sc.size();
    v.cell_prefix = d_cell_prefix; v.col_prefix = d_col_prefix;
    for (int f = 0; f < 2; ++f) {
      v.xtiles[f] = d_xtiles[f]; v.ytiles[f] = d_ytiles[f]; v.htiles[f] = d_htiles[f];
      for (int q = 0; q <= kLineClasses; ++q) {
        v.xcls[f][q] = xcls[f][q]; v.ycls[f][q] = ycls[f][q]; v.hcls[f][q] = hcls[f][q];
        if (q < kLineClasses) {
          v.hslots[f][q] = hslots[f][q];
          v.hone[f][q] = hone[f][q];
        }
      }
    }
    v.cls = d_cls; v.zone = d_zone; v.zcoef = d_zcoef; v.zcoef32 = d_zcoef32; v.setid = d_setid;
    for (const Scen& s : sc) v.zc_cls = std::max(v.zc_cls, std::min(s.ncls, kZcMax));
    if (t) { v.set_off = t->d_set_off; v.coef = t->d_coef; v.pcoef[0] = t->d_pcoef[0]; v.pcoef[1] = t->d_pcoef[1] ? t->d_pcoef[1] : t->d_pcoef[0];
             v.bh = t->d_bh; v.dep = t->d_dep; v.means = t->d_means;
             v.rchunk = t->d_rchunk; v.nrchunk = t->nrchunk; v.partial = t->d_partial;
             v.pchunk = t->d_pchunk; v.pblock = t->d_pblock; v.npblock = t->npblock;
             v.agroup = t->d_agroup; v.nagroup = t->nagroup; }
    v.nbh = nbh_total; v.pcol = d_pcol; v.npcol = (int)pcol.size(); v.pcol_off = d_pcol_off;
    v.tout = d_tout; v.inlet = d_inlet; v.inlet_tab = d_inlet_tab;
    v.step_ctr = d_step_ctr;
    v.nx_max = nx_max; v.ny_max = ny_max; v.nz_max = nz_max; v.countz_max = countz_max;
    v.cells_total = cells_total; v.cols_total = cols_total;
    return v;
  }
};

namespace {

// Classification, zones and z coefficients of one scenario.
void classify(Scen& s) {
  const int nx = s.nx, ny = s.ny, nz = s.nz;
  const size_t nb = s.bh.size();
  // boreholes whose skip radius can contain the column
  std::vector<std::vector<int>> near((size_t)nx * ny);
  for (size_t b = 0; b < nb; ++b) {
    const HostBh& bh = s.bh[b];
    const double R = std::sqrt(std::max(bh.r_backfill_sq, bh.r_and_t_outer_sq)) * 1.000001 + 1e-12;
    const int i0 = (int)(std::lower_bound(s.gx.begin(), s.gx.end(), bh.xc - R) - s.gx.begin());
    const int i1 = (int)(std::upper_bound(s.gx.begin(), s.gx.end(), bh.xc + R) - s.gx.begin());
    const int j0 = (int)(std::lower_bound(s.gy.begin(), s.gy.end(), bh.yc - R) - s.gy.begin());
    const int j1 = (int)(std::upper_bound(s.gy.begin(), s.gy.end(), bh.yc + R) - s.gy.begin());
    for (int i = std::max(i0 - 1, 0); i < std::min(i1 + 1, nx); ++i)
      for (int j = std::max(j0 - 1, 0); j < std::min(j1 + 1, ny); ++j) {
        const double dx = s.gx[i] - bh.xc, dy = s.gy[j] - bh.yc;
        const double r_sq = dx * dx + dy * dy;
        if (!(r_sq > bh.r_backfill_sq) || !(r_sq > bh.r_and_t_outer_sq))
          near[(size_t)j + (size_t)ny * i].push_back((int)b);
      }
  }
  // class 0: plain rock
  std::unordered_map<std::string, int> cmap;
  s.kprof.clear(); s.rprof.clear(); s.rockprof.clear();
  std::vector<double> kp(nz), rp(nz);
  std::vector<uint8_t> rk(nz);
  const std::vector<int> none;
  auto add_profile = [&](void) -> int {
    std::string key(reinterpret_cast<const char*>(kp.data()), nz * sizeof(double));
    key.append(reinterpret_cast<const char*>(rp.data()), nz * sizeof(double));
    key.append(reinterpret_cast<const char*>(rk.data()), nz);
    auto it = cmap.find(key);
    if (it != cmap.end()) return it->second;
    const int id = (int)cmap.size();
    cmap.emplace(std::move(key), id);
    s.kprof.insert(s.kprof.end(), kp.begin(), kp.end());
    s.rprof.insert(s.rprof.end(), rp.begin(), rp.end());
    s.rockprof.insert(s.rockprof.end(), rk.begin(), rk.end());
    return id;
  };
  auto material = [&](double x, double y, int k, const std::vector<int>& nr) {
    bool r = true;
    point_material(x, y, s.gz[k], s.bh, nr, s.mat, kp[k], rp[k], &r);
    rk[k] = r ? 1 : 0;
  };
  for (int k = 0; k < nz; ++k) material(0, 0, k, none);
  add_profile();
  s.cls.assign((size_t)nx * ny, 0);
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      const auto& nr = near[(size_t)j + (size_t)ny * i];
      if (nr.empty()) continue;
      for (int k = 0; k < nz; ++k) material(s.gx[i], s.gy[j], k, nr);
      const int id = add_profile();
      if (id > 65535) throw std::runtime_error("more than 65536 material column classes");
      s.cls[(size_t)j + (size_t)ny * i] = (uint16_t)id;
    }
  s.ncls = (int)cmap.size();
  // zones
  std::unordered_map<std::string, int> zmap;
  s.zone.assign(nz, 0);
  s.zone_rep.clear();
  std::vector<double> key(2 * s.ncls);
  for (int k = 0; k < nz; ++k) {
    for (int c = 0; c < s.ncls; ++c) {
      key[2 * c] = s.kprof[(size_t)c * nz + k];
      key[2 * c + 1] = s.rprof[(size_t)c * nz + k];
    }
    std::string kk(reinterpret_cast<const char*>(key.data()), key.size() * sizeof(double));
    auto it = zmap.find(kk);
    if (it == zmap.end()) {
      const int id = (int)zmap.size();
      zmap.emplace(std::move(kk), id);
      s.zone_rep.push_back(k);
      s.zone[k] = id;
    } else {
      s.zone[k] = it->second;
    }
  }
  s.nzone = (int)zmap.size();
  // z coefficients per (class, level)
  s.zcoef.assign((size_t)s.ncls * nz * 2, 0.0);
  std::vector<double> ea(nz), eb(nz);
  for (int c = 0; c < s.ncls; ++c) {
    line_coefficients(s.gz.data(), nz, &s.kprof[(size_t)c * nz], &s.rprof[(size_t)c * nz], ea.data(), eb.data());
    for (int k = 0; k < nz; ++k) {
      s.zcoef[2 * ((size_t)c * nz + k)] = ea[k];
      s.zcoef[2 * ((size_t)c * nz + k) + 1] = eb[k];
      s.zrho = std::max(s.zrho, 2.0 * (ea[k] + eb[k]));
    }
  }
}

// x- and y-line coefficient sets (deduplicated across lines, zones and scenarios).
void build_line_sets(gw_ctx& ctx, Scen& s, std::unordered_map<std::string, int>& smap) {
  auto add_set = [&](int N, const std::vector<double>& ea, const std::vector<double>& eb) -> int {
    std::string key(reinterpret_cast<const char*>(&N), sizeof N);
    key.append(reinterpret_cast<const char*>(ea.data()), N * sizeof(double));
    key.append(reinterpret_cast<const char*>(eb.data()), N * sizeof(double));
    auto it = smap.find(key);
    if (it != smap.end()) return it->second;
    const int id = (int)ctx.sets.size();
    smap.emplace(std::move(key), id);
    ctx.sets.push_back(LineSet{N, ea, eb});
    return id;
  };
  const int nx = s.nx, ny = s.ny, nz = s.nz;
  s.xset.assign((size_t)ny * s.nzone, 0);
  s.yset.assign((size_t)nx * s.nzone, 0);
  std::vector<double> kc(std::max(nx, ny)), rc(std::max(nx, ny)), ea(std::max(nx, ny)), eb(std::max(nx, ny));
  for (int z = 0; z < s.nzone; ++z) {
    const int kr = s.zone_rep[z];
    for (int j = 0; j < ny; ++j) {  // x lines
      for (int i = 0; i < nx; ++i) {
        const int c = s.cls[(size_t)j + (size_t)ny * i];
        kc[i] = s.kprof[(size_t)c * nz + kr];
        rc[i] = s.rprof[(size_t)c * nz + kr];
      }
      line_coefficients(s.gx.data(), nx, kc.data(), rc.data(), ea.data(), eb.data());
      std::vector<double> a(ea.begin(), ea.begin() + nx), b(eb.begin(), eb.begin() + nx);
      s.xset[(size_t)j * s.nzone + z] = add_set(nx, a, b);
    }
    for (int i = 0; i < nx; ++i) {  // y lines
      for (int j = 0; j < ny; ++j) {
        const int c = s.cls[(size_t)j + (size_t)ny * i];
        kc[j] = s.kprof[(size_t)c * nz + kr];
        rc[j] = s.rprof[(size_t)c * nz + kr];
      }
      line_coefficients(s.gy.data(), ny, kc.data(), rc.data(), ea.data(), eb.data());
      std::vector<double> a(ea.begin(), ea.begin() + ny), b(eb.begin(), eb.begin() + ny);
      s.yset[(size_t)i * s.nzone + z] = add_set(ny, a, b);
    }
  }
}

// Line tiles of one scenario and direction.  Each tile is a k window of lw
// lanes (64 or 32) whose lanes meet at most 3 coefficient sets (plain tiles)
// or 4 (halo tiles).  Plain tiles output all their lanes; halo tiles (halo =
// 3, for the fused z-step) output lanes halo .. lw-1-halo, the others only
// feed the z-stencil of the outputs.  A halo tile needs the sets of 7
// consecutive levels for each output, and a layer boundary a few levels above
// a borehole bottom (the casing's bottom level below it) puts 4 there.

// Offsets (in doubles) of each set's padded copy in DevView::pcoef: SET_NARR
// arrays of its line-length class's longest line.  They depend on the sets'
// lengths only, so every table of the context and the tiles agree on them.
std::vector<int64_t> padded_set_offsets(const gw_ctx& ctx) {
  std::vector<int64_t> off(ctx.sets.size());
  int64_t tot = 0;
  for (size_t q = 0; q < ctx.sets.size(); ++q) {
    off[q] = tot;
    tot += SET_NARR * (int64_t)kLineClassMax[ctx.line_cls(ctx.sets[q].N)];
  }
  return off;
}

// Merges the line-length classes of a ragged batch that hold under
// kClassMergeShare of its line cells, in x and in y, into the next larger
// class present: their lines then run with that class's tile shape (longer
// segments or more waves, the padding rows idle) in its launch, instead of a
// launch of their own whose few tiles run latency-bound.
constexpr double kClassMergeShare = 0.05;
void merge_small_classes(gw_ctx& ctx) {
  double cells[2][kLineClasses] = {}, tot[2] = {};
  for (const Scen& s : ctx.sc) {
    const double n = (double)s.nx * s.ny * s.nz;
    cells[0][line_class(s.nx)] += n;
    cells[1][line_class(s.ny)] += n;
    tot[0] += n;
    tot[1] += n;
  }
  for (int q = 0; q < kLineClasses; ++q) {
    if (cells[0][q] == 0 && cells[1][q] == 0) continue;
    if (cells[0][q] >= kClassMergeShare * tot[0] || cells[1][q] >= kClassMergeShare * tot[1]) continue;
    int q2 = q + 1;
    while (q2 < kLineClasses && cells[0][q2] == 0 && cells[1][q2] == 0) ++q2;
    if (q2 == kLineClasses) break;
    for (int d = 0; d < 2; ++d) { cells[d][q2] += cells[d][q]; cells[d][q] = 0; }
    for (int r = 0; r < kLineClasses; ++r)
      if (ctx.cls_map[r] == q) ctx.cls_map[r] = q2;
  }
}

void line_tiles(const Scen& s, int si, int dir, int halo, int lw, const std::vector<int64_t>& pset_off,
                std::vector<LineTile>& out) {
  const int nlines = dir == 0 ? s.ny : s.nx;
  const std::vector<int32_t>& ids = dir == 0 ? s.xset : s.yset;
  const int width = lw - 2 * halo;
  const int kMaxSets = halo > 0 ? 4 : 3;
  for (int line = 0; line < nlines; ++line) {
    auto sid_at = [&](int k) {
      const int kc = std::min(std::max(k, 0), s.nz - 1);
      return ids[(size_t)line * s.nzone + s.zone[kc]];
    };
    int k0 = 0;
    while (k0 < s.nz) {
      // widest output range [k0, k0+n) whose needed lanes [k0-halo, k0+n+halo)
      // use at most kMaxSets sets
      std::vector<int> sets;
      int n = 0;
      for (int kk = k0 - halo; kk < k0 + halo; ++kk) {
        const int q = sid_at(kk);
        if (std::find(sets.begin(), sets.end(), q) == sets.end()) sets.push_back(q);
      }
      while (n < width && k0 + n < s.nz) {
        const int q = sid_at(k0 + n + halo);
        if (std::find(sets.begin(), sets.end(), q) == sets.end()) {
          if ((int)sets.size() == kMaxSets) break;
          sets.push_back(q);
        }
        ++n;
      }
      if (n == 0 || (int)sets.size() > kMaxSets)
        throw std::runtime_error("more than 4 material zones within a 7-level z window");
      LineTile t{};
      t.scen = si; t.line = line; t.kw = k0 - halo; t.ko0 = k0; t.ko1 = k0 + n;
      t.nset = (int)sets.size();
      for (int q = 0; q < 4; ++q) t.set[q] = q < t.nset ? sets[q] : -1;
      for (int q = 0; q < 4; ++q) t.poff[q] = q < t.nset ? (uint32_t)pset_off[sets[q]] : 0u;
      t.N = dir == 0 ? s.nx : s.ny;
      t.stride = dir == 0 ? s.nz * s.ny : s.nz;
      t.nz = s.nz;
      t.base = s.off + (dir == 0 ? (int64_t)s.nz * line : (int64_t)s.nz * s.ny * line) + std::max(t.kw, 0);
      // each lane's slot in set[] (0 for a lane whose set is not among them:
      // such a lane is past the outputs' halo and feeds no output)
      t.slot_b0 = t.slot_b1 = 0;
      for (int kk = 0; kk < lw; ++kk) {
        const auto it = std::find(sets.begin(), sets.end(), sid_at(t.kw + kk));
        const uint64_t slot = it == sets.end() ? 0 : (uint64_t)(it - sets.begin());
        t.slot_b0 |= (slot & 1) << kk;
        t.slot_b1 |= ((slot >> 1) & 1) << kk;
      }
      out.push_back(t);
      k0 += n;
    }
  }
}

// create_advection_index_lists  (src/cache.jl:9-44)
void index_lists(Scen& s) {
  s.idx_inner.clear(); s.idx_outer.clear();
  for (size_t b = 0; b < s.bh.size(); ++b) {
    const HostBh& bh = s.bh[b];
    for (int i = 0; i < s.nx; ++i)
      for (int j = 0; j < s.ny; ++j) {
        const double dx = s.gx[i] - bh.xc, dy = s.gy[j] - bh.yc;
        const double r_sq = dx * dx + dy * dy;
        const double rti = bh.r_inner + bh.t_inner;
        if (r_sq < bh.r_inner * bh.r_inner) s.idx_inner.push_back({i, j, (int)b});
        else if (rti * rti <= r_sq && r_sq < bh.r_outer * bh.r_outer) s.idx_outer.push_back({i, j, (int)b});
      }
  }
  double hmax = -1e300;
  for (auto& b : s.bh) hmax = std::max(hmax, b.h);
  s.countz = 0;
  for (double z : s.gz) if (z <= hmax) ++s.countz;
  double dzmin = 1e300;
  for (int k = 0; k + 1 < s.nz; ++k) dzmin = std::min(dzmin, s.gz[k + 1] - s.gz[k]);
  s.eigen = 8 * s.mat.dmax() / (dzmin * dzmin);  // src/solver.jl:474-488
}

// interpolation_helper + gpu_searchsortedfirst (src/solver.jl:362-389), 0-based.
void interp(const std::vector<double>& g, double dep, int& kl, int& kr, double& a) {
  int lo = 0, hi = (int)g.size() - 1;
  while (lo < hi) {
    const int mid = lo + (hi - lo) / 2;
    if (g[mid] < dep) lo = mid + 1; else hi = mid;
  }
  kr = lo;
  kl = std::max(kr - 1, 0);
  kr = std::min(kr, (int)g.size() - 1);
  a = (dep - g[kl]) / (g[kr] - g[kl]);
}

// Tables that depend on the sub-step length c.
int build_ctab(gw_ctx& ctx, double c, CTab*& out) {
  auto it = ctx.ctabs.find(c);
  if (it != ctx.ctabs.end()) { out = &it->second; return GW_OK; }
  CTab t;
  t.c = c;
  // LU factors of (I - cA) per line set
  std::vector<int64_t> set_off(ctx.sets.size());
  int64_t tot = 0;
  for (size_t q = 0; q < ctx.sets.size(); ++q) { set_off[q] = tot; tot += SET_NARR * (int64_t)ctx.sets[q].N; }
  std::vector<double> coef((size_t)tot);
  for (size_t q = 0; q < ctx.sets.size(); ++q) {
    const LineSet& L = ctx.sets[q];
    const int N = L.N;
    double* bd = &coef[set_off[q]] + SET_BD * N;
    double* ea = &coef[set_off[q]] + SET_EA * N;
    double* eb = &coef[set_off[q]] + SET_EB * N;
    double* fw = &coef[set_off[q]] + SET_FW * N;
    double* bu = &coef[set_off[q]] + SET_BU * N;
    // (I - cA): lower_l = -c eb_l, diag_l = 1 + c (ea_l + eb_l), upper_l = -c ea_l
    double piv = 1.0 + c * (L.ea[0] + L.eb[0]);
    for (int l = 1; l < N; ++l) {
      const double lower = -c * L.eb[l];
      const double upper_prev = -c * L.ea[l - 1];
      const double w = lower / piv;
      bd[l - 1] = 1.0 / piv;
      piv = (1.0 + c * (L.ea[l] + L.eb[l])) - w * upper_prev;
    }
    bd[N - 1] = 1.0 / piv;
    for (int l = 0; l < N; ++l) {
      ea[l] = L.ea[l];
      eb[l] = L.eb[l];
      fw[l] = (-c * L.eb[l]) * bd[l > 0 ? l - 1 : 0];
      bu[l] = (-c * L.ea[l]) * bd[l];
    }
  }
  // The same values in the layout the line kernels keep in an LDS slot (bd,
  // fw, bu, then the explicit coefficients), each array padded to its
  // line-length class's longest line with identity rows (bd 1, the others 0):
  // a tile copies the arrays it needs with 16-byte LDS DMA loads.  FP64
  // states: ea, eb.  FP32 states: the explicit step's products c ea, c eb,
  // as doubles for the 32-lane family and rounded to float (an array of NM
  // floats at the start of each NM-double array) for the 64-lane family.
  const std::vector<int64_t> pset_off = padded_set_offsets(ctx);
  const int64_t ptot = ctx.sets.empty() ? 0 : pset_off.back() + SET_NARR * (int64_t)kLineClassMax[ctx.line_cls(ctx.sets.back().N)];
  const bool f32 = ctx.prec != GW_F64;
  std::vector<double> pcoef[2];
  for (int f = 0; f < (f32 ? 2 : 1); ++f) {
    pcoef[f].assign((size_t)std::max<int64_t>(ptot, 2), 0.0);
    const double cs = f32 ? c : 1.0;
    for (size_t q = 0; q < ctx.sets.size(); ++q) {
      const int N = ctx.sets[q].N, NM = kLineClassMax[ctx.line_cls(N)];
      const double* g = &coef[set_off[q]];
      double* p = &pcoef[f][pset_off[q]];
      for (int l = 0; l < NM; ++l) {
        const bool ok = l < N;
        p[0 * NM + l] = ok ? g[SET_BD * N + l] : 1.0;
        p[1 * NM + l] = ok ? g[SET_FW * N + l] : 0.0;
        p[2 * NM + l] = ok ? g[SET_BU * N + l] : 0.0;
        const double ea = ok ? cs * g[SET_EA * N + l] : 0.0;
        const double eb = ok ? cs * g[SET_EB * N + l] : 0.0;
        if (f32 && f == 0) {
          reinterpret_cast<float*>(p + 3 * NM)[l] = (float)ea;
          reinterpret_cast<float*>(p + 4 * NM)[l] = (float)eb;
        } else {
          p[3 * NM + l] = ea;
          p[4 * NM + l] = eb;
        }
      }
    }
  }
  // departure tables
  std::vector<Dep> dep;
  std::vector<BhDev> hb(ctx.nbh_total);
  std::vector<RedChunk> chunks, groups;
  // columns per advection block: at most about 2048 blocks per launch (all of
  // them resident at once, 8 per CU, instead of a second partial round), 1 to
  // 8 columns each (a block walks its columns one after another, the next
  // column's loads in flight)
  int64_t npipe = 0;
  for (int b = 0; b < ctx.nbh_total; ++b) npipe += ctx.bh_n_in[b] + ctx.bh_n_an[b];
  const int kAdvGroup = (int)std::min<int64_t>(kAdvGroupMax, std::max<int64_t>(1, (npipe + 2047) / 2048));
  int gb = 0, mean_off = 0, poff = 0;
  for (size_t si = 0; si < ctx.sc.size(); ++si) {
    const Scen& s = ctx.sc[si];
    for (size_t b = 0; b < s.bh.size(); ++b, ++gb) {
      const HostBh& bh = s.bh[b];
      BhDev& B = hb[gb];
      std::memset(&B, 0, sizeof B);
      B.scen = (int)si;
      B.col_in0 = ctx.bh_col_in0[gb]; B.n_in = ctx.bh_n_in[gb];
      B.col_an0 = ctx.bh_col_an0[gb]; B.n_an = ctx.bh_n_an[gb];
      B.inlet_kind = s.inlet.kind;
      B.inlet_value = s.inlet.value;
      // Configurations for which the reference's advection yields NaN: the
      // sub-step refuses them (the diffusion operators of this c still run)
      auto adv_fail = [&](const std::string& msg) { if (t.adv_err.empty()) t.adv_err = msg; };
      if (B.n_in == 0)
        adv_fail("borehole " + std::to_string(b + 1) + " of scenario " + std::to_string(si) +
                 " has no inner-pipe grid columns: its outlet mean would be 0/0 (src/solver.jl:257)");
      const double zmax = bh.v_inner * c;
      int nout = 0;
      for (int k = 0; k < s.nz; ++k) if (s.gz[k] <= zmax) ++nout;
      // gz is sorted, so the outlet levels are k < nout
      B.n_outlev = nout;
      if (nout == 0) adv_fail("no outlet levels (gridz[1] > v_inner*dt)");
      B.dep_off = (int)dep.size();
      int klo = 1 << 30, khi = -1;
      for (int k = 0; k < s.countz; ++k) {
        const double z = s.gz[k];
        Dep di{}, da{};
        if (z > bh.h) {
          di.type = DEP_KEEP; da.type = DEP_KEEP;
        } else {
          const double zd = std::fma(bh.v_inner, c, z);  // src/solver.jl:304
          if (zd > bh.h) {
            const double dt1 = (bh.h - z) / bh.v_inner;
            const double dt2 = c - dt1;
            const double zd2 = std::fma(-bh.v_outer, dt2, bh.h);  // :310
            interp(s.gz, zd2, di.kl, di.kr, di.a);
            di.type = DEP_TURN;
            klo = std::min(klo, di.kl);
            khi = std::max(khi, di.kr);
          } else {
            interp(s.gz, zd, di.kl, di.kr, di.a);
            di.type = DEP_OWN;
          }
          if (di.kl == di.kr || !std::isfinite(di.a)) {
            adv_fail("degenerate inner-pipe departure point (zero flow rate or a turnaround departure above the "
                     "surface): the reference yields NaN (src/solver.jl:310,373)");
            di = Dep{};
            di.type = DEP_KEEP;
          }
          const double za = std::fma(-bh.v_outer, c, z);  // :348
          if (za <= 0.0) {
            da.type = DEP_INLET;
          } else {
            interp(s.gz, za, da.kl, da.kr, da.a);
            da.type = DEP_OWN;
            if (da.kl == da.kr || !std::isfinite(da.a)) {
              adv_fail("degenerate annulus departure point (src/solver.jl:352)");
              da = Dep{};
              da.type = DEP_KEEP;
            }
          }
        }
        dep.push_back(di);
        dep.push_back(da);
      }
      if (khi >= 0) {
        if (B.n_an == 0)
          adv_fail("borehole " + std::to_string(b + 1) +
                   " has no annulus grid columns: the turnaround mean would be 0/0 (src/solver.jl:334)");
        B.klo = klo;
        B.nlev = khi - klo + 1;
      } else {
        B.klo = 0;
        B.nlev = 0;
      }
      B.mean_off = mean_off;
      mean_off += B.nlev;
      // reduction chunks of kRedChunk pipe columns
      constexpr int kChunk = kRedChunk;
      B.red_in0 = (int)chunks.size();
      for (int c0 = 0; c0 < B.n_in; c0 += kChunk) {
        const int nc = std::min(kChunk, B.n_in - c0);
        chunks.push_back(RedChunk{gb, 0, B.col_in0 + c0, nc, poff, 0});
        poff += B.n_outlev;
      }
      B.n_red_in = (int)chunks.size() - B.red_in0;
      B.red_an0 = (int)chunks.size();
      if (B.nlev > 0)
        for (int c0 = 0; c0 < B.n_an; c0 += kChunk) {
          const int nc = std::min(kChunk, B.n_an - c0);
          chunks.push_back(RedChunk{gb, 1, B.col_an0 + c0, nc, poff, 0});
          poff += B.nlev;
        }
      B.n_red_an = (int)chunks.size() - B.red_an0;
      // advection groups: kAdvGroup columns of one kind each
      for (int c0 = 0; c0 < B.n_in; c0 += kAdvGroup)
        groups.push_back(RedChunk{gb, 0, B.col_in0 + c0, std::min(kAdvGroup, B.n_in - c0), 0, 0});
      for (int c0 = 0; c0 < B.n_an; c0 += kAdvGroup)
        groups.push_back(RedChunk{gb, 1, B.col_an0 + c0, std::min(kAdvGroup, B.n_an - c0), 0, 0});
    }
  }
  // the partial-sum kernel's view of the chunks, packed into blocks of
  // consecutive chunks with at most kPartItems levels in all
  std::vector<PartChunk> pch(chunks.size());
  std::vector<int32_t> pblock;
  int items = 0;
  for (size_t q = 0; q < chunks.size(); ++q) {
    const RedChunk& R = chunks[q];
    const BhDev& B = hb[R.bh];
    const int nl = R.part == 0 ? B.n_outlev : B.nlev;
    const int nq = pblock.empty() ? 0 : (int)q - pblock.back();
    if (pblock.empty() || nq == kPartChunks || items + nl > kPartItems || items > kPartItems) {
      pblock.push_back((int32_t)q);
      items = 0;
    }
    PartChunk& P = pch[q];
    P.base = ctx.sc[B.scen].off + (R.part == 0 ? 0 : B.klo);
    P.col0 = R.col0; P.ncol = R.ncol; P.nl = nl; P.poff = R.poff; P.item0 = items; P.pad = 0;
    items += nl;
  }
  const int npblock = (int)pblock.size();
  pblock.push_back((int32_t)chunks.size());
  try {
    t.d_set_off = dev_upload(set_off);
    t.d_coef = dev_upload(coef);
    t.d_pcoef[0] = dev_upload(pcoef[0]);
    t.d_pcoef[1] = f32 ? dev_upload(pcoef[1]) : nullptr;
    t.d_bh = dev_upload(hb);
    t.d_dep = dev_upload(dep);
    std::vector<double> m(std::max(mean_off, 1), 0.0);
    t.d_means = dev_upload(m);
    t.nmeans = mean_off;
    t.d_rchunk = dev_upload(chunks);
    t.nrchunk = (int)chunks.size();
    t.d_pchunk = dev_upload(pch);
    t.d_pblock = dev_upload(pblock);
    t.npblock = npblock;
    t.d_agroup = dev_upload(groups);
    t.nagroup = (int)groups.size();
    std::vector<double> part(std::max(poff, 1), 0.0);
    t.d_partial = dev_upload(part);
    t.npartial = poff;
  } catch (HipFail& e) {
    return fail(GW_ENOMEM, std::string("device allocation failed: ") + hipGetErrorString(e.e));
  }
  t.h_bh = hb;
  t.h_dep = dep;
  auto res = ctx.ctabs.emplace(c, t);
  out = &res.first->second;
  return GW_OK;
}

int hip_status(hipError_t e, const char* where) {
  if (e == hipSuccess) return GW_OK;
  return fail(GW_EHIP, std::string(where) + ": " + hipGetErrorString(e));
}

// Largest x with |R(-y)| <= 1 on [0, x] for R(z) = 1 + z + z^2/2 + c3 z^3.
double stability_bound(double c3) {
  double last = 0.0;
  for (int q = 1; q <= 200000; ++q) {
    const double y = -q * 1e-4;
    const double r = 1 + y + 0.5 * y * y + c3 * y * y * y;
    if (std::fabs(r) > 1 + 1e-12) break;
    last = -y;
  }
  return last;
}

// One advection sub-step (advection!, src/solver.jl:250-264) on phi.
// scratch (may be NULL): at least t->npartial + t->nmeans doubles of device
// memory used for the reduction partials and level means instead of the
// tables' own buffers.
int sub_step_advect(gw_ctx& ctx, const CTab* t, void* phi, double tau, double* jacc = nullptr,
                    double* scratch_mem = nullptr, double* hist = nullptr) {
  if (!t->adv_err.empty()) return fail(GW_EINVAL, t->adv_err);
  DevView v = ctx.view(t);
  if (scratch_mem) {
    v.partial = scratch_mem;
    v.means = scratch_mem + t->npartial;
  }
  if (hist) {  // keep this sub-step's level means and inlets (sub_step_replay)
    v.means = hist;
    v.inlet = hist + t->nmeans;
  }
  int rc = hip_status(launch_pipe_reduce(v, ctx.prec, phi, ctx.stream, jacc), "pipe_reduce");
  if (rc) return rc;
  if (ctx.host_inlet) {
    std::vector<double> tout(ctx.nbh_total), inl(ctx.nbh_total);
    rc = hip_status(hipMemcpyAsync(tout.data(), ctx.d_tout, sizeof(double) * ctx.nbh_total,
                                   hipMemcpyDeviceToHost, ctx.stream), "outlet copy");
    if (rc) return rc;
    rc = hip_status(hipMemcpyAsync(inl.data(), ctx.d_inlet, sizeof(double) * ctx.nbh_total,
                                   hipMemcpyDeviceToHost, ctx.stream), "inlet copy");
    if (rc) return rc;
    rc = hip_status(hipStreamSynchronize(ctx.stream), "sync");
    if (rc) return rc;
    int gb = 0;
    for (auto& s : ctx.sc) {
      const int nb = (int)s.bh.size();
      if (s.inlet.kind == GW_INLET_HOST)
        for (int b = 0; b < nb; ++b) inl[gb + b] = s.inlet.fn(b + 1, &tout[gb], nb, tau, s.inlet.user);
      gb += nb;
    }
    rc = hip_status(hipMemcpyAsync(ctx.d_inlet, inl.data(), sizeof(double) * ctx.nbh_total,
                                   hipMemcpyHostToDevice, ctx.stream), "inlet upload");
    if (rc) return rc;
    rc = hip_status(hipStreamSynchronize(ctx.stream), "sync");
    if (rc) return rc;
  }
  return hip_status(launch_pipe_advect(v, ctx.prec, phi, ctx.stream), "pipe_advect");
}

// The advection of a sub-step whose level means and inlets an earlier
// sub_step_advect kept in hist (the same state: the adjoint's recompute),
// without the reductions.
int sub_step_replay(gw_ctx& ctx, const CTab* t, void* phi, const double* hist) {
  DevView v = ctx.view(t);
  v.means = const_cast<double*>(hist);
  v.inlet = const_cast<double*>(hist) + t->nmeans;
  return hip_status(launch_pipe_advect(v, ctx.prec, phi, ctx.stream), "pipe_advect");
}

// GW_F32_STATE_SR: the stochastic rounding of step n draws random numbers
// keyed by n (the device counter d_step_ctr): a gw_steps call sets it to the
// index of its first step, t / dt, and every step advances it on the device
// (inside the step graphs too), so a run split into several calls rounds like
// one call.  The separate operator entry points advance it after each call,
// so repeated calls draw new numbers.
int set_step_ctr(gw_ctx& ctx, double t, double dt) {
  if (ctx.prec != GW_F32_STATE_SR) return GW_OK;
  const double n = std::nearbyint(t / dt);
  const uint32_t v = (uint32_t)(uint64_t)(int64_t)(std::isfinite(n) ? std::fmod(n, 4294967296.0) : 0.0);
  return hip_status(launch_step_ctr(ctx.d_step_ctr, true, v, ctx.stream), "step counter");
}
int advance_step_ctr(gw_ctx& ctx) {
  if (ctx.prec != GW_F32_STATE_SR) return GW_OK;
  return hip_status(launch_step_ctr(ctx.d_step_ctr, false, 0u, ctx.stream), "step counter");
}

// ADI_and_ADV_callback! with the tables of c = dt/2 (scratch_mem: see
// sub_step_advect).
int adi_adv_tab(gw_ctx& ctx, const CTab* tab, void* phi, double t, double dt, double* scratch_mem) {
  const double c = dt / 2;
  DevView v = ctx.view(tab);
  // explicit y  (src/solver.jl:444)
  int rc = hip_status(launch_line(v, ctx.prec, LINE_EXPLICIT, 1, ctx.lw(LINE_EXPLICIT, 1), phi, phi, c, 0.0, 0.0, ctx.stream, ctx.lsp()), "explicit y");
  if (rc) return rc;
  rc = sub_step_advect(ctx, tab, phi, t, nullptr, scratch_mem);  // :447
  if (rc) return rc;
  // implicit x then explicit x  (:450, :455)
  rc = hip_status(launch_line(v, ctx.prec, LINE_IMPLICIT_EXPLICIT, 0, ctx.lw(LINE_IMPLICIT_EXPLICIT, 0), phi, phi, c, 0.0, 0.0, ctx.stream, ctx.lsp()), "implicit+explicit x");
  if (rc) return rc;
  rc = sub_step_advect(ctx, tab, phi, t + dt / 2, nullptr, scratch_mem);  // :458
  if (rc) return rc;
  // implicit y  (:461)
  rc = hip_status(launch_line(v, ctx.prec, LINE_IMPLICIT, 1, ctx.lw(LINE_IMPLICIT, 1), phi, phi, c, 0.0, 0.0, ctx.stream, ctx.lsp()), "implicit y");
  if (rc) return rc;
  return advance_step_ctr(ctx);
}

int adi_adv(gw_ctx& ctx, void* phi, double t, double dt) {
  CTab* tab = nullptr;
  const int rc = build_ctab(ctx, dt / 2, tab);
  if (rc) return rc;
  return adi_adv_tab(ctx, tab, phi, t, dt, nullptr);
}

// The current device's copy of `scratch`, resolved once per device.
std::mutex g_scratch_mu;
std::map<int, double*> g_scratch_ptr;

int resolve_scratch(int dev, double** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  auto it = g_scratch_ptr.find(dev);
  if (it == g_scratch_ptr.end()) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return fail(GW_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    it = g_scratch_ptr.emplace(dev, static_cast<double*>(p)).first;
  }
  *out = it->second;
  return GW_OK;
}

// The z-step applies R(dt A_z) = 1 + z + z^2/2 + c3 z^3 at every ROCK2 degree:
// ROCK2's longer polynomials agree with it through z^2/2 and differ by
// table coefficients times (dt rho(A_z))^3.., negligible unless dt rho(A_z)
// is O(1), which the stability check below excludes.
int check_zstep(gw_ctx& ctx, double dt) {
  for (auto& s : ctx.sc)
    if (dt * s.zrho > ctx.zstab)
      return fail(GW_EINVAL, "dt * rho(A_z) = " + std::to_string(dt * s.zrho) +
                                 " exceeds the stability interval of the z-step polynomial");
  return GW_OK;
}

int zstep(gw_ctx& ctx, void* phi, double dt) {
  int rc = check_zstep(ctx, dt);
  if (rc) return rc;
  rc = hip_status(launch_zstep(ctx.view(nullptr), ctx.prec, phi, dt, ctx.c3, ctx.stream), "zstep");
  if (rc) return rc;
  return advance_step_ctr(ctx);
}

int create_impl(const gw_problem* p, gw_ctx** out) {
  if (!p || !out) return fail(GW_EINVAL, "null argument");
  if (p->n_scen <= 0 || !p->scen) return fail(GW_EINVAL, "no scenarios");
  if (p->precision != GW_F64 && p->precision != GW_F32_STATE && p->precision != GW_F32_STATE_SR)
    return fail(GW_EINVAL, "unknown precision");
  const int32_t width = p->flags & (GW_FLAG_LINE_W64 | GW_FLAG_LINE_W32);
  const int32_t variant = p->flags & (GW_FLAG_LINE_TWO_TILES | GW_FLAG_HALO_WHOLE | GW_FLAG_LINE_CONCURRENT);
  const int32_t steps = p->flags & (GW_FLAG_STEPS_GRAPH | GW_FLAG_STEPS_DIRECT);
  const int32_t keep_classes = p->flags & GW_FLAG_KEEP_LINE_CLASSES;
  if ((p->flags & ~(width | variant | steps | keep_classes)) || width == (GW_FLAG_LINE_W64 | GW_FLAG_LINE_W32) ||
      (variant && !width) || steps == (GW_FLAG_STEPS_GRAPH | GW_FLAG_STEPS_DIRECT))
    return fail(GW_EINVAL, "unknown or conflicting flags (the launch variants need a forced width)");
  std::unique_ptr<gw_ctx> ctx(new gw_ctx());
  ctx->steps_mode = steps == GW_FLAG_STEPS_GRAPH ? kStepsGraph : (steps == GW_FLAG_STEPS_DIRECT ? kStepsEager : kStepsUndecided);
  ctx->prec = p->precision;
  ctx->force_lw = width == GW_FLAG_LINE_W64 ? 64 : (width == GW_FLAG_LINE_W32 ? 32 : 0);
  if (ctx->force_lw) {
    if (variant & GW_FLAG_LINE_TWO_TILES) ctx->force_lw |= kLwTwoTiles;
    if (variant & GW_FLAG_HALO_WHOLE) ctx->force_lw |= kLwHaloWhole;
    if (variant & GW_FLAG_LINE_CONCURRENT) ctx->force_lw |= kLwConcurrent;
  }
  ctx->c3 = p->rock2_c3 > 0 ? p->rock2_c3 : 1.0 / 16.0;
  ctx->zstab = stability_bound(ctx->c3);
  std::unordered_map<std::string, int> smap;
  int64_t off = 0;
  for (int si = 0; si < p->n_scen; ++si) {
    const gw_scenario& q = p->scen[si];
    if (q.nx < 3 || q.ny < 3 || q.nz < 3) return fail(GW_EINVAL, "every grid needs at least 3 points");
    if (!q.gridx || !q.gridy || !q.gridz) return fail(GW_EINVAL, "null grid");
    if (q.n_bh <= 0 || !q.boreholes) return fail(GW_EINVAL, "no boreholes");
    if (q.nz > 4096) return fail(GW_EINVAL, "nz > 4096 is not supported by the z-step kernel");
    if (q.nx > 640 || q.ny > 640) return fail(GW_EINVAL, "nx or ny > 640 is not supported by the line kernels");
    Scen s;
    s.nx = q.nx; s.ny = q.ny; s.nz = q.nz;
    s.gx.assign(q.gridx, q.gridx + q.nx);
    s.gy.assign(q.gridy, q.gridy + q.ny);
    s.gz.assign(q.gridz, q.gridz + q.nz);
    for (int a = 0; a + 1 < q.nx; ++a) if (!(s.gx[a] < s.gx[a + 1])) return fail(GW_EINVAL, "gridx not increasing");
    for (int a = 0; a + 1 < q.ny; ++a) if (!(s.gy[a] < s.gy[a + 1])) return fail(GW_EINVAL, "gridy not increasing");
    for (int a = 0; a + 1 < q.nz; ++a) if (!(s.gz[a] < s.gz[a + 1])) return fail(GW_EINVAL, "gridz not increasing");
    for (int b = 0; b < q.n_bh; ++b) s.bh.push_back(make_bh(q.boreholes[b]));
    const gw_materials& m = q.materials;
    if (m.n_layers > 0) {
      if (!m.k_rock_layers || !m.rho_c_rock_layers || !m.layer_depths) return fail(GW_EINVAL, "null layer arrays");
      s.mat.kl.assign(m.k_rock_layers, m.k_rock_layers + m.n_layers);
      s.mat.rl.assign(m.rho_c_rock_layers, m.rho_c_rock_layers + m.n_layers);
      s.mat.depth.assign(m.layer_depths, m.layer_depths + m.n_layers);
    }
    s.mat.k_rock = m.k_rock; s.mat.rc_rock = m.rho_c_rock;
    s.mat.k_water = m.k_water; s.mat.rc_water = m.rho_c_water;
    s.mat.k_steel = m.k_steel; s.mat.rc_steel = m.rho_c_steel;
    s.mat.k_ins = m.k_insulating; s.mat.rc_ins = m.rho_c_insulating;
    s.mat.k_bf = m.k_backfill; s.mat.rc_bf = m.rho_c_backfill;
    s.inlet = q.inlet;
    if (q.inlet.kind < 0 || q.inlet.kind > 3) return fail(GW_EINVAL, "unknown inlet kind");
    if (q.inlet.kind == GW_INLET_PER_WELL) {
      if (!q.inlet.per_well) return fail(GW_EINVAL, "per-well inlet table is null");
      s.per_well.assign(q.inlet.per_well, q.inlet.per_well + q.n_bh);
    }
    if (q.inlet.kind == GW_INLET_HOST) {
      if (!q.inlet.fn) return fail(GW_EINVAL, "host inlet callback is null");
      ctx->host_inlet = true;
    }
    try {
      classify(s);
    } catch (std::exception& e) {
      return fail(GW_EINVAL, e.what());
    }
    index_lists(s);
    build_line_sets(*ctx, s, smap);
    s.off = off;
    off += (int64_t)s.nx * s.ny * s.nz;
    ctx->sc.push_back(std::move(s));
  }
  // pipe columns grouped per borehole, inner first
  int gb = 0;
  for (auto& s : ctx->sc) {
    for (size_t b = 0; b < s.bh.size(); ++b, ++gb) {
      ctx->bh_scen.push_back((int)(&s - &ctx->sc[0]));
      ctx->bh_col_in0.push_back((int)ctx->pcol.size());
      int n_in = 0;
      for (auto& e : s.idx_inner) if (e[2] == (int)b) { ctx->pcol.push_back(PipeCol{e[0], e[1], gb, 1}); ++n_in; }
      ctx->bh_n_in.push_back(n_in);
      ctx->bh_col_an0.push_back((int)ctx->pcol.size());
      int n_an = 0;
      for (auto& e : s.idx_outer) if (e[2] == (int)b) { ctx->pcol.push_back(PipeCol{e[0], e[1], gb, 0}); ++n_an; }
      ctx->bh_n_an.push_back(n_an);
    }
  }
  ctx->nbh_total = gb;
  // device upload
  try {
    if (p->stream) {
      ctx->stream = (hipStream_t)p->stream;
    } else {
      hip_check(hipStreamCreate(&ctx->stream));
      ctx->own_stream = true;
    }
    if (p->n_scen > 1) {  // ragged batches launch several line classes per operator
      hip_check(hipEventCreateWithFlags(&ctx->ls.fork, hipEventDisableTiming));
      for (int q = 0; q < kSideStreams; ++q) {
        hip_check(hipStreamCreateWithFlags(&ctx->ls.side[q], hipStreamNonBlocking));
        hip_check(hipEventCreateWithFlags(&ctx->ls.join[q], hipEventDisableTiming));
      }
      ctx->has_ls = true;
    }
    for (auto& s : ctx->sc) {
      ctx->nx_max = std::max(ctx->nx_max, s.nx);
      ctx->ny_max = std::max(ctx->ny_max, s.ny);
      ctx->nz_max = std::max(ctx->nz_max, s.nz);
      ctx->countz_max = std::max(ctx->countz_max, s.countz);
    }
    for (int mode = 0; mode < 5; ++mode)
      for (int dir = 0; dir < 2; ++dir) {
        const int n = dir == 0 ? ctx->nx_max : ctx->ny_max;
        ctx->lw_sel[mode][dir] = ctx->force_lw ? ctx->force_lw : line_window_default(n);
      }
    std::vector<ScenDev> sd;
    std::vector<int64_t> cellp{0}, colp{0};
    if (!keep_classes) merge_small_classes(*ctx);
    const std::vector<int64_t> pset_off = padded_set_offsets(*ctx);
    if (!pset_off.empty() && pset_off.back() > (int64_t)UINT32_MAX / 2)
      throw std::runtime_error("too many coefficient sets");
    std::vector<LineTile> xb[2][kLineClasses], yb[2][kLineClasses], hb[2][kLineClasses];
    std::vector<uint16_t> cls;
    std::vector<int32_t> zone, setid;
    std::vector<double> zcoef, inlet_tab;
    int gbh = 0;
    for (auto& s : ctx->sc) {
      ScenDev d{};
      d.nx = s.nx; d.ny = s.ny; d.nz = s.nz; d.countz = s.countz;
      d.nzone = s.nzone; d.ncls = s.ncls; d.nbh = (int)s.bh.size(); d.bh0 = gbh;
      d.off = s.off;
      d.o_cls = (int64_t)cls.size(); cls.insert(cls.end(), s.cls.begin(), s.cls.end());
      d.o_zone = (int64_t)zone.size(); zone.insert(zone.end(), s.zone.begin(), s.zone.end());
      d.o_zcoef = (int64_t)zcoef.size() / 2; zcoef.insert(zcoef.end(), s.zcoef.begin(), s.zcoef.end());
      d.o_xset = (int64_t)setid.size(); setid.insert(setid.end(), s.xset.begin(), s.xset.end());
      d.o_yset = (int64_t)setid.size(); setid.insert(setid.end(), s.yset.begin(), s.yset.end());
      sd.push_back(d);
      const int64_t ncell = (int64_t)s.nx * s.ny * s.nz;
      cellp.push_back(cellp.back() + ncell);
      colp.push_back(colp.back() + (int64_t)s.nx * s.ny);
      if (line_class(s.nx) < 0 || line_class(s.ny) < 0)
        throw std::runtime_error("grid lines longer than the line kernels support");
      const int qx = ctx->line_cls(s.nx), qy = ctx->line_cls(s.ny);
      for (int f = 0; f < 2; ++f) {
        const int lw = f == 0 ? 64 : 32;
        line_tiles(s, (int)sd.size() - 1, 0, 0, lw, pset_off, xb[f][qx]);
        line_tiles(s, (int)sd.size() - 1, 1, 0, lw, pset_off, yb[f][qy]);
        line_tiles(s, (int)sd.size() - 1, 1, 3, lw, pset_off, hb[f][qy]);
      }
      for (size_t b = 0; b < s.bh.size(); ++b)
        inlet_tab.push_back(s.inlet.kind == GW_INLET_PER_WELL ? s.per_well[b] : 0.0);
      gbh += (int)s.bh.size();
    }
    ctx->cells_total = cellp.back();
    ctx->cols_total = colp.back();
    // each family's tiles in class order, with the class offsets
    std::vector<LineTile> xt[2], yt[2], ht[2];
    auto gather = [](std::vector<LineTile>* b, std::vector<LineTile>& out, int64_t* off) {
      for (int q = 0; q < kLineClasses; ++q) {
        off[q] = (int64_t)out.size();
        out.insert(out.end(), b[q].begin(), b[q].end());
      }
      off[kLineClasses] = (int64_t)out.size();
    };
    for (int f = 0; f < 2; ++f) {
      // halo tiles of one set first within each class (launched with one slot)
      for (int q = 0; q < kLineClasses; ++q) {
        const auto mid = std::stable_partition(hb[f][q].begin(), hb[f][q].end(),
                                               [](const LineTile& t) { return t.nset == 1; });
        int m = 1;
        for (const LineTile& t : hb[f][q]) m = std::max(m, (int)t.nset);
        ctx->hslots[f][q] = m;
        ctx->hone[f][q] = (int64_t)(mid - hb[f][q].begin());  // made absolute below
      }
      gather(xb[f], xt[f], ctx->xcls[f]);
      gather(yb[f], yt[f], ctx->ycls[f]);
      gather(hb[f], ht[f], ctx->hcls[f]);
      for (int q = 0; q < kLineClasses; ++q) ctx->hone[f][q] += ctx->hcls[f][q];
    }
    ctx->d_sc = ctx->up(sd);
    ctx->d_cell_prefix = ctx->up(cellp);
    ctx->d_col_prefix = ctx->up(colp);
    for (int f = 0; f < 2; ++f) {
      ctx->d_xtiles[f] = ctx->up(xt[f]);
      ctx->d_ytiles[f] = ctx->up(yt[f]);
      ctx->d_htiles[f] = ctx->up(ht[f]);
    }
    ctx->fused_ok = true;
    for (auto& s : ctx->sc) ctx->fused_ok = ctx->fused_ok && s.ncls <= kZcMax;
    ctx->d_cls = ctx->up(cls);
    ctx->d_zone = ctx->up(zone);
    ctx->d_zcoef = ctx->up(zcoef);
    if (ctx->prec != GW_F64) {  // the FP32 64-lane kernels' z coefficients
      std::vector<float> z32(zcoef.begin(), zcoef.end());
      if (z32.empty()) z32.push_back(0.f);
      ctx->d_zcoef32 = ctx->up(z32);
    }
    ctx->d_setid = ctx->up(setid);
    ctx->d_pcol = ctx->up(ctx->pcol);
    std::vector<int64_t> pcol_off(ctx->pcol.size());
    for (size_t q = 0; q < ctx->pcol.size(); ++q) {
      const PipeCol& pc = ctx->pcol[q];
      const Scen& s = ctx->sc[ctx->bh_scen[pc.bh]];
      pcol_off[q] = ((int64_t)pc.j + (int64_t)s.ny * pc.i) * s.nz;
    }
    ctx->d_pcol_off = ctx->up(pcol_off);
    std::vector<double> zeros(std::max(ctx->nbh_total, 1), 0.0);
    ctx->d_tout = ctx->up(zeros);
    ctx->d_inlet = ctx->up(zeros);
    ctx->d_step_ctr = ctx->up(std::vector<uint32_t>(1, 0u));
    ctx->d_inlet_tab = ctx->up(inlet_tab);
  } catch (HipFail& e) {
    return fail(GW_EHIP, std::string("device set-up failed: ") + hipGetErrorString(e.e));
  }
  // the no-allocation entry point's scratch on this device; a failure here
  // only disables that entry point
  if (hipGetDevice(&ctx->device) != hipSuccess) return fail(GW_EHIP, "hipGetDevice failed");
  if (resolve_scratch(ctx->device, &ctx->scratch_dev) != GW_OK) {
    ctx->scratch_dev = nullptr;
    ctx->scratch_err = g_err;
  }
  *out = ctx.release();
  return GW_OK;
}

}  // namespace

extern "C" {

const char* gw_last_error(void) { return g_err.c_str(); }

int gw_create(const gw_problem* p, gw_ctx** out) {
  try {
    return create_impl(p, out);
  } catch (std::exception& e) {
    return fail(GW_EINVAL, e.what());
  } catch (...) {
    return fail(GW_EHIP, "unexpected failure in gw_create");
  }
}

int gw_destroy(gw_ctx* ctx) {
  if (!ctx) return GW_OK;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  delete ctx;
  return GW_OK;
}

int gw_rhs_z(gw_ctx* ctx, void* dphi, const void* phi, double /*t*/) {
  if (!ctx || !dphi || !phi) return fail(GW_EINVAL, "null argument");
  return hip_status(launch_zrhs(ctx->view(nullptr), ctx->prec, dphi, phi, ctx->stream), "rhs_z");
}

int gw_adi_adv(gw_ctx* ctx, void* phi, void* /*scratch*/, double t, double dt) {
  if (!ctx || !phi) return fail(GW_EINVAL, "null argument");
  if (!(dt > 0)) return fail(GW_EINVAL, "dt must be positive");
  return adi_adv(*ctx, phi, t, dt);
}

int gw_adi_adv_noalloc(gw_ctx* ctx, void* phi, double t, double dt) {
  if (!ctx || !phi) return fail(GW_EINVAL, "null argument");
  if (!(dt > 0)) return fail(GW_EINVAL, "dt must be positive");
  if (ctx->host_inlet)
    return fail(GW_EINVAL, "gw_adi_adv_noalloc: a host-evaluated inlet (GW_INLET_HOST) synchronises the stream "
                           "every sub-step; use gw_adi_adv");
  if (!ctx->scratch_dev) return fail(GW_EHIP, "gw_adi_adv_noalloc: no scratch on this device (" + ctx->scratch_err + ")");
  const auto it = ctx->ctabs.find(dt / 2);
  if (it == ctx->ctabs.end())
    return fail(GW_ESTATE, "gw_adi_adv_noalloc: no tables for this dt yet; one allocating call with the same dt "
                           "(gw_adi_adv, gw_step, gw_steps) builds them");
  const CTab* tab = &it->second;
  const size_t need = (size_t)tab->npartial + (size_t)tab->nmeans;
  if (need > kNoAllocScratchDoubles)
    return fail(GW_EINVAL, "gw_adi_adv_noalloc: the pipe reductions need " + std::to_string(need) +
                               " doubles of scratch, more than its " + std::to_string(kNoAllocScratchDoubles) +
                               " (8 MiB); use gw_adi_adv");
  return adi_adv_tab(*ctx, tab, phi, t, dt, ctx->scratch_dev);
}

int gw_zstep(gw_ctx* ctx, void* phi, double dt) {
  if (!ctx || !phi) return fail(GW_EINVAL, "null argument");
  if (!(dt > 0)) return fail(GW_EINVAL, "dt must be positive");
  return zstep(*ctx, phi, dt);
}

}  // extern "C"

namespace {

// Fused step pipeline.  Steps n .. n+k-1 are the launch sequence
//   ZY                                    (z-step + explicit y of step n)
//   [adv, X, adv, YZY] for all but the last step
//   adv, X, adv, Yimpl                    (the last step's implicit y)
// where X = implicit x + explicit x and YZY = implicit y of one step, then
// the z-step and explicit y of the next; the field is read and written twice
// per step instead of four times, with the arithmetic of the separate
// operators.  ZY and YZY read a 3-level halo of neighbouring k-windows, so
// they write out of place: the state alternates between the caller's field
// and the context's second buffer (ZY: phi -> buf2, the step ending in YZY
// or Yimpl moves it back), and the last step's implicit y always lands in phi.
int ensure_buf2(gw_ctx& ctx) {
  if (ctx.buf2) return GW_OK;
  const size_t bytes = (size_t)ctx.cells_total * (ctx.prec == GW_F64 ? 8 : 4);
  if (hipMalloc(&ctx.buf2, bytes) != hipSuccess) {
    ctx.buf2 = nullptr;
    return fail(GW_ENOMEM, "cannot allocate the second field buffer of the fused step pipeline");
  }
  return GW_OK;
}

int fused_zy(gw_ctx& ctx, const CTab* tab, const void* in, void* out, double dt) {
  return hip_status(launch_line(ctx.view(tab), ctx.prec, LINE_ZY, 1, ctx.lw(LINE_ZY, 1), out, in, dt / 2, dt, ctx.c3, ctx.stream, ctx.lsp()),
                    "z-step + explicit y");
}

// One step's advection/x part (in place on `cur`) and its closing y kernel
// from `cur` into `next` (YZY, or the plain implicit y when `last`).
// hist (may be NULL): keeps the two sub-steps' level means and inlets at hist
// and hist + (nmeans + nbh).
int fused_rest(gw_ctx& ctx, const CTab* tab, void* cur, void* next, double t_end, double dt, bool last,
               double* jacc = nullptr, double* hist = nullptr) {
  const double c = dt / 2;
  DevView v = ctx.view(tab);
  const size_t hs = (size_t)tab->nmeans + (size_t)ctx.nbh_total;
  int rc = sub_step_advect(ctx, tab, cur, t_end, nullptr, nullptr, hist);
  if (rc) return rc;
  rc = hip_status(launch_line(v, ctx.prec, LINE_IMPLICIT_EXPLICIT, 0, ctx.lw(LINE_IMPLICIT_EXPLICIT, 0), cur, cur, c, 0.0, 0.0, ctx.stream, ctx.lsp()), "x");
  if (rc) return rc;
  rc = sub_step_advect(ctx, tab, cur, t_end + c, jacc, nullptr, hist ? hist + hs : nullptr);
  if (rc) return rc;
  if (last)
    rc = hip_status(launch_line(v, ctx.prec, LINE_IMPLICIT, 1, ctx.lw(LINE_IMPLICIT, 1), next, cur, c, 0.0, 0.0, ctx.stream, ctx.lsp()), "implicit y");
  else
    rc = hip_status(launch_line(v, ctx.prec, LINE_YZY, 1, ctx.lw(LINE_YZY, 1), next, cur, c, dt, ctx.c3, ctx.stream, ctx.lsp()),
                    "implicit y + z-step + explicit y");
  if (rc) return rc;
  return advance_step_ctr(ctx);
}

// Line-tile autotune: each line operator of the fused pipeline is timed in
// both k-window widths, each with its natural LDS and capped at two tiles per
// CU, and for the fused z-step modes with the tiles split by set count and
// whole (kLwHaloWhole), out of place from the caller's field into the second
// buffer (the field is only read); the fastest is kept for the context.  Runs
// once, before the first step graph is captured.
int tune_lines(gw_ctx& ctx, const CTab* tab, const void* phi, double dt) {
  static const int ops[4][2] = {{LINE_ZY, 1}, {LINE_IMPLICIT_EXPLICIT, 0}, {LINE_YZY, 1}, {LINE_IMPLICIT, 1}};
  const DevView v = ctx.view(tab);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = hip_status(hipEventCreate(&e0), "event");
  if (!rc) rc = hip_status(hipEventCreate(&e1), "event");
  for (int q = 0; q < 4 && !rc; ++q) {
    const int mode = ops[q][0], dir = ops[q][1];
    float best = 0.f;
    int best_lw = ctx.lw_sel[mode][dir];
    const bool halo = mode == LINE_ZY || mode == LINE_YZY;
    // batched contexts also time the class launches of an operator one after
    // the other and concurrently on side streams (schedule 0, 1)
    const int nb = halo ? 3 : 2;
    const int nvar = (1 << nb) * (ctx.has_ls ? 2 : 1);
    for (int var0 = 0; var0 < nvar; ++var0) {
      const int var = var0 & ((1 << nb) - 1), sched = var0 >> nb;
      const int lw = (var & 1 ? 32 : 64) | (var & 2 ? kLwTwoTiles : 0) | (var & 4 ? kLwHaloWhole : 0) |
                     (sched == 1 ? kLwConcurrent : 0);
      auto run = [&]() {
        return hip_status(launch_line(v, ctx.prec, mode, dir, lw, ctx.buf2, phi, dt / 2, dt, ctx.c3, ctx.stream, ctx.lsp()),
                          "line tune");
      };
      if ((rc = run())) break;  // first launch: code object load
      // Fastest of five separately timed launches: a single slow launch
      // (clock or cache state) must not decide the variant, and small fields
      // launch in tens of microseconds, close to the timing noise.
      float ms = 0.f;
      for (int r = 0; r < 5 && !rc; ++r) {
        if ((rc = hip_status(hipEventRecord(e0, ctx.stream), "event"))) break;
        if ((rc = run())) break;
        if ((rc = hip_status(hipEventRecord(e1, ctx.stream), "event"))) break;
        if ((rc = hip_status(hipEventSynchronize(e1), "event sync"))) break;
        float t = 0.f;
        if ((rc = hip_status(hipEventElapsedTime(&t, e0, e1), "event time"))) break;
        if (r == 0 || t < ms) ms = t;
      }
      if (rc) break;
      if (best == 0.f || ms < best) { best = ms; best_lw = lw; }
    }
    ctx.lw_sel[mode][dir] = best_lw;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  return rc;
}

int prepare_steps(gw_ctx& ctx, const void* phi, double dt, CTab*& tab) {
  int rc = build_ctab(ctx, dt / 2, tab);
  if (rc) return rc;
  rc = check_zstep(ctx, dt);
  if (rc) return rc;
  if ((rc = ensure_buf2(ctx))) return rc;
  if (!ctx.tuned && !ctx.force_lw) {
    if ((rc = tune_lines(ctx, tab, phi, dt))) return rc;
    ctx.tuned = true;
  }
  return GW_OK;
}

int capture_body(gw_ctx& ctx, const CTab* tab, void* cur, void* next, double t, double dt, hipGraphExec_t* out) {
  hipGraph_t g = nullptr;
  int rc = hip_status(hipStreamBeginCapture(ctx.stream, hipStreamCaptureModeThreadLocal), "begin capture");
  if (rc) return rc;
  const int rc1 = fused_rest(ctx, tab, cur, next, t, dt, false);
  const hipError_t e2 = hipStreamEndCapture(ctx.stream, &g);
  if (rc1) {
    if (g) (void)hipGraphDestroy(g);
    return rc1;
  }
  if ((rc = hip_status(e2, "end capture"))) return rc;
  rc = hip_status(hipGraphInstantiate(out, g, nullptr, nullptr, 0), "graph instantiate");
  (void)hipGraphDestroy(g);
  return rc;
}

}  // namespace

extern "C" {

int gw_step(gw_ctx* ctx, void* phi, double t, double dt) {
  return gw_steps(ctx, phi, t, dt, 1);
}

// Steps 2 .. nsteps-1 of a gw_steps call replay one captured graph per step
// or launch the same kernels directly (identical results).  Unless a flag
// forces one, the first call of at least 2 kStepTune + 3 steps times
// kStepTune steps of each, after one of each (steps 1 .. 2 kStepTune + 2 of
// the call; its last step, which ends with the implicit y alone, follows),
// and keeps the faster: on large
// fields the replays measured 1-7% slower per step than direct launches on
// MI355X (ROCm 7), on small fields the launches' host cost can decide.
constexpr int kStepTune = 3;

int gw_steps(gw_ctx* ctx, void* phi, double t, double dt, int32_t nsteps) {
  if (!ctx || !phi) return fail(GW_EINVAL, "null argument");
  if (!(dt > 0) || nsteps < 0) return fail(GW_EINVAL, "dt must be positive and nsteps >= 0");
  if (nsteps == 0) return GW_OK;
  if (!ctx->fused_ok) {
    for (int n = 0; n < nsteps; ++n) {
      int rc = zstep(*ctx, phi, dt);
      if (rc) return rc;
      if ((rc = adi_adv(*ctx, phi, t + (n + 1) * dt, dt))) return rc;
    }
    return GW_OK;
  }
  CTab* tab = nullptr;
  int rc = prepare_steps(*ctx, phi, dt, tab);
  if (rc) return rc;
  if (!tab->adv_err.empty()) return fail(GW_EINVAL, tab->adv_err);
  if ((rc = set_step_ctr(*ctx, t, dt))) return rc;
  void* B = ctx->buf2;
  rc = fused_zy(*ctx, tab, phi, B, dt);
  if (rc) return rc;
  // step k (1-based) works on B when k is odd, on phi when k is even
  const bool graph_ok = !ctx->host_inlet && nsteps > 2 && ctx->steps_mode != kStepsEager;
  const bool tune = graph_ok && ctx->steps_mode == kStepsUndecided && nsteps >= 2 * kStepTune + 3;
  if (graph_ok && (!ctx->graph_exec || ctx->graph_phi != phi || ctx->graph_dt != dt)) {
    if (ctx->graph_exec) { (void)hipGraphExecDestroy(ctx->graph_exec); ctx->graph_exec = nullptr; }
    if (ctx->graph_exec2) { (void)hipGraphExecDestroy(ctx->graph_exec2); ctx->graph_exec2 = nullptr; }
    if ((rc = capture_body(*ctx, tab, B, phi, t + dt, dt, &ctx->graph_exec2))) return rc;  // B -> phi
    if ((rc = capture_body(*ctx, tab, phi, B, t + dt, dt, &ctx->graph_exec))) return rc;   // phi -> B
    ctx->graph_phi = phi;
    ctx->graph_dt = dt;
  }
  auto step = [&](int k, bool graph) -> int {
    const bool odd = (k & 1) != 0;
    if (graph) return hip_status(hipGraphLaunch(odd ? ctx->graph_exec2 : ctx->graph_exec, ctx->stream), "graph launch");
    return fused_rest(*ctx, tab, odd ? B : phi, odd ? phi : B, t + k * dt, dt, false);
  };
  int k = 1;
  if (tune) {
    hipEvent_t ev[3] = {};
    for (auto& e : ev)
      if (!rc) rc = hip_status(hipEventCreate(&e), "event");
    if (!rc) rc = step(k++, false);
    if (!rc) rc = step(k++, true);
    if (!rc) rc = hip_status(hipEventRecord(ev[0], ctx->stream), "event");
    for (int q = 0; q < kStepTune && !rc; ++q) rc = step(k++, false);
    if (!rc) rc = hip_status(hipEventRecord(ev[1], ctx->stream), "event");
    for (int q = 0; q < kStepTune && !rc; ++q) rc = step(k++, true);
    if (!rc) rc = hip_status(hipEventRecord(ev[2], ctx->stream), "event");
    if (!rc) rc = hip_status(hipEventSynchronize(ev[2]), "event sync");
    float direct = 0.f, graph = 0.f;
    if (!rc) rc = hip_status(hipEventElapsedTime(&direct, ev[0], ev[1]), "event time");
    if (!rc) rc = hip_status(hipEventElapsedTime(&graph, ev[1], ev[2]), "event time");
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    if (rc) return rc;
    ctx->steps_mode = graph < direct ? kStepsGraph : kStepsEager;
  }
  const bool use_graph = graph_ok && ctx->steps_mode != kStepsEager;
  for (; k < nsteps; ++k)
    if ((rc = step(k, use_graph))) return rc;
  void* cur = (nsteps & 1) ? B : phi;
  return fused_rest(*ctx, tab, cur, phi, t + nsteps * dt, dt, true);
}

int gw_line_op(gw_ctx* ctx, void* out, const void* in, double c, double dt_z, int32_t dir, int32_t mode) {
  if (!ctx || !out || !in || (dir != 0 && dir != 1) || mode < 0 || mode > 4) return fail(GW_EINVAL, "bad argument");
  if (mode >= LINE_ZY) {
    if (dir != 1) return fail(GW_EINVAL, "the fused z-step modes run on y-lines");
    if (out == in) return fail(GW_EINVAL, "the fused z-step modes need out != in");
    if (!ctx->fused_ok) return fail(GW_ESTATE, "too many column classes for the fused z-step kernels");
    const int rc = check_zstep(*ctx, dt_z);
    if (rc) return rc;
  }
  CTab* tab = nullptr;
  int rc = build_ctab(*ctx, c, tab);
  if (rc) return rc;
  rc = hip_status(launch_line(ctx->view(tab), ctx->prec, mode, dir, ctx->lw(mode, dir), out, in, c, dt_z, ctx->c3, ctx->stream, ctx->lsp()),
                  "line operator");
  if (rc) return rc;
  return advance_step_ctr(*ctx);
}

int gw_outlet(gw_ctx* ctx, double* T_out) {
  if (!ctx || !T_out) return fail(GW_EINVAL, "null argument");
  int rc = hip_status(hipMemcpyAsync(T_out, ctx->d_tout, sizeof(double) * ctx->nbh_total,
                                     hipMemcpyDeviceToHost, ctx->stream), "outlet copy");
  if (rc) return rc;
  return hip_status(hipStreamSynchronize(ctx->stream), "sync");
}

int gw_set_inlet(gw_ctx* ctx, const double* T_in) {
  if (!ctx || !T_in) return fail(GW_EINVAL, "null argument");
  int gb = 0;
  for (auto& s : ctx->sc) {
    if (s.inlet.kind == GW_INLET_PER_WELL)
      for (size_t b = 0; b < s.bh.size(); ++b) s.per_well[b] = T_in[gb + b];
    gb += (int)s.bh.size();
  }
  std::vector<double> tab(ctx->nbh_total, 0.0);
  gb = 0;
  for (auto& s : ctx->sc) {
    for (size_t b = 0; b < s.bh.size(); ++b) tab[gb + b] = s.inlet.kind == GW_INLET_PER_WELL ? s.per_well[b] : 0.0;
    gb += (int)s.bh.size();
  }
  int rc = hip_status(hipMemcpyAsync(ctx->d_inlet_tab, tab.data(), sizeof(double) * tab.size(),
                                     hipMemcpyHostToDevice, ctx->stream), "inlet table");
  if (rc) return rc;
  return hip_status(hipStreamSynchronize(ctx->stream), "sync");
}

int gw_explicit(gw_ctx* ctx, void* out, const void* in, double c, int32_t dir) {
  if (!ctx || !out || !in || (dir != 0 && dir != 1)) return fail(GW_EINVAL, "bad argument");
  CTab* tab = nullptr;
  int rc = build_ctab(*ctx, c, tab);
  if (rc) return rc;
  rc = hip_status(launch_line(ctx->view(tab), ctx->prec, LINE_EXPLICIT, dir, ctx->lw(LINE_EXPLICIT, dir), out, in, c, 0.0, 0.0, ctx->stream, ctx->lsp()), "explicit");
  if (rc) return rc;
  return advance_step_ctr(*ctx);
}

int gw_implicit(gw_ctx* ctx, void* out, const void* in, double c, int32_t dir) {
  if (!ctx || !out || !in || (dir != 0 && dir != 1)) return fail(GW_EINVAL, "bad argument");
  CTab* tab = nullptr;
  int rc = build_ctab(*ctx, c, tab);
  if (rc) return rc;
  rc = hip_status(launch_line(ctx->view(tab), ctx->prec, LINE_IMPLICIT, dir, ctx->lw(LINE_IMPLICIT, dir), out, in, c, 0.0, 0.0, ctx->stream, ctx->lsp()), "implicit");
  if (rc) return rc;
  return advance_step_ctr(*ctx);
}

int gw_advect(gw_ctx* ctx, void* phi, double c, double t) {
  if (!ctx || !phi) return fail(GW_EINVAL, "null argument");
  CTab* tab = nullptr;
  int rc = build_ctab(*ctx, c, tab);
  if (rc) return rc;
  return sub_step_advect(*ctx, tab, phi, t);
}

int64_t gw_state_elems(gw_ctx* ctx) { return ctx ? ctx->cells_total : -1; }

int64_t gw_scenario_offset(gw_ctx* ctx, int32_t s) {
  if (!ctx || s < 0 || s >= (int)ctx->sc.size()) return -1;
  return ctx->sc[s].off;
}

int gw_counts(gw_ctx* ctx, int32_t s, int32_t* countz, int32_t* n_inner, int32_t* n_outer) {
  if (!ctx || s < 0 || s >= (int)ctx->sc.size()) return fail(GW_EINVAL, "bad scenario");
  if (countz) *countz = ctx->sc[s].countz;
  if (n_inner) *n_inner = (int32_t)ctx->sc[s].idx_inner.size();
  if (n_outer) *n_outer = (int32_t)ctx->sc[s].idx_outer.size();
  return GW_OK;
}

int gw_index_lists(gw_ctx* ctx, int32_t s, int32_t* inner, int32_t* outer) {
  if (!ctx || s < 0 || s >= (int)ctx->sc.size()) return fail(GW_EINVAL, "bad scenario");
  const Scen& sc = ctx->sc[s];
  if (inner)
    for (size_t c = 0; c < sc.idx_inner.size(); ++c)
      for (int q = 0; q < 3; ++q) inner[3 * c + q] = sc.idx_inner[c][q];
  if (outer)
    for (size_t c = 0; c < sc.idx_outer.size(); ++c)
      for (int q = 0; q < 3; ++q) outer[3 * c + q] = sc.idx_outer[c][q];
  return GW_OK;
}

double gw_eigen_estimate(gw_ctx* ctx, int32_t s) {
  if (!ctx || s < 0 || s >= (int)ctx->sc.size()) return -1.0;
  return ctx->sc[s].eigen;
}

int gw_pipe_level_means(gw_ctx* ctx, const void* phi, int32_t s, double* T_inner, double* T_outer, int32_t cap,
                        int32_t* nlev_out) {
  if (!ctx || s < 0 || s >= (int)ctx->sc.size()) return fail(GW_EINVAL, "bad scenario");
  const Scen& sc = ctx->sc[s];
  int nlev = 0;  // Nz_till_h = sum(gridz .< h) of the scenario's first borehole
  for (double z : sc.gz) if (z < sc.bh[0].h) ++nlev;
  if (nlev_out) *nlev_out = nlev;
  if (!T_inner && !T_outer) return GW_OK;
  if (!phi || !T_inner || !T_outer || cap < nlev) return fail(GW_EINVAL, "null argument or output shorter than the level count");
  const int n_in = (int)sc.idx_inner.size(), n_an = (int)sc.idx_outer.size();
  if (n_in == 0 || n_an == 0) return fail(GW_EINVAL, "empty pipe index list: its mean is undefined");
  try {
    auto it = ctx->post_off.find(s);
    if (it == ctx->post_off.end()) {
      std::vector<int64_t> off;
      for (auto& e : sc.idx_inner) off.push_back(sc.off + ((int64_t)e[1] + (int64_t)sc.ny * e[0]) * sc.nz);
      for (auto& e : sc.idx_outer) off.push_back(sc.off + ((int64_t)e[1] + (int64_t)sc.ny * e[0]) * sc.nz);
      it = ctx->post_off.emplace(s, dev_upload(off)).first;
    }
    if (!ctx->post_buf) hip_check(hipMalloc(&ctx->post_buf, sizeof(double) * 2 * std::max(ctx->nz_max, ctx->nx_max)));
  } catch (HipFail& e) {
    return fail(GW_ENOMEM, std::string("device allocation failed: ") + hipGetErrorString(e.e));
  }
  int rc = hip_status(launch_level_means(ctx->prec, phi, ctx->post_off[s], n_in, n_an, nlev, ctx->post_buf,
                                         ctx->stream), "level means");
  if (rc) return rc;
  if ((rc = hip_status(hipMemcpyAsync(T_inner, ctx->post_buf, sizeof(double) * nlev, hipMemcpyDeviceToHost,
                                      ctx->stream), "copy")))
    return rc;
  if ((rc = hip_status(hipMemcpyAsync(T_outer, ctx->post_buf + nlev, sizeof(double) * nlev, hipMemcpyDeviceToHost,
                                      ctx->stream), "copy")))
    return rc;
  return hip_status(hipStreamSynchronize(ctx->stream), "sync");
}

int gw_extract_x_profile(const void* T, int32_t precision, int32_t nx, int32_t ny, int32_t nz, const double* gridy,
                         const double* gridz, double z_target, double yc, double* profile, void* stream) {
  if (!T || !gridy || !gridz || !profile || nx <= 0 || ny <= 0 || nz <= 0 ||
      (precision != GW_F64 && precision != GW_F32_STATE && precision != GW_F32_STATE_SR))
    return fail(GW_EINVAL, "bad argument");
  // y index closest to yc (the first of equals), src/util.jl:276
  int j = 0;
  for (int q = 1; q < ny; ++q)
    if (std::fabs(gridy[q] - yc) < std::fabs(gridy[j] - yc)) j = q;
  // bracketing levels and weights, src/util.jl:279-303
  int k_lo = 0, k_hi = 0;
  double w_lo = 1.0, w_hi = 0.0;
  if (z_target <= gridz[0]) {
    k_lo = k_hi = 0;
  } else if (z_target >= gridz[nz - 1]) {
    k_lo = k_hi = nz - 1;
  } else {
    k_lo = 0;
    for (int q = 0; q < nz; ++q) if (gridz[q] <= z_target) k_lo = q;    // findlast(gridz .<= z)
    k_hi = nz - 1;
    for (int q = nz - 1; q >= 0; --q) if (gridz[q] >= z_target) k_hi = q;  // findfirst(gridz .>= z)
    if (k_lo != k_hi) {
      w_hi = (z_target - gridz[k_lo]) / (gridz[k_hi] - gridz[k_lo]);
      w_lo = 1.0 - w_hi;
    }
  }
  double* d = nullptr;
  if (hipMalloc(&d, sizeof(double) * nx) != hipSuccess) return fail(GW_ENOMEM, "device allocation failed");
  const hipStream_t st = static_cast<hipStream_t>(stream);
  int rc = hip_status(launch_x_profile(precision, T, nx, ny, nz, j, k_lo, k_hi, w_lo, w_hi, d, st), "x profile");
  if (!rc) rc = hip_status(hipMemcpyAsync(profile, d, sizeof(double) * nx, hipMemcpyDeviceToHost, st), "copy");
  if (!rc) rc = hip_status(hipStreamSynchronize(st), "sync");
  (void)hipFree(d);
  return rc;
}

int gw_sync(gw_ctx* ctx) {
  if (!ctx) return fail(GW_EINVAL, "null argument");
  return hip_status(hipStreamSynchronize(ctx->stream), "sync");
}

int gw_steps_mode(gw_ctx* ctx) {
  if (!ctx) return fail(GW_EINVAL, "null argument");
  return ctx->steps_mode;
}

int gw_line_launch(gw_ctx* ctx, int32_t op) {
  if (!ctx || op < 0 || op > 3) return fail(GW_EINVAL, "bad argument");
  static const int ops[4][2] = {{LINE_ZY, 1}, {LINE_IMPLICIT_EXPLICIT, 0}, {LINE_YZY, 1}, {LINE_IMPLICIT, 1}};
  const int lw = ctx->lw(ops[op][0], ops[op][1]);
  return (lw & ~kLwFlags) | (lw & kLwTwoTiles ? GW_FLAG_LINE_TWO_TILES : 0) |
         (lw & kLwHaloWhole ? GW_FLAG_HALO_WHOLE : 0) | (lw & kLwConcurrent ? GW_FLAG_LINE_CONCURRENT : 0);
}

void* gw_stream(gw_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int gw_profile_steps(gw_ctx* ctx, void* phi, double t, double dt, int32_t nsteps, double* ms) {
  if (!ctx || !phi || !ms || nsteps < 0) return fail(GW_EINVAL, "bad argument");
  if (!(dt > 0)) return fail(GW_EINVAL, "dt must be positive");
  for (int q = 0; q < 6; ++q) ms[q] = 0.0;
  if (!ctx->fused_ok) return fail(GW_ESTATE, "profiling follows the fused step pipeline, unavailable here");
  if (nsteps == 0) return GW_OK;
  CTab* tab = nullptr;
  int rc = prepare_steps(*ctx, phi, dt, tab);
  if (rc) return rc;
  if (!tab->adv_err.empty()) return fail(GW_EINVAL, tab->adv_err);
  // Events 0..1 bracket the first z-step; step n records events 2+4n ..
  // 2+4n+4 (the YZY's end is the next step's start).  All are recorded
  // before one synchronisation at the end, so the stream runs the steps back
  // to back as in a multi-step call.
  const int nev = 2 + 4 * nsteps + 1;
  std::vector<hipEvent_t> ev(nev, nullptr);
  for (auto& e : ev)
    if ((rc = hip_status(hipEventCreate(&e), "event create"))) break;
  const double c = dt / 2;
  DevView v = ctx->view(tab);
  auto rec = [&](int q) { return hip_status(hipEventRecord(ev[q], ctx->stream), "event record"); };
  auto acc = [&](int a, int b, int q) -> int {
    float f = 0.0f;
    const int r = hip_status(hipEventElapsedTime(&f, ev[a], ev[b]), "elapsed");
    ms[q] += f;
    return r;
  };
  do {
    if (rc) break;
    void* B = ctx->buf2;
    if ((rc = set_step_ctr(*ctx, t, dt))) break;
    if ((rc = rec(0)) || (rc = fused_zy(*ctx, tab, phi, B, dt)) || (rc = rec(1))) break;
    if ((rc = rec(2))) break;
    for (int n = 0; n < nsteps && rc == GW_OK; ++n) {
      const double te = t + (n + 1) * dt;
      const bool last = n + 1 == nsteps;
      void* cur = (n & 1) ? phi : B;
      void* next = last ? phi : ((n & 1) ? B : phi);
      const int e0 = 2 + 4 * n;
      if ((rc = sub_step_advect(*ctx, tab, cur, te)) || (rc = rec(e0 + 1))) break;
      if ((rc = hip_status(launch_line(v, ctx->prec, LINE_IMPLICIT_EXPLICIT, 0, ctx->lw(LINE_IMPLICIT_EXPLICIT, 0), cur, cur, c, 0.0, 0.0, ctx->stream, ctx->lsp()),
                           "x")) || (rc = rec(e0 + 2)))
        break;
      if ((rc = sub_step_advect(*ctx, tab, cur, te + c)) || (rc = rec(e0 + 3))) break;
      rc = last ? hip_status(launch_line(v, ctx->prec, LINE_IMPLICIT, 1, ctx->lw(LINE_IMPLICIT, 1), next, cur, c, 0.0, 0.0, ctx->stream, ctx->lsp()), "y")
                : hip_status(launch_line(v, ctx->prec, LINE_YZY, 1, ctx->lw(LINE_YZY, 1), next, cur, c, dt, ctx->c3, ctx->stream, ctx->lsp()), "yzy");
      if (rc || (rc = advance_step_ctr(*ctx)) || (rc = rec(e0 + 4))) break;
    }
    if (rc || (rc = hip_status(hipEventSynchronize(ev[nev - 1]), "event sync"))) break;
    if ((rc = acc(0, 1, 0))) break;
    for (int n = 0; n < nsteps && rc == GW_OK; ++n) {
      const int e0 = 2 + 4 * n;
      const bool last = n + 1 == nsteps;
      if ((rc = acc(e0, e0 + 1, 1)) || (rc = acc(e0 + 1, e0 + 2, 2)) || (rc = acc(e0 + 2, e0 + 3, 3)) ||
          (rc = acc(e0 + 3, e0 + 4, last ? 5 : 4)))
        break;
    }
  } while (false);
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Adjoint of the outlet objective (one scenario, FP64).
//
//   J = (1/nsteps) sum_n sum_b w_b T_out_b(n),   T_out_b(n) the outlet of
//   step n's second advection sub-step,
//
// differentiated with respect to the rock conductivity (per layer, or k_rock)
// and the inlet temperatures, by reverse sweeps over checkpointed segments:
// the forward run (the fused pipeline) keeps the state every `seg` steps;
// each segment is recomputed with the separate operators, keeping u_n, the
// z-step result z_n and the implicit-x result e_n of every step, and then
// swept backwards.  The state adjoint runs in mu = lambda / W (see AdjGeo).
// ---------------------------------------------------------------------------
namespace {

// Device buffers of one call, freed on every exit path.
struct DevBufs {
  std::vector<void*> p;
  ~DevBufs() {
    for (void* q : p) (void)hipFree(q);
  }
  template <class T>
  T* alloc(size_t n) {
    void* q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) throw std::bad_alloc();
    p.push_back(q);
    return static_cast<T*>(q);
  }
  template <class T>
  T* up(const std::vector<T>& h) {
    T* d = alloc<T>(h.size());
    if (!h.empty() && hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
      throw HipFail{hipErrorInvalidValue};
    return d;
  }
};

// Control lengths and inverse face spacings of a grid (the denominators of
// line_coefficients: h_l = (dp + dm) / 2 with mirrored ends).
void grid_h(const std::vector<double>& g, std::vector<double>& h, std::vector<double>& inv) {
  const int N = (int)g.size();
  h.assign(N, 0.0);
  inv.assign(N, 0.0);
  for (int l = 0; l < N; ++l) {
    const double dp = (l == N - 1) ? (g[l] - g[l - 1]) : (g[l + 1] - g[l]);
    const double dm = (l == 0) ? dp : (g[l] - g[l - 1]);
    h[l] = (dp + dm) * 0.5;
    if (l + 1 < N) inv[l] = 1.0 / (g[l + 1] - g[l]);
  }
}

// Geometry, materials and zeroed gradient buffers of the adjoint kernels.
AdjGeo make_adj_geo(const gw_ctx& ctx, DevBufs& db) {
  const Scen& s = ctx.sc[0];
  std::vector<double> hx, hy, hz, ix, iy, iz;
  grid_h(s.gx, hx, ix);
  grid_h(s.gy, hy, iy);
  grid_h(s.gz, hz, iz);
  std::vector<double> rock(s.rockprof.begin(), s.rockprof.end());
  AdjGeo ag{};
  ag.hx = db.up(hx); ag.hy = db.up(hy); ag.hz = db.up(hz);
  ag.ix = db.up(ix); ag.iy = db.up(iy); ag.iz = db.up(iz);
  ag.rho = db.up(s.rprof);
  std::vector<double> hzrho(s.rprof.size());
  for (size_t q = 0; q < hzrho.size(); ++q) hzrho[q] = hz[q % (size_t)s.nz] * s.rprof[q];
  ag.hzrho = db.up(hzrho);
  ag.rock = db.up(rock);
  ag.yparts = adj_yparts(ctx.hcls[0][kLineClasses], s.ny);
  ag.ncls = s.ncls;
  ag.gx_part = db.up(std::vector<double>((size_t)s.ny * s.nz, 0.0));
  ag.gy_part = db.up(std::vector<double>((size_t)ag.yparts * s.nx * s.nz, 0.0));
  return ag;
}

// Transposed advection tables of one step length (departure table of c).
AdjPipe build_adj_pipe(const gw_ctx& ctx, const CTab& tab, DevBufs& db) {
  const Scen& s = ctx.sc[0];
  const int nl = s.countz;
  std::vector<AdjBh> abh(ctx.nbh_total);
  std::vector<int32_t> off, toff;
  std::vector<AdjEnt> ent, tent;
  std::vector<RedChunk> ch;
  int poff = 0;
  constexpr int kChunk = kRedChunk;
  for (int b = 0; b < ctx.nbh_total; ++b) {
    const BhDev& B = tab.h_bh[b];
    AdjBh& A = abh[b];
    auto dep = [&](int k, int kind) { return tab.h_dep[(size_t)B.dep_off + 2 * (size_t)k + kind]; };
    // gather lists: lambda_in[k'] += w lambda_out[k] for KEEP (k' = k) and OWN (k' = kl, kr)
    for (int kind = 0; kind < 2; ++kind) {
      std::vector<std::vector<AdjEnt>> lists(nl);
      for (int k = 0; k < nl; ++k) {
        const Dep e = dep(k, kind);
        if (e.type == DEP_KEEP) {
          lists[k].push_back(AdjEnt{k, 0, 1.0});
        } else if (e.type == DEP_OWN) {
          lists[e.kl].push_back(AdjEnt{k, 0, 1.0 - e.a});
          lists[e.kr].push_back(AdjEnt{k, 0, e.a});
        }
      }
      (kind == 0 ? A.csr_in : A.csr_an) = (int32_t)off.size();
      for (int k = 0; k < nl; ++k) {
        off.push_back((int32_t)ent.size());
        ent.insert(ent.end(), lists[k].begin(), lists[k].end());
      }
      off.push_back((int32_t)ent.size());
    }
    // level ranges of the turnaround (inner) and inlet (annulus) departures
    int t0 = nl, t1 = -1, i0 = nl, i1 = -1;
    for (int k = 0; k < nl; ++k) {
      if (dep(k, 0).type == DEP_TURN) { t0 = std::min(t0, k); t1 = std::max(t1, k); }
      if (dep(k, 1).type == DEP_INLET) { i0 = std::min(i0, k); i1 = std::max(i1, k); }
    }
    A.t0 = t1 >= 0 ? t0 : 0;
    A.nt = t1 >= 0 ? t1 - t0 + 1 : 0;
    A.i0 = i1 >= 0 ? i0 : 0;
    A.ni = i1 >= 0 ? i1 - i0 + 1 : 0;
    // turnaround transpose over the annulus mean levels klo .. klo+nlev-1
    std::vector<std::vector<AdjEnt>> tl(std::max(B.nlev, 0));
    for (int k = 0; k < nl; ++k) {
      const Dep e = dep(k, 0);
      if (e.type != DEP_TURN) continue;
      tl[e.kl - B.klo].push_back(AdjEnt{k, 0, 1.0 - e.a});
      tl[e.kr - B.klo].push_back(AdjEnt{k, 0, e.a});
    }
    A.turn_off = (int32_t)toff.size();
    for (int q = 0; q < B.nlev; ++q) {
      toff.push_back((int32_t)tent.size());
      tent.insert(tent.end(), tl[q].begin(), tl[q].end());
    }
    toff.push_back((int32_t)tent.size());
    // reduction chunks: inner columns over the turnaround levels, annulus
    // columns over the inlet levels
    A.red0 = (int32_t)ch.size();
    if (A.nt > 0)
      for (int c0 = 0; c0 < B.n_in; c0 += kChunk) {
        ch.push_back(RedChunk{b, 0, B.col_in0 + c0, std::min(kChunk, B.n_in - c0), poff, 0});
        poff += A.nt;
      }
    A.nred_in = (int32_t)ch.size() - A.red0;
    if (A.ni > 0)
      for (int c0 = 0; c0 < B.n_an; c0 += kChunk) {
        ch.push_back(RedChunk{b, 1, B.col_an0 + c0, std::min(kChunk, B.n_an - c0), poff, 0});
        poff += A.ni;
      }
    A.nred_an = (int32_t)ch.size() - A.red0 - A.nred_in;
  }
  // per pipe column: first cell, class, hx * hy
  std::vector<double> hx, hy, ix, iy;
  grid_h(s.gx, hx, ix);
  grid_h(s.gy, hy, iy);
  std::vector<AdjCol> cols(ctx.pcol.size());
  for (size_t q = 0; q < ctx.pcol.size(); ++q) {
    const PipeCol& pc = ctx.pcol[q];
    cols[q].off = ((int64_t)pc.i * s.ny + pc.j) * s.nz;
    cols[q].hxy = hx[pc.i] * hy[pc.j];
    cols[q].cls = s.cls[(size_t)pc.j + (size_t)s.ny * pc.i];
    cols[q].pad = 0;
  }
  AdjPipe ap{};
  ap.bh = db.up(abh);
  ap.col = db.up(cols);
  ap.off = db.up(off);
  ap.ent = db.up(ent);
  ap.toff = db.up(toff);
  ap.tent = db.up(tent);
  ap.chunk = db.up(ch);
  ap.nchunk = (int32_t)ch.size();
  ap.partial = db.alloc<double>((size_t)poff);
  ap.gacc = db.up(std::vector<double>(ch.size(), 0.0));
  return ap;
}

// n fused steps from t on phi (as gw_steps, without graphs), adding every
// step's second-sub-step outlet to jacc; hist (may be NULL): every sub-step's
// level means and inlets, 2 (nmeans + nbh) doubles per step.
int run_fused(gw_ctx& ctx, const CTab* tab, void* phi, double t, double dt, int n, double* jacc, double* hist) {
  void* B = ctx.buf2;
  const size_t hs = 2 * ((size_t)tab->nmeans + (size_t)ctx.nbh_total);
  int rc = fused_zy(ctx, tab, phi, B, dt);
  if (rc) return rc;
  for (int k = 1; k < n; ++k) {
    const bool odd = (k & 1) != 0;
    if ((rc = fused_rest(ctx, tab, odd ? B : phi, odd ? phi : B, t + k * dt, dt, false, jacc,
                         hist ? hist + (size_t)(k - 1) * hs : nullptr)))
      return rc;
  }
  return fused_rest(ctx, tab, (n & 1) ? B : phi, phi, t + n * dt, dt, true, jacc,
                    hist ? hist + (size_t)(n - 1) * hs : nullptr);
}

// The n steps of a segment from u_0 = U[0]: the z-step + explicit y of the
// first step, then per step m the two advection sub-steps around the fused x
// kernel keeping e_m = (I - cA_x)^-1 b in E[m], and the implicit y keeping
// u_{m+1} in U[m+1], fused with the next step's z-step + explicit y except
// after the last step.  wa, wb: work fields (the fused y kernel writes out of
// place).  hist (may be NULL): the forward run's level means and inlets of
// the segment's sub-steps (run_fused): the advection then skips the
// reductions, which would recompute the same values.
int adj_recompute(gw_ctx& ctx, const CTab* tab, double* const* U, double* const* E, int n, double* wa, double* wb,
                  double t_first_end, double dt, const double* hist) {
  const double c = dt / 2;
  const DevView v = ctx.view(tab);
  const size_t hs = (size_t)tab->nmeans + (size_t)ctx.nbh_total;
  auto advect = [&](int m, int sub, double t_end) {
    return hist ? sub_step_replay(ctx, tab, wa, hist + (2 * (size_t)m + sub) * hs) : sub_step_advect(ctx, tab, wa, t_end);
  };
  int rc = hip_status(launch_line(v, GW_F64, LINE_ZY, 1, ctx.lw(LINE_ZY, 1), wa, U[0], c, dt, ctx.c3, ctx.stream, ctx.lsp()),
                      "z-step + explicit y");
  for (int m = 0; m < n && !rc; ++m) {
    const double t_end = t_first_end + (double)m * dt;
    rc = advect(m, 0, t_end);
    DevView vk = v;
    vk.aux = E[m];
    if (!rc) rc = hip_status(launch_line(vk, GW_F64, LINE_IMPLICIT_EXPLICIT_KEEP, 0, ctx.lw(LINE_IMPLICIT_EXPLICIT, 0),
                                         wa, wa, c, 0.0, 0.0, ctx.stream, ctx.lsp()), "implicit + explicit x");
    if (!rc) rc = advect(m, 1, t_end + c);
    if (rc) break;
    if (m + 1 < n) {
      DevView vy = v;
      vy.aux = U[m + 1];
      rc = hip_status(launch_line(vy, GW_F64, LINE_YZY_KEEP, 1, ctx.lw(LINE_YZY, 1), wb, wa, c, dt, ctx.c3,
                                  ctx.stream, ctx.lsp()), "implicit y + z-step + explicit y");
      std::swap(wa, wb);
    } else {
      rc = hip_status(launch_line(v, GW_F64, LINE_IMPLICIT, 1, ctx.lw(LINE_IMPLICIT, 1), U[m + 1], wa, c, 0.0, 0.0,
                                  ctx.stream, ctx.lsp()), "implicit y");
    }
  }
  return rc;
}

// The adjoint line modes' k-window widths: forced by the problem flags, else
// each mode timed once per context in both widths (out of place on scratch
// fields; the gradient buffers it touches are cleared by the caller).
int adj_tune(gw_ctx& ctx, const CTab* tab, const AdjGeo& ag, double* out, const double* in, const double* x) {
  if (ctx.adj_lw[0]) return GW_OK;
  if (ctx.force_lw) {
    ctx.adj_lw[0] = ctx.adj_lw[1] = ctx.force_lw & ~kLwFlags;
    return GW_OK;
  }
  const DevView v = ctx.view(tab);
  const double c = tab->c;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = hip_status(hipEventCreate(&e0), "event");
  if (!rc) rc = hip_status(hipEventCreate(&e1), "event");
  for (int mode = 0; mode < 2 && !rc; ++mode) {
    float best = 0.f;
    int best_lw = 64;
    for (int lw : {64, 32}) {
      auto run = [&]() {
        return hip_status(lw == 32 ? line_adj_w32(v, ag, mode, out, in, x, nullptr, c, 0.0, 0.0, 0.0, ctx.stream)
                                   : line_adj_w64(v, ag, mode, out, in, x, nullptr, c, 0.0, 0.0, 0.0, ctx.stream),
                          "adjoint line tune");
      };
      if ((rc = run())) break;
      if ((rc = hip_status(hipEventRecord(e0, ctx.stream), "event"))) break;
      for (int r = 0; r < 5 && !rc; ++r) rc = run();
      if (rc || (rc = hip_status(hipEventRecord(e1, ctx.stream), "event"))) break;
      if ((rc = hip_status(hipEventSynchronize(e1), "event sync"))) break;
      float ms = 0.f;
      if ((rc = hip_status(hipEventElapsedTime(&ms, e0, e1), "event time"))) break;
      if (best == 0.f || ms < best) { best = ms; best_lw = lw; }
    }
    ctx.adj_lw[mode] = best_lw;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  return rc;
}

}  // namespace

extern "C" {

int gw_adjoint_op(gw_ctx* ctx, int32_t op, void* out, const void* in, const void* x, double dt) {
  if (!ctx || !out || !in || op < 0 || op > 6) return fail(GW_EINVAL, "bad argument");
  if (!(dt > 0)) return fail(GW_EINVAL, "dt must be positive");
  if (ctx->sc.size() != 1 || ctx->prec != GW_F64 || !ctx->fused_ok)
    return fail(GW_EINVAL, "the adjoint operators take one FP64 scenario");
  if ((op == 1 || op == 2 || op == 4 || op == 6) && !x)
    return fail(GW_EINVAL, "this operator needs the forward field x");
  if ((op == 4 || op == 6) && (out == in || out == x)) return fail(GW_EINVAL, "out must not alias in or x");
  gw_ctx& C = *ctx;
  try {
    CTab* tab = nullptr;
    int rc = prepare_steps(C, in, dt, tab);
    if (rc) return rc;
    const double c = dt / 2;
    DevBufs db;
    const AdjGeo ag = make_adj_geo(C, db);
    const DevView v = C.view(tab);
    double* o = static_cast<double*>(out);
    const double* i = static_cast<const double*>(in);
    const double* xf = static_cast<const double*>(x);
    const size_t bytes = (size_t)C.cells_total * sizeof(double);
    switch (op) {
      case 0: {
        if (!tab->adv_err.empty()) return fail(GW_EINVAL, tab->adv_err);
        const AdjPipe ap = build_adj_pipe(C, *tab, db);
        rc = hip_status(hipMemcpyAsync(o, i, bytes, hipMemcpyDeviceToDevice, C.stream), "copy");
        if (!rc) rc = hip_status(launch_adj_pipe(v, ag, ap, o, nullptr, C.stream), "adjoint advection");
        break;
      }
      case 1:
      case 2: {
        const int mode = op == 1 ? ADJ_Y : ADJ_X;
        const int lw = C.adj_lw[mode] ? C.adj_lw[mode]
                                       : (mode == ADJ_Y ? C.lw(LINE_IMPLICIT, 1) : C.lw(LINE_IMPLICIT_EXPLICIT, 0)) &
                                             ~kLwFlags;
        rc = hip_status(lw == 32 ? line_adj_w32(v, ag, mode, o, i, xf, nullptr, c, 0.0, 0.0, 0.0, C.stream)
                                 : line_adj_w64(v, ag, mode, o, i, xf, nullptr, c, 0.0, 0.0, 0.0, C.stream),
                        "adjoint line");
        break;
      }
      case 3:
        rc = hip_status(launch_line(v, GW_F64, LINE_EXPLICIT, 1, C.lw(LINE_EXPLICIT, 1), o, i, c, 0.0, 0.0,
                                    C.stream, C.lsp()), "explicit y");
        break;
      case 4:
      case 6:
        rc = hip_status(launch_adj_yz(v, ag, op == 6, o, i, xf, c, dt, C.c3, C.stream), "adjoint z");
        break;
      default:
        rc = hip_status(launch_mul_w(v, ag, o, i, C.stream), "W mu");
        break;
    }
    if (!rc) rc = hip_status(hipStreamSynchronize(C.stream), "sync");
    return rc;
  } catch (const std::bad_alloc&) {
    return fail(GW_ENOMEM, "cannot allocate the adjoint tables");
  } catch (const HipFail& e) {
    return fail(GW_EHIP, std::string("adjoint set-up: ") + hipGetErrorString(e.e));
  }
}

int gw_adjoint(gw_ctx* ctx, const void* phi0, double t0, double dt, int32_t nsteps, int32_t seg,
               const double* w_bh, double* J, double* dJ_dk, int32_t n_dk, double* dJ_dTin, void* dJ_du0) {
  if (!ctx || !phi0 || !J || !dJ_dk || !dJ_dTin) return fail(GW_EINVAL, "null argument");
  if (!(dt > 0) || nsteps < 1 || seg < 1) return fail(GW_EINVAL, "dt must be positive, nsteps and seg >= 1");
  if (ctx->sc.size() != 1) return fail(GW_EINVAL, "the adjoint takes one scenario");
  if (ctx->prec != GW_F64) return fail(GW_EINVAL, "the adjoint runs in FP64");
  if (!ctx->fused_ok) return fail(GW_ESTATE, "too many column classes for the adjoint kernels");
  const Scen& s = ctx->sc[0];
  if (s.inlet.kind != GW_INLET_CONSTANT && s.inlet.kind != GW_INLET_PER_WELL)
    return fail(GW_EINVAL, "the adjoint supports constant and per-well inlets");
  const int nlay = s.mat.depth.empty() ? 1 : (int)s.mat.depth.size();
  if (n_dk != nlay) return fail(GW_EINVAL, "dJ_dk must have one entry per rock layer (1 when homogeneous)");
  gw_ctx& C = *ctx;
  try {
    CTab* tab = nullptr;
    int rc = prepare_steps(C, phi0, dt, tab);
    if (rc) return rc;
    const double c = dt / 2;
    const size_t cells = (size_t)C.cells_total, bytes = cells * sizeof(double);
    const int nbh = C.nbh_total;
    if (!tab->adv_err.empty()) return fail(GW_EINVAL, tab->adv_err);
    DevBufs db;
    const AdjGeo ag = make_adj_geo(C, db);
    const AdjPipe ap = build_adj_pipe(C, *tab, db);
    // storage: checkpoints, one segment of u / e, two adjoint fields, two work fields
    const int nseg = (nsteps + seg - 1) / seg;
    const int sl = std::min(seg, nsteps);
    // all of them in one allocation (a few hundred separate hipMalloc/hipFree
    // pairs of field-sized buffers cost ~2% of a 30-day gradient)
    std::vector<double*> ck(nseg), U(sl + 1), E(sl);
    const size_t fstride = (cells + 31) / 32 * 32;  // 256-byte aligned fields
    double* arena = db.alloc<double>(fstride * (size_t)(nseg + (sl + 1) + sl + 4));
    size_t nf = 0;
    for (auto& p : ck) p = arena + fstride * nf++;
    for (auto& p : U) p = arena + fstride * nf++;
    for (auto& p : E) p = arena + fstride * nf++;
    double* MU0 = arena + fstride * nf++;
    double* MU1 = arena + fstride * nf++;
    double* WB = arena + fstride * nf++;
    double* WB2 = arena + fstride * nf++;
    double* jacc = db.up(std::vector<double>(nbh, 0.0));
    // every sub-step's level means and inlets of the forward run, for the
    // recompute (up to 1 GiB; beyond that the recompute reduces again)
    const size_t hstep = 2 * ((size_t)tab->nmeans + (size_t)nbh);
    double* hist = hstep * (size_t)nsteps * sizeof(double) <= ((size_t)1 << 30)
                       ? db.alloc<double>(hstep * (size_t)nsteps) : nullptr;
    // forward with checkpoints
    if ((rc = hip_status(hipMemcpyAsync(U[0], phi0, bytes, hipMemcpyDeviceToDevice, C.stream), "copy"))) return rc;
    for (int sg = 0; sg < nseg; ++sg) {
      const int n = std::min(seg, nsteps - sg * seg);
      if ((rc = hip_status(hipMemcpyAsync(ck[sg], U[0], bytes, hipMemcpyDeviceToDevice, C.stream), "copy")))
        return rc;
      if ((rc = run_fused(C, tab, U[0], t0 + (double)sg * seg * dt, dt, n, jacc,
                          hist ? hist + (size_t)sg * seg * hstep : nullptr)))
        return rc;
    }
    std::vector<double> w(nbh), hj(nbh);
    for (int b = 0; b < nbh; ++b) w[b] = w_bh ? w_bh[b] : 1.0 / nbh;
    if ((rc = hip_status(hipMemcpyAsync(hj.data(), jacc, sizeof(double) * nbh, hipMemcpyDeviceToHost, C.stream),
                         "copy")) ||
        (rc = hip_status(hipStreamSynchronize(C.stream), "sync")))
      return rc;
    double jv = 0.0;
    for (int b = 0; b < nbh; ++b) jv += w[b] * hj[b];
    *J = jv / nsteps;
    // objective seed per borehole: w_b / (nsteps n_in n_out) on the inner outlet levels
    std::vector<double> seedh(nbh);
    for (int b = 0; b < nbh; ++b)
      seedh[b] = w[b] / ((double)nsteps * (double)tab->h_bh[b].n_in * (double)tab->h_bh[b].n_outlev);
    const double* seed = db.up(seedh);
    // backward
    const DevView v = C.view(tab);
    if ((rc = hip_status(hipMemsetAsync(MU0, 0, bytes, C.stream), "memset"))) return rc;
    if ((rc = adj_tune(C, tab, ag, MU1, MU0, static_cast<const double*>(phi0)))) return rc;
    if ((rc = hip_status(hipMemsetAsync(ag.gx_part, 0, sizeof(double) * (size_t)s.ny * s.nz, C.stream), "memset")) ||
        (rc = hip_status(hipMemsetAsync(ag.gy_part, 0, sizeof(double) * (size_t)ag.yparts * s.nx * s.nz, C.stream),
                         "memset")))
      return rc;
    auto adj_line = [&](int mode, double* out, const double* in, const double* x1) {
      const int lw = C.adj_lw[mode];
      return hip_status(lw == 32 ? line_adj_w32(v, ag, mode, out, in, x1, nullptr, c, 0.0, 0.0, 0.0, C.stream)
                                 : line_adj_w64(v, ag, mode, out, in, x1, nullptr, c, 0.0, 0.0, 0.0, C.stream),
                        "adjoint line");
    };
    double* A = MU0;  // mu_{m+1}, then mu_m
    double* Bf = MU1;
    auto reverse = [&](int n) -> int {
      for (int m = n - 1; m >= 0; --m) {
        int r;
        if ((r = adj_line(ADJ_Y, Bf, A, U[m + 1]))) return r;
        if ((r = hip_status(launch_adj_pipe(v, ag, ap, Bf, seed, C.stream), "adjoint advection"))) return r;
        if ((r = adj_line(ADJ_X, A, Bf, E[m]))) return r;
        if ((r = hip_status(launch_adj_pipe(v, ag, ap, A, nullptr, C.stream), "adjoint advection"))) return r;
        if ((r = hip_status(launch_adj_yz(v, ag, true, Bf, A, U[m], c, dt, C.c3, C.stream), "adjoint y + z")))
          return r;
        std::swap(A, Bf);
      }
      return GW_OK;
    };
    // The reverse steps of a full segment launch the same kernels on the same
    // buffers in every segment (U, E, the two adjoint fields, the seed, the
    // gradient accumulators), so with four or more segments they run as one
    // graph replay per segment: one graph per starting adjoint field (two when
    // seg is odd).  Graph or direct follows the context's gw_steps choice; when
    // that is undecided the first two full segments are timed, one each way.
    struct RevGraphs {
      hipStream_t st;
      hipGraphExec_t g[2] = {nullptr, nullptr};
      hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
      ~RevGraphs() {
        if (g[0] || g[1]) (void)hipStreamSynchronize(st);
        for (auto x : g)
          if (x) (void)hipGraphExecDestroy(x);
        for (auto e : ev)
          if (e) (void)hipEventDestroy(e);
      }
    } rg{C.stream};
    int rev_mode = nseg >= 4 ? C.steps_mode : kStepsEager;
    int timed = 0;  // full segments timed so far: the first direct, the second a graph replay
    if (rev_mode == kStepsUndecided)
      for (auto& e : rg.ev)
        if ((rc = hip_status(hipEventCreate(&e), "event"))) return rc;
    for (int sg = nseg - 1; sg >= 0; --sg) {
      const int n = std::min(seg, nsteps - sg * seg);
      const int base = sg * seg;
      if ((rc = hip_status(hipMemcpyAsync(U[0], ck[sg], bytes, hipMemcpyDeviceToDevice, C.stream), "copy")))
        return rc;
      if ((rc = adj_recompute(C, tab, U.data(), E.data(), n, WB, WB2, t0 + (double)(base + 1) * dt, dt,
                              hist ? hist + (size_t)base * hstep : nullptr)))
        return rc;
      const bool full = n == seg && rev_mode != kStepsEager;
      const bool timing = full && rev_mode == kStepsUndecided;
      const bool graph = full && (rev_mode == kStepsGraph || timed == 1);
      hipGraphExec_t& gx = rg.g[A == MU0 ? 0 : 1];
      if (graph && !gx) {
        double* const A0 = A;
        double* const B0 = Bf;
        hipGraph_t g = nullptr;
        if ((rc = hip_status(hipStreamBeginCapture(C.stream, hipStreamCaptureModeThreadLocal), "begin capture")))
          return rc;
        const int rc1 = reverse(n);
        const hipError_t e2 = hipStreamEndCapture(C.stream, &g);
        A = A0;
        Bf = B0;
        if (rc1 || e2 != hipSuccess) {
          if (g) (void)hipGraphDestroy(g);
          return rc1 ? rc1 : hip_status(e2, "end capture");
        }
        rc = hip_status(hipGraphInstantiate(&gx, g, nullptr, nullptr, 0), "graph instantiate");
        (void)hipGraphDestroy(g);
        if (rc) return rc;
      }
      if (timing && (rc = hip_status(hipEventRecord(rg.ev[2 * timed], C.stream), "event"))) return rc;
      if (graph) {
        if ((rc = hip_status(hipGraphLaunch(gx, C.stream), "graph launch"))) return rc;
        if (n & 1) std::swap(A, Bf);
      } else if ((rc = reverse(n))) {
        return rc;
      }
      if (timing) {
        if ((rc = hip_status(hipEventRecord(rg.ev[2 * timed + 1], C.stream), "event"))) return rc;
        if (++timed == 2) {
          float direct = 0.f, replay = 0.f;
          if ((rc = hip_status(hipEventSynchronize(rg.ev[3]), "event sync")) ||
              (rc = hip_status(hipEventElapsedTime(&direct, rg.ev[0], rg.ev[1]), "event time")) ||
              (rc = hip_status(hipEventElapsedTime(&replay, rg.ev[2], rg.ev[3]), "event time")))
            return rc;
          rev_mode = replay < direct ? kStepsGraph : kStepsEager;
        }
      }
    }
    if (dJ_du0 && (rc = hip_status(launch_mul_w(v, ag, static_cast<double*>(dJ_du0), A, C.stream), "W mu")))
      return rc;
    // gradients to the host: per level over the lines in order, then per layer
    std::vector<double> gxp((size_t)s.ny * s.nz), gyp((size_t)ag.yparts * s.nx * s.nz), ga((size_t)ap.nchunk);
    std::vector<RedChunk> chh((size_t)ap.nchunk);
    if ((rc = hip_status(hipMemcpyAsync(gxp.data(), ag.gx_part, sizeof(double) * gxp.size(), hipMemcpyDeviceToHost,
                                        C.stream), "copy")) ||
        (rc = hip_status(hipMemcpyAsync(gyp.data(), ag.gy_part, sizeof(double) * gyp.size(), hipMemcpyDeviceToHost,
                                        C.stream), "copy")) ||
        (rc = hip_status(hipMemcpyAsync(ga.data(), ap.gacc, sizeof(double) * ga.size(), hipMemcpyDeviceToHost,
                                        C.stream), "copy")) ||
        (rc = hip_status(hipMemcpyAsync(chh.data(), ap.chunk, sizeof(RedChunk) * chh.size(), hipMemcpyDeviceToHost,
                                        C.stream), "copy")) ||
        (rc = hip_status(hipStreamSynchronize(C.stream), "sync")))
      return rc;
    for (int q = 0; q < nlay; ++q) dJ_dk[q] = 0.0;
    for (int k = 0; k < s.nz; ++k) {
      double tot = 0.0;
      for (int j = 0; j < s.ny; ++j) tot += gxp[(size_t)j * s.nz + k];
      for (int q = 0; q < ag.yparts * s.nx; ++q) tot += gyp[(size_t)q * s.nz + k];
      int lay = 0;
      if (!s.mat.depth.empty()) {
        lay = nlay - 1;
        for (int q = 0; q < nlay; ++q)
          if (s.gz[k] <= s.mat.depth[q]) { lay = q; break; }
      }
      dJ_dk[lay] += tot;
    }
    // inlet gradient: the annulus chunks' accumulators of each borehole in chunk order
    for (int b = 0; b < nbh; ++b) dJ_dTin[b] = 0.0;
    for (size_t q = 0; q < chh.size(); ++q)
      if (chh[q].part == 1) dJ_dTin[chh[q].bh] += ga[q];
    return GW_OK;
  } catch (const std::bad_alloc&) {
    return fail(GW_ENOMEM, "cannot allocate the adjoint's checkpoints and work fields");
  } catch (const HipFail& e) {
    return fail(GW_EHIP, std::string("adjoint set-up: ") + hipGetErrorString(e.e));
  }
}

}  // extern "C"
