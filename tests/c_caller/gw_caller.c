/* A C caller of libgeowells_hip: no Python, no torch.  It builds the
 * single-borehole problem of the smoke test through the C ABI (grid
 * generators, gw_create), puts the initial field T0 = 2.29 + 0.035 z
 * (src/physics.jl:333-348) in device memory it allocates itself, advances it
 * nsteps steps of 60 s with gw_steps (a heat-exchanger inlet, dT = 5 K, so
 * the outlet feeds back into the inlet on the device) and writes the outlets
 * and the final field:
 *
 *   gw_caller NSTEPS FIELD_OUT   ->  stdout: "grid nx ny nz" and one
 *                                    "outlet <value>" line per borehole;
 *                                    FIELD_OUT: the field, nx*ny*nz doubles
 *
 * This is the sequence a Julia ccall binding runs (create_cache, then the
 * solve loop's steps, then cache.T_outlet).  Exit status 0 on success, 1 with
 * the library's message on stderr otherwise. */
#include <stdio.h>
#include <stdlib.h>

#include <hip/hip_runtime_api.h>

#include "geowells.h"

static int check(int rc, const char *what) {
  if (rc != GW_OK) {
    fprintf(stderr, "%s: %d %s\n", what, rc, gw_last_error());
    exit(1);
  }
  return rc;
}

static double *grid_1d(const gw_borehole *bh, int dir, int *n) {
  const int64_t len = gw_adaptive_grid_1d(-5.0, 5.0, 0.01, 2.0, 2.0, bh, 1, dir, NULL, 0);
  if (len < 3) check((int)len, "adaptive grid");
  double *g = (double *)malloc(sizeof(double) * (size_t)len);
  gw_adaptive_grid_1d(-5.0, 5.0, 0.01, 2.0, 2.0, bh, 1, dir, g, len);
  *n = (int)len;
  return g;
}

int main(int argc, char **argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: %s NSTEPS FIELD_OUT\n", argv[0]);
    return 1;
  }
  const int nsteps = atoi(argv[1]);
  const gw_borehole bh = {0.0, 0.0, 60.0, 0.038, 0.0065, 0.0889, 0.01, 0.108, 0.01 * 998.2, 0.0};
  int nx, ny;
  double *gx = grid_1d(&bh, 0, &nx), *gy = grid_1d(&bh, 1, &ny);
  const int64_t nz = gw_uniform_gridz(0.0, 70.0, 2.0, &bh, 1, NULL, 0);
  if (nz < 3) check((int)nz, "uniform gridz");
  double *gz = (double *)malloc(sizeof(double) * (size_t)nz);
  gw_uniform_gridz(0.0, 70.0, 2.0, &bh, 1, gz, nz);

  gw_scenario sc = {0};
  sc.nx = nx;
  sc.ny = ny;
  sc.nz = (int32_t)nz;
  sc.gridx = gx;
  sc.gridy = gy;
  sc.gridz = gz;
  sc.n_bh = 1;
  sc.boreholes = &bh;
  sc.materials.n_layers = 0;
  sc.materials.k_rock = 2.5;
  sc.materials.rho_c_rock = 2.3e6;
  sc.materials.k_water = 0.6;
  sc.materials.rho_c_water = 4.18e6;
  sc.materials.k_steel = 44.5;
  sc.materials.rho_c_steel = 3.73e6;
  sc.materials.k_insulating = 0.4;
  sc.materials.rho_c_insulating = 1.96e6;
  sc.materials.k_backfill = 1.5;
  sc.materials.rho_c_backfill = 1.76e6;
  sc.inlet.kind = GW_INLET_HEAT_EXCHANGER;
  sc.inlet.value = 5.0;
  gw_problem p = {0};
  p.n_scen = 1;
  p.scen = &sc;
  p.precision = GW_F64;
  gw_ctx *ctx = NULL;
  check(gw_create(&p, &ctx), "gw_create");

  const size_t cells = (size_t)nx * (size_t)ny * (size_t)nz;
  double *h = (double *)malloc(sizeof(double) * cells);
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int64_t k = 0; k < nz; ++k) h[(size_t)k + (size_t)nz * ((size_t)j + (size_t)ny * (size_t)i)] = 2.29 + 0.035 * gz[k];
  void *d = NULL;
  if (hipMalloc(&d, sizeof(double) * cells) != hipSuccess ||
      hipMemcpy(d, h, sizeof(double) * cells, hipMemcpyHostToDevice) != hipSuccess) {
    fprintf(stderr, "device field allocation or copy failed\n");
    return 1;
  }
  check(gw_steps(ctx, d, 0.0, 60.0, nsteps), "gw_steps");
  double T_out = 0.0;
  check(gw_outlet(ctx, &T_out), "gw_outlet");
  if (hipMemcpy(h, d, sizeof(double) * cells, hipMemcpyDeviceToHost) != hipSuccess) {
    fprintf(stderr, "device field copy failed\n");
    return 1;
  }
  FILE *f = fopen(argv[2], "wb");
  if (!f || fwrite(h, sizeof(double), cells, f) != cells) {
    fprintf(stderr, "cannot write %s\n", argv[2]);
    return 1;
  }
  fclose(f);
  printf("grid %d %d %lld\n", nx, ny, (long long)nz);
  printf("outlet %.17g\n", T_out);
  check(gw_destroy(ctx), "gw_destroy");
  (void)hipFree(d);
  free(h);
  free(gx);
  free(gy);
  free(gz);
  return 0;
}
