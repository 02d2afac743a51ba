/*
 * driver.c -- CPU exerciser of the host library (libse2w_host.so's sources)
 * and of the FP64 oracle, built with AddressSanitizer and
 * UndefinedBehaviorSanitizer by tests/test_sanitizers.py.  It walks every
 * host entry point that needs no GPU (basis, mesh, materials, mass, time
 * step, STFs, sources, receivers and their files, options, the -verify_stf
 * program path, error paths) and a short oracle run with every optional
 * feature, so that an out-of-bounds access, use after free, leak or
 * undefined operation in them fails the test.  Returns 0 when every call
 * behaved as expected.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "se2wave.h"

/* The oracle is test infrastructure; it is compiled into this test program. */
#include "../../oracle/ref_cpu.c"

#define EXPECT(c)                                                       \
  do {                                                                  \
    if (!(c)) {                                                         \
      fprintf(stderr, "driver.c:%d: expectation failed: %s\n", __LINE__, #c); \
      return 1;                                                         \
    }                                                                   \
  } while (0)

static int host_layer(const char *dir)
{
  SE2WSpec spec = NULL;
  se2w_spec_info info;
  SE2WSTF stf[3] = {NULL, NULL, NULL};
  SE2WSource src[3] = {NULL, NULL, NULL};
  SE2WReceivers rec = NULL;
  SE2WReceiverUV uv[2] = {NULL, NULL};
  const int mx = 7, my = 5, p = 5, nb = p + 1, npe = nb * nb;
  const double scale[2] = {4000.0, 4000.0}, shift[2] = {0.0, 0.0};
  double vp[35], vs[35], rho[35], D[36], w1d[6], edx[7], edy[5], dt, psi, xi[6], w[6], N[6], dN[6];
  double *coords, *mass, *g, lam[35], mu[35], rr[35], times[3] = {1e-3, 2e-3, 3e-3}, rows[3 * 2 * 7], uvrows[3 * 4];
  double mom[4] = {1.0e9, 2.0e8, -3.0e8, 7.0e8}, frc[4] = {2.0e9, 1.0e10, 0.0, 0.0}, xr[4] = {1234.5, 2987.6, 3999.0, 17.0};
  int32_t *element, nid[2], el[2], gll[2], e, cg;
  double vals[2 * 36], weights[36], mapped[2];
  uint64_t st = 7;
  int k;

  EXPECT(se2w_gll(p, xi, w) == 0 && se2w_deriv_matrix(p, D) == 0);
  EXPECT(se2w_basis1d(p, 1, &xi[2], N) == 0 && se2w_basis1d_deriv(p, 1, &xi[2], dN) == 0);
  EXPECT(se2w_spec_create(mx, my, p, &spec) == 0);
  EXPECT(se2w_spec_scale_coords(spec, scale, shift) == 0);
  EXPECT(se2w_material_random_cells(mx, my, 3, 4000.0, 6000.0, 2500.0, 3000.0, vp, vs, rho) == 0);
  EXPECT(se2w_spec_set_cell_material_velocity(spec, vp, vs, rho) == 0);
  EXPECT(se2w_material_layered(mx, my, 3, 5, 3000.0, 7000.0, 2200.0, 3200.0, vp, vs, rho) == 0);
  EXPECT(se2w_spec_set_perturbed_material_velocity(spec, 5000.0, 250.0, 2900.0, 100.0, 2600.0, 50.0) == 0);
  EXPECT(se2w_spec_set_constant_material_velocity(spec, 6000.0, 3464.0, 2670.0) == 0);
  EXPECT(se2w_spec_get_info(spec, &info) == 0 && info.nnodes == (int64_t)(p * mx + 1) * (p * my + 1));
  coords = (double *)malloc(sizeof(double) * 2 * info.nnodes);
  mass = (double *)malloc(sizeof(double) * info.nnodes);
  g = (double *)malloc(sizeof(double) * 2 * info.nnodes);
  element = (int32_t *)malloc(sizeof(int32_t) * mx * my * npe);
  EXPECT(coords && mass && g && element);
  EXPECT(se2w_spec_get_coords(spec, coords) == 0 && se2w_spec_get_element(spec, element) == 0);
  EXPECT(se2w_spec_get_material(spec, lam, mu, rr) == 0 && se2w_spec_get_tables(spec, D, w1d, edx, edy) == 0);
  EXPECT(se2w_spec_mass(spec, mass) == 0 && se2w_spec_timestep(spec, &dt) == 0 && dt > 0.0);
  EXPECT(se2w_stf_create_ricker(0.02, 25.0, 1.0, &stf[0]) == 0);
  EXPECT(se2w_stf_create_gaussian(0.03, 40.0, 2.0, &stf[1]) == 0);
  EXPECT(se2w_stf_create_yoffe(0.2, 0.3, 1.0, &stf[2]) == 0);
  for (k = 0; k < 3; k++) EXPECT(se2w_stf_evaluate(stf[k], 0.025, &psi) == 0);
  EXPECT(se2w_source_create(spec, SE2W_SOURCE_MOMENT, SE2W_SOURCE_IMPL_NEAREST_QPOINT, &xr[0], mom, 1, &src[0]) == 0);
  EXPECT(se2w_source_create(spec, SE2W_SOURCE_FORCE, SE2W_SOURCE_IMPL_POINTWISE, &xr[2], frc, 1, &src[1]) == 0);
  EXPECT(se2w_source_create(spec, SE2W_SOURCE_MOMENT, SE2W_SOURCE_IMPL_POINTWISE, &xr[0], mom, 0, &src[2]) == 0);
  for (k = 0; k < 3; k++) EXPECT(se2w_source_setup(src[k]) == 0 && se2w_source_get(src[k], &e, &cg, vals) == 0);
  EXPECT(se2w_sources_evaluate(0.025, 3, src, stf, g) == 0);
  /* the legacy cubic-spline (default and set width) and P0 forms */
  {
    SE2WSource leg[3] = {NULL, NULL, NULL};
    int32_t ne = 0, *lel;
    double h, int_w, *lv;
    EXPECT(se2w_source_create(spec, SE2W_SOURCE_MOMENT, SE2W_SOURCE_IMPL_SPLINE, &xr[0], mom, 1, &leg[0]) == 0);
    EXPECT(se2w_source_create(spec, SE2W_SOURCE_MOMENT, SE2W_SOURCE_IMPL_SPLINE, &xr[2], mom, 1, &leg[1]) == 0);
    EXPECT(se2w_source_set_kernel_width(leg[1], 900.0) == 0);
    EXPECT(se2w_source_create(spec, SE2W_SOURCE_MOMENT, SE2W_SOURCE_IMPL_P0, &xr[2], mom, 1, &leg[2]) == 0);
    for (k = 0; k < 3; k++) {
      EXPECT(se2w_source_setup(leg[k]) == 0 && se2w_source_get_elements(leg[k], &ne, NULL, NULL) == 0 && ne >= 1);
      lel = (int32_t *)malloc(sizeof(int32_t) * ne);
      lv = (double *)malloc(sizeof(double) * ne * 2 * npe);
      EXPECT(lel && lv && se2w_source_get_elements(leg[k], &ne, lel, lv) == 0);
      free(lel);
      free(lv);
    }
    EXPECT(se2w_source_get_kernel(leg[1], &h, &int_w) == 0 && h == 900.0);
    EXPECT(se2w_source_get(leg[0], &e, &cg, vals) != 0 && se2w_source_get(leg[2], &e, &cg, vals) == 0);
    EXPECT(se2w_sources_evaluate(0.025, 3, leg, stf, g) == 0);
    EXPECT(se2w_source_set_kernel_width(leg[1], 100.0) != 0); /* after setup */
    for (k = 0; k < 3; k++) se2w_source_destroy(&leg[k]);
  }
  EXPECT(se2w_receivers_create(spec, 2, xr, &rec) == 0 && se2w_receivers_get(rec, nid, el, gll) == 0);
  EXPECT(se2w_receivers_write_header(rec, dir) == 0);
  for (k = 0; k < 3 * 2 * 7; k++) rows[k] = 1.0e-3 * k;
  EXPECT(se2w_receivers_append_rows(rec, 3, times, rows) == 0);
  for (k = 0; k < 2; k++) {
    EXPECT(se2w_receiver_uv_create(spec, &xr[2 * k], k, 1, &uv[k]) == 0);
    EXPECT(se2w_receiver_uv_get(uv[k], &e, weights, mapped) == 0);
    EXPECT(se2w_receiver_uv_write_header(uv[k], dir) == 0);
  }
  for (k = 0; k < 12; k++) uvrows[k] = -2.0e-4 * k;
  EXPECT(se2w_receiver_uv_append_rows(uv[0], 3, times, uvrows) == 0);
  for (k = 0; k < 4; k++) (void)se2w_rng_uniform(&st);
  se2w_rand48_seed(&st, 2);
  (void)se2w_rand48_next(&st);
  /* error paths: each fails with a message and leaks nothing */
  {
    SE2WSource bad = NULL;
    SE2WReceiverUV badr = NULL;
    SE2WSTF bads = NULL;
    const double outside[2] = {-5.0, 100.0}, edge[2] = {4000.0, 100.0};
    EXPECT(se2w_source_create(spec, SE2W_SOURCE_MOMENT, SE2W_SOURCE_IMPL_NEAREST_QPOINT, outside, mom, 1, &bad) != 0);
    EXPECT(se2w_receiver_uv_create(spec, edge, 0, 1, &badr) != 0);
    EXPECT(se2w_stf_create_yoffe(0.3, 0.1, 1.0, &bads) != 0);
    EXPECT(strlen(se2w_host_last_error()) > 0);
  }
  for (k = 0; k < 2; k++) se2w_receiver_uv_destroy(&uv[k]);
  se2w_receivers_destroy(&rec);
  for (k = 0; k < 3; k++) {
    se2w_source_destroy(&src[k]);
    se2w_stf_destroy(&stf[k]);
  }
  se2w_spec_destroy(&spec);
  free(coords); free(mass); free(g); free(element);
  return 0;
}

static int options_and_program(const char *dir)
{
  char out_dir[600];
  char *argv[] = {"se2wave", "-mx", "12", "-bdegree", "3", "-tmax", "0.25", "-verify_stf", "-stf_ricker",
                  "-output_dir", out_dir, NULL};
  char *argv2[] = {"se2wave", "-mx", "9", "-my", "4", "-rec_xy", "100,200,150,-20", "-precision", "fp32k", NULL};
  SE2WOptions opts = NULL;
  double arr[8];
  int n = 8, found = 0, v = 0;
  char s[64];
  snprintf(out_dir, sizeof(out_dir), "%s", dir);
  EXPECT(se2w_options_create(9, argv2, &opts) == 0);
  EXPECT(se2w_options_get_int(opts, "-my", &v, &found) == 0 && found && v == 4);
  EXPECT(se2w_options_get_real_array(opts, "-rec_xy", arr, &n, &found) == 0 && found && n == 4);
  EXPECT(se2w_options_get_string(opts, "-precision", s, sizeof(s), &found) == 0 && found && !strcmp(s, "fp32k"));
  se2w_options_destroy(&opts);
  EXPECT(se2w_main(11, argv) == 0);
  return 0;
}

static int oracle_run(const char *dir)
{
  /* 5 x 4 elements, p = 4, two pointwise / nearest sources, receivers with
   * per-step file output, RecordUV and RecordUV_interp stations */
  const int mx = 5, my = 4, ne = 20;
  double lam[20], mu[20], rho[20], src_xy[4] = {1230.0, 2111.0, 3001.0, 901.0};
  int src_type[4] = {1, 0, 1, 1}, stf_type[4] = {0, 1, 0, 1}, src_impl[4] = {0, 1, 2, 4}, uv_interp[3] = {0, 1, 1};
  double src_values[16] = {1e9, 2e8, -3e8, 7e8, 2e9, 1e10, 0.0, 0.0, 1e9, 0.0, 0.0, 1e9, 0.0, 5e8, 5e8, 0.0};
  double stf_params[16] = {0.02, 25.0, 1.0, 0.0, 0.03, 40.0, 2.0, 0.0, 0.02, 25.0, 1.0, 0.0, 0.03, 40.0, 2.0, 0.0};
  double src_h[4] = {0.0, 0.0, 700.0, 0.0}, Fa[2 * 21 * 17], kh, iw;
  double rec_xy[4] = {1500.3, 2200.7, 3999.0, 3999.0}, uv_xy[6] = {1500.3, 2200.7, 1500.3, 2200.7, 10.0, 3990.0};
  double src_xy4[8] = {1230.0, 2111.0, 3001.0, 901.0, 2000.0, 2000.0, 3700.0, 3600.0};
  double trace[30 * 2 * 7], times[30], uvt[30 * 3 * 4], *u, *F;
  orc_scenario sc;
  int k, done = 0;
  long n2 = 2L * (4 * mx + 1) * (4 * my + 1);
  for (k = 0; k < ne; k++) {
    mu[k] = 3000.0 * 3000.0 * 2600.0;
    lam[k] = 5200.0 * 5200.0 * 2600.0 - 2.0 * mu[k];
    rho[k] = 2600.0 + k;
  }
  memset(&sc, 0, sizeof(sc));
  sc.mx = mx; sc.my = my; sc.p = 4;
  sc.scale[0] = sc.scale[1] = 4000.0;
  sc.lambda = lam; sc.mu = mu; sc.rho = rho;
  sc.nsrc = 4; sc.src_xy = src_xy4; sc.src_type = src_type; sc.src_values = src_values; sc.src_h = src_h;
  sc.stf_type = stf_type; sc.stf_params = stf_params; sc.use_nearest_internal_gll = 1;
  sc.nrec = 2; sc.rec_xy = rec_xy;
  sc.dt = 1.0e-3; sc.nt = 30; sc.tmax = 1.0e30; sc.nthreads = 2;
  sc.src_impl = src_impl; sc.nuv = 3; sc.uv_xy = uv_xy; sc.uv_interp = uv_interp; sc.rec_dir = dir;
  EXPECT(orc_run_ex(&sc, 0, NULL, NULL, NULL, trace, times, &done, NULL, NULL, NULL, uvt) == 0 && done == 30);
  EXPECT(orc_write_receiver_files(&sc, dir, done, times, trace) == 0);
  for (k = 0; k < 4; k++) EXPECT(orc_source_assembled(&sc, k, Fa, &kh, &iw) == 0);
  (void)src_xy;
  u = (double *)calloc(n2, sizeof(double));
  F = (double *)calloc(n2, sizeof(double));
  EXPECT(u && F);
  for (k = 0; k < n2; k++) u[k] = 1e-3 * ((k * 37) % 101);
  EXPECT(orc_assemble_forces(&sc, u, F) == 0);
  EXPECT(orc_receiver_rows(&sc, u, u, u, trace) == 0);
  EXPECT(orc_uv_rows(&sc, u, u, uvt, NULL, NULL) == 0);
  free(u);
  free(F);
  return 0;
}

int main(int argc, char **argv)
{
  const char *dir = argc > 1 ? argv[1] : ".";
  if (host_layer(dir)) return 1;
  if (options_and_program(dir)) return 1;
  if (oracle_run(dir)) return 1;
  printf("sanitized driver ok\n");
  return 0;
}
