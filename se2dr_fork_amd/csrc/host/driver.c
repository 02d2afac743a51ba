/*
 * driver.c -- the se2wave program: main (se2wave.c:4916-4951) and
 * se2wave_demo (4541-4807) with the step loop on the GPU.
 *
 * Reference options (defaults as in se2wave.c):
 *   -mx 8, -my (= mx), -bdegree 2, -tmax 0.4, -dt (CFL estimate x factor),
 *   -nt (int)(tmax/dt)+4, -of 5000, -dump_ic_src, -ignore_receiver_output,
 *   -use_nearest_internal_gll true, -verify_stf (-stf_ricker|-stf_yoffe).
 * Options this program adds to reach the benchmark workloads:
 *   -cfl_factor 0.2         dt = factor * CFL (README: 0.2; the demo code uses 0.5)
 *   -precision fp64|fp32|fp32k, -schedule fused|split
 *   -scale sx,sy -shift tx,ty  domain = [0,1]^2 * scale + shift (demo: 2000,-1000)
 *   -material_vp/-material_vs/-material_rho   constant material
 *   -material_perturbed vp0,dvp,vs0,dvs,rho0,drho    the reference's RAND48 perturbed model
 *   -material_random seed,vp_lo,vp_hi,rho_lo,rho_hi   per-cell, Vs = Vp/sqrt(3)
 *   -material_layers n,seed,vp_lo,vp_hi,rho_lo,rho_hi horizontal layers
 *   -src_xy x0,y0,...  -src_type moment|force  -src_values v0..v3 per source
 *   -src_ricker t0,f,amp per source          (replaces the demo sources)
 *   -rec_xy x0,y0,...                          (replaces the demo receivers)
 *   -source_impl 0|1|2|3   the form of every source, numbered as
 *                          ElastoDynamicsSetSourceImplementation (2202-2229)
 *                          numbers them: pointwise, closest internal basis
 *                          (the demo's), cubic spline, P0; -sm_h h sets the
 *                          cubic-spline kernel width (1975)
 *   -output_dir DIR
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "host_internal.h"
#include "vtk.h"

#define CHK(call)                                                                 \
  do {                                                                            \
    int ierr_ = (call);                                                           \
    if (ierr_) {                                                                  \
      fprintf(stderr, "[se2wave] error %d at %s:%d: %s\n", ierr_, __FILE__, __LINE__, \
              se2w_host_last_error());                                            \
      return ierr_;                                                               \
    }                                                                             \
  } while (0)

/* -source_impl (ElastoDynamicsSourceSetup, 2238-2242): the reference's
 * numbering 0..3 to the SeismicSource implementation, with the lines
 * ElastoDynamicsSetSourceImplementation prints (2211-2222). */
static int source_impl_option(SE2WOptions opts, int *impl)
{
  int v = 1, found = 0;
  *impl = SE2W_SOURCE_IMPL_NEAREST_QPOINT;
  CHK(se2w_options_get_int(opts, "-source_impl", &v, &found));
  if (!found) return 0;
  printf("  [source implementation]: processed from command line args\n");
  switch (v) {
    case 0: printf("  [source implementation]: pointwise evaluation\n"); *impl = SE2W_SOURCE_IMPL_POINTWISE; break;
    case 1: printf("  [source implementation]: closest internal basis\n"); *impl = SE2W_SOURCE_IMPL_NEAREST_QPOINT; break;
    case 2: printf("  [source implementation]: cspline kernel\n"); *impl = SE2W_SOURCE_IMPL_SPLINE; break;
    case 3: printf("  [source implementation]: P0\n"); *impl = SE2W_SOURCE_IMPL_P0; break;
    default: fprintf(stderr, "[se2wave] -source_impl %d: valid choices are 0,1,2,3\n", v); return SE2W_HOST_ERR_SUP;
  }
  return 0;
}

/* CHK for the step loop: a snapshot still being written is waited for
   before the error returns (its thread reads the GPU context). */
#define CHK_JOB(call)                                                             \
  do {                                                                            \
    int ierr_ = (call);                                                           \
    if (ierr_) {                                                                  \
      fprintf(stderr, "[se2wave] error %d at %s:%d: %s\n", ierr_, __FILE__, __LINE__, \
              se2w_host_last_error());                                            \
      se2w_vtk_job_join(&vtk_job);                                                \
      return ierr_;                                                               \
    }                                                                             \
  } while (0)

#define MAX_SRC 64
#define MAX_REC 4096

static int verify_stf(SE2WOptions opts)
{
  SE2WSTF stf = NULL;
  int is_ricker = 0, is_yoffe = 0, k;
  double time = 0.0, tmax = 2.0, dt = 1.0e-6;
  char fname[256];
  FILE *fp;
  CHK(se2w_options_get_bool(opts, "-stf_ricker", &is_ricker, NULL));
  CHK(se2w_options_get_bool(opts, "-stf_yoffe", &is_yoffe, NULL));
  if (!is_ricker && !is_yoffe) {
    fprintf(stderr, "[se2wave] No STF implementation specified. Use one of { -stf_ricker , -stf_yoffe }\n");
    return SE2W_HOST_ERR_USER;
  }
  if (is_ricker) { /* 4830-4851: the options are read and printed, the STF uses 0.15/12/1 */
    double t0 = 0.15, freq = 12.0, amp = 1.0;
    CHK(se2w_options_get_real(opts, "-stf.ricker.t0", &t0, NULL));
    CHK(se2w_options_get_real(opts, "-stf.ricker.freq", &freq, NULL));
    CHK(se2w_options_get_real(opts, "-stf.ricker.amp", &amp, NULL));
    CHK(se2w_stf_create_ricker(0.15, 12.0, 1.0, &stf));
    snprintf(fname, sizeof(fname), "%s", "SeismicSTF_Ricker.log");
    fp = fopen(fname, "w");
    if (!fp) return SE2W_HOST_ERR_FILE;
    fprintf(fp, "# --------------------------------\n# SeismicSTF\n#   Parameters\n#     type  \"Ricker\"\n");
    fprintf(fp, "#     t0    %+1.12e (sec)\n", t0);
    fprintf(fp, "#     freq  %+1.12e (Hz)\n", freq);
    fprintf(fp, "#     amp   %+1.12e\n", amp);
    fprintf(fp, "# --------------------------------\n");
  } else {
    double tau_s = 0.15, tau_r = 1.0, d_max = 1.1;
    CHK(se2w_options_get_real(opts, "-stf.yoffe.tau_s", &tau_s, NULL));
    CHK(se2w_options_get_real(opts, "-stf.yoffe.tau_r", &tau_r, NULL));
    CHK(se2w_options_get_real(opts, "-stf.yoffe.d_max", &d_max, NULL));
    CHK(se2w_stf_create_yoffe(tau_s, tau_r, d_max, &stf));
    snprintf(fname, sizeof(fname), "%s", "SeismicSTF_Yoffe.log");
    fp = fopen(fname, "w");
    if (!fp) return SE2W_HOST_ERR_FILE;
    fprintf(fp, "# --------------------------------\n# SeismicSTF\n#   Parameters\n#     type  \"Yoffe\"\n");
    fprintf(fp, "#     tau_S %+1.12e\n", tau_s);
    fprintf(fp, "#     tau_R %+1.12e\n", tau_r);
    fprintf(fp, "#     D_max %+1.12e\n", d_max);
    fprintf(fp, "# --------------------------------\n");
  }
  CHK(se2w_options_get_real(opts, "-tmax", &tmax, NULL));
  printf("[se2wave] Requested time period: %1.4e\n", tmax);
  CHK(se2w_options_get_real(opts, "-dt", &dt, NULL));
  printf("[se2wave] Using time step size: %1.4e\n", dt);
  for (k = 0;; k++) {
    double v;
    CHK(se2w_stf_evaluate(stf, time, &v));
    fprintf(fp, "%+1.6e %+1.12e\n", time, v);
    time = time + dt;
    if (time >= tmax) break;
  }
  printf("[step %9d] time = %1.4e : dt = %1.4e \n", k, time, dt);
  fclose(fp);
  se2w_stf_destroy(&stf);
  return 0;
}

static void print_norms(se2w_ctx *ctx, int k, double time, double dt)
{
  double n[6];
  printf("[step %9d] time = %1.4e : dt = %1.4e \n", k, time, dt);
  if (se2w_norms(ctx, n) == 0) {
    printf("  [displacement] max = %+1.4e : min = %+1.4e : l2 = %+1.4e \n", n[0], n[1], n[2]);
    printf("  [velocity]     max = %+1.4e : min = %+1.4e : l2 = %+1.4e \n", n[3], n[4], n[5]);
  }
}

static int parse_precision(SE2WOptions opts, int *prec, int *sched)
{
  char buf[64];
  int found;
  *prec = SE2W_FP64;
  *sched = SE2W_SCHED_FUSED;
  CHK(se2w_options_get_string(opts, "-precision", buf, sizeof(buf), &found));
  if (found) {
    if (!strcmp(buf, "fp64")) *prec = SE2W_FP64;
    else if (!strcmp(buf, "fp32")) *prec = SE2W_FP32;
    else if (!strcmp(buf, "fp32k") || !strcmp(buf, "fp32_kahan")) *prec = SE2W_FP32_KAHAN;
    else { fprintf(stderr, "[se2wave] unknown -precision %s\n", buf); return SE2W_HOST_ERR_ARG; }
  }
  CHK(se2w_options_get_string(opts, "-schedule", buf, sizeof(buf), &found));
  if (found) {
    if (!strcmp(buf, "fused")) *sched = SE2W_SCHED_FUSED;
    else if (!strcmp(buf, "split")) *sched = SE2W_SCHED_SPLIT;
    else { fprintf(stderr, "[se2wave] unknown -schedule %s\n", buf); return SE2W_HOST_ERR_ARG; }
  }
  return 0;
}

static int setup_material(SE2WOptions opts, SE2WSpec spec, int mx, int my)
{
  double arr[8];
  int n = 8, found;
  CHK(se2w_options_get_real_array(opts, "-material_perturbed", arr, &n, &found));
  if (found) {
    if (n != 6) { fprintf(stderr, "[se2wave] -material_perturbed vp0,dvp,vs0,dvs,rho0,drho\n"); return SE2W_HOST_ERR_ARG; }
    return se2w_spec_set_perturbed_material_velocity(spec, arr[0], arr[1], arr[2], arr[3], arr[4], arr[5]);
  }
  n = 8;
  CHK(se2w_options_get_real_array(opts, "-material_random", arr, &n, &found));
  if (found) {
    double *vp = malloc(sizeof(double) * mx * my), *vs = malloc(sizeof(double) * mx * my), *rho = malloc(sizeof(double) * mx * my);
    if (n != 5 || !vp || !vs || !rho) { free(vp); free(vs); free(rho); fprintf(stderr, "[se2wave] -material_random seed,vp_lo,vp_hi,rho_lo,rho_hi\n"); return SE2W_HOST_ERR_ARG; }
    CHK(se2w_material_random_cells(mx, my, (uint64_t)arr[0], arr[1], arr[2], arr[3], arr[4], vp, vs, rho));
    CHK(se2w_spec_set_cell_material_velocity(spec, vp, vs, rho));
    printf("  [material] per-cell random: seed %d Vp U[%g,%g] rho U[%g,%g] Vs = Vp/sqrt(3)\n", (int)arr[0], arr[1], arr[2], arr[3], arr[4]);
    free(vp); free(vs); free(rho);
    return 0;
  }
  n = 8;
  CHK(se2w_options_get_real_array(opts, "-material_layers", arr, &n, &found));
  if (found) {
    double *vp = malloc(sizeof(double) * mx * my), *vs = malloc(sizeof(double) * mx * my), *rho = malloc(sizeof(double) * mx * my);
    if (n != 6 || !vp || !vs || !rho) { free(vp); free(vs); free(rho); fprintf(stderr, "[se2wave] -material_layers n,seed,vp_lo,vp_hi,rho_lo,rho_hi\n"); return SE2W_HOST_ERR_ARG; }
    CHK(se2w_material_layered(mx, my, (int)arr[0], (uint64_t)arr[1], arr[2], arr[3], arr[4], arr[5], vp, vs, rho));
    CHK(se2w_spec_set_cell_material_velocity(spec, vp, vs, rho));
    printf("  [material] %d layers: seed %d Vp U[%g,%g] rho U[%g,%g] Vs = Vp/sqrt(3)\n", (int)arr[0], (int)arr[1], arr[2], arr[3], arr[4], arr[5]);
    free(vp); free(vs); free(rho);
    return 0;
  }
  {
    double vp = 4746.3670317412243, vs = 2740.2554625435928, rho = 1000.0; /* 4592 */
    double mu, lambda;
    CHK(se2w_options_get_real(opts, "-material_vp", &vp, NULL));
    CHK(se2w_options_get_real(opts, "-material_vs", &vs, NULL));
    CHK(se2w_options_get_real(opts, "-material_rho", &rho, NULL));
    CHK(se2w_spec_set_constant_material_velocity(spec, vp, vs, rho));
    mu = vs * vs * rho;
    lambda = vp * vp * rho - 2.0 * mu;
    printf("  [material]     Vp = %1.8e\n", vp);
    printf("  [material]     Vs = %1.8e\n", vs);
    printf("  [material] lambda = %1.8e\n", lambda);
    printf("  [material]     mu = %1.8e\n", mu);
    printf("  [material]    rho = %1.8e\n", rho);
  }
  return 0;
}

static int run_demo(SE2WOptions opts, int mx, int my)
{
  SE2WSpec spec = NULL;
  SE2WSource src[MAX_SRC];
  SE2WSTF stf[MAX_SRC];
  SE2WReceivers recv = NULL;
  SE2WSolver solver = NULL;
  double scale[2] = {2.0e3, 2.0e3}, shift[2] = {-1.0e3, -1.0e3}, arr[4 * MAX_SRC];
  double dt, time_max = 0.4, cfl_factor = 0.2, *xr, *times = NULL, *trace = NULL;
  int p = 2, nsources, nrecv, nt, of = 5000, found, n, s, k, prec, sched, use_internal = 1;
  int dump_ic_src = 0, ignore_receiver_output = 0, nsteps_total, impl, found_h = 0;
  double sm_h = 0.0;
  char outdir[512] = ".";
  SE2WVtkJob vtk_job;

  memset(src, 0, sizeof(src));
  memset(stf, 0, sizeof(stf));
  CHK(se2w_options_get_bool(opts, "-dump_ic_src", &dump_ic_src, NULL));
  CHK(se2w_options_get_bool(opts, "-ignore_receiver_output", &ignore_receiver_output, NULL));
  CHK(se2w_options_get_bool(opts, "-use_nearest_internal_gll", &use_internal, NULL));
  CHK(se2w_options_get_string(opts, "-output_dir", outdir, sizeof(outdir), NULL));
  CHK(parse_precision(opts, &prec, &sched));
  CHK(source_impl_option(opts, &impl));
  CHK(se2w_options_get_real(opts, "-sm_h", &sm_h, &found_h));
  CHK(se2w_options_get_int(opts, "-bdegree", &p, NULL));
  CHK(se2w_spec_create(mx, my, p, &spec));
  n = 2;
  CHK(se2w_options_get_real_array(opts, "-scale", scale, &n, &found));
  n = 2;
  CHK(se2w_options_get_real_array(opts, "-shift", shift, &n, &found));
  CHK(se2w_spec_scale_coords(spec, scale, shift));
  CHK(setup_material(opts, spec, mx, my));

  if (dump_ic_src) {
    const int64_t nn = (int64_t)(p * mx + 1) * (p * my + 1);
    double *zero = calloc(2 * nn, sizeof(double));
    if (!zero) return SE2W_HOST_ERR_MEM;
    CHK(se2w_vtk_write(spec, outdir, "uva.vts", "disp", zero, NULL, NULL));
    free(zero);
  }

  /* sources: the demo's two moment sources (4613-4652) unless -src_xy given */
  n = 2 * MAX_SRC;
  CHK(se2w_options_get_real_array(opts, "-src_xy", arr, &n, &found));
  if (!found) {
    const double c0[2] = {0.1, 100.1}, m0[4] = {1000.0, 0.0, 0.0, 1000.0};
    const double c1[2] = {400.1, 400.1}, m1[4] = {0.0, -1000.0, 1000.0, 0.0};
    nsources = 2;
    printf("Moment: [ %+1.2e , %+1.2e ; %+1.2e , %+1.2e ]\n", m0[0], m0[1], m0[2], m0[3]);
    CHK(se2w_source_create(spec, SE2W_SOURCE_MOMENT, impl, c0, m0, use_internal, &src[0]));
    printf("Moment: [ %+1.2e , %+1.2e ; %+1.2e , %+1.2e ]\n", m1[0], m1[1], m1[2], m1[3]);
    CHK(se2w_source_create(spec, SE2W_SOURCE_MOMENT, impl, c1, m1, use_internal, &src[1]));
    CHK(se2w_stf_create_ricker(0.15, 12.0, 1.0, &stf[0]));
    CHK(se2w_stf_create_ricker(0.25, 4.0, 1.0, &stf[1]));
  } else {
    double vals[4 * MAX_SRC], ricker[3 * MAX_SRC];
    char stype[32] = "moment";
    int type, nv = 4 * MAX_SRC, nr3 = 3 * MAX_SRC, fv, fr;
    if (n % 2) { fprintf(stderr, "[se2wave] -src_xy needs pairs\n"); return SE2W_HOST_ERR_ARG; }
    nsources = n / 2;
    CHK(se2w_options_get_string(opts, "-src_type", stype, sizeof(stype), NULL));
    type = !strcmp(stype, "force") ? SE2W_SOURCE_FORCE : SE2W_SOURCE_MOMENT;
    CHK(se2w_options_get_real_array(opts, "-src_values", vals, &nv, &fv));
    CHK(se2w_options_get_real_array(opts, "-src_ricker", ricker, &nr3, &fr));
    if (!fv || nv != 4 * nsources || !fr || nr3 != 3 * nsources) {
      fprintf(stderr, "[se2wave] -src_values needs 4 values and -src_ricker 3 values per source\n");
      return SE2W_HOST_ERR_ARG;
    }
    for (s = 0; s < nsources; s++) {
      CHK(se2w_source_create(spec, type, impl, &arr[2 * s], &vals[4 * s], use_internal, &src[s]));
      CHK(se2w_stf_create_ricker(ricker[3 * s], ricker[3 * s + 1], ricker[3 * s + 2], &stf[s]));
    }
  }
  for (s = 0; s < nsources; s++) {
    int32_t e, g;
    if (impl == SE2W_SOURCE_IMPL_SPLINE && found_h) CHK(se2w_source_set_kernel_width(src[s], sm_h));
    CHK(se2w_source_setup(src[s]));
    CHK(se2w_source_get(src[s], &e, &g, NULL));
    printf("[SeismicSource] source --> element %d closest_gll %d\n", (int)e, (int)g);
    if (impl == SE2W_SOURCE_IMPL_SPLINE) { /* the lines of 1971, 1977 and 2054 */
      double h, int_w;
      int32_t ne;
      CHK(se2w_source_get_kernel(src[s], &h, &int_w));
      CHK(se2w_source_get_elements(src[s], &ne, NULL, NULL));
      if (found_h) printf("kernel_h = %1.4e <from options>\n", h);
      else if (p > 3) printf("kernel_h = %1.4e <with pfac>\n", h);
      printf("\\int W = %1.14e \n", int_w);
      printf("[SeismicSource] cubic-spline source over %d elements\n", (int)ne);
    }
  }
  if (dump_ic_src) { /* 4663-4669: the sources with STF = 1 */
    const int64_t nn = (int64_t)(p * mx + 1) * (p * my + 1);
    double *g = calloc(2 * nn, sizeof(double));
    if (!g) return SE2W_HOST_ERR_MEM;
    CHK(se2w_sources_evaluate(0.0, nsources, src, NULL, g));
    CHK(se2w_vtk_write(spec, outdir, "f.vts", "g", g, NULL, NULL));
    free(g);
  }

  /* receivers (4688-4697) */
  n = 2 * MAX_REC;
  xr = malloc(sizeof(double) * 2 * MAX_REC);
  if (!xr) return SE2W_HOST_ERR_MEM;
  CHK(se2w_options_get_real_array(opts, "-rec_xy", xr, &n, &found));
  if (!found) {
    nrecv = 3;
    xr[0] = 100.0; xr[1] = 200.0; xr[2] = 150.0; xr[3] = 200.0; xr[4] = 180.0; xr[5] = 200.0;
  } else {
    nrecv = n / 2;
  }
  CHK(se2w_receivers_create(spec, nrecv, xr, &recv));

  CHK(se2w_options_get_real(opts, "-tmax", &time_max, NULL));
  printf("[se2wave] Requested time period: %1.4e\n", time_max);
  CHK(se2w_spec_timestep(spec, &dt));
  CHK(se2w_options_get_real(opts, "-cfl_factor", &cfl_factor, NULL));
  dt = dt * cfl_factor;
  CHK(se2w_options_get_real(opts, "-dt", &dt, NULL));
  printf("[se2wave] Using time step size: %1.4e\n", dt);
  nt = (int)(time_max / dt) + 4;
  CHK(se2w_options_get_int(opts, "-nt", &nt, NULL));
  printf("[se2wave] Estimated number of time steps: %d\n", nt);
  CHK(se2w_options_get_int(opts, "-of", &of, NULL));
  if (of < 1) of = 1;

  /* the loop stops after step k when time >= tmax (4772) */
  {
    double t = 0.0;
    nsteps_total = nt;
    for (k = 1; k <= nt; k++) {
      t = t + dt;
      if (t >= time_max) { nsteps_total = k; break; }
    }
  }

  CHK(se2w_solver_create(spec, prec, sched, nsources, src, stf, recv, &solver));
  if (!ignore_receiver_output) CHK(se2w_receivers_write_header(recv, outdir));
  times = malloc(sizeof(double) * 100);
  trace = malloc(sizeof(double) * 100 * (nrecv > 0 ? nrecv : 1) * 7);
  if (!times || !trace) return SE2W_HOST_ERR_MEM;
  k = 0;
  /* Snapshots (step-%.4d.vts every -of steps and at the end, 4764-4770,
     4785-4789) are written by a second thread while the next steps run. */
  memset(&vtk_job, 0, sizeof(vtk_job));
  while (k < nsteps_total) {
    /* advance to the next step at which the reference prints or writes */
    int next = nsteps_total, chunk;
    if ((k / 100 + 1) * 100 < next) next = (k / 100 + 1) * 100;
    if ((k / of + 1) * of < next) next = (k / of + 1) * of;
    chunk = next - k;
    CHK_JOB(se2w_solver_advance(solver, chunk, dt));
    if (!ignore_receiver_output && nrecv > 0) {
      double *tt = malloc(sizeof(double) * chunk), *tr = malloc(sizeof(double) * (size_t)chunk * nrecv * 7);
      if (!tt || !tr) { se2w_vtk_job_join(&vtk_job); return SE2W_HOST_ERR_MEM; }
      CHK_JOB(se2w_solver_get_trace(solver, tt, tr));
      CHK_JOB(se2w_receivers_append_rows(recv, chunk, tt, tr));
      free(tt); free(tr);
    }
    k = next;
    if (k % 100 == 0) print_norms(se2w_solver_context(solver), k, se2w_solver_time(solver), dt);
    if (k % of == 0) {
      char fname[64];
      snprintf(fname, sizeof(fname), "step-%.4d.vts", k);
      CHK_JOB(se2w_vtk_job_start(&vtk_job, spec, se2w_solver_context(solver), outdir, fname));
    }
  }
  print_norms(se2w_solver_context(solver), k, se2w_solver_time(solver), dt);
  {
    char fname[64];
    snprintf(fname, sizeof(fname), "step-%.4d.vts", k);
    CHK_JOB(se2w_vtk_job_start(&vtk_job, spec, se2w_solver_context(solver), outdir, fname));
    CHK(se2w_vtk_job_join(&vtk_job));
  }
  free(times); free(trace); free(xr);
  se2w_solver_destroy(&solver);
  se2w_receivers_destroy(&recv);
  for (s = 0; s < nsources; s++) { se2w_source_destroy(&src[s]); se2w_stf_destroy(&stf[s]); }
  se2w_spec_destroy(&spec);
  return 0;
}

int se2w_main(int argc, char **argv)
{
  SE2WOptions opts = NULL;
  int mx = 8, my, verify = 0, ierr;
  CHK(se2w_options_create(argc, argv, &opts));
  CHK(se2w_options_get_int(opts, "-mx", &mx, NULL));
  my = mx;
  CHK(se2w_options_get_int(opts, "-my", &my, NULL));
  CHK(se2w_options_get_bool(opts, "-verify_stf", &verify, NULL));
  if (verify) ierr = verify_stf(opts);
  else ierr = run_demo(opts, mx, my);
  se2w_options_destroy(&opts);
  return ierr;
}
