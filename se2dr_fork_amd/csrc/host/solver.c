/*
 * solver.c -- drives the GPU step (include/se2w.h) from the host objects:
 * the se2wave_demo time loop (se2wave.c:4722-4775) with the source-time
 * functions evaluated on the host at the accumulated time
 * (time = time + dt, 4724) and the receiver rows collected per advance.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "host_internal.h"

struct se2w_solver {
  struct se2w_spec *spec;
  se2w_ctx *ctx;
  int ns;            /* sources handed to the GPU (those with a source and an STF) */
  SE2WSource *src;   /* [ns] */
  SE2WSTF *stf;      /* [ns] */
  int ngpu;          /* element sources on the GPU: one per source, one per element of a cubic-spline source */
  int *parent;       /* [ngpu] the source each belongs to (its column of the source-time table) */
  int one_each;      /* every source on exactly one element: parent[j] = j */
  double *stfv;      /* [ns] the STF values of one step */
  SE2WReceivers recv;
  double time;
  int steps;
  int last_n;
  double *times, *trace, *psi;
  size_t times_cap, trace_cap, psi_cap;
  int nuv;            /* single-station receivers */
  double *uv_weights; /* [nuv][npe] */
  double *uv_trace;   /* [steps][nuv][4] */
  size_t uv_cap;
};

static int gpu_fail(int code)
{
  return host_fail(code, "GPU: %s (%s)", se2w_strerror(code), se2w_last_error());
}

int se2w_solver_create(SE2WSpec spec, int precision, int schedule, int ns, SE2WSource *src, SE2WSTF *stf,
                       SE2WReceivers recv, SE2WSolver *out)
{
  struct se2w_solver *sv;
  se2w_desc desc;
  double D[HOST_MAX_NB * HOST_MAX_NB], w1d[HOST_MAX_NB], *edx = NULL, *edy = NULL, *mass = NULL, *vals = NULL;
  int32_t *elem = NULL;
  int ierr, p, k;
  if (!spec || !out || ns < 0) return host_fail(SE2W_HOST_ERR_ARG, "bad solver arguments");
  *out = NULL;
  if (!spec->have_material) return host_fail(SE2W_HOST_ERR_USER, "material not set");
  sv = (struct se2w_solver *)calloc(1, sizeof(*sv));
  if (!sv) return host_fail(SE2W_HOST_ERR_MEM, "out of memory");
  sv->spec = spec;
  sv->recv = recv;
  sv->src = (SE2WSource *)calloc(ns > 0 ? ns : 1, sizeof(SE2WSource));
  sv->stf = (SE2WSTF *)calloc(ns > 0 ? ns : 1, sizeof(SE2WSTF));
  if (!sv->src || !sv->stf) { ierr = host_fail(SE2W_HOST_ERR_MEM, "out of memory"); goto bad; }
  /* SeismicSourceEvaluate skips a source without an STF (3485-3490) */
  for (p = 0; p < ns; p++) {
    if (!src || !src[p] || !stf || !stf[p]) continue;
    if (!src[p]->issetup) { ierr = host_fail(SE2W_HOST_ERR_USER, "Seismic source not setup. Must call SeismicSourceSetup()"); goto bad; }
    sv->src[sv->ns] = src[p];
    sv->stf[sv->ns] = stf[p];
    sv->ns++;
  }
  desc.mx = spec->mx; desc.my = spec->my; desc.p = spec->p;
  desc.precision = precision; desc.schedule = schedule; desc.device = 0;
  if ((ierr = se2w_create(&desc, &sv->ctx))) { ierr = gpu_fail(ierr); goto bad; }
  edx = (double *)malloc(sizeof(double) * spec->mx);
  edy = (double *)malloc(sizeof(double) * spec->my);
  mass = (double *)malloc(sizeof(double) * (size_t)spec->nnodes);
  if (!edx || !edy || !mass) { ierr = host_fail(SE2W_HOST_ERR_MEM, "out of memory"); goto bad; }
  se2w_spec_get_tables(spec, D, w1d, edx, edy);
  if ((ierr = se2w_set_tables(sv->ctx, D, w1d, edx, edy))) { ierr = gpu_fail(ierr); goto bad; }
  if ((ierr = se2w_set_cells(sv->ctx, spec->lambda, spec->mu))) { ierr = gpu_fail(ierr); goto bad; }
  if ((ierr = se2w_spec_mass(spec, mass))) goto bad;
  if ((ierr = se2w_set_mass(sv->ctx, mass))) { ierr = gpu_fail(ierr); goto bad; }
  if (sv->ns > 0) {
    /* every source as element sources in source order; the GPU adds them in
       that order, each with its source's STF value */
    int j = 0;
    for (k = 0; k < sv->ns; k++) {
      int32_t ne;
      if ((ierr = se2w_source_get_elements(sv->src[k], &ne, NULL, NULL))) goto bad;
      sv->ngpu += ne;
    }
    elem = (int32_t *)malloc(sizeof(int32_t) * (sv->ngpu > 0 ? sv->ngpu : 1));
    vals = (double *)malloc(sizeof(double) * 2 * spec->npe * (sv->ngpu > 0 ? sv->ngpu : 1));
    sv->parent = (int *)malloc(sizeof(int) * (sv->ngpu > 0 ? sv->ngpu : 1));
    sv->stfv = (double *)malloc(sizeof(double) * sv->ns);
    if (!elem || !vals || !sv->parent || !sv->stfv) { ierr = host_fail(SE2W_HOST_ERR_MEM, "out of memory"); goto bad; }
    for (k = 0; k < sv->ns; k++) {
      int32_t ne, q;
      if ((ierr = se2w_source_get_elements(sv->src[k], &ne, &elem[j], &vals[(size_t)2 * spec->npe * j]))) goto bad;
      for (q = 0; q < ne; q++) sv->parent[j + q] = k;
      j += ne;
    }
    sv->one_each = sv->ngpu == sv->ns;
    for (k = 0; k < sv->ngpu && sv->one_each; k++) sv->one_each = sv->parent[k] == k;
    if (sv->ngpu > 0 && (ierr = se2w_set_sources(sv->ctx, sv->ngpu, elem, vals))) { ierr = gpu_fail(ierr); goto bad; }
  }
  if (recv && recv->nr > 0) {
    if ((ierr = se2w_set_receivers(sv->ctx, recv->nr, recv->elem, recv->gll))) { ierr = gpu_fail(ierr); goto bad; }
  }
  free(edx); free(edy); free(mass); free(elem); free(vals);
  /* Pick the fastest element-kernel schedule for this mesh (SE2W_AUTOTUNE=0
     keeps the default; small meshes keep it too, tuning them costs more than
     it saves). */
  {
    const char *env = getenv("SE2W_AUTOTUNE");
    const int want = env ? atoi(env) : (spec->nnodes >= 65536);
    if (want) {
      int32_t best;
      if ((ierr = se2w_autotune(sv->ctx, 20, 3, &best, NULL))) { ierr = gpu_fail(ierr); goto bad; }
    }
  }
  *out = sv;
  return 0;
bad:
  free(edx); free(edy); free(mass); free(elem); free(vals);
  se2w_solver_destroy(&sv);
  return ierr;
}

int se2w_solver_destroy(SE2WSolver *sp)
{
  if (!sp || !*sp) return 0;
  if ((*sp)->ctx) se2w_destroy((*sp)->ctx);
  free((*sp)->src); free((*sp)->stf); free((*sp)->parent); free((*sp)->stfv);
  free((*sp)->times); free((*sp)->trace); free((*sp)->psi);
  free((*sp)->uv_weights); free((*sp)->uv_trace);
  free(*sp);
  *sp = NULL;
  return 0;
}

se2w_ctx *se2w_solver_context(SE2WSolver sv) { return sv ? sv->ctx : NULL; }
double se2w_solver_time(SE2WSolver sv) { return sv ? sv->time : 0.0; }
int se2w_solver_steps_done(SE2WSolver sv) { return sv ? sv->steps : 0; }

static int grow(double **p, size_t *cap, size_t n)
{
  if (n <= *cap) return 0;
  free(*p);
  *p = (double *)malloc(sizeof(double) * n);
  if (!*p) { *cap = 0; return host_fail(SE2W_HOST_ERR_MEM, "out of memory"); }
  *cap = n;
  return 0;
}

/* Single-station rows of the last run (weighted sums over element nodes). */
static int collect_uv(struct se2w_solver *sv, int nsteps)
{
  int ierr;
  if (sv->nuv == 0) return 0;
  if ((ierr = grow(&sv->uv_trace, &sv->uv_cap, (size_t)nsteps * sv->nuv * 4))) return ierr;
  if ((ierr = se2w_get_element_receivers(sv->ctx, sv->uv_weights, sv->uv_trace))) return gpu_fail(ierr);
  return 0;
}

int se2w_solver_set_uv_receivers(SE2WSolver sv, int n, SE2WReceiverUV *rs)
{
  int32_t *elem;
  double *w;
  int i, ierr;
  if (!sv || n < 0 || (n > 0 && !rs)) return host_fail(SE2W_HOST_ERR_ARG, "bad receiver arguments");
  elem = (int32_t *)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
  w = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1) * sv->spec->npe);
  if (!elem || !w) { free(elem); free(w); return host_fail(SE2W_HOST_ERR_MEM, "out of memory"); }
  for (i = 0; i < n; i++) {
    if (!rs[i] || rs[i]->spec != sv->spec) { free(elem); free(w); return host_fail(SE2W_HOST_ERR_ARG, "receiver %d does not belong to the solver's mesh", i); }
    se2w_receiver_uv_get(rs[i], &elem[i], &w[(size_t)i * sv->spec->npe], NULL);
  }
  if ((ierr = se2w_set_element_receivers(sv->ctx, n, elem))) { free(elem); free(w); return gpu_fail(ierr); }
  free(elem);
  free(sv->uv_weights);
  sv->uv_weights = w;
  sv->nuv = n;
  return 0;
}

int se2w_solver_get_uv_trace(SE2WSolver sv, double *trace)
{
  if (!sv) return host_fail(SE2W_HOST_ERR_ARG, "null solver");
  if (trace && sv->nuv > 0 && sv->last_n) memcpy(trace, sv->uv_trace, sizeof(double) * (size_t)sv->last_n * sv->nuv * 4);
  return 0;
}

int se2w_solver_reserve(SE2WSolver sv, int nsteps)
{
  const int nr = sv && sv->recv ? sv->recv->nr : 0;
  int ierr;
  if (!sv || nsteps < 1) return host_fail(SE2W_HOST_ERR_ARG, "bad reserve arguments");
  if ((ierr = grow(&sv->times, &sv->times_cap, (size_t)nsteps))) return ierr;
  if (sv->ngpu > 0 && (ierr = grow(&sv->psi, &sv->psi_cap, (size_t)nsteps * sv->ngpu))) return ierr;
  if (nr > 0 && (ierr = grow(&sv->trace, &sv->trace_cap, (size_t)nsteps * nr * 7))) return ierr;
  if (sv->nuv > 0 && (ierr = grow(&sv->uv_trace, &sv->uv_cap, (size_t)nsteps * sv->nuv * 4))) return ierr;
  if ((ierr = se2w_reserve(sv->ctx, nsteps))) return gpu_fail(ierr);
  return 0;
}

/* One row of the GPU's source-time table: each source's STF at time t
 * (SeismicSourceEvaluate, 3470-3497), repeated for each of its elements. */
static int stf_row(struct se2w_solver *sv, double t, double *row)
{
  int s, j, ierr;
  if (sv->one_each) {
    for (s = 0; s < sv->ns; s++)
      if ((ierr = se2w_stf_evaluate(sv->stf[s], t, &row[s]))) return ierr;
    return 0;
  }
  for (s = 0; s < sv->ns; s++)
    if ((ierr = se2w_stf_evaluate(sv->stf[s], t, &sv->stfv[s]))) return ierr;
  for (j = 0; j < sv->ngpu; j++) row[j] = sv->stfv[sv->parent[j]];
  return 0;
}

/* Monotonic wall clock in milliseconds (the SE2W_PHASE_TIMES trace). */
static double now_ms(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return 1e3 * (double)ts.tv_sec + 1e-6 * (double)ts.tv_nsec;
}

int se2w_solver_advance(SE2WSolver sv, int nsteps, double dt)
{
  const int nr = sv && sv->recv ? sv->recv->nr : 0;
  double t, tp[4];
  int k, s, ierr;
  if (!sv || nsteps < 1) return host_fail(SE2W_HOST_ERR_ARG, "bad advance arguments");
  if ((ierr = grow(&sv->times, &sv->times_cap, (size_t)nsteps))) return ierr;
  if (sv->ngpu > 0 && (ierr = grow(&sv->psi, &sv->psi_cap, (size_t)nsteps * sv->ngpu))) return ierr;
  if (nr > 0 && (ierr = grow(&sv->trace, &sv->trace_cap, (size_t)nsteps * nr * 7))) return ierr;
  t = sv->time;
  tp[0] = now_ms();
  for (k = 0; k < nsteps; k++) {
    t = t + dt;
    sv->times[k] = t;
    if ((ierr = stf_row(sv, t, &sv->psi[(size_t)k * sv->ngpu]))) return ierr;
  }
  tp[1] = now_ms();
  if ((ierr = se2w_run(sv->ctx, nsteps, dt, sv->psi))) return gpu_fail(ierr);
  tp[2] = now_ms();
  if (nr > 0 && (ierr = se2w_get_receivers(sv->ctx, sv->trace))) return gpu_fail(ierr);
  tp[3] = now_ms();
  if ((ierr = collect_uv(sv, nsteps))) return ierr;
  if (getenv("SE2W_PHASE_TIMES"))
    fprintf(stderr, "[se2w_solver_advance] %d steps: STF table %.3f ms, se2w_run %.3f ms, receiver rows %.3f ms, "
                    "single-station rows %.3f ms\n", nsteps, tp[1] - tp[0], tp[2] - tp[1], tp[3] - tp[2], now_ms() - tp[3]);
  sv->time = t;
  sv->steps += nsteps;
  sv->last_n = nsteps;
  return 0;
}

int se2w_solver_get_trace(SE2WSolver sv, double *times, double *trace)
{
  const int nr = sv && sv->recv ? sv->recv->nr : 0;
  if (!sv) return host_fail(SE2W_HOST_ERR_ARG, "null solver");
  if (times && sv->last_n) memcpy(times, sv->times, sizeof(double) * sv->last_n);
  if (trace && nr > 0 && sv->last_n) memcpy(trace, sv->trace, sizeof(double) * (size_t)sv->last_n * nr * 7);
  return 0;
}

/* Advances n solvers together: the host part of se2w_solver_advance for each
 * (accumulated time, source-time table), one batched GPU step per step for all
 * of them (se2w_ensemble_run), then each solver's receiver rows.  ens must have
 * been created from the solvers' contexts in the same order. */
int se2w_solvers_advance(int n, SE2WSolver *sv, se2w_ensemble *ens, int nsteps, const double *dt)
{
  const double **psi = NULL;
  int i, k, s, ierr;
  if (n < 1 || !sv || !ens || !dt || nsteps < 1) return host_fail(SE2W_HOST_ERR_ARG, "bad batched advance arguments");
  for (i = 0; i < n; i++) {
    se2w_ctx *c = NULL;
    if (!sv[i]) return host_fail(SE2W_HOST_ERR_ARG, "solver %d is null", i);
    if ((ierr = se2w_ensemble_get_member(ens, i, &c))) return gpu_fail(ierr);
    if (c != sv[i]->ctx) return host_fail(SE2W_HOST_ERR_ARG, "ensemble member %d is not solver %d's context", i, i);
  }
  psi = (const double **)calloc((size_t)n, sizeof(double *));
  if (!psi) return host_fail(SE2W_HOST_ERR_MEM, "out of memory");
  for (i = 0; i < n; i++) {
    struct se2w_solver *v = sv[i];
    const int nr = v->recv ? v->recv->nr : 0;
    double t = v->time;
    if ((ierr = grow(&v->times, &v->times_cap, (size_t)nsteps))) goto done;
    if (v->ngpu > 0 && (ierr = grow(&v->psi, &v->psi_cap, (size_t)nsteps * v->ngpu))) goto done;
    if (nr > 0 && (ierr = grow(&v->trace, &v->trace_cap, (size_t)nsteps * nr * 7))) goto done;
    for (k = 0; k < nsteps; k++) {
      t = t + dt[i];
      v->times[k] = t;
      if ((ierr = stf_row(v, t, &v->psi[(size_t)k * v->ngpu]))) goto done;
    }
    psi[i] = v->ngpu > 0 ? v->psi : NULL;
  }
  if ((ierr = se2w_ensemble_run(ens, nsteps, dt, psi))) { ierr = gpu_fail(ierr); goto done; }
  for (i = 0; i < n; i++) {
    struct se2w_solver *v = sv[i];
    const int nr = v->recv ? v->recv->nr : 0;
    if (nr > 0 && (ierr = se2w_get_receivers(v->ctx, v->trace))) { ierr = gpu_fail(ierr); goto done; }
    if ((ierr = collect_uv(v, nsteps))) goto done;
    v->time = v->times[nsteps - 1];
    v->steps += nsteps;
    v->last_n = nsteps;
  }
  ierr = 0;
done:
  free(psi);
  return ierr;
}
