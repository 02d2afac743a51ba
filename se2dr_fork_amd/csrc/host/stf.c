/*
 * stf.c -- extensible source-time-function objects (the reference's
 * SeismicSTF, se2wave.c:59-64, 3431-3468) with the Ricker, Gaussian and
 * triangle-regularised Yoffe implementations (3499-3822).  New STFs are
 * added by filling the {name, data, evaluate, destroy} object.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "host_internal.h"

#define STF_PI 3.14159265358979323846264338327950288

int se2w_stf_create(const char *name, SE2WSTF *out)
{
  SE2WSTF stf;
  if (!out) return host_fail(SE2W_HOST_ERR_ARG, "null output");
  stf = (SE2WSTF)calloc(1, sizeof(*stf));
  if (!stf) return host_fail(SE2W_HOST_ERR_MEM, "out of memory");
  snprintf(stf->name, sizeof(stf->name), "%s", name ? name : "");
  *out = stf;
  return 0;
}

int se2w_stf_evaluate(SE2WSTF stf, double time, double *psi)
{
  if (!stf || !stf->evaluate) return host_fail(SE2W_HOST_ERR_SUP, "A non-NULL source-time function evaluator must be provided");
  return stf->evaluate(stf, time, psi);
}

int se2w_stf_destroy(SE2WSTF *sp)
{
  int ierr = 0;
  if (!sp || !*sp) return 0;
  if ((*sp)->destroy) ierr = (*sp)->destroy(*sp);
  free(*sp);
  *sp = NULL;
  return ierr;
}

static int stf_free_data(SE2WSTF stf)
{
  free(stf->data);
  stf->data = NULL;
  return 0;
}

/* Ricker: psi = amp (1 - 2 a^2) exp(-a^2), a = pi f (t - t0) (3504-3516). */
typedef struct { double t0, freq, amp; } ricker_t;

static int ricker_eval(SE2WSTF stf, double time, double *psi)
{
  const ricker_t *r = (const ricker_t *)stf->data;
  const double arg = STF_PI * r->freq * (time - r->t0);
  const double arg2 = arg * arg;
  const double a = 1.0 - 2.0 * arg2;
  const double b = exp(-arg2);
  *psi = r->amp * a * b;
  return 0;
}

int se2w_stf_create_ricker(double t0, double freq, double amp, SE2WSTF *out)
{
  ricker_t *r;
  int ierr = se2w_stf_create("ricker", out);
  if (ierr) return ierr;
  r = (ricker_t *)malloc(sizeof(*r));
  if (!r) { se2w_stf_destroy(out); return host_fail(SE2W_HOST_ERR_MEM, "out of memory"); }
  r->t0 = t0; r->freq = freq; r->amp = amp;
  (*out)->data = r;
  (*out)->evaluate = ricker_eval;
  (*out)->destroy = stf_free_data;
  return 0;
}

/* Gaussian: psi = amp exp(-omega^2 (t - t0)^2 / 2) (3551-3561). */
typedef struct { double t0, omega, amp; } gauss_t;

static int gauss_eval(SE2WSTF stf, double time, double *psi)
{
  const gauss_t *g = (const gauss_t *)stf->data;
  const double arg = g->omega * g->omega * (time - g->t0) * (time - g->t0) * 0.5;
  *psi = g->amp * exp(-arg);
  return 0;
}

int se2w_stf_create_gaussian(double t0, double omega, double amp, SE2WSTF *out)
{
  gauss_t *g;
  int ierr = se2w_stf_create("gaussian", out);
  if (ierr) return ierr;
  g = (gauss_t *)malloc(sizeof(*g));
  if (!g) { se2w_stf_destroy(out); return host_fail(SE2W_HOST_ERR_MEM, "out of memory"); }
  g->t0 = t0; g->omega = omega; g->amp = amp;
  (*out)->data = g;
  (*out)->evaluate = gauss_eval;
  (*out)->destroy = stf_free_data;
  return 0;
}

/* Triangle-regularised Yoffe (Tinti et al., BSSA 95(4), 2005, appendix),
 * se2wave.c:3592-3822.  K = 2/(pi tau_R tau_S^2); the piecewise C1..C6 terms
 * for the two regimes tau_R > 2 tau_S and tau_S < tau_R < 2 tau_S. */
typedef struct { double K, d_max, tau_s, tau_r; } yoffe_t;

static int yoffe_terms(const yoffe_t *y, double t, int which, double *out)
{
  const double ts = y->tau_s, tr = y->tau_r, tr2 = tr * tr;
  double t1, t2, t3, t4, t5;
  if (which == 1) { /* 3608-3630 */
    t1 = 0.5 * t + 0.25 * tr;
    t2 = t * (tr - t);
    t3 = (t - tr) * tr;
    t4 = t / tr;
    t5 = (fabs(t) < 0.0) ? 1.0e12 : (tr - t) / t;
  } else if (which == 3) { /* 3632-3654 */
    t1 = ts - t - 0.5 * tr;
    t2 = (t - ts) * (tr - t + ts);
    t3 = tr * (2.0 * tr - 2.0 * t + 2.0 * ts);
    t4 = (t - ts) / tr;
    t5 = (fabs(t - ts) < 0.0) ? 1.0e12 : (tr - t + ts) / (t - ts);
  } else { /* 4: 3656-3679 */
    t1 = -ts + 0.5 * t + 0.25 * tr;
    t2 = (t - 2.0 * ts) * (tr - t + 2.0 * ts);
    t3 = tr * (-tr + t - 2.0 * ts);
    t4 = (t - 2.0 * ts) / tr;
    t5 = (fabs(t - 2.0 * ts) < 0.0) ? 1.0e12 : (tr - t + 2.0 * ts) / (t - 2.0 * ts);
  }
  if (t2 < 0.0) return host_fail(SE2W_HOST_ERR_USER, "t2 negative");
  if (t4 < 0.0) return host_fail(SE2W_HOST_ERR_USER, "t4 negative");
  if (t5 < 0.0) return host_fail(SE2W_HOST_ERR_USER, "t5 negative");
  if (which == 1) *out = t1 * sqrt(t2) + t3 * asin(sqrt(t4)) - 0.75 * tr2 * atan(sqrt(t5));
  else if (which == 3) *out = t1 * sqrt(t2) + t3 * asin(sqrt(t4)) + 1.5 * tr2 * atan(sqrt(t5));
  else *out = t1 * sqrt(t2) + t3 * asin(sqrt(t4)) - 0.75 * tr2 * atan(sqrt(t5));
  return 0;
}

static int yoffe_eval(SE2WSTF stf, double t, double *psi)
{
  const yoffe_t *y = (const yoffe_t *)stf->data;
  const double ts = y->tau_s, tr = y->tau_r, tr2 = tr * tr;
  double C1 = 0.0, C2 = 0.0, C3 = 0.0, C4 = 0.0, C5 = 0.0, C6 = 0.0, phi;
  int ierr = 0;
  if (t >= (tr + 2.0 * ts)) { *psi = 0.0; return 0; }
  if (tr > 2.0 * ts) { /* mode 1 (3685-3715) */
    if ((t >= 0.0) && (t < ts)) {
      ierr = yoffe_terms(y, t, 1, &C1); C2 = (3.0 / 8.0) * STF_PI * tr2; phi = C1 + C2;
    } else if ((t >= ts) && (t < 2.0 * ts)) {
      ierr = yoffe_terms(y, t, 1, &C1); C2 = (3.0 / 8.0) * STF_PI * tr2;
      if (!ierr) ierr = yoffe_terms(y, t, 3, &C3);
      phi = C1 - C2 + C3;
    } else if ((t >= 2.0 * ts) && (t < tr)) {
      ierr = yoffe_terms(y, t, 1, &C1);
      if (!ierr) ierr = yoffe_terms(y, t, 3, &C3);
      if (!ierr) ierr = yoffe_terms(y, t, 4, &C4);
      phi = C1 + C3 + C4;
    } else if ((t >= tr) && (t < (tr + ts))) {
      ierr = yoffe_terms(y, t, 3, &C3);
      if (!ierr) ierr = yoffe_terms(y, t, 4, &C4);
      C5 = 0.5 * STF_PI * tr * (t - tr);
      phi = C5 + C3 + C4;
    } else if ((t >= (tr + ts)) && (t < (tr + 2.0 * ts))) {
      ierr = yoffe_terms(y, t, 4, &C4);
      C6 = 0.5 * STF_PI * tr * (2.0 * ts - t + tr);
      phi = C4 + C6;
    } else {
      phi = 0.0;
    }
  } else { /* mode 2 (3741-3770) */
    if ((t >= 0.0) && (t < ts)) {
      ierr = yoffe_terms(y, t, 1, &C1); C2 = (3.0 / 8.0) * STF_PI * tr2; phi = C1 + C2;
    } else if ((t >= ts) && (t < tr)) {
      ierr = yoffe_terms(y, t, 1, &C1); C2 = (3.0 / 8.0) * STF_PI * tr2;
      if (!ierr) ierr = yoffe_terms(y, t, 3, &C3);
      phi = C1 - C2 + C3;
    } else if ((t >= tr) && (t < 2.0 * ts)) {
      C2 = (3.0 / 8.0) * STF_PI * tr2;
      ierr = yoffe_terms(y, t, 3, &C3);
      C5 = 0.5 * STF_PI * tr * (t - tr);
      phi = C5 + C3 - C2;
    } else if ((t >= 2.0 * ts) && (t < (tr + ts))) {
      ierr = yoffe_terms(y, t, 3, &C3);
      if (!ierr) ierr = yoffe_terms(y, t, 4, &C4);
      C5 = 0.5 * STF_PI * tr * (t - tr);
      phi = C5 + C3 + C4;
    } else if ((t >= (tr + ts)) && (t < (tr + 2.0 * ts))) {
      ierr = yoffe_terms(y, t, 4, &C4);
      C6 = 0.5 * STF_PI * tr * (2.0 * ts - t + tr);
      phi = C4 + C6;
    } else {
      phi = 0.0;
    }
  }
  if (ierr) return ierr;
  *psi = y->K * phi;
  return 0;
}

int se2w_stf_create_yoffe(double tau_s, double tau_r, double d_max, SE2WSTF *out)
{
  yoffe_t *y;
  int ierr;
  if (!(tau_r > 2.0 * tau_s) && !((tau_s < tau_r) && (tau_r < 2.0 * tau_s)))
    return host_fail(SE2W_HOST_ERR_USER, "Invalid values for tau_S (%+1.4e) and tau_R(%+1.4e) provided", tau_s, tau_r);
  ierr = se2w_stf_create("reg_yoffe", out);
  if (ierr) return ierr;
  y = (yoffe_t *)malloc(sizeof(*y));
  if (!y) { se2w_stf_destroy(out); return host_fail(SE2W_HOST_ERR_MEM, "out of memory"); }
  y->tau_s = tau_s; y->tau_r = tau_r; y->d_max = d_max;
  y->K = 2.0 / (STF_PI * tau_r * tau_s * tau_s);
  (*out)->data = y;
  (*out)->evaluate = yoffe_eval;
  (*out)->destroy = stf_free_data;
  return 0;
}
