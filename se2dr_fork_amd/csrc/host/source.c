/*
 * source.c -- point sources (the reference's SeismicSource, se2wave.c:66-90,
 * 2995-3429).  A source is located in its element, snapped to the nearest
 * (interior) GLL node for SOURCE_IMPL_NEAREST_QPOINT, and reduced to
 * 2*(p+1)^2 nodal values that the time loop scales by the STF value.
 *
 * Beyond the reference (which rejects them with SETERRQ at 3189-3191 and
 * 3210-3211): SE2W_SOURCE_FORCE gives vals[2k+c] = F_c * N_k(xi), and
 * SE2W_SOURCE_IMPL_POINTWISE evaluates the basis at the source's own local
 * coordinate instead of the snapped GLL node.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "host_internal.h"

/* Corner-node bounding box of the (single) sub-domain,
 * SpecFECtxGetLocalBoundingBox (915-937). */
static void local_bbox(const struct se2w_spec *s, double lo[2], double hi[2])
{
  lo[0] = s->xcol[0];
  lo[1] = s->yrow[0];
  hi[0] = s->xcol[s->mx * s->p];
  hi[1] = s->yrow[s->my * s->p];
}

int se2w_source_create(SE2WSpec s, int type, int impl, const double coor[2], const double *values, int use_internal,
                       SE2WSource *out)
{
  struct se2w_source *src;
  double gmin[2], gmax[2], lo[2], hi[2], dx, dy;
  int ii, jj, e, k;
  if (!s || !coor || !values || !out) return host_fail(SE2W_HOST_ERR_ARG, "null argument");
  *out = NULL;
  if (type != SE2W_SOURCE_FORCE && type != SE2W_SOURCE_MOMENT) return host_fail(SE2W_HOST_ERR_ARG, "unknown source type %d", type);
  if (impl == SE2W_SOURCE_IMPL_DEFAULT) impl = SE2W_SOURCE_IMPL_NEAREST_QPOINT;
  if (impl != SE2W_SOURCE_IMPL_NEAREST_QPOINT && impl != SE2W_SOURCE_IMPL_POINTWISE)
    return host_fail(SE2W_HOST_ERR_SUP, "Source implementation type %d not supported", impl);
  spec_bbox(s, gmin, gmax);
  dx = (gmax[0] - gmin[0]) / ((double)s->mx);
  dy = (gmax[1] - gmin[1]) / ((double)s->my);
  if (coor[0] < gmin[0]) return host_fail(SE2W_HOST_ERR_USER, "Source x-coordinate (%+1.4e) < min(domain).x (%+1.4e)", coor[0], gmin[0]);
  if (coor[1] < gmin[1]) return host_fail(SE2W_HOST_ERR_USER, "Source y-coordinate (%+1.4e) < min(domain).y (%+1.4e)", coor[1], gmin[1]);
  if (coor[0] > gmax[0]) return host_fail(SE2W_HOST_ERR_USER, "Source x-coordinate (%+1.4e) > max(domain).x (%+1.4e)", coor[0], gmax[0]);
  if (coor[1] > gmax[1]) return host_fail(SE2W_HOST_ERR_USER, "Source y-coordinate (%+1.4e) > max(domain).y (%+1.4e)", coor[1], gmax[1]);
  local_bbox(s, lo, hi);
  /* element containing the source, 3035-3048 */
  ii = (int)((coor[0] - lo[0]) / dx);
  jj = (int)((coor[1] - lo[1]) / dy);
  if (ii == s->mx) ii--;
  if (jj == s->my) jj--;
  if (ii < 0 || jj < 0 || ii >= s->mx || jj >= s->my) return host_fail(SE2W_HOST_ERR_SUP, "A source was defined but no rank claimed it");
  e = ii + jj * s->mx;
  src = (struct se2w_source *)calloc(1, sizeof(*src));
  if (!src) return host_fail(SE2W_HOST_ERR_MEM, "out of memory");
  src->spec = s;
  src->type = type;
  src->impl = impl;
  src->use_internal = use_internal;
  src->coor[0] = coor[0];
  src->coor[1] = coor[1];
  src->elem = e;
  src->closest_gll = -1;
  for (k = 0; k < (type == SE2W_SOURCE_MOMENT ? 4 : 2); k++) src->values[k] = values[k];
  /* local coordinate, 3097-3104 */
  src->xi[0] = 2.0 * (coor[0] - (lo[0] + ii * dx)) / dx - 1.0;
  src->xi[1] = 2.0 * (coor[1] - (lo[1] + jj * dy)) / dy - 1.0;
  if (fabs(src->xi[0]) < 1.0e-12) src->xi[0] = 0.0;
  if (fabs(src->xi[1]) < 1.0e-12) src->xi[1] = 0.0;
  if (impl == SE2W_SOURCE_IMPL_NEAREST_QPOINT) { /* 3120-3180 */
    const int lo_k = use_internal ? 1 : 0, hi_k = use_internal ? s->nb - 1 : s->nb;
    double best = 1.7976931348623157e308;
    int ni, nj;
    if (s->p == 1 && use_internal) {
      free(src);
      return host_fail(SE2W_HOST_ERR_USER,
                       "Cannot use basis degree = 1 and approximate the source location based on nearest cell interior GLL point. "
                       "You must provide the command line option -use_nearest_internal_gll false");
    }
    for (nj = lo_k; nj < hi_k; nj++) {
      for (ni = lo_k; ni < hi_k; ni++) {
        const int node = ni + nj * s->nb;
        const double x = spec_x(s, ii * s->p + ni), y = spec_y(s, jj * s->p + nj);
        const double d2 = (x - coor[0]) * (x - coor[0]) + (y - coor[1]) * (y - coor[1]);
        if (d2 < best) { best = d2; src->closest_gll = node; }
      }
    }
    if (src->closest_gll < 0) { free(src); return host_fail(SE2W_HOST_ERR_SUP, "Failed to identify a closest quadrature point"); }
  }
  *out = src;
  return 0;
}

/* SeismicSourceSetup_Pointwise + SeismicSourceSetup_Moment_Pointwise
 * (3310-3429): basis (derivatives) at the source point, scaled to physical
 * derivatives, contracted with the moment tensor. */
int se2w_source_setup(SE2WSource src)
{
  const struct se2w_spec *s;
  double xi[2], dphi_x[HOST_MAX_NB], N_x[HOST_MAX_NB], dphi_y[HOST_MAX_NB], N_y[HOST_MAX_NB];
  double dN_dxi[HOST_MAX_NB * HOST_MAX_NB], dN_deta[HOST_MAX_NB * HOST_MAX_NB], Nk[HOST_MAX_NB * HOST_MAX_NB];
  double dN_dx[HOST_MAX_NB * HOST_MAX_NB], dN_dy[HOST_MAX_NB * HOST_MAX_NB];
  int i, j, k, ierr;
  if (!src) return 0;
  if (src->issetup) return 0;
  s = src->spec;
  if (src->impl == SE2W_SOURCE_IMPL_NEAREST_QPOINT) {
    const int qj = src->closest_gll / s->nb, qi = src->closest_gll - qj * s->nb;
    xi[0] = s->xi[qi];
    xi[1] = s->xi[qj];
  } else {
    xi[0] = src->xi[0];
    xi[1] = src->xi[1];
  }
  /* TabulateBasisDerivativesAtPointTensorProduct2d (605-690) */
  if ((ierr = se2w_basis1d_deriv(s->p, 1, &xi[0], dphi_x)) || (ierr = se2w_basis1d(s->p, 1, &xi[0], N_x)) ||
      (ierr = se2w_basis1d_deriv(s->p, 1, &xi[1], dphi_y)) || (ierr = se2w_basis1d(s->p, 1, &xi[1], N_y)))
    return ierr;
  k = 0;
  for (j = 0; j < s->nb; j++) {
    for (i = 0; i < s->nb; i++) {
      dN_dxi[k] = dphi_x[i] * N_y[j];
      dN_deta[k] = N_x[i] * dphi_y[j];
      Nk[k] = N_x[i] * N_y[j];
      k++;
    }
  }
  spec_point_derivs(s, src->elem, dN_dxi, dN_deta, dN_dx, dN_dy);
  for (k = 0; k < s->npe; k++) {
    if (src->type == SE2W_SOURCE_MOMENT) {
      src->vals[2 * k] = (src->values[0] * dN_dx[k] + src->values[1] * dN_dy[k]);
      src->vals[2 * k + 1] = (src->values[2] * dN_dx[k] + src->values[3] * dN_dy[k]);
    } else {
      src->vals[2 * k] = src->values[0] * Nk[k];
      src->vals[2 * k + 1] = src->values[1] * Nk[k];
    }
  }
  src->issetup = 1;
  return 0;
}

int se2w_source_destroy(SE2WSource *sp)
{
  if (!sp || !*sp) return 0;
  free(*sp);
  *sp = NULL;
  return 0;
}

int se2w_source_get(SE2WSource src, int32_t *elem, int32_t *closest_gll, double *vals)
{
  if (!src) return host_fail(SE2W_HOST_ERR_ARG, "null source");
  if (!src->issetup) return host_fail(SE2W_HOST_ERR_USER, "Seismic source not setup. Must call SeismicSourceSetup()");
  if (elem) *elem = src->elem;
  if (closest_gll) *closest_gll = src->closest_gll;
  if (vals) memcpy(vals, src->vals, sizeof(double) * 2 * src->spec->npe);
  return 0;
}

/* SeismicSourceEvaluate (3470-3497): g = 0; g[dofs] += stf(t) * vals. */
int se2w_sources_evaluate(double time, int ns, SE2WSource *src, SE2WSTF *stf, double *g)
{
  int p, k, ierr;
  const struct se2w_spec *s = NULL;
  for (p = 0; p < ns; p++)
    if (src[p]) { s = src[p]->spec; break; }
  if (!s) return 0;
  memset(g, 0, sizeof(double) * 2 * (size_t)s->nnodes);
  for (p = 0; p < ns; p++) {
    double value = 1.0;
    if (!src[p]) continue;
    if (!src[p]->issetup) return host_fail(SE2W_HOST_ERR_USER, "Seismic source not setup. Must call SeismicSourceSetup()");
    if (stf) {
      if (!stf[p]) continue;
      if ((ierr = se2w_stf_evaluate(stf[p], time, &value))) return ierr;
    }
    for (k = 0; k < s->npe; k++) {
      const int64_t n = spec_node(s, src[p]->elem, k);
      const double b0 = value * src[p]->vals[2 * k], b1 = value * src[p]->vals[2 * k + 1];
      g[2 * n] += b0;
      g[2 * n + 1] += b1;
    }
  }
  return 0;
}
