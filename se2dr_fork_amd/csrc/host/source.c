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
 *
 * The reference's two other moment-source forms, used by its legacy drivers
 * through ElastoDynamicsSourceSetup (2232-2264), are SeismicSource
 * implementations here so that the time loop injects them like any other
 * source: SE2W_SOURCE_IMPL_P0 (2059-2198) on the cell containing the source,
 * and SE2W_SOURCE_IMPL_SPLINE (1925-2057), an assembled vector over every
 * element the kernel reaches, held as one value set per element (each
 * assembled node value on one element) so that injecting the elements in
 * order adds exactly the assembled vector.
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
  if (impl != SE2W_SOURCE_IMPL_NEAREST_QPOINT && impl != SE2W_SOURCE_IMPL_POINTWISE && impl != SE2W_SOURCE_IMPL_SPLINE &&
      impl != SE2W_SOURCE_IMPL_P0)
    return host_fail(SE2W_HOST_ERR_SUP, "Source implementation type %d not supported", impl);
  if ((impl == SE2W_SOURCE_IMPL_SPLINE || impl == SE2W_SOURCE_IMPL_P0) && type != SE2W_SOURCE_MOMENT)
    return host_fail(SE2W_HOST_ERR_SUP, "the cubic-spline and P0 source implementations assemble moment sources");
  spec_bbox(s, gmin, gmax);
  dx = (gmax[0] - gmin[0]) / ((double)s->mx);
  dy = (gmax[1] - gmin[1]) / ((double)s->my);
  if (coor[0] < gmin[0]) return host_fail(SE2W_HOST_ERR_USER, "Source x-coordinate (%+1.4e) < min(domain).x (%+1.4e)", coor[0], gmin[0]);
  if (coor[1] < gmin[1]) return host_fail(SE2W_HOST_ERR_USER, "Source y-coordinate (%+1.4e) < min(domain).y (%+1.4e)", coor[1], gmin[1]);
  if (coor[0] > gmax[0]) return host_fail(SE2W_HOST_ERR_USER, "Source x-coordinate (%+1.4e) > max(domain).x (%+1.4e)", coor[0], gmax[0]);
  if (coor[1] > gmax[1]) return host_fail(SE2W_HOST_ERR_USER, "Source y-coordinate (%+1.4e) > max(domain).y (%+1.4e)", coor[1], gmax[1]);
  local_bbox(s, lo, hi);
  /* element containing the source, 3035-3048; the P0 form locates its cell
     from the bounding box (2089-2096) */
  if (impl == SE2W_SOURCE_IMPL_P0) {
    lo[0] = gmin[0];
    lo[1] = gmin[1];
  }
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
  src->elem = impl == SE2W_SOURCE_IMPL_SPLINE ? -1 : e; /* the spline form spans many elements */
  src->closest_gll = -1;
  src->nsub = impl == SE2W_SOURCE_IMPL_SPLINE ? 0 : 1;
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

/* AssembleLinearForm_ElastoDynamicsMomentDirac2d_P0 (2059-2198): the
 * physical basis derivatives at the cell centre xi = (0, 0) (2133-2148) and
 * fe_i = sum_q (w_q detJ)(M . grad N_i) / (dx dy) over the cell's GLL
 * weights (2160-2166), dx and dy the spacings of the bounding box (2089-2091). */
static int setup_p0(struct se2w_source *src)
{
  const struct se2w_spec *s = src->spec;
  const double *M = src->values, xi0 = 0.0;
  double gmin[2], gmax[2], dx, dy, detJ, dphi[HOST_MAX_NB], N[HOST_MAX_NB];
  double dN_dxi[HOST_MAX_NB * HOST_MAX_NB], dN_deta[HOST_MAX_NB * HOST_MAX_NB];
  double dN_dx[HOST_MAX_NB * HOST_MAX_NB], dN_dy[HOST_MAX_NB * HOST_MAX_NB];
  int i, j, k, q, ierr;
  spec_bbox(s, gmin, gmax);
  dx = (gmax[0] - gmin[0]) / ((double)s->mx);
  dy = (gmax[1] - gmin[1]) / ((double)s->my);
  if ((ierr = se2w_basis1d_deriv(s->p, 1, &xi0, dphi)) || (ierr = se2w_basis1d(s->p, 1, &xi0, N))) return ierr;
  k = 0;
  for (j = 0; j < s->nb; j++)
    for (i = 0; i < s->nb; i++, k++) {
      dN_dxi[k] = dphi[i] * N[j];
      dN_deta[k] = N[i] * dphi[j];
    }
  spec_point_derivs(s, src->elem, dN_dxi, dN_deta, dN_dx, dN_dy);
  detJ = (0.5 * spec_elem_dx(s, src->elem % s->mx)) * (0.5 * spec_elem_dy(s, src->elem / s->mx));
  for (k = 0; k < 2 * s->npe; k++) src->vals[k] = 0.0;
  for (q = 0; q < s->npe; q++)
    for (i = 0; i < s->npe; i++) {
      src->vals[2 * i] += s->w2[q] * detJ * (M[0] * dN_dx[i] + M[1] * dN_dy[i]) / (dx * dy);
      src->vals[2 * i + 1] += s->w2[q] * detJ * (M[2] * dN_dx[i] + M[3] * dN_dy[i]) / (dx * dy);
    }
  return 0;
}

/* The smoothing kernel of 2011-2022: the cubic spline of support 2h,
 * normalised by 10/(7 pi h^2) so that its integral over the plane is 1. */
static double spline_W(double sep, double h)
{
  const double arg = sqrt(sep) / h;
  double w;
  if (arg < 1.0) w = 0.25 * (2.0 - arg) * (2.0 - arg) * (2.0 - arg) - (1.0 - arg) * (1.0 - arg) * (1.0 - arg);
  else if (arg < 2.0) w = 0.25 * (2.0 - arg) * (2.0 - arg) * (2.0 - arg);
  else w = 0.0;
  return 10.0 / (7.0 * M_PI) * w * (1.0 / (h * h));
}

/* Element columns (or rows) [*a, *b] whose node range meets [c - r, c + r]. */
static void band(const double *coord, int nel, int p, double c, double r, int *a, int *b)
{
  int e;
  *a = nel;
  *b = -1;
  for (e = 0; e < nel; e++) {
    const double x0 = coord[e * p], x1 = coord[e * p + p];
    const double lo = x0 < x1 ? x0 : x1, hi = x0 < x1 ? x1 : x0;
    if (hi >= c - r && lo <= c + r) {
      if (e < *a) *a = e;
      *b = e;
    }
  }
}

/* AssembleLinearForm_ElastoDynamicsMomentDirac2d_Kernel_CSpline (1925-2057)
 * for one source.  The kernel width comes from the bounding box (1939-1973)
 * unless set (-sm_h, 1975); every element contributes
 * fe_i = sum_q (w_q detJ)(M . grad N_i(x_q)) W(|x_q - x_s|) over its GLL
 * points, added into the assembled vector in ascending element order
 * (VecSetValues ADD_VALUES, 2046), and int_w = sum (w_q detJ) W (2044, 2053).
 * Elements whose nodes all lie beyond the support 2h add exact zeros, so only
 * the block of elements reaching within 2h (with a 1 % margin) is visited,
 * in the same order. */
static int setup_spline(struct se2w_source *src)
{
  const struct se2w_spec *s = src->spec;
  const int nb = s->nb, npe = s->npe, p = s->p;
  const double *M = src->values, *xs = src->coor;
  double gmin[2], gmax[2], dx, dy, ds, h, fe[2 * HOST_MAX_NB * HOST_MAX_NB];
  double *dN_dxi = NULL, *dN_deta = NULL, *F = NULL;
  int *owner = NULL, i0, i1, j0, j1, LX, LY, ei, ej, q, i, k, ierr = 0, nsub = 0;
  spec_bbox(s, gmin, gmax);
  dx = (gmax[0] - gmin[0]) / ((double)s->mx);
  dy = (gmax[1] - gmin[1]) / ((double)s->my);
  ds = dx;
  if (dy > dx) ds = dy;
  if (p <= 3) h = 1.5 * (0.5 * ds);
  else {
    const double pfac = 3.0 / (p - 2.0);
    h = pfac * (0.5 * ds);
  }
  if (src->h_opt > 0.0) h = src->h_opt;
  src->kernel_h = h;
  src->int_w = 0.0;
  free(src->sub_elem); /* a setup that failed part-way may have left them */
  free(src->sub_vals);
  src->sub_elem = NULL;
  src->sub_vals = NULL;
  band(s->xcol, s->mx, p, xs[0], 2.02 * h, &i0, &i1);
  band(s->yrow, s->my, p, xs[1], 2.02 * h, &j0, &j1);
  src->nsub = 0;
  if (i0 > i1 || j0 > j1) return 0;
  LX = (i1 - i0 + 1) * p + 1;
  LY = (j1 - j0 + 1) * p + 1;
  /* TabulateBasisDerivativesTensorProduct2d (525-603) at the GLL points:
     dN_dxi[q][k] = D[qi][i] * delta(qj, j), dN_deta[q][k] = delta(qi, i) * D[qj][j] */
  dN_dxi = (double *)malloc(sizeof(double) * npe * npe);
  dN_deta = (double *)malloc(sizeof(double) * npe * npe);
  F = (double *)calloc((size_t)2 * LX * LY, sizeof(double));
  owner = (int *)malloc(sizeof(int) * (size_t)LX * LY);
  if (!dN_dxi || !dN_deta || !F || !owner) { ierr = host_fail(SE2W_HOST_ERR_MEM, "out of memory"); goto done; }
  for (q = 0; q < npe; q++)
    for (k = 0; k < npe; k++) {
      const int qi = q % nb, qj = q / nb, ki = k % nb, kj = k / nb;
      dN_dxi[q * npe + k] = s->D[qi * nb + ki] * (qj == kj ? 1.0 : 0.0);
      dN_deta[q * npe + k] = (qi == ki ? 1.0 : 0.0) * s->D[qj * nb + kj];
    }
  for (ej = j0; ej <= j1; ej++)
    for (ei = i0; ei <= i1; ei++) {
      const double J00 = 0.5 * spec_elem_dx(s, ei), J11 = 0.5 * spec_elem_dy(s, ej);
      const double iJ00 = 1.0 / J00, iJ11 = 1.0 / J11, detJ = J00 * J11;
      for (i = 0; i < 2 * npe; i++) fe[i] = 0.0;
      for (q = 0; q < npe; q++) {
        const double xq = spec_x(s, ei * p + q % nb), yq = spec_y(s, ej * p + q / nb);
        double sep = (xq - xs[0]) * (xq - xs[0]), W;
        sep += (yq - xs[1]) * (yq - xs[1]);
        W = spline_W(sep, h);
        for (i = 0; i < npe; i++) {
          const double dNx = iJ00 * dN_dxi[q * npe + i], dNy = iJ11 * dN_deta[q * npe + i];
          fe[2 * i] += s->w2[q] * detJ * (M[0] * dNx + M[1] * dNy) * W;
          fe[2 * i + 1] += s->w2[q] * detJ * (M[2] * dNx + M[3] * dNy) * W;
        }
        src->int_w += s->w2[q] * detJ * W;
      }
      for (k = 0; k < npe; k++) {
        const size_t n = (size_t)((ei - i0) * p + k % nb) + (size_t)((ej - j0) * p + k / nb) * LX;
        F[2 * n] += fe[2 * k];
        F[2 * n + 1] += fe[2 * k + 1];
      }
    }
  /* each node with a nonzero value goes to the first element in ascending
     order that holds it */
  for (k = 0; k < LX * LY; k++) owner[k] = -1;
  src->sub_elem = (int32_t *)malloc(sizeof(int32_t) * (size_t)(i1 - i0 + 1) * (j1 - j0 + 1));
  src->sub_vals = (double *)malloc(sizeof(double) * (size_t)(i1 - i0 + 1) * (j1 - j0 + 1) * 2 * npe);
  if (!src->sub_elem || !src->sub_vals) { ierr = host_fail(SE2W_HOST_ERR_MEM, "out of memory"); goto done; }
  for (ej = j0; ej <= j1; ej++)
    for (ei = i0; ei <= i1; ei++) {
      double *v = &src->sub_vals[(size_t)nsub * 2 * npe];
      int owns = 0;
      for (k = 0; k < npe; k++) {
        const size_t n = (size_t)((ei - i0) * p + k % nb) + (size_t)((ej - j0) * p + k / nb) * LX;
        v[2 * k] = v[2 * k + 1] = 0.0;
        if (owner[n] < 0 && (F[2 * n] != 0.0 || F[2 * n + 1] != 0.0)) {
          owner[n] = ei + ej * s->mx;
          v[2 * k] = F[2 * n];
          v[2 * k + 1] = F[2 * n + 1];
          owns = 1;
        }
      }
      if (owns) src->sub_elem[nsub++] = ei + ej * s->mx;
    }
  src->nsub = nsub;
done:
  free(dN_dxi); free(dN_deta); free(F); free(owner);
  return ierr;
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
  if (src->impl == SE2W_SOURCE_IMPL_SPLINE || src->impl == SE2W_SOURCE_IMPL_P0) {
    if ((ierr = src->impl == SE2W_SOURCE_IMPL_SPLINE ? setup_spline(src) : setup_p0(src))) return ierr;
    src->issetup = 1;
    return 0;
  }
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
  free((*sp)->sub_elem);
  free((*sp)->sub_vals);
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
  if (vals && src->impl == SE2W_SOURCE_IMPL_SPLINE)
    return host_fail(SE2W_HOST_ERR_ARG, "a cubic-spline source spans %d elements: se2w_source_get_elements gives its values",
                     src->nsub);
  if (vals) memcpy(vals, src->vals, sizeof(double) * 2 * src->spec->npe);
  return 0;
}

int se2w_source_set_kernel_width(SE2WSource src, double h)
{
  if (!src) return host_fail(SE2W_HOST_ERR_ARG, "null source");
  if (src->impl != SE2W_SOURCE_IMPL_SPLINE) return host_fail(SE2W_HOST_ERR_ARG, "only a cubic-spline source has a kernel width");
  if (src->issetup) return host_fail(SE2W_HOST_ERR_USER, "set the kernel width before SeismicSourceSetup()");
  if (!(h == h) || h > 1.7976931348623157e308) return host_fail(SE2W_HOST_ERR_ARG, "kernel width must be finite");
  src->h_opt = h;
  return 0;
}

int se2w_source_get_kernel(SE2WSource src, double *h, double *int_w)
{
  if (!src) return host_fail(SE2W_HOST_ERR_ARG, "null source");
  if (src->impl != SE2W_SOURCE_IMPL_SPLINE) return host_fail(SE2W_HOST_ERR_ARG, "only a cubic-spline source has a kernel width");
  if (!src->issetup) return host_fail(SE2W_HOST_ERR_USER, "Seismic source not setup. Must call SeismicSourceSetup()");
  if (h) *h = src->kernel_h;
  if (int_w) *int_w = src->int_w;
  return 0;
}

int se2w_source_get_elements(SE2WSource src, int32_t *ne, int32_t *elem, double *vals)
{
  if (!src || !ne) return host_fail(SE2W_HOST_ERR_ARG, "null argument");
  if (!src->issetup) return host_fail(SE2W_HOST_ERR_USER, "Seismic source not setup. Must call SeismicSourceSetup()");
  *ne = src->nsub;
  if (src->impl == SE2W_SOURCE_IMPL_SPLINE) {
    if (elem && src->nsub) memcpy(elem, src->sub_elem, sizeof(int32_t) * src->nsub);
    if (vals && src->nsub) memcpy(vals, src->sub_vals, sizeof(double) * (size_t)src->nsub * 2 * src->spec->npe);
  } else {
    if (elem) elem[0] = src->elem;
    if (vals) memcpy(vals, src->vals, sizeof(double) * 2 * src->spec->npe);
  }
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
    if (src[p]->impl == SE2W_SOURCE_IMPL_SPLINE) { /* the assembled vector, one element's share at a time */
      int j;
      for (j = 0; j < src[p]->nsub; j++)
        for (k = 0; k < s->npe; k++) {
          const double *v = &src[p]->sub_vals[((size_t)j * s->npe + k) * 2];
          const int64_t n = spec_node(s, src[p]->sub_elem[j], k);
          if (v[0] == 0.0 && v[1] == 0.0) continue;
          g[2 * n] += value * v[0];
          g[2 * n + 1] += value * v[1];
        }
      continue;
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
