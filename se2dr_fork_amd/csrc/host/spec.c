/*
 * spec.c -- mesh, materials, lumped mass and CFL time step of se2wave
 * (the SpecFECtx layer, se2wave.c:692-884, 1148-1279, 1431-1498, 2268-2349).
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "host_internal.h"

/* per thread: the driver writes snapshots from a second thread */
static __thread char host_err[512];

int host_fail(int code, const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(host_err, sizeof(host_err), fmt, ap);
  va_end(ap);
  return code;
}

const char *se2w_host_last_error(void) { return host_err; }

/* SpecFECtxCreateMesh_SEQ + SpecFECtxCreateMeshCoords2d_SEQ (825-884,
 * 739-779): element (ei,ej) writes x = 0.5*(xi_i+1)*dx + ei*dx to its nodes
 * in the loop order ej, ei, j, i, so a node shared by two elements keeps the
 * value of the later (larger-index) element. */
int se2w_spec_create(int mx, int my, int p, SE2WSpec *out)
{
  struct se2w_spec *s;
  int ierr, I, J, i, j;
  double h;
  if (!out) return host_fail(SE2W_HOST_ERR_ARG, "null output");
  *out = NULL;
  if (mx < 1 || my < 1) return host_fail(SE2W_HOST_ERR_ARG, "mx, my must be >= 1");
  if (p < 1 || p + 1 > HOST_MAX_NB) return host_fail(SE2W_HOST_ERR_SUP, "basis degree %d outside 1..%d", p, HOST_MAX_NB - 1);
  s = (struct se2w_spec *)calloc(1, sizeof(*s));
  if (!s) return host_fail(SE2W_HOST_ERR_MEM, "out of memory");
  s->mx = mx; s->my = my; s->p = p; s->nb = p + 1; s->npe = s->nb * s->nb;
  s->nx = p * mx + 1; s->ny = p * my + 1;
  s->nnodes = (int64_t)s->nx * s->ny;
  if ((ierr = se2w_gll(p, s->xi, s->w1d)) || (ierr = se2w_deriv_matrix(p, s->D))) { free(s); return ierr; }
  for (j = 0; j < s->nb; j++)
    for (i = 0; i < s->nb; i++) s->w2[i + j * s->nb] = s->w1d[i] * s->w1d[j];
  s->xcol = (double *)malloc(sizeof(double) * s->nx);
  s->yrow = (double *)malloc(sizeof(double) * s->ny);
  if (!s->xcol || !s->yrow) { se2w_spec_destroy(&s); return host_fail(SE2W_HOST_ERR_MEM, "out of memory"); }
  h = 1.0 / ((double)mx);
  for (I = 0; I < s->nx; I++) {
    const int ei = (I / p < mx) ? I / p : mx - 1; /* last element writing column I */
    const int il = I - ei * p;
    s->xcol[I] = 0.5 * (s->xi[il] + 1.0) * h + (0.0 + ei * h);
  }
  h = 1.0 / ((double)my);
  for (J = 0; J < s->ny; J++) {
    const int ej = (J / p < my) ? J / p : my - 1;
    const int jl = J - ej * p;
    s->yrow[J] = 0.5 * (s->xi[jl] + 1.0) * h + (0.0 + ej * h);
  }
  *out = s;
  return 0;
}

int se2w_spec_destroy(SE2WSpec *sp)
{
  if (!sp || !*sp) return 0;
  free((*sp)->xcol); free((*sp)->yrow);
  free((*sp)->lambda); free((*sp)->mu); free((*sp)->rho);
  free(*sp);
  *sp = NULL;
  return 0;
}

/* SpecFECtxScaleMeshCoords (781-823): VecStrideScale per component, then
 * one VecAXPY(coor, 1.0, ss) per component with ss holding the shift in
 * that component and 0.0 in the other. */
int se2w_spec_scale_coords(SE2WSpec s, const double scale[2], const double shift[2])
{
  int I, J;
  if (!s) return host_fail(SE2W_HOST_ERR_ARG, "null spec");
  if (scale) {
    for (I = 0; I < s->nx; I++) s->xcol[I] *= scale[0];
    for (J = 0; J < s->ny; J++) s->yrow[J] *= scale[1];
  }
  if (shift) {
    for (I = 0; I < s->nx; I++) s->xcol[I] = s->xcol[I] + 1.0 * shift[0];
    for (J = 0; J < s->ny; J++) s->yrow[J] = s->yrow[J] + 1.0 * 0.0;
    for (I = 0; I < s->nx; I++) s->xcol[I] = s->xcol[I] + 1.0 * 0.0;
    for (J = 0; J < s->ny; J++) s->yrow[J] = s->yrow[J] + 1.0 * shift[1];
  }
  return 0;
}

static int ensure_material(struct se2w_spec *s)
{
  const size_t ne = (size_t)s->mx * s->my;
  if (!s->lambda) s->lambda = (double *)malloc(sizeof(double) * ne);
  if (!s->mu) s->mu = (double *)malloc(sizeof(double) * ne);
  if (!s->rho) s->rho = (double *)malloc(sizeof(double) * ne);
  if (!s->lambda || !s->mu || !s->rho) return host_fail(SE2W_HOST_ERR_MEM, "out of memory");
  return 0;
}

int se2w_spec_set_constant_material(SE2WSpec s, double lambda, double mu, double rho)
{
  size_t e, ne;
  int ierr;
  if (!s) return host_fail(SE2W_HOST_ERR_ARG, "null spec");
  if ((ierr = ensure_material(s))) return ierr;
  ne = (size_t)s->mx * s->my;
  for (e = 0; e < ne; e++) { s->lambda[e] = lambda; s->mu[e] = mu; s->rho[e] = rho; }
  s->have_material = 1;
  return 0;
}

/* mu = Vs^2 rho, lambda = Vp^2 rho - 2 mu (1161-1175). */
int se2w_spec_set_constant_material_velocity(SE2WSpec s, double vp, double vs, double rho)
{
  const double mu = vs * vs * rho;
  const double lambda = vp * vp * rho - 2.0 * mu;
  return se2w_spec_set_constant_material(s, lambda, mu, rho);
}

int se2w_spec_set_cell_material(SE2WSpec s, const double *lambda, const double *mu, const double *rho)
{
  size_t ne;
  int ierr;
  if (!s || !lambda || !mu || !rho) return host_fail(SE2W_HOST_ERR_ARG, "null argument");
  if ((ierr = ensure_material(s))) return ierr;
  ne = (size_t)s->mx * s->my;
  memcpy(s->lambda, lambda, sizeof(double) * ne);
  memcpy(s->mu, mu, sizeof(double) * ne);
  memcpy(s->rho, rho, sizeof(double) * ne);
  s->have_material = 1;
  return 0;
}

int se2w_spec_set_cell_material_velocity(SE2WSpec s, const double *vp, const double *vs, const double *rho)
{
  size_t e, ne;
  int ierr;
  if (!s || !vp || !vs || !rho) return host_fail(SE2W_HOST_ERR_ARG, "null argument");
  if ((ierr = ensure_material(s))) return ierr;
  ne = (size_t)s->mx * s->my;
  for (e = 0; e < ne; e++) {
    const double mu = vs[e] * vs[e] * rho[e];
    s->mu[e] = mu;
    s->lambda[e] = vp[e] * vp[e] * rho[e] - 2.0 * mu;
    s->rho[e] = rho[e];
  }
  s->have_material = 1;
  return 0;
}

/* DMGetBoundingBox over the node coordinates. */
void spec_bbox(const struct se2w_spec *s, double gmin[2], double gmax[2])
{
  int I, J;
  gmin[0] = gmax[0] = s->xcol[0];
  gmin[1] = gmax[1] = s->yrow[0];
  for (I = 1; I < s->nx; I++) {
    if (s->xcol[I] < gmin[0]) gmin[0] = s->xcol[I];
    if (s->xcol[I] > gmax[0]) gmax[0] = s->xcol[I];
  }
  for (J = 1; J < s->ny; J++) {
    if (s->yrow[J] < gmin[1]) gmin[1] = s->yrow[J];
    if (s->yrow[J] > gmax[1]) gmax[1] = s->yrow[J];
  }
}

/* ElementEvaluateGeometry_CellWiseConstant2d (1238-1252): dx = x of local
 * node (p,0) - x of local node (0,0); dy = y of (p,p) - y of (0,0). */
double spec_elem_dx(const struct se2w_spec *s, int ei) { return s->xcol[ei * s->p + s->p] - s->xcol[ei * s->p]; }
double spec_elem_dy(const struct se2w_spec *s, int ej) { return s->yrow[ej * s->p + s->p] - s->yrow[ej * s->p]; }

void spec_point_derivs(const struct se2w_spec *s, int e, const double *dN_dxi, const double *dN_deta, double *dN_dx,
                       double *dN_dy)
{
  const double J00 = 0.5 * spec_elem_dx(s, e % s->mx), J11 = 0.5 * spec_elem_dy(s, e / s->mx);
  const double iJ00 = 1.0 / J00, iJ11 = 1.0 / J11;
  int k;
  for (k = 0; k < s->npe; k++) {
    dN_dx[k] = iJ00 * dN_dxi[k];
    dN_dy[k] = iJ11 * dN_deta[k];
  }
}

int se2w_spec_get_info(SE2WSpec s, se2w_spec_info *info)
{
  if (!s || !info) return host_fail(SE2W_HOST_ERR_ARG, "null argument");
  info->mx = s->mx; info->my = s->my; info->p = s->p; info->nx = s->nx; info->ny = s->ny; info->npe = s->npe;
  info->nnodes = s->nnodes;
  spec_bbox(s, info->gmin, info->gmax);
  return 0;
}

int se2w_spec_get_coords(SE2WSpec s, double *coords)
{
  int I, J;
  if (!s || !coords) return host_fail(SE2W_HOST_ERR_ARG, "null argument");
  for (J = 0; J < s->ny; J++)
    for (I = 0; I < s->nx; I++) {
      const int64_t n = (int64_t)I + (int64_t)J * s->nx;
      coords[2 * n] = s->xcol[I];
      coords[2 * n + 1] = s->yrow[J];
    }
  return 0;
}

int se2w_spec_get_material(SE2WSpec s, double *lambda, double *mu, double *rho)
{
  const size_t ne = (size_t)s->mx * s->my;
  if (!s || !s->have_material) return host_fail(SE2W_HOST_ERR_USER, "material not set");
  if (lambda) memcpy(lambda, s->lambda, sizeof(double) * ne);
  if (mu) memcpy(mu, s->mu, sizeof(double) * ne);
  if (rho) memcpy(rho, s->rho, sizeof(double) * ne);
  return 0;
}

int se2w_spec_get_tables(SE2WSpec s, double *D, double *w1d, double *elem_dx, double *elem_dy)
{
  int e;
  if (!s) return host_fail(SE2W_HOST_ERR_ARG, "null spec");
  if (D) memcpy(D, s->D, sizeof(double) * s->nb * s->nb);
  if (w1d) memcpy(w1d, s->w1d, sizeof(double) * s->nb);
  if (elem_dx)
    for (e = 0; e < s->mx; e++) elem_dx[e] = spec_elem_dx(s, e);
  if (elem_dy)
    for (e = 0; e < s->my; e++) elem_dy[e] = spec_elem_dy(s, e);
  return 0;
}

int se2w_spec_get_element(SE2WSpec s, int32_t *element)
{
  int e, k;
  if (!s || !element) return host_fail(SE2W_HOST_ERR_ARG, "null argument");
  for (e = 0; e < s->mx * s->my; e++)
    for (k = 0; k < s->npe; k++) element[(size_t)e * s->npe + k] = (int32_t)spec_node(s, e, k);
  return 0;
}

/* AssembleBilinearForm_Mass2d (1431-1498): M[node] += (detJ*w_q)*rho_e over
 * elements in ascending order. */
int se2w_spec_mass(SE2WSpec s, double *mass)
{
  int e, q;
  if (!s || !mass) return host_fail(SE2W_HOST_ERR_ARG, "null argument");
  if (!s->have_material) return host_fail(SE2W_HOST_ERR_USER, "material not set");
  memset(mass, 0, sizeof(double) * (size_t)s->nnodes);
  for (e = 0; e < s->mx * s->my; e++) {
    const double J00 = 0.5 * spec_elem_dx(s, e % s->mx), J11 = 0.5 * spec_elem_dy(s, e / s->mx);
    const double detJ = J00 * J11;
    for (q = 0; q < s->npe; q++) {
      const double fac = detJ * s->w2[q];
      mass[spec_node(s, e, q)] += fac * (s->rho[e]);
    }
  }
  return 0;
}

/* ElastoDynamicsComputeTimeStep_2d (2276-2349): dt = min_e h*(min dxi)/2 / Vp_e
 * with h = min(dx, dy) from the bounding box. */
int se2w_spec_timestep(SE2WSpec s, double *dt)
{
  double gmin[2], gmax[2], dx, dy, rmin, sepmin = 1.0e32, best = 1.7976931348623157e308;
  int n, e;
  if (!s || !dt) return host_fail(SE2W_HOST_ERR_ARG, "null argument");
  if (!s->have_material) return host_fail(SE2W_HOST_ERR_USER, "material not set");
  spec_bbox(s, gmin, gmax);
  dx = (gmax[0] - gmin[0]) / ((double)s->mx);
  dy = (gmax[1] - gmin[1]) / ((double)s->my);
  rmin = (dx < dy) ? dx : dy;
  for (n = 0; n < s->nb - 1; n++) {
    const double sep = fabs(s->xi[n + 1] - s->xi[n]);
    if (sep < sepmin) sepmin = sep;
  }
  rmin = rmin * (sepmin / 2.0);
  for (e = 0; e < s->mx * s->my; e++) {
    const double vp = sqrt((s->lambda[e] + 2.0 * s->mu[e]) / s->rho[e]);
    const double value = 1.0 * 1.0 * rmin / vp;
    if (value < best) best = value;
  }
  *dt = best;
  return 0;
}
