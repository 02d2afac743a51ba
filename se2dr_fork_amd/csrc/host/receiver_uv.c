/*
 * receiver_uv.c -- single-station displacement/velocity receivers:
 * RecordUV (nearest node of the containing element, se2wave.c:2351-2443) and
 * RecordUV_interp (Lagrange interpolation of the element's nodal values at
 * the station, se2wave.c:2445-2581).  Both are a weighted sum over the
 * nodes of one element (a unit weight at the nearest node, or the tensor
 * basis values at the station), which the GPU side samples every step
 * (se2w_set_element_receivers).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "host_internal.h"

static const char *impl_prefix(int impl)
{
  switch (impl) {
    case -1: return "defaultsource";
    case 0: return "deltasource";
    case 1: return "closestqpsource";
    case 2: return "csplinesource";
    case 3: return "p0source";
    default: return NULL;
  }
}

int se2w_receiver_uv_create(SE2WSpec s, const double xr[2], int interpolate, int source_implementation,
                            SE2WReceiverUV *out)
{
  struct se2w_receiver_uv *r;
  double gmin[2], gmax[2], dx, dy;
  int ei, ej, k, ierr;
  if (!s || !xr || !out) return host_fail(SE2W_HOST_ERR_ARG, "bad receiver arguments");
  *out = NULL;
  if (!impl_prefix(source_implementation)) return host_fail(SE2W_HOST_ERR_ARG, "unknown source implementation %d", source_implementation);
  spec_bbox(s, gmin, gmax);
  /* containing element, 2392-2398 / 2497-2503 (truncation, no clamp) */
  dx = (gmax[0] - gmin[0]) / ((double)s->mx);
  dy = (gmax[1] - gmin[1]) / ((double)s->my);
  ei = (int)((xr[0] - gmin[0]) / dx);
  ej = (int)((xr[1] - gmin[1]) / dy);
  if (xr[0] < gmin[0] || xr[1] < gmin[1] || ei < 0 || ej < 0 || ei >= s->mx || ej >= s->my)
    return host_fail(SE2W_HOST_ERR_USER, "receiver (%+1.8e, %+1.8e) has no containing element", xr[0], xr[1]);
  r = (struct se2w_receiver_uv *)calloc(1, sizeof(*r));
  if (!r) return host_fail(SE2W_HOST_ERR_MEM, "out of memory");
  r->weights = (double *)calloc((size_t)s->npe, sizeof(double));
  if (!r->weights) { free(r); return host_fail(SE2W_HOST_ERR_MEM, "out of memory"); }
  r->spec = s;
  r->xr[0] = xr[0];
  r->xr[1] = xr[1];
  r->interpolate = interpolate != 0;
  r->source_implementation = source_implementation;
  r->elem = ei + ej * s->mx;
  r->nid = -1;
  if (!r->interpolate) {
    /* nearest node of the element, first one on ties (2409-2419) */
    double best = 1.0e32;
    int kbest = -1;
    for (k = 0; k < s->npe; k++) {
      const int I = ei * s->p + k % s->nb, J = ej * s->p + k / s->nb;
      double sep2 = (xr[0] - spec_x(s, I)) * (xr[0] - spec_x(s, I));
      sep2 += (xr[1] - spec_y(s, J)) * (xr[1] - spec_y(s, J));
      if (sep2 < best) {
        best = sep2;
        kbest = k;
      }
    }
    {
      const int I = ei * s->p + kbest % s->nb, J = ej * s->p + kbest / s->nb;
      r->nid = (int32_t)((int64_t)I + (int64_t)J * s->nx);
      r->weights[kbest] = 1.0;
      r->mapped[0] = spec_x(s, I);
      r->mapped[1] = spec_y(s, J);
    }
  } else {
    /* reference coordinates and tensor basis (2511-2532) */
    double xi, eta, x0, y0, Ns1[HOST_MAX_NB], Ns2[HOST_MAX_NB];
    int i, j;
    x0 = gmin[0] + ei * dx;
    y0 = gmin[1] + ej * dy;
    xi = 2.0 * (xr[0] - x0) / dx - 1.0;
    eta = 2.0 * (xr[1] - y0) / dy - 1.0;
    if ((ierr = se2w_basis1d(s->p, 1, &xi, Ns1)) || (ierr = se2w_basis1d(s->p, 1, &eta, Ns2))) {
      se2w_receiver_uv_destroy(&r);
      return ierr;
    }
    k = 0;
    for (j = 0; j < s->nb; j++)
      for (i = 0; i < s->nb; i++) r->weights[k++] = Ns1[i] * Ns2[j];
    /* interpolated coordinate (2534-2542) */
    r->mapped[0] = r->mapped[1] = 0.0;
    for (k = 0; k < s->npe; k++) {
      const int I = ei * s->p + k % s->nb, J = ej * s->p + k / s->nb;
      r->mapped[0] += r->weights[k] * spec_x(s, I);
      r->mapped[1] += r->weights[k] * spec_y(s, J);
    }
    printf("# receiver location: x,y %+1.8e %+1.8e -- interpolated coordinate --> %+1.8e %+1.8e\n", xr[0], xr[1],
           r->mapped[0], r->mapped[1]);
  }
  *out = r;
  return 0;
}

int se2w_receiver_uv_destroy(SE2WReceiverUV *rp)
{
  if (!rp || !*rp) return 0;
  free((*rp)->weights);
  free(*rp);
  *rp = NULL;
  return 0;
}

int se2w_receiver_uv_get(SE2WReceiverUV r, int32_t *elem, double *weights, double mapped[2])
{
  if (!r) return host_fail(SE2W_HOST_ERR_ARG, "null receiver");
  if (elem) *elem = r->elem;
  if (weights) memcpy(weights, r->weights, sizeof(double) * r->spec->npe);
  if (mapped) { mapped[0] = r->mapped[0]; mapped[1] = r->mapped[1]; }
  return 0;
}

/* File name and header (2364-2385, 2421-2432; 2458-2490). */
int se2w_receiver_uv_write_header(SE2WReceiverUV r, const char *dir)
{
  const struct se2w_spec *s;
  FILE *fp;
  if (!r) return host_fail(SE2W_HOST_ERR_ARG, "null receiver");
  s = r->spec;
  if (!dir || !dir[0]) dir = ".";
  snprintf(r->filename, sizeof(r->filename), "%s/%s-receiver%s-%dx%d-p%d.dat", dir, impl_prefix(r->source_implementation),
           r->interpolate ? "" : "CP", s->mx, s->my, s->p);
  fp = fopen(r->filename, "w");
  if (!fp) return host_fail(SE2W_HOST_ERR_FILE, "Failed to open file \"%s\"", r->filename);
  fprintf(fp, "# SpecFECtx meta data\n");
  fprintf(fp, "#   mx %d : my %d : basis order %d\n", s->mx, s->my, s->p);
  fprintf(fp, "#   source implementation %d\n", r->source_implementation);
  fprintf(fp, "# Receiver meta data\n");
  fprintf(fp, "#   + receiver location: x,y %+1.8e %+1.8e\n", r->xr[0], r->xr[1]);
  if (!r->interpolate) {
    fprintf(fp, "#   + takes displ/velo from basis nearest to requested receiver location\n");
    fprintf(fp, "#   + receiver location: x,y %+1.8e %+1.8e --mapped to nearest node --> %+1.8e %+1.8e\n", r->xr[0],
            r->xr[1], r->mapped[0], r->mapped[1]);
  } else {
    fprintf(fp, "#   + records displ/velo at requested receiver location through interpolating the FE solution\n");
  }
  fprintf(fp, "# Time series header\n");
  fprintf(fp, "#   time ux uy vx vy\n");
  fclose(fp);
  r->header_written = 1;
  return 0;
}

/* Rows "%1.4e %+1.8e %+1.8e %+1.8e %+1.8e" (2436, 2573). */
int se2w_receiver_uv_append_rows(SE2WReceiverUV r, int nsteps, const double *times, const double *rows)
{
  FILE *fp;
  int k;
  if (!r || (nsteps > 0 && (!times || !rows))) return host_fail(SE2W_HOST_ERR_ARG, "bad arguments");
  if (!r->header_written) return host_fail(SE2W_HOST_ERR_USER, "receiver header not written");
  fp = fopen(r->filename, "a");
  if (!fp) return host_fail(SE2W_HOST_ERR_FILE, "Failed to open file \"%s\"", r->filename);
  for (k = 0; k < nsteps; k++) {
    const double *o = &rows[(size_t)k * 4];
    fprintf(fp, "%1.4e %+1.8e %+1.8e %+1.8e %+1.8e\n", times[k], o[0], o[1], o[2], o[3]);
  }
  fclose(fp);
  return 0;
}
