/*
 * receiver.c -- receivers and their output files
 * (RecordUVA_MultipleStations_NearestGLL_MPI on one rank, se2wave.c:2687-2981).
 * Each receiver samples the GLL node nearest to it inside its element; the
 * rows hold u, v, a at that node and curl(v) there.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "host_internal.h"

int se2w_receivers_create(SE2WSpec s, int nr, const double *xr, SE2WReceivers *out)
{
  struct se2w_receivers *rc;
  double gmin[2], gmax[2], dx, dy;
  int r;
  if (!s || !out || nr < 0 || (nr > 0 && !xr)) return host_fail(SE2W_HOST_ERR_ARG, "bad receiver arguments");
  *out = NULL;
  spec_bbox(s, gmin, gmax);
  rc = (struct se2w_receivers *)calloc(1, sizeof(*rc));
  if (!rc) return host_fail(SE2W_HOST_ERR_MEM, "out of memory");
  rc->spec = s;
  rc->nr = nr;
  rc->xr = (double *)malloc(sizeof(double) * 2 * (nr > 0 ? nr : 1));
  rc->nid = (int32_t *)malloc(sizeof(int32_t) * (nr > 0 ? nr : 1));
  rc->elem = (int32_t *)malloc(sizeof(int32_t) * (nr > 0 ? nr : 1));
  rc->gll = (int32_t *)malloc(sizeof(int32_t) * (nr > 0 ? nr : 1));
  if (!rc->xr || !rc->nid || !rc->elem || !rc->gll) { se2w_receivers_destroy(&rc); return host_fail(SE2W_HOST_ERR_MEM, "out of memory"); }
  if (nr > 0) memcpy(rc->xr, xr, sizeof(double) * 2 * nr);
  dx = (gmax[0] - gmin[0]) / ((double)s->mx);
  dy = (gmax[1] - gmin[1]) / ((double)s->my);
  for (r = 0; r < nr; r++) {
    const double x = xr[2 * r], y = xr[2 * r + 1];
    double best = 1.0e32;
    int ei, ej, e, k;
    if (x < gmin[0]) { se2w_receivers_destroy(&rc); return host_fail(SE2W_HOST_ERR_USER, "Receiver %d, x-coordinate (%+1.4e) < min(domain).x (%+1.4e)", r, x, gmin[0]); }
    if (y < gmin[1]) { se2w_receivers_destroy(&rc); return host_fail(SE2W_HOST_ERR_USER, "Receiver %d, y-coordinate (%+1.4e) < min(domain).y (%+1.4e)", r, y, gmin[1]); }
    if (x > gmax[0]) { se2w_receivers_destroy(&rc); return host_fail(SE2W_HOST_ERR_USER, "Receiver %d, x-coordinate (%+1.4e) > max(domain).x (%+1.4e)", r, x, gmax[0]); }
    if (y > gmax[1]) { se2w_receivers_destroy(&rc); return host_fail(SE2W_HOST_ERR_USER, "Receiver %d, y-coordinate (%+1.4e) > max(domain).y (%+1.4e)", r, y, gmax[1]); }
    /* element, 2744-2751 (corner-node origin of the sub-domain) */
    ei = (int)((x - s->xcol[0]) / dx);
    if (ei == s->mx) ei--;
    ej = (int)((y - s->yrow[0]) / dy);
    if (ej == s->my) ej--;
    e = ei + ej * s->mx;
    rc->elem[r] = e;
    rc->nid[r] = -1;
    rc->gll[r] = -1;
    /* nearest node of the element, first one on ties, 2762-2772 */
    for (k = 0; k < s->npe; k++) {
      const int I = ei * s->p + k % s->nb, J = ej * s->p + k / s->nb;
      double d2 = (x - spec_x(s, I)) * (x - spec_x(s, I));
      d2 += (y - spec_y(s, J)) * (y - spec_y(s, J));
      if (d2 < best) {
        best = d2;
        rc->nid[r] = (int32_t)((int64_t)I + (int64_t)J * s->nx);
        rc->gll[r] = k;
      }
    }
  }
  *out = rc;
  return 0;
}

int se2w_receivers_destroy(SE2WReceivers *rp)
{
  if (!rp || !*rp) return 0;
  free((*rp)->xr); free((*rp)->nid); free((*rp)->elem); free((*rp)->gll);
  free(*rp);
  *rp = NULL;
  return 0;
}

int se2w_receivers_get(SE2WReceivers rc, int32_t *nid, int32_t *elem, int32_t *gll)
{
  if (!rc) return host_fail(SE2W_HOST_ERR_ARG, "null receivers");
  if (nid) memcpy(nid, rc->nid, sizeof(int32_t) * rc->nr);
  if (elem) memcpy(elem, rc->elem, sizeof(int32_t) * rc->nr);
  if (gll) memcpy(gll, rc->gll, sizeof(int32_t) * rc->nr);
  return 0;
}

/* Header (2806-2842) and .mpimeta file (2850-2877). */
int se2w_receivers_write_header(SE2WReceivers rc, const char *dir)
{
  const struct se2w_spec *s;
  char meta[1100];
  FILE *fp;
  int r;
  if (!rc) return host_fail(SE2W_HOST_ERR_ARG, "null receivers");
  s = rc->spec;
  if (!dir || !dir[0]) dir = ".";
  snprintf(rc->filename, sizeof(rc->filename), "%s/closestqpsource-receiverCP-uva-%dx%d-p%d-rank%d.dat", dir, s->mx, s->my, s->p, 0);
  fp = fopen(rc->filename, "w");
  if (!fp) return host_fail(SE2W_HOST_ERR_FILE, "Failed to open file \"%s\"", rc->filename);
  fprintf(fp, "# SpecFECtx meta data\n");
  fprintf(fp, "#   mx %d : my %d : basis order %d\n", s->mx, s->my, s->p);
  fprintf(fp, "# Receiver meta data\n");
  fprintf(fp, "#   + number receiver locations: %d\n", rc->nr);
  fprintf(fp, "#   + number receiver locations <local>: %d\n", rc->nr);
  fprintf(fp, "#   + takes displ/velo/accel from basis nearest to requested receiver location\n");
  for (r = 0; r < rc->nr; r++) {
    const int I = (int)(rc->nid[r] % s->nx), J = (int)(rc->nid[r] / s->nx);
    fprintf(fp, "#   + receiver location [%d]: x,y %+1.8e %+1.8e\n", r, rc->xr[2 * r], rc->xr[2 * r + 1]);
    fprintf(fp, "#   +   mapped to nearest node --> %+1.8e %+1.8e\n", spec_x(s, I), spec_y(s, J));
  }
  if (rc->nr != 0) {
    fprintf(fp, "# Time series header <field>(<column index>)\n");
    fprintf(fp, "#   time(1)\n");
    for (r = 0; r < rc->nr; r++) {
      const int off = 1 + r * 7;
      fprintf(fp, "#     ux(%d) uy(%d) vx(%d) vy(%d) ax(%d) ay(%d) curl(v) (%d)-> station [%d]\n", off + 1, off + 2, off + 3,
              off + 4, off + 5, off + 6, off + 7, r);
    }
  } else {
    fprintf(fp, "# <note> No receivers found on this sub-domain\n");
    fprintf(fp, "# <note> This file will remain empty\n");
  }
  fclose(fp);
  snprintf(meta, sizeof(meta), "%s/closestqpsource-receiverCP-uva-%dx%d-p%d.mpimeta", dir, s->mx, s->my, s->p);
  fp = fopen(meta, "w");
  if (!fp) return host_fail(SE2W_HOST_ERR_FILE, "Failed to open file \"%s\"", meta);
  fprintf(fp, "# SpecFECtx parallel/MPI meta data\n");
  fprintf(fp, "#   mx %d : my %d : basis order %d\n", s->mx, s->my, s->p);
  fprintf(fp, "# Receiver meta data\n");
  fprintf(fp, "#   + number receiver locations: %d\n", rc->nr);
  for (r = 0; r < rc->nr; r++) fprintf(fp, "#   + receiver [%d]: mapped to MPI rank %d\n", r, 0);
  fclose(fp);
  rc->header_written = 1;
  return 0;
}

/* Rows (2908-2962): "%1.4e" time, then " %+1.8e" x 7 per receiver. */
int se2w_receivers_append_rows(SE2WReceivers rc, int nsteps, const double *times, const double *trace)
{
  FILE *fp;
  int k, r, c;
  if (!rc || (nsteps > 0 && (!times || (rc->nr > 0 && !trace)))) return host_fail(SE2W_HOST_ERR_ARG, "bad arguments");
  if (!rc->header_written) return host_fail(SE2W_HOST_ERR_USER, "receiver header not written");
  if (rc->nr == 0) return 0;
  fp = fopen(rc->filename, "a");
  if (!fp) return host_fail(SE2W_HOST_ERR_FILE, "Failed to open file \"%s\"", rc->filename);
  for (k = 0; k < nsteps; k++) {
    fprintf(fp, "%1.4e", times[k]);
    for (r = 0; r < rc->nr; r++) {
      const double *o = &trace[((size_t)k * rc->nr + r) * 7];
      for (c = 0; c < 7; c++) fprintf(fp, " %+1.8e", o[c]);
    }
    fprintf(fp, "\n");
  }
  fclose(fp);
  return 0;
}
