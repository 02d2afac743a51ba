/* host_internal.h -- private declarations of the se2wave host library. */
#ifndef SE2W_HOST_INTERNAL_H
#define SE2W_HOST_INTERNAL_H

#include <stddef.h>
#include <stdint.h>

#include "se2wave.h"

#define HOST_MAX_NB 11 /* basis degree <= 10 on the host; the GPU kernels take <= 8 */

int host_fail(int code, const char *fmt, ...);

/* Mesh and material context (the reference's SpecFECtx, se2wave.c:35-54).
 * The mesh is a tensor product: node coordinates depend only on the node
 * column (x) or row (y), so they are stored as two 1D arrays. */
struct se2w_spec {
  int mx, my, p, nb, npe, nx, ny;
  int64_t nnodes;
  double xi[HOST_MAX_NB], w1d[HOST_MAX_NB], w2[HOST_MAX_NB * HOST_MAX_NB];
  double D[HOST_MAX_NB * HOST_MAX_NB];
  double *xcol; /* [nx] x of node column I */
  double *yrow; /* [ny] y of node row J */
  double *lambda, *mu, *rho; /* [mx*my] */
  int have_material;
};

/* x, y of node (I, J) */
static inline double spec_x(const struct se2w_spec *s, int I) { return s->xcol[I]; }
static inline double spec_y(const struct se2w_spec *s, int J) { return s->yrow[J]; }
/* global node index of local node k of element e */
static inline int64_t spec_node(const struct se2w_spec *s, int e, int k)
{
  const int ei = e % s->mx, ej = e / s->mx, i = k % s->nb, j = k / s->nb;
  return (int64_t)(ei * s->p + i) + (int64_t)(ej * s->p + j) * s->nx;
}
void spec_bbox(const struct se2w_spec *s, double gmin[2], double gmax[2]);
/* element width along x for element column ei / along y for element row ej */
double spec_elem_dx(const struct se2w_spec *s, int ei);
double spec_elem_dy(const struct se2w_spec *s, int ej);
/* ElementEvaluateDerivatives_CellWiseConstant2d for one point:
 * dN_dx[k] = (1/J00) dN_dxi[k], dN_dy[k] = (1/J11) dN_deta[k]. */
void spec_point_derivs(const struct se2w_spec *s, int e, const double *dN_dxi, const double *dN_deta, double *dN_dx,
                       double *dN_dy);

struct se2w_source {
  struct se2w_spec *spec;
  int type, impl, use_internal;
  double coor[2], xi[2], values[4];
  int elem, closest_gll, issetup;
  double vals[2 * HOST_MAX_NB * HOST_MAX_NB];
  /* cubic-spline source: kernel width (requested, in use), \int W, and the
     elements it injects into with their values [nsub][2*npe] */
  double h_opt, kernel_h, int_w;
  int nsub;
  int32_t *sub_elem;
  double *sub_vals;
};

struct se2w_receivers {
  struct se2w_spec *spec;
  int nr;
  double *xr;
  int32_t *nid, *elem, *gll;
  char filename[1024];
  int header_written;
};

struct se2w_receiver_uv {
  struct se2w_spec *spec;
  double xr[2];
  int interpolate;            /* 0: nearest node (RecordUV), 1: interpolated (RecordUV_interp) */
  int source_implementation;  /* -1..3, names the file as the reference does */
  int32_t elem, nid;
  double *weights;            /* [npe]: basis values at the point, or a unit weight at the nearest node */
  double mapped[2];           /* nearest-node or interpolated coordinate */
  char filename[1024];
  int header_written;
};

#endif
