/* vtk.h -- structured-grid (.vts) snapshots of nodal fields. */
#ifndef SE2W_VTK_H
#define SE2W_VTK_H
#include "se2wave.h"
#ifdef __cplusplus
extern "C" {
#endif
/* Writes the mesh and one or two 2-component nodal fields (interleaved DMDA
 * order) as point arrays <name>_x, <name>_y. */
int se2w_vtk_write(SE2WSpec spec, const char *dir, const char *fname, const char *name0, const double *f0,
                   const char *name1, const double *f1);
/* Writes u ("disp") and v ("velo") of a GPU context. */
int se2w_vtk_write_state(SE2WSpec spec, se2w_ctx *ctx, const char *dir, const char *fname);
#ifdef __cplusplus
}
#endif
#endif
