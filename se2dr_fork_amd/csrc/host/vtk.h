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

/* A snapshot written while the step loop goes on: se2w_vtk_job_start takes
 * the context's u and v (se2w_snapshot_begin) and writes them to dir/fname
 * from a second thread; se2w_vtk_job_join waits for that thread and returns
 * its status.  One job at a time: start joins the previous job first.  Join
 * before the context is destroyed. */
typedef struct {
  SE2WSpec spec;
  se2w_ctx *ctx;
  char path_dir[1024], fname[64];
  int active, ierr;
  char msg[512];
  void *thread;
} SE2WVtkJob;
int se2w_vtk_job_start(SE2WVtkJob *job, SE2WSpec spec, se2w_ctx *ctx, const char *dir, const char *fname);
int se2w_vtk_job_join(SE2WVtkJob *job);
#ifdef __cplusplus
}
#endif
#endif
