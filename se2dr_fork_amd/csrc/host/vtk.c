/*
 * vtk.c -- VTK XML StructuredGrid snapshots (the reference writes u, v with
 * PetscViewerVTKOpen + VecView on the DMDA: uva.vts, f.vts, step-%.4D.vts,
 * se2wave.c:4606-4610, 4665-4669, 4764-4770, 4785-4789).  Raw appended
 * little-endian Float64 arrays; the DMDA field names "_x", "_y" are appended
 * to the vector name as the reference's DMDASetFieldName (1143-1144) does.
 */
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "host_internal.h"
#include "vtk.h"

static int write_block(FILE *fp, const void *data, uint64_t nbytes)
{
  if (fwrite(&nbytes, sizeof(nbytes), 1, fp) != 1) return -1;
  if (nbytes && fwrite(data, 1, (size_t)nbytes, fp) != (size_t)nbytes) return -1;
  return 0;
}

int se2w_vtk_write(SE2WSpec s, const char *dir, const char *fname, const char *name0, const double *f0,
                   const char *name1, const double *f1)
{
  char path[1200];
  FILE *fp;
  const int64_t n = s->nnodes;
  const uint64_t pbytes = (uint64_t)n * 3 * sizeof(double), fbytes = (uint64_t)n * sizeof(double);
  uint64_t off = 0;
  double *buf;
  int64_t i;
  int c, nf = (f1 && name1) ? 2 : 1;
  const char *names[2] = {name0, name1};
  const double *fields[2] = {f0, f1};
  if (!s || !fname || !name0 || !f0) return host_fail(SE2W_HOST_ERR_ARG, "bad vtk arguments");
  snprintf(path, sizeof(path), "%s/%s", (dir && dir[0]) ? dir : ".", fname);
  fp = fopen(path, "wb");
  if (!fp) return host_fail(SE2W_HOST_ERR_FILE, "Failed to open file \"%s\"", path);
  fprintf(fp, "<?xml version=\"1.0\"?>\n");
  fprintf(fp, "<VTKFile type=\"StructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n");
  fprintf(fp, "  <StructuredGrid WholeExtent=\"0 %d 0 %d 0 0\">\n", s->nx - 1, s->ny - 1);
  fprintf(fp, "    <Piece Extent=\"0 %d 0 %d 0 0\">\n", s->nx - 1, s->ny - 1);
  fprintf(fp, "      <Points>\n        <DataArray type=\"Float64\" Name=\"Position\" NumberOfComponents=\"3\" format=\"appended\" offset=\"%llu\"/>\n      </Points>\n",
          (unsigned long long)off);
  off += 8 + pbytes;
  fprintf(fp, "      <PointData>\n");
  for (c = 0; c < nf; c++) {
    fprintf(fp, "        <DataArray type=\"Float64\" Name=\"%s_x\" NumberOfComponents=\"1\" format=\"appended\" offset=\"%llu\"/>\n", names[c], (unsigned long long)off);
    off += 8 + fbytes;
    fprintf(fp, "        <DataArray type=\"Float64\" Name=\"%s_y\" NumberOfComponents=\"1\" format=\"appended\" offset=\"%llu\"/>\n", names[c], (unsigned long long)off);
    off += 8 + fbytes;
  }
  fprintf(fp, "      </PointData>\n    </Piece>\n  </StructuredGrid>\n  <AppendedData encoding=\"raw\">\n_");
  buf = (double *)malloc((size_t)pbytes);
  if (!buf) { fclose(fp); return host_fail(SE2W_HOST_ERR_MEM, "out of memory"); }
  for (i = 0; i < n; i++) {
    buf[3 * i] = spec_x(s, (int)(i % s->nx));
    buf[3 * i + 1] = spec_y(s, (int)(i / s->nx));
    buf[3 * i + 2] = 0.0;
  }
  if (write_block(fp, buf, pbytes)) { free(buf); fclose(fp); return host_fail(SE2W_HOST_ERR_FILE, "write failed"); }
  for (c = 0; c < nf; c++) {
    int comp;
    for (comp = 0; comp < 2; comp++) {
      for (i = 0; i < n; i++) buf[i] = fields[c][2 * i + comp];
      if (write_block(fp, buf, fbytes)) { free(buf); fclose(fp); return host_fail(SE2W_HOST_ERR_FILE, "write failed"); }
    }
  }
  free(buf);
  fprintf(fp, "\n  </AppendedData>\n</VTKFile>\n");
  fclose(fp);
  return 0;
}

int se2w_vtk_write_state(SE2WSpec s, se2w_ctx *ctx, const char *dir, const char *fname)
{
  const size_t n2 = 2 * (size_t)s->nnodes;
  double *u = (double *)malloc(sizeof(double) * n2), *v = (double *)malloc(sizeof(double) * n2);
  int ierr;
  if (!u || !v) { free(u); free(v); return host_fail(SE2W_HOST_ERR_MEM, "out of memory"); }
  ierr = se2w_get_state(ctx, u, v, NULL);
  if (ierr) { free(u); free(v); return host_fail(ierr, "GPU: %s", se2w_last_error()); }
  ierr = se2w_vtk_write(s, dir, fname, "disp", u, "velo", v);
  free(u); free(v);
  return ierr;
}

static void *vtk_job_run(void *arg)
{
  SE2WVtkJob *j = (SE2WVtkJob *)arg;
  const double *u = NULL, *v = NULL;
  int ierr = se2w_snapshot_wait(j->ctx, &u, &v);
  if (ierr) {
    j->ierr = ierr;
    snprintf(j->msg, sizeof(j->msg), "GPU: %s", se2w_last_error());
    return NULL;
  }
  ierr = se2w_vtk_write(j->spec, j->path_dir, j->fname, "disp", u, "velo", v);
  if (ierr) {
    j->ierr = ierr;
    snprintf(j->msg, sizeof(j->msg), "%s", se2w_host_last_error());
  }
  return NULL;
}

int se2w_vtk_job_join(SE2WVtkJob *j)
{
  if (!j || !j->active) return 0;
  pthread_join(*(pthread_t *)j->thread, NULL);
  free(j->thread);
  j->thread = NULL;
  j->active = 0;
  if (j->ierr) return host_fail(j->ierr, "%s", j->msg);
  return 0;
}

int se2w_vtk_job_start(SE2WVtkJob *j, SE2WSpec spec, se2w_ctx *ctx, const char *dir, const char *fname)
{
  int ierr;
  if (!j || !spec || !ctx || !fname) return host_fail(SE2W_HOST_ERR_ARG, "bad vtk job arguments");
  if ((ierr = se2w_vtk_job_join(j))) return ierr;
  if ((ierr = se2w_snapshot_begin(ctx))) return host_fail(ierr, "GPU: %s", se2w_last_error());
  j->spec = spec;
  j->ctx = ctx;
  snprintf(j->path_dir, sizeof(j->path_dir), "%s", (dir && dir[0]) ? dir : ".");
  snprintf(j->fname, sizeof(j->fname), "%s", fname);
  j->ierr = 0;
  j->msg[0] = 0;
  j->thread = malloc(sizeof(pthread_t));
  if (!j->thread) return host_fail(SE2W_HOST_ERR_MEM, "out of memory");
  if (pthread_create((pthread_t *)j->thread, NULL, vtk_job_run, j)) {
    free(j->thread);
    j->thread = NULL;
    vtk_job_run(j); /* no second thread: write the snapshot here */
    return j->ierr ? host_fail(j->ierr, "%s", j->msg) : 0;
  }
  j->active = 1;
  return 0;
}
