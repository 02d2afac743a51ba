/*
 * se2wave.h -- host-side C API of libse2w_host.so: the setup layer of
 * se2wave (basis, mesh, materials, mass, time step, source-time functions,
 * sources, receivers, options) and the solver that drives the GPU step
 * through the C ABI of include/se2w.h.
 *
 * Names map one-to-one onto the reference's (se2wave.c of
 * Nicolucas/se2dr_fork); each declaration cites the function it replaces.
 * All arithmetic is FP64 and follows the reference's operation order, so the
 * tables handed to the GPU are the reference's tables.
 * Status codes: 0 = success, negative = error; se2w_host_last_error() holds
 * the calling thread's last message (the analogue of SETERRQ's text).
 */
#ifndef SE2WAVE_H
#define SE2WAVE_H

#include <stdint.h>
#include <stdio.h>

#include "se2w.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SE2W_HOST_ERR_ARG -101
#define SE2W_HOST_ERR_MEM -102
#define SE2W_HOST_ERR_USER -103 /* PETSC_ERR_USER analogue */
#define SE2W_HOST_ERR_SUP -104  /* PETSC_ERR_SUP analogue */
#define SE2W_HOST_ERR_FILE -105

const char *se2w_host_last_error(void);

/* ---------------------------------------------------------------------- */
/* Basis (se2wave.c:103-690)                                               */
/* ---------------------------------------------------------------------- */
/* CreateGLLCoordsWeights (103-209): xi ascending, w in the unflipped order. */
int se2w_gll(int p, double *xi, double *w);
/* TabulateBasis1d_CLEGENDRE (253-385): N[k*(p+1)+i] = N_i(x_k). */
int se2w_basis1d(int p, int npoints, const double *x, double *N);
/* TabulateBasisDerivatives1d_CLEGENDRE (387-523): dN[k*(p+1)+i] = N_i'(x_k). */
int se2w_basis1d_deriv(int p, int npoints, const double *x, double *dN);
/* Derivative matrix at the GLL nodes, D[q*(p+1)+i] (the table inside
 * TabulateBasisDerivativesTensorProduct2d, 525-603). */
int se2w_deriv_matrix(int p, double *D);

/* ---------------------------------------------------------------------- */
/* SpecFECtx (se2wave.c:35-54, 692-884, 1148-1236)                         */
/* ---------------------------------------------------------------------- */
typedef struct se2w_spec *SE2WSpec;

/* SpecFECtxCreate + SpecFECtxCreateMesh_SEQ (692-704, 825-884): mesh of
 * mx*my elements of degree p on [0,1]^2, DMDA node order. */
int se2w_spec_create(int mx, int my, int p, SE2WSpec *spec);
int se2w_spec_destroy(SE2WSpec *spec);
/* SpecFECtxScaleMeshCoords (781-823): x = x*scale + shift per component. */
int se2w_spec_scale_coords(SE2WSpec spec, const double scale[2], const double shift[2]);
/* SpecFECtxSetConstantMaterialProperties(_Velocity) (1148-1175). */
int se2w_spec_set_constant_material(SE2WSpec spec, double lambda, double mu, double rho);
int se2w_spec_set_constant_material_velocity(SE2WSpec spec, double vp, double vs, double rho);
/* Cell-wise material arrays (element order e = ei + ej*mx). */
int se2w_spec_set_cell_material(SE2WSpec spec, const double *lambda, const double *mu, const double *rho);
int se2w_spec_set_cell_material_velocity(SE2WSpec spec, const double *vp, const double *vs, const double *rho);
/* SpecFECtxSetPerturbedMaterialProperties_Velocity (se2wave.c:1177-1236):
 * per cell Vp ~ U[vp0 - dvp, vp0 + dvp), Vs ~ U[vs0 - dvs, ...), rho ~
 * U[rho0 - drho, ...), drawn in element order from the drand48 stream seeded
 * 1, 2 and 3 respectively; prints the reference's three [material] lines. */
int se2w_spec_set_perturbed_material_velocity(SE2WSpec spec, double vp0, double delta_vp, double vs0, double delta_vs,
                                              double rho0, double delta_rho);

typedef struct {
  int32_t mx, my, p, nx, ny, npe;
  int64_t nnodes;
  double gmin[2], gmax[2]; /* DMGetBoundingBox */
} se2w_spec_info;
int se2w_spec_get_info(SE2WSpec spec, se2w_spec_info *info);
/* Copies of internal arrays (caller-allocated). */
int se2w_spec_get_coords(SE2WSpec spec, double *coords);      /* [2*nnodes] */
int se2w_spec_get_material(SE2WSpec spec, double *lambda, double *mu, double *rho); /* [mx*my] each */
int se2w_spec_get_tables(SE2WSpec spec, double *D, double *w1d, double *elem_dx, double *elem_dy);
int se2w_spec_get_element(SE2WSpec spec, int32_t *element); /* [mx*my*npe] element -> node map */

/* AssembleBilinearForm_Mass2d (1431-1498): M[node] (one value per node). */
int se2w_spec_mass(SE2WSpec spec, double *mass);
/* ElastoDynamicsComputeTimeStep_2d (2276-2349): CFL time step. */
int se2w_spec_timestep(SE2WSpec spec, double *dt);

/* ---------------------------------------------------------------------- */
/* SeismicSTF (se2wave.c:59-64, 3431-3468, 3499-3822)                      */
/* ---------------------------------------------------------------------- */
typedef struct se2w_stf *SE2WSTF;
struct se2w_stf {
  char name[64];
  void *data;
  int (*evaluate)(SE2WSTF stf, double time, double *psi);
  int (*destroy)(SE2WSTF stf);
};
int se2w_stf_create(const char *name, SE2WSTF *stf);                        /* SeismicSTFCreate 3431 */
int se2w_stf_evaluate(SE2WSTF stf, double time, double *psi);                /* SeismicSTFEvaluate 3445 */
int se2w_stf_destroy(SE2WSTF *stf);                                          /* SeismicSTFDestroy 3453 */
int se2w_stf_create_ricker(double t0, double freq, double amp, SE2WSTF *stf); /* 3528-3545 */
int se2w_stf_create_gaussian(double t0, double omega, double amp, SE2WSTF *stf); /* 3573-3590 */
int se2w_stf_create_yoffe(double tau_s, double tau_r, double d_max, SE2WSTF *stf); /* 3795-3822 */

/* ---------------------------------------------------------------------- */
/* SeismicSource (se2wave.c:66-90, 2995-3429)                              */
/* ---------------------------------------------------------------------- */
#define SE2W_SOURCE_FORCE 0  /* SOURCE_TYPE_FORCE (rejected by the reference; point force here) */
#define SE2W_SOURCE_MOMENT 1 /* SOURCE_TYPE_MOMENT */
#define SE2W_SOURCE_IMPL_POINTWISE 0
#define SE2W_SOURCE_IMPL_NEAREST_QPOINT 1
/* SOURCE_IMPL_SPLINE: the cubic-spline smoothed moment source of
 * AssembleLinearForm_ElastoDynamicsMomentDirac2d_Kernel_CSpline (1925-2057),
 * spread over every element its kernel (support 2h) reaches. */
#define SE2W_SOURCE_IMPL_SPLINE 2
#define SE2W_SOURCE_IMPL_DEFAULT 3
/* The P0 moment source of AssembleLinearForm_ElastoDynamicsMomentDirac2d_P0
 * (2059-2198), -source_impl 3 of ElastoDynamicsSetSourceImplementation
 * (2202-2229); 3 is SOURCE_IMPL_DEFAULT in the SeismicSource numbering. */
#define SE2W_SOURCE_IMPL_P0 4

typedef struct se2w_source *SE2WSource;
/* SeismicSourceCreate (2995-3243): locate the element and the nearest
 * (interior) GLL node.  values: force (Fx, Fy) or moment (Mxx, Mxy, Myx, Myy).
 * use_nearest_internal_gll mirrors -use_nearest_internal_gll (default 1). */
int se2w_source_create(SE2WSpec spec, int type, int impl, const double coor[2], const double *values,
                       int use_nearest_internal_gll, SE2WSource *src);
int se2w_source_setup(SE2WSource src);    /* SeismicSourceSetup 3245 */
int se2w_source_destroy(SE2WSource *src); /* SeismicSourceDestroy 3259 */
/* Element index, closest GLL node and the 2*npe nodal values of a source on
 * one element.  A cubic-spline source has no single element: elem and
 * closest_gll are -1 and asking for vals is an error (se2w_source_get_elements
 * gives its values). */
int se2w_source_get(SE2WSource src, int32_t *elem, int32_t *closest_gll, double *vals);
/* Kernel width h of a cubic-spline source: the -sm_h option (1975); h <= 0
 * keeps the default of 1957-1973 (1.5*(ds/2) for p <= 3, else
 * 3/(p-2)*(ds/2), ds the larger mesh spacing).  Call before setup. */
int se2w_source_set_kernel_width(SE2WSource src, double h);
/* The kernel width in use and the quadrature of W over the mesh (the values
 * printed at 1971-1978 and 2054) of a set-up cubic-spline source. */
int se2w_source_get_kernel(SE2WSource src, double *h, double *int_w);
/* The elements a set-up source injects into: *ne elements elem[*ne] and
 * their nodal values vals[*ne][2*npe].  A pointwise, nearest-GLL or P0
 * source has one.  A cubic-spline source assembles its vector over the
 * elements in ascending order (VecSetValues ADD_VALUES, 2046) and lists each
 * element that holds a node where that vector is nonzero first in ascending
 * order: the node's assembled value sits on that element, zero on the
 * others, so that adding the listed elements' values reproduces the
 * assembled vector exactly.  With elem = vals = NULL only *ne is set. */
int se2w_source_get_elements(SE2WSource src, int32_t *ne, int32_t *elem, double *vals);
/* SeismicSourceEvaluate (3470-3497) on a host vector g[2*nnodes]. */
int se2w_sources_evaluate(double time, int ns, SE2WSource *src, SE2WSTF *stf, double *g);

/* ---------------------------------------------------------------------- */
/* Receivers (RecordUVA_MultipleStations_NearestGLL, se2wave.c:2687-2993)   */
/* ---------------------------------------------------------------------- */
typedef struct se2w_receivers *SE2WReceivers;
int se2w_receivers_create(SE2WSpec spec, int nr, const double *xr, SE2WReceivers *recv);
int se2w_receivers_destroy(SE2WReceivers *recv);
int se2w_receivers_get(SE2WReceivers recv, int32_t *nid, int32_t *elem, int32_t *gll);
/* Writes the header of closestqpsource-receiverCP-uva-<mx>x<my>-p<p>-rank0.dat
 * and the .mpimeta file into `dir` (2808-2877). */
int se2w_receivers_write_header(SE2WReceivers recv, const char *dir);
/* Appends rows: time[k] then, per receiver, the 7 values trace[(k*nr+r)*7+c]
 * (2908-2962). */
int se2w_receivers_append_rows(SE2WReceivers recv, int nsteps, const double *times, const double *trace);

/* Single-station receivers of displacement and velocity: RecordUV
 * (interpolate = 0: the nearest node of the containing element,
 * se2wave.c:2351-2443) and RecordUV_interp (interpolate = 1: the element's
 * Lagrange interpolant at the station, se2wave.c:2445-2581).  The containing
 * element is found by truncation without clamping, as there; a station with
 * none is rejected.  source_implementation (-1..3) names the file as the
 * reference does ("<impl>source-receiverCP-..." / "<impl>source-receiver-...").
 * se2w_receiver_uv_get returns the element, the npe weights of its nodes and
 * the mapped (nearest-node or interpolated) coordinate. */
typedef struct se2w_receiver_uv *SE2WReceiverUV;
int se2w_receiver_uv_create(SE2WSpec spec, const double xr[2], int interpolate, int source_implementation,
                            SE2WReceiverUV *receiver);
int se2w_receiver_uv_destroy(SE2WReceiverUV *receiver);
int se2w_receiver_uv_get(SE2WReceiverUV receiver, int32_t *elem, double *weights, double mapped[2]);
int se2w_receiver_uv_write_header(SE2WReceiverUV receiver, const char *dir);
/* rows[nsteps][4] = ux uy vx vy */
int se2w_receiver_uv_append_rows(SE2WReceiverUV receiver, int nsteps, const double *times, const double *rows);

/* ---------------------------------------------------------------------- */
/* Options database (PetscOptionsGet{Int,Real,Bool,RealArray} analogue)     */
/* ---------------------------------------------------------------------- */
typedef struct se2w_options *SE2WOptions;
int se2w_options_create(int argc, char **argv, SE2WOptions *opts);
int se2w_options_destroy(SE2WOptions *opts);
int se2w_options_get_int(SE2WOptions opts, const char *name, int *value, int *found);
int se2w_options_get_real(SE2WOptions opts, const char *name, double *value, int *found);
int se2w_options_get_bool(SE2WOptions opts, const char *name, int *value, int *found);
int se2w_options_get_string(SE2WOptions opts, const char *name, char *value, size_t len, int *found);
/* Comma-separated list; *n is the capacity on input and the count on output. */
int se2w_options_get_real_array(SE2WOptions opts, const char *name, double *values, int *n, int *found);

/* ---------------------------------------------------------------------- */
/* Seeded material models used by the benchmark workloads                   */
/* ---------------------------------------------------------------------- */
/* splitmix64 stream: uniform doubles in [0,1) with 53-bit resolution. */
uint64_t se2w_rng_next(uint64_t *state);
double se2w_rng_uniform(uint64_t *state);
/* Per cell, element order: Vp ~ U[vp_lo,vp_hi], Vs = Vp/sqrt(3),
 * rho ~ U[rho_lo,rho_hi] (draws vp then rho per cell). */
/* POSIX drand48 generator (the reference's PETSCRAND48 random type). */
void se2w_rand48_seed(uint64_t *state, long seed);
double se2w_rand48_next(uint64_t *state);
int se2w_material_random_cells(int mx, int my, uint64_t seed, double vp_lo, double vp_hi, double rho_lo,
                               double rho_hi, double *vp, double *vs, double *rho);
/* nlayers horizontal layers of equal height; one (Vp, rho) draw per layer. */
int se2w_material_layered(int mx, int my, int nlayers, uint64_t seed, double vp_lo, double vp_hi, double rho_lo,
                          double rho_hi, double *vp, double *vs, double *rho);

/* ---------------------------------------------------------------------- */
/* Solver: the se2wave_demo time loop on the GPU (se2wave.c:4541-4807)      */
/* ---------------------------------------------------------------------- */
typedef struct se2w_solver *SE2WSolver;
/* Builds the GPU context from the spec, sources (with their STFs) and
 * receivers; the objects must outlive the solver.  Each source goes to the
 * GPU as its elements (se2w_source_get_elements), in source order, and each
 * step's table holds the source's STF value once per element. */
int se2w_solver_create(SE2WSpec spec, int precision, int schedule, int ns, SE2WSource *src, SE2WSTF *stf,
                       SE2WReceivers recv, SE2WSolver *solver);
int se2w_solver_destroy(SE2WSolver *solver);
se2w_ctx *se2w_solver_context(SE2WSolver solver);
/* Advances nsteps steps of size dt from the solver's current time (time is
 * accumulated as time = time + dt, se2wave.c:4724); receiver rows of these
 * steps are available through se2w_solver_get_trace. */
int se2w_solver_advance(SE2WSolver solver, int nsteps, double dt);
/* Sizes the host and device buffers of advances of up to nsteps steps
 * (se2w_reserve), so that such an advance allocates nothing. */
int se2w_solver_reserve(SE2WSolver solver, int nsteps);
double se2w_solver_time(SE2WSolver solver);
int se2w_solver_steps_done(SE2WSolver solver);
/* times[nsteps] and trace[nsteps][nr][7] of the last advance. */
int se2w_solver_get_trace(SE2WSolver solver, double *times, double *trace);
/* Single-station receivers sampled by the solver from the next advance on;
 * the objects must outlive the solver.  se2w_solver_get_uv_trace returns
 * trace[nsteps][n][4] (ux uy vx vy) of the last advance. */
int se2w_solver_set_uv_receivers(SE2WSolver solver, int n, SE2WReceiverUV *receivers);
int se2w_solver_get_uv_trace(SE2WSolver solver, double *trace);
/* Advances n solvers together, one batched GPU launch per step for all of
 * them; dt[i] is solver i's step.  ens is an ensemble of the solvers'
 * contexts (se2w_solver_context) in the same order.  Each solver's time, step
 * count and trace advance as with se2w_solver_advance. */
int se2w_solvers_advance(int n, SE2WSolver *solvers, se2w_ensemble *ens, int nsteps, const double *dt);

/* Runs the se2wave program (main + se2wave_demo) with reference options. */
int se2w_main(int argc, char **argv);

#ifdef __cplusplus
}
#endif
#endif /* SE2WAVE_H */
