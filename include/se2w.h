/*
 * se2w.h -- C ABI of libse2w_hip.so, the MI355X (gfx950) implementation of
 * se2wave's explicit spectral-element time step.
 *
 * Plain C: pointers, sizes and int status codes only.  All arrays passed in
 * are host arrays owned by the caller and copied during the call.  Nodal
 * arrays are interleaved (x, y) per node in DMDA order, node = I + J*nx with
 * nx = p*mx + 1 (se2wave.c:706-736); per-cell arrays are in element order
 * e = ei + ej*mx.  Status: 0 = success, negative = error (like a
 * PetscErrorCode); se2w_strerror() and se2w_last_error() describe it.
 *
 * Threading: one context per host thread; calls on a context are ordered on
 * the context's own HIP stream; the se2w_get_* calls synchronise it.
 */
#ifndef SE2W_H
#define SE2W_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Arithmetic modes. */
#define SE2W_FP64 0       /* reference FP64 operation order, no FMA contraction */
#define SE2W_FP32 1       /* FP32 operator and update, tables formed in FP64 */
#define SE2W_FP32_KAHAN 2 /* FP32 with compensated displacement accumulation */

/* Step schedules. */
#define SE2W_SCHED_FUSED 0 /* one kernel per step: element operator + update */
#define SE2W_SCHED_SPLIT 1 /* element kernel writes f, update kernel consumes it */

/* Error codes. */
#define SE2W_OK 0
#define SE2W_ERR_ARG -1
#define SE2W_ERR_MEM -2
#define SE2W_ERR_HIP -3
#define SE2W_ERR_STATE -4
#define SE2W_ERR_UNSUPPORTED -5

typedef struct se2w_ctx se2w_ctx;

typedef struct {
  int32_t mx, my;    /* elements in x and y */
  int32_t p;         /* polynomial degree (1..8) */
  int32_t precision; /* SE2W_FP64 | SE2W_FP32 | SE2W_FP32_KAHAN */
  int32_t schedule;  /* SE2W_SCHED_FUSED | SE2W_SCHED_SPLIT */
  int32_t device;    /* HIP device ordinal */
} se2w_desc;

/* Replaces SpecFECtxCreate + SpecFECtxCreateMesh_SEQ + DMCreateGlobalVector
 * (se2wave.c:692-704, 825-884, 4594-4601): device storage for a mesh of
 * mx*my elements of degree p; u = v = a = 0. */
int se2w_create(const se2w_desc *desc, se2w_ctx **ctx);

/* Replaces SpecFECtx's basis tables and per-element geometry
 * (se2wave.c:860-868, 1238-1279): D[q*(p+1)+i] = dN_i/dxi at GLL node q,
 * w1d[p+1] GLL weights, elem_dx[mx] / elem_dy[my] element widths
 * (x of local node (p,0) minus x of local node (0,0), per element column;
 * y of local node (p,p) minus y of local node (0,0), per element row). */
int se2w_set_tables(se2w_ctx *ctx, const double *D, const double *w1d, const double *elem_dx,
                    const double *elem_dy);

/* Replaces SpecFECtx->cell_data (se2wave.c:29-33, 1148-1159): per-cell Lame
 * parameters lambda[mx*my], mu[mx*my]. */
int se2w_set_cells(se2w_ctx *ctx, const double *lambda, const double *mu);

/* Replaces the Md vector of AssembleBilinearForm_Mass2d (se2wave.c:1431-1498):
 * lumped mass, one value per node (the reference stores it on both dofs). */
int se2w_set_mass(se2w_ctx *ctx, const double *mass);

/* Replaces the PointwiseContext of SeismicSourceSetup_Pointwise
 * (se2wave.c:3310-3365, 3367-3429): for each of ns sources the element index
 * elem[s] and the 2*(p+1)^2 nodal values vals[s][2k+c] that
 * SeismicSourceAddValues_Pointwise scales by the STF (se2wave.c:3277-3293). */
int se2w_set_sources(se2w_ctx *ctx, int32_t ns, const int32_t *elem, const double *vals);

/* Replaces the receiver location of RecordUVA_MultipleStations_NearestGLL_MPI
 * (se2wave.c:2700-2805): element index and element-local GLL index of the
 * node nearest to each of nr receivers. */
int se2w_set_receivers(se2w_ctx *ctx, int32_t nr, const int32_t *elem, const int32_t *gll);

/* Replaces the sampling of RecordUV and RecordUV_interp (se2wave.c:2351-2581):
 * each step records u and v at every node of element elem[r] (element order
 * e = ei + ej*mx).  se2w_get_element_receivers returns, for every step k of
 * the last run, trace[k][r][4] = sum over the element's nodes q (local index
 * q = i + j*(p+1), ascending) of weights[r][q] * (ux, uy, vx, vy) at node q:
 * Lagrange basis values interpolate at a point, a unit weight picks the
 * nearest node. */
int se2w_set_element_receivers(se2w_ctx *ctx, int32_t n, const int32_t *elem);
int se2w_get_element_receivers(se2w_ctx *ctx, const double *weights, double *trace);
/* Sets / reads the state vectors u, v, a (each 2*nx*ny doubles; NULL skips).
 * Replaces VecGetArray access to the u, v, a Vecs of se2wave_demo. */
int se2w_set_state(se2w_ctx *ctx, const double *u, const double *v, const double *a);
int se2w_get_state(se2w_ctx *ctx, double *u, double *v, double *a);
/* Snapshot of u and v for the VTK writer that lets the step loop go on
 * (replaces the VecView of u and v every -of steps, se2wave.c:4764-4770,
 * which the reference's loop waits for).  se2w_snapshot_begin enqueues on
 * the context's stream the conversion of the current u and v to interleaved
 * FP64 in DMDA order into a device staging buffer, and on a second stream of
 * the context the copy of that buffer into pinned host memory; it returns at
 * once, and steps enqueued afterwards run while the copy proceeds.
 * se2w_snapshot_wait (callable from another host thread, though not while
 * se2w_snapshot_begin or se2w_destroy runs on the context) blocks until the
 * copy is complete and returns the host arrays, 2*nx*ny doubles each, valid
 * until the next se2w_snapshot_begin or se2w_destroy.  A begin while a
 * snapshot is in flight first waits for it.  The staging buffers (4*nx*ny
 * doubles on the device and as many in pinned host memory) are allocated by
 * the first begin and kept. */
int se2w_snapshot_begin(se2w_ctx *ctx);
int se2w_snapshot_wait(se2w_ctx *ctx, const double **u, const double **v);

/* Replaces the time loop body of se2wave_demo (se2wave.c:4722-4775) for
 * nsteps steps of size dt: predictor, SeismicSourceEvaluate with
 * psi[k*ns + s] = stf_s(time_k) (3470-3497), AssembleLinearForm (1285-1429),
 * f += g, a = f / M, corrector, receiver sampling (2904-2962).  The caller
 * evaluates the source-time functions (its own STF objects) at the
 * accumulated times. */
int se2w_run(se2w_ctx *ctx, int32_t nsteps, double dt, const double *psi);

/* Sizes, ahead of a run, the buffers a run of up to nsteps steps needs (the
 * device source-time table, the device receiver records and the pinned host
 * buffer of the record download), so that se2w_run and se2w_get_receivers
 * of such a run allocate nothing.  No counterpart in the reference, whose
 * loop allocates nothing per step: the buffers here hold a whole run.
 * se2w_get_reserved reports the current capacities {psi values, receiver
 * record values, element-receiver record values, pinned bytes}. */
int se2w_reserve(se2w_ctx *ctx, int32_t nsteps);
int se2w_get_reserved(se2w_ctx *ctx, int64_t caps[4]);

/* counts = {sources, receivers, element receivers, steps of the last run}:
 * the sizes of the psi table a run reads (nsteps * sources values) and of
 * the rows se2w_get_receivers / se2w_get_element_receivers write. */
int se2w_get_counts(se2w_ctx *ctx, int32_t counts[4]);

/* Receiver rows of the last se2w_run: trace[(k*nr + r)*7 + c] with columns
 * ux uy vx vy ax ay curl(v), as RecordUVA_MultipleStations_NearestGLL_MPI
 * writes them (se2wave.c:2908-2959). */
int se2w_get_receivers(se2w_ctx *ctx, double *trace);

/* Replaces AssembleLinearForm_ElastoDynamics2d(c, u, F) (se2wave.c:1285-1429):
 * f = -sum_e B^T C B u_e for the given u (2*nx*ny doubles).  The context's
 * state is not touched. */
int se2w_assemble_forces(se2w_ctx *ctx, const double *u, double *f);

/* se2w_assemble_forces without allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * kernels and the same result as se2w_assemble_forces, with the staging and
 * padded copies of u and f in a fixed 64 MiB device array that each GPU
 * holds for the life of the process (FP64: meshes up to about 1.3 M nodes;
 * FP32 modes: about 2.0 M; larger meshes return SE2W_ERR_UNSUPPORTED).  The
 * array keeps the last call's contents; calls that overlap on one GPU (two
 * host threads, or two contexts' streams) race for it and must be
 * serialised by the caller.  It calls no hipMalloc or hipFree. */
int se2w_assemble_forces_noalloc(se2w_ctx *ctx, const double *u, double *f);

/* max, min and 2-norm over all dofs of u and of v, the quantities the
 * reference prints every 100 steps (VecMax/VecMin/VecNorm, se2wave.c:4747-4755):
 * out = {umax, umin, ul2, vmax, vmin, vl2}. */
int se2w_norms(se2w_ctx *ctx, double out[6]);

/* Wall time of the step loop of the last se2w_run (HIP events on the
 * context's stream, predictor included) and, when profiling is on, the mean
 * duration of the step kernel(s): kernel_ms[0] = fused step kernel (events
 * around the whole sequence of step launches, divided by the step count) or
 * element kernel, kernel_ms[1] = update kernel (split schedule; events
 * around each launch). */
int se2w_set_profiling(se2w_ctx *ctx, int32_t on);
int se2w_get_timing(se2w_ctx *ctx, double *run_ms, double kernel_ms[2]);

/* Element-kernel variants: SE2W_KERNEL_LDS, or SE2W_KERNEL_REG + k for the
 * k-th schedule (workgroup shape, register cap, placement of the node loads,
 * one lane or a lane pair per element) of the register-blocked FP32 kernel
 * (2 <= p <= 8; at p = 7, 8 only the lane-pair schedules);
 * se2w_kernel_count() variants in all. */
#define SE2W_KERNEL_LDS 0 /* LDS-staged tile kernel (all modes, 1 <= p <= 8) */
#define SE2W_KERNEL_REG 1 /* register-blocked FP32 kernel, first schedule */
#define SE2W_KERNEL_MFMA 9 /* FP32, 6 <= p <= 8, fused: element contractions on the matrix cores */
#define SE2W_KERNEL_LINE 10 /* FP32, 6 <= p <= 8, fused: the same layout, contractions one line per thread */
#define SE2W_KERNEL_LDS_WIDE 11 /* the LDS-staged tile kernel with 512 threads per workgroup (all modes, fused) */
#define SE2W_KERNEL_REG64 12 /* FP64, 2 <= p <= 8, fused: register-blocked kernel in the reference's operation
                                order, first schedule; SE2W_KERNEL_REG64 + k for the k-th (nine in all: a lane
                                pair or one lane per element, several tile shapes; above p = 4 the lane pairs;
                                k = 0..6 up to p = 7, k = 7, 8 at p = 8, one wave per SIMD) */

/* Selects / reports the element-kernel variant (no counterpart in the
 * reference: a schedule choice of this implementation, used by the
 * autotuner).  Unavailable variants return SE2W_ERR_UNSUPPORTED. */
int se2w_set_kernel(se2w_ctx *ctx, int32_t variant);
int se2w_get_kernel(se2w_ctx *ctx, int32_t *variant);
int se2w_kernel_count(void);

/* Autotuner (no counterpart in the reference): runs `reps` timed trials of
 * `steps` steps for every element-kernel variant built for this context's
 * degree, precision and schedule, keeps the fastest (median kernel time) and
 * reports it.  The state u, v, a is restored afterwards; receiver rows of the
 * previous se2w_run are discarded.  The choice is cached per (p, precision,
 * schedule, size bucket) for the rest of the process. */
int se2w_autotune(se2w_ctx *ctx, int32_t steps, int32_t reps, int32_t *best_variant, double *best_ms);

/* Tile shape (owned elements per workgroup) used by the step kernels. */
int se2w_get_tile(se2w_ctx *ctx, int32_t *tex, int32_t *tey);

/* ------------------------------------------------------------------------
 * Ensembles (no counterpart in the reference, which runs one instance per
 * program): contexts of one device, degree and precision -- each with its own
 * mesh size, materials, sources, receivers and dt -- advanced together, one
 * batched kernel launch per step, the launch sequence of a run captured in a
 * HIP graph and replayed.  Members keep their own state and receiver record:
 * after a run, se2w_get_receivers / se2w_get_state on each member.  The
 * ensemble does not own its members; destroy it before them.
 * ------------------------------------------------------------------------ */
typedef struct se2w_ensemble se2w_ensemble;

/* members[0..n-1]: fused-schedule contexts with tables, cells and mass set. */
int se2w_ensemble_create(int32_t n, se2w_ctx *const *members, se2w_ensemble **ens);
/* nsteps steps of every member: dt[i] its time step, psi[i] its
 * source-time table [nsteps][ns_i] (may be NULL for members without sources). */
int se2w_ensemble_run(se2w_ensemble *ens, int32_t nsteps, const double *dt, const double *const *psi);
/* Kernel variant of the batched step (SE2W_KERNEL_LDS or SE2W_KERNEL_REG + k),
 * applied to every member. */
int se2w_ensemble_set_kernel(se2w_ensemble *ens, int32_t variant);
int se2w_ensemble_get_kernel(se2w_ensemble *ens, int32_t *variant);
/* Times every built variant over `steps` batched steps (reps repetitions,
 * state restored afterwards) and keeps the fastest. */
int se2w_ensemble_autotune(se2w_ensemble *ens, int32_t steps, int32_t reps, int32_t *best_variant, double *best_ms);
/* on = 1 (default): replay the run's launches from a captured HIP graph;
 * on = 0: launch them one by one. */
int se2w_ensemble_set_graph(se2w_ensemble *ens, int32_t on);
/* Last run: whole run (predictors included) and mean time per batched step. */
int se2w_ensemble_get_timing(se2w_ensemble *ens, double *run_ms, double *step_ms);
/* Tiles of all members and workgroups of the batched launch (one per tile,
 * rounded up to a multiple of 8). */
int se2w_ensemble_get_launch(se2w_ensemble *ens, int32_t *tiles, int32_t *grid);
/* Member i of the ensemble. */
int se2w_ensemble_get_member(se2w_ensemble *ens, int32_t i, se2w_ctx **member);
int se2w_ensemble_destroy(se2w_ensemble *ens);

int se2w_destroy(se2w_ctx *ctx);
const char *se2w_strerror(int code);
const char *se2w_last_error(void);
int se2w_version(void);

#ifdef __cplusplus
}
#endif
#endif /* SE2W_H */
