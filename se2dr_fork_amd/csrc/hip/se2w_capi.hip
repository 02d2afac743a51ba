// se2w_capi.hip -- host side of libse2w_hip.so: the C ABI declared in
// include/se2w.h, device memory management and kernel dispatch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "se2w.h"
#include "se2w_kernels.hpp"
#include "se2w_kernels_reg.hpp"
#include "se2w_kernels_mfma.hpp"
#include "se2w_kernels_reg64.hpp"
#include "se2w_ens.hpp"
#include "se2w_launch.hpp"

using namespace se2w;

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

#define HIPCHK(call)                                                                         \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess) return fail(SE2W_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

// Conversion between the caller's DMDA double layout and the padded device
// layout of the context's precision.
template <typename R>
__global__ void k_from_dmda(const double *src, typename V2<R>::T *dst, int nx, int ny, int pitch) {
  const long n = (long)nx * ny;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long J = i / nx, I = i - J * nx;
    dst[J * pitch + I] = typename V2<R>::T{(R)src[2 * i], (R)src[2 * i + 1]};
  }
}

template <typename R>
__global__ void k_to_dmda(const typename V2<R>::T *src, const typename V2<R>::T *comp, double *dst, int nx, int ny,
                          int pitch) {
  const long n = (long)nx * ny;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long J = i / nx, I = i - J * nx;
    const typename V2<R>::T s = src[J * pitch + I];
    double x = (double)s.x, y = (double)s.y;
    if (comp) {  // compensated value: s - c
      const typename V2<R>::T c = comp[J * pitch + I];
      x -= (double)c.x;
      y -= (double)c.y;
    }
    dst[2 * i] = x;
    dst[2 * i + 1] = y;
  }
}

template <typename R>
__global__ void k_scalar_from_dmda(const double *src, R *dst, int nx, int ny, int pitch, int invert) {
  const long n = (long)nx * ny;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long J = i / nx, I = i - J * nx;
    const double m = src[i];
    dst[J * pitch + I] = invert ? (R)(1.0 / m) : (R)m;
  }
}

// Partial max/min/sum-of-squares of u and v over blocks (double accumulation).
template <typename R>
__global__ __launch_bounds__(256) void k_norms(const typename V2<R>::T *u, const typename V2<R>::T *v, int nx, int ny,
                                               int pitch, double *partial) {
  __shared__ double red[6][256];
  double umax = -1.0e300, umin = 1.0e300, usq = 0.0, vmax = -1.0e300, vmin = 1.0e300, vsq = 0.0;
  const long n = (long)nx * ny;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long J = i / nx, I = i - J * nx;
    const typename V2<R>::T a = u[J * pitch + I], b = v[J * pitch + I];
    const double ax = a.x, ay = a.y, bx = b.x, by = b.y;
    umax = fmax(umax, fmax(ax, ay));
    umin = fmin(umin, fmin(ax, ay));
    usq += ax * ax + ay * ay;
    vmax = fmax(vmax, fmax(bx, by));
    vmin = fmin(vmin, fmin(bx, by));
    vsq += bx * bx + by * by;
  }
  const int t = threadIdx.x;
  red[0][t] = umax; red[1][t] = umin; red[2][t] = usq;
  red[3][t] = vmax; red[4][t] = vmin; red[5][t] = vsq;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
      red[0][t] = fmax(red[0][t], red[0][t + s]);
      red[1][t] = fmin(red[1][t], red[1][t + s]);
      red[2][t] += red[2][t + s];
      red[3][t] = fmax(red[3][t], red[3][t + s]);
      red[4][t] = fmin(red[4][t], red[4][t + s]);
      red[5][t] += red[5][t + s];
    }
    __syncthreads();
  }
  if (t == 0)
    for (int c = 0; c < 6; c++) partial[blockIdx.x * 6 + c] = red[c][0];
}

size_t real_size(int mode) { return mode == SE2W_FP64 ? sizeof(double) : sizeof(float); }

// Zero fill (a kernel rather than hipMemsetAsync, so that it can sit in a
// captured stream).
__global__ void k_zero(double *p, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = 0.0;
}

}  // namespace

// Scratch of se2w_assemble_forces_noalloc: the host-layout staging copy of u
// and f and the padded u and f of the element kernel.  Every GPU holds its
// own copy of this array for the life of the process; host code uses only
// the per-device address from hipGetSymbolAddress (the name itself is a
// host-side stand-in).
static constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on `device`, looked up once per device with
// that device current (the caller has made it current).
int scratch_address(int device, void **out) {
  if (device < 0 || device >= kMaxDevices) return fail(SE2W_ERR_ARG, "device %d out of range", device);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(SE2W_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    g_scratch_dev[device] = p;
  }
  *out = g_scratch_dev[device];
  return SE2W_OK;
}

}  // namespace

struct se2w_ctx {
  int mx = 0, my = 0, p = 0, nb = 0, npe = 0, nx = 0, ny = 0, pitch = 0;
  int mode = 0, schedule = 0, device = 0, variant = 0;
  int tex = 0, tey = 0, tiles_x = 0, tiles_y = 0, ntiles = 0, tiles_per_xcd = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  void *U[2] = {nullptr, nullptr}, *C[2] = {nullptr, nullptr};
  void *V = nullptr, *A = nullptr, *F = nullptr, *Utmp = nullptr;
  // The next run's half-step velocity and (FP32+Kahan) compensation, written
  // by a run's last step together with the next run's predicted displacement
  // in U[cur ^ 1]; valid (pred_ok) until the state is written from the host,
  // for a next run with the same dt (pred_dt).
  void *Vn = nullptr, *Cn = nullptr;
  bool pred_ok = false;
  double pred_dt = 0.0;
  void *mass = nullptr, *lambda = nullptr, *mu = nullptr, *edx = nullptr, *edy = nullptr, *D = nullptr,
       *w2 = nullptr;
  int cur = 0;
  bool have_tables = false, have_cells = false, have_mass = false;
  // host copies needed for receiver post-processing
  std::vector<double> hD, hdx, hdy;
  // sources
  int ns = 0;
  std::vector<int> h_src_elem;
  int *d_src_elem = nullptr, *d_tsrc_ptr = nullptr, *d_tsrc_idx = nullptr;
  void *d_src_vals = nullptr;
  void *d_psi = nullptr;
  void *h_psi = nullptr;  // pinned host copy of a run's source-time table (psi_cap values)
  size_t psi_cap = 0;
  // receivers
  int nr = 0;
  std::vector<int> h_rec_elem, h_rec_gll;
  int *d_rec_elem = nullptr, *d_rec_gll = nullptr, *d_trec_ptr = nullptr, *d_trec_idx = nullptr;
  void *d_rec = nullptr;
  size_t rec_cap = 0;
  void *h_stage = nullptr;  // pinned host staging buffer for the record download
  size_t stage_cap = 0;
  bool rec_staged = false;  // h_stage holds the last run's receiver record
  // element receivers (u, v at every node of an element, for interpolation)
  int ner = 0;
  std::vector<int> h_erec_elem;
  int *d_erec_elem = nullptr, *d_terec_ptr = nullptr, *d_terec_idx = nullptr;
  void *d_erec = nullptr;
  size_t erec_cap = 0;
  int erec_steps = 0;
  int rec_steps = 0;
  std::vector<double> trace;
  // timing
  bool profiling = false;
  double last_ms = 0.0, kern_ms[2] = {0.0, 0.0};
  double *d_partial = nullptr;
  // asynchronous snapshots of u, v (FP64, DMDA order): device staging
  // buffer, pinned host copy, the copy stream and its events
  void *snap_dev = nullptr;
  double *snap_host = nullptr;
  hipStream_t copy_stream = nullptr;
  hipEvent_t snap_ready = nullptr, snap_done = nullptr;
  bool snap_pending = false;
};

namespace {

int dev_alloc(void **p, size_t bytes) {
  if (*p) return SE2W_OK;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(SE2W_ERR_MEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  }
  return SE2W_OK;
}

void dev_free(void *&p) {
  if (p) hipFree(p);
  p = nullptr;
}

// Device scratch of one call, freed on every return path.
struct ScopedDev {
  void *p = nullptr;
  ~ScopedDev() {
    if (p) hipFree(p);
  }
};

template <typename T>
int upload(void **dst, const T *src, size_t n, hipStream_t s) {
  dev_free(*dst);
  int ierr = dev_alloc(dst, sizeof(T) * std::max<size_t>(n, 1));
  if (ierr) return ierr;
  if (n) HIPCHK(hipMemcpyAsync(*dst, src, sizeof(T) * n, hipMemcpyHostToDevice, s));
  HIPCHK(hipStreamSynchronize(s));
  return SE2W_OK;
}

// Real-typed upload of a double host array.
int upload_real(se2w_ctx *c, void **dst, const double *src, size_t n) {
  if (c->mode == SE2W_FP64) return upload<double>(dst, src, n, c->stream);
  std::vector<float> tmp(n);
  for (size_t i = 0; i < n; i++) tmp[i] = (float)src[i];
  return upload<float>(dst, tmp.data(), n, c->stream);
}

// Register-blocked kernel availability: FP32 modes, 2 <= p <= 8 (p = 7, 8:
// the lane-pair schedules only, see reg_schedule_built).
bool reg_available(int p, int mode) { return mode != SE2W_FP64 && p >= 2 && p <= kMaxP; }

// Matrix-core kernel availability: FP32 modes, 6 <= p <= 8, fused schedule.
static_assert(SE2W_KERNEL_MFMA == SE2W_KERNEL_REG + kNumRegSchedules, "variant numbering");
bool mfma_available(int p, int mode, int schedule) {
  return mode != SE2W_FP64 && p >= 6 && p <= 8 && schedule == SE2W_SCHED_FUSED;
}

// variant = SE2W_KERNEL_REG + schedule index into kRegSchedules
int reg_schedule_of(int variant) { return variant - SE2W_KERNEL_REG; }
bool is_reg_variant(int variant) { return variant >= SE2W_KERNEL_REG && variant < SE2W_KERNEL_REG + kNumRegSchedules; }

// variant = SE2W_KERNEL_REG64 + schedule index into kReg64Schedules (FP64)
static_assert(SE2W_KERNEL_REG64 == SE2W_KERNEL_LDS_WIDE + 1, "variant numbering");
bool is_reg64_variant(int variant) {
  return variant >= SE2W_KERNEL_REG64 && variant < SE2W_KERNEL_REG64 + kNumReg64Schedules;
}
bool reg64_variant_built(int p, int mode, int variant, int schedule) {
  return mode == SE2W_FP64 && is_reg64_variant(variant) &&
         reg64_schedule_built(p, variant - SE2W_KERNEL_REG64, schedule == SE2W_SCHED_FUSED);
}

void reg_shape(int variant, int &tex, int &tey) {
  const RegSchedule &r = kRegSchedules[reg_schedule_of(variant)];
  tex = r.ex - 1;
  tey = r.ey - 1;
}

// Tile shapes of the LDS-staged kernels compiled for (p, precision).
void tile_shape(int p, int mode, int &tex, int &tey) {
#define CASE(P_)                                                                                     \
  case P_:                                                                                           \
    if (mode == SE2W_FP64) { tex = TileShape<P_, double>::TEX; tey = TileShape<P_, double>::TEY; } \
    else { tex = TileShape<P_, float>::TEX; tey = TileShape<P_, float>::TEY; }                      \
    break;
  switch (p) { CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) default: tex = tey = 0; }
#undef CASE
}

void mfma_shape(int p, int &tex, int &tey) {
  switch (p) {
    case 6: tex = MfmaShape<6>::TEX; tey = MfmaShape<6>::TEY; break;
    case 7: tex = MfmaShape<7>::TEX; tey = MfmaShape<7>::TEY; break;
    case 8: tex = MfmaShape<8>::TEX; tey = MfmaShape<8>::TEY; break;
    default: tex = tey = 0;
  }
}

// Per-tile CSR of the items (sources or receivers) whose element touches
// nodes owned by the tile.
void build_tile_lists(const se2w_ctx *c, const std::vector<int> &elem, std::vector<int> &ptr, std::vector<int> &idx) {
  const int P = c->p;
  std::vector<std::vector<int>> lists(c->ntiles);
  for (size_t s = 0; s < elem.size(); s++) {
    const int ex = elem[s] % c->mx, ey = elem[s] / c->mx;
    const int x0 = ex * P, x1 = ex * P + P, y0 = ey * P, y1 = ey * P + P;  // node block, inclusive
    for (int ty = 0; ty < c->tiles_y; ty++) {
      const int ey0 = ty * c->tey, ey_end = std::min(ey0 + c->tey, c->my);
      const int oy0 = ey0 * P, oy1 = ey_end * P - 1 + (ey_end == c->my ? 1 : 0);
      if (y1 < oy0 || y0 > oy1) continue;
      for (int tx = 0; tx < c->tiles_x; tx++) {
        const int ex0 = tx * c->tex, ex_end = std::min(ex0 + c->tex, c->mx);
        const int ox0 = ex0 * P, ox1 = ex_end * P - 1 + (ex_end == c->mx ? 1 : 0);
        if (x1 < ox0 || x0 > ox1) continue;
        lists[tx + ty * c->tiles_x].push_back((int)s);
      }
    }
  }
  ptr.assign(c->ntiles + 1, 0);
  idx.clear();
  for (int t = 0; t < c->ntiles; t++) {
    ptr[t + 1] = ptr[t] + (int)lists[t].size();
    idx.insert(idx.end(), lists[t].begin(), lists[t].end());
  }
}

void set_tiling(se2w_ctx *c) {
  if (is_reg_variant(c->variant)) reg_shape(c->variant, c->tex, c->tey);
  else if (is_reg64_variant(c->variant)) {
    c->tex = kReg64Schedules[c->variant - SE2W_KERNEL_REG64].ex - 1;
    c->tey = kReg64Schedules[c->variant - SE2W_KERNEL_REG64].ey - 1;
  }
  else if (c->variant == SE2W_KERNEL_MFMA || c->variant == SE2W_KERNEL_LINE) mfma_shape(c->p, c->tex, c->tey);
  else tile_shape(c->p, c->mode, c->tex, c->tey);
  c->tiles_x = (c->mx + c->tex - 1) / c->tex;
  c->tiles_y = (c->my + c->tey - 1) / c->tey;
  c->ntiles = c->tiles_x * c->tiles_y;
  c->tiles_per_xcd = (c->ntiles + 7) / 8;
}

template <int MODE>
StepParams<typename RealOf<MODE>::T> make_params(se2w_ctx *c, int step, double dt, int last) {
  using R = typename RealOf<MODE>::T;
  using R2 = typename V2<R>::T;
  StepParams<R> p;
  memset(&p, 0, sizeof(p));
  p.mx = c->mx; p.my = c->my; p.nx = c->nx; p.ny = c->ny; p.pitch = c->pitch;
  p.tiles_x = c->tiles_x; p.tiles_y = c->tiles_y; p.tiles_per_xcd = c->tiles_per_xcd;
  p.tex = c->tex; p.tey = c->tey;
  p.u_in = (const R2 *)c->U[c->cur];
  p.u_out = (R2 *)c->U[c->cur ^ 1];
  p.v = (R2 *)c->V;
  p.a = (R2 *)c->A;
  p.f = (R2 *)c->F;
  p.c_in = (const R2 *)c->C[c->cur];
  p.c_out = (R2 *)c->C[c->cur ^ 1];
  p.v_next = (R2 *)c->Vn;
  p.c_next = (R2 *)c->Cn;
  p.mass = (const R *)c->mass;
  p.lambda = (const R *)c->lambda;
  p.mu = (const R *)c->mu;
  p.elem_dx = (const R *)c->edx;
  p.elem_dy = (const R *)c->edy;
  p.D = (const R *)c->D;
  p.w2 = (const R *)c->w2;
  p.tile_src_ptr = c->d_tsrc_ptr; p.tile_src_idx = c->d_tsrc_idx; p.src_elem = c->d_src_elem;
  p.src_vals = (const R *)c->d_src_vals; p.psi = (const R *)c->d_psi; p.ns = c->ns;
  p.tile_rec_ptr = c->d_trec_ptr; p.tile_rec_idx = c->d_trec_idx; p.rec_elem = c->d_rec_elem;
  p.rec_gll = c->d_rec_gll;
  p.rec = (R *)c->d_rec; p.nr = c->nr;
  p.tile_erec_ptr = c->d_terec_ptr; p.tile_erec_idx = c->d_terec_idx; p.erec_elem = c->d_erec_elem;
  p.erec = (R *)c->d_erec; p.ner = c->ner;
  p.step = step;
  // dt, 0.5*dt and 0.5*dt*dt are formed in FP64 as in se2wave.c:4726-4745
  p.dt = (R)dt;
  p.hdt = (R)(0.5 * dt);
  p.hdt2 = (R)(0.5 * dt * dt);
  p.last = last;
  return p;
}


bool reg_variant_built(int p, int variant, int schedule) {
  return is_reg_variant(variant) && reg_schedule_built(p, reg_schedule_of(variant), schedule == SE2W_SCHED_FUSED);
}

// Launch parameters of a context's step kernels.
LaunchCfg launch_cfg(const se2w_ctx *c) { return LaunchCfg{c->p, c->variant, c->schedule, 8 * c->tiles_per_xcd, c->stream}; }

int grid_for(long n) { return (int)std::min<long>((n + 255) / 256, 4096); }

// Pinned host buffer for the receiver-record download of a run of nsteps
// steps.  Grown while the run's kernels execute, so pinning overlaps GPU
// work (se2w_reserve sizes it ahead of a run instead).
int ensure_stage_for(se2w_ctx *c, int nsteps) {
  const size_t bytes = (size_t)nsteps * c->nr * rec_stride(c->p) * real_size(c->mode);
  if (bytes <= c->stage_cap) return SE2W_OK;
  if (c->h_stage) HIPCHK(hipHostFree(c->h_stage));
  c->h_stage = nullptr;
  c->stage_cap = 0;
  HIPCHK(hipHostMalloc(&c->h_stage, bytes, hipHostMallocDefault));
  c->stage_cap = bytes;
  return SE2W_OK;
}

int ensure_stage(se2w_ctx *c) { return ensure_stage_for(c, c->rec_steps); }

// Enqueues on stream s, behind a run's kernels, the download of its receiver
// record into the pinned staging buffer (grown while the kernels execute),
// so that the wait at the run's end covers it and se2w_get_receivers only
// converts host memory.
int stage_record(se2w_ctx *c, hipStream_t s) {
  if (c->nr == 0 || c->rec_steps == 0) return SE2W_OK;
  if (int ierr = ensure_stage(c)) return ierr;
  const size_t bytes = (size_t)c->rec_steps * c->nr * rec_stride(c->p) * real_size(c->mode);
  HIPCHK(hipMemcpyAsync(c->h_stage, c->d_rec, bytes, hipMemcpyDeviceToHost, s));
  c->rec_staged = true;
  return SE2W_OK;
}

// Device buffers of a run of nsteps steps: the source-time table, the
// receiver record and the element-receiver record, grown (never shrunk) to
// the run's length.
int ensure_run_buffers(se2w_ctx *c, int nsteps) {
  const size_t rs = real_size(c->mode);
  int ierr;
  if (c->ns > 0) {
    const size_t n = (size_t)nsteps * c->ns;
    if (n > c->psi_cap) {
      dev_free(c->d_psi);
      if (c->h_psi) HIPCHK(hipHostFree(c->h_psi));
      c->h_psi = nullptr;
      c->psi_cap = 0;
      if ((ierr = dev_alloc(&c->d_psi, n * rs))) return ierr;
      HIPCHK(hipHostMalloc(&c->h_psi, n * rs, hipHostMallocDefault));
      c->psi_cap = n;
    }
  }
  if (c->nr > 0) {
    const size_t n = (size_t)nsteps * c->nr * rec_stride(c->p);
    if (n > c->rec_cap) {
      dev_free(c->d_rec);
      c->rec_cap = 0;
      if ((ierr = dev_alloc(&c->d_rec, n * rs))) return ierr;
      c->rec_cap = n;
    }
  }
  if (c->ner > 0) {
    const size_t n = (size_t)nsteps * c->ner * c->npe * 4;
    if (n > c->erec_cap) {
      dev_free(c->d_erec);
      c->erec_cap = 0;
      if ((ierr = dev_alloc(&c->d_erec, n * rs))) return ierr;
      c->erec_cap = n;
    }
  }
  return SE2W_OK;
}

// Events of one call, destroyed on every return path.
struct ScopedEvents {
  std::vector<hipEvent_t> ev;
  ~ScopedEvents() {
    for (hipEvent_t e : ev)
      if (e) hipEventDestroy(e);
  }
  int create(size_t n) {
    ev.assign(n, nullptr);
    for (auto &e : ev) HIPCHK(hipEventCreate(&e));
    return SE2W_OK;
  }
  hipEvent_t &operator[](size_t i) { return ev[i]; }
};

// Opens a run on stream s: the step kernels start from the predicted state
// (u*, v*) of the run's first step.  The previous run's last step left it in
// U[cur ^ 1], Vn (and Cn) when that run had the same dt and the state has not
// been written since; otherwise the predictor kernel forms it from the plain
// state (U[cur], V, A, C).  Either way U[cur], V and C then hold it.
template <int MODE>
void open_run(se2w_ctx *c, double dt, hipStream_t s) {
  if (c->pred_ok && c->pred_dt == dt) {
    std::swap(c->V, c->Vn);
    if (c->mode == SE2W_FP32_KAHAN) {
      std::swap(c->C[0], c->Cn);
      c->C[1] = c->C[0];
    }
  } else {
    StepParams<typename RealOf<MODE>::T> prm = make_params<MODE>(c, 0, dt, 0);
    launch_predictor<MODE>(prm, (long)c->pitch * c->ny, s);
  }
  c->cur ^= 1;
  c->pred_ok = false;  // until the run's last step has written the next one
}

// Closes a run whose last step read U[cur]: that buffer keeps the plain
// displacement, and the next run's predicted state is in place.
void close_run(se2w_ctx *c, double dt) {
  c->pred_ok = true;
  c->pred_dt = dt;
}

template <int MODE>
int run_mode(se2w_ctx *c, int nsteps, double dt) {
  using R = typename RealOf<MODE>::T;
  // Profiling: the fused schedule is timed over the whole sequence of step
  // launches (one event pair, so no event sits between two launches); the
  // split schedule brackets each of its two kernels.
  const bool fused = c->schedule == SE2W_SCHED_FUSED;
  ScopedEvents evs;
  if (c->profiling)
    if (int ierr = evs.create(fused ? 2 : (size_t)nsteps * 3)) return ierr;
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  open_run<MODE>(c, dt, c->stream);
  HIPCHK(hipGetLastError());
  if (c->profiling && fused) HIPCHK(hipEventRecord(evs[0], c->stream));
  for (int k = 0; k < nsteps; k++) {
    StepParams<R> prm = make_params<MODE>(c, k, dt, k == nsteps - 1);
    launch_step<MODE>(launch_cfg(c), prm, (c->profiling && !fused) ? &evs[(size_t)k * 3] : nullptr);
    if (k < nsteps - 1) c->cur ^= 1;
  }
  if (c->profiling && fused) HIPCHK(hipEventRecord(evs[1], c->stream));
  HIPCHK(hipGetLastError());
  close_run(c, dt);
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  if (int ierr = stage_record(c, c->stream)) return ierr;
  HIPCHK(hipStreamSynchronize(c->stream));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_ms = ms;
  c->kern_ms[0] = c->kern_ms[1] = 0.0;
  if (c->profiling) {
    double s0 = 0.0, s1 = 0.0;
    if (fused) {
      float t0 = 0.f;
      hipEventElapsedTime(&t0, evs[0], evs[1]);
      s0 = t0;
    } else {
      for (int k = 0; k < nsteps; k++) {
        float t0 = 0.f, t1 = 0.f;
        hipEventElapsedTime(&t0, evs[(size_t)k * 3], evs[(size_t)k * 3 + 1]);
        hipEventElapsedTime(&t1, evs[(size_t)k * 3 + 1], evs[(size_t)k * 3 + 2]);
        s0 += t0;
        s1 += t1;
      }
    }
    c->kern_ms[0] = s0 / nsteps;
    c->kern_ms[1] = s1 / nsteps;
  }
  return SE2W_OK;
}

}  // namespace

namespace se2w {
// Error message of the library's other translation units (se2w_last_error).
int set_error(int code, const char *msg) {
  snprintf(g_err, sizeof(g_err), "%s", msg);
  return code;
}

static int ens_dispatch_mode(int p, int mode, int variant, const void *args, int grid, hipStream_t s, int *bpc,
                             int *thr) {
  switch (mode) {
    case SE2W_FP64: return ens_dispatch<MODE_F64>(p, variant, args, grid, s, bpc, thr);
    case SE2W_FP32: return ens_dispatch<MODE_F32>(p, variant, args, grid, s, bpc, thr);
    case SE2W_FP32_KAHAN: return ens_dispatch<MODE_F32K>(p, variant, args, grid, s, bpc, thr);
    default: return -1;
  }
}

int ens_launch(int p, int mode, int variant, const void *args, int grid, hipStream_t s) {
  if (!args || grid < 8 || grid % 8) return -1;
  int bpc = 0, thr = 0;
  return ens_dispatch_mode(p, mode, variant, args, grid, s, &bpc, &thr);
}

int ens_occupancy(int p, int mode, int variant, int *blocks_per_cu, int *threads) {
  return ens_dispatch_mode(p, mode, variant, nullptr, 0, nullptr, blocks_per_cu, threads);
}
}  // namespace se2w

extern "C" {

int se2w_version(void) { return 1; }

const char *se2w_last_error(void) { return g_err; }

const char *se2w_strerror(int code) {
  switch (code) {
    case SE2W_OK: return "success";
    case SE2W_ERR_ARG: return "invalid argument";
    case SE2W_ERR_MEM: return "out of device memory";
    case SE2W_ERR_HIP: return "HIP runtime error";
    case SE2W_ERR_STATE: return "context not fully set up";
    case SE2W_ERR_UNSUPPORTED: return "unsupported configuration";
    default: return "unknown error";
  }
}

int se2w_create(const se2w_desc *desc, se2w_ctx **out) {
  if (!desc || !out) return fail(SE2W_ERR_ARG, "null argument");
  *out = nullptr;
  if (desc->mx < 1 || desc->my < 1) return fail(SE2W_ERR_ARG, "mx, my must be >= 1");
  if (desc->p < 1 || desc->p > kMaxP) return fail(SE2W_ERR_UNSUPPORTED, "p = %d outside 1..%d", desc->p, kMaxP);
  if (desc->precision < SE2W_FP64 || desc->precision > SE2W_FP32_KAHAN)
    return fail(SE2W_ERR_ARG, "unknown precision %d", desc->precision);
  if (desc->schedule != SE2W_SCHED_FUSED && desc->schedule != SE2W_SCHED_SPLIT)
    return fail(SE2W_ERR_ARG, "unknown schedule %d", desc->schedule);
  const long nx = (long)desc->p * desc->mx + 1, ny = (long)desc->p * desc->my + 1;
  if (nx * ny * 2 > 0x7fffffffL) return fail(SE2W_ERR_UNSUPPORTED, "mesh too large for 32-bit dof indices");
  se2w_ctx *c = new se2w_ctx();
  c->mx = desc->mx; c->my = desc->my; c->p = desc->p; c->nb = desc->p + 1; c->npe = c->nb * c->nb;
  c->nx = (int)nx; c->ny = (int)ny;
  c->pitch = (int)((nx + 31) / 32 * 32);
  c->mode = desc->precision;
  c->schedule = desc->schedule;
  c->device = desc->device;
  c->variant = reg_available(c->p, c->mode) && reg_variant_built(c->p, SE2W_KERNEL_REG, c->schedule)
                   ? SE2W_KERNEL_REG
               : reg64_variant_built(c->p, c->mode, SE2W_KERNEL_REG64, c->schedule) ? SE2W_KERNEL_REG64
               : reg64_variant_built(c->p, c->mode, SE2W_KERNEL_REG64 + kReg64FirstP8, c->schedule)
                   ? SE2W_KERNEL_REG64 + kReg64FirstP8
                   : SE2W_KERNEL_LDS;
  set_tiling(c);
  int ierr = SE2W_OK;
  hipError_t he = hipSetDevice(c->device);
  if (he != hipSuccess) { ierr = fail(SE2W_ERR_HIP, "hipSetDevice: %s", hipGetErrorString(he)); goto bad; }
  {  // the device's scratch address, looked up once per device at set-up
    void *sp = nullptr;
    if ((ierr = scratch_address(c->device, &sp))) goto bad;
  }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    ierr = fail(SE2W_ERR_HIP, "stream/event creation failed");
    goto bad;
  }
  {
    const size_t rs = real_size(c->mode);
    const size_t field = (size_t)c->pitch * c->ny * 2 * rs;
    if ((ierr = dev_alloc(&c->U[0], field)) || (ierr = dev_alloc(&c->U[1], field)) || (ierr = dev_alloc(&c->V, field)) ||
        (ierr = dev_alloc(&c->A, field)) || (ierr = dev_alloc(&c->Vn, field)))
      goto bad;
    if (c->mode == SE2W_FP32_KAHAN) {
      // One compensation buffer, updated in place (C[1] aliases C[0]): a
      // node's compensation is read and written only by the thread that
      // updates the node, unlike u, whose halo other tiles read.
      if ((ierr = dev_alloc(&c->C[0], field)) || (ierr = dev_alloc(&c->Cn, field))) goto bad;
      c->C[1] = c->C[0];
      hipMemsetAsync(c->C[0], 0, field, c->stream);
      hipMemsetAsync(c->Cn, 0, field, c->stream);
    }
    if (c->schedule == SE2W_SCHED_SPLIT) {
      if ((ierr = dev_alloc(&c->F, field))) goto bad;
      hipMemsetAsync(c->F, 0, field, c->stream);
    }
    hipMemsetAsync(c->U[0], 0, field, c->stream);
    hipMemsetAsync(c->U[1], 0, field, c->stream);
    hipMemsetAsync(c->V, 0, field, c->stream);
    hipMemsetAsync(c->A, 0, field, c->stream);
    hipMemsetAsync(c->Vn, 0, field, c->stream);
    if ((ierr = dev_alloc((void **)&c->d_partial, sizeof(double) * 6 * 1024))) goto bad;
  }
  {  // empty source / receiver lists
    std::vector<int> ptr(c->ntiles + 1, 0);
    if ((ierr = upload<int>((void **)&c->d_tsrc_ptr, ptr.data(), ptr.size(), c->stream)) ||
        (ierr = upload<int>((void **)&c->d_trec_ptr, ptr.data(), ptr.size(), c->stream)))
      goto bad;
  }
  he = hipStreamSynchronize(c->stream);
  if (he != hipSuccess) { ierr = fail(SE2W_ERR_HIP, "init: %s", hipGetErrorString(he)); goto bad; }
  *out = c;
  return SE2W_OK;
bad:
  se2w_destroy(c);
  return ierr;
}

int se2w_destroy(se2w_ctx *c) {
  if (!c) return SE2W_OK;
  if (c->stream) hipStreamSynchronize(c->stream);
  if (c->C[1] == c->C[0]) c->C[1] = nullptr;  // aliased compensation buffer
  void **bufs[] = {&c->U[0], &c->U[1], &c->C[0], &c->C[1], &c->V, &c->A, &c->Vn, &c->Cn, &c->F, &c->Utmp, &c->mass, &c->lambda,
                   &c->mu, &c->edx, &c->edy, &c->D, &c->w2, &c->d_src_vals, &c->d_psi, &c->d_rec, &c->d_erec};
  for (void **b : bufs) dev_free(*b);
  int **ibufs[] = {&c->d_src_elem, &c->d_tsrc_ptr, &c->d_tsrc_idx, &c->d_rec_elem, &c->d_rec_gll,
                    &c->d_trec_ptr, &c->d_trec_idx, &c->d_erec_elem, &c->d_terec_ptr, &c->d_terec_idx};
  for (int **b : ibufs) {
    void *p = *b;
    dev_free(p);
    *b = nullptr;
  }
  if (c->copy_stream) hipStreamSynchronize(c->copy_stream);
  dev_free(c->snap_dev);
  if (c->snap_host) hipHostFree(c->snap_host);
  if (c->snap_ready) hipEventDestroy(c->snap_ready);
  if (c->snap_done) hipEventDestroy(c->snap_done);
  if (c->copy_stream) hipStreamDestroy(c->copy_stream);
  if (c->h_stage) hipHostFree(c->h_stage);
  if (c->h_psi) hipHostFree(c->h_psi);
  if (c->d_partial) hipFree(c->d_partial);
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return SE2W_OK;
}

int se2w_set_tables(se2w_ctx *c, const double *D, const double *w1d, const double *elem_dx, const double *elem_dy) {
  if (!c || !D || !w1d || !elem_dx || !elem_dy) return fail(SE2W_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(c->device));
  const int nb = c->nb;
  for (int e = 0; e < c->mx; e++)
    if (!(elem_dx[e] > 0.0)) return fail(SE2W_ERR_ARG, "elem_dx[%d] = %g must be > 0", e, elem_dx[e]);
  for (int e = 0; e < c->my; e++)
    if (!(elem_dy[e] > 0.0)) return fail(SE2W_ERR_ARG, "elem_dy[%d] = %g must be > 0", e, elem_dy[e]);
  std::vector<double> w2((size_t)nb * nb);
  for (int j = 0; j < nb; j++)
    for (int i = 0; i < nb; i++) w2[i + j * nb] = w1d[i] * w1d[j];  // se2wave.c:860-865
  int ierr;
  if ((ierr = upload_real(c, &c->D, D, (size_t)nb * nb)) || (ierr = upload_real(c, &c->w2, w2.data(), w2.size())) ||
      (ierr = upload_real(c, &c->edx, elem_dx, c->mx)) || (ierr = upload_real(c, &c->edy, elem_dy, c->my)))
    return ierr;
  c->hD.assign(D, D + (size_t)nb * nb);
  c->hdx.assign(elem_dx, elem_dx + c->mx);
  c->hdy.assign(elem_dy, elem_dy + c->my);
  c->have_tables = true;
  return SE2W_OK;
}

int se2w_set_cells(se2w_ctx *c, const double *lambda, const double *mu) {
  if (!c || !lambda || !mu) return fail(SE2W_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(c->device));
  const size_t ne = (size_t)c->mx * c->my;
  for (size_t e = 0; e < ne; e++)
    if (!std::isfinite(lambda[e]) || !std::isfinite(mu[e])) return fail(SE2W_ERR_ARG, "non-finite material in cell %zu", e);
  int ierr;
  if ((ierr = upload_real(c, &c->lambda, lambda, ne)) || (ierr = upload_real(c, &c->mu, mu, ne))) return ierr;
  c->have_cells = true;
  return SE2W_OK;
}

int se2w_set_mass(se2w_ctx *c, const double *mass) {
  if (!c || !mass) return fail(SE2W_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(c->device));
  const long n = (long)c->nx * c->ny;
  for (long i = 0; i < n; i++)
    if (!(mass[i] > 0.0)) return fail(SE2W_ERR_ARG, "mass[%ld] = %g must be > 0", i, mass[i]);
  const size_t rs = real_size(c->mode);
  int ierr = dev_alloc(&c->mass, (size_t)c->pitch * c->ny * rs);
  if (ierr) return ierr;
  ScopedDev scratch;
  if ((ierr = dev_alloc(&scratch.p, sizeof(double) * n))) return ierr;
  double *tmp = (double *)scratch.p;
  HIPCHK(hipMemcpyAsync(tmp, mass, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  if (c->mode == SE2W_FP64)
    hipLaunchKernelGGL(k_scalar_from_dmda<double>, dim3(grid_for(n)), dim3(256), 0, c->stream, tmp,
                       (double *)c->mass, c->nx, c->ny, c->pitch, 0);
  else  // 1/M is formed in FP64 and rounded once
    hipLaunchKernelGGL(k_scalar_from_dmda<float>, dim3(grid_for(n)), dim3(256), 0, c->stream, tmp, (float *)c->mass,
                       c->nx, c->ny, c->pitch, 1);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  c->have_mass = true;
  return SE2W_OK;
}

int se2w_set_sources(se2w_ctx *c, int32_t ns, const int32_t *elem, const double *vals) {
  if (!c || ns < 0 || (ns > 0 && (!elem || !vals))) return fail(SE2W_ERR_ARG, "bad source arguments");
  HIPCHK(hipSetDevice(c->device));
  const int ne = c->mx * c->my;
  std::vector<int> el(elem, elem + ns);
  for (int s = 0; s < ns; s++)
    if (el[s] < 0 || el[s] >= ne) return fail(SE2W_ERR_ARG, "source %d: element %d out of range", s, el[s]);
  std::vector<int> ptr, idx;
  build_tile_lists(c, el, ptr, idx);
  c->h_src_elem = el;
  int ierr;
  if ((ierr = upload<int>((void **)&c->d_src_elem, el.data(), el.size(), c->stream)) ||
      (ierr = upload<int>((void **)&c->d_tsrc_ptr, ptr.data(), ptr.size(), c->stream)) ||
      (ierr = upload<int>((void **)&c->d_tsrc_idx, idx.data(), idx.size(), c->stream)) ||
      (ierr = upload_real(c, &c->d_src_vals, vals, (size_t)ns * 2 * c->npe)))
    return ierr;
  c->ns = ns;
  return SE2W_OK;
}

int se2w_set_receivers(se2w_ctx *c, int32_t nr, const int32_t *elem, const int32_t *gll) {
  if (!c || nr < 0 || (nr > 0 && (!elem || !gll))) return fail(SE2W_ERR_ARG, "bad receiver arguments");
  HIPCHK(hipSetDevice(c->device));
  const int ne = c->mx * c->my;
  for (int r = 0; r < nr; r++) {
    if (elem[r] < 0 || elem[r] >= ne) return fail(SE2W_ERR_ARG, "receiver %d: element %d out of range", r, elem[r]);
    if (gll[r] < 0 || gll[r] >= c->npe) return fail(SE2W_ERR_ARG, "receiver %d: gll %d out of range", r, gll[r]);
  }
  c->h_rec_elem.assign(elem, elem + nr);
  c->h_rec_gll.assign(gll, gll + nr);
  std::vector<int> ptr, idx;
  build_tile_lists(c, c->h_rec_elem, ptr, idx);
  int ierr;
  if ((ierr = upload<int>((void **)&c->d_rec_elem, c->h_rec_elem.data(), nr, c->stream)) ||
      (ierr = upload<int>((void **)&c->d_rec_gll, c->h_rec_gll.data(), nr, c->stream)) ||
      (ierr = upload<int>((void **)&c->d_trec_ptr, ptr.data(), ptr.size(), c->stream)) ||
      (ierr = upload<int>((void **)&c->d_trec_idx, idx.data(), idx.size(), c->stream)))
    return ierr;
  c->nr = nr;
  c->rec_steps = 0;
  c->trace.clear();
  return SE2W_OK;
}

int se2w_set_element_receivers(se2w_ctx *c, int32_t n, const int32_t *elem) {
  if (!c || n < 0 || (n > 0 && !elem)) return fail(SE2W_ERR_ARG, "bad element-receiver arguments");
  HIPCHK(hipSetDevice(c->device));
  const int ne = c->mx * c->my;
  for (int r = 0; r < n; r++)
    if (elem[r] < 0 || elem[r] >= ne) return fail(SE2W_ERR_ARG, "element receiver %d: element %d out of range", r, elem[r]);
  c->h_erec_elem.assign(elem, elem + n);
  std::vector<int> ptr, idx;
  build_tile_lists(c, c->h_erec_elem, ptr, idx);
  int ierr;
  if ((ierr = upload<int>((void **)&c->d_erec_elem, c->h_erec_elem.data(), n, c->stream)) ||
      (ierr = upload<int>((void **)&c->d_terec_ptr, ptr.data(), ptr.size(), c->stream)) ||
      (ierr = upload<int>((void **)&c->d_terec_idx, idx.data(), idx.size(), c->stream)))
    return ierr;
  c->ner = n;
  c->erec_steps = 0;
  return SE2W_OK;
}

int se2w_get_element_receivers(se2w_ctx *c, const double *weights, double *trace) {
#pragma clang fp contract(off)
  if (!c || (c->ner > 0 && (!weights || !trace))) return fail(SE2W_ERR_ARG, "null argument");
  if (c->ner == 0 || c->erec_steps == 0) return SE2W_OK;
  HIPCHK(hipSetDevice(c->device));
  const int npe = c->npe, ner = c->ner;
  const size_t n = (size_t)c->erec_steps * ner * npe * 4;
  std::vector<double> rec(n);
  if (c->mode == SE2W_FP64) {
    HIPCHK(hipMemcpyAsync(rec.data(), c->d_erec, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  } else {
    std::vector<float> tmp(n);
    HIPCHK(hipMemcpyAsync(tmp.data(), c->d_erec, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < n; i++) rec[i] = (double)tmp[i];
  }
  // sum over the element's nodes in ascending local index (se2wave.c:2563-2571)
  for (int k = 0; k < c->erec_steps; k++)
    for (int r = 0; r < ner; r++) {
      const double *R = &rec[((size_t)k * ner + r) * npe * 4];
      const double *w = weights + (size_t)r * npe;
      double s[4] = {0.0, 0.0, 0.0, 0.0};
      for (int q = 0; q < npe; q++)
        for (int f = 0; f < 4; f++) s[f] = s[f] + w[q] * R[q * 4 + f];
      for (int f = 0; f < 4; f++) trace[((size_t)k * ner + r) * 4 + f] = s[f];
    }
  return SE2W_OK;
}

int se2w_set_state(se2w_ctx *c, const double *u, const double *v, const double *a) {
  if (!c) return fail(SE2W_ERR_ARG, "null context");
  HIPCHK(hipSetDevice(c->device));
  const long n = (long)c->nx * c->ny;
  ScopedDev scratch;
  if (int ierr = dev_alloc(&scratch.p, sizeof(double) * 2 * n)) return ierr;
  double *tmp = (double *)scratch.p;
  const double *src[3] = {u, v, a};
  void *dst[3] = {c->U[c->cur], c->V, c->A};
  for (int f = 0; f < 3; f++) {
    if (!src[f]) continue;
    HIPCHK(hipMemcpyAsync(tmp, src[f], sizeof(double) * 2 * n, hipMemcpyHostToDevice, c->stream));
    if (c->mode == SE2W_FP64)
      hipLaunchKernelGGL(k_from_dmda<double>, dim3(grid_for(n)), dim3(256), 0, c->stream, tmp, (double2 *)dst[f],
                         c->nx, c->ny, c->pitch);
    else
      hipLaunchKernelGGL(k_from_dmda<float>, dim3(grid_for(n)), dim3(256), 0, c->stream, tmp, (float2 *)dst[f], c->nx,
                         c->ny, c->pitch);
    HIPCHK(hipGetLastError());
  }
  if (u && c->mode == SE2W_FP32_KAHAN)
    HIPCHK(hipMemsetAsync(c->C[c->cur], 0, (size_t)c->pitch * c->ny * 2 * sizeof(float), c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->pred_ok = false;
  return SE2W_OK;
}

// Enqueues on the context's stream the conversion of state field f (0 = u,
// with its compensation in FP32+Kahan; 1 = v; 2 = a) to interleaved FP64 in
// DMDA order at the device address dst.
static int enqueue_field_fp64(se2w_ctx *c, int f, double *dst) {
  const long n = (long)c->nx * c->ny;
  void *src = f == 0 ? c->U[c->cur] : f == 1 ? c->V : c->A;
  if (c->mode == SE2W_FP64)
    hipLaunchKernelGGL(k_to_dmda<double>, dim3(grid_for(n)), dim3(256), 0, c->stream, (const double2 *)src,
                       (const double2 *)nullptr, dst, c->nx, c->ny, c->pitch);
  else
    hipLaunchKernelGGL(k_to_dmda<float>, dim3(grid_for(n)), dim3(256), 0, c->stream, (const float2 *)src,
                       (const float2 *)((f == 0 && c->mode == SE2W_FP32_KAHAN) ? c->C[c->cur] : nullptr), dst,
                       c->nx, c->ny, c->pitch);
  HIPCHK(hipGetLastError());
  return SE2W_OK;
}

int se2w_get_state(se2w_ctx *c, double *u, double *v, double *a) {
  if (!c) return fail(SE2W_ERR_ARG, "null context");
  HIPCHK(hipSetDevice(c->device));
  const size_t field = sizeof(double) * 2 * (long)c->nx * c->ny;
  ScopedDev scratch;
  if (int ierr = dev_alloc(&scratch.p, field)) return ierr;
  double *tmp = (double *)scratch.p;
  double *dst[3] = {u, v, a};
  for (int f = 0; f < 3; f++) {
    if (!dst[f]) continue;
    if (int ierr = enqueue_field_fp64(c, f, tmp)) return ierr;
    HIPCHK(hipMemcpyAsync(dst[f], tmp, field, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return SE2W_OK;
}

int se2w_snapshot_begin(se2w_ctx *c) {
  if (!c) return fail(SE2W_ERR_ARG, "null context");
  HIPCHK(hipSetDevice(c->device));
  if (c->snap_pending) {  // the staging buffers are still being copied
    HIPCHK(hipEventSynchronize(c->snap_done));
    c->snap_pending = false;
  }
  const long n = (long)c->nx * c->ny;
  const size_t field = sizeof(double) * 2 * n;
  if (int ierr = dev_alloc(&c->snap_dev, 2 * field)) return ierr;
  if (!c->snap_host) {
    void *h = nullptr;
    if (hipHostMalloc(&h, 2 * field, hipHostMallocDefault) != hipSuccess)
      return fail(SE2W_ERR_MEM, "hipHostMalloc(%zu) failed", 2 * field);
    c->snap_host = (double *)h;
  }
  if (!c->copy_stream) HIPCHK(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
  if (!c->snap_ready) HIPCHK(hipEventCreateWithFlags(&c->snap_ready, hipEventDisableTiming));
  if (!c->snap_done) HIPCHK(hipEventCreateWithFlags(&c->snap_done, hipEventDisableTiming));
  for (int f = 0; f < 2; f++)
    if (int ierr = enqueue_field_fp64(c, f, (double *)c->snap_dev + f * 2 * n)) return ierr;
  HIPCHK(hipEventRecord(c->snap_ready, c->stream));
  HIPCHK(hipStreamWaitEvent(c->copy_stream, c->snap_ready, 0));
  HIPCHK(hipMemcpyAsync(c->snap_host, c->snap_dev, 2 * field, hipMemcpyDeviceToHost, c->copy_stream));
  HIPCHK(hipEventRecord(c->snap_done, c->copy_stream));
  c->snap_pending = true;
  return SE2W_OK;
}

int se2w_snapshot_wait(se2w_ctx *c, const double **u, const double **v) {
  if (!c || !u || !v) return fail(SE2W_ERR_ARG, "null argument");
  if (!c->snap_host) return fail(SE2W_ERR_STATE, "no snapshot taken (se2w_snapshot_begin)");
  HIPCHK(hipSetDevice(c->device));
  if (c->snap_pending) {
    HIPCHK(hipEventSynchronize(c->snap_done));
    c->snap_pending = false;
  }
  *u = c->snap_host;
  *v = c->snap_host + 2 * (long)c->nx * c->ny;
  return SE2W_OK;
}

}  // extern "C"

namespace {

// Checks a run's arguments and stages its inputs: the source-time table and
// the receiver record buffer.
// Checks a run's arguments, sizes its buffers and enqueues its source-time
// table on stream s, the stream its kernels will run on.
int prepare_run(se2w_ctx *c, int32_t nsteps, double dt, const double *psi, hipStream_t s) {
  if (!c) return fail(SE2W_ERR_ARG, "null context");
  if (nsteps < 1) return fail(SE2W_ERR_ARG, "nsteps must be >= 1");
  if (!(dt > 0.0) || !std::isfinite(dt)) return fail(SE2W_ERR_ARG, "dt must be positive and finite");
  if (c->ns > 0 && !psi) return fail(SE2W_ERR_ARG, "psi table required when sources are set");
  if (!c->have_tables || !c->have_cells || !c->have_mass)
    return fail(SE2W_ERR_STATE, "set tables, cells and mass before se2w_run");
  HIPCHK(hipSetDevice(c->device));
  if (int ierr = ensure_run_buffers(c, nsteps)) return ierr;
  // Source-time table for this run, through the pinned buffer: the copy is
  // ordered before the run's kernels on the stream.  A run waits for its
  // stream at its end; the wait here covers one that returned early on an
  // error with the previous copy still queued (on an idle stream it costs a
  // few microseconds).
  if (c->ns > 0) {
    const size_t n = (size_t)nsteps * c->ns;
    HIPCHK(hipStreamSynchronize(s));
    if (c->mode == SE2W_FP64) {
      memcpy(c->h_psi, psi, n * sizeof(double));
    } else {
      float *h = (float *)c->h_psi;
      for (size_t i = 0; i < n; i++) h[i] = (float)psi[i];
    }
    HIPCHK(hipMemcpyAsync(c->d_psi, c->h_psi, n * real_size(c->mode), hipMemcpyHostToDevice, s));
  }
  c->rec_staged = false;
  c->rec_steps = nsteps;
  c->erec_steps = nsteps;
  c->trace.clear();
  return SE2W_OK;
}

}  // namespace

extern "C" {

int se2w_reserve(se2w_ctx *c, int32_t nsteps) {
  if (!c) return fail(SE2W_ERR_ARG, "null context");
  if (nsteps < 1) return fail(SE2W_ERR_ARG, "nsteps must be >= 1");
  HIPCHK(hipSetDevice(c->device));
  if (int ierr = ensure_run_buffers(c, nsteps)) return ierr;
  if (int ierr = ensure_stage_for(c, nsteps)) return ierr;
  // One record download of the reserved length: the process's first
  // device-to-host copy of this size sets up the runtime's copy path
  // (measured at ~6.7 ms once per process), which would otherwise land in
  // the first run's se2w_get_receivers.
  if (c->nr > 0) {
    const size_t bytes = (size_t)nsteps * c->nr * rec_stride(c->p) * real_size(c->mode);
    HIPCHK(hipMemcpyAsync(c->h_stage, c->d_rec, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  return SE2W_OK;
}

int se2w_get_counts(se2w_ctx *c, int32_t counts[4]) {
  if (!c || !counts) return fail(SE2W_ERR_ARG, "null argument");
  counts[0] = c->ns;
  counts[1] = c->nr;
  counts[2] = c->ner;
  counts[3] = c->rec_steps;
  return SE2W_OK;
}

int se2w_get_reserved(se2w_ctx *c, int64_t caps[4]) {
  if (!c || !caps) return fail(SE2W_ERR_ARG, "null argument");
  caps[0] = (int64_t)c->psi_cap;
  caps[1] = (int64_t)c->rec_cap;
  caps[2] = (int64_t)c->erec_cap;
  caps[3] = (int64_t)c->stage_cap;
  return SE2W_OK;
}

int se2w_run(se2w_ctx *c, int32_t nsteps, double dt, const double *psi) {
  // SE2W_PHASE_TIMES=1 prints, per run, the host time of the argument checks
  // and table upload, of the launches and the wait, and the GPU time between
  // the run's first and last events (stderr).
  static const bool trace = getenv("SE2W_PHASE_TIMES") != nullptr;
  const auto t0 = std::chrono::steady_clock::now();
  int ierr = prepare_run(c, nsteps, dt, psi, c->stream);
  if (ierr) return ierr;
  const auto t1 = std::chrono::steady_clock::now();
  if (c->mode == SE2W_FP64) ierr = run_mode<MODE_F64>(c, nsteps, dt);
  else if (c->mode == SE2W_FP32) ierr = run_mode<MODE_F32>(c, nsteps, dt);
  else ierr = run_mode<MODE_F32K>(c, nsteps, dt);
  if (trace) {
    const auto t2 = std::chrono::steady_clock::now();
    const double a = std::chrono::duration<double, std::milli>(t1 - t0).count();
    const double b = std::chrono::duration<double, std::milli>(t2 - t1).count();
    fprintf(stderr, "[se2w_run] %d steps: prepare %.3f ms, launch+wait %.3f ms, GPU events %.3f ms\n", nsteps, a, b,
            c->last_ms);
  }
  return ierr;
}

int se2w_get_receivers(se2w_ctx *c, double *trace) {
#pragma clang fp contract(off)
  if (!c || (!trace && c->nr > 0)) return fail(SE2W_ERR_ARG, "null argument");
  if (c->nr == 0 || c->rec_steps == 0) return SE2W_OK;
  HIPCHK(hipSetDevice(c->device));
  const int nb = c->nb, nr = c->nr, rs = rec_stride(c->p);
  const size_t n = (size_t)c->rec_steps * nr * rs;
  const size_t bytes = n * real_size(c->mode);
  if (!c->rec_staged) {
    if (int ierr = ensure_stage(c)) return ierr;
    HIPCHK(hipMemcpyAsync(c->h_stage, c->d_rec, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  const double *rd = (const double *)c->h_stage;
  const float *rf = (const float *)c->h_stage;
  const bool f64 = c->mode == SE2W_FP64;
  for (int k = 0; k < c->rec_steps; k++) {
    for (int r = 0; r < nr; r++) {
      const size_t base = ((size_t)k * nr + r) * rs;
      double R[6 + 2 * (kMaxP + 1)];
      for (int q = 0; q < rs; q++) R[q] = f64 ? rd[base + q] : (double)rf[base + q];
      double *o = &trace[((size_t)k * nr + r) * 7];
      const int g = c->h_rec_gll[r], gi = g % nb, gj = g / nb;
      for (int q = 0; q < 6; q++) o[q] = R[q];
      // curl(v) at the receiver's GLL node (se2wave.c:2922-2959): dN_dx on
      // row gj, dN_dy on column gi, each sum in ascending node order.
      const int ex = c->h_rec_elem[r] % c->mx, ey = c->h_rec_elem[r] / c->mx;
      const double J00 = 0.5 * c->hdx[ex], J11 = 0.5 * c->hdy[ey];
      const double iJ00 = 1.0 / J00, iJ11 = 1.0 / J11;
      double dvxdy = 0.0, dvydx = 0.0;
      for (int j = 0; j < nb; j++) dvxdy = dvxdy + (iJ11 * c->hD[gj * nb + j]) * R[6 + j];
      for (int i = 0; i < nb; i++) dvydx = dvydx + (iJ00 * c->hD[gi * nb + i]) * R[6 + nb + i];
      o[6] = dvydx - dvxdy;
    }
  }
  return SE2W_OK;
}

}  // extern "C"

namespace {

// The element kernel that f = K(u) runs on: the context's variant when that
// variant has an element-only form (the LDS tile kernel, or the register
// kernel's first schedule), in FP64 the FP64 register kernel's element-only
// form where it is built, else the LDS tile kernel.  The fused-only
// variants (the other register schedules, the 512-thread tile, matrix-core
// and line kernels) tile the mesh differently from the element-only kernel
// they would fall back to, so the call switches the tiling with the kernel.
int element_only_variant(const se2w_ctx *c) {
  if (is_reg_variant(c->variant) && reg_available(c->p, c->mode) &&
      reg_variant_built(c->p, c->variant, SE2W_SCHED_SPLIT))
    return c->variant;
  if (reg64_variant_built(c->p, c->mode, SE2W_KERNEL_REG64, SE2W_SCHED_SPLIT)) return SE2W_KERNEL_REG64;
  return SE2W_KERNEL_LDS;
}

// f = K(u) with u, f in the padded device layout of the context (Utmp -> F),
// the launch tiled for the element-only kernel it runs.  The context's tiling
// is restored before returning; the per-tile source and receiver lists are
// not read by the element-only kernels.
int enqueue_forces(se2w_ctx *c, void *u_dev, void *f_dev) {
  const int v0 = c->variant;
  c->variant = element_only_variant(c);
  set_tiling(c);
  if (c->mode == SE2W_FP64) {
    StepParams<double> prm = make_params<MODE_F64>(c, 0, 1.0, 0);
    prm.u_in = (const double2 *)u_dev;
    prm.f = (double2 *)f_dev;
    launch_element_only<MODE_F64>(launch_cfg(c), prm);
  } else {
    StepParams<float> prm = make_params<MODE_F32>(c, 0, 1.0, 0);
    prm.u_in = (const float2 *)u_dev;
    prm.f = (float2 *)f_dev;
    launch_element_only<MODE_F32>(launch_cfg(c), prm);
  }
  c->variant = v0;
  set_tiling(c);
  HIPCHK(hipGetLastError());
  return SE2W_OK;
}

}  // namespace

extern "C" {

int se2w_assemble_forces(se2w_ctx *c, const double *u, double *f) {
  if (!c || !u || !f) return fail(SE2W_ERR_ARG, "null argument");
  if (!c->have_tables || !c->have_cells) return fail(SE2W_ERR_STATE, "set tables and cells first");
  HIPCHK(hipSetDevice(c->device));
  const size_t rs = real_size(c->mode);
  const size_t field = (size_t)c->pitch * c->ny * 2 * rs;
  int ierr;
  if ((ierr = dev_alloc(&c->Utmp, field))) return ierr;
  if ((ierr = dev_alloc(&c->F, field))) return ierr;
  const long n = (long)c->nx * c->ny;
  ScopedDev scratch;
  if ((ierr = dev_alloc(&scratch.p, sizeof(double) * 2 * n))) return ierr;
  double *tmp = (double *)scratch.p;
  HIPCHK(hipMemcpyAsync(tmp, u, sizeof(double) * 2 * n, hipMemcpyHostToDevice, c->stream));
  // every owned node is written by the element kernel; the clear keeps a
  // missed node from returning the previous call's value
  HIPCHK(hipMemsetAsync(c->F, 0, field, c->stream));
  if (c->mode == SE2W_FP64)
    hipLaunchKernelGGL(k_from_dmda<double>, dim3(grid_for(n)), dim3(256), 0, c->stream, tmp, (double2 *)c->Utmp,
                       c->nx, c->ny, c->pitch);
  else
    hipLaunchKernelGGL(k_from_dmda<float>, dim3(grid_for(n)), dim3(256), 0, c->stream, tmp, (float2 *)c->Utmp, c->nx,
                       c->ny, c->pitch);
  if ((ierr = enqueue_forces(c, c->Utmp, c->F))) return ierr;
  if (c->mode == SE2W_FP64)
    hipLaunchKernelGGL(k_to_dmda<double>, dim3(grid_for(n)), dim3(256), 0, c->stream, (const double2 *)c->F,
                       (const double2 *)nullptr, tmp, c->nx, c->ny, c->pitch);
  else
    hipLaunchKernelGGL(k_to_dmda<float>, dim3(grid_for(n)), dim3(256), 0, c->stream, (const float2 *)c->F,
                       (const float2 *)nullptr, tmp, c->nx, c->ny, c->pitch);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(f, tmp, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return SE2W_OK;
}

int se2w_assemble_forces_noalloc(se2w_ctx *c, const double *u, double *f) {
  if (!c || !u || !f) return fail(SE2W_ERR_ARG, "null argument");
  if (!c->have_tables || !c->have_cells) return fail(SE2W_ERR_STATE, "set tables and cells first");
  auto al = [](size_t x) { return (x + 255) / 256 * 256; };
  const long n = (long)c->nx * c->ny;
  const size_t stage = sizeof(double) * 2 * (size_t)n;
  const size_t field = (size_t)c->pitch * c->ny * 2 * real_size(c->mode);
  const size_t need = al(stage) + 2 * al(field);
  if (need > kScratchBytes)
    return fail(SE2W_ERR_UNSUPPORTED,
                "se2w_assemble_forces_noalloc: a mesh of %ld nodes needs %zu bytes of scratch, above its %zu-byte limit "
                "(use se2w_assemble_forces)",
                n, need, kScratchBytes);
  HIPCHK(hipSetDevice(c->device));
  void *base = nullptr;
  if (int ierr = scratch_address(c->device, &base)) return ierr;
  double *tmp = (double *)base;
  void *u_dev = (char *)base + al(stage);
  void *f_dev = (char *)base + al(stage) + al(field);
  HIPCHK(hipMemcpyAsync(tmp, u, stage, hipMemcpyHostToDevice, c->stream));
  // the scratch keeps the previous call's contents: clear f, which the
  // element kernel writes only at the mesh's nodes (not the row padding)
  hipLaunchKernelGGL(k_zero, dim3(grid_for((long)(field / sizeof(double)))), dim3(256), 0, c->stream, (double *)f_dev,
                     (long)(field / sizeof(double)));
  if (c->mode == SE2W_FP64)
    hipLaunchKernelGGL(k_from_dmda<double>, dim3(grid_for(n)), dim3(256), 0, c->stream, tmp, (double2 *)u_dev, c->nx,
                       c->ny, c->pitch);
  else
    hipLaunchKernelGGL(k_from_dmda<float>, dim3(grid_for(n)), dim3(256), 0, c->stream, tmp, (float2 *)u_dev, c->nx,
                       c->ny, c->pitch);
  if (int ierr = enqueue_forces(c, u_dev, f_dev)) return ierr;
  if (c->mode == SE2W_FP64)
    hipLaunchKernelGGL(k_to_dmda<double>, dim3(grid_for(n)), dim3(256), 0, c->stream, (const double2 *)f_dev,
                       (const double2 *)nullptr, tmp, c->nx, c->ny, c->pitch);
  else
    hipLaunchKernelGGL(k_to_dmda<float>, dim3(grid_for(n)), dim3(256), 0, c->stream, (const float2 *)f_dev,
                       (const float2 *)nullptr, tmp, c->nx, c->ny, c->pitch);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(f, tmp, stage, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return SE2W_OK;
}

int se2w_norms(se2w_ctx *c, double out[6]) {
  if (!c || !out) return fail(SE2W_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(c->device));
  const int blocks = 1024;
  if (c->mode == SE2W_FP64)
    hipLaunchKernelGGL(k_norms<double>, dim3(blocks), dim3(256), 0, c->stream, (const double2 *)c->U[c->cur],
                       (const double2 *)c->V, c->nx, c->ny, c->pitch, c->d_partial);
  else
    hipLaunchKernelGGL(k_norms<float>, dim3(blocks), dim3(256), 0, c->stream, (const float2 *)c->U[c->cur],
                       (const float2 *)c->V, c->nx, c->ny, c->pitch, c->d_partial);
  HIPCHK(hipGetLastError());
  std::vector<double> part((size_t)blocks * 6);
  HIPCHK(hipMemcpyAsync(part.data(), c->d_partial, part.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  double r[6] = {-1.0e300, 1.0e300, 0.0, -1.0e300, 1.0e300, 0.0};
  for (int b = 0; b < blocks; b++) {
    r[0] = std::max(r[0], part[b * 6 + 0]);
    r[1] = std::min(r[1], part[b * 6 + 1]);
    r[2] += part[b * 6 + 2];
    r[3] = std::max(r[3], part[b * 6 + 3]);
    r[4] = std::min(r[4], part[b * 6 + 4]);
    r[5] += part[b * 6 + 5];
  }
  out[0] = r[0]; out[1] = r[1]; out[2] = std::sqrt(r[2]);
  out[3] = r[3]; out[4] = r[4]; out[5] = std::sqrt(r[5]);
  return SE2W_OK;
}

int se2w_set_profiling(se2w_ctx *c, int32_t on) {
  if (!c) return fail(SE2W_ERR_ARG, "null context");
  c->profiling = on != 0;
  return SE2W_OK;
}

int se2w_get_timing(se2w_ctx *c, double *run_ms, double kernel_ms[2]) {
  if (!c) return fail(SE2W_ERR_ARG, "null context");
  if (run_ms) *run_ms = c->last_ms;
  if (kernel_ms) { kernel_ms[0] = c->kern_ms[0]; kernel_ms[1] = c->kern_ms[1]; }
  return SE2W_OK;
}

int se2w_set_kernel(se2w_ctx *c, int32_t variant) {
  if (!c) return fail(SE2W_ERR_ARG, "null context");
  if (variant != SE2W_KERNEL_LDS && !is_reg_variant(variant) && variant != SE2W_KERNEL_MFMA &&
      variant != SE2W_KERNEL_LINE && variant != SE2W_KERNEL_LDS_WIDE && !is_reg64_variant(variant))
    return fail(SE2W_ERR_ARG, "unknown kernel %d", variant);
  if (is_reg64_variant(variant) && !reg64_variant_built(c->p, c->mode, variant, c->schedule))
    return fail(SE2W_ERR_UNSUPPORTED,
                "the FP64 register-blocked kernel needs FP64, 2 <= p <= %d (a lane pair per element above p = 4; "
                "schedules %d and %d at p = 8, the others up to p = 7) and the fused schedule (the split "
                "schedule: its first form only)",
                kReg64MaxP, SE2W_KERNEL_REG64 + kReg64FirstP8, SE2W_KERNEL_REG64 + kNumReg64Schedules - 1);
  if (variant == SE2W_KERNEL_LDS_WIDE && c->schedule != SE2W_SCHED_FUSED)
    return fail(SE2W_ERR_UNSUPPORTED, "the 512-thread tile kernel is built for the fused schedule");
  if (variant == SE2W_KERNEL_MFMA && !mfma_available(c->p, c->mode, c->schedule))
    return fail(SE2W_ERR_UNSUPPORTED, "matrix-core kernel needs FP32, 6 <= p <= 8 and the fused schedule");
  if (variant == SE2W_KERNEL_LINE && !mfma_available(c->p, c->mode, c->schedule))
    return fail(SE2W_ERR_UNSUPPORTED, "line kernel needs FP32, 6 <= p <= 8 and the fused schedule");
  if (is_reg_variant(variant) && !reg_available(c->p, c->mode))
    return fail(SE2W_ERR_UNSUPPORTED, "register-blocked kernel needs FP32 and 2 <= p <= 8");
  if (is_reg_variant(variant) && !reg_variant_built(c->p, variant, c->schedule))
    return fail(SE2W_ERR_UNSUPPORTED, "kernel variant %d is not built for p = %d and this schedule", variant, c->p);
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->variant = variant;
  set_tiling(c);
  std::vector<int> ptr, idx;
  int ierr;
  build_tile_lists(c, c->h_src_elem, ptr, idx);
  if ((ierr = upload<int>((void **)&c->d_tsrc_ptr, ptr.data(), ptr.size(), c->stream)) ||
      (ierr = upload<int>((void **)&c->d_tsrc_idx, idx.data(), idx.size(), c->stream)))
    return ierr;
  build_tile_lists(c, c->h_rec_elem, ptr, idx);
  if ((ierr = upload<int>((void **)&c->d_trec_ptr, ptr.data(), ptr.size(), c->stream)) ||
      (ierr = upload<int>((void **)&c->d_trec_idx, idx.data(), idx.size(), c->stream)))
    return ierr;
  build_tile_lists(c, c->h_erec_elem, ptr, idx);
  if ((ierr = upload<int>((void **)&c->d_terec_ptr, ptr.data(), ptr.size(), c->stream)) ||
      (ierr = upload<int>((void **)&c->d_terec_idx, idx.data(), idx.size(), c->stream)))
    return ierr;
  return SE2W_OK;
}

int se2w_kernel_count(void) { return SE2W_KERNEL_REG64 + kNumReg64Schedules; }

// Autotuner: times every variant built for the context's (p, precision,
// schedule) on the context's own mesh and keeps the fastest.  The state is
// snapshotted and restored around the trials, so tuning does not perturb the
// simulation.  Choices are cached per (p, precision, schedule, size bucket)
// for the life of the process.
namespace {
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, int, int>, int> g_tune_cache;
int size_bucket(int n) {
  int b = 0;
  while (n > 1) { n >>= 1; b++; }
  return b;
}
// Copy of a context's state (u, v, a and the compensation of the current
// buffer) kept on the device while trial runs overwrite it.
struct Snapshot {
  void *copy[4] = {nullptr, nullptr, nullptr, nullptr};
  size_t bytes = 0;
  int cur = 0, ns = 0, nr = 0;
};

void snapshot_free(Snapshot &sn) {
  for (void *&x : sn.copy) dev_free(x);
}

int snapshot_take(se2w_ctx *c, Snapshot &sn) {
  sn.bytes = (size_t)c->pitch * c->ny * 2 * real_size(c->mode);
  void *srcs[4] = {c->U[c->cur], c->V, c->A, c->C[c->cur]};
  for (int f = 0; f < 4; f++) {
    if (!srcs[f]) continue;
    int ierr = dev_alloc(&sn.copy[f], sn.bytes);
    if (ierr) { snapshot_free(sn); return ierr; }
    HIPCHK(hipMemcpyAsync(sn.copy[f], srcs[f], sn.bytes, hipMemcpyDeviceToDevice, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  sn.cur = c->cur;
  sn.ns = c->ns;
  sn.nr = c->nr;
  return SE2W_OK;
}

int snapshot_restore(se2w_ctx *c, Snapshot &sn) {
  c->cur = sn.cur;
  c->pred_ok = false;  // the restored plain state; the next run forms its predictor
  // the runs since the snapshot may have exchanged V and C with Vn and Cn:
  // restore into the buffers that hold the plain state now
  void *dsts[4] = {c->U[c->cur], c->V, c->A, c->C[c->cur]};
  for (int f = 0; f < 4; f++)
    if (sn.copy[f]) HIPCHK(hipMemcpyAsync(dsts[f], sn.copy[f], sn.bytes, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  snapshot_free(sn);
  c->ns = sn.ns;
  c->nr = sn.nr;
  c->rec_steps = 0;
  c->erec_steps = 0;
  return SE2W_OK;
}

}  // namespace


int se2w_autotune(se2w_ctx *c, int32_t steps, int32_t reps, int32_t *best_variant, double *best_ms) {
  if (!c || steps < 1 || reps < 1) return fail(SE2W_ERR_ARG, "bad autotune arguments");
  if (!c->have_tables || !c->have_cells || !c->have_mass) return fail(SE2W_ERR_STATE, "set tables, cells and mass first");
  HIPCHK(hipSetDevice(c->device));
  const auto key = std::make_tuple(c->p, c->mode, c->schedule, size_bucket(c->mx), size_bucket(c->my));
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      int ierr = se2w_set_kernel(c, it->second);
      if (ierr) return ierr;
      if (best_variant) *best_variant = it->second;
      if (best_ms) *best_ms = -1.0;
      return SE2W_OK;
    }
  }
  std::vector<int> cands = {SE2W_KERNEL_LDS};
  if (c->schedule == SE2W_SCHED_FUSED) cands.push_back(SE2W_KERNEL_LDS_WIDE);
  if (reg_available(c->p, c->mode))
    for (int k = 0; k < kNumRegSchedules; k++)
      if (reg_variant_built(c->p, SE2W_KERNEL_REG + k, c->schedule)) cands.push_back(SE2W_KERNEL_REG + k);
  for (int k = 0; k < kNumReg64Schedules; k++)
    if (reg64_variant_built(c->p, c->mode, SE2W_KERNEL_REG64 + k, c->schedule)) cands.push_back(SE2W_KERNEL_REG64 + k);
  if (mfma_available(c->p, c->mode, c->schedule)) {
    cands.push_back(SE2W_KERNEL_MFMA);
    cands.push_back(SE2W_KERNEL_LINE);
  }
  Snapshot sn;
  if (int ierr = snapshot_take(c, sn)) return ierr;
  const int variant0 = c->variant;
  std::vector<double> zeros((size_t)steps * std::max(c->ns, 1), 0.0);
  int best = variant0;
  double best_t = 1e300;
  const bool prof0 = c->profiling;
  c->profiling = true;
  for (int v : cands) {
    if (se2w_set_kernel(c, v)) continue;
    std::vector<double> t;
    int ierr = SE2W_OK;
    for (int r = 0; r < reps + 1 && !ierr; r++) {  // first repetition is a warm-up
      ierr = se2w_run(c, steps, 1.0e-30, zeros.data());
      if (!ierr && r > 0) t.push_back(c->kern_ms[0] + c->kern_ms[1]);
    }
    if (ierr || t.empty()) continue;
    std::sort(t.begin(), t.end());
    const double med = t[t.size() / 2];
    if (med < best_t) { best_t = med; best = v; }
  }
  c->profiling = prof0;
  if (int ierr = snapshot_restore(c, sn)) return ierr;
  int ierr = se2w_set_kernel(c, best);
  if (ierr) return ierr;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = best;
  }
  if (best_variant) *best_variant = best;
  if (best_ms) *best_ms = best_t;
  return SE2W_OK;
}

int se2w_get_kernel(se2w_ctx *c, int32_t *variant) {
  if (!c || !variant) return fail(SE2W_ERR_ARG, "null argument");
  *variant = c->variant;
  return SE2W_OK;
}

int se2w_get_tile(se2w_ctx *c, int32_t *tex, int32_t *tey) {
  if (!c) return fail(SE2W_ERR_ARG, "null context");
  if (tex) *tex = c->tex;
  if (tey) *tey = c->tey;
  return SE2W_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Ensembles: many contexts of one device, degree and precision advanced
// together, one batched launch per step.
// ---------------------------------------------------------------------------
struct se2w_ensemble {
  std::vector<se2w_ctx *> members;
  int device = 0, p = 0, mode = 0, variant = SE2W_KERNEL_LDS;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
  void *d_prms = nullptr;
  int *d_tile_off = nullptr;
  int total = 0, grid = 0;
  bool use_graph = true;
  hipGraphExec_t exec = nullptr;
  int exec_steps = 0;
  double run_ms = 0.0, step_ms = 0.0;
};

namespace {

void ens_drop_graph(se2w_ensemble *e) {
  if (e->exec) hipGraphExecDestroy(e->exec);
  e->exec = nullptr;
  e->exec_steps = 0;
}

bool ens_variant_built(const se2w_ensemble *e, int variant) {
  int bpc = 0, thr = 0;
  return ens_occupancy(e->p, e->mode, variant, &bpc, &thr) == 0 && bpc > 0;
}

// Sets the members' tiling for `variant`, the global tile offsets and the
// grid: one workgroup per tile, rounded up to a multiple of 8.
int ens_set_variant(se2w_ensemble *e, int variant) {
  if (!ens_variant_built(e, variant))
    return fail(SE2W_ERR_UNSUPPORTED, "kernel variant %d is not built for ensembles at p = %d", variant, e->p);
  for (se2w_ctx *c : e->members)
    if (int ierr = se2w_set_kernel(c, variant)) return ierr;
  std::vector<int> off(e->members.size() + 1, 0);
  for (size_t i = 0; i < e->members.size(); i++) off[i + 1] = off[i] + e->members[i]->ntiles;
  e->total = off.back();
  if (int ierr = upload<int>((void **)&e->d_tile_off, off.data(), off.size(), e->stream)) return ierr;
  e->grid = (int)std::max<long>(8, ((long)e->total + 7) / 8 * 8);
  e->variant = variant;
  ens_drop_graph(e);
  return SE2W_OK;
}

template <int MODE>
int ens_step_params(se2w_ensemble *e, const double *dt) {
  using R = typename RealOf<MODE>::T;
  const size_t n = e->members.size();
  std::vector<StepParams<R>> prm(2 * n);
  for (size_t i = 0; i < n; i++) {
    se2w_ctx *c = e->members[i];
    prm[2 * i] = make_params<MODE>(c, 0, dt[i], 0);
    c->cur ^= 1;
    prm[2 * i + 1] = make_params<MODE>(c, 0, dt[i], 0);
    c->cur ^= 1;
  }
  HIPCHK(hipMemcpyAsync(e->d_prms, prm.data(), prm.size() * sizeof(StepParams<R>), hipMemcpyHostToDevice, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  return SE2W_OK;
}

template <int MODE>
int ens_launch_steps(se2w_ensemble *e, int nsteps) {
  EnsArgs<typename RealOf<MODE>::T> a;
  a.prms = (const StepParams<typename RealOf<MODE>::T> *)e->d_prms;
  a.tile_off = e->d_tile_off;
  a.n = (int)e->members.size();
  a.total = e->total;
  for (int k = 0; k < nsteps; k++) {
    a.step = k;
    a.last = k == nsteps - 1;
    if (ens_launch(e->p, e->mode, e->variant, &a, e->grid, e->stream))
      return fail(SE2W_ERR_UNSUPPORTED, "no batched kernel for variant %d", e->variant);
  }
  HIPCHK(hipGetLastError());
  return SE2W_OK;
}

template <int MODE>
int ens_run_mode(se2w_ensemble *e, int nsteps, const double *dt) {
  HIPCHK(hipEventRecord(e->ev0, e->stream));
  // every member's predicted state (carried over or formed), then the step
  // parameters of both parities
  for (size_t i = 0; i < e->members.size(); i++) open_run<MODE>(e->members[i], dt[i], e->stream);
  HIPCHK(hipGetLastError());
  if (int ierr = ens_step_params<MODE>(e, dt)) return ierr;
  HIPCHK(hipEventRecord(e->ev1, e->stream));
  if (e->use_graph) {
    if (!e->exec || e->exec_steps != nsteps) {
      ens_drop_graph(e);
      hipGraph_t g = nullptr;
      HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
      int ierr = ens_launch_steps<MODE>(e, nsteps);
      hipError_t ce = hipStreamEndCapture(e->stream, &g);
      if (ierr) { if (g) hipGraphDestroy(g); return ierr; }
      if (ce != hipSuccess) return fail(SE2W_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ce));
      hipError_t ie = hipGraphInstantiate(&e->exec, g, nullptr, nullptr, 0);
      hipGraphDestroy(g);
      if (ie != hipSuccess) { e->exec = nullptr; return fail(SE2W_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(ie)); }
      e->exec_steps = nsteps;
    }
    HIPCHK(hipGraphLaunch(e->exec, e->stream));
  } else {
    if (int ierr = ens_launch_steps<MODE>(e, nsteps)) return ierr;
  }
  HIPCHK(hipEventRecord(e->ev2, e->stream));
  for (size_t i = 0; i < e->members.size(); i++) {
    se2w_ctx *c = e->members[i];
    c->cur ^= ((nsteps - 1) & 1);  // the last step's input: the plain displacement
    close_run(c, dt[i]);
    if (int ierr = stage_record(c, e->stream)) return ierr;
  }
  HIPCHK(hipStreamSynchronize(e->stream));
  float a = 0.f, b = 0.f;
  HIPCHK(hipEventElapsedTime(&a, e->ev0, e->ev2));
  HIPCHK(hipEventElapsedTime(&b, e->ev1, e->ev2));
  e->run_ms = a;
  e->step_ms = b / nsteps;
  return SE2W_OK;
}

int ens_run(se2w_ensemble *e, int nsteps, const double *dt, const double *const *psi) {
  HIPCHK(hipSetDevice(e->device));
  for (size_t i = 0; i < e->members.size(); i++) {
    se2w_ctx *c = e->members[i];
    if (int ierr = prepare_run(c, nsteps, dt[i], psi ? psi[i] : nullptr, e->stream)) return ierr;
  }
  if (e->mode == SE2W_FP64) return ens_run_mode<MODE_F64>(e, nsteps, dt);
  if (e->mode == SE2W_FP32) return ens_run_mode<MODE_F32>(e, nsteps, dt);
  return ens_run_mode<MODE_F32K>(e, nsteps, dt);
}

}  // namespace

extern "C" {

int se2w_ensemble_create(int32_t n, se2w_ctx *const *members, se2w_ensemble **out) {
  if (!out || n < 1 || !members) return fail(SE2W_ERR_ARG, "bad ensemble arguments");
  *out = nullptr;
  se2w_ctx *c0 = members[0];
  for (int i = 0; i < n; i++) {
    se2w_ctx *c = members[i];
    if (!c) return fail(SE2W_ERR_ARG, "member %d is null", i);
    for (int j = 0; j < i; j++)
      if (members[j] == c) return fail(SE2W_ERR_ARG, "member %d appears twice", i);
    if (c->device != c0->device || c->p != c0->p || c->mode != c0->mode)
      return fail(SE2W_ERR_ARG, "member %d: device, degree and precision must match member 0", i);
    if (c->schedule != SE2W_SCHED_FUSED) return fail(SE2W_ERR_UNSUPPORTED, "member %d: ensembles use the fused schedule", i);
    if (!c->have_tables || !c->have_cells || !c->have_mass)
      return fail(SE2W_ERR_STATE, "member %d: set tables, cells and mass first", i);
  }
  HIPCHK(hipSetDevice(c0->device));
  se2w_ensemble *e = new se2w_ensemble;
  e->members.assign(members, members + n);
  e->device = c0->device;
  e->p = c0->p;
  e->mode = c0->mode;
  int ierr = SE2W_OK;
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&e->ev0) != hipSuccess || hipEventCreate(&e->ev1) != hipSuccess ||
      hipEventCreate(&e->ev2) != hipSuccess) {
    se2w_ensemble_destroy(e);
    return fail(SE2W_ERR_HIP, "stream/event creation failed");
  }
  const size_t pbytes = 2 * (size_t)n * (e->mode == SE2W_FP64 ? sizeof(StepParams<double>) : sizeof(StepParams<float>));
  if ((ierr = dev_alloc(&e->d_prms, pbytes))) { se2w_ensemble_destroy(e); return ierr; }
  const int v0 = ens_variant_built(e, c0->variant) ? c0->variant : SE2W_KERNEL_LDS;
  if ((ierr = ens_set_variant(e, v0))) { se2w_ensemble_destroy(e); return ierr; }
  *out = e;
  return SE2W_OK;
}

int se2w_ensemble_run(se2w_ensemble *e, int32_t nsteps, const double *dt, const double *const *psi) {
  if (!e || !dt) return fail(SE2W_ERR_ARG, "null argument");
  if (nsteps < 1) return fail(SE2W_ERR_ARG, "nsteps must be >= 1");
  return ens_run(e, nsteps, dt, psi);
}

int se2w_ensemble_set_kernel(se2w_ensemble *e, int32_t variant) {
  if (!e) return fail(SE2W_ERR_ARG, "null ensemble");
  HIPCHK(hipSetDevice(e->device));
  return ens_set_variant(e, variant);
}

int se2w_ensemble_get_kernel(se2w_ensemble *e, int32_t *variant) {
  if (!e || !variant) return fail(SE2W_ERR_ARG, "null argument");
  *variant = e->variant;
  return SE2W_OK;
}

int se2w_ensemble_set_graph(se2w_ensemble *e, int32_t on) {
  if (!e) return fail(SE2W_ERR_ARG, "null ensemble");
  e->use_graph = on != 0;
  if (!e->use_graph) ens_drop_graph(e);
  return SE2W_OK;
}

int se2w_ensemble_get_timing(se2w_ensemble *e, double *run_ms, double *step_ms) {
  if (!e) return fail(SE2W_ERR_ARG, "null ensemble");
  if (run_ms) *run_ms = e->run_ms;
  if (step_ms) *step_ms = e->step_ms;
  return SE2W_OK;
}

int se2w_ensemble_get_launch(se2w_ensemble *e, int32_t *tiles, int32_t *grid) {
  if (!e) return fail(SE2W_ERR_ARG, "null ensemble");
  if (tiles) *tiles = e->total;
  if (grid) *grid = e->grid;
  return SE2W_OK;
}

int se2w_ensemble_autotune(se2w_ensemble *e, int32_t steps, int32_t reps, int32_t *best_variant, double *best_ms) {
  if (!e || steps < 1 || reps < 1) return fail(SE2W_ERR_ARG, "bad autotune arguments");
  HIPCHK(hipSetDevice(e->device));
  std::vector<int> cands;
  for (int v = 0; v < SE2W_KERNEL_REG64 + kNumReg64Schedules; v++)
    if (ens_variant_built(e, v)) cands.push_back(v);
  const size_t n = e->members.size();
  std::vector<Snapshot> sn(n);
  for (size_t i = 0; i < n; i++) {
    if (int ierr = snapshot_take(e->members[i], sn[i])) {
      for (size_t j = 0; j < i; j++) snapshot_free(sn[j]);
      return ierr;
    }
  }
  std::vector<double> dts(n, 1.0e-30);
  std::vector<std::vector<double>> zeros(n);
  std::vector<const double *> psi(n);
  for (size_t i = 0; i < n; i++) {
    zeros[i].assign((size_t)steps * std::max(e->members[i]->ns, 1), 0.0);
    psi[i] = zeros[i].data();
  }
  const int variant0 = e->variant;
  const bool graph0 = e->use_graph;
  e->use_graph = false;
  int best = variant0;
  double best_t = 1e300;
  for (int v : cands) {
    if (ens_set_variant(e, v)) continue;
    std::vector<double> t;
    int ierr = SE2W_OK;
    for (int r = 0; r < reps + 1 && !ierr; r++) {  // first repetition is a warm-up
      ierr = ens_run(e, steps, dts.data(), psi.data());
      if (!ierr && r > 0) t.push_back(e->step_ms);
    }
    if (ierr || t.empty()) continue;
    std::sort(t.begin(), t.end());
    const double med = t[t.size() / 2];
    if (med < best_t) { best_t = med; best = v; }
  }
  e->use_graph = graph0;
  int ierr = SE2W_OK;
  for (size_t i = 0; i < n; i++) {
    int r = snapshot_restore(e->members[i], sn[i]);
    if (r && !ierr) ierr = r;
  }
  if (ierr) return ierr;
  if ((ierr = ens_set_variant(e, best))) return ierr;
  if (best_variant) *best_variant = best;
  if (best_ms) *best_ms = best_t;
  return SE2W_OK;
}

int se2w_ensemble_get_member(se2w_ensemble *e, int32_t i, se2w_ctx **member) {
  if (!e || !member) return fail(SE2W_ERR_ARG, "null argument");
  if (i < 0 || i >= (int)e->members.size()) return fail(SE2W_ERR_ARG, "member %d out of range", i);
  *member = e->members[i];
  return SE2W_OK;
}

int se2w_ensemble_destroy(se2w_ensemble *e) {
  if (!e) return SE2W_OK;
  hipSetDevice(e->device);
  if (e->stream) hipStreamSynchronize(e->stream);
  ens_drop_graph(e);
  dev_free(e->d_prms);
  void *t = e->d_tile_off;
  dev_free(t);
  if (e->ev0) hipEventDestroy(e->ev0);
  if (e->ev1) hipEventDestroy(e->ev1);
  if (e->ev2) hipEventDestroy(e->ev2);
  if (e->stream) hipStreamDestroy(e->stream);
  delete e;
  return SE2W_OK;
}

}  // extern "C"
