// se2w_kernels.hpp -- CDNA4 (gfx950) kernels for se2wave's explicit
// spectral-element time step.
//
// Storage (HBM): nodal fields are interleaved 2-vectors (x, y) in DMDA node
// order (node = I + J*nx, se2wave.c:706-736) with the row pitch padded to a
// multiple of 32 nodes so that every tile row starts on a 128-byte line.
// Per-cell lambda/mu are indexed e = ei + ej*mx.  Element widths depend only
// on the element column (elem_dx[ei]) or row (elem_dy[ej]) because the
// reference's mesh is a tensor product (se2wave.c:739-779).
//
// Work decomposition: one workgroup per tile of TEX x TEY elements.  The
// workgroup owns the nodes I in [ex0*P, ex_end*P) (+ the last node column at
// the right boundary) and likewise in y.  To finish the internal force of its
// left/bottom boundary nodes it also evaluates the one element column to the
// left and the element row below (halo elements).  Every owned node then sums
// the contributions of its (up to four) elements in ascending element index,
// the accumulation order of the reference's assembly (se2wave.c:1417-1422),
// so the result is deterministic and, in FP64 mode, bit-identical to it.
//
// Modes: MODE_F64 reproduces the reference's FP64 operation order with FMA
// contraction disabled; MODE_F32 evaluates the sum-factorised operator in FP32
// with FMA; MODE_F32K adds a compensated (Kahan) displacement accumulation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace se2w {

enum { MODE_F64 = 0, MODE_F32 = 1, MODE_F32K = 2 };

template <int MODE> struct RealOf { using T = float; };
template <> struct RealOf<MODE_F64> { using T = double; };

template <typename R> struct V2 { using T = float2; };
template <> struct V2<double> { using T = double2; };

// Tile shapes: the largest tile whose LDS footprint (u tile + stress tile)
// stays under 48 KB, favouring few halo elements.
template <int P, typename R> struct TileShape;
#define SE2W_TILE(P_, R_, X_, Y_) \
  template <> struct TileShape<P_, R_> { static constexpr int TEX = X_, TEY = Y_; };
SE2W_TILE(1, float, 32, 24)
SE2W_TILE(2, float, 24, 12)
SE2W_TILE(3, float, 12, 12)
SE2W_TILE(4, float, 10, 8)
SE2W_TILE(5, float, 8, 6)
SE2W_TILE(6, float, 8, 5)
SE2W_TILE(7, float, 5, 5)
SE2W_TILE(8, float, 5, 4)
SE2W_TILE(1, double, 24, 16)
SE2W_TILE(2, double, 12, 12)
SE2W_TILE(3, double, 12, 6)
SE2W_TILE(4, double, 8, 5)
SE2W_TILE(5, double, 5, 5)
SE2W_TILE(6, double, 4, 4)
SE2W_TILE(7, double, 4, 3)
SE2W_TILE(8, double, 3, 3)
#undef SE2W_TILE

constexpr int kThreads = 256;
constexpr int kMaxP = 8;

template <typename R> struct StepParams {
  using R2 = typename V2<R>::T;
  int mx, my, nx, ny, pitch;
  int tiles_x, tiles_y, tiles_per_xcd;
  int tex, tey;    // owned elements per tile of the kernel in use
  const R2 *u_in;  // displacement read by the element operator
  R2 *u_out;       // displacement written by the update (may alias u_in in split mode)
  R2 *v;           // velocity, updated in place
  R2 *a;           // acceleration, written when `last`
  R2 *f;           // split mode: internal force between the two kernels
  const R2 *c_in;  // Kahan compensation (MODE_F32K)
  R2 *c_out;
  R2 *v_next;      // last step: the next run's half-step velocity (null: not formed)
  R2 *c_next;      // last step, MODE_F32K: the next run's compensation
  const R *mass;   // MODE_F64: lumped mass M; FP32 modes: 1/M (per node, padded)
  const R *lambda, *mu;
  const R *elem_dx, *elem_dy;
  const R *D;      // D[q*(P+1) + i] = dN_i/dxi at GLL node q
  const R *w2;     // w2[i + j*(P+1)] = w_i * w_j
  // sources (CSR per tile)
  const int *tile_src_ptr, *tile_src_idx, *src_elem;
  const R *src_vals;  // [ns][2*npe]
  const R *psi;       // [nsteps_run][ns]
  int ns;
  // receivers (CSR per tile)
  const int *tile_rec_ptr, *tile_rec_idx, *rec_elem, *rec_gll;
  const int *tile_erec_ptr, *tile_erec_idx, *erec_elem;  // element receivers (all nodes of an element)
  R *rec;             // [nsteps_run][nr][rec_stride(p)]
  int nr;
  R *erec;            // [nsteps_run][ner][npe][4]: ux uy vx vy at every node of the element
  int ner;
  int step;           // step index inside the current run
  R dt, hdt, hdt2;    // dt, 0.5*dt, 0.5*dt*dt (formed in FP64 on the host)
  // Last step of a run: v and a receive the corrected state of the step, u_in
  // keeps its displacement; with v_next set, the predictor of the next run's
  // first step is applied as on every other step, into u_out, v_next and
  // c_next, so that a next run with the same dt needs no predictor pass;
  // without, u_out receives a copy of the displacement.
  int last;
};

// XCD-aware tile order: hardware places workgroup b on XCD (b mod 8).  Give
// each XCD a contiguous band of tiles so that halo re-reads of a neighbouring
// tile hit the same XCD's L2.
__device__ __forceinline__ int tile_of_block(int b, int tiles_per_xcd) {
  return (b & 7) * tiles_per_xcd + (b >> 3);
}

// The tile's ranges in the per-tile source, receiver and element-receiver
// lists, read once per tile through the constant address space (the lists are
// set up before a run and never written by a kernel): scalar loads, so the
// per-node update carries no memory access for them; and the launch's step
// index and last-step flag.
struct TileLists {
  int s0, s1;  // sources
  int r0, r1;  // receivers
  int q0, q1;  // element receivers
  int step;    // step index inside the run
  int last;    // last step of the run (see StepParams::last)
  __device__ __forceinline__ bool empty() const { return s0 >= s1 && r0 >= r1 && q0 >= q1; }
};

// The step kernels and their device functions take the step parameters as
// `PRM &prm`: the kernel argument itself (StepParams<R>) in the single-instance
// kernels, and in the batched kernels a reference into the members' parameter
// table in the constant address space, so that every field is a scalar load
// where it is used and the pointers loaded from it address global memory.
template <typename PRM>
__device__ __forceinline__ TileLists tile_lists(PRM &prm, int tile, int step, int last) {
  using CI = const __attribute__((address_space(4))) int;
  TileLists t;
  t.step = step;
  t.last = last;
  t.s0 = ((CI *)(uintptr_t)prm.tile_src_ptr)[tile];
  t.s1 = ((CI *)(uintptr_t)prm.tile_src_ptr)[tile + 1];
  t.r0 = ((CI *)(uintptr_t)prm.tile_rec_ptr)[tile];
  t.r1 = ((CI *)(uintptr_t)prm.tile_rec_ptr)[tile + 1];
  if (prm.ner > 0) {
    t.q0 = ((CI *)(uintptr_t)prm.tile_erec_ptr)[tile];
    t.q1 = ((CI *)(uintptr_t)prm.tile_erec_ptr)[tile + 1];
  } else {
    t.q0 = t.q1 = 0;
  }
  return t;
}

// ---------------------------------------------------------------------------
// Phase B: stress at every GLL point of every tile element.
// ---------------------------------------------------------------------------

// FP64, operation order of the reference's dense loop (se2wave.c:1357-1400),
// with the zero products of the padded tensor tables left out (they only add
// signed zeros).  dN_dx = iJ00*Di[i] on row qj, dN_dy = iJ11*Dj[j] on column
// qi, where Di = D[qi][.] and Dj = D[qj][.].
template <int P>
__device__ __forceinline__ void stress_f64(const double *ux_row, const double *ux_col, const double *uy_row,
                                           const double *uy_col, int pitch, int qi, int qj, const double *Di,
                                           const double *Dj, double iJ00, double iJ11, double lam, double mu,
                                           double &sxx, double &syy, double &sxy) {
#pragma clang fp contract(off)
  constexpr int NB = P + 1;
  double e0 = 0.0, e1 = 0.0, e2 = 0.0;
#pragma unroll
  for (int i = 0; i < NB; i++) e0 = e0 + (iJ00 * Di[i]) * ux_row[i];
#pragma unroll
  for (int j = 0; j < NB; j++) e1 = e1 + (iJ11 * Dj[j]) * uy_col[j * pitch];
#pragma unroll
  for (int j = 0; j < NB; j++) {
    if (j != qj) {
      e2 = e2 + (iJ11 * Dj[j]) * ux_col[j * pitch];
    } else {
#pragma unroll
      for (int i = 0; i < NB; i++) {
        double t = (iJ00 * Di[i]) * uy_row[i];
        if (i == qi) t = t + (iJ11 * Dj[j]) * ux_col[j * pitch];  // j == qj
        e2 = e2 + t;
      }
    }
  }
  const double c11 = 2.0 * mu + lam, c12 = lam, c21 = lam, c22 = 2.0 * mu + lam, c33 = mu;
  sxx = c11 * e0 + c12 * e1;
  syy = c21 * e0 + c22 * e1;
  sxy = c33 * e2;
}

// FP32: sum-factorised contractions with the Jacobian applied once, stress
// pre-multiplied by the quadrature factor detJ*w_q.  Di = D[qi][.], Dj =
// D[qj][.].
template <int P>
__device__ __forceinline__ void stress_f32(const float *ux_row, const float *ux_col, const float *uy_row,
                                           const float *uy_col, int pitch, const float *Di, const float *Dj,
                                           float iJ00, float iJ11, float lam, float mu, float fac, float &sxx,
                                           float &syy, float &sxy) {
  constexpr int NB = P + 1;
  float dxux = 0.f, dxuy = 0.f, dyux = 0.f, dyuy = 0.f;
#pragma unroll
  for (int i = 0; i < NB; i++) {
    const float d = Di[i];
    dxux = fmaf(d, ux_row[i], dxux);
    dxuy = fmaf(d, uy_row[i], dxuy);
  }
#pragma unroll
  for (int j = 0; j < NB; j++) {
    const float d = Dj[j];
    dyux = fmaf(d, ux_col[j * pitch], dyux);
    dyuy = fmaf(d, uy_col[j * pitch], dyuy);
  }
  const float e0 = iJ00 * dxux, e1 = iJ11 * dyuy, e2 = iJ11 * dyux + iJ00 * dxuy;
  const float c11 = 2.f * mu + lam;
  sxx = fac * (c11 * e0 + lam * e1);
  syy = fac * (lam * e0 + c11 * e1);
  sxy = fac * (mu * e2);
}

// ---------------------------------------------------------------------------
// Phase C: internal force of node (i,j) of one element from the stresses.
// ---------------------------------------------------------------------------

// FP64: fe[2k] += -fac*(dNdx*sxx + dNdy*sxy); fe[2k+1] += -fac*(dNdy*syy +
// dNdx*sxy) over q ascending (se2wave.c:1402-1413); nonzero terms only.
template <int P>
__device__ __forceinline__ void force_f64(const double *Sxx, const double *Syy, const double *Sxy, int qpitch, int i,
                                          int j, const double *D, const double *w2, double iJ00, double iJ11,
                                          double detJ, double &fx, double &fy) {
#pragma clang fp contract(off)
  constexpr int NB = P + 1;
  double ax = 0.0, ay = 0.0;
#pragma unroll
  for (int qj = 0; qj < NB; qj++) {
    if (qj != j) {
      const int q = i + qj * qpitch;
      const double fac = detJ * w2[i + qj * NB];
      const double dNy = iJ11 * D[qj * NB + j];
      ax = ax + (-fac) * (dNy * Sxy[q]);
      ay = ay + (-fac) * (dNy * Syy[q]);
    } else {
#pragma unroll
      for (int qi = 0; qi < NB; qi++) {
        const int q = qi + j * qpitch;
        const double fac = detJ * w2[qi + j * NB];
        const double dNx = iJ00 * D[qi * NB + i];
        if (qi != i) {
          ax = ax + (-fac) * (dNx * Sxx[q]);
          ay = ay + (-fac) * (dNx * Sxy[q]);
        } else {
          const double dNy = iJ11 * D[j * NB + j];
          ax = ax + (-fac) * (dNx * Sxx[q] + dNy * Sxy[q]);
          ay = ay + (-fac) * (dNy * Syy[q] + dNx * Sxy[q]);
        }
      }
    }
  }
  fx = ax;
  fy = ay;
}

// FP32: transposed contractions of the fac-weighted stresses.
template <int P>
__device__ __forceinline__ void force_f32(const float *Sxx, const float *Syy, const float *Sxy, int qpitch, int i,
                                          int j, const float *D, float iJ00, float iJ11, float &fx, float &fy) {
  constexpr int NB = P + 1;
  float rx = 0.f, ry = 0.f, cx = 0.f, cy = 0.f;
#pragma unroll
  for (int qi = 0; qi < NB; qi++) {
    const float d = D[qi * NB + i];
    rx = fmaf(d, Sxx[qi + j * qpitch], rx);
    ry = fmaf(d, Sxy[qi + j * qpitch], ry);
  }
#pragma unroll
  for (int qj = 0; qj < NB; qj++) {
    const float d = D[qj * NB + j];
    cx = fmaf(d, Sxy[i + qj * qpitch], cx);
    cy = fmaf(d, Syy[i + qj * qpitch], cy);
  }
  fx = -(iJ00 * rx + iJ11 * cx);
  fy = -(iJ00 * ry + iJ11 * cy);
}

// Receiver sampling.  Per receiver and step the record holds what the
// reference's RecordUVA row needs (se2wave.c:2922-2959): u, v, a at the
// receiver's GLL node (gi, gj) of its element, vx on the element's column gi
// and vy on its row gj (for curl v = dvy/dx - dvx/dy, formed on the host).
// Layout per (step, receiver): [ux uy vx vy ax ay | vx(gi, j), j = 0..p |
// vy(i, gj), i = 0..p], i.e. rec_stride(P) = 6 + 2 (P + 1) values.
__host__ __device__ constexpr int rec_stride(int p) { return 6 + 2 * (p + 1); }

template <int P, typename R, typename PRM>
__device__ __forceinline__ void record_receivers(PRM &prm, const TileLists &tl, int gx, int gy, R ux,
                                                 R uy, R vx, R vy, R ax, R ay) {
  constexpr int NB = P + 1;
  for (int t = tl.r0; t < tl.r1; t++) {
    const int r = prm.tile_rec_idx[t];
    const int e = prm.rec_elem[r], g = prm.rec_gll[r];
    const int li = gx - (e % prm.mx) * P, lj = gy - (e / prm.mx) * P;
    const int gi = g % NB, gj = g / NB;
    R *o = prm.rec + ((long)tl.step * prm.nr + r) * rec_stride(P);
    if (li == gi && lj >= 0 && lj <= P) o[6 + lj] = vx;
    if (lj == gj && li >= 0 && li <= P) o[6 + NB + li] = vy;
    if (li == gi && lj == gj) {
      o[0] = ux; o[1] = uy; o[2] = vx; o[3] = vy; o[4] = ax; o[5] = ay;
    }
  }
  {  // element receivers: every node of the element
    for (int t = tl.q0; t < tl.q1; t++) {
      const int r = prm.tile_erec_idx[t];
      const int e = prm.erec_elem[r];
      const int li = gx - (e % prm.mx) * P, lj = gy - (e / prm.mx) * P;
      if (li >= 0 && li <= P && lj >= 0 && lj <= P) {
        R *o = prm.erec + (((long)tl.step * prm.ner + r) * (NB * NB) + li + lj * NB) * 4;
        o[0] = ux; o[1] = uy; o[2] = vx; o[3] = vy;
      }
    }
  }
}

// FP64: a = f / M; v += (0.5 dt) a; record; u += dt v; u += (0.5 dt^2) a;
// v += (0.5 dt) a -- each operation rounded as in the reference's AXPYs.
template <int P, typename PRM>
__device__ __forceinline__ void node_finish_f64(PRM &prm, const TileLists &tl, int gx, int gy, long idx,
                                                double fx, double fy, double ux, double uy, double2 vv, double mass) {
#pragma clang fp contract(off)
  const double ax = fx / mass, ay = fy / mass;
  const double vx = vv.x + prm.hdt * ax, vy = vv.y + prm.hdt * ay;
  record_receivers<P, double>(prm, tl, gx, gy, ux, uy, vx, vy, ax, ay);
  double2 *vdst = prm.v;
  if (tl.last) {
    prm.v[idx] = double2{vx, vy};
    prm.a[idx] = double2{ax, ay};
    if (!prm.v_next) {
      prm.u_out[idx] = double2{ux, uy};
      return;
    }
    vdst = prm.v_next;
  }
  double nux = ux + prm.dt * vx, nuy = uy + prm.dt * vy;
  nux = nux + prm.hdt2 * ax;
  nuy = nuy + prm.hdt2 * ay;
  prm.u_out[idx] = double2{nux, nuy};
  vdst[idx] = double2{vx + prm.hdt * ax, vy + prm.hdt * ay};
}

// FP32: same sequence with fused multiply-adds; MODE_F32K accumulates the
// displacement with a compensated (Kahan) sum, one per reference AXPY.
template <int P, int MODE, typename PRM>
__device__ __forceinline__ void node_finish_f32(PRM &prm, const TileLists &tl, int gx, int gy, long idx,
                                                float fx, float fy, float ux, float uy, float2 vv, float minv,
                                                float2 cc) {
  const float ax = fx * minv, ay = fy * minv;
  const float vx = fmaf(prm.hdt, ax, vv.x), vy = fmaf(prm.hdt, ay, vv.y);
  record_receivers<P, float>(prm, tl, gx, gy, ux, uy, vx, vy, ax, ay);
  float2 *vdst = prm.v;
  [[maybe_unused]] float2 *cdst = prm.c_out;
  if (tl.last) {
    prm.v[idx] = float2{vx, vy};
    prm.a[idx] = float2{ax, ay};
    if (!prm.v_next) {
      prm.u_out[idx] = float2{ux, uy};
      if constexpr (MODE == MODE_F32K) prm.c_out[idx] = cc;
      return;
    }
    vdst = prm.v_next;  // the compensation of u stays in c_in (c_out aliases it)
    cdst = prm.c_next;
  }
  float nux, nuy;
  if constexpr (MODE == MODE_F32K) {
    float2 c = cc;
    float y, t;
    y = fmaf(prm.dt, vx, -c.x); t = ux + y; c.x = (t - ux) - y; nux = t;
    y = fmaf(prm.dt, vy, -c.y); t = uy + y; c.y = (t - uy) - y; nuy = t;
    const float m0 = nux, m1 = nuy;
    y = fmaf(prm.hdt2, ax, -c.x); t = m0 + y; c.x = (t - m0) - y; nux = t;
    y = fmaf(prm.hdt2, ay, -c.y); t = m1 + y; c.y = (t - m1) - y; nuy = t;
    cdst[idx] = c;
  } else {
    nux = fmaf(prm.hdt2, ax, fmaf(prm.dt, vx, ux));
    nuy = fmaf(prm.hdt2, ay, fmaf(prm.dt, vy, uy));
  }
  prm.u_out[idx] = float2{nux, nuy};
  vdst[idx] = float2{fmaf(prm.hdt, ax, vx), fmaf(prm.hdt, ay, vy)};
}

// ---------------------------------------------------------------------------
// Node update with the nodal v, mass and compensation already loaded.
// ---------------------------------------------------------------------------
template <int P, int MODE, typename PRM>
__device__ __forceinline__ void node_update_core(PRM &prm, const TileLists &tl,
                                                 int gx, int gy, long idx, typename RealOf<MODE>::T fx,
                                                 typename RealOf<MODE>::T fy, typename RealOf<MODE>::T ux,
                                                 typename RealOf<MODE>::T uy, typename V2<typename RealOf<MODE>::T>::T vv,
                                                 typename RealOf<MODE>::T mass,
                                                 typename V2<typename RealOf<MODE>::T>::T cc) {
#pragma clang fp contract(off)
  using R = typename RealOf<MODE>::T;
  constexpr int NB = P + 1;
  // g = sum_s psi_s * vals_s in source order, then f = f + g (3476-3495, 4739)
  {
    if (tl.s0 < tl.s1) {
      R gxv = R(0), gyv = R(0);
      bool hit = false;
      for (int t = tl.s0; t < tl.s1; t++) {
        const int s = prm.tile_src_idx[t];
        const int e = prm.src_elem[s];
        const int li = gx - (e % prm.mx) * P, lj = gy - (e / prm.mx) * P;
        if (li >= 0 && li <= P && lj >= 0 && lj <= P) {
          const int k = li + lj * NB;
          const R ps = prm.psi[(long)tl.step * prm.ns + s];
          gxv = gxv + ps * prm.src_vals[(long)s * 2 * NB * NB + 2 * k];
          gyv = gyv + ps * prm.src_vals[(long)s * 2 * NB * NB + 2 * k + 1];
          hit = true;
        }
      }
      if (hit) {
        fx = fx + gxv;
        fy = fy + gyv;
      }
    }
  }
  if constexpr (MODE == MODE_F64) {
    node_finish_f64<P>(prm, tl, gx, gy, idx, fx, fy, ux, uy, vv, mass);
  } else {
    node_finish_f32<P, MODE>(prm, tl, gx, gy, idx, fx, fy, ux, uy, vv, mass, cc);
  }
}

// Phase D: corrector of step n+1, receiver sampling, predictor of step n+2.
// Sequence per node (se2wave.c:4739-4745 then 4726-4730):
//   f = f + g;  a = f / M;  v = v + (0.5 dt) a;  [record u, v, a]
//   u = u + dt v;  u = u + (0.5 dt^2) a;  v = v + (0.5 dt) a.
template <int P, int MODE, typename PRM>
__device__ __forceinline__ void node_update(PRM &prm, const TileLists &tl,
                                            int gx, int gy, typename RealOf<MODE>::T fx, typename RealOf<MODE>::T fy,
                                            typename RealOf<MODE>::T ux, typename RealOf<MODE>::T uy) {
  using R = typename RealOf<MODE>::T;
  using R2 = typename V2<R>::T;
  const long idx = (long)gy * prm.pitch + gx;
  R2 cc = R2{R(0), R(0)};
  if constexpr (MODE == MODE_F32K) cc = prm.c_in[idx];
  node_update_core<P, MODE>(prm, tl, gx, gy, idx, fx, fy, ux, uy, prm.v[idx], prm.mass[idx], cc);
}

// Owned node k of a tile (OX x OY owned nodes starting on an element corner)
// in class order: first the nodes inside one element column and row, then
// those on element-column edges (x a multiple of p), on element-row edges,
// and the corners.  Within a class the same number of elements meets at every
// node (except on the mesh boundary), so the threads of a wave take the same
// path through the sum over elements.
template <int P>
__device__ __forceinline__ void owned_node(int k, int OX, int OY, int &ox, int &oy) {
  const int nxe = (OX - 1) / P + 1, nye = (OY - 1) / P + 1;  // edge columns / rows
  const int nxi = OX - nxe, nyi = OY - nye;                  // interior columns / rows
  int c = k;
  if constexpr (P > 1) {
    auto inner = [](int i) { return (i / (P - 1)) * P + i % (P - 1) + 1; };
    if (c < nxi * nyi) {
      ox = inner(c % nxi);
      oy = inner(c / nxi);
      return;
    }
    c -= nxi * nyi;
    if (c < nxe * nyi) {
      ox = (c % nxe) * P;
      oy = inner(c / nxe);
      return;
    }
    c -= nxe * nyi;
    if (c < nxi * nye) {
      ox = inner(c % nxi);
      oy = (c / nxi) * P;
      return;
    }
    c -= nxi * nye;
  }
  ox = (c % nxe) * P;
  oy = (c / nxe) * P;
}

// ---------------------------------------------------------------------------
// Tile kernel.  FUSED: element operator + update in one pass (one launch per
// step).  !FUSED: element operator only, f written for the update kernel.
// ---------------------------------------------------------------------------
// One workgroup processes one tile (all threads call it with the same tile).
// A workgroup that calls it for several tiles in a row must put a barrier
// between the calls: the last phase of one call reads `su`, which the first
// phase of the next call overwrites.
template <int P, int MODE, bool FUSED, int NT = kThreads, typename PRM>
__device__ __forceinline__ void tile_body(PRM &prm, int tile, int step, int last) {
  using R = typename RealOf<MODE>::T;
  using R2 = typename V2<R>::T;
  constexpr int NB = P + 1;
  constexpr int TEX = TileShape<P, R>::TEX, TEY = TileShape<P, R>::TEY;
  constexpr int NXL = (TEX + 1) * P + 1, NYL = (TEY + 1) * P + 1;
  constexpr int QX = (TEX + 1) * NB, QY = (TEY + 1) * NB;
  constexpr int NE = (TEX + 1) * (TEY + 1);
  __shared__ R su[2][NYL][NXL];
  __shared__ R ss[3][QY][QX];
  __shared__ R sD[NB * NB], sW[NB * NB];
  // per tile element (+ halo): 1/J00, 1/J11, J00 J11, lambda, mu
  __shared__ R econ[5][NE];

  const int tx = tile % prm.tiles_x, ty = tile / prm.tiles_x;
  const int ex0 = tx * TEX, ey0 = ty * TEY;
  const int gx0 = (ex0 - 1) * P, gy0 = (ey0 - 1) * P;
  const int tid = threadIdx.x;

  for (int k = tid; k < NB * NB; k += NT) {
    sD[k] = prm.D[k];
    sW[k] = prm.w2[k];
  }
  for (int k = tid; k < NE; k += NT) {
    const int ex = ex0 - 1 + k % (TEX + 1), ey = ey0 - 1 + k / (TEX + 1);
    if (ex >= 0 && ex < prm.mx && ey >= 0 && ey < prm.my) {
#pragma clang fp contract(off)
      const R J00 = R(0.5) * prm.elem_dx[ex], J11 = R(0.5) * prm.elem_dy[ey];
      econ[0][k] = R(1) / J00;
      econ[1][k] = R(1) / J11;
      econ[2][k] = J00 * J11;
      econ[3][k] = prm.lambda[ex + ey * prm.mx];
      econ[4][k] = prm.mu[ex + ey * prm.mx];
    }
  }
  // Node data of the update (v, M, compensation) for the thread's owned
  // nodes.  For p <= 6 all of them are loaded at once before phase A, so that
  // the latency overlaps phases A and B; slots past the owned block load its
  // last node (in bounds, unused), so the loads carry no branches and all
  // issue before the first wait.  For p >= 7 the registers would lower the
  // occupancy: each node's data is loaded in phase D.
  const int ex_end = min(ex0 + TEX, prm.mx), ey_end = min(ey0 + TEY, prm.my);
  const int OX = (ex_end - ex0) * P + (ex_end == prm.mx ? 1 : 0);
  const int OY = (ey_end - ey0) * P + (ey_end == prm.my ? 1 : 0);
  constexpr int NSLOT = ((TEX * P + 1) * (TEY * P + 1) + NT - 1) / NT;  // owned nodes per thread (max)
  constexpr bool EARLY = P <= 6;
  // the thread's owned nodes, in class order (see owned_node); slots past the
  // owned block repeat its last node
  int sox[NSLOT], soy[NSLOT];
#pragma unroll
  for (int sl = 0; sl < NSLOT; sl++) owned_node<P>(min(tid + sl * NT, OX * OY - 1), OX, OY, sox[sl], soy[sl]);
  R2 pv[NSLOT], pc[NSLOT];
  R pm[NSLOT];
  auto node_loads = [&]() {
#pragma unroll
    for (int sl = 0; sl < NSLOT; sl++) {
      const long idx = (long)(ey0 * P + soy[sl]) * prm.pitch + ex0 * P + sox[sl];
      pv[sl] = prm.v[idx];
      pm[sl] = prm.mass[idx];
      if constexpr (MODE == MODE_F32K) pc[sl] = prm.c_in[idx];
      else pc[sl] = R2{R(0), R(0)};
    }
  };
  if constexpr (FUSED && EARLY) node_loads();
  (void)node_loads;
  // Phase A: displacement of the tile's nodes (+ halo) into LDS.
  for (int k = tid; k < NXL * NYL; k += NT) {
    const int ly = k / NXL, lx = k - ly * NXL;
    const int gx = gx0 + lx, gy = gy0 + ly;
    R2 uv = R2{R(0), R(0)};
    if (gx >= 0 && gx < prm.nx && gy >= 0 && gy < prm.ny) uv = prm.u_in[(long)gy * prm.pitch + gx];
    su[0][ly][lx] = uv.x;
    su[1][ly][lx] = uv.y;
  }
  __syncthreads();
  // Phase B: stress at the GLL points of the tile elements (+ halo).
  // FP32 modes: each thread keeps one quadrature point (qi, qj) for all its
  // elements, so the two rows of D it needs stay in registers and only the
  // displacements are read from LDS.  FP64: the points of the tile are dealt
  // out evenly (the heavier per-point work favours balance over the table
  // reads).
  if constexpr (MODE == MODE_F64) {
    for (unsigned k = tid; k < (unsigned)(QX * QY); k += NT) {
      const unsigned qy = k / QX, qx = k - qy * QX;
      const unsigned elx = qx / NB, qi = qx - elx * NB;
      const unsigned ely = qy / NB, qj = qy - ely * NB;
      const int ex = ex0 - 1 + (int)elx, ey = ey0 - 1 + (int)ely;
      if (ex < 0 || ex >= prm.mx || ey < 0 || ey >= prm.my) continue;
      const unsigned ei = ely * (TEX + 1) + elx;
      const unsigned lx0 = elx * P, ly0 = ely * P;
      R sxx, syy, sxy;
      stress_f64<P>(&su[0][ly0 + qj][lx0], &su[0][ly0][lx0 + qi], &su[1][ly0 + qj][lx0], &su[1][ly0][lx0 + qi], NXL,
                    qi, qj, &sD[qi * NB], &sD[qj * NB], econ[0][ei], econ[1][ei], econ[3][ei], econ[4][ei], sxx, syy,
                    sxy);
      ss[0][qy][qx] = sxx;
      ss[1][qy][qx] = syy;
      ss[2][qy][qx] = sxy;
    }
  } else {
    constexpr int NQ = NB * NB, GROUPS = NT / NQ;
    static_assert(GROUPS >= 1, "a thread per quadrature point of an element");
    if (tid < GROUPS * NQ) {
      const int q = tid % NQ, g = tid / NQ;
      const int qi = q % NB, qj = q / NB;
      R Di[NB], Dj[NB];
#pragma unroll
      for (int k = 0; k < NB; k++) {
        Di[k] = sD[qi * NB + k];
        Dj[k] = sD[qj * NB + k];
      }
      const R wq = sW[q];
      for (int ei = g; ei < NE; ei += GROUPS) {
        const int elx = ei % (TEX + 1), ely = ei / (TEX + 1);
        const int ex = ex0 - 1 + elx, ey = ey0 - 1 + ely;
        if (ex < 0 || ex >= prm.mx || ey < 0 || ey >= prm.my) continue;
        const int lx0 = elx * P, ly0 = ely * P;
        const R *uxr = &su[0][ly0 + qj][lx0];
        const R *uyr = &su[1][ly0 + qj][lx0];
        const R *uxc = &su[0][ly0][lx0 + qi];
        const R *uyc = &su[1][ly0][lx0 + qi];
        R sxx, syy, sxy;
        const float fac = econ[2][ei] * wq;
        stress_f32<P>(uxr, uxc, uyr, uyc, NXL, Di, Dj, econ[0][ei], econ[1][ei], econ[3][ei], econ[4][ei], fac, sxx,
                      syy, sxy);
        const int qy = ely * NB + qj, qx = elx * NB + qi;
        ss[0][qy][qx] = sxx;
        ss[1][qy][qx] = syy;
        ss[2][qy][qx] = sxy;
      }
    }
  }
  __syncthreads();
  // Phase C (+D): owned nodes.
  const TileLists tl = tile_lists(prm, tile, step, last);
#pragma unroll
  for (int sl = 0; sl < NSLOT; sl++) {
    if (tid + sl * NT >= OX * OY) break;
    const int gx = ex0 * P + sox[sl], gy = ey0 * P + soy[sl];
    int exs[2], eys[2], nex = 0, ney = 0;
    if (gx % P == 0) {
      if (gx / P - 1 >= 0) exs[nex++] = gx / P - 1;
      if (gx / P < prm.mx) exs[nex++] = gx / P;
    } else {
      exs[nex++] = gx / P;
    }
    if (gy % P == 0) {
      if (gy / P - 1 >= 0) eys[ney++] = gy / P - 1;
      if (gy / P < prm.my) eys[ney++] = gy / P;
    } else {
      eys[ney++] = gy / P;
    }
    R fx = R(0), fy = R(0);
    for (int b = 0; b < ney; b++) {
      for (int a = 0; a < nex; a++) {
        const int ex = exs[a], ey = eys[b];
        const int elx = ex - ex0 + 1, ely = ey - ey0 + 1;
        const int i = gx - ex * P, j = gy - ey * P;
        const R *Sxx = &ss[0][ely * NB][elx * NB];
        const R *Syy = &ss[1][ely * NB][elx * NB];
        const R *Sxy = &ss[2][ely * NB][elx * NB];
        R cfx, cfy;
        const int ei = ely * (TEX + 1) + elx;
        if constexpr (MODE == MODE_F64) {
#pragma clang fp contract(off)
          force_f64<P>(Sxx, Syy, Sxy, QX, i, j, sD, sW, econ[0][ei], econ[1][ei], econ[2][ei], cfx, cfy);
          fx = fx + cfx;
          fy = fy + cfy;
        } else {
          force_f32<P>(Sxx, Syy, Sxy, QX, i, j, sD, econ[0][ei], econ[1][ei], cfx, cfy);
          fx += cfx;
          fy += cfy;
        }
      }
    }
    if constexpr (FUSED) {
      const int lx = gx - gx0, ly = gy - gy0;
      if constexpr (EARLY) {
        const long idx = (long)gy * prm.pitch + gx;
        node_update_core<P, MODE>(prm, tl, gx, gy, idx, fx, fy, su[0][ly][lx], su[1][ly][lx], pv[sl], pm[sl], pc[sl]);
      } else {
        node_update<P, MODE>(prm, tl, gx, gy, fx, fy, su[0][ly][lx], su[1][ly][lx]);
      }
    } else {
      prm.f[(long)gy * prm.pitch + gx] = R2{fx, fy};
    }
  }
}

// NT threads per workgroup: 256, or 512 for more waves in flight per tile.
template <int P, int MODE, bool FUSED, int NT = kThreads>
__global__ __launch_bounds__(NT) void tile_kernel(StepParams<typename RealOf<MODE>::T> prm) {
  const int tile = tile_of_block(blockIdx.x, prm.tiles_per_xcd);
  if (tile >= prm.tiles_x * prm.tiles_y) return;  // uniform per workgroup
  tile_body<P, MODE, FUSED, NT>(prm, tile, prm.step, prm.last);
}

// Update kernel of the split schedule: one workgroup per tile of owned nodes.
template <int P, int MODE>
__global__ __launch_bounds__(kThreads) void update_kernel(StepParams<typename RealOf<MODE>::T> prm) {
  using R = typename RealOf<MODE>::T;
  using R2 = typename V2<R>::T;
  const int TEX = prm.tex, TEY = prm.tey;
  const int ntiles = prm.tiles_x * prm.tiles_y;
  const int tile = tile_of_block(blockIdx.x, prm.tiles_per_xcd);
  if (tile >= ntiles) return;
  const int tx = tile % prm.tiles_x, ty = tile / prm.tiles_x;
  const int ex0 = tx * TEX, ey0 = ty * TEY;
  const int ex_end = min(ex0 + TEX, prm.mx), ey_end = min(ey0 + TEY, prm.my);
  const int OX = (ex_end - ex0) * P + (ex_end == prm.mx ? 1 : 0);
  const int OY = (ey_end - ey0) * P + (ey_end == prm.my ? 1 : 0);
  const TileLists tl = tile_lists(prm, tile, prm.step, prm.last);
  for (int k = threadIdx.x; k < OX * OY; k += kThreads) {
    const int oy = k / OX, ox = k - oy * OX;
    const int gx = ex0 * P + ox, gy = ey0 * P + oy;
    const long idx = (long)gy * prm.pitch + gx;
    const R2 f = prm.f[idx];
    const R2 u = prm.u_in[idx];
    node_update<P, MODE>(prm, tl, gx, gy, f.x, f.y, u.x, u.y);
  }
}

// Predictor that opens a run from a plain state (u_n, v_n, a_n):
// u* = u + dt v + (0.5 dt^2) a;  v* = v + (0.5 dt) a   (se2wave.c:4726-4730),
// with the operations of the step kernels' own predictor (FP64: each product
// and sum rounded; FP32 modes: fused multiply-adds), so that a run split into
// chunks gives the bits of the unsplit run.
template <int MODE>
__global__ __launch_bounds__(kThreads) void predictor_kernel(StepParams<typename RealOf<MODE>::T> prm) {
#pragma clang fp contract(off)
  using R = typename RealOf<MODE>::T;
  using R2 = typename V2<R>::T;
  const long n = (long)prm.pitch * prm.ny;
  for (long idx = (long)blockIdx.x * kThreads + threadIdx.x; idx < n; idx += (long)gridDim.x * kThreads) {
    const int gx = (int)(idx % prm.pitch);
    if (gx >= prm.nx) continue;
    const R2 u = prm.u_in[idx];
    const R2 v = prm.v[idx];
    const R2 a = prm.a[idx];
    R nux, nuy;
    if constexpr (MODE == MODE_F32K) {
      R2 c = prm.c_in[idx];
      R y, t;
      y = fmaf(prm.dt, v.x, -c.x); t = u.x + y; c.x = (t - u.x) - y; nux = t;
      y = fmaf(prm.dt, v.y, -c.y); t = u.y + y; c.y = (t - u.y) - y; nuy = t;
      R m0 = nux, m1 = nuy;
      y = fmaf(prm.hdt2, a.x, -c.x); t = m0 + y; c.x = (t - m0) - y; nux = t;
      y = fmaf(prm.hdt2, a.y, -c.y); t = m1 + y; c.y = (t - m1) - y; nuy = t;
      prm.c_out[idx] = c;
    } else if constexpr (MODE == MODE_F32) {
      nux = fmaf(prm.hdt2, a.x, fmaf(prm.dt, v.x, u.x));
      nuy = fmaf(prm.hdt2, a.y, fmaf(prm.dt, v.y, u.y));
    } else {
      nux = u.x + prm.dt * v.x;
      nuy = u.y + prm.dt * v.y;
      nux = nux + prm.hdt2 * a.x;
      nuy = nuy + prm.hdt2 * a.y;
    }
    prm.u_out[idx] = R2{nux, nuy};
    if constexpr (MODE == MODE_F64) prm.v[idx] = R2{v.x + prm.hdt * a.x, v.y + prm.hdt * a.y};
    else prm.v[idx] = R2{fmaf(prm.hdt, a.x, v.x), fmaf(prm.hdt, a.y, v.y)};
  }
}

}  // namespace se2w
