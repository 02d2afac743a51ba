// se2w_kernels_reg.hpp -- register-blocked FP32 step kernel.
//
// One element per thread (LPE = 1) or per pair of adjacent lanes (LPE = 2).
// The element's (p+1)^2 displacements are loaded from HBM/L2 into registers;
// the tensor-product contractions run row by row with the derivative matrix
// in scalar registers (the same for every lane; the quadrature weights too
// for p <= 4, from LDS above), and the nodal forces accumulate in
// registers.  With LPE = 2 each lane of a pair
// holds one displacement component and accumulates one force component: the
// lane computes the x and y derivatives of its component, swaps them with its
// partner (one DPP move each) and forms the two stress combinations its force
// component needs, so the pair does the work of one LPE = 1 thread with half
// the registers per lane.  Both forms perform the same floating-point
// operations in the same order and give identical bits.
//
// Node ownership: element (ex, ey) owns its nodes with local i < p and j < p
// (plus i = p on the domain's right boundary and j = p on its top boundary),
// so every node has exactly one owner.  The owner already holds its own
// contribution; the contributions of the left, lower and lower-left
// neighbours come through LDS, where each element publishes its right edge
// (i = p, j = 0..p) and top edge (j = p, i = 0..p).  The owner writes the
// summed force into a node-major LDS tile, and a node phase (one wave per
// node row, one node per lane) applies the Verlet update with coalesced
// global loads and stores.
//
// Workgroup = EX x EY elements: (EX-1) x (EY-1) owned elements plus the halo
// column to the left and the halo row below, which only publish their edges.
#pragma once
#include "se2w_kernels.hpp"

namespace se2w {

// Schedules of the register-blocked kernel: workgroup shape (EX x EY
// elements), a minimum number of waves per SIMD (0 = compiler's choice, which
// caps the registers per lane), the placement of the node-phase loads (pf, see
// reg_body) and the lanes per element.  The autotuner picks one per
// problem.
struct RegSchedule { int ex, ey, waves, pf, lpe; };
constexpr RegSchedule kRegSchedules[] = {{32, 4, 0, 4, 1}, {64, 2, 0, 4, 1}, {32, 2, 0, 2, 1}, {32, 4, 0, 0, 2},
                                         {16, 4, 0, 4, 2}, {16, 8, 0, 4, 2}, {16, 4, 0, 2, 2}, {16, 4, 0, 0, 2}};
constexpr int kNumRegSchedules = sizeof(kRegSchedules) / sizeof(kRegSchedules[0]);

// Schedules instantiated for (p, fused): all of them for p = 4, the first
// six and the last for p = 2, 3, 5, 6, only the lane-pair schedules 3, 4 and
// the last for p = 7, 8 (one lane per element would need over 400
// registers; the 16 x 8 tile's LDS, 81 KB at p = 8, would leave one
// workgroup per CU); the first one for the split schedule's element kernel
// (p <= 6).
constexpr bool reg_schedule_built(int P, int S, bool fused) {
  return P >= 7 ? (fused && (S == 3 || S == 4 || S == kNumRegSchedules - 1))
                : fused ? (P == 4 || S < 6 || S == kNumRegSchedules - 1) : S == 0;
}

// Partner lane's value (lanes 2k and 2k+1 swap): DPP quad_perm [1, 0, 3, 2].
__device__ __forceinline__ float pair_swap(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0xB1, 0xF, 0xF, false));
}

// Component c (0 = x, 1 = y) of an LDS float2.
__device__ __forceinline__ float &comp(float2 &v, int c) { return reinterpret_cast<float *>(&v)[c]; }

// Loads element (ex, ey)'s nodal displacements, [j][i]: both components
// (LPE = 1) or component c into ux (LPE = 2).
template <int P, int LPE, typename PRM>
__device__ __forceinline__ void reg_load_element(PRM &prm, int ex, int ey, int c, float (&ux)[P + 1][P + 1],
                                                 float (&uy)[P + 1][P + 1]) {
  constexpr int NB = P + 1;
  const long base = (long)(ey * P) * prm.pitch + ex * P;
  if constexpr (LPE == 1) {
    const float2 *ub = prm.u_in + base;
#pragma unroll
    for (int j = 0; j < NB; j++) {
#pragma unroll
      for (int i = 0; i < NB; i++) {
        const float2 t = ub[(long)j * prm.pitch + i];
        ux[j][i] = t.x;
        uy[j][i] = t.y;
      }
    }
  } else {
    const float *ub = reinterpret_cast<const float *>(prm.u_in + base) + c;
#pragma unroll
    for (int j = 0; j < NB; j++)
#pragma unroll
      for (int i = 0; i < NB; i++) ux[j][i] = ub[2 * ((long)j * prm.pitch + i)];
  }
}

// The element operator of one lane: the nodal forces of element (ex, ey)
// from its displacements (LPE = 1: both components; LPE = 2: component c,
// the partner lane holding the other).  sWq: the quadrature weights in LDS
// (W_LDS, p >= 5) or unused.
template <int P, int LPE, bool W_LDS>
struct RegOperator {
  static constexpr int NB = P + 1;
  // D and the weights are read through the constant address space: uniform,
  // never written by a kernel, so they are scalar loads into SGPRs also when
  // prm itself is not a kernel argument (the batched kernels).
  using CF = const __attribute__((address_space(4))) float;
  CF *D;
  CF *W;
  const float *sWq;
  // Jacobians, determinant and the internal-force sign folded into the
  // material constants: A = iJ00*fac*sxx etc. with fac = -detJ*w_q.
  float a1, a2, a3, b1, b2, b3;
  float rA, rB, cA, cB;

  template <typename PRM>
  __device__ __forceinline__ RegOperator(PRM &prm, int ex, int ey, int c, const float *sw) : sWq(sw) {
    const int e = ex + ey * prm.mx;
    const float J00 = 0.5f * prm.elem_dx[ex], J11 = 0.5f * prm.elem_dy[ey];
    const float iJ00 = 1.f / J00, iJ11 = 1.f / J11, detJ = J00 * J11;
    const float lam = prm.lambda[e], mu = prm.mu[e], c11 = 2.f * mu + lam;
    const float s = -detJ;
    a1 = s * c11 * iJ00 * iJ00, a2 = s * lam * iJ00 * iJ11, a3 = s * c11 * iJ11 * iJ11;
    b1 = s * mu * iJ11 * iJ11, b2 = s * mu * iJ00 * iJ11, b3 = s * mu * iJ00 * iJ00;
    D = (CF *)(uintptr_t)prm.D;
    W = (CF *)(uintptr_t)prm.w2;
    rA = c ? b3 : a1, rB = c ? b2 : a2, cA = c ? a3 : b1, cB = c ? a2 : b2;
  }

  __device__ __forceinline__ float weight(int qi, int qj) const { return W_LDS ? sWq[qi + qj * NB] : W[qi + qj * NB]; }

  // Fac-weighted stresses of quadrature row qj, pre-multiplied by the
  // Jacobian factor of the test-function derivative they meet:
  //   A  = w (a1 dx ux + a2 dy uy)  iJ00 fac sxx  (row terms of fx)
  //   B  = w (b1 dy ux + b2 dx uy)  iJ11 fac sxy  (column terms of fx)
  //   Cx = w (b3 dx uy + b2 dy ux)  iJ00 fac sxy  (row terms of fy)
  //   E  = w (a3 dy uy + a2 dx ux)  iJ11 fac syy  (column terms of fy)
  // The first product of each is the one of the lane's own component, so
  // LPE = 2 forms R (A or Cx) and Cq (B or E) from its own derivatives and
  // the partner's with the same operations: R = w (rA own_dx + rB other_dy),
  // Cq = w (cA own_dy + cB other_dx).
  __device__ __forceinline__ void row_stress(int qj, const float (&ux)[NB][NB], const float (&uy)[NB][NB], float *A,
                                             float *B, float *Cx, float *E) const {
#pragma unroll
    for (int qi = 0; qi < NB; qi++) {
      float dxux = 0.f, dxuy = 0.f, dyux = 0.f, dyuy = 0.f;
#pragma unroll
      for (int k = 0; k < NB; k++) {
        dxux = fmaf(D[qi * NB + k], ux[qj][k], dxux);
        dxuy = fmaf(D[qi * NB + k], uy[qj][k], dxuy);
        dyux = fmaf(D[qj * NB + k], ux[k][qi], dyux);
        dyuy = fmaf(D[qj * NB + k], uy[k][qi], dyuy);
      }
      const float w = weight(qi, qj);
      A[qi] = w * fmaf(a1, dxux, a2 * dyuy);
      E[qi] = w * fmaf(a3, dyuy, a2 * dxux);
      B[qi] = w * fmaf(b1, dyux, b2 * dxuy);
      Cx[qi] = w * fmaf(b3, dxuy, b2 * dyux);
    }
  }
  __device__ __forceinline__ void row_stress_pair(int qj, const float (&ux)[NB][NB], float *R, float *Cq) const {
#pragma unroll
    for (int qi = 0; qi < NB; qi++) {
      float dx = 0.f, dy = 0.f;
#pragma unroll
      for (int k = 0; k < NB; k++) {
        dx = fmaf(D[qi * NB + k], ux[qj][k], dx);
        dy = fmaf(D[qj * NB + k], ux[k][qi], dy);
      }
      const float odx = pair_swap(dx), ody = pair_swap(dy);
      const float w = weight(qi, qj);
      R[qi] = w * fmaf(rA, dx, rB * ody);
      Cq[qi] = w * fmaf(cA, dy, cB * odx);
    }
  }

  // All nodal forces, [j][i] (LPE = 2: fx only).
  __device__ __forceinline__ void forces(const float (&ux)[NB][NB], const float (&uy)[NB][NB], float (&fx)[NB][NB],
                                         float (&fy)[NB][NB]) const {
#pragma unroll
    for (int j = 0; j < NB; j++)
#pragma unroll
      for (int i = 0; i < NB; i++) fx[j][i] = fy[j][i] = 0.f;
#pragma unroll
    for (int qj = 0; qj < NB; qj++) {
      float A[NB], B[NB], Cx[NB], E[NB];
      if constexpr (LPE == 1) row_stress(qj, ux, uy, A, B, Cx, E);
      else row_stress_pair(qj, ux, A, B);
      // row terms: nodes (i, qj) from the points of row qj
#pragma unroll
      for (int i = 0; i < NB; i++) {
#pragma unroll
        for (int qi = 0; qi < NB; qi++) {
          fx[qj][i] = fmaf(D[qi * NB + i], A[qi], fx[qj][i]);
          if constexpr (LPE == 1) fy[qj][i] = fmaf(D[qi * NB + i], Cx[qi], fy[qj][i]);
        }
      }
      // column terms: nodes (qi, j) from the point (qi, qj)
#pragma unroll
      for (int j = 0; j < NB; j++) {
#pragma unroll
        for (int qi = 0; qi < NB; qi++) {
          fx[j][qi] = fmaf(D[qj * NB + j], B[qi], fx[j][qi]);
          if constexpr (LPE == 1) fy[j][qi] = fmaf(D[qj * NB + j], E[qi], fy[j][qi]);
        }
      }
    }
  }

  // Only the top edge (j = p) of the forces, with the same operations in the
  // same order as forces(), so its values are identical to forces()'s row p.
  __device__ __forceinline__ void top_edge(const float (&ux)[NB][NB], const float (&uy)[NB][NB], float (&tx)[NB],
                                           float (&ty)[NB]) const {
#pragma unroll
    for (int i = 0; i < NB; i++) tx[i] = ty[i] = 0.f;
#pragma unroll
    for (int qj = 0; qj < NB; qj++) {
      float A[NB], B[NB], Cx[NB], E[NB];
      if constexpr (LPE == 1) row_stress(qj, ux, uy, A, B, Cx, E);
      else row_stress_pair(qj, ux, A, B);
      if (qj == P) {
#pragma unroll
        for (int i = 0; i < NB; i++) {
#pragma unroll
          for (int qi = 0; qi < NB; qi++) {
            tx[i] = fmaf(D[qi * NB + i], A[qi], tx[i]);
            if constexpr (LPE == 1) ty[i] = fmaf(D[qi * NB + i], Cx[qi], ty[i]);
          }
        }
      }
#pragma unroll
      for (int qi = 0; qi < NB; qi++) {
        tx[qi] = fmaf(D[qj * NB + P], B[qi], tx[qi]);
        if constexpr (LPE == 1) ty[qi] = fmaf(D[qj * NB + P], E[qi], ty[qi]);
      }
    }
  }
};


// One workgroup of EX x EY x LPE threads processes one tile of (EX-1) x
// (EY-1) elements (all threads call it with the same tile).  A workgroup that
// calls it for several tiles in a row puts a barrier between the calls.
template <int P, int MODE, bool FUSED, int EX, int EY, int PF, int LPE, typename PRM>
__device__ __forceinline__ void reg_body(PRM &prm, int tile, int step, int last) {
  static_assert(MODE == MODE_F32 || MODE == MODE_F32K, "FP32 kernel");
  static_assert(LPE == 1 || LPE == 2, "one or two lanes per element");
  constexpr int NB = P + 1;
  constexpr int TEX = EX - 1, TEY = EY - 1;
  constexpr int NT = EX * EY * LPE;
  __shared__ float2 redge[EY][EX][NB];  // right edge (i = p), j = 0..p
  __shared__ float2 tedge[EY][EX][NB];  // top edge (j = p), i = 0..p
  // nodal forces of the owned nodes; column I stored at I + I/P
  __shared__ float2 sf[TEY * P + 1][TEX * P + 1 + TEX + 1];
  // p >= 5: the quadrature weights in LDS (uniform reads) rather than in
  // SGPRs, which the (p+1)^2 derivative values already nearly fill.
  constexpr bool W_LDS = P >= 5;
  __shared__ float sWq[W_LDS ? NB * NB : 1];
  if constexpr (W_LDS) {
    for (int k = threadIdx.x; k < NB * NB; k += EX * EY * LPE) sWq[k] = prm.w2[k];
    __syncthreads();
  }

  const int tx = tile % prm.tiles_x, ty = tile / prm.tiles_x;
  const int ex0 = tx * TEX, ey0 = ty * TEY;
  const int ex_end = min(ex0 + TEX, prm.mx), ey_end = min(ey0 + TEY, prm.my);
  const int c = LPE == 2 ? (int)(threadIdx.x & 1) : 0;  // component of this lane (LPE = 2)
  const int el = threadIdx.x / LPE;
  const int lx = el % EX, ly = el / EX;
  const int ex = ex0 - 1 + lx, ey = ey0 - 1 + ly;
  const bool valid = ex >= 0 && ey >= 0 && ex < ex_end && ey < ey_end;
  const bool owner = valid && lx >= 1 && ly >= 1;
  // owned node block of the tile and its node-phase assignment: one wave per
  // node row, one node per lane
  const int ox1 = ex_end * P - 1 + (ex_end == prm.mx ? 1 : 0), oy1 = ey_end * P - 1 + (ey_end == prm.my ? 1 : 0);
  const int OX = ox1 - ex0 * P + 1, OY = oy1 - ey0 * P + 1;
  constexpr int LANES = NT < 64 ? NT : 64;                   // threads per node row
  constexpr int NWAVES = NT / LANES;                          // node rows in flight
  static_assert(NWAVES * LANES == NT, "workgroup must be a multiple of its row width");
  constexpr int RPW = (TEY * P + 1 + NWAVES - 1) / NWAVES;  // node rows per row group (max)
  constexpr int CPL = (TEX * P + 1 + LANES - 1) / LANES;    // node columns per thread (max)
  const int wave = threadIdx.x / LANES, lane = threadIdx.x % LANES;
  // The halo row below the tile fills whole waves when a row of elements is
  // a multiple of 64 lanes: those waves take a shorter path (top edge only).
  constexpr bool HALO_WAVES = (EX * LPE) % 64 == 0;

  // The node phase's loads do not depend on the forces.  A thread's node
  // loads are issued together, ahead of the updates: pf 3 before the
  // element's loads, pf 2 after the element operator (their latency overlaps
  // the force exchange), pf 0 after the exchange barrier (overlapping the
  // owners' sums); pf 1 splits u, v, M (as pf 2) from the compensation (as
  // pf 0); pf 4 issues them in the node phase, CHUNK slots at a time, which
  // caps the registers they hold (high p, many node slots per thread).
  float2 pu[RPW][CPL], pv[RPW][CPL], pc[RPW][CPL];
  float pm[RPW][CPL];
  // Slots past the tile's owned block load its last owned node instead (in
  // bounds, unused): no branch between the loads, so they all issue before
  // the first wait.
  auto node_loads = [&](int r, int cidx, bool uvm, bool compensation) {
    const int lyn = min(wave + r * NWAVES, OY - 1), lxn = min(lane + cidx * LANES, OX - 1);
    const long idx = (long)(ey0 * P + lyn) * prm.pitch + ex0 * P + lxn;
    if (uvm) {
      pu[r][cidx] = prm.u_in[idx];
      pv[r][cidx] = prm.v[idx];
      pm[r][cidx] = prm.mass[idx];
    }
    if (compensation) {
      if constexpr (MODE == MODE_F32K) pc[r][cidx] = prm.c_in[idx];
      else pc[r][cidx] = float2{0.f, 0.f};
    }
  };
  if constexpr (FUSED && PF == 3) {
#pragma unroll
    for (int r = 0; r < RPW; r++)
#pragma unroll
      for (int cidx = 0; cidx < CPL; cidx++) node_loads(r, cidx, true, true);
  }
  // LPE = 1: ux, uy, fx, fy hold both components; LPE = 2: ux and fx hold
  // this lane's component (uy, fy unused).
  float ux[NB][NB], uy[NB][NB];  // [j][i]
  float fx[NB][NB], fy[NB][NB];
  if (valid) {
    reg_load_element<P, LPE>(prm, ex, ey, c, ux, uy);
    const RegOperator<P, LPE, W_LDS> op(prm, ex, ey, c, sWq);
    if (HALO_WAVES && ly == 0) {
      // Halo row: only the top edge (j = p; node (p, p) is also the corner
      // the owners read from redge).
      float tx[NB], ty[NB];
      op.top_edge(ux, uy, tx, ty);
      if constexpr (LPE == 1) {
#pragma unroll
        for (int k = 0; k < NB; k++) tedge[ly][lx][k] = float2{tx[k], ty[k]};
        redge[ly][lx][P] = float2{tx[P], ty[P]};
      } else {
#pragma unroll
        for (int k = 0; k < NB; k++) comp(tedge[ly][lx][k], c) = tx[k];
        comp(redge[ly][lx][P], c) = tx[P];
      }
    } else {
      op.forces(ux, uy, fx, fy);
      // publish the right and top edges for the neighbouring owners
#pragma unroll
      for (int k = 0; k < NB; k++) {
        if constexpr (LPE == 1) {
          redge[ly][lx][k] = float2{fx[k][P], fy[k][P]};
          tedge[ly][lx][k] = float2{fx[P][k], fy[P][k]};
        } else {
          comp(redge[ly][lx][k], c) = fx[k][P];
          comp(tedge[ly][lx][k], c) = fx[P][k];
        }
      }
    }
  }
  if constexpr (FUSED && (PF == 1 || PF == 2)) {
#pragma unroll
    for (int r = 0; r < RPW; r++)
#pragma unroll
      for (int cidx = 0; cidx < CPL; cidx++) node_loads(r, cidx, true, PF == 2);
  }
  __syncthreads();
  const TileLists tl = tile_lists(prm, tile, step, last);
  constexpr int NSL = RPW * CPL, CHUNK = 8;
  constexpr bool CHUNKED = PF == 4;
  if constexpr (FUSED && PF < 2 && !CHUNKED) {
#pragma unroll
    for (int r = 0; r < RPW; r++)
#pragma unroll
      for (int cidx = 0; cidx < CPL; cidx++) node_loads(r, cidx, PF == 0, true);
  }
  // Owners combine their own and the neighbours' contributions in ascending
  // element order and store the nodal force in a node-major LDS tile.
  if (owner) {
    const bool left = ex > 0, below = ey > 0;
    const bool rb = ex == prm.mx - 1, tb = ey == prm.my - 1;  // domain right / top boundary
#pragma unroll
    for (int j = 0; j < NB; j++) {
#pragma unroll
      for (int i = 0; i < NB; i++) {
        if ((i < P || rb) && (j < P || tb)) {
          const int lxn = (lx - 1) * P + i, lyn = (ly - 1) * P + j;
          if constexpr (LPE == 1) {
            float gx = fx[j][i], gy = fy[j][i];
            if (i == 0 && j == 0 && left && below) {  // (ex-1,ey-1), (ex,ey-1), (ex-1,ey), own
              const float2 d = redge[ly - 1][lx - 1][P], b = tedge[ly - 1][lx][0], l = redge[ly][lx - 1][0];
              gx = ((d.x + b.x) + l.x) + gx;
              gy = ((d.y + b.y) + l.y) + gy;
            } else if (i == 0 && left) {
              const float2 l = redge[ly][lx - 1][j];
              gx = l.x + gx;
              gy = l.y + gy;
            } else if (j == 0 && below) {
              const float2 b = tedge[ly - 1][lx][i];
              gx = b.x + gx;
              gy = b.y + gy;
            }
            sf[lyn][lxn + lxn / P] = float2{gx, gy};
          } else {
            float g = fx[j][i];
            if (i == 0 && j == 0 && left && below) {
              g = ((comp(redge[ly - 1][lx - 1][P], c) + comp(tedge[ly - 1][lx][0], c)) + comp(redge[ly][lx - 1][0], c)) + g;
            } else if (i == 0 && left) {
              g = comp(redge[ly][lx - 1][j], c) + g;
            } else if (j == 0 && below) {
              g = comp(tedge[ly - 1][lx][i], c) + g;
            }
            comp(sf[lyn][lxn + lxn / P], c) = g;
          }
        }
      }
    }
  }
  __syncthreads();
  // Node phase: coalesced update of the owned nodes from the loaded values.
  auto node_step = [&](int r, int cidx) {
    const int lyn = wave + r * NWAVES, lxn = lane + cidx * LANES;
    if (lyn < OY && lxn < OX) {
      const int gx = ex0 * P + lxn, gy = ey0 * P + lyn;
      const long idx = (long)gy * prm.pitch + gx;
      const float2 f = sf[lyn][lxn + lxn / P];
      if constexpr (FUSED) {
        node_update_core<P, MODE>(prm, tl, gx, gy, idx, f.x, f.y, pu[r][cidx].x, pu[r][cidx].y, pv[r][cidx],
                                  pm[r][cidx], pc[r][cidx]);
      } else {
        prm.f[idx] = f;
      }
    }
  };
  if constexpr (FUSED && CHUNKED) {
#pragma unroll
    for (int c0 = 0; c0 < NSL; c0 += CHUNK) {
#pragma unroll
      for (int sl = c0; sl < c0 + CHUNK && sl < NSL; sl++) node_loads(sl / CPL, sl % CPL, true, true);
#pragma unroll
      for (int sl = c0; sl < c0 + CHUNK && sl < NSL; sl++) node_step(sl / CPL, sl % CPL);
    }
  } else {
#pragma unroll
    for (int r = 0; r < RPW; r++)
#pragma unroll
      for (int cidx = 0; cidx < CPL; cidx++) node_step(r, cidx);
  }
}

template <int P, int MODE, bool FUSED, int EX, int EY, int WAVES, int PF, int LPE>
__global__ __launch_bounds__(EX * EY * LPE) __attribute__((amdgpu_waves_per_eu(WAVES > 0 ? WAVES : 1, 8)))
void reg_kernel(StepParams<float> prm) {
  const int tile = tile_of_block(blockIdx.x, prm.tiles_per_xcd);
  if (tile >= prm.tiles_x * prm.tiles_y) return;  // uniform per workgroup
  reg_body<P, MODE, FUSED, EX, EY, PF, LPE>(prm, tile, prm.step, prm.last);
}

}  // namespace se2w
