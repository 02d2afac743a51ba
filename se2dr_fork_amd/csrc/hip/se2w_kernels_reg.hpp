// se2w_kernels_reg.hpp -- register-blocked FP32 step kernel.
//
// One thread per element.  The element's (p+1)^2 displacements are loaded
// from HBM/L2 into registers; the tensor-product contractions run row by row
// with the derivative matrix and quadrature weights in scalar registers (the
// same for every lane), and the nodal forces accumulate in registers.
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
// Workgroup = EX x EY element threads: (EX-1) x (EY-1) owned elements plus
// the halo column to the left and the halo row below, which only publish
// their edges.
#pragma once
#include "se2w_kernels.hpp"

namespace se2w {

// Schedules of the register-blocked kernel: workgroup shape (element threads
// EX x EY) and a minimum number of waves per SIMD (0 = compiler's choice),
// which caps the registers per lane.  The autotuner picks one per problem.
struct RegSchedule { int ex, ey, waves; };
constexpr RegSchedule kRegSchedules[] = {{32, 4, 0}, {16, 4, 0}, {32, 2, 0}, {16, 8, 0}, {64, 2, 0}, {16, 4, 3},
                                         {32, 4, 3}, {16, 16, 0}};
constexpr int kNumRegSchedules = sizeof(kRegSchedules) / sizeof(kRegSchedules[0]);

// Schedules instantiated for (p, fused): all of them for p = 4, the first
// three for the other degrees, the first one for the split schedule's
// element kernel.
constexpr bool reg_schedule_built(int P, int S, bool fused) { return fused ? (P == 4 || S < 3) : S == 0; }

// One workgroup of EX x EY threads processes one tile of (EX-1) x (EY-1)
// elements (all threads call it with the same tile).  A workgroup that calls
// it for several tiles in a row puts a barrier between the calls.
template <int P, int MODE, bool FUSED, int EX, int EY>
__device__ __forceinline__ void reg_body(const StepParams<float> &prm, int tile) {
  static_assert(MODE == MODE_F32 || MODE == MODE_F32K, "FP32 kernel");
  constexpr int NB = P + 1;
  constexpr int TEX = EX - 1, TEY = EY - 1;
  __shared__ float2 redge[EY][EX][NB];  // right edge (i = p), j = 0..p
  __shared__ float2 tedge[EY][EX][NB];  // top edge (j = p), i = 0..p
  // nodal forces of the owned nodes; column I stored at I + I/P
  __shared__ float2 sf[TEY * P + 1][TEX * P + 1 + TEX + 1];

  const int tx = tile % prm.tiles_x, ty = tile / prm.tiles_x;
  const int ex0 = tx * TEX, ey0 = ty * TEY;
  const int ex_end = min(ex0 + TEX, prm.mx), ey_end = min(ey0 + TEY, prm.my);
  const int lx = threadIdx.x % EX, ly = threadIdx.x / EX;
  const int ex = ex0 - 1 + lx, ey = ey0 - 1 + ly;
  const bool valid = ex >= 0 && ey >= 0 && ex < ex_end && ey < ey_end;
  const bool owner = valid && lx >= 1 && ly >= 1;
  // owned node block of the tile and its node-phase assignment: one wave per
  // node row, one node per lane
  const int ox1 = ex_end * P - 1 + (ex_end == prm.mx ? 1 : 0), oy1 = ey_end * P - 1 + (ey_end == prm.my ? 1 : 0);
  const int OX = ox1 - ex0 * P + 1, OY = oy1 - ey0 * P + 1;
  constexpr int LANES = EX * EY < 64 ? EX * EY : 64;         // threads per node row
  constexpr int NWAVES = (EX * EY) / LANES;                   // node rows in flight
  static_assert(NWAVES * LANES == EX * EY, "workgroup must be a multiple of its row width");
  constexpr int RPW = (TEY * P + 1 + NWAVES - 1) / NWAVES;  // node rows per row group (max)
  constexpr int CPL = (TEX * P + 1 + LANES - 1) / LANES;    // node columns per thread (max)
  const int wave = threadIdx.x / LANES, lane = threadIdx.x % LANES;

  float ux[NB][NB], uy[NB][NB];  // [j][i]
  float fx[NB][NB], fy[NB][NB];
  long base = 0;
  if (valid) {
    base = (long)(ey * P) * prm.pitch + ex * P;
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
    const int e = ex + ey * prm.mx;
    const float J00 = 0.5f * prm.elem_dx[ex], J11 = 0.5f * prm.elem_dy[ey];
    const float iJ00 = 1.f / J00, iJ11 = 1.f / J11, detJ = J00 * J11;
    const float lam = prm.lambda[e], mu = prm.mu[e], c11 = 2.f * mu + lam;
    // Jacobians, determinant and the internal-force sign folded into the
    // material constants: A = iJ00*fac*sxx etc. with fac = -detJ*w_q.
    const float s = -detJ;
    const float a1 = s * c11 * iJ00 * iJ00, a2 = s * lam * iJ00 * iJ11, a3 = s * c11 * iJ11 * iJ11;
    const float b1 = s * mu * iJ11 * iJ11, b2 = s * mu * iJ00 * iJ11, b3 = s * mu * iJ00 * iJ00;
    const float *__restrict__ D = prm.D;
    const float *__restrict__ W = prm.w2;
#pragma unroll
    for (int j = 0; j < NB; j++)
#pragma unroll
      for (int i = 0; i < NB; i++) fx[j][i] = fy[j][i] = 0.f;
#pragma unroll
    for (int qj = 0; qj < NB; qj++) {
      float A[NB], B[NB], Cx[NB], E[NB];
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
        const float w = W[qi + qj * NB];
        A[qi] = w * fmaf(a1, dxux, a2 * dyuy);   // iJ00 * fac * sxx
        E[qi] = w * fmaf(a2, dxux, a3 * dyuy);   // iJ11 * fac * syy
        B[qi] = w * fmaf(b1, dyux, b2 * dxuy);   // iJ11 * fac * sxy
        Cx[qi] = w * fmaf(b2, dyux, b3 * dxuy);  // iJ00 * fac * sxy
      }
      // row terms: nodes (i, qj) from the points of row qj
#pragma unroll
      for (int i = 0; i < NB; i++) {
#pragma unroll
        for (int qi = 0; qi < NB; qi++) {
          fx[qj][i] = fmaf(D[qi * NB + i], A[qi], fx[qj][i]);
          fy[qj][i] = fmaf(D[qi * NB + i], Cx[qi], fy[qj][i]);
        }
      }
      // column terms: nodes (qi, j) from the point (qi, qj)
#pragma unroll
      for (int j = 0; j < NB; j++) {
#pragma unroll
        for (int qi = 0; qi < NB; qi++) {
          fx[j][qi] = fmaf(D[qj * NB + j], B[qi], fx[j][qi]);
          fy[j][qi] = fmaf(D[qj * NB + j], E[qi], fy[j][qi]);
        }
      }
    }
    // publish the right and top edges for the neighbouring owners
#pragma unroll
    for (int k = 0; k < NB; k++) {
      redge[ly][lx][k] = float2{fx[k][P], fy[k][P]};
      tedge[ly][lx][k] = float2{fx[P][k], fy[P][k]};
    }
  }
  // The node phase's loads do not depend on the forces: issue them all now so
  // that their latency overlaps the force exchange.
  float2 pu[RPW][CPL], pv[RPW][CPL], pc[RPW][CPL];
  float pm[RPW][CPL];
  if constexpr (FUSED) {
#pragma unroll
    for (int r = 0; r < RPW; r++) {
#pragma unroll
      for (int cidx = 0; cidx < CPL; cidx++) {
        const int lyn = wave + r * NWAVES, lxn = lane + cidx * LANES;
        if (lyn < OY && lxn < OX) {
          const long idx = (long)(ey0 * P + lyn) * prm.pitch + ex0 * P + lxn;
          pu[r][cidx] = prm.u_in[idx];
          pv[r][cidx] = prm.v[idx];
          pm[r][cidx] = prm.mass[idx];
          if constexpr (MODE == MODE_F32K) pc[r][cidx] = prm.c_in[idx];
          else pc[r][cidx] = float2{0.f, 0.f};
        }
      }
    }
  }
  __syncthreads();
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
          const int lxn = (lx - 1) * P + i, lyn = (ly - 1) * P + j;
          sf[lyn][lxn + lxn / P] = float2{gx, gy};
        }
      }
    }
  }
  __syncthreads();
  // Node phase: coalesced update of the owned nodes from the prefetched values.
#pragma unroll
  for (int r = 0; r < RPW; r++) {
#pragma unroll
    for (int cidx = 0; cidx < CPL; cidx++) {
      const int lyn = wave + r * NWAVES, lxn = lane + cidx * LANES;
      if (lyn < OY && lxn < OX) {
        const int gx = ex0 * P + lxn, gy = ey0 * P + lyn;
        const long idx = (long)gy * prm.pitch + gx;
        const float2 f = sf[lyn][lxn + lxn / P];
        if constexpr (FUSED) {
          node_update_core<P, MODE>(prm, tile, gx, gy, idx, f.x, f.y, pu[r][cidx].x, pu[r][cidx].y, pv[r][cidx],
                                    pm[r][cidx], pc[r][cidx]);
        } else {
          prm.f[idx] = f;
        }
      }
    }
  }
}

template <int P, int MODE, bool FUSED, int EX, int EY, int WAVES>
__global__ __launch_bounds__(EX * EY) __attribute__((amdgpu_waves_per_eu(WAVES > 0 ? WAVES : 1, 8)))
void reg_kernel(StepParams<float> prm) {
  const int tile = tile_of_block(blockIdx.x, prm.tiles_per_xcd);
  if (tile >= prm.tiles_x * prm.tiles_y) return;  // uniform per workgroup
  reg_body<P, MODE, FUSED, EX, EY>(prm, tile);
}

}  // namespace se2w
