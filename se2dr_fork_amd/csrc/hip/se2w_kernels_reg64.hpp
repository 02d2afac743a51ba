// se2w_kernels_reg64.hpp -- register-blocked FP64 step kernel in the
// reference's operation order.
//
// A pair of adjacent lanes holds one element: lane 2k the x components
// (displacements ux and force accumulators fx), lane 2k+1 the y components,
// each (p+1)^2 doubles in VGPRs.  For every quadrature point q = (qi, qj), in
// ascending q, each lane forms the products of its own component that the
// reference's dense loop sums (se2wave.c:1357-1377, zero entries of the
// tensor tables left out):
//   row products    dN_dx[q][(i,qj)] * u(i, qj),  i = 0..p
//   column products dN_dy[q][(qi,j)] * u(qi, j),  j = 0..p
// with dN_dx = iJ00 * D[qi][i] and dN_dy = iJ11 * D[qj][j] rounded as the
// reference rounds them (1270-1277).  e0 is the x lane's row sum, e1 the y
// lane's column sum, and e2 interleaves the x lane's column products with the
// y lane's row products in ascending node order k = i + j (p+1): one DPP move
// per 32-bit half broadcasts a value of one lane of the pair to both, so both
// lanes sum e2 in the same order and hold the same e0, e1, e2.  The stress
// (1391-1400) and the force terms -fac * (dN_dx sxx + dN_dy sxy) and
// -fac * (dN_dy syy + dN_dx sxy) (1402-1413) follow with FMA contraction off;
// each lane accumulates its own component for the nodes of row qj and
// column qi, so every node's sum runs over q in ascending order.
//
// Tiling, edge exchange and node phase are those of the FP32 register kernel
// (se2w_kernels_reg.hpp): a workgroup of EX x EY elements whose left column
// and bottom row are halo elements that only publish their edges; owners sum
// the contributions of their (up to four) elements in ascending element
// order ((e_ll + e_lr) + e_ul) + e_ur, as the reference's assembly into a
// zeroed vector does, so the result is bit-identical to the oracle.
#pragma once
#include "se2w_kernels.hpp"

namespace se2w {

// Schedules of the FP64 register kernel: workgroup shape (EX x EY elements,
// two lanes each), waves per SIMD (0 = the compiler's choice; 3 caps the
// registers at 168 per lane, which at p = 4 spills a few to scratch; 1 gives
// a lane the whole register file, 256 VGPRs and 256 AGPRs, the AGPRs taking
// the spills) and the placement of the node-phase loads (0: after the force
// exchange barrier, all slots at once; 4: in the node phase, CHUNK slots at
// a time).
struct Reg64Schedule { int ex, ey, waves, pf, lpe; };
constexpr Reg64Schedule kReg64Schedules[] = {{32, 4, 3, 0, 2}, {32, 4, 0, 0, 2}, {16, 8, 0, 4, 2}, {32, 4, 3, 4, 2},
                                             {32, 4, 2, 0, 1}, {64, 2, 2, 0, 1}, {32, 4, 2, 4, 1}, {16, 4, 1, 4, 2},
                                             {32, 4, 1, 4, 2}};
constexpr int kNumReg64Schedules = sizeof(kReg64Schedules) / sizeof(kReg64Schedules[0]);
// The first schedule of the p = 8 pair.
constexpr int kReg64FirstP8 = 7;
// Degrees with a register-resident element (two (p+1)^2 arrays of doubles
// per lane): 2 <= p <= 8; above p = 4 only the lane pairs (one lane per
// element would need over 300 VGPRs at p = 5).  The first seven schedules
// for 2 <= p <= 7; at p = 8, where the lane pair's 2 x 162 VGPRs of nodal
// values and forces exceed 256, the two one-wave schedules, which keep the
// spills in AGPRs (48 bytes per lane go to scratch; the compiler's choice of
// registers sent kilobytes there and ran slower than the LDS tile kernel).
// Every schedule for the fused step, the first one also as the element
// kernel of the split schedule (f = K(u) stored, no update).
constexpr int kReg64MaxP = 8;
constexpr bool reg64_schedule_built(int P, int S, bool fused) {
  return P >= 2 && P <= kReg64MaxP && S >= 0 && S < kNumReg64Schedules && (P <= 4 || kReg64Schedules[S].lpe == 2) &&
         (S < kReg64FirstP8 ? P <= 7 && (fused || S == 0) : P == 8 && fused);
}

// The value that lane 2k (x) or lane 2k+1 (y) of a pair holds, in both
// lanes of the pair: DPP quad_perm [0,0,2,2] / [1,1,3,3] on each 32-bit half.
template <int CTRL>
__device__ __forceinline__ double pair_bcast(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double from_x(double v) { return pair_bcast<0xA0>(v); }
__device__ __forceinline__ double from_y(double v) { return pair_bcast<0xF5>(v); }

__device__ __forceinline__ double &comp64(double2 &v, int c) { return reinterpret_cast<double *>(&v)[c]; }

// One workgroup of EX x EY x 2 threads processes one tile of (EX-1) x (EY-1)
// elements (all threads call it with the same tile).  A workgroup that calls
// it for several tiles in a row puts a barrier between the calls.
template <int P, int EX, int EY, int PF, bool FUSED = true, typename PRM>
__device__ __forceinline__ void reg64_body(PRM &prm, int tile, int step, int last) {
#pragma clang fp contract(off)
  constexpr int NB = P + 1;
  constexpr int TEX = EX - 1, TEY = EY - 1;
  constexpr int NT = EX * EY * 2;
  __shared__ double2 redge[EY][EX][NB];  // right edge (i = p), j = 0..p
  __shared__ double2 tedge[EY][EX][NB];  // top edge (j = p), i = 0..p
  // nodal forces of the owned nodes; column I stored at I + I/P
  __shared__ double2 sf[TEY * P + 1][TEX * P + 1 + TEX + 1];
  // quadrature weights (uniform reads; D alone takes 2 (p+1)^2 SGPRs)
  __shared__ double sW[NB * NB];
  for (int k = threadIdx.x; k < NB * NB; k += NT) sW[k] = prm.w2[k];
  __syncthreads();

  const int tx = tile % prm.tiles_x, ty = tile / prm.tiles_x;
  const int ex0 = tx * TEX, ey0 = ty * TEY;
  const int ex_end = min(ex0 + TEX, prm.mx), ey_end = min(ey0 + TEY, prm.my);
  const int c = (int)(threadIdx.x & 1);  // component of this lane
  const int el = threadIdx.x >> 1;
  const int lx = el % EX, ly = el / EX;
  const int ex = ex0 - 1 + lx, ey = ey0 - 1 + ly;
  const bool valid = ex >= 0 && ey >= 0 && ex < ex_end && ey < ey_end;
  const bool owner = valid && lx >= 1 && ly >= 1;
  const int ox1 = ex_end * P - 1 + (ex_end == prm.mx ? 1 : 0), oy1 = ey_end * P - 1 + (ey_end == prm.my ? 1 : 0);
  const int OX = ox1 - ex0 * P + 1, OY = oy1 - ey0 * P + 1;
  constexpr int LANES = NT < 64 ? NT : 64;
  constexpr int NWAVES = NT / LANES;
  static_assert(NWAVES * LANES == NT, "workgroup must be a multiple of its row width");
  constexpr int RPW = (TEY * P + 1 + NWAVES - 1) / NWAVES;
  constexpr int CPL = (TEX * P + 1 + LANES - 1) / LANES;
  const int wave = threadIdx.x / LANES, lane = threadIdx.x % LANES;
  // A row of EX = 32 elements is one wave: the halo row's wave needs only
  // its top edge and skips the other nodes' force terms.
  constexpr bool HALO_WAVES = EX * 2 == 64;

  double U[NB][NB];  // this lane's displacement component, [j][i]
  double F[NB][NB];  // this lane's force component, [j][i]
  if (valid) {
    const long base = (long)(ey * P) * prm.pitch + ex * P;
    const double *ub = reinterpret_cast<const double *>(prm.u_in + base) + c;
#pragma unroll
    for (int j = 0; j < NB; j++)
#pragma unroll
      for (int i = 0; i < NB; i++) U[j][i] = ub[2 * ((long)j * prm.pitch + i)];
    const int e = ex + ey * prm.mx;
    // ElementEvaluateGeometry/Derivatives (1245-1277) and the stiffness
    // constants (1391-1395)
    const double J00 = 0.5 * prm.elem_dx[ex], J11 = 0.5 * prm.elem_dy[ey];
    const double iJ00 = 1.0 / J00, iJ11 = 1.0 / J11, detJ = J00 * J11;
    const double lam = prm.lambda[e], mu = prm.mu[e];
    const double c11 = 2.0 * mu + lam;
    // x lane: R = sxx = c11 e0 + c12 e1, Cq = sxy; y lane: R = sxy,
    // Cq = syy = c21 e0 + c22 e1 (c12 = c21 = lambda, c22 = c11)
    const double cA = c ? lam : c11, cB = c ? c11 : lam;
    // D is read through the constant address space: uniform, never written
    // by a kernel, so scalar loads.
    using CD = const __attribute__((address_space(4))) double;
    CD *D = (CD *)(uintptr_t)prm.D;
    // uniform per wave when a wave is one element row: a scalar branch
    // The Jacobian factors are made opaque per row and point by an empty asm,
    // so that the compiler forms each row of derivative products where it is
    // used instead of keeping all 2 (p+1)^2 of them live (the reference forms
    // the same products once per element; the bits are the same).
    auto row_factors = [&](int qj, double *dny) {
      double iy = iJ11;
      asm volatile("" : "+v"(iy));
#pragma unroll
      for (int j = 0; j < NB; j++) dny[j] = iy * D[qj * NB + j];
    };
    // Stress at q = (qi, qj): R, Cq of this lane's component (x: sxx, sxy;
    // y: sxy, syy) and -fac; dnx receives dN_dx on row qj.  The products are
    // consumed as they are formed: e1 and the column half of e2 in ascending
    // j, e0 and the row half of e2 in ascending i, e2 in ascending node order.
    auto point_stress = [&](int qi, int qj, const double *dny, double *dnx, double &R, double &Cq, double &nfac) {
      double ix = iJ00;
      asm volatile("" : "+v"(ix));
#pragma unroll
      for (int i = 0; i < NB; i++) dnx[i] = ix * D[qi * NB + i];
      // e2 is summed by the x lane from its own column products of ux and
      // the y lane's row products of uy (one broadcast each), then handed to
      // the y lane; the y lane's running e2 is discarded.
      double srow = 0.0, scol = 0.0, e2 = 0.0;
#pragma unroll
      for (int j = 0; j < qj; j++) {  // column products below the point's row
        const double t = dny[j] * U[j][qi];
        scol = scol + t;
        e2 = e2 + t;
      }
      const double tq = dny[qj] * U[qj][qi];
      scol = scol + tq;
#pragma unroll
      for (int i = 0; i < NB; i++) {  // the point's row
        const double t = dnx[i] * U[qj][i];
        srow = srow + t;
        double r = from_y(t);
        if (i == qi) r = r + tq;
        e2 = e2 + r;
      }
#pragma unroll
      for (int j = qj + 1; j < NB; j++) {  // column products above the row
        const double t = dny[j] * U[j][qi];
        scol = scol + t;
        e2 = e2 + t;
      }
      e2 = from_x(e2);
      // e0 = the x lane's row sum, e1 = the y lane's column sum (1372-1377)
      const double e0 = from_x(srow), e1 = from_y(scol);
      const double T1 = cA * e0 + cB * e1, T2 = mu * e2;
      R = c ? T2 : T1;
      Cq = c ? T1 : T2;
      nfac = -(detJ * sW[qi + qj * NB]);
    };
    // uniform per wave when a wave is one element row: a scalar branch
    if (HALO_WAVES && __builtin_amdgcn_readfirstlane(ly) == 0) {
      // Halo row: only the top edge (j = p) is needed, with the same terms
      // in the same order as the full path below.
      double FT[NB];
#pragma unroll
      for (int i = 0; i < NB; i++) FT[i] = 0.0;
#pragma unroll
      for (int qj = 0; qj < NB; qj++) {
        double dny[NB];
        row_factors(qj, dny);
#pragma unroll
        for (int qi = 0; qi < NB; qi++) {
          double dnx[NB], R, Cq, nfac;
          point_stress(qi, qj, dny, dnx, R, Cq, nfac);
          if (qj == P) {
#pragma unroll
            for (int i = 0; i < NB; i++)
              if (i != qi) FT[i] = FT[i] + nfac * (dnx[i] * R);
            FT[qi] = FT[qi] + nfac * (dnx[qi] * R + dny[qj] * Cq);
          } else {
            FT[qi] = FT[qi] + nfac * (dny[P] * Cq);
          }
        }
      }
#pragma unroll
      for (int k = 0; k < NB; k++) comp64(tedge[ly][lx][k], c) = FT[k];
      comp64(redge[ly][lx][P], c) = FT[P];
    } else {
#pragma unroll
      for (int j = 0; j < NB; j++)
#pragma unroll
        for (int i = 0; i < NB; i++) F[j][i] = 0.0;
#pragma unroll
      for (int qj = 0; qj < NB; qj++) {
        double dny[NB];
        row_factors(qj, dny);
#pragma unroll
        for (int qi = 0; qi < NB; qi++) {
          double dnx[NB], R, Cq, nfac;
          point_stress(qi, qj, dny, dnx, R, Cq, nfac);
          // row terms (nodes (i, qj)), the node (qi, qj), column terms
          // (nodes (qi, j)); each node receives one term per q
#pragma unroll
          for (int i = 0; i < NB; i++)
            if (i != qi) F[qj][i] = F[qj][i] + nfac * (dnx[i] * R);
          F[qj][qi] = F[qj][qi] + nfac * (dnx[qi] * R + dny[qj] * Cq);
#pragma unroll
          for (int j = 0; j < NB; j++)
            if (j != qj) F[j][qi] = F[j][qi] + nfac * (dny[j] * Cq);
        }
      }
      // publish the right and top edges for the neighbouring owners
#pragma unroll
      for (int k = 0; k < NB; k++) {
        comp64(redge[ly][lx][k], c) = F[k][P];
        comp64(tedge[ly][lx][k], c) = F[P][k];
      }
    }
  }
  // Node-phase loads (the predicted u, v and M of the owned nodes): slots
  // past the owned block load its last node (in bounds, unused), so the
  // loads carry no branch.
  double2 pu[RPW][CPL], pv[RPW][CPL];
  double pm[RPW][CPL];
  auto node_loads = [&](int r, int cidx) {
    const int lyn = min(wave + r * NWAVES, OY - 1), lxn = min(lane + cidx * LANES, OX - 1);
    const long idx = (long)(ey0 * P + lyn) * prm.pitch + ex0 * P + lxn;
    pu[r][cidx] = prm.u_in[idx];
    pv[r][cidx] = prm.v[idx];
    pm[r][cidx] = prm.mass[idx];
  };
  __syncthreads();
  const TileLists tl = tile_lists(prm, tile, step, last);
  constexpr int NSL = RPW * CPL, CHUNK = 8;
  if constexpr (FUSED && PF == 0) {
#pragma unroll
    for (int r = 0; r < RPW; r++)
#pragma unroll
      for (int cidx = 0; cidx < CPL; cidx++) node_loads(r, cidx);
  }
  // Owners combine the contributions in ascending element order and store
  // the nodal force in a node-major LDS tile.
  if (owner) {
    const bool left = ex > 0, below = ey > 0;
    const bool rb = ex == prm.mx - 1, tb = ey == prm.my - 1;
#pragma unroll
    for (int j = 0; j < NB; j++) {
#pragma unroll
      for (int i = 0; i < NB; i++) {
        if ((i < P || rb) && (j < P || tb)) {
          const int lxn = (lx - 1) * P + i, lyn = (ly - 1) * P + j;
          double g = F[j][i];
          if (i == 0 && j == 0 && left && below) {  // (ex-1,ey-1), (ex,ey-1), (ex-1,ey), own
            g = ((comp64(redge[ly - 1][lx - 1][P], c) + comp64(tedge[ly - 1][lx][0], c)) +
                 comp64(redge[ly][lx - 1][0], c)) + g;
          } else if (i == 0 && left) {
            g = comp64(redge[ly][lx - 1][j], c) + g;
          } else if (j == 0 && below) {
            g = comp64(tedge[ly - 1][lx][i], c) + g;
          }
          comp64(sf[lyn][lxn + lxn / P], c) = g;
        }
      }
    }
  }
  __syncthreads();
  // Node phase: coalesced update of the owned nodes.
  auto node_step = [&](int r, int cidx) {
    const int lyn = wave + r * NWAVES, lxn = lane + cidx * LANES;
    if (lyn < OY && lxn < OX) {
      const int gx = ex0 * P + lxn, gy = ey0 * P + lyn;
      const long idx = (long)gy * prm.pitch + gx;
      const double2 f = sf[lyn][lxn + lxn / P];
      if constexpr (FUSED)
        node_update_core<P, MODE_F64>(prm, tl, gx, gy, idx, f.x, f.y, pu[r][cidx].x, pu[r][cidx].y, pv[r][cidx],
                                      pm[r][cidx], double2{0.0, 0.0});
      else
        prm.f[idx] = f;
    }
  };
  if constexpr (FUSED && PF == 4) {
#pragma unroll
    for (int c0 = 0; c0 < NSL; c0 += CHUNK) {
#pragma unroll
      for (int sl = c0; sl < c0 + CHUNK && sl < NSL; sl++) node_loads(sl / CPL, sl % CPL);
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

// One lane per element: the lane holds both displacement components and both
// force accumulators of its element and forms every product, sum and force
// term of the reference's loop itself (no cross-lane moves), in the same
// order as the lane-pair body above.  A workgroup of EX x EY threads
// processes one tile of (EX-1) x (EY-1) elements.
template <int P, int EX, int EY, int PF, bool FUSED = true, typename PRM>
__device__ __forceinline__ void reg64s_body(PRM &prm, int tile, int step, int last) {
#pragma clang fp contract(off)
  constexpr int NB = P + 1;
  constexpr int TEX = EX - 1, TEY = EY - 1;
  constexpr int NT = EX * EY;
  __shared__ double2 redge[EY][EX][NB];  // right edge (i = p), j = 0..p
  __shared__ double2 tedge[EY][EX][NB];  // top edge (j = p), i = 0..p
  __shared__ double2 sf[TEY * P + 1][TEX * P + 1 + TEX + 1];
  __shared__ double sW[NB * NB];
  for (int k = threadIdx.x; k < NB * NB; k += NT) sW[k] = prm.w2[k];
  __syncthreads();

  const int tx = tile % prm.tiles_x, ty = tile / prm.tiles_x;
  const int ex0 = tx * TEX, ey0 = ty * TEY;
  const int ex_end = min(ex0 + TEX, prm.mx), ey_end = min(ey0 + TEY, prm.my);
  const int el = threadIdx.x;
  const int lx = el % EX, ly = el / EX;
  const int ex = ex0 - 1 + lx, ey = ey0 - 1 + ly;
  const bool valid = ex >= 0 && ey >= 0 && ex < ex_end && ey < ey_end;
  const bool owner = valid && lx >= 1 && ly >= 1;
  const int ox1 = ex_end * P - 1 + (ex_end == prm.mx ? 1 : 0), oy1 = ey_end * P - 1 + (ey_end == prm.my ? 1 : 0);
  const int OX = ox1 - ex0 * P + 1, OY = oy1 - ey0 * P + 1;
  constexpr int LANES = NT < 64 ? NT : 64;
  constexpr int NWAVES = NT / LANES;
  static_assert(NWAVES * LANES == NT, "workgroup must be a multiple of its row width");
  constexpr int RPW = (TEY * P + 1 + NWAVES - 1) / NWAVES;
  constexpr int CPL = (TEX * P + 1 + LANES - 1) / LANES;
  const int wave = threadIdx.x / LANES, lane = threadIdx.x % LANES;
  // a row of elements that is a whole number of waves: the halo row's waves
  // need only its top edge
  constexpr bool HALO_WAVES = EX % 64 == 0;

  double ux[NB][NB], uy[NB][NB];  // [j][i]
  double fx[NB][NB], fy[NB][NB];
  if (valid) {
    const long base = (long)(ey * P) * prm.pitch + ex * P;
    const double2 *ub = prm.u_in + base;
#pragma unroll
    for (int j = 0; j < NB; j++)
#pragma unroll
      for (int i = 0; i < NB; i++) {
        const double2 t = ub[(long)j * prm.pitch + i];
        ux[j][i] = t.x;
        uy[j][i] = t.y;
      }
    const int e = ex + ey * prm.mx;
    const double J00 = 0.5 * prm.elem_dx[ex], J11 = 0.5 * prm.elem_dy[ey];
    const double iJ00 = 1.0 / J00, iJ11 = 1.0 / J11, detJ = J00 * J11;
    const double lam = prm.lambda[e], mu = prm.mu[e];
    const double c11 = 2.0 * mu + lam;  // = c22; c12 = c21 = lambda, c33 = mu
    using CD = const __attribute__((address_space(4))) double;
    CD *D = (CD *)(uintptr_t)prm.D;
    auto row_factors = [&](int qj, double *dny) {
      double iy = iJ11;
      asm volatile("" : "+v"(iy));
#pragma unroll
      for (int j = 0; j < NB; j++) dny[j] = iy * D[qj * NB + j];
    };
    // sxx, syy, sxy and -fac at q = (qi, qj), 1357-1402; dnx receives
    // dN_dx on row qj.  e1 and the column half of e2 run in ascending j, e0
    // and the row half of e2 in ascending i, e2 in ascending node order.
    auto point_stress = [&](int qi, int qj, const double *dny, double *dnx, double &sxx, double &syy, double &sxy,
                            double &nfac) {
      double ix = iJ00;
      asm volatile("" : "+v"(ix));
#pragma unroll
      for (int i = 0; i < NB; i++) dnx[i] = ix * D[qi * NB + i];
      double e0 = 0.0, e1 = 0.0, e2 = 0.0;
#pragma unroll
      for (int j = 0; j < qj; j++) {
        e1 = e1 + dny[j] * uy[j][qi];
        e2 = e2 + dny[j] * ux[j][qi];
      }
      e1 = e1 + dny[qj] * uy[qj][qi];
      const double tq = dny[qj] * ux[qj][qi];
#pragma unroll
      for (int i = 0; i < NB; i++) {
        e0 = e0 + dnx[i] * ux[qj][i];
        double r = dnx[i] * uy[qj][i];
        if (i == qi) r = r + tq;
        e2 = e2 + r;
      }
#pragma unroll
      for (int j = qj + 1; j < NB; j++) {
        e1 = e1 + dny[j] * uy[j][qi];
        e2 = e2 + dny[j] * ux[j][qi];
      }
      sxx = c11 * e0 + lam * e1;
      syy = lam * e0 + c11 * e1;
      sxy = mu * e2;
      nfac = -(detJ * sW[qi + qj * NB]);
    };
    if (HALO_WAVES && __builtin_amdgcn_readfirstlane(ly) == 0) {
      double gx[NB], gy[NB];  // top edge only
#pragma unroll
      for (int i = 0; i < NB; i++) gx[i] = gy[i] = 0.0;
#pragma unroll
      for (int qj = 0; qj < NB; qj++) {
        double dny[NB];
        row_factors(qj, dny);
#pragma unroll
        for (int qi = 0; qi < NB; qi++) {
          double dnx[NB], sxx, syy, sxy, nfac;
          point_stress(qi, qj, dny, dnx, sxx, syy, sxy, nfac);
          if (qj == P) {
#pragma unroll
            for (int i = 0; i < NB; i++)
              if (i != qi) {
                gx[i] = gx[i] + nfac * (dnx[i] * sxx);
                gy[i] = gy[i] + nfac * (dnx[i] * sxy);
              }
            gx[qi] = gx[qi] + nfac * (dnx[qi] * sxx + dny[qj] * sxy);
            gy[qi] = gy[qi] + nfac * (dny[qj] * syy + dnx[qi] * sxy);
          } else {
            gx[qi] = gx[qi] + nfac * (dny[P] * sxy);
            gy[qi] = gy[qi] + nfac * (dny[P] * syy);
          }
        }
      }
#pragma unroll
      for (int k = 0; k < NB; k++) tedge[ly][lx][k] = double2{gx[k], gy[k]};
      redge[ly][lx][P] = double2{gx[P], gy[P]};
    } else {
#pragma unroll
      for (int j = 0; j < NB; j++)
#pragma unroll
        for (int i = 0; i < NB; i++) fx[j][i] = fy[j][i] = 0.0;
#pragma unroll
      for (int qj = 0; qj < NB; qj++) {
        double dny[NB];
        row_factors(qj, dny);
#pragma unroll
        for (int qi = 0; qi < NB; qi++) {
          double dnx[NB], sxx, syy, sxy, nfac;
          point_stress(qi, qj, dny, dnx, sxx, syy, sxy, nfac);
#pragma unroll
          for (int i = 0; i < NB; i++)
            if (i != qi) {
              fx[qj][i] = fx[qj][i] + nfac * (dnx[i] * sxx);
              fy[qj][i] = fy[qj][i] + nfac * (dnx[i] * sxy);
            }
          fx[qj][qi] = fx[qj][qi] + nfac * (dnx[qi] * sxx + dny[qj] * sxy);
          fy[qj][qi] = fy[qj][qi] + nfac * (dny[qj] * syy + dnx[qi] * sxy);
#pragma unroll
          for (int j = 0; j < NB; j++)
            if (j != qj) {
              fx[j][qi] = fx[j][qi] + nfac * (dny[j] * sxy);
              fy[j][qi] = fy[j][qi] + nfac * (dny[j] * syy);
            }
        }
      }
#pragma unroll
      for (int k = 0; k < NB; k++) {
        redge[ly][lx][k] = double2{fx[k][P], fy[k][P]};
        tedge[ly][lx][k] = double2{fx[P][k], fy[P][k]};
      }
    }
  }
  double2 pu[RPW][CPL], pv[RPW][CPL];
  double pm[RPW][CPL];
  auto node_loads = [&](int r, int cidx) {
    const int lyn = min(wave + r * NWAVES, OY - 1), lxn = min(lane + cidx * LANES, OX - 1);
    const long idx = (long)(ey0 * P + lyn) * prm.pitch + ex0 * P + lxn;
    pu[r][cidx] = prm.u_in[idx];
    pv[r][cidx] = prm.v[idx];
    pm[r][cidx] = prm.mass[idx];
  };
  __syncthreads();
  const TileLists tl = tile_lists(prm, tile, step, last);
  constexpr int NSL = RPW * CPL, CHUNK = 8;
  if constexpr (FUSED && PF == 0) {
#pragma unroll
    for (int r = 0; r < RPW; r++)
#pragma unroll
      for (int cidx = 0; cidx < CPL; cidx++) node_loads(r, cidx);
  }
  if (owner) {
    const bool left = ex > 0, below = ey > 0;
    const bool rb = ex == prm.mx - 1, tb = ey == prm.my - 1;
#pragma unroll
    for (int j = 0; j < NB; j++) {
#pragma unroll
      for (int i = 0; i < NB; i++) {
        if ((i < P || rb) && (j < P || tb)) {
          const int lxn = (lx - 1) * P + i, lyn = (ly - 1) * P + j;
          double gx = fx[j][i], gy = fy[j][i];
          if (i == 0 && j == 0 && left && below) {  // (ex-1,ey-1), (ex,ey-1), (ex-1,ey), own
            const double2 d = redge[ly - 1][lx - 1][P], b = tedge[ly - 1][lx][0], l = redge[ly][lx - 1][0];
            gx = ((d.x + b.x) + l.x) + gx;
            gy = ((d.y + b.y) + l.y) + gy;
          } else if (i == 0 && left) {
            const double2 l = redge[ly][lx - 1][j];
            gx = l.x + gx;
            gy = l.y + gy;
          } else if (j == 0 && below) {
            const double2 b = tedge[ly - 1][lx][i];
            gx = b.x + gx;
            gy = b.y + gy;
          }
          sf[lyn][lxn + lxn / P] = double2{gx, gy};
        }
      }
    }
  }
  __syncthreads();
  auto node_step = [&](int r, int cidx) {
    const int lyn = wave + r * NWAVES, lxn = lane + cidx * LANES;
    if (lyn < OY && lxn < OX) {
      const int gx = ex0 * P + lxn, gy = ey0 * P + lyn;
      const long idx = (long)gy * prm.pitch + gx;
      const double2 f = sf[lyn][lxn + lxn / P];
      if constexpr (FUSED)
        node_update_core<P, MODE_F64>(prm, tl, gx, gy, idx, f.x, f.y, pu[r][cidx].x, pu[r][cidx].y, pv[r][cidx],
                                      pm[r][cidx], double2{0.0, 0.0});
      else
        prm.f[idx] = f;
    }
  };
  if constexpr (FUSED && PF == 4) {
#pragma unroll
    for (int c0 = 0; c0 < NSL; c0 += CHUNK) {
#pragma unroll
      for (int sl = c0; sl < c0 + CHUNK && sl < NSL; sl++) node_loads(sl / CPL, sl % CPL);
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

template <int P, int EX, int EY, int WAVES, int PF, int LPE>
__global__ __launch_bounds__(EX * EY * LPE) __attribute__((amdgpu_waves_per_eu(WAVES > 0 ? WAVES : 1, WAVES == 1 ? 1 : 8)))
void reg64_kernel(StepParams<double> prm) {
  const int tile = tile_of_block(blockIdx.x, prm.tiles_per_xcd);
  if (tile >= prm.tiles_x * prm.tiles_y) return;  // uniform per workgroup
  if constexpr (LPE == 2) reg64_body<P, EX, EY, PF>(prm, tile, prm.step, prm.last);
  else reg64s_body<P, EX, EY, PF>(prm, tile, prm.step, prm.last);
}

// f = K(u) alone (the split schedule's element kernel and
// se2w_assemble_forces): the same element operator and owner sums, the
// nodal force stored instead of the update.
template <int P, int EX, int EY, int WAVES, int PF, int LPE>
__global__ __launch_bounds__(EX * EY * LPE) __attribute__((amdgpu_waves_per_eu(WAVES > 0 ? WAVES : 1, WAVES == 1 ? 1 : 8)))
void reg64_element_kernel(StepParams<double> prm) {
  const int tile = tile_of_block(blockIdx.x, prm.tiles_per_xcd);
  if (tile >= prm.tiles_x * prm.tiles_y) return;  // uniform per workgroup
  if constexpr (LPE == 2) reg64_body<P, EX, EY, PF, false>(prm, tile, 0, 0);
  else reg64s_body<P, EX, EY, PF, false>(prm, tile, 0, 0);
}

}  // namespace se2w
