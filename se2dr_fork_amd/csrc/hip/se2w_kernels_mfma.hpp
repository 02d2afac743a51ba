// se2w_kernels_mfma.hpp -- fused FP32 step for high degrees (6 <= p <= 8)
// with the element contractions on the matrix cores.
//
// The tile, its halo and the node update are those of tile_body; the eight
// sum-factorised contractions of the element operator run as batched small
// matrix products on v_mfma_f32_16x16x4_f32 (FP32 in, FP32 accumulate):
//
//   B1  dxi u (row, el, qi)  = sum_k D[qi][k] u(row, el P + k)     M = qi, N = 16 node rows
//   B2  deta u (er, qj, col) = sum_k D[qj][k] u(er P + k, col)     M = qj, N = 16 node columns
//   C1  R(el, i, qrow)      = sum_qi D[qi][i] S(qrow, el, qi)     M = i,  N = 16 quadrature rows
//   C2  C(er, j, qcol)      = sum_qj D[qj][j] S(er, qj, qcol)     M = j,  N = 16 quadrature columns
//
// with K = the p+1 contraction points padded to 12 (three k-steps of 4).
// Every node row or column of the tile is one N column of a product, so a
// product covers whole rows of elements at once and the tiny D (or D^T) is
// the constant A operand, held in three registers per lane.
//
// MFMA 16x16x4 f32 operand layout (lane l of 64): A[m][k] at m = l % 16,
// k = l / 16; B[k][n] at k = l / 16, n = l % 16; result D[m][n] in
// accumulator r at m = 4 (l / 16) + r, n = l % 16.
#pragma once

#include "se2w_kernels.hpp"

namespace se2w {

typedef __attribute__((__vector_size__(4 * sizeof(float)))) float mfma_f4;

// Tile shapes (owned elements TEX x TEY) keeping the LDS of one workgroup
// under 80 KB, so that two fit a CU.
template <int P> struct MfmaShape;
template <> struct MfmaShape<6> { static constexpr int TEX = 6, TEY = 5; };
template <> struct MfmaShape<7> { static constexpr int TEX = 5, TEY = 5; };
template <> struct MfmaShape<8> { static constexpr int TEX = 4, TEY = 4; };

template <int P>
struct MfmaLayout {
  static constexpr int NB = P + 1;
  static constexpr int TEX = MfmaShape<P>::TEX, TEY = MfmaShape<P>::TEY;
  static constexpr int NXL = (TEX + 1) * P + 1, NYL = (TEY + 1) * P + 1;  // tile nodes (+ halo)
  static constexpr int QX = (TEX + 1) * NB, QY = (TEY + 1) * NB;          // quadrature points (+ halo)
  // LDS: the weighted stresses S, and a work area that holds first the
  // derivatives (Dx, Dy), then the transposed contractions (R, C)
  static constexpr int DXY = 2 * NYL * QX + 2 * QY * NXL;
  static constexpr int SS = 3 * QY * QX;
  static constexpr int RC = 4 * QX * QY;
  static constexpr int WORK = DXY > RC ? DXY : RC;
  // products per phase and per wave (4 waves)
  static constexpr int JB1 = 2 * (TEX + 1) * ((NYL + 15) / 16);
  static constexpr int JB2 = 2 * (TEY + 1) * ((NXL + 15) / 16);
  static constexpr int JC1 = 2 * (TEX + 1) * ((QY + 15) / 16);
  static constexpr int JC2 = 2 * (TEY + 1) * ((QX + 15) / 16);
  static constexpr int JC = JC1 + JC2;
  static constexpr int PTS = QX * QY;
};

// The products of one phase: jobs wave, wave + 4, ... (4 waves), taken G at
// a time with independent accumulators so that the LDS loads of the B
// operands and the matrix-core latencies of the G products overlap.
// fetch(job, s) is this lane's B operand of k-step s; store(job, acc) writes
// this lane's four results.  `job < NJOB` is uniform across a wave.
template <int NJOB, int G, typename Fetch, typename Store>
__device__ __forceinline__ void mfma_phase(int wave, const float (&A)[3], Fetch fetch, Store store) {
  for (int base = wave; base < NJOB; base += 4 * G) {
    float b[G][3];
#pragma unroll
    for (int u = 0; u < G; u++) {
      const int job = base + 4 * u;
#pragma unroll
      for (int s = 0; s < 3; s++) b[u][s] = job < NJOB ? fetch(job, s) : 0.f;
    }
    mfma_f4 acc[G];
#pragma unroll
    for (int u = 0; u < G; u++) acc[u] = mfma_f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < 3; s++) {
#pragma unroll
      for (int u = 0; u < G; u++) acc[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[s], b[u][s], acc[u], 0, 0, 0);
    }
#pragma unroll
    for (int u = 0; u < G; u++) {
      const int job = base + 4 * u;
      if (job < NJOB) store(job, acc[u]);
    }
  }
}

// The same products on the vector ALUs, one line (one N column) per thread:
// out[m] = sum_k A[m][k] in[k] for m, k < NB, with A = D or D^T read from the
// uniform table (scalar loads) and the NB inputs and outputs in registers.
template <int NB, int NLINES, bool TRANS, typename Load, typename Store>
__device__ __forceinline__ void line_phase(const float *__restrict__ D, Load load, Store store) {
  for (int line = threadIdx.x; line < NLINES; line += kThreads) {
    float in[NB];
#pragma unroll
    for (int k = 0; k < NB; k++) in[k] = load(line, k);
#pragma unroll
    for (int m = 0; m < NB; m++) {
      float acc = 0.f;
#pragma unroll
      for (int k = 0; k < NB; k++) acc = fmaf(TRANS ? D[k * NB + m] : D[m * NB + k], in[k], acc);
      store(line, m, acc);
    }
  }
}

template <int P, int MODE, bool LINES, typename PRM>
__device__ __forceinline__ void mfma_body(PRM &prm, int tile, int step, int last) {
  // LINES = false: contractions on the matrix cores; true: on the vector ALUs,
  // one line per thread (same data layout and phases).
  static_assert(MODE == MODE_F32 || MODE == MODE_F32K, "FP32 kernel");
  using Lay = MfmaLayout<P>;
  constexpr int NB = Lay::NB, TEX = Lay::TEX, TEY = Lay::TEY;
  constexpr int NXL = Lay::NXL, NYL = Lay::NYL, QX = Lay::QX, QY = Lay::QY;
  constexpr int NE = (TEX + 1) * (TEY + 1);
  static_assert(NE <= kThreads, "one thread per tile element");
  __shared__ float su[2][NYL][NXL];
  __shared__ float econ[5][NE];        // 1/J00, 1/J11, J00 J11, lambda, mu (0 outside the mesh)
  __shared__ float sW[NB * NB];
  __shared__ float S[Lay::SS];         // [3][QY][QX]: sxx, syy, sxy times detJ w_q
  __shared__ float work[Lay::WORK];
  float *Dx = work;                    // [2][NYL][QX]
  float *Dy = work + 2 * NYL * QX;     // [2][QY][NXL]
  float *Rr = work;                    // [2][QX][QY]
  float *Cc = work + 2 * QX * QY;      // [2][QY][QX]

  const int tx = tile % prm.tiles_x, ty = tile / prm.tiles_x;
  const int ex0 = tx * TEX, ey0 = ty * TEY;
  const int gx0 = (ex0 - 1) * P, gy0 = (ey0 - 1) * P;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int lm = lane & 15, lk = lane >> 4;

  // A operands: D[m][k] and D^T[m][k] = D[k][m] for the three k-steps
  float Ad[3], At[3];
#pragma unroll
  for (int s = 0; s < 3; s++) {
    const int k = 4 * s + lk;
    Ad[s] = (lm < NB && k < NB) ? prm.D[lm * NB + k] : 0.f;
    At[s] = (lm < NB && k < NB) ? prm.D[k * NB + lm] : 0.f;
  }
  // Phase A: displacement of the tile's nodes (+ halo) into LDS; per-element
  // constants (inverse Jacobians, determinant, Lame parameters).
  for (int k = tid; k < NXL * NYL; k += kThreads) {
    const int ly = k / NXL, lx = k - ly * NXL;
    const int gx = gx0 + lx, gy = gy0 + ly;
    float2 uv = float2{0.f, 0.f};
    if (gx >= 0 && gx < prm.nx && gy >= 0 && gy < prm.ny) uv = prm.u_in[(long)gy * prm.pitch + gx];
    su[0][ly][lx] = uv.x;
    su[1][ly][lx] = uv.y;
  }
  if (tid < NE) {
    const int el = tid % (TEX + 1), er = tid / (TEX + 1);
    const int ex = ex0 - 1 + el, ey = ey0 - 1 + er;
    float c0 = 0.f, c1 = 0.f, c2 = 0.f, c3 = 0.f, c4 = 0.f;
    if (ex >= 0 && ex < prm.mx && ey >= 0 && ey < prm.my) {
      const float J00 = 0.5f * prm.elem_dx[ex], J11 = 0.5f * prm.elem_dy[ey];
      const int e = ex + ey * prm.mx;
      c0 = 1.f / J00;
      c1 = 1.f / J11;
      c2 = J00 * J11;
      c3 = prm.lambda[e];
      c4 = prm.mu[e];
    }
    econ[0][tid] = c0;
    econ[1][tid] = c1;
    econ[2][tid] = c2;
    econ[3][tid] = c3;
    econ[4][tid] = c4;
  }
  for (int k = tid; k < NB * NB; k += kThreads) sW[k] = prm.w2[k];
  __syncthreads();
  // B1: xi-derivatives along every node row, per element column.
  if constexpr (LINES) {
    line_phase<NB, 2 * (TEX + 1) * NYL, false>(
        prm.D,
        [&](int line, int k) {
          const int row = line % NYL, ce = line / NYL, el = ce % (TEX + 1), c = ce / (TEX + 1);
          return su[c][row][el * P + k];
        },
        [&](int line, int m, float v) {
          const int row = line % NYL, ce = line / NYL, el = ce % (TEX + 1), c = ce / (TEX + 1);
          Dx[(c * NYL + row) * QX + el * NB + m] = v;
        });
  } else {
  mfma_phase<Lay::JB1, 4>(
      wave, Ad,
      [&](int job, int s) {
        const int c = job & 1, rest = job >> 1;
        const int el = rest % (TEX + 1), row = 16 * (rest / (TEX + 1)) + lm, k = 4 * s + lk;
        return (k < NB && row < NYL) ? su[c][row][el * P + k] : 0.f;
      },
      [&](int job, const mfma_f4 &acc) {
        const int c = job & 1, rest = job >> 1;
        const int el = rest % (TEX + 1), row = 16 * (rest / (TEX + 1)) + lm;
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int qi = 4 * lk + r;
          if (qi < NB && row < NYL) Dx[(c * NYL + row) * QX + el * NB + qi] = acc[r];
        }
      });
  }
  // B2: eta-derivatives along every node column, per element row.
  if constexpr (LINES) {
    line_phase<NB, 2 * (TEY + 1) * NXL, false>(
        prm.D,
        [&](int line, int k) {
          const int col = line % NXL, ce = line / NXL, er = ce % (TEY + 1), c = ce / (TEY + 1);
          return su[c][er * P + k][col];
        },
        [&](int line, int m, float v) {
          const int col = line % NXL, ce = line / NXL, er = ce % (TEY + 1), c = ce / (TEY + 1);
          Dy[(c * QY + er * NB + m) * NXL + col] = v;
        });
  } else {
  mfma_phase<Lay::JB2, 4>(
      wave, Ad,
      [&](int job, int s) {
        const int c = job & 1, rest = job >> 1;
        const int er = rest % (TEY + 1), col = 16 * (rest / (TEY + 1)) + lm, k = 4 * s + lk;
        return (k < NB && col < NXL) ? su[c][er * P + k][col] : 0.f;
      },
      [&](int job, const mfma_f4 &acc) {
        const int c = job & 1, rest = job >> 1;
        const int er = rest % (TEY + 1), col = 16 * (rest / (TEY + 1)) + lm;
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int qj = 4 * lk + r;
          if (qj < NB && col < NXL) Dy[(c * QY + er * NB + qj) * NXL + col] = acc[r];
        }
      });
  }
  __syncthreads();
  // B3: weighted stresses at the quadrature points.
  for (int pnt = tid; pnt < Lay::PTS; pnt += kThreads) {
    const int qrow = pnt / QX, qcol = pnt - qrow * QX;
    const int er = qrow / NB, qj = qrow - er * NB, el = qcol / NB, qi = qcol - el * NB;
    const int row = er * P + qj, col = el * P + qi, ei = er * (TEX + 1) + el;
    const float dxux = Dx[(0 * NYL + row) * QX + qcol], dxuy = Dx[(1 * NYL + row) * QX + qcol];
    const float dyux = Dy[(0 * QY + qrow) * NXL + col], dyuy = Dy[(1 * QY + qrow) * NXL + col];
    const float iJ00 = econ[0][ei], iJ11 = econ[1][ei], lam = econ[3][ei], mu = econ[4][ei];
    const float fac = econ[2][ei] * sW[qi + qj * NB], c11 = 2.f * mu + lam;  // zero outside the mesh
    const float e0 = iJ00 * dxux, e1 = iJ11 * dyuy, e2 = iJ11 * dyux + iJ00 * dxuy;
    S[0 * QY * QX + pnt] = fac * (c11 * e0 + lam * e1);
    S[1 * QY * QX + pnt] = fac * (lam * e0 + c11 * e1);
    S[2 * QY * QX + pnt] = fac * (mu * e2);
  }
  __syncthreads();  // S complete; the derivatives are dead, R and C reuse their space
  // C1: row terms R(c, el, i, qrow); c = 0 from sxx (x force), c = 1 from sxy (y force)
  if constexpr (LINES) {
    line_phase<NB, 2 * (TEX + 1) * QY, true>(
        prm.D,
        [&](int line, int k) {
          const int qrow = line % QY, ce = line / QY, el = ce % (TEX + 1), c = ce / (TEX + 1);
          return S[(c == 0 ? 0 : 2) * QY * QX + qrow * QX + el * NB + k];
        },
        [&](int line, int m, float v) {
          const int qrow = line % QY, ce = line / QY, el = ce % (TEX + 1), c = ce / (TEX + 1);
          Rr[(c * QX + el * NB + m) * QY + qrow] = v;
        });
  } else {
  mfma_phase<Lay::JC1, 4>(
      wave, At,
      [&](int job, int s) {
        const int c = job & 1, rest = job >> 1;
        const int el = rest % (TEX + 1), qrow = 16 * (rest / (TEX + 1)) + lm, k = 4 * s + lk;
        const float *Src = S + (c == 0 ? 0 : 2) * QY * QX;
        return (k < NB && qrow < QY) ? Src[qrow * QX + el * NB + k] : 0.f;
      },
      [&](int job, const mfma_f4 &acc) {
        const int c = job & 1, rest = job >> 1;
        const int el = rest % (TEX + 1), qrow = 16 * (rest / (TEX + 1)) + lm;
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int i = 4 * lk + r;
          if (i < NB && qrow < QY) Rr[(c * QX + el * NB + i) * QY + qrow] = acc[r];
        }
      });
  }
  // C2: column terms C(c, er, j, qcol); c = 0 from sxy (x force), c = 1 from syy (y force)
  if constexpr (LINES) {
    line_phase<NB, 2 * (TEY + 1) * QX, true>(
        prm.D,
        [&](int line, int k) {
          const int qcol = line % QX, ce = line / QX, er = ce % (TEY + 1), c = ce / (TEY + 1);
          return S[(c == 0 ? 2 : 1) * QY * QX + (er * NB + k) * QX + qcol];
        },
        [&](int line, int m, float v) {
          const int qcol = line % QX, ce = line / QX, er = ce % (TEY + 1), c = ce / (TEY + 1);
          Cc[(c * QY + er * NB + m) * QX + qcol] = v;
        });
  } else {
  mfma_phase<Lay::JC2, 4>(
      wave, At,
      [&](int job, int s) {
        const int c = job & 1, rest = job >> 1;
        const int er = rest % (TEY + 1), qcol = 16 * (rest / (TEY + 1)) + lm, k = 4 * s + lk;
        const float *Src = S + (c == 0 ? 2 : 1) * QY * QX;
        return (k < NB && qcol < QX) ? Src[(er * NB + k) * QX + qcol] : 0.f;
      },
      [&](int job, const mfma_f4 &acc) {
        const int c = job & 1, rest = job >> 1;
        const int er = rest % (TEY + 1), qcol = 16 * (rest / (TEY + 1)) + lm;
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int j = 4 * lk + r;
          if (j < NB && qcol < QX) Cc[(c * QY + er * NB + j) * QX + qcol] = acc[r];
        }
      });
  }
  __syncthreads();
  // D: owned nodes: element contributions in ascending element order, update.
  const TileLists tl = tile_lists(prm, tile, step, last);
  const int ex_end = min(ex0 + TEX, prm.mx), ey_end = min(ey0 + TEY, prm.my);
  const int OX = (ex_end - ex0) * P + (ex_end == prm.mx ? 1 : 0);
  const int OY = (ey_end - ey0) * P + (ey_end == prm.my ? 1 : 0);
  for (int k = tid; k < OX * OY; k += kThreads) {
    int ox, oy;
    owned_node<P>(k, OX, OY, ox, oy);  // class order: uniform element count per wave
    const int gx = ex0 * P + ox, gy = ey0 * P + oy;
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
    float fx = 0.f, fy = 0.f;
    for (int b = 0; b < ney; b++) {
      for (int a = 0; a < nex; a++) {
        const int ex = exs[a], ey = eys[b];
        const int el = ex - ex0 + 1, er = ey - ey0 + 1;
        const int i = gx - ex * P, j = gy - ey * P;
        const int ei = er * (TEX + 1) + el;
        const float iJ00 = econ[0][ei], iJ11 = econ[1][ei];
        const int qr = er * NB + j, qc = el * NB + i;
        fx += -(iJ00 * Rr[(0 * QX + qc) * QY + qr] + iJ11 * Cc[(0 * QY + qr) * QX + qc]);
        fy += -(iJ00 * Rr[(1 * QX + qc) * QY + qr] + iJ11 * Cc[(1 * QY + qr) * QX + qc]);
      }
    }
    const int lx = gx - gx0, ly = gy - gy0;
    node_update<P, MODE>(prm, tl, gx, gy, fx, fy, su[0][ly][lx], su[1][ly][lx]);
  }
}

template <int P, int MODE>
__global__ __launch_bounds__(kThreads) void mfma_kernel(StepParams<float> prm) {
  const int tile = tile_of_block(blockIdx.x, prm.tiles_per_xcd);
  if (tile >= prm.tiles_x * prm.tiles_y) return;  // uniform per workgroup
  mfma_body<P, MODE, false>(prm, tile, prm.step, prm.last);
}

template <int P, int MODE>
__global__ __launch_bounds__(kThreads) void line_kernel(StepParams<float> prm) {
  const int tile = tile_of_block(blockIdx.x, prm.tiles_per_xcd);
  if (tile >= prm.tiles_x * prm.tiles_y) return;  // uniform per workgroup
  mfma_body<P, MODE, true>(prm, tile, prm.step, prm.last);
}

}  // namespace se2w
