// se2w_ens.hip -- batched (ensemble) step kernels: one workgroup per tile of
// every member; each workgroup runs the single-instance tile body with the
// step parameters of the member that owns the tile.  Compiled once per
// precision mode with -DSE2W_MODE_TU=<mode>.
#include "se2w_ens.hpp"

#include "se2w.h"
#include "se2w_kernels_reg.hpp"
#include "se2w_kernels_reg64.hpp"
#include "se2w_kernels_mfma.hpp"

#ifndef SE2W_MODE_TU
#error "compile with -DSE2W_MODE_TU=<mode>"
#endif

namespace se2w {

// Workgroup b -> global tile t of the ensemble -> (member, member tile).  The
// grid has one workgroup per tile (8 * ceil(total / 8)); workgroup b runs on
// XCD b % 8, and the eight XCDs each take one contiguous eighth of the tile
// range, so neighbouring tiles of one member, which re-read each other's halo
// elements, share an L2.  One tile per workgroup, with no loop over tiles:
// the dispatcher starts a workgroup as soon as a slot frees, and no value
// derived from the thread index is kept live from one tile to the next (a
// persistent grid walking the tiles measured 10 % slower at p = 6, with 20 %
// more VGPRs for the same tile body).
template <typename R, typename Body>
__device__ __forceinline__ void ens_tile(const EnsArgs<R> &a, Body body) {
  const int xcd = blockIdx.x & 7, per = (a.total + 7) >> 3;
  const int t = xcd * per + (blockIdx.x >> 3);
  if (t >= min((xcd + 1) * per, a.total)) return;  // uniform per workgroup
  int lo = 0, hi = a.n - 1;  // largest i with tile_off[i] <= t
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (a.tile_off[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  const int i = __builtin_amdgcn_readfirstlane(lo);
  // The member's parameters are read in place through the constant address
  // space (the table is written before the run and never by a kernel): each
  // field is a scalar load where the body uses it, and the pointers loaded
  // from it address global memory.  A copy of the ~64-dword block into
  // registers instead held it in SGPRs for the whole tile (spilling them
  // beside the derivative table) and lost the pointers' address space, so
  // every state access became a flat load or store.
  using CP = const __attribute__((address_space(4))) StepParams<R>;
  CP &prm = *(CP *)(uintptr_t)(a.prms + 2 * i + (a.step & 1));
  body(prm, t - a.tile_off[i], a.step, a.last);
}

template <int P, int MODE, int NT>
__global__ __launch_bounds__(NT) void ens_tile_kernel(EnsArgs<typename RealOf<MODE>::T> a) {
  ens_tile(a, [](auto &prm, int tile, int step, int last) { tile_body<P, MODE, true, NT>(prm, tile, step, last); });
}

template <int P, int MODE, int EX, int EY, int WAVES, int PF, int LPE>
__global__ __launch_bounds__(EX * EY * LPE) __attribute__((amdgpu_waves_per_eu(WAVES > 0 ? WAVES : 1, 8)))
void ens_reg_kernel(EnsArgs<float> a) {
  ens_tile(a, [](auto &prm, int tile, int step, int last) {
    reg_body<P, MODE, true, EX, EY, PF, LPE>(prm, tile, step, last);
  });
}

template <int P, int EX, int EY, int WAVES, int PF, int LPE>
__global__ __launch_bounds__(EX * EY * LPE) __attribute__((amdgpu_waves_per_eu(WAVES > 0 ? WAVES : 1, WAVES == 1 ? 1 : 8)))
void ens_reg64_kernel(EnsArgs<double> a) {
  ens_tile(a, [](auto &prm, int tile, int step, int last) {
    if constexpr (LPE == 2) reg64_body<P, EX, EY, PF>(prm, tile, step, last);
    else reg64s_body<P, EX, EY, PF>(prm, tile, step, last);
  });
}

template <int P, int MODE>
__global__ __launch_bounds__(kThreads) void ens_mfma_kernel(EnsArgs<float> a) {
  ens_tile(a, [](auto &prm, int tile, int step, int last) { mfma_body<P, MODE, false>(prm, tile, step, last); });
}

template <int P, int MODE>
__global__ __launch_bounds__(kThreads) void ens_line_kernel(EnsArgs<float> a) {
  ens_tile(a, [](auto &prm, int tile, int step, int last) { mfma_body<P, MODE, true>(prm, tile, step, last); });
}

namespace {

template <int P, int MODE, int S>
int reg_dispatch(int sched, const void *args, int grid, hipStream_t s, int *bpc, int *thr) {
  if constexpr (S < kNumRegSchedules) {
    if constexpr (reg_schedule_built(P, S, true)) {
      if (sched == S) {
        constexpr RegSchedule r = kRegSchedules[S];
        auto k = ens_reg_kernel<P, MODE, r.ex, r.ey, r.waves, r.pf, r.lpe>;
        const int nt = r.ex * r.ey * r.lpe;
        if (!args) {
          int n = 0;
          if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, nt, 0) != hipSuccess) return -1;
          *bpc = n;
          *thr = nt;
          return 0;
        }
        hipLaunchKernelGGL(k, dim3(grid), dim3(nt), 0, s, *(const EnsArgs<float> *)args);
        return 0;
      }
    }
    return reg_dispatch<P, MODE, S + 1>(sched, args, grid, s, bpc, thr);
  }
  return -1;
}

template <int P, int S>
int reg64_dispatch(int sched, const void *args, int grid, hipStream_t s, int *bpc, int *thr) {
  if constexpr (S < kNumReg64Schedules) {
    if constexpr (reg64_schedule_built(P, S, true)) {
      if (sched == S) {
        constexpr Reg64Schedule r = kReg64Schedules[S];
        auto k = ens_reg64_kernel<P, r.ex, r.ey, r.waves, r.pf, r.lpe>;
        const int nt = r.ex * r.ey * r.lpe;
        if (!args) {
          int n = 0;
          if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, nt, 0) != hipSuccess) return -1;
          *bpc = n;
          *thr = nt;
          return 0;
        }
        hipLaunchKernelGGL(k, dim3(grid), dim3(nt), 0, s, *(const EnsArgs<double> *)args);
        return 0;
      }
    }
    return reg64_dispatch<P, S + 1>(sched, args, grid, s, bpc, thr);
  }
  return -1;
}

template <int P, int MODE>
int dispatch_pm(int variant, const void *args, int grid, hipStream_t s, int *bpc, int *thr) {
  using R = typename RealOf<MODE>::T;
  if (variant == SE2W_KERNEL_LDS || variant == SE2W_KERNEL_LDS_WIDE) {
    const int nt = variant == SE2W_KERNEL_LDS ? kThreads : 512;
    auto k = variant == SE2W_KERNEL_LDS ? ens_tile_kernel<P, MODE, kThreads> : ens_tile_kernel<P, MODE, 512>;
    if (!args) {
      int n = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, nt, 0) != hipSuccess) return -1;
      *bpc = n;
      *thr = nt;
      return 0;
    }
    hipLaunchKernelGGL(k, dim3(grid), dim3(nt), 0, s, *(const EnsArgs<R> *)args);
    return 0;
  }
  if constexpr (MODE != MODE_F64 && P >= 6) {
    if (variant == SE2W_KERNEL_MFMA || variant == SE2W_KERNEL_LINE) {
      auto k = variant == SE2W_KERNEL_MFMA ? ens_mfma_kernel<P, MODE> : ens_line_kernel<P, MODE>;
      if (!args) {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, kThreads, 0) != hipSuccess) return -1;
        *bpc = n;
        *thr = kThreads;
        return 0;
      }
      hipLaunchKernelGGL(k, dim3(grid), dim3(kThreads), 0, s, *(const EnsArgs<float> *)args);
      return 0;
    }
  }
  if constexpr (MODE != MODE_F64 && P >= 2)
    return reg_dispatch<P, MODE, 0>(variant - SE2W_KERNEL_REG, args, grid, s, bpc, thr);
  if constexpr (MODE == MODE_F64 && P >= 2)
    return reg64_dispatch<P, 0>(variant - SE2W_KERNEL_REG64, args, grid, s, bpc, thr);
  return -1;
}

}  // namespace

template <>
int ens_dispatch<SE2W_MODE_TU>(int p, int variant, const void *args, int grid, hipStream_t s, int *bpc, int *thr) {
  constexpr int M = SE2W_MODE_TU;
  switch (p) {
    case 1: return dispatch_pm<1, M>(variant, args, grid, s, bpc, thr);
    case 2: return dispatch_pm<2, M>(variant, args, grid, s, bpc, thr);
    case 3: return dispatch_pm<3, M>(variant, args, grid, s, bpc, thr);
    case 4: return dispatch_pm<4, M>(variant, args, grid, s, bpc, thr);
    case 5: return dispatch_pm<5, M>(variant, args, grid, s, bpc, thr);
    case 6: return dispatch_pm<6, M>(variant, args, grid, s, bpc, thr);
    case 7: return dispatch_pm<7, M>(variant, args, grid, s, bpc, thr);
    case 8: return dispatch_pm<8, M>(variant, args, grid, s, bpc, thr);
    default: return -1;
  }
}

}  // namespace se2w
