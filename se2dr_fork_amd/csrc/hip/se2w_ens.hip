// se2w_ens.hip -- batched (ensemble) step kernels: a persistent grid walks the
// tiles of all members; each workgroup runs the single-instance tile bodies
// with the step parameters of the member that owns the tile.  Compiled once
// per precision mode with -DSE2W_MODE_TU=<mode>.
#include "se2w_ens.hpp"

#include "se2w.h"
#include "se2w_kernels_reg.hpp"
#include "se2w_kernels_mfma.hpp"

#ifndef SE2W_MODE_TU
#error "compile with -DSE2W_MODE_TU=<mode>"
#endif

namespace se2w {

// Global tile t of the ensemble -> (member, member tile), walked by a
// persistent grid.  The eight XCDs (workgroup b runs on XCD b % 8) each take
// one contiguous eighth of the tile range, so neighbouring tiles of one
// member, which re-read each other's halo elements, share an L2.
template <typename R, typename Body>
__device__ __forceinline__ void ens_loop(const EnsArgs<R> &a, Body body) {
  const int xcd = blockIdx.x & 7, lb = blockIdx.x >> 3, nbx = gridDim.x >> 3;
  const int per = (a.total + 7) >> 3;
  const int t0 = xcd * per, t1 = min(t0 + per, a.total);
  bool first = true;
  for (int t = t0 + lb; t < t1; t += nbx) {
    int lo = 0, hi = a.n - 1;  // largest i with tile_off[i] <= t
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (a.tile_off[mid] <= t) lo = mid;
      else hi = mid - 1;
    }
    const int i = __builtin_amdgcn_readfirstlane(lo);
    if (!first) __syncthreads();  // the previous tile's last phase has read its LDS
    first = false;
    // The parameter table is read through the constant address space: the
    // loads are uniform and never clobbered, so they become scalar loads into
    // SGPRs (a copy through a global pointer would hold the ~60-dword struct in
    // VGPRs and lower the occupancy).
    using CWord = const __attribute__((address_space(4))) uint32_t;
    static_assert(sizeof(StepParams<R>) % 4 == 0, "whole dwords");
    StepParams<R> prm;
    {
      CWord *src = (CWord *)(uintptr_t)(a.prms + 2 * i + (a.step & 1));
      uint32_t *dst = reinterpret_cast<uint32_t *>(&prm);
#pragma unroll
      for (unsigned k = 0; k < sizeof(StepParams<R>) / 4; k++) dst[k] = src[k];
    }
    prm.step = a.step;
    prm.last = a.last;
    body(prm, t - a.tile_off[i]);
  }
}

template <int P, int MODE, int NT>
__global__ __launch_bounds__(NT) void ens_tile_kernel(EnsArgs<typename RealOf<MODE>::T> a) {
  ens_loop(a, [](const StepParams<typename RealOf<MODE>::T> &prm, int tile) { tile_body<P, MODE, true, NT>(prm, tile); });
}

template <int P, int MODE, int EX, int EY, int WAVES, int PF, int LPE>
__global__ __launch_bounds__(EX * EY * LPE) __attribute__((amdgpu_waves_per_eu(WAVES > 0 ? WAVES : 1, 8)))
void ens_reg_kernel(EnsArgs<float> a) {
  ens_loop(a, [](const StepParams<float> &prm, int tile) { reg_body<P, MODE, true, EX, EY, PF, LPE>(prm, tile); });
}

template <int P, int MODE>
__global__ __launch_bounds__(kThreads) void ens_mfma_kernel(EnsArgs<float> a) {
  ens_loop(a, [](const StepParams<float> &prm, int tile) { mfma_body<P, MODE, false>(prm, tile); });
}

template <int P, int MODE>
__global__ __launch_bounds__(kThreads) void ens_line_kernel(EnsArgs<float> a) {
  ens_loop(a, [](const StepParams<float> &prm, int tile) { mfma_body<P, MODE, true>(prm, tile); });
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
