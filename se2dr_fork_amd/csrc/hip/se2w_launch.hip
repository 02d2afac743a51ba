// se2w_launch.hip -- the single-instance step kernels of one precision mode
// and their launches.  Compiled once per mode with -DSE2W_MODE_TU=<mode>
// (MODE_F64 = 0, MODE_F32 = 1, MODE_F32K = 2).
#include "se2w_launch.hpp"

#include "se2w.h"
#include "se2w_kernels_mfma.hpp"
#include "se2w_kernels_reg.hpp"
#include "se2w_kernels_reg64.hpp"

#ifndef SE2W_MODE_TU
#error "compile with -DSE2W_MODE_TU=<mode>"
#endif

namespace se2w {
namespace {

constexpr int kMode = SE2W_MODE_TU;
using R = typename RealOf<kMode>::T;

[[maybe_unused]] bool is_reg(int variant) { return variant >= SE2W_KERNEL_REG && variant < SE2W_KERNEL_REG + kNumRegSchedules; }

template <int P, int S>
bool launch_reg64_s(const LaunchCfg &c, const StepParams<R> &prm) {
  if constexpr (kMode == MODE_F64 && S < kNumReg64Schedules) {
    if constexpr (reg64_schedule_built(P, S, true)) {
      if (c.variant - SE2W_KERNEL_REG64 == S) {
        constexpr Reg64Schedule r = kReg64Schedules[S];
        hipLaunchKernelGGL((reg64_kernel<P, r.ex, r.ey, r.waves, r.pf, r.lpe>), dim3(c.grid), dim3(r.ex * r.ey * r.lpe),
                           0, c.stream, prm);
        return true;
      }
    }
    return launch_reg64_s<P, S + 1>(c, prm);
  }
  return false;
}

template <int P, bool FUSED, int S>
bool launch_reg_s(const LaunchCfg &c, const StepParams<R> &prm) {
  if constexpr (kMode != MODE_F64 && S < kNumRegSchedules) {
    if constexpr (reg_schedule_built(P, S, FUSED)) {
      if (c.variant - SE2W_KERNEL_REG == S) {
        constexpr RegSchedule r = kRegSchedules[S];
        hipLaunchKernelGGL((reg_kernel<P, kMode, FUSED, r.ex, r.ey, r.waves, r.pf, r.lpe>), dim3(c.grid),
                           dim3(r.ex * r.ey * r.lpe), 0, c.stream, prm);
        return true;
      }
    }
    return launch_reg_s<P, FUSED, S + 1>(c, prm);
  }
  return false;
}

// Element (+ update when fused) kernel of the variant; variants without a
// form for (P, FUSED) fall back to the LDS tile kernel (the callers set the
// tiling of the kernel that runs).
template <int P, bool FUSED>
void launch_element_p(const LaunchCfg &c, const StepParams<R> &prm) {
  const dim3 grid(c.grid);
  if constexpr (kMode != MODE_F64 && P >= 2) {
    if (is_reg(c.variant) && launch_reg_s<P, FUSED, 0>(c, prm)) return;
  }
  if constexpr (kMode != MODE_F64 && P >= 6 && FUSED) {
    if (c.variant == SE2W_KERNEL_MFMA) {
      hipLaunchKernelGGL((mfma_kernel<P, kMode>), grid, dim3(kThreads), 0, c.stream, prm);
      return;
    }
    if (c.variant == SE2W_KERNEL_LINE) {
      hipLaunchKernelGGL((line_kernel<P, kMode>), grid, dim3(kThreads), 0, c.stream, prm);
      return;
    }
  }
  if constexpr (kMode == MODE_F64 && FUSED) {
    if (c.variant >= SE2W_KERNEL_REG64 && launch_reg64_s<P, 0>(c, prm)) return;
  }
  if constexpr (kMode == MODE_F64 && !FUSED) {
    if constexpr (reg64_schedule_built(P, 0, false)) {
      if (c.variant == SE2W_KERNEL_REG64) {
        constexpr Reg64Schedule r = kReg64Schedules[0];
        hipLaunchKernelGGL((reg64_element_kernel<P, r.ex, r.ey, r.waves, r.pf, r.lpe>), grid,
                           dim3(r.ex * r.ey * r.lpe), 0, c.stream, prm);
        return;
      }
    }
  }
  if constexpr (FUSED) {
    if (c.variant == SE2W_KERNEL_LDS_WIDE) {
      hipLaunchKernelGGL((tile_kernel<P, kMode, true, 512>), grid, dim3(512), 0, c.stream, prm);
      return;
    }
  }
  hipLaunchKernelGGL((tile_kernel<P, kMode, FUSED>), grid, dim3(kThreads), 0, c.stream, prm);
}

template <int P>
void launch_step_p(const LaunchCfg &c, const StepParams<R> &prm, hipEvent_t *evs) {
  if (evs) hipEventRecord(evs[0], c.stream);
  if (c.schedule == SE2W_SCHED_FUSED) {
    launch_element_p<P, true>(c, prm);
    if (evs) hipEventRecord(evs[1], c.stream);
  } else {
    launch_element_p<P, false>(c, prm);
    if (evs) hipEventRecord(evs[1], c.stream);
    hipLaunchKernelGGL((update_kernel<P, kMode>), dim3(c.grid), dim3(kThreads), 0, c.stream, prm);
    if (evs) hipEventRecord(evs[2], c.stream);
  }
}

}  // namespace

template <>
void launch_step<kMode>(const LaunchCfg &c, const StepParams<R> &prm, hipEvent_t *evs) {
  switch (c.p) {
    case 1: launch_step_p<1>(c, prm, evs); break;
    case 2: launch_step_p<2>(c, prm, evs); break;
    case 3: launch_step_p<3>(c, prm, evs); break;
    case 4: launch_step_p<4>(c, prm, evs); break;
    case 5: launch_step_p<5>(c, prm, evs); break;
    case 6: launch_step_p<6>(c, prm, evs); break;
    case 7: launch_step_p<7>(c, prm, evs); break;
    case 8: launch_step_p<8>(c, prm, evs); break;
    default: break;
  }
}

template <>
void launch_element_only<kMode>(const LaunchCfg &c, const StepParams<R> &prm) {
  switch (c.p) {
    case 1: launch_element_p<1, false>(c, prm); break;
    case 2: launch_element_p<2, false>(c, prm); break;
    case 3: launch_element_p<3, false>(c, prm); break;
    case 4: launch_element_p<4, false>(c, prm); break;
    case 5: launch_element_p<5, false>(c, prm); break;
    case 6: launch_element_p<6, false>(c, prm); break;
    case 7: launch_element_p<7, false>(c, prm); break;
    case 8: launch_element_p<8, false>(c, prm); break;
    default: break;
  }
}

template <>
void launch_predictor<kMode>(const StepParams<R> &prm, long n_padded, hipStream_t s) {
  const int grid = (int)std::min<long>((n_padded + kThreads - 1) / kThreads, 4096);
  hipLaunchKernelGGL((predictor_kernel<kMode>), dim3(grid), dim3(kThreads), 0, s, prm);
}

}  // namespace se2w
