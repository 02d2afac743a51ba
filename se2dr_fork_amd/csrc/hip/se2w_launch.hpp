// se2w_launch.hpp -- launches of the single-instance step kernels.  Each
// precision's kernels are instantiated in a translation unit of their own
// (se2w_launch.hip compiled once per mode), so the three build in parallel.
#pragma once

#include <hip/hip_runtime.h>

#include "se2w_kernels.hpp"

namespace se2w {

// What a launch needs from a context: degree, element-kernel variant
// (SE2W_KERNEL_*), schedule, workgroups of the tiled grid (8 x tiles per
// XCD) and the stream.
struct LaunchCfg {
  int p, variant, schedule, grid;
  hipStream_t stream;
};

// One step (SE2W_SCHED_FUSED: the fused kernel; SE2W_SCHED_SPLIT: element
// kernel, then update kernel).  evs (may be null) are recorded around the
// launches: evs[0] before, evs[1] after the (element) kernel, evs[2] after
// the update kernel of the split schedule.
template <int MODE>
void launch_step(const LaunchCfg &c, const StepParams<typename RealOf<MODE>::T> &prm, hipEvent_t *evs);

// f = K(u): the element-only kernel of the variant (prm.u_in -> prm.f).
template <int MODE>
void launch_element_only(const LaunchCfg &c, const StepParams<typename RealOf<MODE>::T> &prm);

// The predictor that opens a run over n_padded = pitch * ny nodes.
template <int MODE>
void launch_predictor(const StepParams<typename RealOf<MODE>::T> &prm, long n_padded, hipStream_t s);

}  // namespace se2w
