// se2w_ens.hpp -- the batched step of an ensemble of uneven instances: one
// launch advances every member by one step.
#pragma once

#include <hip/hip_runtime.h>

#include "se2w_kernels.hpp"

namespace se2w {

// Arguments of one batched step launch.  Member i owns the global tiles
// [tile_off[i], tile_off[i+1]); its step parameters for the two parities of
// the double-buffered displacement are prms[2 i] (even steps) and
// prms[2 i + 1] (odd steps).
template <typename R>
struct EnsArgs {
  const StepParams<R> *prms;
  const int *tile_off;
  int n;      // members
  int total;  // tiles of all members
  int step;   // step index within the run
  int last;   // last step of the run: write the plain state
};

// Launches the batched step kernel of (p, mode, variant) with `grid`
// workgroups (8 * ceil(tiles / 8): one per tile) on stream s.  args points to an
// EnsArgs<double> (FP64) or EnsArgs<float>.  Returns 0, or -1 when the
// variant is not built for (p, mode).
int ens_launch(int p, int mode, int variant, const void *args, int grid, hipStream_t s);

// Workgroups of the batched kernel that fit one CU at once, and its
// workgroup size; -1 when the variant is not built for (p, mode).
int ens_occupancy(int p, int mode, int variant, int *blocks_per_cu, int *threads);

// The batched kernels of one precision mode (one translation unit each):
// with args, launches the kernel of (p, variant); without, reports its
// occupancy in *bpc and *thr.  Returns 0, or -1 when the variant is not built.
template <int MODE>
int ens_dispatch(int p, int variant, const void *args, int grid, hipStream_t s, int *bpc, int *thr);

}  // namespace se2w
