// Launches of the mixed form's kernels (FP32 flux evaluation, FP64 RK
// registers), defined in mixed.hip (single domain) and mixed_ensemble.hip
// (ensembles).  Those translation units alone instantiate the FP32 kernels:
// they allow contraction into FMA within an expression, under one pragma,
// while the FP64 kernels keep the reference's operation order, so a
// mixed-form ensemble member still advances bit-for-bit as it would alone.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels_ensemble.hpp"

namespace mhd {

// Host stub of the interior stage-2 kernel of the given variant (MHD or
// hydro, with rotation), for occupancy queries.
const void* mixed_stream_fn(bool mag, int waves, int sched);

// Interior planes of one stage (launch_stream_stage<float, ...>).
void mixed_stage(bool mag, bool rot, int stage, const StageArgs& a, hipStream_t s, int waves,
                 int sched, int target_blocks);

// Wall planes of one stage (stage_wall_kernel<float, ...>).
void mixed_wall(bool mag, bool rot, int stage, const StageArgs& a, dim3 grid, dim3 block,
                hipStream_t s);

// Host stub of the ensemble's interior stage-2 kernel of the given variant.
const void* mixed_ens_stage_fn(int waves, int sched);

// One ensemble stage: the persistent interior launch over `items` (grid_int
// blocks) and the persistent wall launch over `walls` (grid_wall blocks); an
// empty list launches nothing.
void mixed_ens_stage(int stage, int waves, int sched, const StageArgs* args, const WorkItem* items,
                     int nitems, int grid_int, const WallItem* walls, int nwalls, int grid_wall,
                     hipStream_t s);

}  // namespace mhd
