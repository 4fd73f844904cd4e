// The mixed form's ensemble kernels (FP32 flux evaluation, FP64 RK
// registers): the persistent interior and wall launches over many members.
// Same contraction pragma as mixed.hip (so a member rounds exactly as its
// run alone), built with the default instruction scheduler, which measured
// faster for these kernels than the one mixed.hip uses.
#define MHD3D_FP32_FORM_TU 1

#include <hip/hip_runtime.h>

#include "kernels_basic.hpp"
#include "kernels_ensemble.hpp"
#include "kernels_stream.hpp"
#include "mixed_launch.hpp"

namespace mhd {

template <int STAGE, int WAVES>
static const void* ens_fn(int sched)
{
    switch (sched & 3) {
    case 0: return (const void*)ensemble_stage_kernel<float, STAGE, WAVES, false, false>;
    case 1: return (const void*)ensemble_stage_kernel<float, STAGE, WAVES, true, false>;
    case 2: return (const void*)ensemble_stage_kernel<float, STAGE, WAVES, false, true>;
    default: return (const void*)ensemble_stage_kernel<float, STAGE, WAVES, true, true>;
    }
}

const void* mixed_ens_stage_fn(int waves, int sched)
{
    return waves == 1 ? ens_fn<2, 1>(sched) : ens_fn<2, 2>(sched);
}

template <int STAGE, int WAVES>
static void ens_launch(int sched, const StageArgs* args, const WorkItem* items, int nitems,
                       int grid, hipStream_t s)
{
    const dim3 blk(STX, STY, 1);
    switch (sched & 3) {
    case 0: ensemble_stage_kernel<float, STAGE, WAVES, false, false><<<grid, blk, 0, s>>>(args, items, nitems); break;
    case 1: ensemble_stage_kernel<float, STAGE, WAVES, true, false><<<grid, blk, 0, s>>>(args, items, nitems); break;
    case 2: ensemble_stage_kernel<float, STAGE, WAVES, false, true><<<grid, blk, 0, s>>>(args, items, nitems); break;
    default: ensemble_stage_kernel<float, STAGE, WAVES, true, true><<<grid, blk, 0, s>>>(args, items, nitems); break;
    }
}

void mixed_ens_stage(int stage, int waves, int sched, const StageArgs* args, const WorkItem* items,
                     int nitems, int grid_int, const WallItem* walls, int nwalls, int grid_wall,
                     hipStream_t s)
{
    if (nitems == 0) {
    } else if (waves == 1) {
        if (stage == 1) ens_launch<1, 1>(sched, args, items, nitems, grid_int, s);
        else if (stage == 2) ens_launch<2, 1>(sched, args, items, nitems, grid_int, s);
        else ens_launch<3, 1>(sched, args, items, nitems, grid_int, s);
    } else {
        if (stage == 1) ens_launch<1, 2>(sched, args, items, nitems, grid_int, s);
        else if (stage == 2) ens_launch<2, 2>(sched, args, items, nitems, grid_int, s);
        else ens_launch<3, 2>(sched, args, items, nitems, grid_int, s);
    }
    const dim3 wblk(64, 4, 1);
    if (nwalls == 0) {
    } else if (stage == 1) ensemble_wall_kernel<float, 1><<<grid_wall, wblk, 0, s>>>(args, walls, nwalls);
    else if (stage == 2) ensemble_wall_kernel<float, 2><<<grid_wall, wblk, 0, s>>>(args, walls, nwalls);
    else ensemble_wall_kernel<float, 3><<<grid_wall, wblk, 0, s>>>(args, walls, nwalls);
}

}  // namespace mhd
