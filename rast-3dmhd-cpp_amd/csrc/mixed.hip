// The mixed form's single-domain kernels: FP32 flux evaluation from FP32
// stencil copies, FP64 RK registers, FP64 CFL (interior and wall stages).
// This translation unit and mixed_ensemble.hip (the ensemble's FP32 kernels)
// instantiate every FP32 stage kernel and nothing else.  Both allow
// contraction into FMA within an expression (the FP64 kernels of mhd3d.hip
// never contract, to keep the reference's rounding) under the same pragma,
// so every FP32 kernel rounds the same way and an ensemble member matches
// its run alone bit for bit; they differ only in the instruction scheduler
// they are built with, which does not change results.
#define MHD3D_FP32_FORM_TU 1

#include <hip/hip_runtime.h>

#include "kernels_basic.hpp"
#include "kernels_ensemble.hpp"
#include "kernels_stream.hpp"
#include "mixed_launch.hpp"

namespace mhd {

const void* mixed_stream_fn(bool mag, int waves, int sched)
{
    return mag ? stream_fn<float, true, true>(waves, sched) : stream_fn<float, false, true>(waves, sched);
}

template <bool MAG, bool ROT>
static void stage_mr(int stage, const StageArgs& a, hipStream_t s, int waves, int sched, int tb)
{
    if (stage == 1)
        launch_stream_stage<float, MAG, ROT, 1>(a, s, waves, sched, tb);
    else if (stage == 2)
        launch_stream_stage<float, MAG, ROT, 2>(a, s, waves, sched, tb);
    else
        launch_stream_stage<float, MAG, ROT, 3>(a, s, waves, sched, tb);
}

void mixed_stage(bool mag, bool rot, int stage, const StageArgs& a, hipStream_t s, int waves,
                 int sched, int target_blocks)
{
    if (mag && rot) stage_mr<true, true>(stage, a, s, waves, sched, target_blocks);
    else if (mag) stage_mr<true, false>(stage, a, s, waves, sched, target_blocks);
    else if (rot) stage_mr<false, true>(stage, a, s, waves, sched, target_blocks);
    else stage_mr<false, false>(stage, a, s, waves, sched, target_blocks);
}

template <bool MAG, bool ROT>
static void wall_mr(int stage, const StageArgs& a, dim3 grid, dim3 block, hipStream_t s)
{
    if (stage == 1)
        stage_wall_kernel<float, MAG, ROT, 1><<<grid, block, 0, s>>>(a);
    else if (stage == 2)
        stage_wall_kernel<float, MAG, ROT, 2><<<grid, block, 0, s>>>(a);
    else
        stage_wall_kernel<float, MAG, ROT, 3><<<grid, block, 0, s>>>(a);
}

void mixed_wall(bool mag, bool rot, int stage, const StageArgs& a, dim3 grid, dim3 block,
                hipStream_t s)
{
    if (mag && rot) wall_mr<true, true>(stage, a, grid, block, s);
    else if (mag) wall_mr<true, false>(stage, a, grid, block, s);
    else if (rot) wall_mr<false, true>(stage, a, grid, block, s);
    else wall_mr<false, false>(stage, a, grid, block, s);
}

}  // namespace mhd
