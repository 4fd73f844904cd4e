// z-streaming stage kernel (placeholder: forwards to the point kernel).
#pragma once
#include "kernels_basic.hpp"

namespace mhd {

template <bool MAG, bool ROT, int STAGE>
static void launch_stream_stage(const StageArgs& a, hipStream_t s)
{
    dim3 blk(64, 4, 1);
    dim3 grd((a.g.nPx + 63) / 64, (a.g.nPy + 3) / 4, a.g.nPz - 2);
    stage_point_kernel<MAG, ROT, STAGE><<<grd, blk, 0, s>>>(a);
}

}  // namespace mhd
