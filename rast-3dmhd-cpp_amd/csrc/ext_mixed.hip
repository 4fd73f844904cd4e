// The off-default branches in the mixed form (S32): FP64 fluxes from the
// FP32 copies of the stage values with the background added back, the
// copies written as the mixed form writes them.  Single domains take the
// streaming kernel with the branches of one ExtVariant compiled in (as the
// FP64 form does); ensemble members take the kernel with all of them.  A
// translation unit of its own, so its 75 instantiations build in parallel
// with the others; the arithmetic is that of mhd3d.hip's kernels (no
// contraction).
#define MHD3D_SECONDARY_TU 1

#include <hip/hip_runtime.h>

#include "ext_launch.hpp"
#include "kernels_basic.hpp"
#include "kernels_ensemble.hpp"
#include "kernels_stream.hpp"

namespace mhd {

namespace {

// interior branch sets by ExtVariant (EXTV_NONE: a run whose branches act on
// the walls only still evaluates its fluxes from the FP32 copies in FP64)
constexpr int kMaskS32[8] = {0,        EXT_DH | EXT_DV, EXT_REM,  EXT_SHR,
                             EXT_HEAT, EXT_INTERIOR,    EXT_RLAX, EXT_DH | EXT_DV | EXT_REM};

// Schedule of each kernel, as for the FP64 form's: two barriers per plane
// (single-buffered planes, 76-82 KB of LDS instead of 106-114 KB) where the
// registers then allow two MHD blocks per CU -- every stage of the wall-only,
// iD, iTCon 2, RLAX and lRem kernels -- and one barrier per plane where
// registers hold them to one block anyway.
template <bool MAG, int STAGE, int M>
constexpr bool kOneBarrierS32 = !MAG || M == kMaskS32[EXTV_SHR] || M == kMaskS32[EXTV_ALL];
// lRem's stage 3 under a two-wave register bound (it needs 263 registers)
template <bool MAG, int STAGE, int M>
constexpr int kWavesS32 = (MAG && STAGE == 3 && M == kMaskS32[EXTV_REM]) ? 2 : 1;

template <bool MAG, bool ROT, int STAGE, int M>
void launch_s32(const StageArgs& a, hipStream_t s, int target_blocks)
{
    const StreamPlan p = stream_plan(a.g, target_blocks);
    stage_stream_kernel<double, MAG, ROT, STAGE, kWavesS32<MAG, STAGE, M>, kOneBarrierS32<MAG, STAGE, M>, true, true, M, true>
        <<<p.nblocks, dim3(STX, STY, 1), 0, s>>>(a, p.zchunk, p.nchunks, p.ntx, p.nty, p.xcd_remap);
}

template <int M>
void stage_s32(bool mag, bool rot, int stage, const StageArgs& a, hipStream_t s, int tb)
{
    if (mag && rot) {
        if (stage == 1) launch_s32<true, true, 1, M>(a, s, tb);
        else if (stage == 2) launch_s32<true, true, 2, M>(a, s, tb);
        else launch_s32<true, true, 3, M>(a, s, tb);
    } else if (mag) {
        if (stage == 1) launch_s32<true, false, 1, M>(a, s, tb);
        else if (stage == 2) launch_s32<true, false, 2, M>(a, s, tb);
        else launch_s32<true, false, 3, M>(a, s, tb);
    } else if (rot) {
        if (stage == 1) launch_s32<false, true, 1, M>(a, s, tb);
        else if (stage == 2) launch_s32<false, true, 2, M>(a, s, tb);
        else launch_s32<false, true, 3, M>(a, s, tb);
    } else {
        if (stage == 1) launch_s32<false, false, 1, M>(a, s, tb);
        else if (stage == 2) launch_s32<false, false, 2, M>(a, s, tb);
        else launch_s32<false, false, 3, M>(a, s, tb);
    }
}

template <int M>
const void* fn_s32(bool mag)
{
    return mag ? (const void*)stage_stream_kernel<double, true, true, 2, 1, kOneBarrierS32<true, 2, M>, true, true, M, true>
               : (const void*)stage_stream_kernel<double, false, true, 2, 1, kOneBarrierS32<false, 2, M>, true, true, M, true>;
}

}  // namespace

const void* ext_stream_fn_s32(bool mag, int variant)
{
    switch (variant) {
    case EXTV_NONE: return fn_s32<kMaskS32[EXTV_NONE]>(mag);
    case EXTV_ID: return fn_s32<kMaskS32[EXTV_ID]>(mag);
    case EXTV_REM: return fn_s32<kMaskS32[EXTV_REM]>(mag);
    case EXTV_SHR: return fn_s32<kMaskS32[EXTV_SHR]>(mag);
    case EXTV_HEAT: return fn_s32<kMaskS32[EXTV_HEAT]>(mag);
    case EXTV_RLAX: return fn_s32<kMaskS32[EXTV_RLAX]>(mag);
    case EXTV_IDREM: return fn_s32<kMaskS32[EXTV_IDREM]>(mag);
    default: return fn_s32<kMaskS32[EXTV_ALL]>(mag);
    }
}

void ext_stage_s32(bool mag, bool rot, int stage, int variant, const StageArgs& a, hipStream_t s,
                   int target_blocks)
{
    switch (variant) {
    case EXTV_NONE: stage_s32<kMaskS32[EXTV_NONE]>(mag, rot, stage, a, s, target_blocks); break;
    case EXTV_ID: stage_s32<kMaskS32[EXTV_ID]>(mag, rot, stage, a, s, target_blocks); break;
    case EXTV_REM: stage_s32<kMaskS32[EXTV_REM]>(mag, rot, stage, a, s, target_blocks); break;
    case EXTV_SHR: stage_s32<kMaskS32[EXTV_SHR]>(mag, rot, stage, a, s, target_blocks); break;
    case EXTV_HEAT: stage_s32<kMaskS32[EXTV_HEAT]>(mag, rot, stage, a, s, target_blocks); break;
    case EXTV_RLAX: stage_s32<kMaskS32[EXTV_RLAX]>(mag, rot, stage, a, s, target_blocks); break;
    case EXTV_IDREM: stage_s32<kMaskS32[EXTV_IDREM]>(mag, rot, stage, a, s, target_blocks); break;
    default: stage_s32<kMaskS32[EXTV_ALL]>(mag, rot, stage, a, s, target_blocks); break;
    }
}

const void* ext_ens_fn_s32()
{
    return (const void*)ensemble_stage_ext_kernel<2, true>;
}

void ext_ens_stage_s32(int stage, const StageArgs* args, const WorkItem* items, int nitems, int grid_int,
                       const WallItem* walls, int nwalls, int grid_wall, hipStream_t s)
{
    if (nitems > 0) {
        const dim3 blk(STX, STY, 1);
        if (stage == 1) ensemble_stage_ext_kernel<1, true><<<grid_int, blk, 0, s>>>(args, items, nitems);
        else if (stage == 2) ensemble_stage_ext_kernel<2, true><<<grid_int, blk, 0, s>>>(args, items, nitems);
        else ensemble_stage_ext_kernel<3, true><<<grid_int, blk, 0, s>>>(args, items, nitems);
    }
    if (nwalls > 0) {
        const dim3 wblk(64, 4, 1);
        if (stage == 1) ensemble_wall_kernel<double, 1, true, true><<<grid_wall, wblk, 0, s>>>(args, walls, nwalls);
        else if (stage == 2) ensemble_wall_kernel<double, 2, true, true><<<grid_wall, wblk, 0, s>>>(args, walls, nwalls);
        else ensemble_wall_kernel<double, 3, true, true><<<grid_wall, wblk, 0, s>>>(args, walls, nwalls);
    }
}

}  // namespace mhd
