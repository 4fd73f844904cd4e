// Launches of the off-default interior stage kernels (FP64 fluxes), defined
// in ext_stream.hip (FP64 form) and ext_mixed.hip (mixed form).  Each variant compiles in one set of branches (EXTM mask),
// so a run with one off-default switch does not carry the registers of the
// others; a run whose interior switches fit no single set takes the variant
// with all of them.  Every variant has one wave per SIMD, one barrier per
// plane and the deferred update.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels_basic.hpp"

namespace mhd {

enum ExtVariant : int {
    EXTV_NONE = 0,   // no interior branch: the default interior kernel
    EXTV_ID = 1,     // density diffusion (iD 1/2; DH, DV)
    EXTV_REM = 2,    // lRem conduction
    EXTV_SHR = 3,    // lShr carried-over FT
    EXTV_HEAT = 4,   // iTCon 2 heat loss
    EXTV_ALL = 5,    // any combination
    EXTV_RLAX = 6,   // the RLAX correction of FT
    EXTV_IDREM = 7,  // density diffusion with lRem (both plane means)
};

// The narrowest variant whose branch set holds the interior flags.
inline int ext_variant(int flags)
{
    const int f = flags & EXT_INTERIOR;
    if (f == 0) return EXTV_NONE;
    if ((f & ~(EXT_DH | EXT_DV)) == 0) return EXTV_ID;
    if (f == EXT_REM) return EXTV_REM;
    if (f == EXT_SHR) return EXTV_SHR;
    if (f == EXT_HEAT) return EXTV_HEAT;
    if (f == EXT_RLAX) return EXTV_RLAX;
    if ((f & EXT_REM) && (f & ~(EXT_DH | EXT_DV | EXT_REM)) == 0) return EXTV_IDREM;
    return EXTV_ALL;
}

// Host stub of the stage-2 kernel of a variant (MHD or hydro, with
// rotation), for occupancy queries.
const void* ext_stream_fn(bool mag, int variant);

// Interior planes of one stage with the variant's kernel.
void ext_stage(bool mag, bool rot, int stage, int variant, const StageArgs& a, hipStream_t s,
               int target_blocks);

// The mixed form with off-default branches (ext_mixed.hip): the branch set
// of the variant compiled in, FP64 fluxes from the FP32 copies of the stage
// values (stage 1 from the FP64 state), the copies written as in the mixed
// form.
const void* ext_stream_fn_s32(bool mag, int variant);
void ext_stage_s32(bool mag, bool rot, int stage, int variant, const StageArgs& a, hipStream_t s,
                   int target_blocks);

struct WorkItem;
struct WallItem;

// Mixed-form ensemble members with off-default branches: the persistent
// interior launch over `items` (grid_int blocks, ensemble_stage_ext_kernel
// with S32) and the persistent wall launch over `walls` (grid_wall blocks);
// an empty list launches nothing.  The host stub of the interior stage-2
// kernel is for occupancy queries.
const void* ext_ens_fn_s32();
void ext_ens_stage_s32(int stage, const StageArgs* args, const WorkItem* items, int nitems, int grid_int,
                       const WallItem* walls, int nwalls, int grid_wall, hipStream_t s);

}  // namespace mhd
