// The off-default interior stage kernels of the FP64 form: the streaming
// kernel with the branches of one ExtVariant compiled in.  Kept in a
// translation unit of its own so the 60 instantiations build in parallel
// with the main one; the arithmetic is that of mhd3d.hip's kernels (no
// contraction).
#define MHD3D_SECONDARY_TU 1

#include <hip/hip_runtime.h>

#include "ext_launch.hpp"
#include "kernels_basic.hpp"
#include "kernels_stream.hpp"

namespace mhd {

namespace {

constexpr int kMask[8] = {0,        EXT_DH | EXT_DV, EXT_REM,      EXT_SHR,
                          EXT_HEAT, EXT_INTERIOR,    EXT_RLAX,     EXT_DH | EXT_DV | EXT_REM};

// Schedule and register budget of each kernel: one barrier per plane
// (double-buffered planes) or two (single-buffered).  MHD's double-buffered
// planes take 106-114 KB of LDS, one block per CU; single-buffered, 76-82 KB,
// so MHD kernels take two barriers where their registers then allow two
// waves per SIMD: iD, iD with lRem, lRem, iTCon 2 and RLAX, the stage-3
// kernels of the last four under a two-wave register bound (239-245 VGPRs,
// no spills, with their branch tests constant; it held them at 256 with a
// few spills before, still faster than one block per CU).  lShr (256 VGPRs
// and 77-101 AGPRs) and the all-branch combinations stay register-bound at
// one block per CU, where one barrier per plane is the faster.
template <bool MAG, int STAGE, int M>
constexpr bool kOneBarrier = !MAG || M == kMask[EXTV_SHR] || M == kMask[EXTV_ALL];
template <bool MAG, int STAGE, int M>
constexpr int kWaves = (MAG && STAGE == 3 && M != kMask[EXTV_ID] && !kOneBarrier<MAG, STAGE, M>) ? 2 : 1;

template <bool MAG, bool ROT, int STAGE, int M>
void launch_m(const StageArgs& a, hipStream_t s, int target_blocks)
{
    const StreamPlan p = stream_plan(a.g, target_blocks);
    stage_stream_kernel<double, MAG, ROT, STAGE, kWaves<MAG, STAGE, M>, kOneBarrier<MAG, STAGE, M>, true, true, M>
        <<<p.nblocks, dim3(STX, STY, 1), 0, s>>>(a, p.zchunk, p.nchunks, p.ntx, p.nty, p.xcd_remap);
}

template <int M>
void stage_mask(bool mag, bool rot, int stage, const StageArgs& a, hipStream_t s, int tb)
{
    if (mag && rot) {
        if (stage == 1) launch_m<true, true, 1, M>(a, s, tb);
        else if (stage == 2) launch_m<true, true, 2, M>(a, s, tb);
        else launch_m<true, true, 3, M>(a, s, tb);
    } else if (mag) {
        if (stage == 1) launch_m<true, false, 1, M>(a, s, tb);
        else if (stage == 2) launch_m<true, false, 2, M>(a, s, tb);
        else launch_m<true, false, 3, M>(a, s, tb);
    } else if (rot) {
        if (stage == 1) launch_m<false, true, 1, M>(a, s, tb);
        else if (stage == 2) launch_m<false, true, 2, M>(a, s, tb);
        else launch_m<false, true, 3, M>(a, s, tb);
    } else {
        if (stage == 1) launch_m<false, false, 1, M>(a, s, tb);
        else if (stage == 2) launch_m<false, false, 2, M>(a, s, tb);
        else launch_m<false, false, 3, M>(a, s, tb);
    }
}

template <int M>
const void* fn_mask(bool mag)
{
    return mag ? (const void*)stage_stream_kernel<double, true, true, 2, 1, kOneBarrier<true, 2, M>, true, true, M>
               : (const void*)stage_stream_kernel<double, false, true, 2, 1, kOneBarrier<false, 2, M>, true, true, M>;
}

}  // namespace

const void* ext_stream_fn(bool mag, int variant)
{
    switch (variant) {
    case EXTV_ID: return fn_mask<kMask[EXTV_ID]>(mag);
    case EXTV_REM: return fn_mask<kMask[EXTV_REM]>(mag);
    case EXTV_SHR: return fn_mask<kMask[EXTV_SHR]>(mag);
    case EXTV_HEAT: return fn_mask<kMask[EXTV_HEAT]>(mag);
    case EXTV_RLAX: return fn_mask<kMask[EXTV_RLAX]>(mag);
    case EXTV_IDREM: return fn_mask<kMask[EXTV_IDREM]>(mag);
    default: return fn_mask<kMask[EXTV_ALL]>(mag);
    }
}

void ext_stage(bool mag, bool rot, int stage, int variant, const StageArgs& a, hipStream_t s,
               int target_blocks)
{
    switch (variant) {
    case EXTV_ID: stage_mask<kMask[EXTV_ID]>(mag, rot, stage, a, s, target_blocks); break;
    case EXTV_REM: stage_mask<kMask[EXTV_REM]>(mag, rot, stage, a, s, target_blocks); break;
    case EXTV_SHR: stage_mask<kMask[EXTV_SHR]>(mag, rot, stage, a, s, target_blocks); break;
    case EXTV_HEAT: stage_mask<kMask[EXTV_HEAT]>(mag, rot, stage, a, s, target_blocks); break;
    case EXTV_RLAX: stage_mask<kMask[EXTV_RLAX]>(mag, rot, stage, a, s, target_blocks); break;
    case EXTV_IDREM: stage_mask<kMask[EXTV_IDREM]>(mag, rot, stage, a, s, target_blocks); break;
    default: stage_mask<kMask[EXTV_ALL]>(mag, rot, stage, a, s, target_blocks); break;
    }
}

}  // namespace mhd
