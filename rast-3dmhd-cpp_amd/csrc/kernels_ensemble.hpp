// Kernels of an ensemble: many independent domains (uneven sizes, flags and
// parameters) advanced together, one persistent launch per RK3 stage for the
// planes between the walls and one for the wall planes of every member.
//
// Each launch walks a work list sorted by cost (largest first) with a static
// stride of the grid size; a work item names its member, whose operands sit
// in a table of StageArgs.  The arithmetic of a point is exactly that of the
// single-domain kernels (the same stage_item / wall_point code), so a member
// advances bit-for-bit as it would alone.  TSTEP partials go to the member's
// own partial records; ensemble_finish_kernel closes each member's step
// (TIMC, nIt) and forms its next DT, one block per member.
#pragma once
// FP64 arithmetic follows the reference's operation order, so nothing is
// contracted into FMA; the translation unit of the FP32 (mixed-form) kernels
// alone allows contraction within an expression.
#ifdef MHD3D_FP32_FORM_TU
#pragma clang fp contract(on)
#else
#pragma clang fp contract(off)
#endif

#include "kernels_basic.hpp"
#include "kernels_stream.hpp"

namespace mhd {

// One interior work item: tile (tile_x, tile_y) of member `member` over the
// planes zb .. ze-1; flags bit 0 = lMag, bit 1 = lRot; `part` = index of its
// TSTEP partial record in the member's array.
struct WorkItem {
    int member, flags, tile_x, tile_y, zb, ze, part, pad;
};

// One wall work item: a 64 x 4 block of points at (x0, y0) of the top
// (top = 1) or bottom wall of member `member`.
struct WallItem {
    int member, flags, x0, y0, top, part, pad0, pad1;
};

// Per-member closing data of a step: the member's partial records and
// scalars, and the slot of the ensemble's contiguous copy of the scalars
// (one readback for all members).
struct FinishArgs {
    CflPart* parts;
    int nparts;
    int mag;
    DevScalars* ds;
    DevScalars* mirror;
    Phys c;
};

template <typename T, int STAGE, bool ONE_BARRIER, bool DEFER>
__device__ __forceinline__ void ens_item(const StageArgs& a, const WorkItem& w, char* lds,
                                         CflAcc& cfl)
{
    switch (w.flags & 3) {
    case 0:
        stage_item<T, false, false, STAGE, ONE_BARRIER, DEFER>(a, w.tile_x, w.tile_y, w.zb, w.ze, lds, cfl);
        break;
    case 1:
        stage_item<T, true, false, STAGE, ONE_BARRIER, DEFER>(a, w.tile_x, w.tile_y, w.zb, w.ze, lds, cfl);
        break;
    case 2:
        stage_item<T, false, true, STAGE, ONE_BARRIER, DEFER>(a, w.tile_x, w.tile_y, w.zb, w.ze, lds, cfl);
        break;
    default:
        stage_item<T, true, true, STAGE, ONE_BARRIER, DEFER>(a, w.tile_x, w.tile_y, w.zb, w.ze, lds, cfl);
        break;
    }
}

// Persistent interior stage over all members: block b takes items b, b + G, ...
template <typename T, int STAGE, int WAVES, bool ONE_BARRIER, bool DEFER>
__global__ void __launch_bounds__(256, WAVES)
    ensemble_stage_kernel(const StageArgs* __restrict__ args, const WorkItem* __restrict__ items,
                          int nitems)
{
    // sized for MHD items; a hydro item's planes and ring fit inside
    __shared__ __align__(16) char lds[StageLds<T, true, STAGE, ONE_BARRIER, DEFER>::bytes];
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const WorkItem w = items[it];
        const StageArgs& a = args[w.member];
        if (domain_halted(a)) continue;   // uniform over the block
        CflAcc cfl;
        ens_item<T, STAGE, ONE_BARRIER, DEFER>(a, w, lds, cfl);
        if (STAGE == 3 && a.cfl_parts) cfl_block_reduce(cfl, a.cfl_parts, w.part);
        __syncthreads();   // the next item reuses the LDS ring
    }
}

// Persistent interior stage over the members whose interior points take
// off-default branches (FP64 fluxes): every branch compiled in, as the
// single-domain kernel that holds them all (one wave per SIMD, one barrier per
// plane, the deferred update); each point takes its member's branches from its
// ExtArgs.  S32: members of the mixed form (fluxes from the FP32 copies of the
// stage values with the background added back, as their runs alone).
template <int STAGE, bool S32 = false>
__global__ void __launch_bounds__(256, 1)
    ensemble_stage_ext_kernel(const StageArgs* __restrict__ args, const WorkItem* __restrict__ items,
                              int nitems)
{
    __shared__ __align__(16) char lds[StreamLds<double, true, true, true>::bytes];
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const WorkItem w = items[it];
        const StageArgs& a = args[w.member];
        if (domain_halted(a)) continue;   // uniform over the block
        CflAcc cfl;
        switch (w.flags & 3) {
        case 0:
            stage_item<double, false, false, STAGE, true, true, true, EXT_INTERIOR, S32>(a, w.tile_x, w.tile_y, w.zb,
                                                                                   w.ze, lds, cfl);
            break;
        case 1:
            stage_item<double, true, false, STAGE, true, true, true, EXT_INTERIOR, S32>(a, w.tile_x, w.tile_y, w.zb,
                                                                                  w.ze, lds, cfl);
            break;
        case 2:
            stage_item<double, false, true, STAGE, true, true, true, EXT_INTERIOR, S32>(a, w.tile_x, w.tile_y, w.zb,
                                                                                  w.ze, lds, cfl);
            break;
        default:
            stage_item<double, true, true, STAGE, true, true, true, EXT_INTERIOR, S32>(a, w.tile_x, w.tile_y, w.zb,
                                                                                 w.ze, lds, cfl);
            break;
        }
        if (STAGE == 3 && a.cfl_parts) cfl_block_reduce(cfl, a.cfl_parts, w.part);
        __syncthreads();   // the next item reuses the LDS ring
    }
}

// Persistent wall stage over the members (blocks of 64 x 4 threads).  EXT:
// the members with off-default branches (FP64 fluxes; S32: of the mixed
// form), in a list and a launch of their own so the others keep the smaller
// register budget.
template <typename T, int STAGE, bool EXT = false, bool S32 = false>
__global__ void __launch_bounds__(256)
    ensemble_wall_kernel(const StageArgs* __restrict__ args, const WallItem* __restrict__ items,
                         int nitems)
{
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const WallItem w = items[it];
        const StageArgs& a = args[w.member];
        if (domain_halted(a)) continue;   // uniform over the block
        const int x = w.x0 + threadIdx.x, y = w.y0 + threadIdx.y;
        CflAcc cfl;
        if (x < a.g.nPx && y < a.g.nPy) {
            const bool top = w.top != 0;
            switch (w.flags & 3) {
            case 0: wall_point<T, false, false, STAGE, EXT, S32>(a, x, y, top, cfl); break;
            case 1: wall_point<T, true, false, STAGE, EXT, S32>(a, x, y, top, cfl); break;
            case 2: wall_point<T, false, true, STAGE, EXT, S32>(a, x, y, top, cfl); break;
            default: wall_point<T, true, true, STAGE, EXT, S32>(a, x, y, top, cfl); break;
            }
        }
        if (STAGE == 3 && a.cfl_parts) cfl_block_reduce(cfl, a.cfl_parts, w.part);
    }
}

// Closes the step of every member (block m = member m): TIMC += DT,
// nIt += 1, the DT guard and the step budget, then -- unless the guard
// halted it -- the next DT and the non-finite flag from the member's TSTEP
// partials (as step_close_kernel of a single domain).  A stopped member
// keeps its scalars.
MHD_TU_LOCAL __global__ void __launch_bounds__(256) ensemble_finish_kernel(const FinishArgs* __restrict__ f)
{
    const FinishArgs& m = f[blockIdx.x];
    if (m.ds->halted) {
        if (threadIdx.x == 0) *m.mirror = *m.ds;
        return;
    }
    __syncthreads();   // every thread has read the flag before thread 0 may set it
    if (threadIdx.x == 0) clock_and_guard(m.ds);
    __syncthreads();
    if (!(m.ds->halted & HALT_GUARD)) cfl_finalize_block(m.parts, m.nparts, m.ds, m.c, m.mag);
    if (threadIdx.x == 0) *m.mirror = *m.ds;   // thread 0 wrote the scalars
}

// Start of a step call for every member: non-finite flag cleared, nIt at
// the start recorded (as call_begin_kernel of a single domain).
MHD_TU_LOCAL __global__ void ensemble_clear_kernel(const FinishArgs* __restrict__ f, int n)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < n) {
        f[m].ds->nonfinite = 0u;
        f[m].ds->nit_call0 = f[m].ds->nit;
        *f[m].mirror = *f[m].ds;
    }
}

}  // namespace mhd
