// C ABI of the MI355X RK3 timestep (include/mhd3d.h).
//
// One context = one domain on one GPU: the device registers, the metric
// arrays, a block of device scalars and one HIP stream.  FP64 form: two U
// buffers used as ping-pong and one Z.  Mixed form: one FP64 U (the state,
// read by stage 1 and rewritten by stage 3), the FP64 Z and two FP32 copies
// of the stage values (written by stages 1 and 2, read by stages 2 and 3).
// A step is: CFL (unless fixed DT or already evaluated inside the previous
// stage 3) -> 3 x (interior stage kernel, wall stage kernel) -> clock update,
// with DT kept on the device so the host never waits inside a step.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mhd3d.h"
#include "kernels_basic.hpp"
#include "kernels_stream.hpp"
#include "kernels_ensemble.hpp"
#include "mixed_launch.hpp"
#include "ext_launch.hpp"

using namespace mhd;

struct mhd3d_ctx {
    mhd3d_desc d;
    int nf = 5;
    Geom g{};
    Phys c{};
    Metric m{};
    double* dmet = nullptr;
    double* U[2] = {nullptr, nullptr};
    double* Z = nullptr;
    bool mixed = false;             // MHD3D_MIXED: FP32 stencils, FP64 registers
    float* U32[2] = {nullptr, nullptr};
    float* bgf = nullptr;           // mixed: z-only background (BG_*), FP32
    double* bg64 = nullptr;         //        and FP64 (BG64_*)
    std::vector<double> h_dzz, h_d2zz;   // host copies of dzzdz, d2zzdz2
    double* F = nullptr;
    DevScalars* ds = nullptr;
    DevScalars* hs = nullptr;       // pinned host copy of *ds (read back after a step call)
    CflPart* cfl_parts = nullptr;
    int cfl_parts_cap = 0;
    bool dt_ready = false;          // ds->dt holds TSTEP of the current state
    double* host_stage = nullptr;   // pinned, one field
    hipEvent_t slab_ev[4] = {};     // last copy through each slab of host_stage
    hipStream_t stream = nullptr;
    int cur = 0;
    int device = 0;
    int kernel_mode = 0;            // 0 = z-streaming LDS kernel, 1 = point kernel
    // streaming-kernel variant per [form][physics] (form 0 = FP64, 1 = mixed;
    // physics 0 = hydro, 1 = MHD): register budget in waves per SIMD, and
    // schedule bits (1: one barrier per plane, 2: deferred RK update).
    // Defaults: the fastest variants measured on MI355X at 256x256x128 and
    // 512x512x256 (launches sized to the resident block count).
    int tune_waves[2][2] = {{1, 1}, {2, 2}};
    int tune_sched[2][2] = {{3, 2}, {2, 2}};
    bool fuse_cfl = true;           // TSTEP evaluated inside the stage-3 launch
    // off-default branches (iD, lRem, lShr, iTCon 2, TP plume, RLAX): FP64, evaluated
    // by the point and wall kernels
    ExtArgs xa{};
    bool ext = false;
    bool ext_timedep = false;       // heat loss or plume with TC != 0: factors of TIMT per step
    double* ext_buf = nullptr;      // means [2 nz] | wwz2 [nz] | heat x, y, z | plume x, y
    double* ft = nullptr;           // lShr: the FT register (field layout)
    // RLAX: FCONM [nz] | CORRECT [nz] + RLAX*TFAC | rank end values [nz], the
    // relaxation statics, and their constants
    double* relax_buf = nullptr;
    RelaxState* rs = nullptr;
    RelaxArgs ra{};
    bool relax_exit_seen = false;   // the exit's register has been settled on the host
    int relax_exit_code = 0;        // RelaxState::exited as last read
    int target_blocks = 0;          // work items of one streaming launch (z chunking);
                                    // 0: the resident blocks of the kernel variant
    int64_t bytes = 0;
    std::string err;
    // timing
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    double stage_ms = 0.0;
    int64_t stage_launches = 0;
    double step_ms = 0.0;
    int64_t steps_timed = 0;
};

static std::string g_create_error;

#define HIPCHK(ctx, call)                                                              \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);           \
            return MHD3D_EHIP;                                                         \
        }                                                                              \
    } while (0)

static int fail(mhd3d_ctx* ctx, int code, const std::string& msg)
{
    if (ctx) ctx->err = msg;
    return code;
}

// ---------------------------------------------------------------------------

// The off-default branches a descriptor switches on (ExtArgs::flags).
static int ext_flags(const mhd3d_desc* d)
{
    int f = 0;
    if (d->flag_iD != 0 && d->DH != 0.0) f |= EXT_DH;   // `#if DH != 0.0` (3dmhdsub.h:253)
    if (d->flag_iD != 0 && d->DV != 0.0) f |= EXT_DV;   // `#if DV != 0.0` (:259)
    if (d->flag_lRem) f |= EXT_REM;
    if (d->flag_lShr) f |= EXT_SHR;
    if (d->flag_iTCon == 2) f |= EXT_HEAT;
    if ((d->flag_iTCon == 0 || d->flag_iTCon == 1) && d->TP != 0.0) f |= EXT_PLUME;
    if (d->flag_iD != 0 || d->flag_lRem) f |= EXT_ROCOPY;
    if (d->RLAX > 0.0) f |= EXT_RLAX;                   // `if (RLAX > 0.0)` (3dmhdsub.h:322)
    return f;
}

// Branches whose factors of TIMT the host forms before every step (the heat
// loss and the plume with TC != 0): single domains only, as ensembles and
// the server advance many steps without the host.
static int timedep_flags(const mhd3d_desc* d)
{
    return d->TC != 0.0 ? (ext_flags(d) & (EXT_HEAT | EXT_PLUME)) : 0;
}
static const char* const kTimedepWhy =
    "the heat loss and the plume with TC != 0 take factors of TIMT formed on the host before every "
    "step (3dmhdsub.h:446-460,752-757): single domains only";

static int validate(const mhd3d_desc* d, std::string& why)
{
    if (!d) { why = "null descriptor"; return MHD3D_EINVAL; }
    if (d->nPx < 1 || d->nPy < 1) { why = "nPx and nPy must be >= 1"; return MHD3D_EINVAL; }
    if (d->nPz < 5) { why = "nPz must be >= 5 (one-sided wall stencils reach 3 planes in)"; return MHD3D_EINVAL; }
    if (!d->dxxdx || !d->d2xxdx2 || !d->DDX || !d->dyydy || !d->d2yydy2 || !d->DDY ||
        !d->dzzdz || !d->d2zzdz2 || !d->DDZ || !d->RKAPA || !d->DKAPA) {
        why = "metric array pointer missing";
        return MHD3D_EINVAL;
    }
    if (d->flag_ixCon != 0 || d->flag_iyCon != 0) { why = "fluxes: Non-periodic horizontal boundaries (3dmhdsub.h:208)"; return MHD3D_EUNSUPPORTED; }
    if (d->RLAX > 0.0) {
        if (d->relax_nPEz < 2 || d->nPz % d->relax_nPEz != 0 || d->nPz / d->relax_nPEz < 3) {
            why = "RLAX > 0 needs relax_nPEz: the z ranks of the reference run (>= 2, dividing nPz, "
                  "at least 3 planes each; the relaxation scan of 3dmhdsub.h:354-421 runs per rank)";
            return MHD3D_EINVAL;
        }
        if (!d->ZEE) { why = "RLAX > 0 needs the coordinate array ZEE (3dmhdsub.h:387)"; return MHD3D_EINVAL; }
    }
    if (d->flag_iD < 0 || d->flag_iD > 2) { why = "iD must be 0, 1 or 2 (3dmhdsub.h:243-252)"; return MHD3D_EINVAL; }
    if (d->flag_iTCon < 0 || d->flag_iTCon > 4) { why = "invalid iTCon"; return MHD3D_EINVAL; }
    if (d->flag_izCon != 0 && d->flag_izCon != 1) { why = "BCON:  Invalid izCon (3dmhdsub.h:863)"; return MHD3D_EINVAL; }
    const int ef = ext_flags(d);
    if (((ef & (EXT_DH | EXT_DV)) && d->flag_iD == 2 && !d->ZEE) ||
        ((ef & EXT_HEAT) && (!d->EXX || !d->WYY || !d->ZEE)) || ((ef & EXT_PLUME) && (!d->EXX || !d->WYY))) {
        why = "coordinate array (EXX, WYY or ZEE) missing";
        return MHD3D_EINVAL;
    }
    if (d->flag_lMag && d->flag_iBCon != 0) { why = "iBCon != 0 is not built"; return MHD3D_EUNSUPPORTED; }
    if (d->precision != MHD3D_FP64 && d->precision != MHD3D_MIXED) { why = "invalid precision"; return MHD3D_EINVAL; }
    return MHD3D_OK;
}

static Phys make_phys(const mhd3d_desc* d)
{
    Phys c{};
    c.C13 = d->C13; c.C23 = d->C23; c.C43 = d->C43;
    c.GAM[0] = d->GAM1; c.GAM[1] = d->GAM2; c.GAM[2] = d->GAM3;
    c.ZETA[0] = d->ZETA1; c.ZETA[1] = d->ZETA2;
    c.SFF = d->SFF;
    c.hx = d->hx; c.h2x = d->h2x; c.hy = d->hy; c.h2y = d->h2y; c.hz = d->hz; c.h2z = d->h2z;
    c.ORZF = d->ORZF;
    c.RE = d->RE; c.ORE = d->ORE; c.REPR = d->REPR; c.CV = d->CV; c.OCV = d->OCV;
    c.GAMMA = d->GAMMA; c.OGAMMA = 1.0 / d->GAMMA; c.THETA = d->THETA; c.GRAV = d->GRAV;
    c.OMX = d->OMX; c.OMZ = d->OMZ; c.RM = d->RM; c.ORM = d->ORM; c.OBETA = d->OBETA;
    c.TP = d->TP; c.TB = d->TB;
    c.ORE_C43 = d->ORE * d->C43;
    c.C13_ORE = d->C13 * d->ORE;
    c.OCV_REPR = d->OCV / d->REPR;
    c.OCV_ORE = d->OCV * d->ORE;
    c.OCV_ORE_C43 = c.OCV_ORE * d->C43;
    c.OCV_ORE_2 = c.OCV_ORE * 2;
    c.C43_OCV_ORE = d->C43 * d->OCV * d->ORE;
    c.KJ = d->OBETA * d->ORM * d->OCV;
    c.KJ2 = -2.0 * d->OBETA * d->ORM * d->OCV;
    c.izCon = d->flag_izCon; c.iTCon = d->flag_iTCon; c.iBCon = d->flag_iBCon;
    c.xdim = (d->nPx + 2) > 3;   // `nx > ix + 1` (3dmhdsub.h:77)
    return c;
}

// Factors of TIMT the off-default branches take for one step, formed on the
// host with the reference's expressions: the heat-loss amplitude
// TP*OCV*0.5f*(1+tanh(log(2.0f)/QFH*(TIMT-TC))) (TP*OCV when TC = 0,
// 3dmhdsub.h:446-460) and the plume's TPR = TP + (1-TP)*exp(-TIMT/TC) (TP
// when TC = 0, :752-757).  TIMT = TIMI + TIMC is constant within a step.
static void step_factors(mhd3d_ctx* ctx, double TIMT)
{
    const mhd3d_desc& d = ctx->d;
    ExtArgs& X = ctx->xa;
    if (d.TC != 0.0) {
        X.heat_amp = d.TP * d.OCV * 0.5f * (1.0f + std::tanh(std::log(2.0f) / d.QFH * (TIMT - d.TC)));
        X.tpr = d.TP + (1.0f - d.TP) * std::exp(-TIMT / d.TC);
    } else {
        X.heat_amp = d.TP * d.OCV;
        X.tpr = d.TP;
    }
}

// Tables and registers of the off-default branches.  The factors that hold
// exp() and log() of the coordinates are formed here on the host with the
// reference's expressions (so with the same library function values): the
// iD 2 weights (3dmhdsub.h:244-252), the iTCon 2 heat-loss Gaussians
// exp(CLN*(EXX-XP)^2/HH) etc. (:443-461) and the plume Gaussians
// exp(CLN*(EXX-XP)^2) (:758-772), CLN = -4.0f*log(2.0f)/HH/HH.  log(2.0f) is
// the float logarithm: the reference's translation unit includes <math.h>
// (3dmhdphsic.h:11) before these functions, which brings the float overload
// into the global namespace.
static int setup_ext(mhd3d_ctx* ctx, const mhd3d_desc* d)
{
    ExtArgs& X = ctx->xa;
    X = ExtArgs{};
    X.flags = ext_flags(d);
    X.iD = d->flag_iD;
    X.lRem = d->flag_lRem ? 1 : 0;
    X.lShr = d->flag_lShr ? 1 : 0;
    X.TU = d->TU;
    X.DZTU = d->DZTU;
    X.DZTB = d->DZTB;
    X.ORZD = d->ORZD;
    X.HH = d->HH;
    X.TIMI = d->TIMI;
    X.cdh = d->DH * d->ORE;
    X.cdv = d->DV * d->ORE;
    ctx->ext = X.flags != 0;
    if (!ctx->ext) return MHD3D_OK;
    ctx->ext_timedep = d->TC != 0.0 && (X.flags & (EXT_HEAT | EXT_PLUME));
    const Geom& g = ctx->g;
    const int nx = g.nPx + 2, ny = g.nPy + 2, nz = g.nz;
    // host image: means [2nz] | wwz2 [nz] | heat_x [nx] heat_y [ny] heat_z [nz] | plume_x [nx] plume_y [ny]
    const size_t n = 2 * (size_t)nz + nz + nx + ny + nz + nx + ny;
    std::vector<double> h(n, 0.0);
    double* wwz2 = h.data() + 2 * nz;
    double* hx = wwz2 + nz;
    double* hy = hx + nx;
    double* hz = hy + ny;
    double* px = hz + nz;
    double* py = px + nx;
    auto sq = [](double v) { return v * v; };
    if ((X.flags & (EXT_DH | EXT_DV)) && d->flag_iD == 2) {
        const double SGM = 1.0f;
        const double CLN = -4.0f * std::log(2.0f) / SGM / SGM;
        for (int z = 0; z < nz; ++z)
            wwz2[z] = std::exp(CLN * sq(d->ZEE[z])) + std::exp(CLN * sq(d->ZEE[z] - d->zMax));
    }
    if (X.flags & EXT_HEAT) {
        const double HH = d->HH;
        const double CLN = -4.0f * std::log(2.0f) / HH / HH;
        for (int x = 0; x < nx; ++x) hx[x] = std::exp(CLN * sq(d->EXX[x] - d->XP) / HH);
        for (int y = 0; y < ny; ++y) hy[y] = std::exp(CLN * sq(d->WYY[y] - d->YP) / HH);
        for (int z = 0; z < nz; ++z) hz[z] = std::exp(CLN * sq(d->ZEE[z] - d->ZP) / HH);
    }
    if (X.flags & EXT_PLUME) {
        const double HH = d->HH;
        const double CLN = -4.0f * std::log(2.0f) / HH / HH;
        for (int x = 0; x < nx; ++x) px[x] = std::exp(CLN * sq(d->EXX[x] - d->XP));
        for (int y = 0; y < ny; ++y) py[y] = std::exp(CLN * sq(d->WYY[y] - d->YP));
    }
    HIPCHK(ctx, hipMalloc(&ctx->ext_buf, n * sizeof(double)));
    HIPCHK(ctx, hipMemcpy(ctx->ext_buf, h.data(), n * sizeof(double), hipMemcpyHostToDevice));
    ctx->bytes += (int64_t)(n * sizeof(double));
    double* b = ctx->ext_buf;
    X.means = b;
    X.wwz2 = b + 2 * nz;
    X.heat_x = X.wwz2 + nz;
    X.heat_y = X.heat_x + nx;
    X.heat_z = X.heat_y + ny;
    X.plume_x = X.heat_z + nz;
    X.plume_y = X.plume_x + nx;
    if (X.flags & EXT_SHR) {
        // FT is a static array of the reference, zero at program start
        const size_t bytes = (size_t)g.fs * sizeof(double);
        HIPCHK(ctx, hipMalloc(&ctx->ft, bytes));
        HIPCHK(ctx, hipMemset(ctx->ft, 0, bytes));
        ctx->bytes += (int64_t)bytes;
        X.ft = ctx->ft;
    }
    if (X.flags & EXT_RLAX) {
        // FCONM [nz] | CORRECT [nz] and RLAX*TFAC | rank end values [nz] | ZEE [nz]
        // | relax_kernel's work arrays, when they exceed its LDS
        const size_t nb = 4 * (size_t)nz + 1 + RELAX_WORK(nz, d->relax_nPEz);
        std::vector<double> hb(nb, 0.0);
        std::memcpy(&hb[3 * (size_t)nz + 1], d->ZEE, nz * sizeof(double));
        HIPCHK(ctx, hipMalloc(&ctx->relax_buf, nb * sizeof(double)));
        HIPCHK(ctx, hipMemcpy(ctx->relax_buf, hb.data(), nb * sizeof(double), hipMemcpyHostToDevice));
        RelaxState r0{};
        r0.TFAC = 1.0;          // the statics' initial values (3dmhdsub.h:196-198)
        r0.TTM_MAX = -1e99;
        r0.TTM_MIN = 1e99;
        HIPCHK(ctx, hipMalloc(&ctx->rs, sizeof(RelaxState)));
        HIPCHK(ctx, hipMemcpy(ctx->rs, &r0, sizeof(RelaxState), hipMemcpyHostToDevice));
        ctx->bytes += (int64_t)(nb * sizeof(double) + sizeof(RelaxState));
        X.relax = ctx->relax_buf + nz;
        RelaxArgs& ra = ctx->ra;
        ra.RLAX = d->RLAX;
        ra.TSTART = d->TSTART;
        ra.TOFF = d->TOFF;
        ra.REPR = d->REPR;
        ra.hz = d->hz;
        ra.TIMI = d->TIMI;
        if (d->flag_lRem) {
            const float k = 8.07f * 0.4f;   // FTOT = 8.07f * 0.4f * THETA: a float product first (:379)
            ra.FTOT = (double)k * d->THETA;
        } else {
            ra.FTOT = d->THETA / d->REPR;   // :381
        }
        ra.combo = (d->flag_iTCon == 0 && d->flag_izCon == 1) ? 1 : (d->flag_iTCon == 1 && d->flag_izCon == 0) ? 2 : 0;
        ra.P = d->relax_nPEz;
        ra.nRz = d->nPz / d->relax_nPEz;
        ra.dzzdz = ctx->m.dzzdz;
        ra.RKAPA = ctx->m.RKAPA;
        ra.ZEE = ctx->relax_buf + 3 * (size_t)nz + 1;
    }
    step_factors(ctx, d->TIMI);
    return MHD3D_OK;
}

// Resident blocks of the streaming stage kernel a context runs with the
// register budget `w` (waves per SIMD) and schedule `sc`, on the whole device.
static int variant_resident(const mhd3d_ctx* ctx, int w, int sc)
{
    const bool mag = ctx->d.flag_lMag != 0;
    const void* fn;
    if (ctx->mixed && ctx->ext)
        fn = ext_stream_fn_s32(mag, ext_variant(ctx->xa.flags));
    else if (ctx->mixed)
        fn = mixed_stream_fn(mag, w, sc);
    else if (ext_variant(ctx->xa.flags) != EXTV_NONE)
        fn = ext_stream_fn(mag, ext_variant(ctx->xa.flags));
    else
        fn = mag ? stream_fn<double, true, true>(w, sc) : stream_fn<double, false, true>(w, sc);
    return resident_blocks(fn, STX * STY);
}

// Blocks of one streaming launch: the resident blocks, or (planes > 0) z
// chunks of at most `planes` planes in rounds of them -- at most 8 rounds,
// which bounds the TSTEP partial records.
static int launch_blocks(const mhd3d_ctx* ctx, int resident, int planes)
{
    if (planes <= 0) return resident;
    const StreamPlan one = stream_plan(ctx->g, 1);
    const int want = one.ntx * one.nty * ((ctx->g.nPz - 2 + planes - 1) / planes);
    return std::max(resident, std::min(want, 8 * resident));
}

extern "C" int mhd3d_create(const mhd3d_desc* desc, mhd3d_ctx** out)
{
    if (!out) return MHD3D_EINVAL;
    *out = nullptr;
    std::string why;
    int rc = validate(desc, why);
    if (rc != MHD3D_OK) { g_create_error = why; return rc; }
    mhd3d_ctx* ctx = new mhd3d_ctx();
    ctx->d = *desc;
    ctx->nf = desc->flag_lMag ? NF_MAG : NF_HYDRO;
    Geom& g = ctx->g;
    g.nPx = desc->nPx; g.nPy = desc->nPy; g.nPz = desc->nPz;
    g.nz = desc->nPz + 2;
    g.pitch = (desc->nPx + 7) / 8 * 8;
    g.plane = (int64_t)g.pitch * g.nPy;
    g.fs = (g.plane * g.nz + 31) / 32 * 32;   // 256-byte aligned fields
    ctx->c = make_phys(desc);
    auto cleanup = [&](int code, const std::string& m) {
        g_create_error = m;
        mhd3d_destroy(ctx);
        return code;
    };
    if (hipGetDevice(&ctx->device) != hipSuccess) return cleanup(MHD3D_EHIP, "hipGetDevice failed");
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(MHD3D_EHIP, "hipStreamCreate failed");
    const int nx = g.nPx + 2, ny = g.nPy + 2, nz = g.nz;
    // metric block: TSTEP records (16 doubles per plane) | dxxdx d2xxdx2 DDX |
    // dyydy d2yydy2 DDY | dzzdz d2zzdz2 DDZ RKAPA DKAPA KCFL
    const size_t nrec = (size_t)(sizeof(CflRec) / sizeof(double)) * nz;
    const size_t nmet = nrec + 3 * (size_t)nx + 3 * (size_t)ny + 6 * (size_t)nz;
    std::vector<double> hm(nmet, 0.0);
    size_t o = nrec;
    auto put = [&](const double* src, int n) { std::memcpy(&hm[o], src, n * sizeof(double)); o += n; };
    put(desc->dxxdx, nx); put(desc->d2xxdx2, nx); put(desc->DDX, nx);
    put(desc->dyydy, ny); put(desc->d2yydy2, ny); put(desc->DDY, ny);
    put(desc->dzzdz, nz); put(desc->d2zzdz2, nz); put(desc->DDZ, nz);
    put(desc->RKAPA, nz); put(desc->DKAPA, nz);
    // TSTEP conduction denominator: RKAPA[z], or 1.0f + RKAPA[z] with lRem (3dmhdsub.h:78-82)
    for (int z = 0; z < nz; ++z) hm[o++] = desc->flag_lRem ? 1.0f + desc->RKAPA[z] : desc->RKAPA[z];
    for (int z = 0; z < nz; ++z) {
        CflRec r{};
        r.DDZ = desc->DDZ[z];
        r.KCFL = hm[o - nz + z];
        r.OGAMMA = ctx->c.OGAMMA;
        r.GAMMA = ctx->c.GAMMA;
        r.OBETA = ctx->c.OBETA;
        r.REPR = ctx->c.REPR;
        r.CV = ctx->c.CV;
        r.RE = ctx->c.RE;
        r.RM = ctx->c.RM;
        std::memcpy(&hm[(size_t)z * (sizeof(CflRec) / sizeof(double))], &r, sizeof(CflRec));
    }
    if (hipMalloc(&ctx->dmet, nmet * sizeof(double)) != hipSuccess) return cleanup(MHD3D_ENOMEM, "hipMalloc(metrics) failed");
    if (hipMemcpy(ctx->dmet, hm.data(), nmet * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(MHD3D_EHIP, "metric upload failed");
    double* p = ctx->dmet;
    ctx->m.cflz = reinterpret_cast<const CflRec*>(p);
    p += nrec;
    ctx->m.dxxdx = p; p += nx; ctx->m.d2xxdx2 = p; p += nx; ctx->m.DDX = p; p += nx;
    ctx->m.dyydy = p; p += ny; ctx->m.d2yydy2 = p; p += ny; ctx->m.DDY = p; p += ny;
    ctx->m.dzzdz = p; p += nz; ctx->m.d2zzdz2 = p; p += nz; ctx->m.DDZ = p; p += nz;
    ctx->m.RKAPA = p; p += nz; ctx->m.DKAPA = p; p += nz;
    ctx->m.KCFL = p; p += nz;
    rc = setup_ext(ctx, desc);
    if (rc != MHD3D_OK) return cleanup(rc, ctx->err);
    ctx->mixed = desc->precision == MHD3D_MIXED;
    const size_t reg = (size_t)g.fs * ctx->nf * sizeof(double);
    const size_t reg32 = (size_t)g.fs * ctx->nf * sizeof(float);
    const int nU64 = ctx->mixed ? 1 : 2;
    for (int k = 0; k < nU64; ++k) {
        if (hipMalloc(&ctx->U[k], reg) != hipSuccess) return cleanup(MHD3D_ENOMEM, "hipMalloc(U) failed");
        if (hipMemset(ctx->U[k], 0, reg) != hipSuccess) return cleanup(MHD3D_EHIP, "hipMemset failed");
    }
    if (ctx->mixed) {
        for (int k = 0; k < 2; ++k) {
            if (hipMalloc(&ctx->U32[k], reg32) != hipSuccess) return cleanup(MHD3D_ENOMEM, "hipMalloc(U32) failed");
            if (hipMemset(ctx->U32[k], 0, reg32) != hipSuccess) return cleanup(MHD3D_EHIP, "hipMemset failed");
        }
        if (hipMalloc(&ctx->bgf, sizeof(float) * NBGF * nz) != hipSuccess ||
            hipMalloc(&ctx->bg64, sizeof(double) * NBG64 * nz) != hipSuccess)
            return cleanup(MHD3D_ENOMEM, "hipMalloc(background) failed");
        if (hipMemset(ctx->bgf, 0, sizeof(float) * NBGF * nz) != hipSuccess ||
            hipMemset(ctx->bg64, 0, sizeof(double) * NBG64 * nz) != hipSuccess)
            return cleanup(MHD3D_EHIP, "hipMemset failed");
        ctx->h_dzz.assign(desc->dzzdz, desc->dzzdz + nz);
        ctx->h_d2zz.assign(desc->d2zzdz2, desc->d2zzdz2 + nz);
    }
    if (hipMalloc(&ctx->Z, reg) != hipSuccess) return cleanup(MHD3D_ENOMEM, "hipMalloc(Z) failed");
    if (hipMalloc(&ctx->ds, sizeof(DevScalars)) != hipSuccess) return cleanup(MHD3D_ENOMEM, "hipMalloc(scalars) failed");
    if (hipMemset(ctx->Z, 0, reg) != hipSuccess || hipMemset(ctx->ds, 0, sizeof(DevScalars)) != hipSuccess)
        return cleanup(MHD3D_EHIP, "hipMemset failed");
    if (hipHostMalloc(&ctx->host_stage, (size_t)g.fs * sizeof(double)) != hipSuccess ||
        hipHostMalloc(&ctx->hs, sizeof(DevScalars)) != hipSuccess)
        return cleanup(MHD3D_ENOMEM, "hipHostMalloc failed");
    ctx->bytes += (int64_t)((nU64 + 1) * reg + (ctx->mixed ? 2 * reg32 : 0) + nmet * sizeof(double) +
                            sizeof(DevScalars));
    const char* km = std::getenv("MHD3D_KERNEL");
    if (km && std::strcmp(km, "point") == 0) ctx->kernel_mode = 1;
    // MHD3D_TUNE_<FP64|MIXED>_<HYDRO|MHD>=waves,sched overrides a variant
    static const char* tune_env[2][2] = {{"MHD3D_TUNE_FP64_HYDRO", "MHD3D_TUNE_FP64_MHD"},
                                         {"MHD3D_TUNE_MIXED_HYDRO", "MHD3D_TUNE_MIXED_MHD"}};
    for (int fm = 0; fm < 2; ++fm)
        for (int ph = 0; ph < 2; ++ph) {
            const char* v = std::getenv(tune_env[fm][ph]);
            int w = 0, sc = 0;
            if (v && std::sscanf(v, "%d,%d", &w, &sc) == 2 && (w == 1 || w == 2) && sc >= 0 && sc <= 3) {
                ctx->tune_waves[fm][ph] = w;
                ctx->tune_sched[fm][ph] = sc;
            }
        }
    const char* tb = std::getenv("MHD3D_TARGET_BLOCKS");
    if (tb && std::atoi(tb) >= 1) ctx->target_blocks = std::atoi(tb);
    const char* fc = std::getenv("MHD3D_FUSE_CFL");
    if (fc && std::atoi(fc) == 0) ctx->fuse_cfl = false;
    if (ctx->mixed) ctx->kernel_mode = 0;   // the mixed form has only the streaming kernel
    if (ctx->target_blocks == 0) {
        const int fm = ctx->mixed ? 1 : 0, ph = desc->flag_lMag ? 1 : 0;
        // FP64: z chunks of at most 64 planes (MI355X, FP64 MHD 512x512x256
        // stage kernel: one chunk 3.47 ms, two 3.40, four 3.35, eight 3.74;
        // 256x256x128 is best at its resident count, two chunks of 63); the
        // mixed form: one round of its resident blocks (its best)
        ctx->target_blocks = launch_blocks(ctx, variant_resident(ctx, ctx->tune_waves[fm][ph], ctx->tune_sched[fm][ph]),
                                           ctx->mixed ? 0 : 64);
    }
    *out = ctx;
    return MHD3D_OK;
}

extern "C" int mhd3d_destroy(mhd3d_ctx* ctx)
{
    if (!ctx) return MHD3D_OK;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto e : ctx->ev_pool) (void)hipEventDestroy(e);
    (void)hipFree(ctx->dmet);
    (void)hipFree(ctx->U[0]);
    (void)hipFree(ctx->U[1]);
    (void)hipFree(ctx->Z);
    (void)hipFree(ctx->U32[0]);
    (void)hipFree(ctx->U32[1]);
    (void)hipFree(ctx->bgf);
    (void)hipFree(ctx->bg64);
    (void)hipFree(ctx->F);
    (void)hipFree(ctx->ds);
    (void)hipFree(ctx->cfl_parts);
    (void)hipFree(ctx->ext_buf);
    (void)hipFree(ctx->ft);
    (void)hipFree(ctx->relax_buf);
    (void)hipFree(ctx->rs);
    for (hipEvent_t& ev : ctx->slab_ev)
        if (ev) (void)hipEventDestroy(ev);
    if (ctx->host_stage) (void)hipHostFree(ctx->host_stage);
    if (ctx->hs) (void)hipHostFree(ctx->hs);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return MHD3D_OK;
}

extern "C" const char* mhd3d_last_error(const mhd3d_ctx* ctx)
{
    return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

extern "C" int64_t mhd3d_device_bytes(const mhd3d_ctx* ctx) { return ctx ? ctx->bytes : 0; }

extern "C" int mhd3d_geometry(const mhd3d_ctx* ctx, int32_t out[4])
{
    if (!ctx || !out) return MHD3D_EINVAL;
    out[0] = ctx->g.nPx;
    out[1] = ctx->g.nPy;
    out[2] = ctx->g.nPz;
    out[3] = ctx->nf;
    return MHD3D_OK;
}

// ---------------------------------------------------------------------------
// host <-> device field transfer

// A field moves between the host array ([nz][ny][nx], ghosts included) and
// the device layout ([nz][nPy][pitch]) through the pinned staging buffer in
// kSlabs slabs of z planes: a team of host threads repacks one slab while
// the copy of another is in flight, and each slab's event orders the reuse
// of its part of the buffer.
constexpr int kSlabs = 4;

// fn(begin, end) over parts of [0, n) on up to 16 host threads (the CPU
// share of one GPU on a shared host), at least `grain` items a thread.
template <typename F>
static void host_parallel(int64_t n, int64_t grain, const F& fn)
{
    const unsigned hw = std::thread::hardware_concurrency();
    int64_t T = hw ? std::min<unsigned>(hw, 16u) : 4;
    T = std::min<int64_t>(T, (n + grain - 1) / grain);
    if (T <= 1) {
        fn((int64_t)0, n);
        return;
    }
    std::vector<std::thread> team;
    int64_t started = 1;
    try {
        for (; started < T; ++started) {
            const int64_t t = started;
            team.emplace_back([&fn, n, T, t] { fn(n * t / T, n * (t + 1) / T); });
        }
    } catch (...) {
        // fewer threads than asked: this thread takes the parts left over
    }
    fn((int64_t)0, n / T);
    for (int64_t t = started; t < T; ++t) fn(n * t / T, n * (t + 1) / T);
    for (std::thread& th : team) th.join();
}

static int slab_events(mhd3d_ctx* ctx)
{
    for (hipEvent_t& ev : ctx->slab_ev)
        if (!ev) HIPCHK(ctx, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    return MHD3D_OK;
}

// Planes [z0, z1) of slab s.
static void slab_planes(const Geom& g, int s, int& z0, int& z1)
{
    const int zs = (g.nz + kSlabs - 1) / kSlabs;
    z0 = std::min(g.nz, s * zs);
    z1 = std::min(g.nz, z0 + zs);
}

// The host array src into the device field dst (interior x, y of every
// plane; the stream orders the copies, the caller synchronizes it).
static int upload_field(mhd3d_ctx* ctx, double* dst, const double* src)
{
    const Geom& g = ctx->g;
    const int nx = g.nPx + 2, ny = g.nPy + 2;
    double* h = ctx->host_stage;
    int rc = slab_events(ctx);
    if (rc) return rc;
    for (int s = 0; s < kSlabs; ++s) {
        int z0, z1;
        slab_planes(g, s, z0, z1);
        if (z0 >= z1) break;
        HIPCHK(ctx, hipEventSynchronize(ctx->slab_ev[s]));   // the slab's last copy is done
        host_parallel(z1 - z0, 1, [&](int64_t p0, int64_t p1) {
            for (int64_t z = z0 + p0; z < z0 + p1; ++z)
                for (int y = 0; y < g.nPy; ++y) {
                    double* d = h + (size_t)z * g.plane + (size_t)y * g.pitch;
                    std::memcpy(d, src + ((size_t)z * ny + (y + 1)) * nx + 1, g.nPx * sizeof(double));
                    if (g.pitch > g.nPx) std::memset(d + g.nPx, 0, (g.pitch - g.nPx) * sizeof(double));
                }
        });
        const size_t off = (size_t)z0 * g.plane;
        HIPCHK(ctx, hipMemcpyAsync(dst + off, h + off, (size_t)(z1 - z0) * g.plane * sizeof(double),
                                   hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(ctx, hipEventRecord(ctx->slab_ev[s], ctx->stream));
    }
    return MHD3D_OK;
}

// nf device fields from src (field stride fs) into the host arrays dst[f]:
// the slab copies run ahead of the unpacking by one buffer.  periodic: the
// x and y ghosts take the periodic copies (3dmhdmpi.h:65-68,118-121,135-142)
// and every plane is written (the state); otherwise ghosts and the z ghost
// planes are zero (fluxes).
static int download_fields(mhd3d_ctx* ctx, double* const* dst, const double* src, int nf, bool periodic)
{
    const Geom& g = ctx->g;
    const int nx = g.nPx + 2, ny = g.nPy + 2;
    double* h = ctx->host_stage;
    int rc = slab_events(ctx);
    if (rc) return rc;
    int ns = 0;   // slabs with planes
    for (int s = 0; s < kSlabs; ++s) {
        int z0, z1;
        slab_planes(g, s, z0, z1);
        if (z0 < z1) ns = s + 1;
    }
    const int jobs = nf * ns;
    auto issue = [&](int j) -> int {
        const int f = j / ns, s = j % ns;
        int z0, z1;
        slab_planes(g, s, z0, z1);
        const size_t off = (size_t)z0 * g.plane;
        HIPCHK(ctx, hipMemcpyAsync(h + off, src + (size_t)f * g.fs + off, (size_t)(z1 - z0) * g.plane * sizeof(double),
                                   hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipEventRecord(ctx->slab_ev[s], ctx->stream));
        return MHD3D_OK;
    };
    for (int j = 0; j < std::min(ns, jobs); ++j) {
        rc = issue(j);
        if (rc) return rc;
    }
    for (int j = 0; j < jobs; ++j) {
        const int f = j / ns, s = j % ns;
        int z0, z1;
        slab_planes(g, s, z0, z1);
        HIPCHK(ctx, hipEventSynchronize(ctx->slab_ev[s]));
        double* out = dst[f];
        host_parallel(z1 - z0, 1, [&](int64_t p0, int64_t p1) {
            for (int64_t z = z0 + p0; z < z0 + p1; ++z) {
                double* pl = out + (size_t)z * ny * nx;
                if (!periodic && (z == 0 || z == g.nz - 1)) {
                    std::memset(pl, 0, (size_t)ny * nx * sizeof(double));
                    continue;
                }
                for (int y = 0; y < g.nPy; ++y) {
                    double* row = pl + (size_t)(y + 1) * nx;
                    std::memcpy(row + 1, h + (size_t)z * g.plane + (size_t)y * g.pitch, g.nPx * sizeof(double));
                    row[0] = periodic ? row[nx - 2] : 0.0;
                    row[nx - 1] = periodic ? row[1] : 0.0;
                }
                if (periodic) {
                    std::memcpy(pl, pl + (size_t)(ny - 2) * nx, nx * sizeof(double));
                    std::memcpy(pl + (size_t)(ny - 1) * nx, pl + (size_t)nx, nx * sizeof(double));
                } else {
                    std::memset(pl, 0, nx * sizeof(double));
                    std::memset(pl + (size_t)(ny - 1) * nx, 0, nx * sizeof(double));
                }
            }
        });
        if (j + ns < jobs) {
            rc = issue(j + ns);
            if (rc) return rc;
        }
    }
    return MHD3D_OK;
}

// Background of the mixed form from the state being set: RO0(z), TT0(z) (and
// BX0, BY0, BZ0 with lMag) are the horizontal means of the fields on every
// plane (ghost planes included), P0 = RO0*TT0.  For the planes 1..nPz the
// z differences the deviation-form fluxes add back are evaluated here in FP64
// with the stencils of 3dmhdutils.h:195-220 (and, for D2z(BX0), D2z(BY0) on
// the wall planes, the one-sided forms of 3dmhdsub.h:638-690); HB is the
// hydrostatic residual RO0*GRAV - D1z(P0).  The background stays fixed until
// the next set_state.
static int set_background_profiles(mhd3d_ctx* ctx, const std::vector<double>* mean);

static int set_background(mhd3d_ctx* ctx, const double* const* fields)
{
    const Geom& g = ctx->g;
    const int nx = g.nPx + 2, ny = g.nPy + 2, nz = g.nz;
    const bool mag = ctx->nf == NF_MAG;
    // means[k][z], k = RO, TT, BX, BY, BZ
    std::vector<double> mean[5];
    const int src_field[5] = {3, 4, 5, 6, 7};
    const int nk = mag ? 5 : 2;
    for (int k = 0; k < 5; ++k) mean[k].assign(nz, 0.0);
    // one plane's sum on one thread, in a fixed order (the same doubles
    // whatever the number of threads)
    host_parallel((int64_t)nk * nz, 1, [&](int64_t j0, int64_t j1) {
        for (int64_t j = j0; j < j1; ++j) {
            const int k = (int)(j / nz), z = (int)(j % nz);
            const double* F = fields[src_field[k]];
            double sum = 0.0;
            for (int y = 1; y <= g.nPy; ++y)
                for (int x = 1; x <= g.nPx; ++x) sum += F[((size_t)z * ny + y) * nx + x];
            mean[k][z] = sum / ((double)g.nPx * g.nPy);
        }
    });
    return set_background_profiles(ctx, mean);
}

// Background from given z profiles mean[k][z] (k = RO, TT, BX, BY, BZ).
static int set_background_profiles(mhd3d_ctx* ctx, const std::vector<double>* mean)
{
    const Geom& g = ctx->g;
    const int nz = g.nz;
    const std::vector<double>& R0 = mean[0];
    const std::vector<double>& T0 = mean[1];
    std::vector<float> bf((size_t)NBGF * nz, 0.0f);
    std::vector<double> b64((size_t)NBG64 * nz, 0.0);
    const Phys& c = ctx->c;
    auto d1z = [&](const std::vector<double>& f, int z) {
        return ((f[z + 1] - f[z - 1]) * ctx->h_dzz[z]) * c.hz;
    };
    auto d2z = [&](const std::vector<double>& f, int z) {
        const double dzz = ctx->h_dzz[z];
        return (f[z + 1] - f[z - 1]) * ctx->h_d2zz[z] * c.hz +
               ((f[z + 1] - 2.0 * f[z]) + f[z - 1]) * (dzz * dzz) * c.h2z;
    };
    auto d2z_wall = [&](const std::vector<double>& f, int z, bool top) {
        const double dzz = ctx->h_dzz[z];
        const int s = top ? 1 : -1;
        const double b0 = f[z], b1 = f[z + s], b2 = f[z + 2 * s], b3 = f[z + 3 * s];
        const double first = top ? (-3.0 * b0 + 4.0 * b1 - b2) : (3.0 * b0 - 4.0 * b1 + b2);
        return first * c.hz * ctx->h_d2zz[z] +
               (2.0 * b0 - 5.0 * b1 + 4.0 * b2 - b3) * c.h2z * dzz * dzz;
    };
    std::vector<double> P0(nz);
    for (int z = 0; z < nz; ++z) P0[z] = R0[z] * T0[z];
    for (int z = 0; z < nz; ++z) {
        bf[BG_R0 * nz + z] = (float)R0[z];
        bf[BG_T0 * nz + z] = (float)T0[z];
        b64[BG64_R0 * nz + z] = R0[z];
        b64[BG64_T0 * nz + z] = T0[z];
        for (int k = 0; k < 3; ++k) {
            bf[(BG_B0 + k) * nz + z] = (float)mean[2 + k][z];
            b64[(BG64_B0 + k) * nz + z] = mean[2 + k][z];
        }
        if (z >= 1 && z <= g.nPz) {
            bf[BG_DZT0 * nz + z] = (float)d1z(T0, z);
            bf[BG_D2ZT0 * nz + z] = (float)d2z(T0, z);
            bf[BG_HB * nz + z] = (float)(R0[z] * c.GRAV - d1z(P0, z));
            const bool wall = z == 1 || z == g.nPz;
            for (int k = 0; k < 3; ++k) {
                const std::vector<double>& B = mean[2 + k];
                bf[(BG_DZB + k) * nz + z] = (float)d1z(B, z);
                bf[(BG_D2ZB + k) * nz + z] =
                    (float)((wall && k < 2) ? d2z_wall(B, z, z == 1) : d2z(B, z));
            }
        }
    }
    HIPCHK(ctx, hipMemcpyAsync(ctx->bgf, bf.data(), bf.size() * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(ctx->bg64, b64.data(), b64.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return MHD3D_OK;
}

extern "C" int mhd3d_set_state(mhd3d_ctx* ctx, const double* const* fields)
{
    if (!ctx || !fields) return MHD3D_EINVAL;
    ctx->dt_ready = false;
    for (int f = 0; f < ctx->nf; ++f)
        if (!fields[f]) return fail(ctx, MHD3D_EINVAL, "null field pointer");
    for (int f = 0; f < ctx->nf; ++f) {
        int rc = upload_field(ctx, ctx->U[ctx->cur] + (size_t)f * ctx->g.fs, fields[f]);
        if (rc) return rc;
    }
    // the other U buffers need the z ghost planes too (stage kernels never
    // write them)
    const int64_t n = ctx->g.fs * ctx->nf;
    if (ctx->mixed) {
        int rc = set_background(ctx, fields);
        if (rc) return rc;
        for (int k = 0; k < 2; ++k)
            to_f32_dev_kernel<<<1024, 256, 0, ctx->stream>>>(ctx->U[0], ctx->U32[k], ctx->g, ctx->nf,
                                                            ctx->bg64);
        HIPCHK(ctx, hipGetLastError());
    } else {
        HIPCHK(ctx, hipMemcpyAsync(ctx->U[1 - ctx->cur], ctx->U[ctx->cur], (size_t)n * sizeof(double),
                                   hipMemcpyDeviceToDevice, ctx->stream));
    }
    clear_halt_kernel<<<1, 1, 0, ctx->stream>>>(ctx->ds);   // a new state runs again
    if (ctx->rs) relax_clear_exit_kernel<<<1, 1, 0, ctx->stream>>>(ctx->rs);
    ctx->relax_exit_seen = false;
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return MHD3D_OK;
}

// Device-side synthetic stratified state (init_stratified_kernel); the mixed
// form takes its background from the same profiles (B0 = (b0, 0, 0)).
extern "C" int mhd3d_init_stratified(mhd3d_ctx* ctx, const double* TT_z, const double* RO_z,
                                     uint64_t seed, double amp, double b0, double b_amp)
{
    if (!ctx || !TT_z || !RO_z) return MHD3D_EINVAL;
    const Geom& g = ctx->g;
    if (g.nPx >= (1 << 18) || g.nPy >= (1 << 18)) return fail(ctx, MHD3D_EINVAL, "init_stratified: nPx, nPy < 262144");
    ctx->dt_ready = false;
    const int nz = g.nz;
    double* prof = nullptr;
    HIPCHK(ctx, hipMalloc(&prof, sizeof(double) * 2 * nz));
    hipError_t e = hipMemcpyAsync(prof, TT_z, sizeof(double) * nz, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(prof + nz, RO_z, sizeof(double) * nz, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        dim3 blk(64, 4, 1);
        dim3 grd((g.nPx + 63) / 64, (g.nPy + 3) / 4, nz);
        init_stratified_kernel<<<grd, blk, 0, ctx->stream>>>(g, ctx->nf, ctx->U[ctx->cur], prof, prof + nz,
                                                             (unsigned long long)seed, amp, b0, b_amp);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(prof);
    if (e != hipSuccess) return fail(ctx, MHD3D_EHIP, std::string("init_stratified: ") + hipGetErrorString(e));
    const int64_t n = g.fs * ctx->nf;
    if (ctx->mixed) {
        std::vector<double> mean[5];
        mean[0].assign(RO_z, RO_z + nz);
        mean[1].assign(TT_z, TT_z + nz);
        mean[2].assign(nz, ctx->nf == NF_MAG ? b0 : 0.0);
        mean[3].assign(nz, 0.0);
        mean[4].assign(nz, 0.0);
        int rc = set_background_profiles(ctx, mean);
        if (rc) return rc;
        for (int k = 0; k < 2; ++k)
            to_f32_dev_kernel<<<1024, 256, 0, ctx->stream>>>(ctx->U[0], ctx->U32[k], ctx->g, ctx->nf, ctx->bg64);
        HIPCHK(ctx, hipGetLastError());
    } else {
        HIPCHK(ctx, hipMemcpyAsync(ctx->U[1 - ctx->cur], ctx->U[ctx->cur], (size_t)n * sizeof(double),
                                   hipMemcpyDeviceToDevice, ctx->stream));
    }
    clear_halt_kernel<<<1, 1, 0, ctx->stream>>>(ctx->ds);   // a new state runs again
    if (ctx->rs) relax_clear_exit_kernel<<<1, 1, 0, ctx->stream>>>(ctx->rs);
    ctx->relax_exit_seen = false;
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return MHD3D_OK;
}

extern "C" int mhd3d_get_state(mhd3d_ctx* ctx, double* const* fields)
{
    if (!ctx || !fields) return MHD3D_EINVAL;
    for (int f = 0; f < ctx->nf; ++f)
        if (!fields[f]) return fail(ctx, MHD3D_EINVAL, "null field pointer");
    return download_fields(ctx, fields, ctx->U[ctx->cur], ctx->nf, true);
}

// ---------------------------------------------------------------------------
// launches

// Operands of stage k (0: no stage, the current state only).
static StageArgs stage_args(mhd3d_ctx* ctx, int k = 0)
{
    StageArgs a;
    a.g = ctx->g;
    a.c = ctx->c;
    a.m = ctx->m;
    a.Uin = ctx->U[ctx->cur];
    a.Uout = ctx->mixed ? ctx->U[0] : ctx->U[1 - ctx->cur];
    a.Z = ctx->Z;
    a.dt = &ctx->ds->dt;
    a.ds = ctx->ds;
    a.Uin32 = nullptr;
    a.Uout32 = nullptr;
    a.bgf = ctx->bgf;
    a.bg64 = ctx->bg64;
    if (ctx->mixed) {
        // stage 1: U64 -> U32[0]; stage 2: U32[0] -> U32[1]; stage 3: U32[1] -> U64
        if (k == 1) a.Uout32 = ctx->U32[0];
        if (k == 2) { a.Uin32 = ctx->U32[0]; a.Uout32 = ctx->U32[1]; }
        if (k == 3) a.Uin32 = ctx->U32[1];
    }
    a.cfl_parts = nullptr;
    a.cfl_base = 0;
    a.x = ctx->xa;
    return a;
}

static hipEvent_t next_event(mhd3d_ctx* ctx)
{
    if (ctx->ev_used == ctx->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        ctx->ev_pool.push_back(e);
    }
    return ctx->ev_pool[ctx->ev_used++];
}

template <bool MAG, bool ROT, int STAGE>
static void launch_stage_t(mhd3d_ctx* ctx, const StageArgs& a)
{
    const Geom& g = ctx->g;
    dim3 blk(64, 4, 1);
    const int fm = ctx->mixed ? 1 : 0, ph = MAG ? 1 : 0;
    if (ctx->mixed && ctx->ext) {
        ext_stage_s32(MAG, ROT, STAGE, ext_variant(ctx->xa.flags), a, ctx->stream, ctx->target_blocks);
    } else if (ctx->mixed) {
        mixed_stage(MAG, ROT, STAGE, a, ctx->stream, ctx->tune_waves[fm][ph], ctx->tune_sched[fm][ph],
                    ctx->target_blocks);
    } else if (ctx->kernel_mode == 0 && ext_variant(ctx->xa.flags) != EXTV_NONE) {
        ext_stage(MAG, ROT, STAGE, ext_variant(ctx->xa.flags), a, ctx->stream, ctx->target_blocks);
    } else if (ctx->kernel_mode == 1) {
        dim3 grd((g.nPx + 63) / 64, (g.nPy + 3) / 4, g.nPz - 2);
        if (ctx->ext)
            stage_point_kernel<MAG, ROT, STAGE, true><<<grd, blk, 0, ctx->stream>>>(a);
        else
            stage_point_kernel<MAG, ROT, STAGE><<<grd, blk, 0, ctx->stream>>>(a);
    } else {
        launch_stream_stage<double, MAG, ROT, STAGE>(a, ctx->stream, ctx->tune_waves[fm][ph],
                                                     ctx->tune_sched[fm][ph], ctx->target_blocks);
    }
}

template <bool MAG, bool ROT, int STAGE>
static void launch_wall_t(mhd3d_ctx* ctx, const StageArgs& a)
{
    const Geom& g = ctx->g;
    dim3 blk(64, 4, 1);
    dim3 grd((g.nPx + 63) / 64, (g.nPy + 3) / 4, 2);
    if (ctx->mixed && ctx->ext)
        stage_wall_kernel<double, MAG, ROT, STAGE, true, true><<<grd, blk, 0, ctx->stream>>>(a);
    else if (ctx->mixed)
        mixed_wall(MAG, ROT, STAGE, a, grd, blk, ctx->stream);
    else if (ctx->ext)
        stage_wall_kernel<double, MAG, ROT, STAGE, true><<<grd, blk, 0, ctx->stream>>>(a);
    else
        stage_wall_kernel<double, MAG, ROT, STAGE><<<grd, blk, 0, ctx->stream>>>(a);
}

template <int STAGE>
static void launch_stage_k(mhd3d_ctx* ctx, const StageArgs& a, bool wall)
{
    const bool mag = ctx->d.flag_lMag, rot = ctx->d.flag_lRot;
    if (!wall) {
        if (mag && rot) launch_stage_t<true, true, STAGE>(ctx, a);
        else if (mag) launch_stage_t<true, false, STAGE>(ctx, a);
        else if (rot) launch_stage_t<false, true, STAGE>(ctx, a);
        else launch_stage_t<false, false, STAGE>(ctx, a);
    } else {
        if (mag && rot) launch_wall_t<true, true, STAGE>(ctx, a);
        else if (mag) launch_wall_t<true, false, STAGE>(ctx, a);
        else if (rot) launch_wall_t<false, true, STAGE>(ctx, a);
        else launch_wall_t<false, false, STAGE>(ctx, a);
    }
}

// Number of CFL partial records the fused stage-3 launch writes: one per
// block of the streaming kernel (or of the point kernel) and of the wall kernel.
static void cfl_part_layout(const mhd3d_ctx* ctx, int& nstream, int& nwall)
{
    const Geom& g = ctx->g;
    if (ctx->kernel_mode == 1)
        nstream = ((g.nPx + 63) / 64) * ((g.nPy + 3) / 4) * (g.nPz - 2);
    else
        nstream = stream_plan(g, ctx->target_blocks).nblocks;
    nwall = ((g.nPx + 63) / 64) * ((g.nPy + 3) / 4) * 2;
}

static int ensure_cfl_parts(mhd3d_ctx* ctx, int n)
{
    if (n > ctx->cfl_parts_cap) {
        (void)hipFree(ctx->cfl_parts);
        ctx->cfl_parts = nullptr;
        HIPCHK(ctx, hipMalloc(&ctx->cfl_parts, sizeof(CflPart) * n));
        ctx->cfl_parts_cap = n;
    }
    return MHD3D_OK;
}

// Whether fluxes() of this domain needs plane means or the relaxation pass.
static bool needs_prepass(const mhd3d_ctx* ctx)
{
    return (ctx->xa.flags & (EXT_DH | EXT_DV | EXT_REM | EXT_RLAX)) != 0;
}

// What a fluxes() evaluation on register U needs before its point kernels:
// the horizontal means of RO and TT when a branch reads them (iD: ROM;
// lRem, RLAX: TTM -- which under lShr is never taken, 3dmhdsub.h:293-302, so
// it stays zero), then for RLAX the mean convective flux FCONM and the
// relaxation's scalar pass (TFAC, CORRECT, the exit test).  stage = 1..3
// inside a step (a stopped domain skips it), 0 for mhd3d_fluxes.
// The plane sums follow the reference's order in the FP64 form and in
// mhd3d_fluxes: the RO and TT means by the exact block-parallel walk
// (walk_plane_sum), the convective-flux means, sums of both signs near zero,
// by one chain per plane with its summands formed beside it
// (ordered_plane_sum); inside the mixed form's steps a fixed-order tree of
// partial sums (tree_plane_sum).
template <typename TI>
static int run_prepass_t(mhd3d_ctx* ctx, const TI* U, int stage, hipStream_t stream)
{
    const int f = ctx->xa.flags;
    const int want_ro = (f & (EXT_DH | EXT_DV)) ? 1 : 0;
    const int want_tt = ((f & (EXT_REM | EXT_RLAX)) && !(f & EXT_SHR)) ? 1 : 0;
    const bool tree = ctx->mixed && stage > 0;
    if (want_ro || want_tt) {
        const dim3 grd(ctx->g.nz, 2, 1);
        if (tree)
            plane_mean_tree_kernel<TI><<<grd, PT_T, 0, stream>>>(ctx->g, U, ctx->bg64, want_ro, want_tt,
                                                                 ctx->ext_buf);
        else
            plane_mean_walk_kernel<TI><<<grd, PW_T, 0, stream>>>(ctx->g, U, ctx->bg64, want_ro, want_tt,
                                                                 ctx->ext_buf);
        HIPCHK(ctx, hipGetLastError());
    }
    if (f & EXT_RLAX) {
        const int nz = ctx->g.nz;
        double* fconm = ctx->relax_buf;
        const int shr = (f & EXT_SHR) ? 1 : 0;
        if (tree)
            conv_flux_mean_tree_kernel<TI><<<ctx->g.nPz, PT_T, 0, stream>>>(ctx->g, U, ctx->bg64, ctx->ext_buf + nz,
                                                                            shr, ctx->ds, 1, fconm);
        else
            conv_flux_mean_kernel<TI><<<ctx->g.nPz, PM_T, 0, stream>>>(ctx->g, U, ctx->bg64, ctx->ext_buf + nz,
                                                                     shr, ctx->ds, stage ? 1 : 0, fconm);
        HIPCHK(ctx, hipGetLastError());
        const size_t wbytes = RELAX_WORK(nz, ctx->ra.P) * sizeof(double);
        const int use_lds = wbytes <= 65536 ? 1 : 0;
        relax_kernel<<<1, 256, use_lds ? wbytes : 0, stream>>>(
            ctx->ra, ctx->g, ctx->ext_buf + nz, fconm, ctx->ds, ctx->rs, ctx->relax_buf + nz,
            ctx->relax_buf + 2 * (size_t)nz + 1, stage, ctx->relax_buf + 4 * (size_t)nz + 1, use_lds);
        HIPCHK(ctx, hipGetLastError());
    }
    return MHD3D_OK;
}

// The prepass on the input of stage k's fluxes (k = 0: mhd3d_fluxes, on
// the state): the FP64 register, or in the mixed form's stages 2-3 the FP32
// copy of the stage value.
static int run_prepass(mhd3d_ctx* ctx, const StageArgs& a, int k, hipStream_t stream)
{
    if (ctx->mixed && k >= 2) return run_prepass_t<float>(ctx, a.Uin32, k, stream);
    return run_prepass_t<double>(ctx, a.Uin, k, stream);
}

static int run_stage(mhd3d_ctx* ctx, int k, bool fused_cfl = false)
{
    StageArgs a = stage_args(ctx, k);
    if (ctx->ext) {
        int rc = run_prepass(ctx, a, k, ctx->stream);
        if (rc) return rc;
    }
    if (fused_cfl && k == 3 && ctx->kernel_mode == 0) {
        int ns, nw;
        cfl_part_layout(ctx, ns, nw);
        int rc = ensure_cfl_parts(ctx, ns + nw);
        if (rc) return rc;
        a.cfl_parts = ctx->cfl_parts;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ctx->timing) {
        e0 = next_event(ctx);
        e1 = next_event(ctx);
        if (!e0 || !e1) return fail(ctx, MHD3D_EHIP, "hipEventCreate failed");
        HIPCHK(ctx, hipEventRecord(e0, ctx->stream));
    }
    if (k == 1) launch_stage_k<1>(ctx, a, false);
    else if (k == 2) launch_stage_k<2>(ctx, a, false);
    else launch_stage_k<3>(ctx, a, false);
    if (ctx->timing) HIPCHK(ctx, hipEventRecord(e1, ctx->stream));
    if (a.cfl_parts) {
        int ns, nw;
        cfl_part_layout(ctx, ns, nw);
        a.cfl_base = ns;
    }
    if (k == 1) launch_stage_k<1>(ctx, a, true);
    else if (k == 2) launch_stage_k<2>(ctx, a, true);
    else launch_stage_k<3>(ctx, a, true);
    HIPCHK(ctx, hipGetLastError());
    if (!ctx->mixed) ctx->cur = 1 - ctx->cur;
    return MHD3D_OK;
}

static int run_cfl(mhd3d_ctx* ctx)
{
    const Geom& g = ctx->g;
    dim3 blk(64, 4, 1);
    const int zchunks = g.nPz < 8 ? g.nPz : 8;
    dim3 grd((g.nPx + 63) / 64, (g.nPy + 3) / 4, zchunks);
    const int nparts = (int)(grd.x * grd.y * grd.z);
    int rc = ensure_cfl_parts(ctx, nparts);
    if (rc) return rc;
    if (ctx->d.flag_lMag)
        cfl_kernel<true><<<grd, blk, 0, ctx->stream>>>(g, ctx->c, ctx->m, ctx->U[ctx->cur], ctx->cfl_parts);
    else
        cfl_kernel<false><<<grd, blk, 0, ctx->stream>>>(g, ctx->c, ctx->m, ctx->U[ctx->cur], ctx->cfl_parts);
    cfl_finalize_kernel<<<1, 256, 0, ctx->stream>>>(ctx->cfl_parts, nparts, ctx->ds, ctx->c,
                                                   ctx->d.flag_lMag ? 1 : 0);
    HIPCHK(ctx, hipGetLastError());
    return MHD3D_OK;
}

static int collect_timing(mhd3d_ctx* ctx)
{
    if (!ctx->timing) return MHD3D_OK;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    for (size_t k = 0; k + 1 < ctx->ev_used; k += 2) {
        float ms = 0.f;
        HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev_pool[k], ctx->ev_pool[k + 1]));
        ctx->stage_ms += ms;
        ctx->stage_launches += 1;
    }
    ctx->ev_used = 0;
    return MHD3D_OK;
}

// The message of the reference's exit_err at a relaxation exit point.
static std::string relax_exit_text(const mhd3d_ctx* ctx)
{
    return ctx->relax_exit_code == 2
               ? "Relax: check boundary temperature (3dmhdsub.h:344,416: the reference exits)"
               : "yk not implement : sub 350 (3dmhdsub.h:427-439: the reference exits at TIMT "
                 "within DT after a multiple of 25)";
}

// MHD3D_EEXIT for the exit recorded in rs.
static int relax_exit_status(mhd3d_ctx* ctx, const RelaxState& rs)
{
    ctx->relax_exit_code = (int)rs.exited;
    return fail(ctx, MHD3D_EEXIT, relax_exit_text(ctx));
}

static int step_core(mhd3d_ctx* ctx, int nsteps, int fixed_dt, double* dt_inout)
{
    if (!ctx || nsteps < 0) return MHD3D_EINVAL;
    if (fixed_dt && !dt_inout) return fail(ctx, MHD3D_EINVAL, "fixed_dt needs *dt_inout");
    const int cur0 = ctx->cur;
    call_begin_kernel<<<1, 1, 0, ctx->stream>>>(ctx->ds);
    if (fixed_dt) {
        set_dt_kernel<<<1, 1, 0, ctx->stream>>>(ctx->ds, *dt_inout);
        HIPCHK(ctx, hipGetLastError());
    }
    // TSTEP of each step is evaluated on the state the previous step's
    // stage 3 produced, inside that launch (fused); a standalone pass runs
    // only when no fused value for the current state exists.  The fused
    // partials carry the non-finite flag of the state they were taken on,
    // so the last step's close tells whether the final state is finite.
    const bool fuse = !fixed_dt && ctx->kernel_mode == 0 && ctx->fuse_cfl;
    for (int n = 0; n < nsteps; ++n) {
        if (ctx->ext_timedep) {
            // the heat-loss / plume factors of this step need TIMT on the host
            HIPCHK(ctx, hipMemcpyAsync(ctx->hs, ctx->ds, sizeof(DevScalars), hipMemcpyDeviceToHost, ctx->stream));
            HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
            step_factors(ctx, ctx->d.TIMI + ctx->hs->timc);
        }
        if (!fixed_dt && !(fuse && ctx->dt_ready)) {
            int rc = run_cfl(ctx);
            if (rc) return rc;
        }
        for (int k = 1; k <= 3; ++k) {
            int rc = run_stage(ctx, k, fuse);
            if (rc) return rc;
        }
        if (fuse) {
            int ns, nw;
            cfl_part_layout(ctx, ns, nw);
            step_close_kernel<<<1, 256, 0, ctx->stream>>>(ctx->cfl_parts, ns + nw, ctx->ds, ctx->c,
                                                         ctx->d.flag_lMag ? 1 : 0);
        } else {
            clock_kernel<<<1, 1, 0, ctx->stream>>>(ctx->ds);
        }
        ctx->dt_ready = fuse;
        // bound the event pool: fold timings every 64 steps
        if (ctx->timing && ctx->ev_used >= 384) {
            int rc = collect_timing(ctx);
            if (rc) return rc;
        }
    }
    HIPCHK(ctx, hipGetLastError());
    if (!fuse || nsteps == 0) {
        // no fused partials of the final state: scan it
        const Geom& g = ctx->g;
        dim3 blk(64, 4, 1);
        dim3 grd((g.nPx + 63) / 64, (g.nPy + 3) / 4, g.nPz < 8 ? g.nPz : 8);
        nonfinite_kernel<<<grd, blk, 0, ctx->stream>>>(g, ctx->nf, ctx->U[ctx->cur], ctx->ds);
        HIPCHK(ctx, hipGetLastError());
    }
    HIPCHK(ctx, hipMemcpyAsync(ctx->hs, ctx->ds, sizeof(DevScalars), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    const DevScalars hs = *ctx->hs;
    // the register holding the state: each step taken flips it (three
    // stages); kernels of a stopped domain return at once, so only the steps
    // actually taken count
    const int64_t taken = (int64_t)(hs.nit - hs.nit_call0);
    if (!ctx->mixed) ctx->cur = cur0 ^ (int)(taken & 1);
    if (taken < nsteps) ctx->dt_ready = false;
    int rc = collect_timing(ctx);
    if (rc) return rc;
    if (dt_inout && nsteps > 0) *dt_inout = hs.dt_last;
    if ((ctx->xa.flags & EXT_RLAX) && (hs.halted & HALT_EXIT)) {
        RelaxState rs;
        HIPCHK(ctx, hipMemcpy(&rs, ctx->rs, sizeof(RelaxState), hipMemcpyDeviceToHost));
        // the state stays the input of the stage whose fluxes() exited:
        // stage 2 reads the register stage 1 wrote, stages 1 and 3 the
        // register the step started from (the one the taken steps name).
        // The mixed form has one FP64 register, which holds the step's
        // start: its stage values live only in the FP32 copies.
        if (!ctx->mixed && !ctx->relax_exit_seen && rs.exit_stage == 2) ctx->cur ^= 1;
        ctx->relax_exit_seen = true;
        ctx->dt_ready = false;
        return relax_exit_status(ctx, rs);
    }
    if (hs.halted & HALT_GUARD) return fail(ctx, MHD3D_EHALTED, "DT too small, halt (DT guard)");
    if (hs.nonfinite) return fail(ctx, MHD3D_ENONFINITE, "state is not finite after the step");
    return MHD3D_OK;
}

extern "C" int mhd3d_step(mhd3d_ctx* ctx, int nsteps, int fixed_dt, double* dt_inout)
{
    return step_core(ctx, nsteps, fixed_dt, dt_inout);
}

// ---------------------------------------------------------------------------
// Step without device allocation (mhd3d_step_noalloc): the TSTEP partial
// records of the standalone and fused passes live in `scratch`, a file-scope
// device array each GPU holds once for the whole process, instead of the
// context's records that mhd3d_step allocates on first use.  Host code
// never passes `scratch` itself (its host-side name is no device address):
// the device address comes from hipGetSymbolAddress, once per device, and
// is kept in g_scratch_dev.  Every partial record a kernel reads was written
// earlier in the same call, so nothing in it needs clearing; calls that
// overlap on one GPU (other streams, other threads) share the array and
// race.

#include <mutex>

constexpr int kScratchParts = 32768;   // 32768 records x 48 B = 1.5 MiB per GPU
static __device__ __align__(256) CflPart scratch[kScratchParts];

static std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;
static CflPart* g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up once).
static int scratch_records(mhd3d_ctx* ctx, CflPart** out)
{
    int dev = -1;
    HIPCHK(ctx, hipGetDevice(&dev));
    if (dev != ctx->device)
        return fail(ctx, MHD3D_EINVAL, "mhd3d_step_noalloc: the context's device is not the current device");
    if (dev < 0 || dev >= kMaxDevices) return fail(ctx, MHD3D_EINVAL, "mhd3d_step_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(ctx, MHD3D_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_dev[dev] = static_cast<CflPart*>(p);
    }
    *out = g_scratch_dev[dev];
    return MHD3D_OK;
}

extern "C" int mhd3d_step_noalloc(mhd3d_ctx* ctx, int nsteps, int fixed_dt, double* dt_inout)
{
    if (!ctx || nsteps < 0) return MHD3D_EINVAL;
    if (ctx->timing)
        return fail(ctx, MHD3D_EINVAL, "mhd3d_step_noalloc: kernel timing creates events; disable it first");
    // records the call needs: the fused pass (streaming blocks + wall
    // blocks) or the standalone TSTEP pass, whichever is larger
    int ns, nw;
    cfl_part_layout(ctx, ns, nw);
    const Geom& g = ctx->g;
    const int nsa = ((g.nPx + 63) / 64) * ((g.nPy + 3) / 4) * (g.nPz < 8 ? g.nPz : 8);
    const int need = std::max(ns + nw, nsa);
    if (need > kScratchParts)
        return fail(ctx, MHD3D_EINVAL,
                    "mhd3d_step_noalloc: this grid needs " + std::to_string(need) +
                        " TSTEP partial records; the no-alloc step holds at most " +
                        std::to_string(kScratchParts) + " (nPx*nPy up to about 1M points)");
    CflPart* rec = nullptr;
    int rc = scratch_records(ctx, &rec);
    if (rc) return rc;
    // the context's records are set aside for the call (no allocation:
    // the capacity covers every pass of the call)
    CflPart* own = ctx->cfl_parts;
    const int own_cap = ctx->cfl_parts_cap;
    ctx->cfl_parts = rec;
    ctx->cfl_parts_cap = kScratchParts;
    rc = step_core(ctx, nsteps, fixed_dt, dt_inout);
    ctx->cfl_parts = own;
    ctx->cfl_parts_cap = own_cap;
    return rc;
}

extern "C" int mhd3d_stage(mhd3d_ctx* ctx, int stage, double dt)
{
    if (!ctx || stage < 1 || stage > 3) return MHD3D_EINVAL;
    if (ctx->mixed)
        return fail(ctx, MHD3D_EUNSUPPORTED,
                    "mhd3d_stage: the mixed form keeps stage values in FP32 copies; use mhd3d_step");
    ctx->dt_ready = false;
    if (ctx->ext_timedep) {
        DevScalars hs;
        HIPCHK(ctx, hipMemcpyAsync(&hs, ctx->ds, sizeof(DevScalars), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        step_factors(ctx, ctx->d.TIMI + hs.timc);
    }
    set_dt_kernel<<<1, 1, 0, ctx->stream>>>(ctx->ds, dt);
    int rc = run_stage(ctx, stage);
    if (rc) return rc;
    HIPCHK(ctx, hipMemcpyAsync(ctx->hs, ctx->ds, sizeof(DevScalars), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->hs->halted) ctx->cur = 1 - ctx->cur;   // the kernels of a stopped domain did nothing
    if ((ctx->xa.flags & EXT_RLAX) && (ctx->hs->halted & HALT_EXIT)) {
        RelaxState rs;
        HIPCHK(ctx, hipMemcpy(&rs, ctx->rs, sizeof(RelaxState), hipMemcpyDeviceToHost));
        ctx->relax_exit_seen = true;
        return relax_exit_status(ctx, rs);
    }
    return collect_timing(ctx);
}

extern "C" int mhd3d_fluxes(mhd3d_ctx* ctx, double* const* flux)
{
    if (!ctx || !flux) return MHD3D_EINVAL;
    for (int f = 0; f < ctx->nf; ++f)
        if (!flux[f]) return fail(ctx, MHD3D_EINVAL, "null flux pointer");
    const Geom& g = ctx->g;
    const size_t reg = (size_t)g.fs * ctx->nf * sizeof(double);
    if (!ctx->F) {
        HIPCHK(ctx, hipMalloc(&ctx->F, reg));
        ctx->bytes += (int64_t)reg;
    }
    HIPCHK(ctx, hipMemsetAsync(ctx->F, 0, reg, ctx->stream));
    if (ctx->ext_timedep) {
        DevScalars hs;
        HIPCHK(ctx, hipMemcpyAsync(&hs, ctx->ds, sizeof(DevScalars), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        step_factors(ctx, ctx->d.TIMI + hs.timc);
    }
    StageArgs a = stage_args(ctx);
    dim3 blk(64, 4, 1);
    dim3 grd((g.nPx + 63) / 64, (g.nPy + 3) / 4, g.nPz);
    const bool mag = ctx->d.flag_lMag, rot = ctx->d.flag_lRot;
    if (ctx->ext) {
        int rc = run_prepass(ctx, a, 0, ctx->stream);
        if (rc) return rc;
        if (mag && rot) flux_only_kernel<true, true, true><<<grd, blk, 0, ctx->stream>>>(a, ctx->F);
        else if (mag) flux_only_kernel<true, false, true><<<grd, blk, 0, ctx->stream>>>(a, ctx->F);
        else if (rot) flux_only_kernel<false, true, true><<<grd, blk, 0, ctx->stream>>>(a, ctx->F);
        else flux_only_kernel<false, false, true><<<grd, blk, 0, ctx->stream>>>(a, ctx->F);
    } else if (mag && rot) flux_only_kernel<true, true><<<grd, blk, 0, ctx->stream>>>(a, ctx->F);
    else if (mag) flux_only_kernel<true, false><<<grd, blk, 0, ctx->stream>>>(a, ctx->F);
    else if (rot) flux_only_kernel<false, true><<<grd, blk, 0, ctx->stream>>>(a, ctx->F);
    else flux_only_kernel<false, false><<<grd, blk, 0, ctx->stream>>>(a, ctx->F);
    HIPCHK(ctx, hipGetLastError());
    {
        int rc = download_fields(ctx, flux, ctx->F, ctx->nf, false);
        if (rc) return rc;
    }
    if (ctx->xa.flags & EXT_RLAX) {
        // the reference's fluxes() ends in exit_err at an exit point: the
        // fluxes are complete, then the domain stops
        RelaxState rs;
        HIPCHK(ctx, hipMemcpyAsync(&rs, ctx->rs, sizeof(RelaxState), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        if (rs.exited) {
            relax_stop_kernel<<<1, 1, 0, ctx->stream>>>(ctx->ds, ctx->rs);
            HIPCHK(ctx, hipGetLastError());
            return relax_exit_status(ctx, rs);
        }
    }
    return MHD3D_OK;
}

extern "C" int mhd3d_bcon(mhd3d_ctx* ctx)
{
    if (!ctx) return MHD3D_EINVAL;
    ctx->dt_ready = false;
    if (ctx->ext_timedep) {
        DevScalars hs;
        HIPCHK(ctx, hipMemcpyAsync(&hs, ctx->ds, sizeof(DevScalars), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        step_factors(ctx, ctx->d.TIMI + hs.timc);
    }
    const Geom& g = ctx->g;
    dim3 blk(64, 4, 1);
    dim3 grd((g.nPx + 63) / 64, (g.nPy + 3) / 4, 2);
    const bool mag = ctx->d.flag_lMag;
    if (ctx->ext) {
        if (mag) bcon_kernel<true, true><<<grd, blk, 0, ctx->stream>>>(g, ctx->c, ctx->m, ctx->xa, ctx->ds, ctx->U[ctx->cur]);
        else bcon_kernel<false, true><<<grd, blk, 0, ctx->stream>>>(g, ctx->c, ctx->m, ctx->xa, ctx->ds, ctx->U[ctx->cur]);
    } else {
        if (mag) bcon_kernel<true, false><<<grd, blk, 0, ctx->stream>>>(g, ctx->c, ctx->m, ctx->xa, ctx->ds, ctx->U[ctx->cur]);
        else bcon_kernel<false, false><<<grd, blk, 0, ctx->stream>>>(g, ctx->c, ctx->m, ctx->xa, ctx->ds, ctx->U[ctx->cur]);
    }
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return MHD3D_OK;
}

extern "C" int mhd3d_cfl(mhd3d_ctx* ctx, mhd3d_cfl_info* out)
{
    if (!ctx || !out) return MHD3D_EINVAL;
    int rc = run_cfl(ctx);
    if (rc) return rc;
    DevScalars hs;
    HIPCHK(ctx, hipMemcpyAsync(&hs, ctx->ds, sizeof(DevScalars), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < 4; ++k) out->wmin[k] = hs.wmin[k];
    out->umach = hs.umach;
    out->dt = hs.dt;
    return MHD3D_OK;
}

// Autotuner of a single domain: times `steps` CFL steps of every streaming
// variant (register budget x schedule x launch size) on the current state,
// restoring the state, the clock and the TSTEP of the state after each trial,
// and keeps the fastest.  The results do not depend on the variant.
extern "C" int mhd3d_autotune(mhd3d_ctx* ctx, int steps, double* report, int* ncand_out)
{
    if (!ctx || steps < 1) return MHD3D_EINVAL;
    if (ctx->kernel_mode != 0)
        return fail(ctx, MHD3D_EUNSUPPORTED, "mhd3d_autotune: the point kernel has no variants");
    if (ctx->ext)
        return fail(ctx, MHD3D_EUNSUPPORTED,
                    "mhd3d_autotune: the off-default branches keep state between steps (lShr's FT, RLAX) "
                    "and run one variant");
    if (ctx->timing) return fail(ctx, MHD3D_EINVAL, "mhd3d_autotune: disable kernel timing first");
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    const size_t reg = (size_t)ctx->g.fs * ctx->nf * sizeof(double);
    double* snap = nullptr;
    DevScalars* ds_snap = nullptr;
    if (hipMalloc(&snap, reg) != hipSuccess || hipMalloc(&ds_snap, sizeof(DevScalars)) != hipSuccess) {
        (void)hipFree(snap);
        return fail(ctx, MHD3D_ENOMEM, "mhd3d_autotune: snapshot");
    }
    const int cur_save = ctx->cur;
    const bool ready_save = ctx->dt_ready;
    const int fm = ctx->mixed ? 1 : 0, ph = ctx->d.flag_lMag ? 1 : 0;
    int best_w = ctx->tune_waves[fm][ph], best_s = ctx->tune_sched[fm][ph], best_t = ctx->target_blocks;
    int rc = MHD3D_OK;
    if (hipMemcpyAsync(snap, ctx->U[ctx->cur], reg, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(ds_snap, ctx->ds, sizeof(DevScalars), hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess)
        rc = fail(ctx, MHD3D_EHIP, "mhd3d_autotune: snapshot copy");
    auto restore = [&]() -> int {
        ctx->cur = cur_save;
        HIPCHK(ctx, hipMemcpyAsync(ctx->U[ctx->cur], snap, reg, hipMemcpyDeviceToDevice, ctx->stream));
        HIPCHK(ctx, hipMemcpyAsync(ctx->ds, ds_snap, sizeof(DevScalars), hipMemcpyDeviceToDevice, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        ctx->dt_ready = ready_save;
        return MHD3D_OK;
    };
    hipEvent_t t0 = nullptr, t1 = nullptr;
    (void)hipEventCreate(&t0);
    (void)hipEventCreate(&t1);
    const int wv[2] = {1, 2};
    const int planes[3] = {0, 64, 32};
    double best = 1e300;
    int ncand = 0;
    for (int iw = 0; iw < 2 && !rc; ++iw)
        for (int is = 0; is < 4 && !rc; ++is)
            for (int ip = 0; ip < 3 && !rc; ++ip) {
                ctx->tune_waves[fm][ph] = wv[iw];
                ctx->tune_sched[fm][ph] = is;
                ctx->target_blocks = launch_blocks(ctx, variant_resident(ctx, wv[iw], is), planes[ip]);
                ctx->dt_ready = false;   // the partial records follow the launch size
                double dt = 0.0;
                int r = step_core(ctx, 1, 0, &dt);   // warm-up
                double per = -1.0;
                if (r == MHD3D_OK) {
                    (void)hipEventRecord(t0, ctx->stream);
                    r = step_core(ctx, steps, 0, &dt);
                    (void)hipEventRecord(t1, ctx->stream);
                    (void)hipEventSynchronize(t1);
                    float ms = 0.f;
                    (void)hipEventElapsedTime(&ms, t0, t1);
                    if (r == MHD3D_OK) per = ms / steps;
                }
                // a trial the DT guard stops or whose state turns non-finite
                // does not count; any other failure ends the search
                if (r != MHD3D_OK && r != MHD3D_EHALTED && r != MHD3D_ENONFINITE) rc = r;
                if (report) {
                    report[4 * ncand + 0] = wv[iw];
                    report[4 * ncand + 1] = is;
                    report[4 * ncand + 2] = planes[ip];
                    report[4 * ncand + 3] = per;
                }
                ++ncand;
                if (per > 0.0 && per < best) {
                    best = per;
                    best_w = wv[iw];
                    best_s = is;
                    best_t = ctx->target_blocks;
                }
                const int rr = restore();
                if (!rc) rc = rr;
            }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    ctx->tune_waves[fm][ph] = best_w;
    ctx->tune_sched[fm][ph] = best_s;
    ctx->target_blocks = best_t;
    ctx->dt_ready = false;
    (void)hipFree(snap);
    (void)hipFree(ds_snap);
    if (ncand_out) *ncand_out = ncand;
    return rc;
}

// The variant a context runs: waves per SIMD, schedule bits, blocks per launch.
extern "C" int mhd3d_get_variant(mhd3d_ctx* ctx, int* waves, int* sched, int* blocks)
{
    if (!ctx) return MHD3D_EINVAL;
    const int fm = ctx->mixed ? 1 : 0, ph = ctx->d.flag_lMag ? 1 : 0;
    if (waves) *waves = ctx->tune_waves[fm][ph];
    if (sched) *sched = ctx->tune_sched[fm][ph];
    if (blocks) *blocks = ctx->target_blocks;
    return MHD3D_OK;
}

extern "C" int mhd3d_relax_state(mhd3d_ctx* ctx, double out[3], int64_t* icall, int* exited)
{
    if (!ctx) return MHD3D_EINVAL;
    RelaxState rs{};
    rs.TFAC = 1.0;
    rs.TTM_MAX = -1e99;
    rs.TTM_MIN = 1e99;
    if (ctx->rs) {
        HIPCHK(ctx, hipMemcpyAsync(&rs, ctx->rs, sizeof(RelaxState), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    }
    if (out) {
        out[0] = rs.TFAC;
        out[1] = rs.TTM_MAX;
        out[2] = rs.TTM_MIN;
    }
    if (icall) *icall = (int64_t)rs.icall;
    if (exited) *exited = rs.exited ? 1 : 0;
    return MHD3D_OK;
}

extern "C" int mhd3d_clock(mhd3d_ctx* ctx, int64_t* nIt, double* timt, double* timc)
{
    if (!ctx) return MHD3D_EINVAL;
    DevScalars hs;
    HIPCHK(ctx, hipMemcpyAsync(&hs, ctx->ds, sizeof(DevScalars), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (nIt) *nIt = (int64_t)hs.nit;
    if (timc) *timc = hs.timc;
    if (timt) *timt = ctx->d.TIMI + hs.timc;    // TIMT = TIMI + TIMC (3dmhdsub.h:152)
    return MHD3D_OK;
}

extern "C" int mhd3d_diagnostics(mhd3d_ctx* ctx, double out[3])
{
    if (!ctx || !out) return MHD3D_EINVAL;
    const Geom& g = ctx->g;
    const int nblk = 1024;
    double* part = nullptr;
    HIPCHK(ctx, hipMalloc(&part, sizeof(double) * 3 * nblk));
    diag_kernel<<<nblk, 256, 0, ctx->stream>>>(g, ctx->c, ctx->m, ctx->nf, ctx->U[ctx->cur], part);
    diag_finalize_kernel<<<1, 64, 0, ctx->stream>>>(part, nblk, part);
    double h[3];
    hipError_t e = hipMemcpyAsync(h, part, sizeof(h), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(part);
    if (e != hipSuccess) return fail(ctx, MHD3D_EHIP, hipGetErrorString(e));
    out[0] = h[0];
    out[1] = h[1];
    out[2] = h[2];
    return MHD3D_OK;
}

extern "C" int mhd3d_timing_enable(mhd3d_ctx* ctx, int on)
{
    if (!ctx) return MHD3D_EINVAL;
    ctx->timing = on != 0;
    ctx->stage_ms = 0.0;
    ctx->stage_launches = 0;
    ctx->step_ms = 0.0;
    ctx->steps_timed = 0;
    ctx->ev_used = 0;
    return MHD3D_OK;
}

extern "C" int mhd3d_timing_read(mhd3d_ctx* ctx, mhd3d_timing* out)
{
    if (!ctx || !out) return MHD3D_EINVAL;
    out->stage_ms_total = ctx->stage_ms;
    out->stage_launches = ctx->stage_launches;
    out->step_ms_total = ctx->step_ms;
    out->steps = ctx->steps_timed;
    return MHD3D_OK;
}

extern "C" int mhd3d_synchronize(mhd3d_ctx* ctx)
{
    if (!ctx) return MHD3D_EINVAL;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return MHD3D_OK;
}

extern "C" int mhd3d_set_dt_guard(mhd3d_ctx* ctx, double dt_min)
{
    if (!ctx || !(dt_min >= 0.0)) return MHD3D_EINVAL;
    // a halted domain's DT is that of the step that tripped the guard, not
    // the TSTEP of its state: resuming takes a fresh TSTEP
    ctx->dt_ready = false;
    set_guard_kernel<<<1, 1, 0, ctx->stream>>>(ctx->ds, dt_min);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return MHD3D_OK;
}

// ===========================================================================
// Ensembles: independent domains advanced together (kernels_ensemble.hpp).

struct mhd3d_ensemble {
    std::vector<mhd3d_ctx*> mem;       // members (their scalars: mem[m]->ds)
    bool owns = true;                  // destroy the members with the ensemble
    int n = 0;
    bool mixed = false;
    hipStream_t stream = nullptr;
    DevScalars* mirror = nullptr;      // contiguous copy of the members' scalars
    int mirror_cap = 0;
    StageArgs* d_args = nullptr;       // [parity at step start][stage][member]
    WorkItem* d_items = nullptr;
    WorkItem* d_items_x = nullptr;     // interior items of members with off-default branches
    WallItem* d_walls = nullptr;
    WallItem* d_walls_x = nullptr;     // wall items of members with off-default branches
    FinishArgs* d_fin = nullptr;
    int nitems = 0, nwalls = 0, nitems_x = 0, nwalls_x = 0;
    int grid_int = 0, grid_wall = 0, grid_x = 0, grid_wall_x = 0;
    bool prepass = false;              // some member needs plane means / the relaxation pass
    int waves = 1, sched = 0, chunk = 0;   // kernel variant; chunk: planes per item (0 = auto)
    int parity = 0;                    // FP64: register parity of the step tables
    std::vector<int> off;              // member m's state is U[parity ^ off[m]] (FP64)
    size_t cap_items = 0, cap_walls = 0, cap_fin = 0, cap_args = 0, cap_items_x = 0, cap_walls_x = 0;   // device table capacities
    bool use_graph = false;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    int64_t cells = 0;
    std::vector<DevScalars> last;      // members' scalars after the last step call
    std::string err;
};

static std::string g_ens_create_error;

static int ens_fail(mhd3d_ensemble* e, int code, const std::string& msg)
{
    if (e) e->err = msg;
    return code;
}

#define ENSCHK(e, call)                                                                \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            (e)->err = std::string(#call) + ": " + hipGetErrorString(e_);             \
            return MHD3D_EHIP;                                                         \
        }                                                                              \
    } while (0)

template <typename T, int STAGE, int WAVES>
static void ens_stage_launch(const mhd3d_ensemble* e, const StageArgs* args, hipStream_t s, int sched)
{
    dim3 blk(STX, STY, 1);
    switch (sched & 3) {
    case 0: ensemble_stage_kernel<T, STAGE, WAVES, false, false><<<e->grid_int, blk, 0, s>>>(args, e->d_items, e->nitems); break;
    case 1: ensemble_stage_kernel<T, STAGE, WAVES, true, false><<<e->grid_int, blk, 0, s>>>(args, e->d_items, e->nitems); break;
    case 2: ensemble_stage_kernel<T, STAGE, WAVES, false, true><<<e->grid_int, blk, 0, s>>>(args, e->d_items, e->nitems); break;
    default: ensemble_stage_kernel<T, STAGE, WAVES, true, true><<<e->grid_int, blk, 0, s>>>(args, e->d_items, e->nitems); break;
    }
}

template <typename T, int STAGE, int WAVES>
static const void* ens_stage_fn(int sched)
{
    switch (sched & 3) {
    case 0: return (const void*)ensemble_stage_kernel<T, STAGE, WAVES, false, false>;
    case 1: return (const void*)ensemble_stage_kernel<T, STAGE, WAVES, true, false>;
    case 2: return (const void*)ensemble_stage_kernel<T, STAGE, WAVES, false, true>;
    default: return (const void*)ensemble_stage_kernel<T, STAGE, WAVES, true, true>;
    }
}

template <typename T>
static void ens_stage(const mhd3d_ensemble* e, int k, const StageArgs* args)
{
    if (e->nitems == 0) {
    } else if (e->waves == 1) {
        if (k == 1) ens_stage_launch<T, 1, 1>(e, args, e->stream, e->sched);
        else if (k == 2) ens_stage_launch<T, 2, 1>(e, args, e->stream, e->sched);
        else ens_stage_launch<T, 3, 1>(e, args, e->stream, e->sched);
    } else {
        if (k == 1) ens_stage_launch<T, 1, 2>(e, args, e->stream, e->sched);
        else if (k == 2) ens_stage_launch<T, 2, 2>(e, args, e->stream, e->sched);
        else ens_stage_launch<T, 3, 2>(e, args, e->stream, e->sched);
    }
    if (e->nitems_x > 0) {
        dim3 blk(STX, STY, 1);
        if (k == 1) ensemble_stage_ext_kernel<1><<<e->grid_x, blk, 0, e->stream>>>(args, e->d_items_x, e->nitems_x);
        else if (k == 2) ensemble_stage_ext_kernel<2><<<e->grid_x, blk, 0, e->stream>>>(args, e->d_items_x, e->nitems_x);
        else ensemble_stage_ext_kernel<3><<<e->grid_x, blk, 0, e->stream>>>(args, e->d_items_x, e->nitems_x);
    }
    dim3 wblk(64, 4, 1);
    if (e->nwalls > 0) {
        if (k == 1) ensemble_wall_kernel<T, 1><<<e->grid_wall, wblk, 0, e->stream>>>(args, e->d_walls, e->nwalls);
        else if (k == 2) ensemble_wall_kernel<T, 2><<<e->grid_wall, wblk, 0, e->stream>>>(args, e->d_walls, e->nwalls);
        else ensemble_wall_kernel<T, 3><<<e->grid_wall, wblk, 0, e->stream>>>(args, e->d_walls, e->nwalls);
    }
    if (e->nwalls_x > 0) {
        const WallItem* wx = e->d_walls_x;
        if (k == 1) ensemble_wall_kernel<double, 1, true><<<e->grid_wall_x, wblk, 0, e->stream>>>(args, wx, e->nwalls_x);
        else if (k == 2) ensemble_wall_kernel<double, 2, true><<<e->grid_wall_x, wblk, 0, e->stream>>>(args, wx, e->nwalls_x);
        else ensemble_wall_kernel<double, 3, true><<<e->grid_wall_x, wblk, 0, e->stream>>>(args, wx, e->nwalls_x);
    }
}

// Resident blocks of an interior kernel on the whole device: the chosen
// variant, or (ext) the kernel of the members with off-default branches.
static int ens_resident_blocks(const mhd3d_ensemble* e, bool ext = false)
{
    const void* fn;
    if (ext)
        fn = e->mixed ? ext_ens_fn_s32() : (const void*)ensemble_stage_ext_kernel<2>;
    else if (e->mixed)
        fn = mixed_ens_stage_fn(e->waves, e->sched);
    else
        fn = e->waves == 1 ? ens_stage_fn<double, 2, 1>(e->sched) : ens_stage_fn<double, 2, 2>(e->sched);
    int per_cu = 0, dev = 0, ncu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, STX * STY, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
        ncu = 256;
    return per_cu * ncu;
}

static void ens_drop_graphs(mhd3d_ensemble* e)
{
    for (int p = 0; p < 2; ++p)
        if (e->gexec[p]) {
            (void)hipGraphExecDestroy(e->gexec[p]);
            e->gexec[p] = nullptr;
        }
}

// Device room for n records of `size` bytes in *p (capacity cap, in records).
static int ens_reserve(mhd3d_ensemble* e, void** p, size_t& cap, size_t size, size_t n)
{
    if (n <= cap && *p) return MHD3D_OK;
    const size_t want = std::max(n, cap + cap / 2);
    (void)hipFree(*p);
    *p = nullptr;
    cap = 0;
    ENSCHK(e, hipMalloc(p, size * (want ? want : 1)));
    cap = want;
    return MHD3D_OK;
}

// Operand tables [p0][k-1][m]: during stage k of a step that starts with
// table parity p0, member m reads U[p0 ^ off[m] ^ ((k-1) & 1)] (FP64).
static int ens_upload_args(mhd3d_ensemble* e)
{
    std::vector<StageArgs> args((size_t)2 * 3 * e->n);
    for (int p0 = 0; p0 < 2; ++p0)
        for (int k = 1; k <= 3; ++k)
            for (int m = 0; m < e->n; ++m) {
                mhd3d_ctx* c = e->mem[m];
                const int save = c->cur;
                c->cur = e->mixed ? 0 : (p0 ^ e->off[m] ^ ((k - 1) & 1));
                StageArgs a = stage_args(c, k);
                c->cur = save;
                a.dt = &c->ds->dt;
                a.ds = c->ds;
                a.cfl_parts = k == 3 ? c->cfl_parts : nullptr;
                a.cfl_base = 0;
                args[((size_t)p0 * 3 + (k - 1)) * e->n + m] = a;
            }
    ENSCHK(e, hipMemcpy(e->d_args, args.data(), sizeof(StageArgs) * args.size(), hipMemcpyHostToDevice));
    return MHD3D_OK;
}

// Work lists, operand tables and grid sizes for the current variant.
static int ens_build(mhd3d_ensemble* e)
{
    ens_drop_graphs(e);
    if (e->n > e->mirror_cap) {
        (void)hipFree(e->mirror);
        e->mirror = nullptr;
        e->mirror_cap = 0;
        ENSCHK(e, hipMalloc(&e->mirror, sizeof(DevScalars) * e->n));
        e->mirror_cap = e->n;
    }
    std::vector<WorkItem> items, items_x;
    std::vector<WallItem> walls, walls_x;
    std::vector<FinishArgs> fin(e->n);
    e->prepass = false;
    for (int m = 0; m < e->n; ++m) {
        mhd3d_ctx* c = e->mem[m];
        const Geom& g = c->g;
        const int flags = (c->d.flag_lMag ? 1 : 0) | (c->d.flag_lRot ? 2 : 0);
        // (a mixed-form member with any branch evaluates every flux in FP64,
        // as alone)
        const bool ext_interior = e->mixed ? c->ext : ext_variant(c->xa.flags) != EXTV_NONE;
        if (needs_prepass(c)) e->prepass = true;
        StreamPlan p = stream_plan(g, 256);
        if (e->chunk > 0) {
            const int nint = g.nPz - 2;
            p.nchunks = nint / e->chunk > 0 ? nint / e->chunk : 1;
            p.zchunk = nint / p.nchunks;
        }
        int part = 0;
        for (int ch = 0; ch < p.nchunks; ++ch)
            for (int ty = 0; ty < p.nty; ++ty)
                for (int tx = 0; tx < p.ntx; ++tx) {
                    WorkItem w{};
                    w.member = m; w.flags = flags; w.tile_x = tx; w.tile_y = ty;
                    w.zb = 2 + ch * p.zchunk;
                    w.ze = ch == p.nchunks - 1 ? g.nPz : w.zb + p.zchunk;
                    w.part = part++;
                    (ext_interior ? items_x : items).push_back(w);
                }
        for (int top = 1; top >= 0; --top)
            for (int y0 = 0; y0 < g.nPy; y0 += 4)
                for (int x0 = 0; x0 < g.nPx; x0 += 64) {
                    WallItem w{};
                    w.member = m; w.flags = flags; w.x0 = x0; w.y0 = y0; w.top = top;
                    w.part = part++;
                    (c->ext ? walls_x : walls).push_back(w);
                }
        // room for the fused partials and for the standalone TSTEP pass
        const int nsa = ((g.nPx + 63) / 64) * ((g.nPy + 3) / 4) * (g.nPz < 8 ? g.nPz : 8);
        int rc = ensure_cfl_parts(c, part > nsa ? part : nsa);
        if (rc) return ens_fail(e, rc, c->err);
        fin[m].parts = c->cfl_parts;
        fin[m].nparts = part;
        fin[m].mag = c->d.flag_lMag ? 1 : 0;
        fin[m].ds = c->ds;
        fin[m].mirror = e->mirror + m;
        fin[m].c = c->c;
    }
    // largest items first (a z chunk of an MHD member costs about twice a hydro one)
    auto by_cost = [](const WorkItem& a, const WorkItem& b) {
        const int ca = (a.ze - a.zb) * ((a.flags & 1) ? 2 : 1);
        const int cb = (b.ze - b.zb) * ((b.flags & 1) ? 2 : 1);
        return ca > cb;
    };
    std::stable_sort(items.begin(), items.end(), by_cost);
    std::stable_sort(items_x.begin(), items_x.end(), by_cost);
    // device tables grow only: a server's ensemble changes membership often
    // and reuses the room it has
    int rc = ens_reserve(e, (void**)&e->d_items, e->cap_items, sizeof(WorkItem), items.size());
    if (!rc) rc = ens_reserve(e, (void**)&e->d_items_x, e->cap_items_x, sizeof(WorkItem), items_x.size());
    if (!rc) rc = ens_reserve(e, (void**)&e->d_walls, e->cap_walls, sizeof(WallItem), walls.size());
    if (!rc) rc = ens_reserve(e, (void**)&e->d_walls_x, e->cap_walls_x, sizeof(WallItem), walls_x.size());
    if (!rc) rc = ens_reserve(e, (void**)&e->d_fin, e->cap_fin, sizeof(FinishArgs), fin.size());
    if (!rc) rc = ens_reserve(e, (void**)&e->d_args, e->cap_args, sizeof(StageArgs), (size_t)6 * e->n);
    if (rc) return rc;
    if (!items.empty())
        ENSCHK(e, hipMemcpy(e->d_items, items.data(), sizeof(WorkItem) * items.size(), hipMemcpyHostToDevice));
    if (!items_x.empty())
        ENSCHK(e, hipMemcpy(e->d_items_x, items_x.data(), sizeof(WorkItem) * items_x.size(), hipMemcpyHostToDevice));
    if (!walls.empty())
        ENSCHK(e, hipMemcpy(e->d_walls, walls.data(), sizeof(WallItem) * walls.size(), hipMemcpyHostToDevice));
    if (!walls_x.empty())
        ENSCHK(e, hipMemcpy(e->d_walls_x, walls_x.data(), sizeof(WallItem) * walls_x.size(), hipMemcpyHostToDevice));
    ENSCHK(e, hipMemcpy(e->d_fin, fin.data(), sizeof(FinishArgs) * fin.size(), hipMemcpyHostToDevice));
    e->off.assign(e->n, 0);
    for (int m = 0; m < e->n; ++m) e->off[m] = e->mixed ? 0 : (e->parity ^ e->mem[m]->cur);
    rc = ens_upload_args(e);
    if (rc) return rc;
    e->nitems = (int)items.size();
    e->nitems_x = (int)items_x.size();
    e->nwalls = (int)walls.size();
    e->nwalls_x = (int)walls_x.size();
    e->grid_wall_x = e->nwalls_x < 2048 ? e->nwalls_x : 2048;
    const int res = ens_resident_blocks(e);
    e->grid_int = e->nitems < res ? e->nitems : res;
    if (e->nitems_x > 0) {
        const int rx = ens_resident_blocks(e, true);
        e->grid_x = e->nitems_x < rx ? e->nitems_x : rx;
    }
    e->grid_wall = e->nwalls < 2048 ? e->nwalls : 2048;
    return MHD3D_OK;
}

// The ensemble kernel variant an ensemble or a server starts with, the
// fastest measured on MI355X for the 40-run served trace (the autotuner may
// pick another): one wave per SIMD with the deferred update, two barriers
// per plane in FP64 (6296 Mcell-steps/s; two waves 6186-6225, one barrier
// 5014) and one in the mixed form (6488; two waves and two barriers 6076).
// The override MHD3D_TUNE_ENS_<FP64|MIXED>=waves,sched[,chunk] replaces it.
static void ens_default_variant(mhd3d_ensemble* e)
{
    e->waves = 1;
    e->sched = e->mixed ? 3 : 2;
    const char* tv = std::getenv(e->mixed ? "MHD3D_TUNE_ENS_MIXED" : "MHD3D_TUNE_ENS_FP64");
    int w = 0, sc = 0, ch = 0;
    if (tv && std::sscanf(tv, "%d,%d,%d", &w, &sc, &ch) >= 2 && (w == 1 || w == 2) && sc >= 0 && sc <= 3) {
        e->waves = w;
        e->sched = sc;
        e->chunk = ch > 0 ? ch : 0;
    }
}

extern "C" int mhd3d_ensemble_create(const mhd3d_desc* const* descs, int n, mhd3d_ensemble** out)
{
    if (!out || !descs || n < 1) return MHD3D_EINVAL;
    *out = nullptr;
    mhd3d_ensemble* e = new mhd3d_ensemble();
    auto bail = [&](int code, const std::string& m) {
        g_ens_create_error = m;
        mhd3d_ensemble_destroy(e);
        return code;
    };
    e->n = n;
    e->mixed = descs[0] && descs[0]->precision == MHD3D_MIXED;
    for (int m = 0; m < n; ++m) {
        if (!descs[m]) return bail(MHD3D_EINVAL, "null descriptor");
        if ((descs[m]->precision == MHD3D_MIXED) != e->mixed)
            return bail(MHD3D_EINVAL, "ensemble members must share one precision form");
        if (timedep_flags(descs[m]))
            return bail(MHD3D_EUNSUPPORTED, "member " + std::to_string(m) + ": " + kTimedepWhy);
        mhd3d_ctx* c = nullptr;
        int rc = mhd3d_create(descs[m], &c);
        if (rc) return bail(rc, "member " + std::to_string(m) + ": " + g_create_error);
        c->kernel_mode = 0;
        e->mem.push_back(c);
        e->cells += (int64_t)descs[m]->nPx * descs[m]->nPy * descs[m]->nPz;
    }
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(MHD3D_EHIP, "hipStreamCreate failed");
    ens_default_variant(e);
    int rc = ens_build(e);
    if (rc) return bail(rc, e->err);
    *out = e;
    return MHD3D_OK;
}

extern "C" int mhd3d_ensemble_destroy(mhd3d_ensemble* e)
{
    if (!e) return MHD3D_OK;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    ens_drop_graphs(e);
    if (e->owns)
        for (auto* c : e->mem) mhd3d_destroy(c);
    (void)hipFree(e->mirror);
    (void)hipFree(e->d_items);
    (void)hipFree(e->d_items_x);
    (void)hipFree(e->d_walls);
    (void)hipFree(e->d_walls_x);
    (void)hipFree(e->d_fin);
    (void)hipFree(e->d_args);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return MHD3D_OK;
}

extern "C" const char* mhd3d_ensemble_last_error(const mhd3d_ensemble* e)
{
    return e ? e->err.c_str() : g_ens_create_error.c_str();
}

extern "C" int mhd3d_ensemble_size(const mhd3d_ensemble* e) { return e ? e->n : 0; }

extern "C" int mhd3d_ensemble_set_state(mhd3d_ensemble* e, int member, const double* const* fields)
{
    if (!e || member < 0 || member >= e->n) return MHD3D_EINVAL;
    ENSCHK(e, hipStreamSynchronize(e->stream));
    mhd3d_ctx* c = e->mem[member];
    int rc = mhd3d_set_state(c, fields);
    if (rc) return ens_fail(e, rc, c->err);
    // a fresh state needs its TSTEP (set_state cleared the member's
    // dt_ready); reset the member's clock, keep its DT guard
    DevScalars was{};
    ENSCHK(e, hipMemcpy(&was, c->ds, sizeof(was), hipMemcpyDeviceToHost));
    DevScalars z{};
    z.dt_min = was.dt_min;
    ENSCHK(e, hipMemcpy(c->ds, &z, sizeof(z), hipMemcpyHostToDevice));
    return MHD3D_OK;
}

extern "C" int mhd3d_ensemble_get_state(mhd3d_ensemble* e, int member, double* const* fields)
{
    if (!e || member < 0 || member >= e->n) return MHD3D_EINVAL;
    ENSCHK(e, hipStreamSynchronize(e->stream));
    mhd3d_ctx* c = e->mem[member];
    int rc = mhd3d_get_state(c, fields);
    if (rc) return ens_fail(e, rc, c->err);
    return MHD3D_OK;
}

// TSTEP of every member on its current state (the first step of a call
// after set_state; later steps take it from the fused stage-3 partials).
static int ens_cfl_all(mhd3d_ensemble* e)
{
    for (int m = 0; m < e->n; ++m) {
        mhd3d_ctx* c = e->mem[m];
        if (c->dt_ready) continue;
        const Geom& g = c->g;
        dim3 blk(64, 4, 1);
        const int zchunks = g.nPz < 8 ? g.nPz : 8;
        dim3 grd((g.nPx + 63) / 64, (g.nPy + 3) / 4, zchunks);
        const int nparts = (int)(grd.x * grd.y * grd.z);   // within the capacity ens_build ensured
        const double* U = c->U[c->cur];
        if (c->d.flag_lMag)
            cfl_kernel<true><<<grd, blk, 0, e->stream>>>(g, c->c, c->m, U, c->cfl_parts);
        else
            cfl_kernel<false><<<grd, blk, 0, e->stream>>>(g, c->c, c->m, U, c->cfl_parts);
        cfl_finalize_kernel<<<1, 256, 0, e->stream>>>(c->cfl_parts, nparts, c->ds, c->c,
                                                      c->d.flag_lMag ? 1 : 0);
    }
    ENSCHK(e, hipGetLastError());
    return MHD3D_OK;
}

static void ens_enqueue_step(mhd3d_ensemble* e, int p0)
{
    for (int k = 1; k <= 3; ++k) {
        const StageArgs* args = e->d_args + ((size_t)p0 * 3 + (k - 1)) * e->n;
        if (e->prepass) {
            // plane means and the relaxation pass of the members whose
            // fluxes() need them, on the register the stage reads (a
            // failed launch shows in the caller's hipGetLastError)
            // (the mixed form: stage 1 reads the FP64 state, stages 2-3 the
            // FP32 copies of the stage values)
            for (int m = 0; m < e->n; ++m) {
                mhd3d_ctx* c = e->mem[m];
                if (!needs_prepass(c)) continue;
                if (e->mixed && k >= 2) (void)run_prepass_t<float>(c, c->U32[k - 2], k, e->stream);
                else if (e->mixed) (void)run_prepass_t<double>(c, c->U[0], k, e->stream);
                else (void)run_prepass_t<double>(c, c->U[p0 ^ e->off[m] ^ ((k - 1) & 1)], k, e->stream);
            }
        }
        if (e->mixed) {
            mixed_ens_stage(k, e->waves, e->sched, args, e->d_items, e->nitems, e->grid_int, e->d_walls,
                            e->nwalls, e->grid_wall, e->stream);
            ext_ens_stage_s32(k, args, e->d_items_x, e->nitems_x, e->grid_x, e->d_walls_x, e->nwalls_x,
                              e->grid_wall_x, e->stream);
        } else {
            ens_stage<double>(e, k, args);
        }
    }
    ensemble_finish_kernel<<<e->n, 256, 0, e->stream>>>(e->d_fin);
}

// Enqueues nsteps steps of every member (no host synchronisation).
static int ens_step_enqueue(mhd3d_ensemble* e, int nsteps)
{
    {
        int rc = ens_cfl_all(e);   // members without a TSTEP of their current state
        if (rc) return rc;
    }
    ensemble_clear_kernel<<<(e->n + 255) / 256, 256, 0, e->stream>>>(e->d_fin, e->n);
    for (int s = 0; s < nsteps; ++s) {
        const int p0 = e->mixed ? 0 : e->parity;
        if (e->use_graph) {
            if (!e->gexec[p0]) {
                hipGraph_t gr = nullptr;
                ENSCHK(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
                ens_enqueue_step(e, p0);
                ENSCHK(e, hipStreamEndCapture(e->stream, &gr));
                hipError_t ie = hipGraphInstantiate(&e->gexec[p0], gr, nullptr, nullptr, 0);
                (void)hipGraphDestroy(gr);
                if (ie != hipSuccess) return ens_fail(e, MHD3D_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
            }
            ENSCHK(e, hipGraphLaunch(e->gexec[p0], e->stream));
        } else {
            ens_enqueue_step(e, p0);
        }
        if (!e->mixed) e->parity ^= 1;
    }
    ENSCHK(e, hipGetLastError());
    e->last.resize(e->n);
    ENSCHK(e, hipMemcpyAsync(e->last.data(), e->mirror, sizeof(DevScalars) * e->n, hipMemcpyDeviceToHost, e->stream));
    return MHD3D_OK;
}

// Waits for the enqueued steps and settles each member's register: a
// member the DT guard or its step budget stopped took fewer steps than the
// others, and its state sits in the register of the steps it took.
static int ens_step_finish(mhd3d_ensemble* e, int nsteps)
{
    ENSCHK(e, hipStreamSynchronize(e->stream));
    bool refresh = false;
    for (int m = 0; m < e->n; ++m) {
        mhd3d_ctx* c = e->mem[m];
        const int64_t taken = (int64_t)(e->last[m].nit - e->last[m].nit_call0);
        c->dt_ready = taken == nsteps;
        if (!e->mixed) {
            const int cur0 = e->parity ^ (nsteps & 1) ^ e->off[m];   // the register at the call's start
            c->cur = cur0 ^ (int)(taken & 1);
        }
        if ((c->xa.flags & EXT_RLAX) && (e->last[m].halted & HALT_EXIT)) {
            // stopped at the reference's exit point: the state is the input
            // of the stage that exited (stage 2 reads what stage 1 wrote;
            // the mixed form keeps the step's start in its one FP64 register)
            RelaxState rs;
            ENSCHK(e, hipMemcpy(&rs, c->rs, sizeof(RelaxState), hipMemcpyDeviceToHost));
            // (off[] below keeps the flip for the member's later calls)
            if (!e->mixed && !c->relax_exit_seen && rs.exit_stage == 2) c->cur ^= 1;
            c->relax_exit_seen = true;
            c->relax_exit_code = (int)rs.exited;
        }
        if (e->mixed) continue;
        const int off = e->parity ^ c->cur;
        if (off != e->off[m]) {
            e->off[m] = off;
            refresh = true;
        }
    }
    if (refresh) {
        // captured prepass launches name the registers they read
        if (e->prepass) ens_drop_graphs(e);
        int rc = ens_upload_args(e);
        if (rc) return rc;
    }
    return MHD3D_OK;
}

extern "C" int mhd3d_ensemble_step(mhd3d_ensemble* e, int nsteps)
{
    if (!e || nsteps < 0) return MHD3D_EINVAL;
    int rc = ens_step_enqueue(e, nsteps);
    if (!rc) rc = ens_step_finish(e, nsteps);
    if (rc) return rc;
    for (int m = 0; m < e->n; ++m)
        if (e->last[m].halted & HALT_EXIT)
            return ens_fail(e, MHD3D_EEXIT, "member " + std::to_string(m) + ": " + relax_exit_text(e->mem[m]));
    for (int m = 0; m < e->n; ++m)
        if (e->last[m].nonfinite)
            return ens_fail(e, MHD3D_ENONFINITE, "member " + std::to_string(m) + ": state is not finite after the step");
    return MHD3D_OK;
}

extern "C" int mhd3d_ensemble_clock(mhd3d_ensemble* e, int member, int64_t* nIt, double* timt,
                                    double* timc, double* dt_last)
{
    if (!e || member < 0 || member >= e->n) return MHD3D_EINVAL;
    DevScalars hs;
    ENSCHK(e, hipMemcpyAsync(&hs, e->mem[member]->ds, sizeof(hs), hipMemcpyDeviceToHost, e->stream));
    ENSCHK(e, hipStreamSynchronize(e->stream));
    if (nIt) *nIt = (int64_t)hs.nit;
    if (timc) *timc = hs.timc;
    if (timt) *timt = e->mem[member]->d.TIMI + hs.timc;
    if (dt_last) *dt_last = hs.dt_last;
    return MHD3D_OK;
}

extern "C" int mhd3d_ensemble_graph(mhd3d_ensemble* e, int on)
{
    if (!e) return MHD3D_EINVAL;
    e->use_graph = on != 0;
    if (!e->use_graph) ens_drop_graphs(e);
    return MHD3D_OK;
}

extern "C" int mhd3d_ensemble_set_variant(mhd3d_ensemble* e, int waves, int sched, int chunk)
{
    if (!e || (waves != 1 && waves != 2) || sched < 0 || sched > 3 || chunk < 0) return MHD3D_EINVAL;
    ENSCHK(e, hipStreamSynchronize(e->stream));
    e->waves = waves;
    e->sched = sched;
    e->chunk = chunk;
    return ens_build(e);
}

extern "C" int mhd3d_ensemble_get_variant(const mhd3d_ensemble* e, int* waves, int* sched, int* chunk)
{
    if (!e) return MHD3D_EINVAL;
    if (waves) *waves = e->waves;
    if (sched) *sched = e->sched;
    if (chunk) *chunk = e->chunk;
    return MHD3D_OK;
}

extern "C" int64_t mhd3d_ensemble_cells(const mhd3d_ensemble* e) { return e ? e->cells : 0; }

extern "C" int mhd3d_ensemble_set_dt_guard(mhd3d_ensemble* e, double dt_min)
{
    if (!e || !(dt_min >= 0.0)) return MHD3D_EINVAL;
    ENSCHK(e, hipStreamSynchronize(e->stream));
    for (auto* c : e->mem) {
        int rc = mhd3d_set_dt_guard(c, dt_min);
        if (rc) return ens_fail(e, rc, c->err);
    }
    return MHD3D_OK;
}

extern "C" int mhd3d_ensemble_set_dt_guard_member(mhd3d_ensemble* e, int member, double dt_min)
{
    if (!e || member < 0 || member >= e->n || !(dt_min >= 0.0)) return MHD3D_EINVAL;
    ENSCHK(e, hipStreamSynchronize(e->stream));
    mhd3d_ctx* c = e->mem[member];
    int rc = mhd3d_set_dt_guard(c, dt_min);
    if (rc) return ens_fail(e, rc, c->err);
    return MHD3D_OK;
}

// Whether the DT guard stopped member `member` (after its last step call).
extern "C" int mhd3d_ensemble_halted(mhd3d_ensemble* e, int member, int* halted)
{
    if (!e || !halted || member < 0 || member >= e->n) return MHD3D_EINVAL;
    DevScalars hs;
    ENSCHK(e, hipMemcpyAsync(&hs, e->mem[member]->ds, sizeof(hs), hipMemcpyDeviceToHost, e->stream));
    ENSCHK(e, hipStreamSynchronize(e->stream));
    *halted = (hs.halted & HALT_GUARD) ? 1 : 0;
    return MHD3D_OK;
}

extern "C" int mhd3d_ensemble_synchronize(mhd3d_ensemble* e)
{
    if (!e) return MHD3D_EINVAL;
    ENSCHK(e, hipStreamSynchronize(e->stream));
    return MHD3D_OK;
}

// Autotuner: times `steps` ensemble steps for every kernel variant in the
// candidate set (register budget x barrier/update schedule x planes per work
// item) on the current states, restores the states and clocks after each
// trial, and keeps the fastest variant.  Results do not depend on the
// variant; only the time does.  report (optional, 4 * ncand doubles): waves,
// sched, chunk, ms per step of each trial.
extern "C" int mhd3d_ensemble_autotune(mhd3d_ensemble* e, int steps, double* report, int* ncand_out)
{
    if (!e || steps < 1) return MHD3D_EINVAL;
    ENSCHK(e, hipStreamSynchronize(e->stream));
    // snapshot: every member's state register, and the device scalars
    std::vector<double*> snap(e->n, nullptr);
    const int p_save = e->parity;
    std::vector<char> ready_save(e->n);
    std::vector<int> cur_save(e->n);
    // RLAX members: the relaxation's statics (TFAC, the temperature range,
    // the call count) and the host's record of an exit
    std::vector<RelaxState> rs_save(e->n);
    std::vector<char> exit_seen_save(e->n);
    std::vector<int> exit_code_save(e->n);
    for (int m = 0; m < e->n; ++m) {
        mhd3d_ctx* c = e->mem[m];
        ready_save[m] = c->dt_ready;
        cur_save[m] = c->cur;
        exit_seen_save[m] = c->relax_exit_seen;
        exit_code_save[m] = c->relax_exit_code;
        if (c->xa.flags & EXT_RLAX)
            ENSCHK(e, hipMemcpy(&rs_save[m], c->rs, sizeof(RelaxState), hipMemcpyDeviceToHost));
    }
    DevScalars* ds_snap = nullptr;
    auto cleanup = [&]() {
        for (double* p : snap) (void)hipFree(p);
        (void)hipFree(ds_snap);
    };
    for (int m = 0; m < e->n; ++m) {
        mhd3d_ctx* c = e->mem[m];
        const size_t reg = (size_t)c->g.fs * c->nf * sizeof(double);
        if (hipMalloc(&snap[m], reg) != hipSuccess) { cleanup(); return ens_fail(e, MHD3D_ENOMEM, "autotune snapshot"); }
        if (hipMemcpyAsync(snap[m], c->U[c->cur], reg, hipMemcpyDeviceToDevice, e->stream) != hipSuccess) {
            cleanup();
            return ens_fail(e, MHD3D_EHIP, "autotune snapshot copy");
        }
    }
    if (hipMalloc(&ds_snap, sizeof(DevScalars) * e->n) != hipSuccess) { cleanup(); return ens_fail(e, MHD3D_ENOMEM, "autotune snapshot"); }
    for (int m = 0; m < e->n; ++m)
        (void)hipMemcpyAsync(ds_snap + m, e->mem[m]->ds, sizeof(DevScalars), hipMemcpyDeviceToDevice, e->stream);
    auto restore = [&]() -> int {
        e->parity = p_save;
        for (int m = 0; m < e->n; ++m) {
            mhd3d_ctx* c = e->mem[m];
            const size_t reg = (size_t)c->g.fs * c->nf * sizeof(double);
            c->cur = cur_save[m];
            ENSCHK(e, hipMemcpyAsync(c->U[c->cur], snap[m], reg, hipMemcpyDeviceToDevice, e->stream));
        }
        for (int m = 0; m < e->n; ++m)
            ENSCHK(e, hipMemcpyAsync(e->mem[m]->ds, ds_snap + m, sizeof(DevScalars), hipMemcpyDeviceToDevice, e->stream));
        ENSCHK(e, hipStreamSynchronize(e->stream));
        for (int m = 0; m < e->n; ++m) {
            mhd3d_ctx* c = e->mem[m];
            c->dt_ready = ready_save[m] != 0;
            c->relax_exit_seen = exit_seen_save[m] != 0;
            c->relax_exit_code = exit_code_save[m];
            if (c->xa.flags & EXT_RLAX)
                ENSCHK(e, hipMemcpy(c->rs, &rs_save[m], sizeof(RelaxState), hipMemcpyHostToDevice));
        }
        return MHD3D_OK;
    };
    const int wv[2] = {1, 2};
    const int sc[4] = {0, 1, 2, 3};
    const int chs[4] = {0, 8, 16, 32};
    int best_w = e->waves, best_s = e->sched, best_c = e->chunk;
    double best = 1e300;
    int ncand = 0;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    (void)hipEventCreate(&t0);
    (void)hipEventCreate(&t1);
    const bool graph_save = e->use_graph;
    int rc = MHD3D_OK;
    for (int iw = 0; iw < 2 && !rc; ++iw)
        for (int is = 0; is < 4 && !rc; ++is)
            for (int ic = 0; ic < 4 && !rc; ++ic) {
                e->waves = wv[iw];
                e->sched = sc[is];
                e->chunk = chs[ic];
                rc = ens_build(e);
                if (rc) break;
                // a member that reaches an RLAX exit point stops (the trial
                // goes on with the others; the restore below undoes it)
                rc = mhd3d_ensemble_step(e, 1);   // warm-up (and the TSTEP pass if needed)
                if (rc == MHD3D_EEXIT) rc = MHD3D_OK;
                if (rc) break;
                (void)hipEventRecord(t0, e->stream);
                rc = mhd3d_ensemble_step(e, steps);
                if (rc == MHD3D_EEXIT) rc = MHD3D_OK;
                (void)hipEventRecord(t1, e->stream);
                (void)hipEventSynchronize(t1);
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, t0, t1);
                const double per = ms / steps;
                if (report) {
                    report[4 * ncand + 0] = wv[iw];
                    report[4 * ncand + 1] = sc[is];
                    report[4 * ncand + 2] = chs[ic];
                    report[4 * ncand + 3] = per;
                }
                ++ncand;
                if (!rc && per < best) {
                    best = per;
                    best_w = wv[iw];
                    best_s = sc[is];
                    best_c = chs[ic];
                }
                int rr = restore();
                if (!rc) rc = rr;
            }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    e->use_graph = graph_save;
    e->waves = best_w;
    e->sched = best_s;
    e->chunk = best_c;
    int rb = ens_build(e);
    cleanup();
    if (ncand_out) *ncand_out = ncand;
    return rc ? rc : rb;
}

// ===========================================================================
// Request server: runs (grid, flags, parameters, step count) arrive at any
// time; they are admitted first-come-first-served while their device memory
// fits the HBM budget, their state is built and kept on the device, and every
// tick advances all admitted runs together by one quantum of steps through an
// ensemble over them (continuous batching: the ensemble is rebuilt whenever
// runs join or leave).  A run advances exactly as it would alone.

#include <chrono>
#include <deque>
#include <unordered_map>

struct ServerReq {
    int64_t id = 0;
    mhd3d_desc d{};
    std::vector<double> met;          // the descriptor's metric arrays
    std::vector<double> Tz, Rz;       // initial z profiles
    uint64_t seed = 0;
    double amp = 0, b0 = 0, b_amp = 0;
    int nsteps = 0, done = 0;
    int cur = 0;                      // register holding the final state (DONE)
    int status = MHD3D_REQ_QUEUED;
    mhd3d_ctx* ctx = nullptr;
    int64_t bytes = 0;
    double t_submit = 0, t_admit = -1, t_finish = -1, dt_last = 0;
};

struct mhd3d_server {
    int precision = 0;
    int64_t budget = 0, used = 0;
    int quantum = 10;
    mhd3d_ensemble* ens = nullptr;    // view over the admitted runs (owns = false)
    std::unordered_map<int64_t, ServerReq> reqs;
    std::deque<int64_t> queue;
    std::vector<int64_t> active;
    bool dirty = false;
    int64_t next_id = 1;
    int64_t cell_steps = 0;
    std::chrono::steady_clock::time_point t0;
    std::string err;
};

static std::string g_srv_create_error;

static double srv_now(const mhd3d_server* s)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - s->t0).count();
}

// Device bytes a context for this descriptor allocates (mhd3d_create, plus
// the TSTEP partial records of the fused and standalone passes).
static int64_t ctx_bytes_estimate(const mhd3d_desc* d)
{
    const int nf = d->flag_lMag ? NF_MAG : NF_HYDRO;
    const int64_t pitch = (d->nPx + 7) / 8 * 8;
    const int64_t plane = pitch * d->nPy;
    const int64_t fs = (plane * (d->nPz + 2) + 31) / 32 * 32;
    const int64_t reg = fs * nf * 8;
    const bool mixed = d->precision == MHD3D_MIXED;
    const int64_t nmet = 3 * (int64_t)(d->nPx + 2) + 3 * (int64_t)(d->nPy + 2) + 5 * (int64_t)(d->nPz + 2);
    const int64_t tiles = (int64_t)((d->nPx + STX - 1) / STX) * ((d->nPy + STY - 1) / STY);
    const int64_t parts = tiles * (d->nPz / 8 + 1) + (int64_t)((d->nPx + 63) / 64) * ((d->nPy + 3) / 4) * 10;
    return (mixed ? 2 : 3) * reg + (mixed ? reg : 0) + nmet * 8 + (int64_t)sizeof(DevScalars) +
           parts * (int64_t)sizeof(CflPart) + (mixed ? (int64_t)(NBGF * 4 + NBG64 * 8) * (d->nPz + 2) : 0);
}

extern "C" int mhd3d_server_create(int precision, int64_t hbm_budget_bytes, int quantum_steps,
                                   mhd3d_server** out)
{
    if (!out || hbm_budget_bytes <= 0 || quantum_steps < 1 ||
        (precision != MHD3D_FP64 && precision != MHD3D_MIXED))
        return MHD3D_EINVAL;
    mhd3d_server* s = new mhd3d_server();
    s->precision = precision;
    s->budget = hbm_budget_bytes;
    s->quantum = quantum_steps;
    s->t0 = std::chrono::steady_clock::now();
    s->ens = new mhd3d_ensemble();
    s->ens->owns = false;
    s->ens->mixed = precision == MHD3D_MIXED;
    ens_default_variant(s->ens);
    if (hipStreamCreateWithFlags(&s->ens->stream, hipStreamNonBlocking) != hipSuccess) {
        g_srv_create_error = "hipStreamCreate failed";
        mhd3d_server_destroy(s);
        return MHD3D_EHIP;
    }
    *out = s;
    return MHD3D_OK;
}

extern "C" int mhd3d_server_destroy(mhd3d_server* s)
{
    if (!s) return MHD3D_OK;
    if (s->ens) mhd3d_ensemble_destroy(s->ens);   // a view: leaves the runs
    for (auto& kv : s->reqs)
        if (kv.second.ctx) mhd3d_destroy(kv.second.ctx);
    delete s;
    return MHD3D_OK;
}

extern "C" const char* mhd3d_server_last_error(const mhd3d_server* s)
{
    return s ? s->err.c_str() : g_srv_create_error.c_str();
}

extern "C" double mhd3d_server_now(const mhd3d_server* s) { return s ? srv_now(s) : 0.0; }

extern "C" int64_t mhd3d_server_bytes_in_use(const mhd3d_server* s) { return s ? s->used : 0; }

extern "C" int64_t mhd3d_server_cell_steps(const mhd3d_server* s) { return s ? s->cell_steps : 0; }

extern "C" int mhd3d_server_submit(mhd3d_server* s, const mhd3d_desc* desc, const double* TT_z,
                                   const double* RO_z, uint64_t seed, double amp, double b0,
                                   double b_amp, int nsteps, int64_t* id)
{
    if (!s || !desc || !TT_z || !RO_z || nsteps < 1 || !id) return MHD3D_EINVAL;
    std::string why;
    int rc = validate(desc, why);
    if (rc) { s->err = why; return rc; }
    if (desc->precision != s->precision) { s->err = "request precision differs from the server's"; return MHD3D_EINVAL; }
    if (timedep_flags(desc)) {
        s->err = kTimedepWhy;
        return MHD3D_EUNSUPPORTED;
    }
    ServerReq r;
    r.id = s->next_id++;
    r.d = *desc;
    // the descriptor's arrays are copied: the run is created at admission
    const int nx = desc->nPx + 2, ny = desc->nPy + 2, nz = desc->nPz + 2;
    constexpr int NA = 14;
    const double* src[NA] = {desc->dxxdx, desc->d2xxdx2, desc->DDX, desc->dyydy, desc->d2yydy2, desc->DDY,
                             desc->dzzdz, desc->d2zzdz2, desc->DDZ, desc->RKAPA, desc->DKAPA,
                             desc->EXX, desc->WYY, desc->ZEE};
    const int len[NA] = {nx, nx, nx, ny, ny, ny, nz, nz, nz, nz, nz, nx, ny, nz};
    size_t tot = 0;
    for (int k = 0; k < NA; ++k) tot += src[k] ? len[k] : 0;
    r.met.resize(tot);
    size_t o = 0;
    const double** dst[NA] = {&r.d.dxxdx, &r.d.d2xxdx2, &r.d.DDX, &r.d.dyydy, &r.d.d2yydy2, &r.d.DDY,
                              &r.d.dzzdz, &r.d.d2zzdz2, &r.d.DDZ, &r.d.RKAPA, &r.d.DKAPA,
                              &r.d.EXX, &r.d.WYY, &r.d.ZEE};
    std::vector<size_t> offs(NA, SIZE_MAX);
    for (int k = 0; k < NA; ++k) {
        if (!src[k]) continue;   // an optional coordinate array the run does not read
        std::memcpy(&r.met[o], src[k], sizeof(double) * len[k]);
        offs[k] = o;
        o += len[k];
    }
    r.Tz.assign(TT_z, TT_z + nz);
    r.Rz.assign(RO_z, RO_z + nz);
    r.seed = seed;
    r.amp = amp;
    r.b0 = b0;
    r.b_amp = b_amp;
    r.nsteps = nsteps;
    r.bytes = ctx_bytes_estimate(desc);
    if (r.bytes > s->budget) { s->err = "request alone exceeds the HBM budget"; return MHD3D_ENOMEM; }
    r.t_submit = srv_now(s);
    auto& slot = s->reqs[r.id];
    slot = std::move(r);
    for (int k = 0; k < NA; ++k) *dst[k] = nullptr;
    // point the stored descriptor at the stored copies
    const double** sd[NA] = {&slot.d.dxxdx, &slot.d.d2xxdx2, &slot.d.DDX, &slot.d.dyydy, &slot.d.d2yydy2,
                             &slot.d.DDY, &slot.d.dzzdz, &slot.d.d2zzdz2, &slot.d.DDZ, &slot.d.RKAPA,
                             &slot.d.DKAPA, &slot.d.EXX, &slot.d.WYY, &slot.d.ZEE};
    for (int k = 0; k < NA; ++k) *sd[k] = offs[k] == SIZE_MAX ? nullptr : slot.met.data() + offs[k];
    s->queue.push_back(slot.id);
    *id = slot.id;
    return MHD3D_OK;
}

// Admits queued runs in arrival order while they fit the budget.
static int srv_admit(mhd3d_server* s)
{
    while (!s->queue.empty()) {
        ServerReq& r = s->reqs[s->queue.front()];
        if (s->used + r.bytes > s->budget) break;
        mhd3d_ctx* c = nullptr;
        int rc = mhd3d_create(&r.d, &c);
        if (rc) { s->err = g_create_error; return rc; }
        rc = mhd3d_init_stratified(c, r.Tz.data(), r.Rz.data(), r.seed, r.amp, r.b0, r.b_amp);
        if (!rc) rc = mhd3d_set_dt_guard(c, MHD3D_REFERENCE_DT_MIN);
        if (!rc) {
            // the run stops on the device after its own step count, whatever
            // the quantum of the ensemble it runs in
            set_stop_kernel<<<1, 1, 0, c->stream>>>(c->ds, (double)r.nsteps);
            if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess) {
                c->err = "step budget setup failed";
                rc = MHD3D_EHIP;
            }
        }
        if (rc) { s->err = c->err; mhd3d_destroy(c); return rc; }
        r.ctx = c;
        r.status = MHD3D_REQ_RUNNING;
        r.t_admit = srv_now(s);
        s->used += r.bytes;
        s->active.push_back(r.id);
        s->queue.pop_front();
        s->dirty = true;
    }
    return MHD3D_OK;
}

// Rebuilds the server's ensemble over the active runs after joins or leaves.
static int srv_rebuild(mhd3d_server* s)
{
    mhd3d_ensemble* e = s->ens;
    e->mem.clear();
    e->cells = 0;
    for (int64_t id : s->active) {
        mhd3d_ctx* c = s->reqs[id].ctx;
        e->mem.push_back(c);
        e->cells += (int64_t)c->g.nPx * c->g.nPy * c->g.nPz;
    }
    e->n = (int)e->mem.size();
    int rc = ens_build(e);
    if (rc) { s->err = e->err; return rc; }
    s->dirty = false;
    return MHD3D_OK;
}

extern "C" int mhd3d_server_tick(mhd3d_server* s, int* n_active, int* n_queued)
{
    if (!s) return MHD3D_EINVAL;
    int rc = MHD3D_OK;
    if (s->active.empty()) {
        rc = srv_admit(s);
        if (rc) return rc;
    }
    if (!s->active.empty()) {
        mhd3d_ensemble* e = s->ens;
        if (s->dirty) {
            rc = srv_rebuild(s);
            if (rc) return rc;
        }
        // one quantum for every active run; each stops on the device at its
        // own step count (set_stop_kernel), so a run near its end does not
        // shorten the others' quantum
        const int k = s->quantum;
        rc = ens_step_enqueue(e, k);
        if (rc) { s->err = e->err; return rc; }
        // admission overlaps the quantum on the device: new runs are created
        // and initialised on their own streams and join at the next tick
        int rca = srv_admit(s);
        rc = ens_step_finish(e, k);
        if (rc) { s->err = e->err; return rc; }
        const double now = srv_now(s);
        std::vector<int64_t> keep;
        for (int m = 0; m < e->n; ++m) {
            ServerReq& r = s->reqs[s->active[m]];
            const DevScalars& hs = e->last[m];
            r.dt_last = hs.dt_last;
            const int now_done = (int)hs.nit;
            s->cell_steps += (int64_t)(now_done - r.done) * r.ctx->g.nPx * r.ctx->g.nPy * r.ctx->g.nPz;
            r.done = now_done;
            if (hs.halted & HALT_EXIT) {    // the reference's program exits there (RLAX, 3dmhdsub.h:344,416,427-439)
                r.status = MHD3D_REQ_EXITED;
                r.t_finish = now;
            } else if (hs.halted & HALT_GUARD) {   // the reference's driver halts such a run (3dmhd.cpp:292-330)
                r.status = MHD3D_REQ_FAILED;
                r.t_finish = now;
            } else if (hs.nonfinite) {      // evicted: its state holds NaN/Inf
                r.status = MHD3D_REQ_NONFINITE;
                r.t_finish = now;
            } else if (r.done >= r.nsteps) {
                r.status = MHD3D_REQ_DONE;
                r.cur = r.ctx->cur;
                r.t_finish = now;
            } else {
                keep.push_back(r.id);
            }
        }
        for (int64_t id : s->active) {
            ServerReq& r = s->reqs[id];
            if (r.status == MHD3D_REQ_RUNNING) continue;
            s->dirty = true;
        }
        // runs admitted during the quantum follow the runs still active
        std::vector<int64_t> joined(s->active.begin() + e->n, s->active.end());
        s->active.swap(keep);
        s->active.insert(s->active.end(), joined.begin(), joined.end());
        if (!joined.empty()) s->dirty = true;
        if (rca) return rca;
    }
    if (n_active) *n_active = (int)s->active.size();
    if (n_queued) *n_queued = (int)s->queue.size();
    return MHD3D_OK;
}

extern "C" int mhd3d_server_poll(mhd3d_server* s, int64_t id, mhd3d_request_info* info)
{
    if (!s || !info) return MHD3D_EINVAL;
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) { s->err = "unknown request"; return MHD3D_EINVAL; }
    const ServerReq& r = it->second;
    info->status = r.status;
    info->steps_done = r.done;
    info->t_submit = r.t_submit;
    info->t_admit = r.t_admit;
    info->t_finish = r.t_finish;
    info->dt_last = r.dt_last;
    return MHD3D_OK;
}

extern "C" int mhd3d_server_result(mhd3d_server* s, int64_t id, double* const* fields)
{
    if (!s) return MHD3D_EINVAL;
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) { s->err = "unknown request"; return MHD3D_EINVAL; }
    ServerReq& r = it->second;
    if (r.status != MHD3D_REQ_DONE && r.status != MHD3D_REQ_FAILED && r.status != MHD3D_REQ_NONFINITE &&
        r.status != MHD3D_REQ_EXITED) {
        s->err = "request not finished";
        return MHD3D_EINVAL;
    }
    int rc = MHD3D_OK;
    if (fields && r.status == MHD3D_REQ_DONE) {
        r.ctx->cur = r.cur;   // other runs' ticks since it finished do not move it
        rc = mhd3d_get_state(r.ctx, fields);
        if (rc) s->err = r.ctx->err;
    }
    if (r.ctx) {
        mhd3d_destroy(r.ctx);
        r.ctx = nullptr;
        s->used -= r.bytes;
    }
    r.status = r.status == MHD3D_REQ_DONE ? MHD3D_REQ_RELEASED : r.status;
    return rc;
}

// Calibration of HBM byte counters: `reps` launches of a plain copy of
// `bytes` bytes between two fresh device buffers (so a profiler's per-
// dispatch FETCH_SIZE / WRITE_SIZE can be compared with a known traffic of
// `bytes` read and `bytes` written); *ms_per_copy: event-timed average.
extern "C" int mhd3d_calibration_copy(int64_t bytes, int reps, double* ms_per_copy)
{
    if (bytes < 8 || reps < 1) return MHD3D_EINVAL;
    const int64_t n = bytes / 8;
    double *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = MHD3D_OK;
    if (hipMalloc(&a, n * 8) != hipSuccess || hipMalloc(&b, n * 8) != hipSuccess) rc = MHD3D_ENOMEM;
    if (!rc && (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
                hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = MHD3D_EHIP;
    if (!rc && hipMemsetAsync(a, 0, n * 8, st) != hipSuccess) rc = MHD3D_EHIP;
    if (!rc) {
        calibration_copy_kernel<<<4096, 256, 0, st>>>(a, b, n);   // warm-up
        (void)hipEventRecord(e0, st);
        for (int r = 0; r < reps; ++r) calibration_copy_kernel<<<4096, 256, 0, st>>>(a, b, n);
        (void)hipEventRecord(e1, st);
        if (hipStreamSynchronize(st) != hipSuccess) rc = MHD3D_EHIP;
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms_per_copy) *ms_per_copy = ms / reps;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}

// Streaming ceiling (stream_ceiling_kernel): `reps` launches moving nr input
// and nw output arrays of `bytes_per_array` bytes each, 4 independent 16-byte
// loads per thread and array in flight, 2048 blocks of 256 threads (8 per CU
// on 256 CUs); *ms_per_launch: event-timed average.
extern "C" int mhd3d_stream_ceiling(int64_t bytes_per_array, int nr, int nw, int reps, double* ms_per_launch)
{
    if (bytes_per_array < 16 || reps < 1 || nr < 1 || nr > 2 || nw < 0 || nw > 2) return MHD3D_EINVAL;
    const int64_t n = bytes_per_array / 16;
    dvec2 *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = MHD3D_OK;
    if (hipMalloc(&a, n * 16 * nr) != hipSuccess || hipMalloc(&b, n * 16 * (nw > 0 ? nw : 1)) != hipSuccess)
        rc = MHD3D_ENOMEM;
    if (!rc && (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
                hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = MHD3D_EHIP;
    if (!rc && (hipMemsetAsync(a, 0, n * 16 * nr, st) != hipSuccess ||
                hipMemsetAsync(b, 0, n * 16 * (nw > 0 ? nw : 1), st) != hipSuccess))
        rc = MHD3D_EHIP;
    auto launch = [&]() {
        const dim3 grd(2048), blk(256);
        if (nr == 1 && nw == 0) stream_ceiling_kernel<1, 0, 4><<<grd, blk, 0, st>>>(a, b, n);
        else if (nr == 1 && nw == 1) stream_ceiling_kernel<1, 1, 4><<<grd, blk, 0, st>>>(a, b, n);
        else if (nr == 1 && nw == 2) stream_ceiling_kernel<1, 2, 4><<<grd, blk, 0, st>>>(a, b, n);
        else if (nr == 2 && nw == 0) stream_ceiling_kernel<2, 0, 4><<<grd, blk, 0, st>>>(a, b, n);
        else if (nr == 2 && nw == 1) stream_ceiling_kernel<2, 1, 4><<<grd, blk, 0, st>>>(a, b, n);
        else stream_ceiling_kernel<2, 2, 4><<<grd, blk, 0, st>>>(a, b, n);
    };
    if (!rc) {
        launch();   // warm-up
        (void)hipEventRecord(e0, st);
        for (int r = 0; r < reps; ++r) launch();
        (void)hipEventRecord(e1, st);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) rc = MHD3D_EHIP;
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms_per_launch) *ms_per_launch = ms / reps;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}
