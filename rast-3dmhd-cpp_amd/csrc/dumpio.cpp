// The reference's dump/restart files (3dmhdio.h:20-72), written and read
// from a context's state.
//
// .dat0 file: records of nCore = nPx*nPy*nPz doubles; dump number idx holds
// the fields RU, RV, RW, TT, RO (then BX, BY, BZ with lMag) as records
// n_f*idx .. n_f*idx + n_f - 1, each the planes 1..nPz, rows 1..nPy,
// columns 1..nPx of the [nz][ny][nx] array, x fastest (save(), 3dmhdutils.h
// :105-110).  .par file: records of 96 doubles, record idx = par1 of dump idx
// (par1[0..31] = ipar, par1[32..95] = par; 3dmhd.cpp:271-289).
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/mhd3d.h"

namespace {

// field order of a record group: RU RV RW TT RO BX BY BZ (3dmhdio.h:24-33)
const int kRecordField[8] = {MHD3D_RU, MHD3D_RV, MHD3D_RW, MHD3D_TT, MHD3D_RO,
                             MHD3D_BX, MHD3D_BY, MHD3D_BZ};

FILE* open_rw(const char* path)
{
    FILE* f = std::fopen(path, "r+b");
    if (!f) f = std::fopen(path, "w+b");
    return f;
}

struct Shape {
    int nPx, nPy, nPz, nf;
    size_t nx, ny, nz, arr, core;
};

int shape_of(mhd3d_ctx* ctx, Shape& s)
{
    int32_t g[4];
    int rc = mhd3d_geometry(ctx, g);
    if (rc) return rc;
    s.nPx = g[0]; s.nPy = g[1]; s.nPz = g[2]; s.nf = g[3];
    s.nx = (size_t)s.nPx + 2; s.ny = (size_t)s.nPy + 2; s.nz = (size_t)s.nPz + 2;
    s.arr = s.nx * s.ny * s.nz;
    s.core = (size_t)s.nPx * s.nPy * s.nPz;
    return MHD3D_OK;
}

int get_all(mhd3d_ctx* ctx, const Shape& s, std::vector<double>& buf, double** ptrs)
{
    buf.assign(s.arr * s.nf, 0.0);
    for (int f = 0; f < s.nf; ++f) ptrs[f] = buf.data() + (size_t)f * s.arr;
    return mhd3d_get_state(ctx, ptrs);
}

}  // namespace

extern "C" int mhd3d_dump_data(mhd3d_ctx* ctx, const char* path, int idx)
{
    if (!ctx || !path || idx < 0) return MHD3D_EINVAL;
    Shape s;
    int rc = shape_of(ctx, s);
    if (rc) return rc;
    std::vector<double> buf;
    double* ptrs[8];
    rc = get_all(ctx, s, buf, ptrs);
    if (rc) return rc;
    FILE* f = open_rw(path);
    if (!f) return MHD3D_EINVAL;
    std::vector<double> rec(s.core);
    for (int r = 0; r < s.nf; ++r) {
        const double* a = ptrs[kRecordField[r]];
        size_t o = 0;
        for (size_t z = 1; z <= (size_t)s.nPz; ++z)
            for (size_t y = 1; y <= (size_t)s.nPy; ++y) {
                std::memcpy(&rec[o], a + (z * s.ny + y) * s.nx + 1, sizeof(double) * s.nPx);
                o += s.nPx;
            }
        const long long off = ((long long)s.nf * idx + r) * (long long)s.core * 8;
        if (fseeko(f, (off_t)off, SEEK_SET) != 0 || std::fwrite(rec.data(), 8, s.core, f) != s.core) {
            std::fclose(f);
            return MHD3D_EINVAL;
        }
    }
    return std::fclose(f) == 0 ? MHD3D_OK : MHD3D_EINVAL;
}

extern "C" int mhd3d_load_data(mhd3d_ctx* ctx, const char* path, int idx)
{
    if (!ctx || !path || idx < 0) return MHD3D_EINVAL;
    Shape s;
    int rc = shape_of(ctx, s);
    if (rc) return rc;
    std::vector<double> buf;
    double* ptrs[8];
    rc = get_all(ctx, s, buf, ptrs);   // keeps the context's z ghost planes
    if (rc) return rc;
    FILE* f = std::fopen(path, "rb");
    if (!f) return MHD3D_EINVAL;
    std::vector<double> rec(s.core);
    for (int r = 0; r < s.nf; ++r) {
        const long long off = ((long long)s.nf * idx + r) * (long long)s.core * 8;
        if (fseeko(f, (off_t)off, SEEK_SET) != 0 || std::fread(rec.data(), 8, s.core, f) != s.core) {
            std::fclose(f);
            return MHD3D_EINVAL;
        }
        double* a = ptrs[kRecordField[r]];
        size_t o = 0;
        for (size_t z = 1; z <= (size_t)s.nPz; ++z)
            for (size_t y = 1; y <= (size_t)s.nPy; ++y) {
                std::memcpy(a + (z * s.ny + y) * s.nx + 1, &rec[o], sizeof(double) * s.nPx);
                o += s.nPx;
            }
    }
    std::fclose(f);
    return mhd3d_set_state(ctx, ptrs);   // x, y ghosts are periodic on the device
}

extern "C" int mhd3d_dump_param(const char* path, int idx, const double* par1)
{
    if (!path || !par1 || idx < 0) return MHD3D_EINVAL;
    FILE* f = open_rw(path);
    if (!f) return MHD3D_EINVAL;
    const bool ok = fseeko(f, (off_t)idx * 96 * 8, SEEK_SET) == 0 && std::fwrite(par1, 8, 96, f) == 96;
    return (std::fclose(f) == 0 && ok) ? MHD3D_OK : MHD3D_EINVAL;
}

extern "C" int mhd3d_load_param(const char* path, int idx, double* par1)
{
    if (!path || !par1 || idx < 0) return MHD3D_EINVAL;
    FILE* f = std::fopen(path, "rb");
    if (!f) return MHD3D_EINVAL;
    const bool ok = fseeko(f, (off_t)idx * 96 * 8, SEEK_SET) == 0 && std::fread(par1, 8, 96, f) == 96;
    std::fclose(f);
    return ok ? MHD3D_OK : MHD3D_EINVAL;
}
