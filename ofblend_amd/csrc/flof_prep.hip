// flof_prep.hip -- input preprocessing of the FlOF pipeline and the .uni
// grid file format.
//
// Reference: source/plugin/optflow4d.cpp -- repeatFrame4d :1254-1266,
// knExtrap4dLsSimple / knSetRemaining4d / extrap4dLsSimple :1323-1406;
// source/fileio.cpp -- UniHeader :38-44, getUniFileSize :572-599,
// writeGridUni / readGridUni (MNT2) :608-731, writeGrid4dUni /
// readGrid4dUni (M4T2) :834-999.

#include <zlib.h>

#include <algorithm>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "flof_internal.h"

namespace flof {

// ---------------------------------------------------------------------------
// repeatFrame4d: every (bnd-interior) column takes its quadrilinear sample at
// time srct, written to slices [t_lo, t_hi].  A thread reads its column
// before writing it; other columns enter the sample with weight 0 only.
__global__ void k_repeat_frame(float *__restrict__ phi, Dims4 D, float srct, int t_lo,
                               int t_hi, int bnd)
{
    const int nx = D.X - 2 * bnd, ny = D.Y - 2 * bnd, nz = D.Z - 2 * bnd;
    if (nx <= 0 || ny <= 0 || nz <= 0) return;
    const int64_t n = (int64_t)nx * ny * nz;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n;
         c += (int64_t)gridDim.x * blockDim.x) {
        const int i = bnd + (int)(c % nx);
        const int j = bnd + (int)((c / nx) % ny);
        const int k = bnd + (int)(c / ((int64_t)nx * ny));
        const float v = interp4<float>(phi, D, (float)i + 0.5f, (float)j + 0.5f,
                                       (float)k + 0.5f, srct + 0.5f);
        const int64_t c3 = (int64_t)i + (int64_t)D.X * j + D.sz * k;
        for (int t = t_lo; t <= t_hi; ++t) phi[c3 + D.st * t] = v;
    }
}

// extrap4dLsSimple
__device__ __forceinline__ bool inner(const Dims4 &D, int i, int j, int k, int t)
{
    return i >= 1 && j >= 1 && k >= 1 && t >= 1 && i < D.X - 1 && j < D.Y - 1 &&
           k < D.Z - 1 && t < D.T - 1;
}

__global__ void k_extrap_mark(const float *__restrict__ phi, int *__restrict__ tmp,
                              Dims4 D, Coord4 C, int inside)
{
    const int64_t n = (int64_t)D.st * D.T;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n;
         c += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k, t;
        decompose((uint32_t)c, C, i, j, k, t);
        int m = 0;
        if (inner(D, i, j, k, t)) m = (!inside ? phi[c] < 0.f : phi[c] > 0.f) ? 1 : 0;
        tmp[c] = m;
    }
}

// first layer: unmarked interior cells next to a marked one (reads ==1 only,
// so cells turned to 2 in this pass do not interfere)
__global__ void k_extrap_first(int *tmp, Dims4 D, Coord4 C)
{
    const int64_t n = (int64_t)D.st * D.T;
    const int64_t off[8] = {-1, 1, -(int64_t)D.X, (int64_t)D.X, -D.sz, D.sz, -D.st, D.st};
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n;
         c += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k, t;
        decompose((uint32_t)c, C, i, j, k, t);
        if (!inner(D, i, j, k, t) || tmp[c] != 0) continue;
        bool nb = false;
        for (int q = 0; q < 8; ++q) nb = nb || (tmp[c + off[q]] == 1);
        if (nb) tmp[c] = 2;
    }
}

// layer d+1 from layer d: mean of the layer-d neighbours (x-, x+, y-, y+, z-,
// z+, t-, t+ order) plus `direction`
__global__ void k_extrap_layer(float *phi, int *tmp, Dims4 D, Coord4 C, int d,
                               float direction)
{
    const int64_t n = (int64_t)D.st * D.T;
    const int64_t off[8] = {-1, 1, -(int64_t)D.X, (int64_t)D.X, -D.sz, D.sz, -D.st, D.st};
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n;
         c += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k, t;
        decompose((uint32_t)c, C, i, j, k, t);
        if (!inner(D, i, j, k, t) || tmp[c] != 0) continue;
        int nbs = 0;
        float avg = 0.f;
        for (int q = 0; q < 8; ++q)
            if (tmp[c + off[q]] == d) {
                avg = avg + phi[c + off[q]];
                nbs++;
            }
        if (nbs > 0) {
            tmp[c] = d + 1;
            phi[c] = avg / (float)nbs + direction;
        }
    }
}

__global__ void k_extrap_remaining(float *phi, const int *tmp, Dims4 D, Coord4 C, float set)
{
    const int64_t n = (int64_t)D.st * D.T;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n;
         c += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k, t;
        decompose((uint32_t)c, C, i, j, k, t);
        if (inner(D, i, j, k, t) && tmp[c] == 0) phi[c] = set;
    }
}

// ---------------------------------------------------------------------------
// .uni format (fileio.cpp:38-44): "MNT2"/"M4T2" + header; M4T2 adds int dimT,
// then the slices; gzip level 1 on write.
struct UniHeader {
    int dimX, dimY, dimZ;
    int gridType, elementType, bytesPerElement;
    char info[256];
    unsigned long long timestamp;
};
static_assert(sizeof(UniHeader) == 288, "UniHeader layout");

struct Gz {
    gzFile f = nullptr;
    ~Gz()
    {
        if (f) gzclose(f);
    }
};

static void gz_read_all(gzFile f, void *dst, size_t bytes, const std::string &name)
{
    char *p = (char *)dst;
    while (bytes > 0) {
        const unsigned chunk = (unsigned)std::min<size_t>(bytes, 1u << 30);
        const int got = gzread(f, p, chunk);
        if (got <= 0) throw Error{FLOF_ERR_ARG, "can't read file " + name + ": truncated"};
        p += got;
        bytes -= (size_t)got;
    }
}
static void gz_write_all(gzFile f, const void *src, size_t bytes, const std::string &name)
{
    const char *p = (const char *)src;
    while (bytes > 0) {
        const unsigned chunk = (unsigned)std::min<size_t>(bytes, 1u << 30);
        const int put = gzwrite(f, p, chunk);
        if (put <= 0) throw Error{FLOF_ERR_ARG, "can't write file " + name};
        p += put;
        bytes -= (size_t)put;
    }
}

// grid type and element type codes written by the reference for 4D grids
// (Grid4dBase::Grid4dType, fileio.cpp:848-857)
static void uni_types(int elem, int &gridType, int &elementType, int &bpe)
{
    switch (elem) {
    case FLOF_INT: gridType = 2; elementType = 0; bpe = 4; break;
    case FLOF_VEC4: gridType = 8; elementType = 2; bpe = 16; break;
    default: gridType = 1; elementType = 1; bpe = 4; break;
    }
}

void uni_write4d(const std::string &name, const void *host, int X, int Y, int Z, int T,
                 int elem)
{
    UniHeader h;
    std::memset(&h, 0, sizeof(h));
    h.dimX = X;
    h.dimY = Y;
    h.dimZ = Z;
    uni_types(elem, h.gridType, h.elementType, h.bytesPerElement);
    std::snprintf(h.info, sizeof(h.info), "%s", "ofblend_amd flof (gfx950)");
    h.timestamp = (unsigned long long)std::time(nullptr);
    Gz gz;
    gz.f = gzopen(name.c_str(), "wb1");
    FLOF_REQUIRE(gz.f, FLOF_ERR_ARG, "can't open file " + name);
    gz_write_all(gz.f, "M4T2", 4, name);
    gz_write_all(gz.f, &h, sizeof(h), name);
    gz_write_all(gz.f, &T, sizeof(int), name);
    const size_t slice = (size_t)X * Y * Z * h.bytesPerElement;
    for (int t = 0; t < T; ++t) gz_write_all(gz.f, (const char *)host + slice * t, slice, name);
    const int rc = gzclose(gz.f);
    gz.f = nullptr;
    FLOF_REQUIRE(rc == Z_OK, FLOF_ERR_ARG, "error closing " + name);
}

// reads the header: returns "MNT2" or "M4T2" files' sizes (T = 1 for 3D)
static void uni_header(gzFile f, const std::string &name, UniHeader &h, int &T, bool &is4d)
{
    char id[5] = {0, 0, 0, 0, 0};
    gz_read_all(f, id, 4, name);
    if (!std::strcmp(id, "M4T2")) is4d = true;
    else if (!std::strcmp(id, "MNT2")) is4d = false;
    else throw Error{FLOF_ERR_ARG, "Unknown header in " + name};
    gz_read_all(f, &h, sizeof(h), name);
    T = 1;
    if (is4d) gz_read_all(f, &T, sizeof(int), name);
}

void uni_size(const std::string &name, int dims[4])
{
    Gz gz;
    gz.f = gzopen(name.c_str(), "rb");
    FLOF_REQUIRE(gz.f, FLOF_ERR_ARG, "can't open file " + name);
    UniHeader h;
    int T;
    bool is4d;
    uni_header(gz.f, name, h, T, is4d);
    dims[0] = h.dimX;
    dims[1] = h.dimY;
    dims[2] = h.dimZ;
    dims[3] = T;
}

// unifyGridType (fileio.cpp:470-479): Real <> levelset, Vec3 <> MAC bits
static int unify(int t)
{
    if (t & 1) t |= 16;
    if (t & 16) t |= 1;
    if (t & 4) t |= 8;
    if (t & 8) t |= 4;
    return t;
}

void uni_read(const std::string &name, void *host, int X, int Y, int Z, int T, int elem)
{
    Gz gz;
    gz.f = gzopen(name.c_str(), "rb");
    FLOF_REQUIRE(gz.f, FLOF_ERR_ARG, "can't open file " + name);
    UniHeader h;
    int fT;
    bool is4d;
    uni_header(gz.f, name, h, fT, is4d);
    int gt, et, bpe;
    uni_types(elem, gt, et, bpe);
    FLOF_REQUIRE(h.dimX == X && h.dimY == Y && h.dimZ == Z && fT == T, FLOF_ERR_DIM,
                 "grid dim doesn't match in " + name);
    FLOF_REQUIRE(h.bytesPerElement == bpe, FLOF_ERR_TYPE,
                 "grid element size doesn't match in " + name);
    FLOF_REQUIRE(unify(h.gridType) == unify(gt), FLOF_ERR_TYPE,
                 "grid type doesn't match in " + name);
    gz_read_all(gz.f, host, (size_t)X * Y * Z * T * bpe, name);
}

}  // namespace flof

using namespace flof;

#define API_TRY try {
#define API_CATCH(ctx)                                                        \
    }                                                                         \
    catch (const flof::Error &e)                                              \
    {                                                                         \
        flof::set_error(ctx, e);                                              \
        return e.code;                                                        \
    }                                                                         \
    catch (const std::exception &e)                                           \
    {                                                                         \
        flof::set_error(ctx, flof::Error{FLOF_ERR_ARG, e.what()});            \
        return FLOF_ERR_ARG;                                                  \
    }                                                                         \
    return FLOF_OK;

extern "C" {

int flof_repeat_frame4d(flof_grid4d *phi, float srct, float range, int bnd)
{
    flof_ctx *ctx = phi ? phi->ctx : nullptr;
    API_TRY
    FLOF_REQUIRE(phi && phi->elem == FLOF_REAL && phi->dtype == FLOF_F32, FLOF_ERR_TYPE,
                 "repeatFrame4d: REAL FP32 grid");
    const int t_lo = int(srct - range + 0.5f), t_hi = int(srct);
    if (t_hi < t_lo) return FLOF_OK;
    FLOF_REQUIRE(t_lo >= 0 && t_hi < phi->T && bnd >= 0, FLOF_ERR_RANGE,
                 "repeatFrame4d: slices outside the grid");
    FLOF_REQUIRE(phi->X >= 2 && phi->Y >= 2 && phi->Z >= 2 && phi->T >= 2, FLOF_ERR_DIM,
                 "repeatFrame4d: grid too small");
    Dims4 D = dims_of(phi);
    const int64_t n = (int64_t)std::max(0, D.X - 2 * bnd) * std::max(0, D.Y - 2 * bnd) *
                      std::max(0, D.Z - 2 * bnd);
    hipLaunchKernelGGL(k_repeat_frame, dim3(grid_blocks(ctx, n)), dim3(256), 0, ctx->stream,
                       (float *)phi->data, D, srct, t_lo, t_hi, bnd);
    check_launch("k_repeat_frame");
    API_CATCH(ctx)
}

int flof_extrap4d_ls_simple(flof_grid4d *phi, int distance, int inside)
{
    flof_ctx *ctx = phi ? phi->ctx : nullptr;
    API_TRY
    FLOF_REQUIRE(phi && phi->elem == FLOF_REAL && phi->dtype == FLOF_F32, FLOF_ERR_TYPE,
                 "extrap4dLsSimple: REAL FP32 grid");
    Dims4 D = dims_of(phi);
    Coord4 C = make_coord4(D);
    const int64_t n = phi->n();
    const int nb = grid_blocks(ctx, n);
    Tmp tmp(ctx, (size_t)n * 4);
    float *p = (float *)phi->data;
    int *t = tmp.as<int>();
    hipLaunchKernelGGL(k_extrap_mark, dim3(nb), dim3(256), 0, ctx->stream, p, t, D, C,
                       inside ? 1 : 0);
    hipLaunchKernelGGL(k_extrap_first, dim3(nb), dim3(256), 0, ctx->stream, t, D, C);
    const float direction = inside ? -1.f : 1.f;
    for (int d = 2; d < 1 + distance; ++d)
        hipLaunchKernelGGL(k_extrap_layer, dim3(nb), dim3(256), 0, ctx->stream, p, t, D, C,
                           d, direction);
    hipLaunchKernelGGL(k_extrap_remaining, dim3(nb), dim3(256), 0, ctx->stream, p,
                       (const int *)t, D, C, direction * (float)(distance + 2));
    check_launch("k_extrap_*");
    API_CATCH(ctx)
}

int flof_shift_forw_grid4d(flof_grid4d *phi, int overrideGoodRegion)
{
    flof_ctx *ctx = phi ? phi->ctx : nullptr;
    API_TRY
    FLOF_REQUIRE(phi && phi->elem == FLOF_REAL, FLOF_ERR_TYPE, "shiftForwGrid4d: REAL grid");
    FLOF_REQUIRE(overrideGoodRegion >= 0 && overrideGoodRegion <= phi->T, FLOF_ERR_RANGE,
                 "shiftForwGrid4d: good region outside the grid");
    // optflow4d.cpp:1598-1616: slices [0, good) take slices [shift, shift+good)
    const int shift = phi->T - overrideGoodRegion;
    const size_t sl = (size_t)phi->X * phi->Y * phi->Z * 4;
    for (int s = 0; s < overrideGoodRegion; s += std::max(1, shift)) {
        int len = shift;
        if (s + shift >= overrideGoodRegion) len = overrideGoodRegion - s;
        if (len <= 0 || shift <= 0) break;
        FLOF_HIP(hipMemcpyAsync((char *)phi->data + sl * s, (char *)phi->data + sl * (s + shift),
                                sl * len, hipMemcpyDeviceToDevice, ctx->stream));
    }
    API_CATCH(ctx)
}

int flof_uni_file_size(const char *name, int dims[4])
{
    try {
        FLOF_REQUIRE(name && dims, FLOF_ERR_ARG, "getUniFileSize: null argument");
        uni_size(name, dims);
    } catch (const flof::Error &e) {
        return e.code;
    }
    return FLOF_OK;
}

int flof_uni_write4d(const char *name, const void *host, int X, int Y, int Z, int T, int elem)
{
    try {
        FLOF_REQUIRE(name && host && X > 0 && Y > 0 && Z > 0 && T > 0, FLOF_ERR_ARG,
                     "writeGrid4dUni: bad argument");
        uni_write4d(name, host, X, Y, Z, T, elem);
    } catch (const flof::Error &e) {
        return e.code;
    }
    return FLOF_OK;
}

int flof_uni_read(const char *name, void *host, int X, int Y, int Z, int T, int elem)
{
    try {
        FLOF_REQUIRE(name && host, FLOF_ERR_ARG, "readGridUni: bad argument");
        uni_read(name, host, X, Y, Z, T, elem);
    } catch (const flof::Error &e) {
        return e.code;
    }
    return FLOF_OK;
}

// the reference formats file names with snprintf(pattern, index): accept
// exactly one integer conversion (%d / %i with optional flags and width)
static std::string format_index(const char *pattern, int idx)
{
    int conv = 0;
    for (const char *c = pattern; *c; ++c) {
        if (*c != '%') continue;
        if (c[1] == '%') {
            ++c;
            continue;
        }
        const char *e = c + 1;
        while (*e && std::strchr("0123456789-+ #", *e)) ++e;
        FLOF_REQUIRE(*e == 'd' || *e == 'i', FLOF_ERR_ARG,
                     std::string("file pattern needs one integer conversion: ") + pattern);
        ++conv;
        c = e;
    }
    FLOF_REQUIRE(conv == 1, FLOF_ERR_ARG,
                 std::string("file pattern needs one integer conversion: ") + pattern);
    char buf[4096];
    std::snprintf(buf, sizeof(buf), pattern, idx);
    return buf;
}

int flof_load_place_grid4d_uni(const char *fname, flof_grid4d *phi, const float offset[4],
                               const float scale[4], int fileIdxStart, int fileIdxEnd,
                               int debugSkipLoad, float spread, const float overrideSize[4],
                               float overrideTimeOff, int overrideGoodRegion,
                               float loadTimeScale, int rescaleSdfValues, float sdfIsoOff,
                               float repeatStartFrame)
{
    flof_ctx *ctx = phi ? phi->ctx : nullptr;
    API_TRY
    FLOF_REQUIRE(fname && phi && offset && scale, FLOF_ERR_ARG,
                 "loadPlaceGrid4d: null argument");
    // the size comes from the file of index fileIdxStart as passed
    int dims[4];
    uni_size(format_index(fname, fileIdxStart), dims);
    FLOF_REQUIRE(dims[0] >= 1 && dims[1] >= 1 && dims[2] >= 1, FLOF_ERR_DIM,
                 std::string("Invalid src size from ") + fname);
    const size_t fsz = (size_t)dims[0] * dims[1] * dims[2];
    std::vector<float> host(fsz);
    Tmp dev(ctx, fsz * 4);
    place_frames(
        phi, dims[0], dims[1], dims[2],
        [&](int f) -> const float * {
            // kernels of the previous frame may still read the buffer
            FLOF_HIP(hipStreamSynchronize(ctx->stream));
            uni_read(format_index(fname, f), host.data(), dims[0], dims[1], dims[2], 1,
                     FLOF_REAL);
            FLOF_HIP(hipMemcpyAsync(dev.p, host.data(), fsz * 4, hipMemcpyHostToDevice,
                                    ctx->stream));
            FLOF_HIP(hipStreamSynchronize(ctx->stream));
            return dev.as<float>();
        },
        offset, scale, fileIdxStart, fileIdxEnd, debugSkipLoad, spread, overrideSize,
        overrideTimeOff, overrideGoodRegion, loadTimeScale, rescaleSdfValues, sdfIsoOff,
        repeatStartFrame);
    API_CATCH(ctx)
}

int flof_warp_state_open(flof_ctx *ctx, const char *fname, int useDefoVols, int doAligned,
                         float partialLoadFac, flof_warp_state **out)
{
    API_TRY
    FLOF_REQUIRE(ctx && fname && out, FLOF_ERR_ARG, "loadAdvectTimeSlice_OptInit: null argument");
    *out = nullptr;
    flof_warp_state *s = new flof_warp_state();
    s->ctx = ctx;
    s->use_vols = useDefoVols != 0;
    s->aligned = doAligned != 0;
    s->partial = partialLoadFac;
    const int rc = flof_warp_state_add(s, fname);
    if (rc != FLOF_OK) {
        delete s;
        return rc;
    }
    *out = s;
    API_CATCH(ctx)
}

int flof_warp_state_add(flof_warp_state *s, const char *fname)
{
    flof_ctx *ctx = s ? s->ctx : nullptr;
    API_TRY
    FLOF_REQUIRE(s && fname, FLOF_ERR_ARG, "loadAdvectTimeSlice_OptAdd: null argument");
    // without deformation volumes the reference ignores the call
    if (!s->defos.empty() && !s->use_vols) return FLOF_OK;
    FLOF_REQUIRE((int)s->defos.size() < 9, FLOF_ERR_RANGE, "Too many defovolumes loaded!");
    int d[4];
    uni_size(fname, d);
    FLOF_REQUIRE(d[0] >= 2 && d[1] >= 2 && d[2] >= 2 && d[3] >= 1, FLOF_ERR_DIM,
                 std::string("Invalid src size from ") + fname);
    flof_grid4d *g = nullptr;
    int rc = flof_grid4d_create(ctx, d[0], d[1], d[2], d[3], FLOF_VEC4, FLOF_F32, &g);
    if (rc != FLOF_OK) return rc;
    rc = flof_grid4d_load(g, fname);
    if (rc != FLOF_OK) {
        flof_grid4d_destroy(g);
        return rc;
    }
    s->defos.push_back(g);
    API_CATCH(ctx)
}

int flof_warp_state_run(flof_warp_state *s, float *dst, int dx, int dy, int dz,
                        flof_grid4d *phi, float time, float blendAlpha, float loadTimeScale,
                        const float defoOffset[4], const float defoScale[4],
                        const float defoFactor[4], const float overrideSize[4],
                        float overrideTimeOff, int zeroVel, float thirdAlpha, int bordSkip,
                        float fourthAlpha, float defoAniFac)
{
    if (!s || s->defos.empty()) return FLOF_ERR_ARG;
    if (!s->use_vols)
        return flof_advect_time_slice_opt(s->defos[0], dst, dx, dy, dz, phi, time, blendAlpha,
                                          loadTimeScale, defoOffset, defoScale, defoFactor,
                                          overrideSize, overrideTimeOff, zeroVel, bordSkip,
                                          defoAniFac);
    return flof_advect_time_slice_vols(s->defos.data(), (int)s->defos.size(), dst, dx, dy, dz,
                                       phi, time, blendAlpha, loadTimeScale, defoOffset,
                                       defoScale, defoFactor, overrideSize, overrideTimeOff,
                                       zeroVel, bordSkip, defoAniFac, s->aligned ? 1 : 0,
                                       s->partial, thirdAlpha, fourthAlpha);
}

int flof_warp_state_close(flof_warp_state *s)
{
    if (!s) return FLOF_ERR_ARG;
    for (flof_grid4d *g : s->defos) flof_grid4d_destroy(g);
    delete s;
    return FLOF_OK;
}

int flof_load_vel_rescaled(flof_grid4d *vel, const char *fname, const int defoSize[4],
                           const float defoOffset[4], const float defoScale[4],
                           const float defoFactor[4])
{
    flof_ctx *ctx = vel ? vel->ctx : nullptr;
    API_TRY
    FLOF_REQUIRE(vel && fname && defoOffset && defoScale && defoFactor, FLOF_ERR_ARG,
                 "loadVelRescaled: null argument");
    FLOF_REQUIRE(vel->elem == FLOF_VEC4, FLOF_ERR_TYPE, "loadVelRescaled: vel must be VEC4");
    int d[4];
    if (defoSize) {
        for (int i = 0; i < 4; ++i) d[i] = defoSize[i];
    } else {
        uni_size(fname, d);
    }
    // the deformation grid of the defo solver's size; load() rejects a file
    // of another size, as readGrid4dUni does
    flof_grid4d *ipol = nullptr;
    int rc = flof_grid4d_create(ctx, d[0], d[1], d[2], d[3], FLOF_VEC4, FLOF_F32, &ipol);
    if (rc != FLOF_OK) return rc;
    const float whole[4] = {-1.f, -1.f, -1.f, -1.f};
    rc = flof_grid4d_load(ipol, fname);
    if (rc == FLOF_OK) rc = flof_interpolate_grid4d(vel, ipol, defoOffset, defoScale, whole);
    if (rc == FLOF_OK) rc = flof_grid4d_mult_const(vel, defoFactor);
    flof_grid4d_destroy(ipol);  // waits for the stream before freeing
    return rc;
    API_CATCH(ctx)
}

int flof_grid4d_save(flof_grid4d *g, const char *name)
{
    flof_ctx *ctx = g ? g->ctx : nullptr;
    API_TRY
    FLOF_REQUIRE(g && name && g->dtype == FLOF_F32, FLOF_ERR_ARG, "Grid4d::save: bad argument");
    std::vector<char> host(g->bytes);
    FLOF_HIP(hipMemcpyAsync(host.data(), g->data, g->bytes, hipMemcpyDeviceToHost,
                            ctx->stream));
    FLOF_HIP(hipStreamSynchronize(ctx->stream));
    uni_write4d(name, host.data(), g->X, g->Y, g->Z, g->T, g->elem);
    API_CATCH(ctx)
}

int flof_grid4d_load(flof_grid4d *g, const char *name)
{
    flof_ctx *ctx = g ? g->ctx : nullptr;
    API_TRY
    FLOF_REQUIRE(g && name && g->dtype == FLOF_F32, FLOF_ERR_ARG, "Grid4d::load: bad argument");
    std::vector<char> host(g->bytes);
    uni_read(name, host.data(), g->X, g->Y, g->Z, g->T, g->elem);
    FLOF_HIP(hipMemcpyAsync(g->data, host.data(), g->bytes, hipMemcpyHostToDevice,
                            ctx->stream));
    FLOF_HIP(hipStreamSynchronize(ctx->stream));
    API_CATCH(ctx)
}

}  // extern "C"
