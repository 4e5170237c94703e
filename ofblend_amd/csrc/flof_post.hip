// flof_post.hip -- output post-processing of warped 3D frames (the step after
// the warp in scenes/flof.py): level-set union (join), the smoothing of the
// positive side simpleBlurSpecial, boundary values, and the scalar grid
// operations the union blends and the temporal merge are composed of.
//
// Reference: source/levelset.cpp:114-118 (KnJoin), source/test.cpp:94-127
// (knSimpleBlurSpecial, simpleBlurSpecial), source/grid.cpp:462-473
// (knSetBoundary), source/grid.h:735-743 (gridMultScalar, gridScaledAdd),
// source/grid.cpp:198-245 (copyFrom, setConst), scenes/flof.py:737-782
// (union blends, temporal merge), scenes/ofHelpers.py:108-115 (output
// smoothing).  A frame is a float[Z][Y][X] device buffer; Z == 1 is a 2D
// grid (the reference's is3D() false).

#include <algorithm>

#include "flof_internal.h"

namespace flof {

__global__ void k_frame_op(int op, float *__restrict__ a, const float *__restrict__ b,
                           int64_t n, float s)
{
    FOR_XCD(c, n)
    {
        switch (op) {
        case FLOF_FRAME_COPY: a[c] = b[c]; break;
        case FLOF_FRAME_SET_CONST: a[c] = s; break;
        case FLOF_FRAME_MULT_CONST: a[c] = a[c] * s; break;
        case FLOF_FRAME_ADD_SCALED: a[c] = a[c] + s * b[c]; break;
        default: {  // FLOF_FRAME_JOIN: std::min(a, b) = b < a ? b : a
            const float x = a[c], y = b[c];
            a[c] = y < x ? y : x;
        }
        }
    }
}

// one pass of knSimpleBlurSpecial over the kernel range (bnd = 1: the outer
// layer is not written); dst keeps what it held there
__global__ void k_blur_special(const float *__restrict__ a, float *__restrict__ dst, int X,
                               int Y, int Z, float thresh, int bord)
{
    const bool is3d = Z > 1;
    const int nx = X - 2, ny = Y - 2, nz = is3d ? Z - 2 : 1;
    const int64_t n = (int64_t)nx * ny * nz;
    const int64_t sY = X, sZ = (int64_t)X * Y;
    FOR_XCD(q, n)
    {
        const int i = 1 + (int)(q % nx);
        const int j = 1 + (int)((q / nx) % ny);
        const int k = (is3d ? 1 : 0) + (int)(q / ((int64_t)nx * ny));
        const int64_t c = i + sY * j + sZ * k;
        bool in = i >= bord && j >= bord && i < X - bord && j < Y - bord;
        in = in && (is3d ? (k >= bord && k < Z - bord) : k == 0);
        const float ac = a[c];
        if (!in || ac < thresh) {
            dst[c] = ac;
            continue;
        }
        float v = a[c + 1] + a[c - 1] + a[c + sY] + a[c - sY];
        if (is3d) {
            v += a[c + sZ] + a[c - sZ];
            v = (float)((double)v * (1. / 6.));
        } else {
            v = (float)((double)v * (1. / 4.));
        }
        dst[c] = v;
    }
}

__global__ void k_frame_set_bound(float *__restrict__ a, int X, int Y, int Z, float value,
                                  int w)
{
    const int64_t n = (int64_t)X * Y * Z;
    FOR_XCD(c, n)
    {
        const int i = (int)(c % X), j = (int)((c / X) % Y), k = (int)(c / ((int64_t)X * Y));
        const bool bnd = i <= w || i >= X - 1 - w || j <= w || j >= Y - 1 - w ||
                         (Z > 1 && (k <= w || k >= Z - 1 - w));
        if (bnd) a[c] = value;
    }
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

int flof_frame_op(flof_ctx *ctx, int op, float *a, const float *b, int64_t n, float s)
{
    API_TRY
    FLOF_REQUIRE(ctx && a && n >= 0 && op >= FLOF_FRAME_COPY && op <= FLOF_FRAME_JOIN,
                 FLOF_ERR_ARG, "frame op: bad argument");
    FLOF_REQUIRE(b || op == FLOF_FRAME_SET_CONST || op == FLOF_FRAME_MULT_CONST, FLOF_ERR_ARG,
                 "frame op: second frame missing");
    if (n == 0) return FLOF_OK;
    hipLaunchKernelGGL(k_frame_op, dim3(grid_blocks_xcd(ctx, n)), dim3(256), 0, ctx->stream,
                       op, a, b, n, s);
    check_launch("k_frame_op");
    API_CATCH(ctx)
}

int flof_simple_blur_special(flof_ctx *ctx, float *a, int X, int Y, int Z, int iter,
                             float thresh, int bord)
{
    API_TRY
    FLOF_REQUIRE(ctx && a && X >= 3 && Y >= 3 && (Z == 1 || Z >= 3) && iter >= 0,
                 FLOF_ERR_ARG, "simpleBlurSpecial: bad grid");
    const int64_t n = (int64_t)X * Y * Z;
    FLOF_REQUIRE(n < (1ll << 31), FLOF_ERR_DIM, "simpleBlurSpecial: grid too large");
    if (iter == 0) return FLOF_OK;
    // the reference's temporary grid starts zeroed and the two swap every
    // pass, so the outer layer alternates between 0 and the input's values
    Tmp tmp(ctx, (size_t)n * 4);
    FLOF_HIP(hipMemsetAsync(tmp.p, 0, (size_t)n * 4, ctx->stream));
    float *cur = a, *oth = tmp.as<float>();
    const int64_t ni = (int64_t)(X - 2) * (Y - 2) * (Z > 1 ? Z - 2 : 1);
    for (int it = 0; it < iter; ++it) {
        hipLaunchKernelGGL(k_blur_special, dim3(grid_blocks_xcd(ctx, ni)), dim3(256), 0,
                           ctx->stream, (const float *)cur, oth, X, Y, Z, thresh, bord);
        check_launch("k_blur_special");
        std::swap(cur, oth);
    }
    if (cur != a)
        FLOF_HIP(hipMemcpyAsync(a, cur, (size_t)n * 4, hipMemcpyDeviceToDevice, ctx->stream));
    API_CATCH(ctx)
}

int flof_frame_set_bound(flof_ctx *ctx, float *a, int X, int Y, int Z, float value,
                         int width)
{
    API_TRY
    FLOF_REQUIRE(ctx && a && X >= 1 && Y >= 1 && Z >= 1, FLOF_ERR_ARG,
                 "setBound: bad grid");
    const int64_t n = (int64_t)X * Y * Z;
    FLOF_REQUIRE(n < (1ll << 31), FLOF_ERR_DIM, "setBound: grid too large");
    hipLaunchKernelGGL(k_frame_set_bound, dim3(grid_blocks_xcd(ctx, n)), dim3(256), 0,
                       ctx->stream, a, X, Y, Z, value, width);
    check_launch("k_frame_set_bound");
    API_CATCH(ctx)
}

}  // extern "C"
