// flof_core.hip -- context, Grid4d device storage, elementwise operators,
// reductions, resampling, advection, boundary conditions, error metric and
// the seeded drop generator of libflof_hip.
//
// Reference counterparts: source/grid4d.{h,cpp}, source/util/vector4d.h
// (interpolation), source/plugin/optflow4d.cpp (advection 1275-1321,
// resampling 40-57, calcLsDiff 895-927).

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <mutex>
#include <new>

#include "flof_internal.h"

namespace flof {

// ---------------------------------------------------------------------------
// pool / errors / timing

void *Pool::get(size_t bytes)
{
    auto it = free_.find(bytes);
    if (it != free_.end() && !it->second.empty()) {
        void *p = it->second.back();
        it->second.pop_back();
        cached -= bytes;
        return p;
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        trim();
        e = hipMalloc(&p, bytes);
    }
    if (e != hipSuccess)
        throw Error{FLOF_ERR_NOMEM,
                    "device allocation of " + std::to_string(bytes) +
                        " bytes failed: " + hipGetErrorString(e)};
    return p;
}

void Pool::put(void *p, size_t bytes)
{
    free_[bytes].push_back(p);
    cached += bytes;
}

void Pool::trim()
{
    for (auto &kv : free_)
        for (void *p : kv.second) (void)hipFree(p);
    free_.clear();
    cached = 0;
}

void set_error(flof_ctx *ctx, const Error &e)
{
    if (ctx) ctx->err = e.msg;
}

void check_launch(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        throw Error{FLOF_ERR_HIP,
                    std::string("launch of ") + what + ": " + hipGetErrorString(e)};
}

void *reduce_scratch(flof_ctx *ctx, size_t bytes)
{
    if (ctx->red_bytes < bytes) {
        if (ctx->red) {
            FLOF_HIP(hipStreamSynchronize(ctx->stream));
            FLOF_HIP(hipFree(ctx->red));
        }
        ctx->red = nullptr;
        ctx->red_bytes = 0;
        FLOF_HIP(hipMalloc(&ctx->red, bytes));
        FLOF_HIP(hipMemsetAsync(ctx->red, 0, bytes, ctx->stream));
        ctx->red_bytes = bytes;
    }
    return ctx->red;
}

// Layout of the reduction scratch: [0,64) counters (zeroed, reset by the
// last block of every reduction), [64,256) result slots, [256, ...) partials.
static const size_t RED_COUNTER = 0;
static const size_t RED_RESULT = 64;
static const size_t RED_PART = 256;

struct RedPtrs {
    unsigned int *counter;
    double *result;
    void *part;
};
static RedPtrs red_ptrs(flof_ctx *ctx, size_t part_bytes)
{
    char *base = (char *)reduce_scratch(ctx, RED_PART + part_bytes + 4096);
    return RedPtrs{(unsigned int *)(base + RED_COUNTER),
                   (double *)(base + RED_RESULT), base + RED_PART};
}

// ---------------------------------------------------------------------------
// elementwise kernels

enum EwOp { EW_SET, EW_ADDC, EW_MULC, EW_ADD, EW_SUB, EW_MULT, EW_ADDSCALED, EW_CLAMP, EW_COPY };

template <class T> struct Ew;
template <> struct Ew<float> {
    __device__ static float apply(int op, float a, float b, float c, float c2)
    {
        switch (op) {
        case EW_SET: return c;
        case EW_ADDC: return a + c;
        case EW_MULC: return a * c;
        case EW_ADD: return a + b;
        case EW_SUB: return a - b;
        case EW_MULT: return a * b;
        case EW_ADDSCALED: return a + c * b;
        case EW_CLAMP: return a < c ? c : (a > c2 ? c2 : a);
        default: return b;
        }
    }
};
template <> struct Ew<int> {
    __device__ static int apply(int op, int a, int b, int c, int c2)
    {
        switch (op) {
        case EW_SET: return c;
        case EW_ADDC: return a + c;
        case EW_MULC: return a * c;
        case EW_ADD: return a + b;
        case EW_SUB: return a - b;
        case EW_MULT: return a * b;
        case EW_ADDSCALED: return a + c * b;
        case EW_CLAMP: return a < c ? c : (a > c2 ? c2 : a);
        default: return b;
        }
    }
};
template <> struct Ew<float4> {
    __device__ static float4 apply(int op, float4 a, float4 b, float4 c, float4 c2)
    {
        float4 r;
        r.x = Ew<float>::apply(op, a.x, b.x, c.x, c2.x);
        r.y = Ew<float>::apply(op, a.y, b.y, c.y, c2.y);
        r.z = Ew<float>::apply(op, a.z, b.z, c.z, c2.z);
        r.w = Ew<float>::apply(op, a.w, b.w, c.w, c2.w);
        return r;
    }
};

template <class T>
__global__ void k_ew(T *__restrict__ a, const T *__restrict__ b, T c, T c2, int op,
                     int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        T bv = b ? b[i] : c;
        T av = (op == EW_SET || op == EW_COPY) ? c : a[i];
        a[i] = Ew<T>::apply(op, av, bv, c, c2);
    }
}

template <class T>
static void launch_ew(flof_ctx *ctx, void *a, const void *b, T c, T c2, int op,
                      int64_t n)
{
    hipLaunchKernelGGL(k_ew<T>, dim3(grid_blocks(ctx, n)), dim3(256), 0,
                       ctx->stream, (T *)a, (const T *)b, c, c2, op, n);
    check_launch("k_ew");
}

void launch_fill(flof_ctx *ctx, void *p, size_t bytes)
{
    // zero fill by kernel (usable under stream capture)
    launch_ew<float>(ctx, p, nullptr, 0.f, 0.f, EW_SET, (int64_t)(bytes / 4));
}

void launch_axpy4(flof_ctx *ctx, float4 *y, const float4 *x, float4 a, int64_t n)
{
    launch_ew<float4>(ctx, y, x, a, a, EW_ADDSCALED, n);
}
void launch_scale4(flof_ctx *ctx, float4 *y, float4 a, int64_t n)
{
    launch_ew<float4>(ctx, y, nullptr, a, a, EW_MULC, n);
}

// ---------------------------------------------------------------------------
// min / max reductions.  REAL: value; VEC4: squared norm (x*x+y*y+z*z+w*w in
// that order); INT: value.  Result slot: [min, max] as doubles.

template <class T> struct RedVal;
template <> struct RedVal<float> {
    __device__ static float get(float v) { return v; }
};
template <> struct RedVal<int> {
    __device__ static float get(int v) { return (float)v; }
};
template <> struct RedVal<float4> {
    __device__ static float get(float4 v) { return v.x * v.x + v.y * v.y + v.z * v.z + v.w * v.w; }
};

template <class T>
__global__ void k_minmax(const T *__restrict__ a, int64_t n, float *part)
{
    float mn = INFINITY, mx = -INFINITY;
    FOR_XCD(i, n)
    {
        const float v = RedVal<T>::get(a[i]);
        mn = fminf(mn, v);
        mx = fmaxf(mx, v);
    }
    mn = block_min(mn);
    mx = block_max(mx);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = mn;
        part[2 * blockIdx.x + 1] = mx;
    }
}
// one block: reduce the per-block partials (launched after the main kernel)
__global__ void k_minmax_fin(const float *part, int nb, double *result)
{
    float bmn = INFINITY, bmx = -INFINITY;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        bmn = fminf(bmn, part[2 * b]);
        bmx = fmaxf(bmx, part[2 * b + 1]);
    }
    bmn = block_min(bmn);
    bmx = block_max(bmx);
    if (threadIdx.x == 0) {
        result[0] = bmn;
        result[1] = bmx;
    }
}

// integer grids reduce exactly in int
__global__ void k_minmax_int(const int *__restrict__ a, int64_t n, int *part)
{
    __shared__ int smn[256], smx[256];
    int mn = 2147483647, mx = -2147483647 - 1;
    FOR_XCD(i, n)
    {
        mn = min(mn, a[i]);
        mx = max(mx, a[i]);
    }
    smn[threadIdx.x] = mn;
    smx[threadIdx.x] = mx;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            smn[threadIdx.x] = min(smn[threadIdx.x], smn[threadIdx.x + s]);
            smx[threadIdx.x] = max(smx[threadIdx.x], smx[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = smn[0];
        part[2 * blockIdx.x + 1] = smx[0];
    }
}
__global__ void k_minmax_int_fin(const int *part, int nb, double *result)
{
    if (threadIdx.x == 0) {
        int bmn = 2147483647, bmx = -2147483647 - 1;
        for (int b = 0; b < nb; ++b) {
            bmn = min(bmn, part[2 * b]);
            bmx = max(bmx, part[2 * b + 1]);
        }
        result[0] = bmn;
        result[1] = bmx;
    }
}

static void grid_minmax(flof_grid4d *g, double &mn, double &mx)
{
    flof_ctx *ctx = g->ctx;
    const int64_t n = g->n();
    const int nb = grid_blocks_xcd(ctx, n);
    RedPtrs r = red_ptrs(ctx, (size_t)nb * 16);
    if (g->elem == FLOF_REAL)
        hipLaunchKernelGGL(k_minmax<float>, dim3(nb), dim3(256), 0, ctx->stream,
                           (const float *)g->data, n, (float *)r.part);
    else if (g->elem == FLOF_VEC4)
        hipLaunchKernelGGL(k_minmax<float4>, dim3(nb), dim3(256), 0, ctx->stream,
                           (const float4 *)g->data, n, (float *)r.part);
    else
        hipLaunchKernelGGL(k_minmax_int, dim3(nb), dim3(256), 0, ctx->stream,
                           (const int *)g->data, n, (int *)r.part);
    check_launch("k_minmax");
    if (g->elem == FLOF_INT)
        hipLaunchKernelGGL(k_minmax_int_fin, dim3(1), dim3(256), 0, ctx->stream,
                           (const int *)r.part, nb, r.result);
    else
        hipLaunchKernelGGL(k_minmax_fin, dim3(1), dim3(256), 0, ctx->stream,
                           (const float *)r.part, nb, r.result);
    check_launch("k_minmax_fin");
    double h[2];
    FLOF_HIP(hipMemcpyAsync(h, r.result, 16, hipMemcpyDeviceToHost, ctx->stream));
    FLOF_HIP(hipStreamSynchronize(ctx->stream));
    mn = h[0];
    mx = h[1];
}

float host_max_norm_vec4(flof_ctx *ctx, const float4 *v, int64_t n)
{
    const int nb = grid_blocks_xcd(ctx, n);
    RedPtrs r = red_ptrs(ctx, (size_t)nb * 16);
    hipLaunchKernelGGL(k_minmax<float4>, dim3(nb), dim3(256), 0, ctx->stream, v, n,
                       (float *)r.part);
    check_launch("k_minmax");
    hipLaunchKernelGGL(k_minmax_fin, dim3(1), dim3(256), 0, ctx->stream,
                       (const float *)r.part, nb, r.result);
    check_launch("k_minmax_fin");
    double h[2];
    FLOF_HIP(hipMemcpyAsync(h, r.result, 16, hipMemcpyDeviceToHost, ctx->stream));
    FLOF_HIP(hipStreamSynchronize(ctx->stream));
    return std::sqrt((float)h[1]);
}

// ---------------------------------------------------------------------------
// boundary conditions (grid4d.cpp:355-412)

template <class T>
__global__ void k_set_bound(T *g, Dims4 D, Coord4 C, T value, int w)
{
    const int64_t n = (int64_t)D.st * D.T;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n;
         c += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k, t;
        decompose((uint32_t)c, C, i, j, k, t);
        const bool b = i <= w || i >= D.X - 1 - w || j <= w || j >= D.Y - 1 - w ||
                       k <= w || k >= D.Z - 1 - w || t <= w || t >= D.T - 1 - w;
        if (b) g[c] = value;
    }
}

template <class T> __global__ void k_neumann(T *g, Dims4 D, Coord4 C, int w)
{
    const int64_t n = (int64_t)D.st * D.T;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n;
         c += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k, t;
        decompose((uint32_t)c, C, i, j, k, t);
        bool set = false;
        int si = i, sj = j, sk = k, st = t;
        if (i <= w) { si = w + 1; set = true; }
        if (i >= D.X - 1 - w) { si = D.X - 2 - w; set = true; }
        if (j <= w) { sj = w + 1; set = true; }
        if (j >= D.Y - 1 - w) { sj = D.Y - 2 - w; set = true; }
        if (k <= w) { sk = w + 1; set = true; }
        if (k >= D.Z - 1 - w) { sk = D.Z - 2 - w; set = true; }
        if (t <= w) { st = w + 1; set = true; }
        if (t >= D.T - 1 - w) { st = D.T - 2 - w; set = true; }
        if (set) g[c] = g[(int64_t)si + (int64_t)D.X * sj + D.sz * sk + D.st * st];
    }
}

void launch_neumann(flof_ctx *ctx, void *g, const Dims4 &D, bool vec4, int w)
{
    // the copy sources must lie inside the grid and outside the band
    FLOF_REQUIRE(w >= 0 && D.X >= w + 2 && D.Y >= w + 2 && D.Z >= w + 2 &&
                     D.T >= w + 2,
                 FLOF_ERR_DIM, "setBoundNeumann: grid too small for width");
    const int64_t n = (int64_t)D.st * D.T;
    Coord4 C = make_coord4(D);
    if (vec4)
        hipLaunchKernelGGL(k_neumann<float4>, dim3(grid_blocks(ctx, n)), dim3(256),
                           0, ctx->stream, (float4 *)g, D, C, w);
    else
        hipLaunchKernelGGL(k_neumann<float>, dim3(grid_blocks(ctx, n)), dim3(256), 0,
                           ctx->stream, (float *)g, D, C, w);
    check_launch("k_neumann");
}

void launch_set_bound(flof_ctx *ctx, void *g, const Dims4 &D, bool vec4,
                      float4 value, int w)
{
    const int64_t n = (int64_t)D.st * D.T;
    Coord4 C = make_coord4(D);
    if (vec4)
        hipLaunchKernelGGL(k_set_bound<float4>, dim3(grid_blocks(ctx, n)), dim3(256),
                           0, ctx->stream, (float4 *)g, D, C, value, w);
    else
        hipLaunchKernelGGL(k_set_bound<float>, dim3(grid_blocks(ctx, n)), dim3(256),
                           0, ctx->stream, (float *)g, D, C, value.x, w);
    check_launch("k_set_bound");
}

// ---------------------------------------------------------------------------
// resampling: target(c) = interp4(source, c*fac + off) * post
// (grid4d.cpp:531-547, grid4d.h:463-471)

template <class T, class S = T>
__global__ void k_resample(S *__restrict__ dst, Dims4 Dd, Coord4 Cd,
                           const S *__restrict__ src, Dims4 Ds, float4 fac,
                           float4 off, float post, int use_post)
{
    const int64_t n = (int64_t)Dd.st * Dd.T;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n;
         c += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k, t;
        decompose((uint32_t)c, Cd, i, j, k, t);
        const float px = (float)i * fac.x + off.x;
        const float py = (float)j * fac.y + off.y;
        const float pz = (float)k * fac.z + off.z;
        const float pt = (float)t * fac.w + off.w;
        T v = interp4<T, S>(src, Ds, px, py, pz, pt);
        if (use_post) v = VOps<T>::muls(v, post);
        dst[c] = Stor<T, S>::st(v);
    }
}

void launch_resample(flof_ctx *ctx, void *dst, const Dims4 &Dd, const void *src,
                     const Dims4 &Ds, bool vec4, const float fac[4],
                     const float off[4])
{
    FLOF_REQUIRE(Ds.X >= 2 && Ds.Y >= 2 && Ds.Z >= 2 && Ds.T >= 2, FLOF_ERR_DIM,
                 "interpolation needs at least 2 cells per dimension");
    const int64_t n = (int64_t)Dd.st * Dd.T;
    Coord4 C = make_coord4(Dd);
    float4 f = make_float4(fac[0], fac[1], fac[2], fac[3]);
    float4 o = make_float4(off[0], off[1], off[2], off[3]);
    if (vec4)
        hipLaunchKernelGGL(k_resample<float4>, dim3(grid_blocks(ctx, n)), dim3(256),
                           0, ctx->stream, (float4 *)dst, Dd, C, (const float4 *)src,
                           Ds, f, o, 1.f, 0);
    else
        hipLaunchKernelGGL(k_resample<float>, dim3(grid_blocks(ctx, n)), dim3(256),
                           0, ctx->stream, (float *)dst, Dd, C, (const float *)src,
                           Ds, f, o, 1.f, 0);
    check_launch("k_resample");
}

// interpolGridTempl (optflow4d.cpp:40-57): factor = source/target size,
// offset = factor * 0.5; post_scale != 1 fuses the following multConst.
void launch_pyramid(flof_ctx *ctx, void *dst, const Dims4 &Dd, const void *src,
                    const Dims4 &Ds, bool vec4, float post_scale)
{
    FLOF_REQUIRE(Ds.X >= 2 && Ds.Y >= 2 && Ds.Z >= 2 && Ds.T >= 2, FLOF_ERR_DIM,
                 "interpolation needs at least 2 cells per dimension");
    const float fx = (float)Ds.X / (float)Dd.X, fy = (float)Ds.Y / (float)Dd.Y,
                fz = (float)Ds.Z / (float)Dd.Z, ft = (float)Ds.T / (float)Dd.T;
    float4 f = make_float4(fx, fy, fz, ft);
    float4 o = make_float4((float)(fx * 0.5), (float)(fy * 0.5), (float)(fz * 0.5),
                           (float)(ft * 0.5));
    const int64_t n = (int64_t)Dd.st * Dd.T;
    Coord4 C = make_coord4(Dd);
    const int use = post_scale != 1.f;
    if (vec4)
        hipLaunchKernelGGL(k_resample<float4>, dim3(grid_blocks(ctx, n)), dim3(256),
                           0, ctx->stream, (float4 *)dst, Dd, C, (const float4 *)src,
                           Ds, f, o, post_scale, use);
    else
        hipLaunchKernelGGL(k_resample<float>, dim3(grid_blocks(ctx, n)), dim3(256),
                           0, ctx->stream, (float *)dst, Dd, C, (const float *)src,
                           Ds, f, o, post_scale, use);
    check_launch("k_pyramid");
}

// the pyramid restriction of a bfloat16 image (FP32 interpolation of the
// stored values, rounded once when stored)
void launch_pyramid_bf16(flof_ctx *ctx, void *dst, const Dims4 &Dd, const void *src,
                         const Dims4 &Ds)
{
    FLOF_REQUIRE(Ds.X >= 2 && Ds.Y >= 2 && Ds.Z >= 2 && Ds.T >= 2, FLOF_ERR_DIM,
                 "interpolation needs at least 2 cells per dimension");
    const float fx = (float)Ds.X / (float)Dd.X, fy = (float)Ds.Y / (float)Dd.Y,
                fz = (float)Ds.Z / (float)Dd.Z, ft = (float)Ds.T / (float)Dd.T;
    float4 f = make_float4(fx, fy, fz, ft);
    float4 o = make_float4((float)(fx * 0.5), (float)(fy * 0.5), (float)(fz * 0.5),
                           (float)(ft * 0.5));
    const int64_t n = (int64_t)Dd.st * Dd.T;
    hipLaunchKernelGGL((k_resample<float, bf16_t>), dim3(grid_blocks(ctx, n)), dim3(256), 0,
                       ctx->stream, (bf16_t *)dst, Dd, make_coord4(Dd), (const bf16_t *)src,
                       Ds, f, o, 1.f, 0);
    check_launch("k_pyramid (bfloat16)");
}

__global__ void k_to_bf16(const float *__restrict__ src, bf16_t *__restrict__ dst, int64_t n)
{
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n;
         c += (int64_t)gridDim.x * blockDim.x)
        dst[c] = Stor<float, bf16_t>::st(src[c]);
}

void launch_to_bf16(flof_ctx *ctx, void *dst, const float *src, int64_t n)
{
    hipLaunchKernelGGL(k_to_bf16, dim3(grid_blocks(ctx, n)), dim3(256), 0, ctx->stream, src,
                       (bf16_t *)dst, n);
    check_launch("k_to_bf16");
}

void launch_neumann_bf16(flof_ctx *ctx, void *g, const Dims4 &D, int w)
{
    FLOF_REQUIRE(w >= 0 && D.X >= w + 2 && D.Y >= w + 2 && D.Z >= w + 2 &&
                     D.T >= w + 2,
                 FLOF_ERR_DIM, "setBoundNeumann: grid too small for width");
    const int64_t n = (int64_t)D.st * D.T;
    hipLaunchKernelGGL(k_neumann<bf16_t>, dim3(grid_blocks(ctx, n)), dim3(256), 0,
                       ctx->stream, (bf16_t *)g, D, make_coord4(D), w);
    check_launch("k_neumann (bfloat16)");
}

// ---------------------------------------------------------------------------
// semi-Lagrangian advection (optflow4d.cpp:1275-1290): interior cells sample
// the source at c + 1/2 - (vel*vel_scale)*dt, the outer layer becomes 0.

template <class T, class S = T>
__global__ void k_advect(S *__restrict__ dst, const S *__restrict__ src,
                         const float4 *__restrict__ vel, float4 vscale,
                         int use_vscale, Dims4 D, Coord4 C, float dt)
{
    const int64_t n = (int64_t)D.st * D.T;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n;
         c += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k, t;
        decompose((uint32_t)c, C, i, j, k, t);
        if (i < 1 || j < 1 || k < 1 || t < 1 || i >= D.X - 1 || j >= D.Y - 1 ||
            k >= D.Z - 1 || t >= D.T - 1) {
            dst[c] = Stor<T, S>::st(VOps<T>::zero());
            continue;
        }
        float4 v = vel[c];
        if (use_vscale) v = f4mul(v, vscale);
        v = f4muls(v, dt);
        const float px = ((float)i + 0.5f) - v.x;
        const float py = ((float)j + 0.5f) - v.y;
        const float pz = ((float)k + 0.5f) - v.z;
        const float pt = ((float)t + 0.5f) - v.w;
        dst[c] = Stor<T, S>::st(interp4<T, S>(src, D, px, py, pz, pt));
    }
}

void launch_advect(flof_ctx *ctx, const float4 *vel, float vel_scale, void *grid,
                   void *tmp, const Dims4 &D, bool vec4, float dt)
{
    FLOF_REQUIRE(D.X >= 2 && D.Y >= 2 && D.Z >= 2 && D.T >= 2, FLOF_ERR_DIM,
                 "advection needs at least 2 cells per dimension");
    const int64_t n = (int64_t)D.st * D.T;
    Coord4 C = make_coord4(D);
    float4 vs = make_float4(vel_scale, vel_scale, vel_scale, vel_scale);
    const int use = vel_scale != 1.f;
    if (vec4)
        klaunch(ctx, FLOF_KT_ADVECT, k_advect<float4>, dim3(grid_blocks(ctx, n)), dim3(256),
                (float4 *)tmp, (const float4 *)grid, vel, vs, use, D, C, dt);
    else
        klaunch(ctx, FLOF_KT_ADVECT, k_advect<float>, dim3(grid_blocks(ctx, n)), dim3(256),
                (float *)tmp, (const float *)grid, vel, vs, use, D, C, dt);
    check_launch("k_advect");
}

// advectCflTemplate<Grid4d> + advectCflHelper4d (optflow4d.cpp:1309-1321,
// 2170-2180): int(|velFactor*vel|max*dt/cfl)+1 substeps of dt/steps.
void advect_cfl(flof_ctx *ctx, const float4 *vel, void *grid, const Dims4 &D,
                bool vec4, float cfl, float velFactor, float dt)
{
    const int64_t n = (int64_t)D.st * D.T;
    const float4 *v = vel;
    Tmp scaled;
    if (velFactor != 1.f) {
        scaled = Tmp(ctx, (size_t)n * 16);
        FLOF_HIP(hipMemcpyAsync(scaled.p, vel, (size_t)n * 16,
                                hipMemcpyDeviceToDevice, ctx->stream));
        launch_scale4(ctx, scaled.as<float4>(),
                      make_float4(velFactor, velFactor, velFactor, velFactor), n);
        v = scaled.as<float4>();
    }
    const float maxVel = host_max_norm_vec4(ctx, v, n) * dt;
    const int steps = int(maxVel / cfl) + 1;
    const float sdt = dt / (float)steps;
    const size_t bytes = (size_t)n * (vec4 ? 16 : 4);
    Tmp tmp(ctx, bytes);
    for (int s = 0; s < steps; ++s) {
        launch_advect(ctx, v, 1.f, grid, tmp.p, D, vec4, sdt);
        FLOF_HIP(hipMemcpyAsync(grid, tmp.p, bytes, hipMemcpyDeviceToDevice,
                                ctx->stream));
    }
}

// the same for a bfloat16 image: every substep interpolates the stored
// values in FP32 and rounds its result once when it stores it
void advect_cfl_bf16(flof_ctx *ctx, const float4 *vel, void *grid, const Dims4 &D,
                     float cfl, float dt)
{
    FLOF_REQUIRE(D.X >= 2 && D.Y >= 2 && D.Z >= 2 && D.T >= 2, FLOF_ERR_DIM,
                 "advection needs at least 2 cells per dimension");
    const int64_t n = (int64_t)D.st * D.T;
    const float maxVel = host_max_norm_vec4(ctx, vel, n) * dt;
    const int steps = int(maxVel / cfl) + 1;
    const float sdt = dt / (float)steps;
    const size_t bytes = (size_t)n * 2;
    Tmp tmp(ctx, bytes);
    Coord4 C = make_coord4(D);
    const float4 one = make_float4(1.f, 1.f, 1.f, 1.f);
    for (int s = 0; s < steps; ++s) {
        klaunch(ctx, FLOF_KT_ADVECT, k_advect<float, bf16_t>, dim3(grid_blocks(ctx, n)),
                dim3(256), tmp.as<bf16_t>(), (const bf16_t *)grid, vel, one, 0, D, C, sdt);
        check_launch("k_advect (bfloat16)");
        FLOF_HIP(hipMemcpyAsync(grid, tmp.p, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    }
}

// ---------------------------------------------------------------------------
// calcLsDiff (optflow4d.cpp:895-927): correctly rounded FP64 sum of
// min(1, |a-b|*corr) over sign-mismatching cells of the bnd interior.

template <class S>
__global__ void k_lsdiff(const S *__restrict__ a, const S *__restrict__ b,
                         float *__restrict__ out, Dims4 D, Coord4 C, float corr,
                         int bnd, DD *part)
{
    const int64_t n = (int64_t)D.st * D.T;
    DD acc{0., 0.};
    FOR_XCD(c, n)
    {
        int i, j, k, t;
        decompose((uint32_t)c, C, i, j, k, t);
        if (i < bnd || j < bnd || k < bnd || t < bnd || i >= D.X - bnd ||
            j >= D.Y - bnd || k >= D.Z - bnd || t >= D.T - bnd)
            continue;
        const float va = Stor<float, S>::ld(a[c]), vb = Stor<float, S>::ld(b[c]);
        if ((va < 0.f && vb < 0.f) || (va >= 0.f && vb >= 0.f)) {
            if (out) out[c] = 0.f;
        } else {
            float d = fabsf(va - vb) * corr;
            if (d > 1.f) d = 1.f;
            dd_add(acc, (double)d);
            if (out) out[c] = d;
        }
    }
    acc = block_dd(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ void k_dd_fin(const DD *part, int nb, double *result)
{
    DD r{0., 0.};
    for (int bb = threadIdx.x; bb < nb; bb += blockDim.x) dd_merge(r, part[bb]);
    r = block_dd(r);
    if (threadIdx.x == 0) result[0] = r.hi + r.lo;
}

// Enqueues the sum into *dst_dev (device double); the caller scales it.
static void enqueue_ls_sum(flof_ctx *ctx, const float *a, const float *b, float *out,
                           const Dims4 &D, float corr, int bnd, double *dst_dev)
{
    const int64_t n = (int64_t)D.st * D.T;
    const int nb = grid_blocks_xcd(ctx, n);
    RedPtrs r = red_ptrs(ctx, (size_t)nb * sizeof(DD));
    hipLaunchKernelGGL(k_lsdiff<float>, dim3(nb), dim3(256), 0, ctx->stream, a, b, out, D,
                       make_coord4(D), corr, bnd, (DD *)r.part);
    check_launch("k_lsdiff");
    hipLaunchKernelGGL(k_dd_fin, dim3(1), dim3(256), 0, ctx->stream, (const DD *)r.part, nb,
                       dst_dev ? dst_dev : r.result);
    check_launch("k_dd_fin");
}

static double ls_scale(double accu, const Dims4 &D, int bnd)
{
    const int64_t sx = D.X - 2 * bnd, sy = D.Y - 2 * bnd, sz = D.Z - 2 * bnd,
                  st = D.T - 2 * bnd;
    accu *= 1000.;
    if (D.T > 1) accu *= 1000.;
    accu *= 1. / double(sx * sy * sz * st);
    return accu;
}

double host_ls_diff_bf16(flof_ctx *ctx, const void *a, const void *b, const Dims4 &D,
                         float corr, int bnd)
{
    const int64_t n = (int64_t)D.st * D.T;
    const int nb = grid_blocks_xcd(ctx, n);
    RedPtrs r = red_ptrs(ctx, (size_t)nb * sizeof(DD));
    hipLaunchKernelGGL(k_lsdiff<bf16_t>, dim3(nb), dim3(256), 0, ctx->stream,
                       (const bf16_t *)a, (const bf16_t *)b, nullptr, D, make_coord4(D), corr,
                       bnd, (DD *)r.part);
    check_launch("k_lsdiff (bfloat16)");
    hipLaunchKernelGGL(k_dd_fin, dim3(1), dim3(256), 0, ctx->stream, (const DD *)r.part, nb,
                       r.result);
    check_launch("k_dd_fin");
    double h;
    FLOF_HIP(hipMemcpyAsync(&h, r.result, 8, hipMemcpyDeviceToHost, ctx->stream));
    FLOF_HIP(hipStreamSynchronize(ctx->stream));
    return (double)(float)ls_scale(h, D, bnd);
}

double host_ls_diff(flof_ctx *ctx, const float *a, const float *b, float *out,
                    const Dims4 &D, float corr, int bnd)
{
    enqueue_ls_sum(ctx, a, b, out, D, corr, bnd, nullptr);
    RedPtrs r = red_ptrs(ctx, 0);  // after the launch: scratch may have grown
    double h;
    FLOF_HIP(hipMemcpyAsync(&h, r.result, 8, hipMemcpyDeviceToHost, ctx->stream));
    FLOF_HIP(hipStreamSynchronize(ctx->stream));
    return (double)(float)ls_scale(h, D, bnd);
}

// ---------------------------------------------------------------------------
// region / slices / max diff (grid4d.cpp:419-524)

template <class T>
__global__ void k_set_region(T *g, Dims4 D, Coord4 C, float4 s, float4 e, T value)
{
    const int64_t n = (int64_t)D.st * D.T;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n;
         c += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k, t;
        decompose((uint32_t)c, C, i, j, k, t);
        const float p[4] = {(float)i, (float)j, (float)k, (float)t};
        const float ss[4] = {s.x, s.y, s.z, s.w}, ee[4] = {e.x, e.y, e.z, e.w};
        bool in = true;
        for (int q = 0; q < 4; ++q)
            if (p[q] < ss[q] || p[q] > ee[q]) in = false;
        if (in) g[c] = value;
    }
}

template <class T>
__global__ void k_maxdiff(const T *a, const T *b, int64_t n, double *part)
{
    // per cell: |a-b| of the FP32 difference (REAL) or the sum of the 4
    // component differences formed in double (VEC4), as grid4d.cpp:419-464
    double mx = 0.;
    FOR_XCD(i, n)
    {
        double d;
        if constexpr (sizeof(T) == 16) {
            const float4 x = ((const float4 *)a)[i], y = ((const float4 *)b)[i];
            d = fabs((double)x.x - (double)y.x) + fabs((double)x.y - (double)y.y) +
                fabs((double)x.z - (double)y.z) + fabs((double)x.w - (double)y.w);
        } else {
            d = (double)fabsf(((const float *)a)[i] - ((const float *)b)[i]);
        }
        mx = fmax(mx, d);
    }
    __shared__ double sh[256];
    sh[threadIdx.x] = mx;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}
__global__ void k_maxdiff_fin(const double *part, int nb, double *result)
{
    if (threadIdx.x == 0) {
        double m = 0.;
        for (int bb = 0; bb < nb; ++bb) m = fmax(m, part[bb]);
        result[0] = (double)(float)m;  // the reference returns Real
    }
}

__global__ void k_get_slice(const float *src, Dims4 D, int srct, float *dst, int dx,
                            int dy, int dz)
{
    const int64_t n = D.st;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n;
         c += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(c % D.X), j = (int)((c / D.X) % D.Y), k = (int)(c / D.sz);
        if (i < dx && j < dy && k < dz)
            dst[i + (int64_t)dx * j + (int64_t)dx * dy * k] = src[D.st * srct + c];
    }
}

__global__ void k_place3d(const float *src, int sx, int sy, int sz, float *dst,
                          Dims4 D, int dstt)
{
    const int64_t n = (int64_t)sx * sy * sz;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n;
         c += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(c % sx), j = (int)((c / sx) % sy),
                  k = (int)(c / ((int64_t)sx * sy));
        if (i < D.X && j < D.Y && k < D.Z)
            dst[i + (int64_t)D.X * j + D.sz * k + D.st * dstt] = src[c];
    }
}

// ---------------------------------------------------------------------------
// seeded drop generator (no reference counterpart)

static inline uint64_t splitmix64(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static inline double u01(uint64_t &s)
{
    return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
}

struct Drops {
    int n;
    double cx[8], y0[8], cz[8], r[8];
};

static Drops make_drops(int X, int Y, int Z, uint64_t seed, int nspheres,
                        double rmin, double rmax)
{
    Drops d;
    uint64_t s = seed;
    d.n = std::min(nspheres, 8);
    for (int i = 0; i < d.n; ++i) {
        d.cx[i] = X * (0.3 + 0.4 * u01(s));
        d.y0[i] = Y * (0.6 + 0.2 * u01(s));
        d.cz[i] = Z * (0.3 + 0.4 * u01(s));
        d.r[i] = (rmax > rmin) ? X * (rmin + (rmax - rmin) * u01(s)) : X * rmin;
    }
    return d;
}

__global__ void k_drops(float *out, int fx, int fy, int fz, int T, int Y, double F,
                        double cl, Drops d)
{
    const int64_t n = (int64_t)fx * fy * fz * T;
    const double g = 0.4 * Y;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n;
         c += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(c % fx);
        const int j = (int)((c / fx) % fy);
        const int k = (int)((c / ((int64_t)fx * fy)) % fz);
        const int t = (int)(c / ((int64_t)fx * fy * fz));
        const double s = T > 1 ? (double)t / (double)(T - 1) : 0.0;
        const double px = i + 0.5, py = j + 0.5, pz = k + 0.5;
        double v = py - 0.2 * Y * F;
        for (int q = 0; q < d.n; ++q) {
            const double cy = d.y0[q] - 0.5 * g * s * s;
            const double ddx = px - F * d.cx[q], ddy = py - F * cy, ddz = pz - F * d.cz[q];
            const double sd = sqrt(ddx * ddx + ddy * ddy + ddz * ddz) - F * d.r[q];
            if (sd < v) v = sd;
        }
        if (v > cl) v = cl;
        if (v < -cl) v = -cl;
        out[c] = (float)v;
    }
}

// ---------------------------------------------------------------------------
// Allocation-free calcLsDiff.  Its workgroup partials live in `scratch`, a
// file-scope device array of NOALLOC_PARTS double-doubles (32 KiB) that every
// GPU the library runs on holds until the process ends.  Host code never
// passes the array's name to a kernel (on the host it is a stand-in symbol,
// not a device address): noalloc_scratch() looks up each device's own
// address once with hipGetSymbolAddress, while that device is current.
// Every workgroup writes its own partial before the finalizer reads it, so
// nothing needs clearing; two calls that overlap on one GPU (different
// streams) share the array and race.
static const int NOALLOC_PARTS = 2048;
static __device__ __align__(256) DD scratch[NOALLOC_PARTS];
static const int NOALLOC_MAX_DEVICES = 64;

struct NoallocDev {
    DD *part = nullptr;
    int num_cu = 0;
};
static std::mutex g_noalloc_mu;
static NoallocDev g_noalloc[NOALLOC_MAX_DEVICES];

static NoallocDev noalloc_scratch()
{
    int dev = -1;
    FLOF_HIP(hipGetDevice(&dev));
    FLOF_REQUIRE(dev >= 0 && dev < NOALLOC_MAX_DEVICES, FLOF_ERR_RANGE,
                 "calcLsDiff4d (no allocation): device index above 63");
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    NoallocDev &d = g_noalloc[dev];
    if (!d.part) {
        void *p = nullptr;
        FLOF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        FLOF_REQUIRE(p != nullptr, FLOF_ERR_HIP,
                     "calcLsDiff4d (no allocation): no device address for the scratch array");
        int cu = 0;
        FLOF_HIP(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev));
        d.num_cu = cu > 0 ? cu : 256;
        d.part = (DD *)p;
    }
    return d;
}

// finalizer of the allocation-free form: the sum of the partials scaled as
// ls_scale() does on the host, stored as the reference's Real
__global__ void k_dd_fin_ls(const DD *part, int nb, int four_d, double inv_cells,
                            float *result)
{
    DD r{0., 0.};
    for (int bb = threadIdx.x; bb < nb; bb += blockDim.x) dd_merge(r, part[bb]);
    r = block_dd(r);
    if (threadIdx.x == 0) {
        double accu = r.hi + r.lo;
        accu *= 1000.;
        if (four_d) accu *= 1000.;
        accu *= inv_cells;
        result[0] = (float)accu;
    }
}

static void noalloc_ls_diff(const float *a, const float *b, float *out, const Dims4 &D,
                            float corr, int bnd, hipStream_t stream, float *result_dev)
{
    const NoallocDev nd = noalloc_scratch();
    const int64_t n = (int64_t)D.st * D.T;
    // the workgroup count of the context form (8 per CU, a multiple of 8)
    int64_t nb = (n + 255) / 256;
    const int64_t cap = (int64_t)nd.num_cu * 8;
    if (nb > cap) nb = cap;
    if (nb < 1) nb = 1;
    nb = ((nb + FLOF_NXCD - 1) / FLOF_NXCD) * FLOF_NXCD;
    FLOF_REQUIRE(nb <= NOALLOC_PARTS, FLOF_ERR_RANGE,
                 "calcLsDiff4d (no allocation): more workgroups than scratch partials");
    const int64_t sx = D.X - 2 * bnd, sy = D.Y - 2 * bnd, sz = D.Z - 2 * bnd,
                  st = D.T - 2 * bnd;
    const double inv_cells = 1. / double(sx * sy * sz * st);
    hipLaunchKernelGGL(k_lsdiff<float>, dim3((unsigned)nb), dim3(256), 0, stream, a, b, out,
                       D, make_coord4(D), corr, bnd, nd.part);
    check_launch("k_lsdiff (no allocation)");
    hipLaunchKernelGGL(k_dd_fin_ls, dim3(1), dim3(256), 0, stream, (const DD *)nd.part,
                       (int)nb, D.T > 1 ? 1 : 0, inv_cells, result_dev);
    check_launch("k_dd_fin_ls");
}

// message of the last failed call that has no context, per host thread
thread_local std::string g_noctx_err;

}  // namespace flof

// ===========================================================================
// C-ABI

using namespace flof;

#define API_TRY try {
#define API_CATCH(ctx)                                                        \
    }                                                                         \
    catch (const flof::Error &e)                                              \
    {                                                                         \
        flof::set_error(ctx, e);                                              \
        return e.code;                                                        \
    }                                                                         \
    catch (const std::bad_alloc &)                                            \
    {                                                                         \
        flof::set_error(ctx, flof::Error{FLOF_ERR_NOMEM, "host allocation failed"}); \
        return FLOF_ERR_NOMEM;                                                \
    }                                                                         \
    catch (const std::exception &e)                                           \
    {                                                                         \
        flof::set_error(ctx, flof::Error{FLOF_ERR_ARG, e.what()});            \
        return FLOF_ERR_ARG;                                                  \
    }                                                                         \
    return FLOF_OK;

static void need_same(const flof_grid4d *a, const flof_grid4d *b, const char *what)
{
    FLOF_REQUIRE(a && b, FLOF_ERR_ARG, std::string(what) + ": null grid");
    FLOF_REQUIRE(a->X == b->X && a->Y == b->Y && a->Z == b->Z && a->T == b->T,
                 FLOF_ERR_DIM, std::string(what) + ": different Grid4d resolutions");
    FLOF_REQUIRE(a->elem == b->elem && a->dtype == b->dtype, FLOF_ERR_TYPE,
                 std::string(what) + ": different element types");
}
static void need_f32(const flof_grid4d *g, const char *what)
{
    FLOF_REQUIRE(g, FLOF_ERR_ARG, std::string(what) + ": null grid");
    FLOF_REQUIRE(g->dtype == FLOF_F32, FLOF_ERR_TYPE,
                 std::string(what) + ": FP32 storage required");
}

extern "C" {

void flof_of_params_default(flof_of_params *p)
{
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->wSmooth = 0.;
    p->wEnergy = 0.;
    p->level = 0;
    p->postVelBlur = 0.;
    p->cgAccuracy = 1e-4;
    p->blurType = 1;
    p->cfl = 999.;
    p->orderTime = 1;
    p->orderSpace = 1;
    p->resetBndWidth = -1.;
    p->multiStep = 1;
    p->projSizeThresh = 9999;
    p->minGridSize = 10;
    p->doFinalProject = 0;
    p->cgMaxIter = 1000;
    p->warmStart = 0;
}

size_t flof_abi_struct_size(int which)
{
    switch (which) {
    case 0: return sizeof(flof_of_params);
    case 1: return sizeof(flof_solve_info);
    default: return 0;
    }
}

int flof_ctx_create(flof_ctx **out, int device)
{
    flof_ctx *ctx = nullptr;
    API_TRY
    FLOF_REQUIRE(out, FLOF_ERR_ARG, "flof_ctx_create: null out");
    *out = nullptr;
    ctx = new flof_ctx();
    ctx->device = device;
    FLOF_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    FLOF_HIP(hipGetDeviceProperties(&prop, device));
    ctx->num_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    FLOF_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    FLOF_HIP(hipHostMalloc(&ctx->host_scratch, 4096, hipHostMallocDefault));
    {
        const char *ng = std::getenv("FLOF_NO_GRAPH");
        ctx->use_graphs = !(ng && ng[0] == '1');
        const char *bp = std::getenv("FLOF_BLOCKS_PER_CU");
        if (bp && std::atoi(bp) > 0 && std::atoi(bp) <= 64) ctx->blocks_per_cu = std::atoi(bp);
        const char *nt = std::getenv("FLOF_NO_TMARCH");
        ctx->use_tmarch = !(nt && nt[0] == '1');
        const char *ls = std::getenv("FLOF_LDS_STENCIL");
        if (ls && ls[0] == '0') ctx->stencil = FLOF_STENCIL_TMARCH;
        if (ls && ls[0] == '1') ctx->stencil = FLOF_STENCIL_LDS;
        const char *lx = std::getenv("FLOF_LDS_XCD");
        if (lx) ctx->lds_xcd = lx[0] == '1';
        const char *lb = std::getenv("FLOF_LDS_BPC");
        if (lb && std::atoi(lb) > 0 && std::atoi(lb) <= 64) ctx->lds_blocks_per_cu = std::atoi(lb);
        const char *tb = std::getenv("FLOF_TM_BPC");
        if (tb && std::atoi(tb) > 0 && std::atoi(tb) <= 64) ctx->tm_blocks_per_cu = std::atoi(tb);
    }
    *out = ctx;
    }
    catch (const flof::Error &e)
    {
        if (ctx) {
            if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
            delete ctx;
        }
        return e.code;
    }
    return FLOF_OK;
}

int flof_ctx_destroy(flof_ctx *ctx)
{
    if (!ctx) return FLOF_ERR_ARG;
    for (flof_ctx *w : ctx->workers) flof_ctx_destroy(w);
    ctx->workers.clear();
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &t : ctx->timers)
        for (auto &ev : t.pending) {
            (void)hipEventDestroy(ev.e0);
            (void)hipEventDestroy(ev.e1);
        }
    ctx->pool.trim();
    if (ctx->red) (void)hipFree(ctx->red);
    if (ctx->host_scratch) (void)hipHostFree(ctx->host_scratch);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return FLOF_OK;
}

int flof_ctx_set_autotune(flof_ctx *ctx, int enable)
{
    if (!ctx) return FLOF_ERR_ARG;
    ctx->autotune = enable != 0;
    for (flof_ctx *w : ctx->workers) w->autotune = ctx->autotune;
    return FLOF_OK;
}

int flof_ctx_set_stencil(flof_ctx *ctx, int kind)
{
    if (!ctx || kind < FLOF_STENCIL_AUTO || kind > FLOF_STENCIL_QARRAY) return FLOF_ERR_ARG;
    const bool tm = kind != FLOF_STENCIL_QARRAY;
    ctx->stencil = tm ? kind : FLOF_STENCIL_AUTO;
    ctx->use_tmarch = tm;
    for (flof_ctx *w : ctx->workers) {
        w->stencil = ctx->stencil;
        w->use_tmarch = tm;
    }
    return FLOF_OK;
}

int flof_ctx_tuned_blocks(const flof_ctx *ctx, int X, int Y, int Z, int T)
{
    if (!ctx) return FLOF_ERR_ARG;
    auto it = ctx->tm_tuned.find(
        std::array<int, 5>{X, Y, Z, T, tune_variant(ctx)});
    return it == ctx->tm_tuned.end() ? 0 : it->second.bpc;
}

int flof_ctx_tuned_shape(const flof_ctx *ctx, int X, int Y, int Z, int T, int *tm_bpc,
                         int *lds_bpc, int *lds_seglen)
{
    if (!ctx) return FLOF_ERR_ARG;
    auto it = ctx->tm_tuned.find(
        std::array<int, 5>{X, Y, Z, T, tune_variant(ctx)});
    if (it == ctx->tm_tuned.end()) return 1;
    if (tm_bpc) *tm_bpc = it->second.bpc;
    if (lds_bpc) *lds_bpc = it->second.lds_bpc;
    if (lds_seglen) {
        *lds_seglen = 0;
        if (it->second.lds_bpc > 0) {
            const Dims4 D{X, Y, Z, T, (int64_t)X * Y, (int64_t)X * Y * Z};
            *lds_seglen = flof::lds_seglen(ctx, D, it->second.lds_bpc);
        }
    }
    return FLOF_OK;
}

const char *flof_last_error(const flof_ctx *ctx)
{
    if (ctx) return ctx->err.c_str();
    return flof::g_noctx_err.empty() ? "null context" : flof::g_noctx_err.c_str();
}

int flof_sync(flof_ctx *ctx)
{
    API_TRY
    FLOF_REQUIRE(ctx, FLOF_ERR_ARG, "null context");
    FLOF_HIP(hipStreamSynchronize(ctx->stream));
    API_CATCH(ctx)
}

void *flof_ctx_stream(flof_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int flof_ctx_trim(flof_ctx *ctx)
{
    API_TRY
    FLOF_REQUIRE(ctx, FLOF_ERR_ARG, "null context");
    FLOF_HIP(hipStreamSynchronize(ctx->stream));
    ctx->pool.trim();
    API_CATCH(ctx)
}

int flof_ctx_enable_timing(flof_ctx *ctx, int kind, int enable)
{
    API_TRY
    FLOF_REQUIRE(ctx && kind >= 0 && kind < FLOF_KT_COUNT, FLOF_ERR_ARG,
                 "flof_ctx_enable_timing: bad kind");
    ctx->timers[kind].on = enable != 0;
    API_CATCH(ctx)
}

int flof_ctx_timing_tag(flof_ctx *ctx, int kind, int64_t tag, int64_t *launches,
                        double *ms_total)
{
    API_TRY
    FLOF_REQUIRE(ctx && kind >= 0 && kind < FLOF_KT_COUNT, FLOF_ERR_ARG,
                 "flof_ctx_timing: bad kind");
    FLOF_HIP(hipStreamSynchronize(ctx->stream));
    int64_t n = 0;
    double ms_sum = 0.0;
    for (auto &ev : ctx->timers[kind].pending) {
        if (tag >= 0 && ev.tag != tag) continue;
        float ms = 0.f;
        FLOF_HIP(hipEventElapsedTime(&ms, ev.e0, ev.e1));
        ms_sum += ms;
        n += 1;
    }
    if (launches) *launches = n;
    if (ms_total) *ms_total = ms_sum;
    API_CATCH(ctx)
}

int flof_ctx_timing(flof_ctx *ctx, int kind, int64_t *launches, double *ms_total)
{
    const int rc = flof_ctx_timing_tag(ctx, kind, -1, launches, ms_total);
    if (rc != FLOF_OK) return rc;
    KTimer &t = ctx->timers[kind];
    for (auto &ev : t.pending) {
        (void)hipEventDestroy(ev.e0);
        (void)hipEventDestroy(ev.e1);
    }
    t.pending.clear();
    return FLOF_OK;
}

// ---------------------------------------------------------------------------
// Streaming kernels with the byte mixes of the PCG kernels and no stencil
// work: the bandwidth HBM sustains for a mix (reads and writes interleave
// differently from a copy), against which the PCG kernels are compared.
namespace flof {
namespace {
typedef float f4s __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 nt_ld(const float4 *a)
{
    const f4s v = __builtin_nontemporal_load((const f4s *)a);
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void nt_st(float4 *a, const float4 &v)
{
    f4s w = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(w, (f4s *)a);
}
// update mix: read x, r, p, g; write x, r (84 B per cell)
__global__ void __launch_bounds__(256)
k_stream_update(float4 *__restrict__ x, float4 *__restrict__ r, const float4 *__restrict__ p,
                const float *__restrict__ g, int64_t n, float a)
{
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n;
         c += (int64_t)gridDim.x * blockDim.x) {
        const float4 pc = nt_ld(p + c);
        const float gc = __builtin_nontemporal_load(g + c);
        float4 xc = nt_ld(x + c), rc = nt_ld(r + c);
        xc = make_float4(xc.x + a * pc.x, xc.y + a * pc.y, xc.z + a * pc.z, xc.w + a * pc.w);
        rc = make_float4(rc.x - gc * pc.x, rc.y - gc * pc.y, rc.z - gc * pc.z, rc.w - gc * pc.w);
        nt_st(x + c, xc);
        nt_st(r + c, rc);
    }
}
// direction mix: read r, p, g; write p (52 B per cell)
__global__ void __launch_bounds__(256)
k_stream_dir(float4 *__restrict__ p, const float4 *__restrict__ r, const float *__restrict__ g,
             int64_t n, float b)
{
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n;
         c += (int64_t)gridDim.x * blockDim.x) {
        const float4 rc = nt_ld(r + c), pc = nt_ld(p + c);
        const float gc = __builtin_nontemporal_load(g + c);
        nt_st(p + c, make_float4(rc.x * gc + b * pc.x, rc.y * gc + b * pc.y,
                                 rc.z * gc + b * pc.z, rc.w * gc + b * pc.w));
    }
}
// direction mix of odd iterations: read x, r, p, p_prev, g; write x, p
// (100 B per cell)
__global__ void __launch_bounds__(256)
k_stream_dirx(float4 *__restrict__ x, float4 *__restrict__ p, const float4 *__restrict__ r,
              const float4 *__restrict__ pc, const float *__restrict__ g, int64_t n, float b)
{
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n;
         c += (int64_t)gridDim.x * blockDim.x) {
        const float4 rc = nt_ld(r + c), pk = nt_ld(pc + c), pp = nt_ld(p + c);
        const float4 xc = nt_ld(x + c);
        const float gc = __builtin_nontemporal_load(g + c);
        nt_st(x + c, make_float4(xc.x + b * pp.x + b * pk.x, xc.y + b * pp.y + b * pk.y,
                                 xc.z + b * pp.z + b * pk.z, xc.w + b * pp.w + b * pk.w));
        nt_st(p + c, make_float4(rc.x * gc + b * pk.x, rc.y * gc + b * pk.y,
                                 rc.z * gc + b * pk.z, rc.w * gc + b * pk.w));
    }
}
// bitwise comparison of rcp_newton(d) with the IEEE division over bit patterns
__global__ void k_check_rcp(uint32_t lo, uint32_t hi, unsigned long long *count)
{
    const uint64_t n = (uint64_t)hi - lo + 1;
    unsigned long long bad = 0;
    for (uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < n;
         q += (uint64_t)gridDim.x * blockDim.x) {
        const float d = __uint_as_float(lo + (uint32_t)q);
        bad += __float_as_uint(rcp_newton(d)) != __float_as_uint(1.0f / d);
    }
    if (bad) atomicAdd(count, bad);
}
// copy: read a, write b (32 B per cell)
__global__ void __launch_bounds__(256)
k_stream_copy(float4 *__restrict__ b, const float4 *__restrict__ a, int64_t n)
{
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n;
         c += (int64_t)gridDim.x * blockDim.x)
        nt_st(b + c, nt_ld(a + c));
}
}  // namespace
}  // namespace flof

int flof_stream_bandwidth(flof_ctx *ctx, int mix, int64_t cells, int reps, double *gbps)
{
    API_TRY
    FLOF_REQUIRE(ctx && mix >= 0 && mix <= 3 && cells > 0 && reps > 0 && gbps, FLOF_ERR_ARG,
                 "flof_stream_bandwidth: bad argument");
    const size_t v4 = (size_t)cells * 16;
    Tmp a(ctx, v4), b(ctx, v4), c(ctx, v4), d(ctx, mix == FLOF_MIX_DIR_X ? v4 : 16),
        g(ctx, (size_t)cells * 4);
    FLOF_HIP(hipMemsetAsync(a.p, 0, v4, ctx->stream));
    FLOF_HIP(hipMemsetAsync(b.p, 0, v4, ctx->stream));
    FLOF_HIP(hipMemsetAsync(c.p, 0, v4, ctx->stream));
    if (mix == FLOF_MIX_DIR_X) FLOF_HIP(hipMemsetAsync(d.p, 0, v4, ctx->stream));
    FLOF_HIP(hipMemsetAsync(g.p, 0, (size_t)cells * 4, ctx->stream));
    const double per_cell = mix == FLOF_MIX_UPDATE ? 84.0
                            : mix == FLOF_MIX_DIR  ? 52.0
                            : mix == FLOF_MIX_COPY ? 32.0
                                                   : 100.0;
    const double bytes = per_cell * (double)cells;
    hipEvent_t e0, e1;
    FLOF_HIP(hipEventCreate(&e0));
    FLOF_HIP(hipEventCreate(&e1));
    double best = 0.0;
    hipError_t err = hipSuccess;
    // the best over the workgroups per CU: which count HBM serves best
    // depends on the mix (2 for the update's, measured on MI355X)
    for (int bpc : {1, 2, 4, 8, 16}) {
        const int grid = ctx->num_cu * bpc;
        auto launch = [&]() {
            if (mix == 0)
                hipLaunchKernelGGL(k_stream_update, dim3(grid), dim3(256), 0, ctx->stream,
                                   a.as<float4>(), b.as<float4>(), (const float4 *)c.as<float4>(),
                                   (const float *)g.as<float>(), cells, 0.5f);
            else if (mix == 1)
                hipLaunchKernelGGL(k_stream_dir, dim3(grid), dim3(256), 0, ctx->stream,
                                   a.as<float4>(), (const float4 *)b.as<float4>(),
                                   (const float *)g.as<float>(), cells, 0.5f);
            else if (mix == FLOF_MIX_DIR_X)
                hipLaunchKernelGGL(k_stream_dirx, dim3(grid), dim3(256), 0, ctx->stream,
                                   a.as<float4>(), b.as<float4>(), (const float4 *)c.as<float4>(),
                                   (const float4 *)d.as<float4>(),
                                   (const float *)g.as<float>(), cells, 0.5f);
            else
                hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, ctx->stream,
                                   a.as<float4>(), (const float4 *)b.as<float4>(), cells);
        };
        launch();  // warm-up
        launch();
        err = hipEventRecord(e0, ctx->stream);
        for (int k = 0; k < reps && err == hipSuccess; ++k) launch();
        if (err == hipSuccess) err = hipEventRecord(e1, ctx->stream);
        if (err == hipSuccess) err = hipEventSynchronize(e1);
        float ms = 0.f;
        if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
        if (err != hipSuccess) break;
        if (ms > 0.f) best = std::max(best, bytes * reps / (ms * 1e-3) / 1e9);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    FLOF_HIP(err);
    FLOF_HIP(hipGetLastError());
    *gbps = best;
    API_CATCH(ctx)
}

int flof_check_reciprocal(flof_ctx *ctx, uint32_t lo, uint32_t hi, uint64_t *mismatches)
{
    API_TRY
    FLOF_REQUIRE(ctx && mismatches && lo <= hi, FLOF_ERR_ARG,
                 "flof_check_reciprocal: bad argument");
    Tmp cnt(ctx, sizeof(unsigned long long));
    FLOF_HIP(hipMemsetAsync(cnt.p, 0, sizeof(unsigned long long), ctx->stream));
    hipLaunchKernelGGL(k_check_rcp, dim3(ctx->num_cu * 8), dim3(256), 0, ctx->stream, lo, hi,
                       cnt.as<unsigned long long>());
    check_launch("k_check_rcp");
    unsigned long long h = 0;
    FLOF_HIP(hipMemcpyAsync(&h, cnt.p, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    FLOF_HIP(hipStreamSynchronize(ctx->stream));
    *mismatches = h;
    API_CATCH(ctx)
}

int flof_grid4d_create(flof_ctx *ctx, int X, int Y, int Z, int T, int elem, int dtype,
                       flof_grid4d **out)
{
    flof_grid4d *g = nullptr;
    API_TRY
    FLOF_REQUIRE(ctx && out, FLOF_ERR_ARG, "flof_grid4d_create: null argument");
    *out = nullptr;
    FLOF_REQUIRE(X > 0 && Y > 0 && Z > 0 && T > 0, FLOF_ERR_DIM,
                 "flof_grid4d_create: sizes must be positive");
    FLOF_REQUIRE((int64_t)X * Y * Z * T < (1ll << 31), FLOF_ERR_DIM,
                 "flof_grid4d_create: more than 2^31 cells");
    FLOF_REQUIRE(elem == FLOF_REAL || elem == FLOF_VEC4 || elem == FLOF_INT,
                 FLOF_ERR_TYPE, "flof_grid4d_create: unknown element type");
    FLOF_REQUIRE(dtype == FLOF_F32, FLOF_ERR_TYPE,
                 "flof_grid4d_create: only FP32 grid storage is available");
    g = new flof_grid4d();
    g->ctx = ctx;
    g->X = X;
    g->Y = Y;
    g->Z = Z;
    g->T = T;
    g->elem = elem;
    g->dtype = dtype;
    g->bytes = (size_t)g->n() * elem_bytes(elem);
    FLOF_HIP(hipMalloc(&g->data, g->bytes));
    FLOF_HIP(hipMemsetAsync(g->data, 0, g->bytes, ctx->stream));
    *out = g;
    }
    catch (const flof::Error &e)
    {
        delete g;
        flof::set_error(ctx, e);
        return e.code;
    }
    return FLOF_OK;
}

int flof_grid4d_destroy(flof_grid4d *g)
{
    if (!g) return FLOF_ERR_ARG;
    (void)hipStreamSynchronize(g->ctx->stream);
    if (g->owned && g->data) (void)hipFree(g->data);
    delete g;
    return FLOF_OK;
}

int flof_grid4d_dims(const flof_grid4d *g, int dims[4], int *elem, int *dtype)
{
    if (!g) return FLOF_ERR_ARG;
    if (dims) {
        dims[0] = g->X;
        dims[1] = g->Y;
        dims[2] = g->Z;
        dims[3] = g->T;
    }
    if (elem) *elem = g->elem;
    if (dtype) *dtype = g->dtype;
    return FLOF_OK;
}

void *flof_grid4d_device_ptr(flof_grid4d *g) { return g ? g->data : nullptr; }

int flof_grid4d_upload(flof_grid4d *g, const void *host, size_t bytes)
{
    flof_ctx *ctx = g ? g->ctx : nullptr;
    API_TRY
    FLOF_REQUIRE(g && host, FLOF_ERR_ARG, "flof_grid4d_upload: null argument");
    FLOF_REQUIRE(bytes == g->bytes, FLOF_ERR_DIM, "flof_grid4d_upload: size mismatch");
    FLOF_HIP(hipMemcpyAsync(g->data, host, bytes, hipMemcpyHostToDevice, ctx->stream));
    FLOF_HIP(hipStreamSynchronize(ctx->stream));
    API_CATCH(ctx)
}

int flof_grid4d_download(const flof_grid4d *g, void *host, size_t bytes)
{
    flof_ctx *ctx = g ? g->ctx : nullptr;
    API_TRY
    FLOF_REQUIRE(g && host, FLOF_ERR_ARG, "flof_grid4d_download: null argument");
    FLOF_REQUIRE(bytes == g->bytes, FLOF_ERR_DIM, "flof_grid4d_download: size mismatch");
    FLOF_HIP(hipMemcpyAsync(host, g->data, bytes, hipMemcpyDeviceToHost, ctx->stream));
    FLOF_HIP(hipStreamSynchronize(ctx->stream));
    API_CATCH(ctx)
}

int flof_grid4d_clear(flof_grid4d *g)
{
    flof_ctx *ctx = g ? g->ctx : nullptr;
    API_TRY
    FLOF_REQUIRE(g, FLOF_ERR_ARG, "null grid");
    launch_fill(ctx, g->data, g->bytes);
    API_CATCH(ctx)
}

int flof_grid4d_copy_from(flof_grid4d *dst, const flof_grid4d *src)
{
    flof_ctx *ctx = dst ? dst->ctx : nullptr;
    API_TRY
    need_same(dst, src, "copyFrom");
    FLOF_HIP(hipMemcpyAsync(dst->data, src->data, dst->bytes, hipMemcpyDeviceToDevice,
                            ctx->stream));
    API_CATCH(ctx)
}

int flof_grid4d_swap(flof_grid4d *a, flof_grid4d *b)
{
    flof_ctx *ctx = a ? a->ctx : nullptr;
    API_TRY
    need_same(a, b, "Grid4d::swap()");
    std::swap(a->data, b->data);
    std::swap(a->owned, b->owned);
    API_CATCH(ctx)
}

static int ew_op(flof_grid4d *g, const flof_grid4d *b, const float *value,
                 const float *value2, int op, const char *what)
{
    flof_ctx *ctx = g ? g->ctx : nullptr;
    API_TRY
    FLOF_REQUIRE(g, FLOF_ERR_ARG, std::string(what) + ": null grid");
    if (b) need_same(g, b, what);
    const int64_t n = g->n();
    if (g->elem == FLOF_VEC4) {
        float4 c = value ? make_float4(value[0], value[1], value[2], value[3])
                         : make_float4(0, 0, 0, 0);
        float4 c2 = value2 ? make_float4(value2[0], value2[1], value2[2], value2[3]) : c;
        launch_ew<float4>(ctx, g->data, b ? b->data : nullptr, c, c2, op, n);
    } else if (g->elem == FLOF_INT) {
        int c = value ? (int)value[0] : 0, c2 = value2 ? (int)value2[0] : c;
        launch_ew<int>(ctx, g->data, b ? b->data : nullptr, c, c2, op, n);
    } else {
        float c = value ? value[0] : 0.f, c2 = value2 ? value2[0] : c;
        launch_ew<float>(ctx, g->data, b ? b->data : nullptr, c, c2, op, n);
    }
    API_CATCH(ctx)
}

int flof_grid4d_set_const(flof_grid4d *g, const float *v)
{
    return ew_op(g, nullptr, v, nullptr, EW_SET, "setConst");
}
int flof_grid4d_add_const(flof_grid4d *g, const float *v)
{
    return ew_op(g, nullptr, v, nullptr, EW_ADDC, "addConst");
}
int flof_grid4d_mult_const(flof_grid4d *g, const float *v)
{
    return ew_op(g, nullptr, v, nullptr, EW_MULC, "multConst");
}
int flof_grid4d_add(flof_grid4d *g, const flof_grid4d *a)
{
    return ew_op(g, a, nullptr, nullptr, EW_ADD, "add");
}
int flof_grid4d_sub(flof_grid4d *g, const flof_grid4d *a)
{
    return ew_op(g, a, nullptr, nullptr, EW_SUB, "sub");
}
int flof_grid4d_mult(flof_grid4d *g, const flof_grid4d *a)
{
    return ew_op(g, a, nullptr, nullptr, EW_MULT, "mult");
}
int flof_grid4d_add_scaled(flof_grid4d *g, const flof_grid4d *a, const float *f)
{
    return ew_op(g, a, f, nullptr, EW_ADDSCALED, "addScaled");
}
int flof_grid4d_clamp(flof_grid4d *g, float vmin, float vmax)
{
    float a[4] = {vmin, vmin, vmin, vmin}, b[4] = {vmax, vmax, vmax, vmax};
    return ew_op(g, nullptr, a, b, EW_CLAMP, "clamp");
}

static int get_red(flof_grid4d *g, float *out, int which)
{
    flof_ctx *ctx = g ? g->ctx : nullptr;
    API_TRY
    FLOF_REQUIRE(g && out, FLOF_ERR_ARG, "null argument");
    double mn, mx;
    grid_minmax(g, mn, mx);
    float r;
    if (g->elem == FLOF_VEC4) {
        // getMax/getMin/getMaxAbs: sqrt of max/min squared norm (grid4d.cpp:259-270)
        r = which == 1 ? std::sqrt((float)mn) : std::sqrt((float)mx);
    } else {
        if (which == 0) r = (float)mx;
        else if (which == 1) r = (float)mn;
        else r = std::max(std::fabs((float)mn), std::fabs((float)mx));
    }
    *out = r;
    API_CATCH(ctx)
}
int flof_grid4d_get_max(flof_grid4d *g, float *out) { return get_red(g, out, 0); }
int flof_grid4d_get_min(flof_grid4d *g, float *out) { return get_red(g, out, 1); }
int flof_grid4d_get_max_abs(flof_grid4d *g, float *out) { return get_red(g, out, 2); }

int flof_grid4d_set_bound(flof_grid4d *g, const float *value, int width)
{
    flof_ctx *ctx = g ? g->ctx : nullptr;
    API_TRY
    need_f32(g, "setBound");
    FLOF_REQUIRE(value && g->elem != FLOF_INT, FLOF_ERR_TYPE, "setBound: REAL or VEC4 grid");
    float4 v = g->elem == FLOF_VEC4 ? make_float4(value[0], value[1], value[2], value[3])
                                    : make_float4(value[0], 0, 0, 0);
    launch_set_bound(ctx, g->data, dims_of(g), g->elem == FLOF_VEC4, v, width);
    API_CATCH(ctx)
}

int flof_grid4d_set_bound_neumann(flof_grid4d *g, int width)
{
    flof_ctx *ctx = g ? g->ctx : nullptr;
    API_TRY
    need_f32(g, "setBoundNeumann");
    FLOF_REQUIRE(g->elem != FLOF_INT, FLOF_ERR_TYPE, "setBoundNeumann: REAL or VEC4 grid");
    launch_neumann(ctx, g->data, dims_of(g), g->elem == FLOF_VEC4, width);
    API_CATCH(ctx)
}

int flof_grid4d_max_diff(const flof_grid4d *a, const flof_grid4d *b, double *out)
{
    flof_ctx *ctx = a ? a->ctx : nullptr;
    API_TRY
    need_same(a, b, "grid4dMaxDiff");
    FLOF_REQUIRE(out && a->elem != FLOF_INT, FLOF_ERR_TYPE, "grid4dMaxDiff: REAL or VEC4");
    const int64_t n = a->n();
    const int nb = grid_blocks_xcd(ctx, n);
    RedPtrs r = red_ptrs(ctx, (size_t)nb * 8);
    if (a->elem == FLOF_VEC4)
        hipLaunchKernelGGL(k_maxdiff<float4>, dim3(nb), dim3(256), 0, ctx->stream,
                           (const float4 *)a->data, (const float4 *)b->data, n,
                           (double *)r.part);
    else
        hipLaunchKernelGGL(k_maxdiff<float>, dim3(nb), dim3(256), 0, ctx->stream,
                           (const float *)a->data, (const float *)b->data, n,
                           (double *)r.part);
    check_launch("k_maxdiff");
    hipLaunchKernelGGL(k_maxdiff_fin, dim3(1), dim3(256), 0, ctx->stream,
                       (const double *)r.part, nb, r.result);
    check_launch("k_maxdiff_fin");
    FLOF_HIP(hipMemcpyAsync(out, r.result, 8, hipMemcpyDeviceToHost, ctx->stream));
    FLOF_HIP(hipStreamSynchronize(ctx->stream));
    API_CATCH(ctx)
}

int flof_set_region4d(flof_grid4d *dst, const float start[4], const float end[4],
                      const float *value)
{
    flof_ctx *ctx = dst ? dst->ctx : nullptr;
    API_TRY
    need_f32(dst, "setRegion4d");
    FLOF_REQUIRE(start && end && value, FLOF_ERR_ARG, "setRegion4d: null argument");
    Dims4 D = dims_of(dst);
    const int64_t n = dst->n();
    float4 s = make_float4(start[0], start[1], start[2], start[3]);
    float4 e = make_float4(end[0], end[1], end[2], end[3]);
    if (dst->elem == FLOF_VEC4)
        hipLaunchKernelGGL(k_set_region<float4>, dim3(grid_blocks(ctx, n)), dim3(256),
                           0, ctx->stream, (float4 *)dst->data, D, make_coord4(D), s, e,
                           make_float4(value[0], value[1], value[2], value[3]));
    else
        hipLaunchKernelGGL(k_set_region<float>, dim3(grid_blocks(ctx, n)), dim3(256),
                           0, ctx->stream, (float *)dst->data, D, make_coord4(D), s, e,
                           value[0]);
    check_launch("k_set_region");
    API_CATCH(ctx)
}

int flof_get_slice_from4d(flof_grid4d *src, int srct, float *dst3d, int dx, int dy,
                          int dz)
{
    flof_ctx *ctx = src ? src->ctx : nullptr;
    API_TRY
    need_f32(src, "getSliceFrom4d");
    FLOF_REQUIRE(src->elem == FLOF_REAL && dst3d, FLOF_ERR_TYPE,
                 "getSliceFrom4d: REAL grid");
    if (srct < 0 || srct >= src->T) return FLOF_OK;  // reference: silently nothing
    Dims4 D = dims_of(src);
    hipLaunchKernelGGL(k_get_slice, dim3(grid_blocks(ctx, D.st)), dim3(256), 0,
                       ctx->stream, (const float *)src->data, D, srct, dst3d, dx, dy, dz);
    check_launch("k_get_slice");
    API_CATCH(ctx)
}

int flof_place_grid3d(const float *src3d, int sx, int sy, int sz, flof_grid4d *dst,
                      int dstt)
{
    flof_ctx *ctx = dst ? dst->ctx : nullptr;
    API_TRY
    need_f32(dst, "placeGrid3d");
    FLOF_REQUIRE(dst->elem == FLOF_REAL && src3d, FLOF_ERR_TYPE, "placeGrid3d: REAL grid");
    if (dstt < 0 || dstt >= dst->T) return FLOF_OK;
    Dims4 D = dims_of(dst);
    const int64_t n = (int64_t)sx * sy * sz;
    hipLaunchKernelGGL(k_place3d, dim3(grid_blocks(ctx, n)), dim3(256), 0, ctx->stream,
                       src3d, sx, sy, sz, (float *)dst->data, D, dstt);
    check_launch("k_place3d");
    API_CATCH(ctx)
}

// gridFactor4d (grid4d.cpp:559-569) in the reference's float arithmetic
static void grid_factor4d(const float s1[4], float s2[4], const float *optSize,
                          const float *scale, float srcFac[4], float retOff[4])
{
    for (int c = 0; c < 4; ++c)
        if (optSize && optSize[c] > 0.f) s2[c] = optSize[c];
    for (int c = 0; c < 4; ++c) {
        const float f = s1[c] / s2[c];
        srcFac[c] = f / (scale ? scale[c] : 1.f);
    }
    for (int c = 0; c < 4; ++c)
        retOff[c] = (-retOff[c]) * srcFac[c] + (float)(srcFac[c] * 0.5);
}

int flof_interpolate_grid4d(flof_grid4d *target, flof_grid4d *source,
                            const float offset[4], const float scale[4],
                            const float size[4])
{
    flof_ctx *ctx = target ? target->ctx : nullptr;
    API_TRY
    need_f32(target, "interpolateGrid4d");
    need_f32(source, "interpolateGrid4d");
    FLOF_REQUIRE(target->elem == source->elem && target->elem != FLOF_INT, FLOF_ERR_TYPE,
                 "interpolateGrid4d: REAL->REAL or VEC4->VEC4");
    float s1[4] = {(float)source->X, (float)source->Y, (float)source->Z, (float)source->T};
    float s2[4] = {(float)target->X, (float)target->Y, (float)target->Z, (float)target->T};
    float fac[4], off[4] = {0, 0, 0, 0};
    if (offset)
        for (int c = 0; c < 4; ++c) off[c] = offset[c];
    grid_factor4d(s1, s2, size, scale, fac, off);
    launch_resample(ctx, target->data, dims_of(target), source->data, dims_of(source),
                    target->elem == FLOF_VEC4, fac, off);
    API_CATCH(ctx)
}

int flof_advect4d(flof_grid4d *vel, flof_grid4d *grid, float dt, float dtFac)
{
    flof_ctx *ctx = grid ? grid->ctx : nullptr;
    API_TRY
    need_f32(vel, "advect4d");
    need_f32(grid, "advect4d");
    FLOF_REQUIRE(vel->elem == FLOF_VEC4, FLOF_ERR_TYPE, "advect4d: vel must be Vec4");
    FLOF_REQUIRE(grid->elem == FLOF_REAL || grid->elem == FLOF_VEC4, FLOF_ERR_TYPE,
                 "AdvectSemiLagrange4d: Grid Type is not supported (only Real, Vec4)");
    FLOF_REQUIRE(vel->X == grid->X && vel->Y == grid->Y && vel->Z == grid->Z &&
                     vel->T == grid->T,
                 FLOF_ERR_DIM, "advect4d: size mismatch");
    Tmp tmp(ctx, grid->bytes);
    launch_advect(ctx, (const float4 *)vel->data, 1.f, grid->data, tmp.p, dims_of(grid),
                  grid->elem == FLOF_VEC4, dt * dtFac);
    std::swap(grid->data, tmp.p);  // the grid takes the fresh buffer (swap)
    API_CATCH(ctx)
}

int flof_advect_cfl4d(flof_grid4d *vel, flof_grid4d *grid, float cfl, float velFactor,
                      float dt)
{
    flof_ctx *ctx = grid ? grid->ctx : nullptr;
    API_TRY
    need_f32(vel, "advectCfl4d");
    need_f32(grid, "advectCfl4d");
    FLOF_REQUIRE(vel->elem == FLOF_VEC4, FLOF_ERR_TYPE, "advectCfl4d: vel must be Vec4");
    FLOF_REQUIRE(grid->elem == FLOF_REAL || grid->elem == FLOF_VEC4, FLOF_ERR_TYPE,
                 "advectCfl4d: Real or Vec4 grid");
    advect_cfl(ctx, (const float4 *)vel->data, grid->data, dims_of(grid),
               grid->elem == FLOF_VEC4, cfl, velFactor, dt);
    API_CATCH(ctx)
}

int flof_calc_ls_diff4d(flof_grid4d *i0, flof_grid4d *i1, flof_grid4d *out,
                        float correction, int bnd, float *result)
{
    flof_ctx *ctx = i0 ? i0->ctx : nullptr;
    API_TRY
    need_same(i0, i1, "calcLsDiff4d");
    need_f32(i0, "calcLsDiff4d");
    FLOF_REQUIRE(i0->elem == FLOF_REAL && result, FLOF_ERR_TYPE, "calcLsDiff4d: REAL grids");
    if (out) need_same(i0, out, "calcLsDiff4d(out)");
    *result = (float)host_ls_diff(ctx, (const float *)i0->data, (const float *)i1->data,
                                  out ? (float *)out->data : nullptr, dims_of(i0),
                                  correction, bnd);
    API_CATCH(ctx)
}

int flof_calc_ls_diff4d_noalloc(const float *i0, const float *i1, float *out, int X,
                                int Y, int Z, int T, float correction, int bnd,
                                void *stream, float *result_dev)
{
    try {
        FLOF_REQUIRE(i0 && i1 && result_dev, FLOF_ERR_ARG,
                     "calcLsDiff4d (no allocation): null argument");
        FLOF_REQUIRE(X > 0 && Y > 0 && Z > 0 && T > 0, FLOF_ERR_DIM,
                     "calcLsDiff4d (no allocation): sizes must be positive");
        FLOF_REQUIRE((int64_t)X * Y * Z * T < (1ll << 31), FLOF_ERR_RANGE,
                     "calcLsDiff4d (no allocation): more than 2^31 - 1 cells");
        FLOF_REQUIRE(bnd >= 0, FLOF_ERR_ARG, "calcLsDiff4d (no allocation): negative bnd");
        const Dims4 D{X, Y, Z, T, (int64_t)X * Y, (int64_t)X * Y * Z};
        noalloc_ls_diff(i0, i1, out, D, correction, bnd, (hipStream_t)stream, result_dev);
    } catch (const flof::Error &e) {
        flof::g_noctx_err = e.msg;
        return e.code;
    } catch (const std::exception &e) {
        flof::g_noctx_err = e.what();
        return FLOF_ERR_ARG;
    }
    return FLOF_OK;
}

int flof_drop_sdf4d(flof_grid4d *phi, uint64_t seed, int nspheres, double rmin,
                    double rmax, double clampCells)
{
    flof_ctx *ctx = phi ? phi->ctx : nullptr;
    API_TRY
    need_f32(phi, "drop_sdf4d");
    FLOF_REQUIRE(phi->elem == FLOF_REAL, FLOF_ERR_TYPE, "drop_sdf4d: REAL grid");
    FLOF_REQUIRE(nspheres >= 0 && nspheres <= 8, FLOF_ERR_RANGE, "drop_sdf4d: 0..8 spheres");
    Drops d = make_drops(phi->X, phi->Y, phi->Z, seed, nspheres, rmin, rmax);
    const int64_t n = phi->n();
    hipLaunchKernelGGL(k_drops, dim3(grid_blocks(ctx, n)), dim3(256), 0, ctx->stream,
                       (float *)phi->data, phi->X, phi->Y, phi->Z, phi->T, phi->Y, 1.0,
                       clampCells, d);
    check_launch("k_drops");
    API_CATCH(ctx)
}

int flof_drop_frames3d(flof_ctx *ctx, float *frames, int X, int Y, int Z, int T,
                       int scale, uint64_t seed, int nspheres, double rmin, double rmax,
                       double clampCells)
{
    API_TRY
    FLOF_REQUIRE(ctx && frames && scale >= 1 && X > 0 && Y > 0 && Z > 0 && T > 0,
                 FLOF_ERR_ARG, "drop_frames3d: bad argument");
    FLOF_REQUIRE(nspheres >= 0 && nspheres <= 8, FLOF_ERR_RANGE,
                 "drop_frames3d: 0..8 spheres");
    Drops d = make_drops(X, Y, Z, seed, nspheres, rmin, rmax);
    const int64_t n = (int64_t)X * Y * Z * scale * scale * scale * T;
    hipLaunchKernelGGL(k_drops, dim3(grid_blocks(ctx, n)), dim3(256), 0, ctx->stream,
                       frames, X * scale, Y * scale, Z * scale, T, Y, (double)scale,
                       clampCells * (double)scale, d);
    check_launch("k_drops");
    API_CATCH(ctx)
}

int flof_dev_alloc(flof_ctx *ctx, size_t bytes, void **out)
{
    API_TRY
    FLOF_REQUIRE(ctx && out, FLOF_ERR_ARG, "flof_dev_alloc: null argument");
    *out = nullptr;
    FLOF_HIP(hipMalloc(out, bytes));
    FLOF_HIP(hipMemsetAsync(*out, 0, bytes, ctx->stream));
    API_CATCH(ctx)
}

int flof_dev_free(flof_ctx *ctx, void *p)
{
    API_TRY
    FLOF_REQUIRE(ctx, FLOF_ERR_ARG, "flof_dev_free: null context");
    FLOF_HIP(hipStreamSynchronize(ctx->stream));
    if (p) FLOF_HIP(hipFree(p));
    API_CATCH(ctx)
}

int flof_dev_upload(flof_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    API_TRY
    FLOF_REQUIRE(ctx && dst && src, FLOF_ERR_ARG, "flof_dev_upload: null argument");
    FLOF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    FLOF_HIP(hipStreamSynchronize(ctx->stream));
    API_CATCH(ctx)
}

int flof_dev_download(flof_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    API_TRY
    FLOF_REQUIRE(ctx && dst && src, FLOF_ERR_ARG, "flof_dev_download: null argument");
    FLOF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    FLOF_HIP(hipStreamSynchronize(ctx->stream));
    API_CATCH(ctx)
}

}  // extern "C"
