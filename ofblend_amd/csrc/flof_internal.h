// flof_internal.h -- shared host/device definitions of libflof_hip.
//
// Device code follows the reference's FP32 arithmetic operation by operation
// (the library is compiled with -ffp-contract=off, so no a*b+c is fused), so
// that every kernel reproduces the CPU result bit for bit; the only
// deliberate difference is the order of FP64 sums, which are computed as
// correctly rounded sums (double-double accumulation, see dd_* below).
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <functional>
#include <array>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/flof.h"

namespace flof {

// ---------------------------------------------------------------------------
// error handling
struct Error {
    int code;
    std::string msg;
};

#define FLOF_HIP(call)                                                        \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess)                                                 \
            throw ::flof::Error{FLOF_ERR_HIP,                                 \
                                std::string(#call) + ": " +                   \
                                    hipGetErrorString(e_)};                   \
    } while (0)

#define FLOF_REQUIRE(cond, code, msg)                                         \
    do {                                                                      \
        if (!(cond)) throw ::flof::Error{(code), (msg)};                      \
    } while (0)

// ---------------------------------------------------------------------------
// device memory pool: cached blocks by exact size, stream ordered (all work
// of a context runs on its one stream, so a released block may be reused by
// the next enqueued kernel without a sync)
struct Pool {
    std::unordered_map<size_t, std::vector<void *>> free_;
    size_t cached = 0;
    void *get(size_t bytes);
    void put(void *p, size_t bytes);
    void trim();
};

struct KTimer {
    bool on = false;
    struct Rec {
        hipEvent_t e0, e1;
        int64_t tag;  // cells of the grid the launch worked on
    };
    std::vector<Rec> pending;
};

}  // namespace flof

struct flof_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    flof::Pool pool;
    int num_cu = 256;
    flof::KTimer timers[FLOF_KT_COUNT];
    // pinned host staging for scalar read-backs
    void *host_scratch = nullptr;
    // capture CG iteration chunks as hipGraphs (FLOF_NO_GRAPH=1 disables)
    bool use_graphs = true;
    // workgroups per CU of the streaming kernels (FLOF_BLOCKS_PER_CU)
    int blocks_per_cu = 8;
    // t-marching stencil (FLOF_NO_TMARCH=1 disables) and its workgroups per
    // CU (FLOF_TM_BPC)
    bool use_tmarch = true;
    int tm_blocks_per_cu = 4;
    // kernels of the two PCG stencil passes: FLOF_STENCIL_AUTO / _TMARCH /
    // _LDS (FLOF_LDS_STENCIL=0 or 1 forces the t-marching or the LDS-tiled
    // ones), work items per CU of the LDS-tiled ones (FLOF_LDS_BPC) and
    // XCD-contiguous work items (FLOF_LDS_XCD=0 disables)
    int stencil = FLOF_STENCIL_AUTO;
    int lds_blocks_per_cu = 2;
    bool lds_xcd = true;
    // device scratch for reductions (partials + state), grown on demand
    void *red = nullptr;
    size_t red_bytes = 0;
    // tag recorded with timed launches (the PCG sets the level's cell count)
    int64_t timer_tag = 0;
    // worker contexts of batched solves (same device, own streams)
    std::vector<flof_ctx *> workers;
    // launch-shape autotuner: on/off and the choice per grid shape
    bool autotune = false;
    struct TunedShape {
        int bpc;      // workgroups per CU of the t-marching kernels
        int lds_bpc;  // work items per CU of the LDS-tiled stencil kernels
                      // (sets their t-segment length); 0: the t-marching
                      // stencil kernels
    };
    std::map<std::array<int, 5>, TunedShape> tm_tuned;  // X, Y, Z, T, variant
    // storage of the CG search direction for the call in progress
    // (FLOF_PSTORE_*), set from flof_of_params.precision by the entry points
    int p_store = FLOF_PSTORE_F32;
    // images of the call in progress stored as bfloat16 with binary16
    // gradients (the mixed-precision solve, flof_of_params.precision)
    bool img_bf16 = false;
};

// the kernel variant part of the autotuner's key of a grid shape
inline int tune_variant(const flof_ctx *ctx)
{
    return ctx->p_store + 2 * ctx->stencil + 16 * (ctx->img_bf16 ? 1 : 0);
}

struct flof_warp_state {
    flof_ctx *ctx = nullptr;
    std::vector<flof_grid4d *> defos;  // deformation volumes, device resident
    bool use_vols = false;             // useDefoVols of loadAdvectTimeSlice_OptInit
    bool aligned = false;              // doAligned
    float partial = 0.2f;              // partialLoadFac
};

struct flof_grid4d {
    flof_ctx *ctx = nullptr;
    int X = 0, Y = 0, Z = 0, T = 0;
    int elem = FLOF_REAL;
    int dtype = FLOF_F32;
    void *data = nullptr;
    size_t bytes = 0;
    bool owned = true;
    int64_t n() const { return (int64_t)X * Y * Z * T; }
};

namespace flof {

inline size_t elem_bytes(int elem)
{
    return elem == FLOF_VEC4 ? 16 : 4;
}

// RAII temporary device buffer from the context pool
struct Tmp {
    flof_ctx *ctx = nullptr;
    void *p = nullptr;
    size_t bytes = 0;
    Tmp() {}
    Tmp(flof_ctx *c, size_t b) : ctx(c), bytes(b) { p = c->pool.get(b); }
    Tmp(const Tmp &) = delete;
    Tmp &operator=(const Tmp &) = delete;
    Tmp(Tmp &&o) noexcept : ctx(o.ctx), p(o.p), bytes(o.bytes) { o.p = nullptr; }
    Tmp &operator=(Tmp &&o) noexcept
    {
        release();
        ctx = o.ctx;
        p = o.p;
        bytes = o.bytes;
        o.p = nullptr;
        return *this;
    }
    ~Tmp() { release(); }
    void release()
    {
        if (p) ctx->pool.put(p, bytes);
        p = nullptr;
    }
    template <class T> T *as() const { return (T *)p; }
};

// Launch on the context stream.  While timing of kernel class `kind` is on,
// the start/stop events are attached to the dispatch packet itself
// (hipExtLaunchKernelGGL), so they bracket the kernel's execution and not the
// surrounding launch overhead.
template <typename... KArgs, typename... Args>
inline void klaunch(flof_ctx *ctx, int kind, void (*kernel)(KArgs...), dim3 grid,
                    dim3 block, Args... args)
{
    if (kind >= 0 && kind < FLOF_KT_COUNT && ctx->timers[kind].on) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        FLOF_HIP(hipEventCreate(&e0));
        FLOF_HIP(hipEventCreate(&e1));
        hipExtLaunchKernelGGL(kernel, grid, block, 0, ctx->stream, e0, e1, 0, args...);
        ctx->timers[kind].pending.push_back({e0, e1, ctx->timer_tag});
    } else {
        hipLaunchKernelGGL(kernel, grid, block, 0, ctx->stream, args...);
    }
}

// grid size for grid-stride kernels: enough waves to fill the chip
inline int grid_blocks(flof_ctx *ctx, int64_t work, int block = 256,
                       int per_cu = 8)
{
    int64_t need = (work + block - 1) / block;
    int64_t cap = (int64_t)ctx->num_cu * per_cu;
    if (need < 1) need = 1;
    return (int)(need < cap ? need : cap);
}
// same, rounded up to a multiple of the XCD count (for FOR_XCD kernels)
#define FLOF_NXCD 8
inline int grid_blocks_xcd(flof_ctx *ctx, int64_t work, int block = 256,
                           int per_cu = 0)
{
    int64_t g = grid_blocks(ctx, work, block, per_cu > 0 ? per_cu : ctx->blocks_per_cu);
    g = ((g + FLOF_NXCD - 1) / FLOF_NXCD) * FLOF_NXCD;
    return (int)g;
}

// ---------------------------------------------------------------------------
// device helpers

struct Dims4 {
    int X, Y, Z, T;
    int64_t sz, st;  // X*Y, X*Y*Z
};
inline Dims4 dims_of(const flof_grid4d *g)
{
    Dims4 d{g->X, g->Y, g->Z, g->T, (int64_t)g->X * g->Y,
            (int64_t)g->X * g->Y * g->Z};
    return d;
}

// unsigned 32-bit division by a runtime constant (round-up method of
// Granlund-Montgomery; exact for every 32-bit n, d == 1 special-cased)
struct FastDiv {
    uint32_t d, m, s;
};
inline FastDiv make_fastdiv(uint32_t d)
{
    FastDiv f;
    f.d = d;
    uint32_t s = 0;
    while ((1ull << s) < d) ++s;
    f.s = s;
    f.m = (uint32_t)(((1ull << 32) * ((1ull << s) - d)) / d + 1);
    return f;
}
__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv &f)
{
    if (f.d == 1) return n;
    uint32_t t = __umulhi(n, f.m);
    return (t + ((n - t) >> 1)) >> (f.s - 1);
}

struct Coord4 {
    FastDiv fx, fy, fz;
};
inline Coord4 make_coord4(const Dims4 &D)
{
    return Coord4{make_fastdiv(D.X), make_fastdiv(D.Y), make_fastdiv(D.Z)};
}
__device__ __forceinline__ void decompose(uint32_t n, const Coord4 &c, int &i,
                                          int &j, int &k, int &t)
{
    uint32_t q = fdiv(n, c.fx);
    i = (int)(n - q * c.fx.d);
    uint32_t q2 = fdiv(q, c.fy);
    j = (int)(q - q2 * c.fy.d);
    uint32_t q3 = fdiv(q2, c.fz);
    k = (int)(q2 - q3 * c.fz.d);
    t = (int)q3;
}

// float4 helpers (componentwise, single rounding per op)
// The correctly rounded FP32 reciprocal 1/d (what the IEEE division 1.f / d
// returns) for d in [2^-126, 2^126): the hardware reciprocal refined by one
// FMA Newton step, y + y (1 - d y), 3 instead of about 10 instructions.
// Every float of the range was compared with the division on gfx950
// (flof_check_reciprocal: 0 mismatches); outside it (0, denormals, d at or
// above 2^126, inf, NaN) the two differ, so callers use it only where
// rcp_exact_range(d) holds.
__device__ __forceinline__ float rcp_newton(float d)
{
    const float y = __builtin_amdgcn_rcpf(d);
    return __builtin_fmaf(y, __builtin_fmaf(-d, y, 1.0f), y);
}
__device__ __forceinline__ bool rcp_exact_range(float d)
{
    return d >= 0x1p-126f && d < 0x1p126f;
}

__device__ __forceinline__ float4 f4add(float4 a, float4 b)
{
    return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}
__device__ __forceinline__ float4 f4sub(float4 a, float4 b)
{
    return make_float4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w);
}
__device__ __forceinline__ float4 f4mul(float4 a, float4 b)
{
    return make_float4(a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w);
}
__device__ __forceinline__ float4 f4muls(float4 a, float s)
{
    return make_float4(a.x * s, a.y * s, a.z * s, a.w * s);
}
__device__ __forceinline__ float f4get(const float4 &v, int c)
{
    return c == 0 ? v.x : c == 1 ? v.y : c == 2 ? v.z : v.w;
}

template <class T> struct VOps;
template <> struct VOps<float> {
    __device__ static __forceinline__ float add(float a, float b) { return a + b; }
    __device__ static __forceinline__ float muls(float a, float s) { return a * s; }
    __device__ static __forceinline__ float zero() { return 0.f; }
};
template <> struct VOps<float4> {
    __device__ static __forceinline__ float4 add(float4 a, float4 b) { return f4add(a, b); }
    __device__ static __forceinline__ float4 muls(float4 a, float s) { return f4muls(a, s); }
    __device__ static __forceinline__ float4 zero() { return make_float4(0.f, 0.f, 0.f, 0.f); }
};

// bfloat16 storage (round to nearest even on every store; NaN stays NaN)
__device__ __forceinline__ float bf16_to_f32(unsigned short u)
{
    return __uint_as_float((unsigned)u << 16);
}
__device__ __forceinline__ unsigned short f32_to_bf16(float f)
{
    unsigned u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);  // NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
// a Real grid value stored as bfloat16 (the mixed-precision images)
struct bf16_t {
    unsigned short u;
};
// value of storage type S as arithmetic type T, and back (rounded)
template <class T, class S> struct Stor {
    __device__ static __forceinline__ T ld(const S &s) { return s; }
    __device__ static __forceinline__ S st(const T &v) { return v; }
};
template <> struct Stor<float, bf16_t> {
    __device__ static __forceinline__ float ld(const bf16_t &s) { return bf16_to_f32(s.u); }
    __device__ static __forceinline__ bf16_t st(float v) { return bf16_t{f32_to_bf16(v)}; }
};
// x rounded to the nearest binary16 value (ties to even; subnormals kept;
// magnitudes from 65520 on become infinite), returned as FP32: the FP16
// gradients of the mixed-precision solve, by bit operations so the result
// does not depend on the device's FP16 denormal mode
__device__ __forceinline__ float round_f16(float x)
{
    const unsigned u = __float_as_uint(x);
    const unsigned sign = u & 0x80000000u;
    unsigned a = u & 0x7fffffffu;
    if (a >= 0x7f800000u) return x;                               // inf, NaN
    if (a >= 0x477ff000u) return __uint_as_float(sign | 0x7f800000u);  // >= 65520
    if (a < 0x38800000u) {                                        // < 2^-14
        const float q = rintf(fabsf(x) * 16777216.0f) * 5.9604644775390625e-8f;
        return __uint_as_float(sign | __float_as_uint(q));
    }
    a += 0x0fffu + ((a >> 13) & 1u);
    return __uint_as_float(sign | (a & ~0x1fffu));
}

// Quadrilinear sample of a cell-centred 4D grid at p (cells at c + 1/2).
// Per dimension q = p - 1/2, i = (int)q (truncation), w1 = q - i; q < 0
// clamps to (0; 1, 0), i >= n - 1 clamps to (n - 2; 0, 1); combination order
// y, x, z, t -- util/vector4d.h:422-516.
template <class T, class S = T>
__device__ __forceinline__ T interp4(const S *__restrict__ d, const Dims4 &D,
                                     float px, float py, float pz, float pt)
{
    typedef VOps<T> O;
    typedef Stor<T, S> L;
    px = px - 0.5f;
    py = py - 0.5f;
    pz = pz - 0.5f;
    pt = pt - 0.5f;
    int xi = (int)px, yi = (int)py, zi = (int)pz, ti = (int)pt;
    float s1 = px - (float)xi, s0 = 1.f - s1;
    float t1 = py - (float)yi, t0 = 1.f - t1;
    float f1 = pz - (float)zi, f0 = 1.f - f1;
    float g1 = pt - (float)ti, g0 = 1.f - g1;
    if (px < 0.f) { xi = 0; s0 = 1.f; s1 = 0.f; }
    if (py < 0.f) { yi = 0; t0 = 1.f; t1 = 0.f; }
    if (pz < 0.f) { zi = 0; f0 = 1.f; f1 = 0.f; }
    if (pt < 0.f) { ti = 0; g0 = 1.f; g1 = 0.f; }
    if (xi >= D.X - 1) { xi = D.X - 2; s0 = 0.f; s1 = 1.f; }
    if (yi >= D.Y - 1) { yi = D.Y - 2; t0 = 0.f; t1 = 1.f; }
    if (zi >= D.Z - 1) { zi = D.Z - 2; f0 = 0.f; f1 = 1.f; }
    if (ti >= D.T - 1) { ti = D.T - 2; g0 = 0.f; g1 = 1.f; }
    const int64_t sY = D.X, sZ = D.sz, sT = D.st;
    const S *b = d + (int64_t)xi + sY * yi + sZ * zi + sT * ti;
    // y first, then x, z, t
    T a00 = O::add(O::muls(L::ld(b[0]), t0), O::muls(L::ld(b[sY]), t1));
    T a01 = O::add(O::muls(L::ld(b[1]), t0), O::muls(L::ld(b[1 + sY]), t1));
    T a02 = O::add(O::muls(L::ld(b[sZ]), t0), O::muls(L::ld(b[sY + sZ]), t1));
    T a03 = O::add(O::muls(L::ld(b[1 + sZ]), t0), O::muls(L::ld(b[1 + sY + sZ]), t1));
    T z0 = O::add(O::muls(O::add(O::muls(a00, s0), O::muls(a01, s1)), f0),
                  O::muls(O::add(O::muls(a02, s0), O::muls(a03, s1)), f1));
    T a10 = O::add(O::muls(L::ld(b[sT]), t0), O::muls(L::ld(b[sT + sY]), t1));
    T a11 = O::add(O::muls(L::ld(b[sT + 1]), t0), O::muls(L::ld(b[sT + 1 + sY]), t1));
    T a12 = O::add(O::muls(L::ld(b[sT + sZ]), t0), O::muls(L::ld(b[sT + sY + sZ]), t1));
    T a13 = O::add(O::muls(L::ld(b[sT + 1 + sZ]), t0), O::muls(L::ld(b[sT + 1 + sY + sZ]), t1));
    T z1 = O::add(O::muls(O::add(O::muls(a10, s0), O::muls(a11, s1)), f0),
                  O::muls(O::add(O::muls(a12, s0), O::muls(a13, s1)), f1));
    return O::add(O::muls(z0, g0), O::muls(z1, g1));
}

// Trilinear sample of a 3D grid (util/interpol.h:57-116; z is not clamped
// when nz == 1).
template <class T>
__device__ __forceinline__ T interp3(const T *__restrict__ d, int nx, int ny,
                                     int nz, int64_t sZ, float px, float py,
                                     float pz)
{
    typedef VOps<T> O;
    px = px - 0.5f;
    py = py - 0.5f;
    pz = pz - 0.5f;
    int xi = (int)px, yi = (int)py, zi = (int)pz;
    float s1 = px - (float)xi, s0 = 1.f - s1;
    float t1 = py - (float)yi, t0 = 1.f - t1;
    float f1 = pz - (float)zi, f0 = 1.f - f1;
    if (px < 0.f) { xi = 0; s0 = 1.f; s1 = 0.f; }
    if (py < 0.f) { yi = 0; t0 = 1.f; t1 = 0.f; }
    if (pz < 0.f) { zi = 0; f0 = 1.f; f1 = 0.f; }
    if (xi >= nx - 1) { xi = nx - 2; s0 = 0.f; s1 = 1.f; }
    if (yi >= ny - 1) { yi = ny - 2; t0 = 0.f; t1 = 1.f; }
    if (nz > 1) {
        if (zi >= nz - 1) { zi = nz - 2; f0 = 0.f; f1 = 1.f; }
    }
    const int64_t sY = nx;
    const T *b = d + (int64_t)xi + sY * yi + sZ * zi;
    T a00 = O::add(O::muls(b[0], t0), O::muls(b[sY], t1));
    T a01 = O::add(O::muls(b[1], t0), O::muls(b[1 + sY], t1));
    T a02 = O::add(O::muls(b[sZ], t0), O::muls(b[sY + sZ], t1));
    T a03 = O::add(O::muls(b[1 + sZ], t0), O::muls(b[1 + sY + sZ], t1));
    return O::add(O::muls(O::add(O::muls(a00, s0), O::muls(a01, s1)), f0),
                  O::muls(O::add(O::muls(a02, s0), O::muls(a03, s1)), f1));
}

// ---------------------------------------------------------------------------
// Correctly rounded FP64 sums: each partial is an unevaluated double-double
// (hi + lo); additions use TwoSum and carry the rounding error in lo, and the
// final value is rounded once (hi + lo).  The result equals the exact sum
// rounded to double except when the exact sum lies within ~1e-25 (relative)
// of a rounding boundary, independently of the order of the terms.
struct DD {
    double hi, lo;
};
__device__ __forceinline__ void dd_add(DD &a, double x)
{
    const double s = a.hi + x;
    const double bp = s - a.hi;
    a.lo += (a.hi - (s - bp)) + (x - bp);
    a.hi = s;
}
__device__ __forceinline__ void dd_merge(DD &a, const DD &b)
{
    const double s = a.hi + b.hi;
    const double bp = s - a.hi;
    const double e = (a.hi - (s - bp)) + (b.hi - bp);
    a.lo = a.lo + b.lo + e;
    a.hi = s;
}
__device__ __forceinline__ double shfl_xor_d(double v, int m)
{
    return __shfl_xor(v, m, 64);
}
// wavefront (64 lanes) reductions
__device__ __forceinline__ DD wave_dd(DD v)
{
    for (int m = 32; m >= 1; m >>= 1) {
        DD o{shfl_xor_d(v.hi, m), shfl_xor_d(v.lo, m)};
        dd_merge(v, o);
    }
    return v;
}
__device__ __forceinline__ float wave_max(float v)
{
    for (int m = 32; m >= 1; m >>= 1) v = fmaxf(v, __shfl_xor(v, m, 64));
    return v;
}
__device__ __forceinline__ float wave_min(float v)
{
    for (int m = 32; m >= 1; m >>= 1) v = fminf(v, __shfl_xor(v, m, 64));
    return v;
}

// Cell ranges per XCD.  Workgroups are dispatched round-robin over the 8
// XCDs (blockIdx % 8), each with its own L2; giving every XCD one contiguous
// 1/8 of the cells keeps a stencil's neighbour rows in the L2 that already
// holds them.  Requires gridDim.x to be a multiple of 8 (grid_blocks_xcd).
// No grid-wide fences are used anywhere: an agent-scope fence writes back
// and invalidates the XCD's L2, so partial results go to memory with plain
// stores and a following kernel reduces them.
struct XRange {
    int64_t cur_, end_, stride_;
};
__device__ __forceinline__ XRange xcd_range(int64_t n)
{
    XRange r;
    if (gridDim.x % FLOF_NXCD != 0) {
        // plain grid-stride loop (never a zero stride)
        r.cur_ = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        r.end_ = n;
        r.stride_ = (int64_t)gridDim.x * blockDim.x;
        return r;
    }
    const int xcd = blockIdx.x % FLOF_NXCD;
    const int lb = blockIdx.x / FLOF_NXCD;
    const int gb = gridDim.x / FLOF_NXCD;
    const int64_t chunk = (n + FLOF_NXCD - 1) / FLOF_NXCD;
    const int64_t b = (int64_t)xcd * chunk;
    r.cur_ = b + (int64_t)lb * blockDim.x + threadIdx.x;
    r.end_ = b + chunk < n ? b + chunk : n;
    r.stride_ = (int64_t)gb * blockDim.x;
    return r;
}
#define FOR_XCD(IDX, N)                                                        \
    for (XRange xr_ = xcd_range(N); xr_.cur_ < xr_.end_; xr_.cur_ += xr_.stride_) \
        if (const int64_t IDX = xr_.cur_; true)

// block-wide reductions for blockDim.x == 256 (4 wavefronts)
__device__ __forceinline__ DD block_dd(DD v)
{
    __shared__ double sh_hi[4], sh_lo[4];
    v = wave_dd(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) {
        sh_hi[w] = v.hi;
        sh_lo[w] = v.lo;
    }
    __syncthreads();
    DD r{sh_hi[0], sh_lo[0]};
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) dd_merge(r, DD{sh_hi[i], sh_lo[i]});
    __syncthreads();
    return r;
}
__device__ __forceinline__ float block_max(float v)
{
    __shared__ float sh[4];
    v = wave_max(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) sh[w] = v;
    __syncthreads();
    float r = sh[0];
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) r = fmaxf(r, sh[i]);
    __syncthreads();
    return r;
}
__device__ __forceinline__ float block_min(float v)
{
    __shared__ float sh[4];
    v = wave_min(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) sh[w] = v;
    __syncthreads();
    float r = sh[0];
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) r = fminf(r, sh[i]);
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------------------
// host-side helpers implemented in flof_core.hip
void set_error(flof_ctx *ctx, const Error &e);
void *reduce_scratch(flof_ctx *ctx, size_t bytes);
void check_launch(const char *what);
float host_max_norm_vec4(flof_ctx *ctx, const float4 *v, int64_t n);
double host_ls_diff(flof_ctx *ctx, const float *a, const float *b, float *out,
                    const Dims4 &D, float corr, int bnd);
void place_frames(flof_grid4d *phi, int sx, int sy, int sz,
                  const std::function<const float *(int)> &frame_of, const float offset[4],
                  const float scale[4], int fileIdxStart, int fileIdxEnd, int debugSkipLoad,
                  float spread, const float overrideSize[4], float overrideTimeOff,
                  int overrideGoodRegion, float loadTimeScale, int rescaleSdfValues,
                  float sdfIsoOff, float repeatStartFrame);
void launch_resample(flof_ctx *ctx, void *dst, const Dims4 &Dd, const void *src,
                     const Dims4 &Ds, bool vec4, const float fac[4],
                     const float off[4]);
void launch_pyramid(flof_ctx *ctx, void *dst, const Dims4 &Dd, const void *src,
                    const Dims4 &Ds, bool vec4, float post_scale);
void launch_advect(flof_ctx *ctx, const float4 *vel, float vel_scale,
                   void *grid, void *tmp, const Dims4 &D, bool vec4, float dt);
void advect_cfl(flof_ctx *ctx, const float4 *vel, void *grid, const Dims4 &D,
                bool vec4, float cfl, float velFactor, float dt);
void launch_neumann(flof_ctx *ctx, void *g, const Dims4 &D, bool vec4, int w);
void launch_set_bound(flof_ctx *ctx, void *g, const Dims4 &D, bool vec4,
                      float4 value, int w);
void launch_axpy4(flof_ctx *ctx, float4 *y, const float4 *x, float4 a,
                  int64_t n);
void launch_scale4(flof_ctx *ctx, float4 *y, float4 a, int64_t n);
void launch_fill(flof_ctx *ctx, void *p, size_t bytes);
// bfloat16 images of the mixed-precision solve (Real grids, 2 B per cell)
void launch_to_bf16(flof_ctx *ctx, void *dst, const float *src, int64_t n);
void launch_pyramid_bf16(flof_ctx *ctx, void *dst, const Dims4 &Dd, const void *src,
                         const Dims4 &Ds);
void advect_cfl_bf16(flof_ctx *ctx, const float4 *vel, void *grid, const Dims4 &D,
                     float cfl, float dt);
void launch_neumann_bf16(flof_ctx *ctx, void *g, const Dims4 &D, int w);
double host_ls_diff_bf16(flof_ctx *ctx, const void *a, const void *b, const Dims4 &D,
                         float corr, int bnd);
// the warm-start network's flow from (i0, i1) of one grid, on ctx's stream
// (the network must belong to ctx); keeps no activations for a backward pass
void warmnet_infer(flof_ctx *ctx, flof_warmnet *net, const float *i0, const float *i1,
                   const Dims4 &D, float4 *flow);
// t-segment length of the LDS-tiled stencil kernels at lds_bpc work items per CU
int lds_seglen(const flof_ctx *ctx, const Dims4 &D, int lds_bpc);

}  // namespace flof
