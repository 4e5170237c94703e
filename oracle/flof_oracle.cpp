// flof_oracle.cpp -- CPU restatement of the FlOF 4D deformation path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU library is
// compared against; only tests/, the smoke check and the bench's CPU-baseline
// leg load it.  The product library never links or calls it.
//
// It restates, from the arithmetic of the reference, the functions of
// source/plugin/optflow4d.cpp, source/grid4d.{h,cpp} and
// source/util/{vector4d.h,interpol.h} that the FlOF deformation path runs.
// Every function is templated on the storage type R:
//   R = float  : "FP32-faithful" -- the reference's default build (Real=float,
//                source/util/vectorbase.h:42-51), including where the
//                reference computes in double and stores to float;
//   R = double : the reference's -DDOUBLEPRECISION build (vectorbase.h:50-57),
//                used as the FP64 oracle.
// Build with -ffp-contract=off so a*b+c is never fused (the reference build
// for x86-64 has no FMA either).
//
// Loops that the reference runs as KERNEL(...) are OpenMP-parallel over the
// outermost t (as the reference's OpenMP build, preprocessor/
// codegen_kernel.cpp:213-290); the CG and the matrix set-up are serial, as in
// the reference (optflow4d.cpp:211-355,398-493).
//
// Parity status per function (what pins it) is listed in the header comment
// of oracle/oracle.py.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <type_traits>
#include <vector>

#include "../include/flof.h"

namespace orc {

typedef int64_t idx_t;

template <class R> struct Eps;
template <> struct Eps<float> {
    static constexpr float v = 1e-6f;  // VECTOR_EPSILON, vectorbase.h:53
};
template <> struct Eps<double> {
    static constexpr double v = 1e-10;  // vectorbase.h:57
};

template <class R> struct V4 {
    R x, y, z, t;
    V4() : x(0), y(0), z(0), t(0) {}
    explicit V4(R s) : x(s), y(s), z(s), t(s) {}
    V4(R a, R b, R c, R d) : x(a), y(b), z(c), t(d) {}
    R &operator[](int i) { return (&x)[i]; }
    const R &operator[](int i) const { return (&x)[i]; }
};
template <class R> inline V4<R> operator+(const V4<R> &a, const V4<R> &b)
{
    return V4<R>(a.x + b.x, a.y + b.y, a.z + b.z, a.t + b.t);
}
template <class R> inline V4<R> operator-(const V4<R> &a, const V4<R> &b)
{
    return V4<R>(a.x - b.x, a.y - b.y, a.z - b.z, a.t - b.t);
}
// componentwise product
template <class R> inline V4<R> operator*(const V4<R> &a, const V4<R> &b)
{
    return V4<R>(a.x * b.x, a.y * b.y, a.z * b.z, a.t * b.t);
}
// vector times scalar of the storage type
template <class R> inline V4<R> operator*(const V4<R> &a, R s)
{
    return V4<R>(a.x * s, a.y * s, a.z * s, a.t * s);
}
// vector times a double computed in double and stored back (Vec4 * double)
template <class R> inline V4<R> muld(const V4<R> &a, double s)
{
    return V4<R>((R)(a.x * s), (R)(a.y * s), (R)(a.z * s), (R)(a.t * s));
}
template <class R> inline R muld(R a, double s) { return (R)(a * s); }
// division by a scalar of the storage type (Vec4 / Real, vector4d.h:258)
template <class R> inline V4<R> divs(const V4<R> &a, R s)
{
    return V4<R>(a.x / s, a.y / s, a.z / s, a.t / s);
}
template <class R> inline R divs(R a, R s) { return a / s; }
template <class R> inline R normSq(const V4<R> &v)
{
    return v.x * v.x + v.y * v.y + v.z * v.z + v.t * v.t;
}

template <class T> struct Scal;
template <> struct Scal<float> { typedef float R; };
template <> struct Scal<double> { typedef double R; };
template <> struct Scal<V4<float>> { typedef float R; };
template <> struct Scal<V4<double>> { typedef double R; };

struct Dims {
    int X, Y, Z, T;
    idx_t sz() const { return (idx_t)X * Y; }
    idx_t st() const { return (idx_t)X * Y * Z; }
    idx_t n() const { return (idx_t)X * Y * Z * T; }
    idx_t at(int i, int j, int k, int t) const
    {
        return (idx_t)i + (idx_t)X * j + sz() * k + st() * t;
    }
    bool inb(int i, int j, int k, int t, int b) const
    {
        return i >= b && j >= b && i < X - b && j < Y - b && k >= b &&
               k < Z - b && t >= b && t < T - b;
    }
};

// Summation of FP64 accumulations.  Mode 0 follows the reference: a serial
// double sum (optflow4d.cpp:234-241, :899-913).  Mode 1 returns the
// correctly rounded value of the exact sum, accumulated as an unevaluated
// double-double (TwoSum with the rounding errors carried in a second double)
// and rounded once at the end; this is the order-independent sum the GPU
// library computes, so mode 1 isolates every other difference.
static int g_sum_mode = 0;

struct DSum {
    double hi = 0., lo = 0.;
    inline void add(double x)
    {
        if (g_sum_mode == 0) {
            hi += x;
            return;
        }
        const double s = hi + x;
        const double bp = s - hi;
        lo += (hi - (s - bp)) + (x - bp);
        hi = s;
    }
    inline double value() const { return g_sum_mode == 0 ? hi : hi + lo; }
};

// Storage of the CG search direction (extension; no reference counterpart):
// 0 keeps it in R, 1 rounds every stored value to bfloat16 (round to nearest
// even, through float) and the rounded value is the one used.
static int g_pstore = 0;

// float -> bfloat16 -> float, round to nearest even (NaN stays NaN)
static float round_bf16(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return f;
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}

template <class R> static inline R round_p(R v)
{
    if (g_pstore == 1) return (R)round_bf16((float)v);
    return v;
}

// Mixed-precision images (extension; no reference counterpart): with
// g_imgmode set, the FP32 solve stores its images (i0, i1, their pyramid
// levels and every warp of i0) as bfloat16 -- each stored value rounded to
// nearest even -- and rounds each component of the gradient of i1 to the
// nearest binary16 value where it is formed.  The FP64 variant is never
// rounded (it is the accuracy yardstick).
static int g_imgmode = 0;

// float -> binary16 -> float, round to nearest even, subnormals kept,
// magnitudes from 65520 on to infinity (the IEEE 754 conversion)
static float round_f16(float x)
{
    uint32_t u;
    std::memcpy(&u, &x, 4);
    const uint32_t sign = u & 0x80000000u;
    uint32_t a = u & 0x7fffffffu;
    if (a >= 0x7f800000u) return x;
    uint32_t r;
    if (a >= 0x477ff000u) {
        r = sign | 0x7f800000u;
    } else if (a < 0x38800000u) {
        // below 2^-14: multiples of 2^-24 (exact scalings by powers of 2)
        const float q = std::nearbyint(std::fabs(x) * 16777216.0f) * 5.9604644775390625e-8f;
        std::memcpy(&r, &q, 4);
        r |= sign;
    } else {
        a += 0x0fffu + ((a >> 13) & 1u);
        r = sign | (a & ~0x1fffu);
    }
    float o;
    std::memcpy(&o, &r, 4);
    return o;
}

template <class R> static inline R img_round(R v)
{
    if constexpr (std::is_same<R, float>::value)
        if (g_imgmode) return round_bf16(v);
    return v;
}
template <class R> static inline R grad_round(R v)
{
    if constexpr (std::is_same<R, float>::value)
        if (g_imgmode) return round_f16(v);
    return v;
}
template <class R> static void img_round_all(R *a, idx_t n)
{
    if constexpr (std::is_same<R, float>::value)
        if (g_imgmode)
            for (idx_t c = 0; c < n; ++c) a[c] = round_bf16(a[c]);
}

// ---------------------------------------------------------------------------
// Quadrilinear sampling, util/vector4d.h:422-516.  Cell centres are at c+1/2;
// per dimension q = p - 1/2, i = (int)q, w1 = q - i; q < 0 clamps to (0; 1,0),
// i >= n-1 clamps to (n-2; 0,1).  The combination order (y, x, z, t) is the
// reference's.
template <class R, class T>
inline T interp4(const T *d, const Dims &D, R px_, R py_, R pz_, R pt_)
{
    R px = px_ - (R)0.5f, py = py_ - (R)0.5f, pz = pz_ - (R)0.5f,
      pt = pt_ - (R)0.5f;
    int xi = (int)px, yi = (int)py, zi = (int)pz, ti = (int)pt;
    R s1 = px - (R)xi, s0 = (R)(1. - s1);
    R t1 = py - (R)yi, t0 = (R)(1. - t1);
    R f1 = pz - (R)zi, f0 = (R)(1. - f1);
    R g1 = pt - (R)ti, g0 = (R)(1. - g1);
    if (px < 0.) { xi = 0; s0 = 1; s1 = 0; }
    if (py < 0.) { yi = 0; t0 = 1; t1 = 0; }
    if (pz < 0.) { zi = 0; f0 = 1; f1 = 0; }
    if (pt < 0.) { ti = 0; g0 = 1; g1 = 0; }
    if (xi >= D.X - 1) { xi = D.X - 2; s0 = 0; s1 = 1; }
    if (yi >= D.Y - 1) { yi = D.Y - 2; t0 = 0; t1 = 1; }
    if (zi >= D.Z - 1) { zi = D.Z - 2; f0 = 0; f1 = 1; }
    if (ti >= D.T - 1) { ti = D.T - 2; g0 = 0; g1 = 1; }
    const idx_t sX = 1, sY = D.X, sZ = D.sz(), sT = D.st();
    const idx_t i = (idx_t)xi + sY * yi + sZ * zi + sT * ti;
    return (((d[i] * t0 + d[i + sY] * t1) * s0 +
             (d[i + sX] * t0 + d[i + sX + sY] * t1) * s1) *
                f0 +
            ((d[i + sZ] * t0 + d[i + sY + sZ] * t1) * s0 +
             (d[i + sX + sZ] * t0 + d[i + sX + sY + sZ] * t1) * s1) *
                f1) *
               g0 +
           (((d[i + sT] * t0 + d[i + sT + sY] * t1) * s0 +
             (d[i + sT + sX] * t0 + d[i + sT + sX + sY] * t1) * s1) *
                f0 +
            ((d[i + sT + sZ] * t0 + d[i + sT + sY + sZ] * t1) * s0 +
             (d[i + sT + sX + sZ] * t0 + d[i + sT + sX + sY + sZ] * t1) *
                 s1) *
                f1) *
               g1;
}

// Trilinear sampling, util/interpol.h:57-116 (z is not clamped when nz == 1).
template <class R, class T>
inline T interp3(const T *d, int nx, int ny, int nz, idx_t strideZ, R px_,
                 R py_, R pz_)
{
    R px = px_ - (R)0.5f, py = py_ - (R)0.5f, pz = pz_ - (R)0.5f;
    int xi = (int)px, yi = (int)py, zi = (int)pz;
    R s1 = px - (R)xi, s0 = (R)(1. - s1);
    R t1 = py - (R)yi, t0 = (R)(1. - t1);
    R f1 = pz - (R)zi, f0 = (R)(1. - f1);
    if (px < 0.) { xi = 0; s0 = 1; s1 = 0; }
    if (py < 0.) { yi = 0; t0 = 1; t1 = 0; }
    if (pz < 0.) { zi = 0; f0 = 1; f1 = 0; }
    if (xi >= nx - 1) { xi = nx - 2; s0 = 0; s1 = 1; }
    if (yi >= ny - 1) { yi = ny - 2; t0 = 0; t1 = 1; }
    if (nz > 1) {
        if (zi >= nz - 1) { zi = nz - 2; f0 = 0; f1 = 1; }
    }
    const idx_t X = 1, Y = nx, Z = strideZ;
    const idx_t i = (idx_t)xi + Y * yi + Z * zi;
    return ((d[i] * t0 + d[i + Y] * t1) * s0 +
            (d[i + X] * t0 + d[i + X + Y] * t1) * s1) *
               f0 +
           ((d[i + Z] * t0 + d[i + Y + Z] * t1) * s0 +
            (d[i + X + Z] * t0 + d[i + X + Y + Z] * t1) * s1) *
               f1;
}

// ---------------------------------------------------------------------------
// advection, optflow4d.cpp:1275-1321: a fresh (zeroed) grid receives the
// backtraced samples on the bnd=1 interior and is swapped in.
template <class R, class T>
void advect4d(const V4<R> *vel, T *grid, const Dims &D, R dt)
{
    std::vector<T> fwd(D.n(), T(0));
#pragma omp parallel for schedule(static)
    for (int t = 1; t < D.T - 1; ++t)
        for (int k = 1; k < D.Z - 1; ++k)
            for (int j = 1; j < D.Y - 1; ++j)
                for (int i = 1; i < D.X - 1; ++i) {
                    const V4<R> v = vel[D.at(i, j, k, t)] * dt;
                    fwd[D.at(i, j, k, t)] = interp4<R, T>(
                        grid, D, (R)(i + 0.5f) - v.x, (R)(j + 0.5f) - v.y,
                        (R)(k + 0.5f) - v.z, (R)(t + 0.5f) - v.t);
                }
    std::copy(fwd.begin(), fwd.end(), grid);
}

template <class R> R maxNormVec(const V4<R> *v, idx_t n)
{
    R m = -std::numeric_limits<R>::max();
    for (idx_t i = 0; i < n; ++i) {
        const R s = normSq(v[i]);
        if (s > m) m = s;
    }
    return (R)std::sqrt(m);
}

// advectCflTemplate<Grid4d> + advectCflHelper4d, optflow4d.cpp:1309-1321,
// 2170-2180: advect with a scaled copy of vel, in int(|v|max*dt/cfl)+1
// substeps of dt/steps.
template <class R, class T>
void advectCfl4d(const V4<R> *vel, T *grid, const Dims &D, R cfl,
                 R velFactor, R dt, bool image = false)
{
    std::vector<V4<R>> vt(vel, vel + D.n());
    for (auto &v : vt) v = v * V4<R>(velFactor);
    const R maxVel = maxNormVec(vt.data(), D.n()) * dt;
    const int steps = int(maxVel / cfl) + 1;
    const R sdt = dt / R(steps);
    for (int s = 0; s < steps; ++s) {
        advect4d<R, T>(vt.data(), grid, D, sdt);
        if constexpr (std::is_same<T, R>::value)
            if (image) img_round_all<R>(grid, D.n());
    }
}

// ---------------------------------------------------------------------------
// resampling, grid4d.h:463-471 and grid4d.cpp:531-569
template <class R, class T>
void resample4d(T *dst, const Dims &Dd, const T *src, const Dims &Ds,
                const V4<R> &fac, const V4<R> &off)
{
#pragma omp parallel for schedule(static)
    for (int t = 0; t < Dd.T; ++t)
        for (int k = 0; k < Dd.Z; ++k)
            for (int j = 0; j < Dd.Y; ++j)
                for (int i = 0; i < Dd.X; ++i) {
                    const V4<R> p =
                        V4<R>((R)i, (R)j, (R)k, (R)t) * fac + off;
                    dst[Dd.at(i, j, k, t)] =
                        interp4<R, T>(src, Ds, p.x, p.y, p.z, p.t);
                }
}

// gridFactor4d, grid4d.cpp:559-569
template <class R>
void gridFactor4d(V4<R> s1, V4<R> s2, const V4<R> &optSize, const V4<R> &scale,
                  V4<R> &srcFac, V4<R> &retOff)
{
    for (int c = 0; c < 4; ++c)
        if (optSize[c] > 0.) s2[c] = optSize[c];
    V4<R> f(s1.x / s2.x, s1.y / s2.y, s1.z / s2.z, s1.t / s2.t);
    srcFac = V4<R>(f.x / scale.x, f.y / scale.y, f.z / scale.z, f.t / scale.t);
    retOff = V4<R>(-retOff.x, -retOff.y, -retOff.z, -retOff.t) * srcFac +
             muld(srcFac, 0.5);
}

// interpolGridTempl<Grid4d>, optflow4d.cpp:40-57: factor = src/dst size,
// offset = factor/2 (restriction = 16-cell mean, prolongation 3/4,1/4).
template <class R, class T>
void pyramidResample(T *dst, const Dims &Dd, const T *src, const Dims &Ds)
{
    V4<R> f((R)Ds.X / Dd.X, (R)Ds.Y / Dd.Y, (R)Ds.Z / Dd.Z, (R)Ds.T / Dd.T);
    resample4d<R, T>(dst, Dd, src, Ds, f, muld(f, 0.5));
}

// setBound / setBoundNeumann, grid4d.cpp:355-412
template <class T> void setBound(T *g, const Dims &D, T value, int w)
{
#pragma omp parallel for schedule(static)
    for (int t = 0; t < D.T; ++t)
        for (int k = 0; k < D.Z; ++k)
            for (int j = 0; j < D.Y; ++j)
                for (int i = 0; i < D.X; ++i) {
                    bool b = i <= w || i >= D.X - 1 - w || j <= w ||
                             j >= D.Y - 1 - w || k <= w || k >= D.Z - 1 - w ||
                             t <= w || t >= D.T - 1 - w;
                    if (b) g[D.at(i, j, k, t)] = value;
                }
}
template <class T> void setBoundNeumann(T *g, const Dims &D, int w)
{
    // sources lie strictly inside the w band, so the in-place update is
    // order independent (as the reference's parallel kernel assumes)
#pragma omp parallel for schedule(static)
    for (int t = 0; t < D.T; ++t)
        for (int k = 0; k < D.Z; ++k)
            for (int j = 0; j < D.Y; ++j)
                for (int i = 0; i < D.X; ++i) {
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
                    if (set) g[D.at(i, j, k, t)] = g[D.at(si, sj, sk, st)];
                }
}

// ---------------------------------------------------------------------------
// sign-mismatch error, optflow4d.cpp:895-927 (returned as Real)
template <class R>
R lsDiff(const R *a, const R *b, R *out, const Dims &D, R corr, int bnd)
{
    DSum acc;
    for (int t = bnd; t < D.T - bnd; ++t)
        for (int k = bnd; k < D.Z - bnd; ++k)
            for (int j = bnd; j < D.Y - bnd; ++j)
                for (int i = bnd; i < D.X - bnd; ++i) {
                    const idx_t c = D.at(i, j, k, t);
                    if ((a[c] < 0. && b[c] < 0.) ||
                        (a[c] >= 0. && b[c] >= 0.)) {
                        if (out) out[c] = 0;
                    } else {
                        R d = std::fabs(a[c] - b[c]) * corr;
                        if (d > 1.) d = 1.;
                        acc.add(d);
                        if (out) out[c] = d;
                    }
                }
    int sx = D.X - 2 * bnd, sy = D.Y - 2 * bnd, sz = D.Z - 2 * bnd,
        st = D.T - 2 * bnd;
    double accu = acc.value();
    accu *= 1000.;
    if (D.T > 1) accu *= 1000.;
    accu *= 1. / double(sx * sy * sz * st);
    return (R)accu;
}

// ---------------------------------------------------------------------------
// Gaussian blur, optflow4d.cpp:127-174: normalised (2s+1)^4 window clipped to
// the grid, bnd=1 writes into a zeroed temp that is swapped in, 2*iter passes.
template <class R, class T>
void gaussianBlur(T *a, const Dims &D, R sigma, int iter)
{
    int s = int(1. * sigma + 0.5);
    if (s == 0) s = 1;
    std::vector<T> tmpv(D.n(), T(0));
    T *cur = a, *tmp = tmpv.data();
    for (int pass = 0; pass < 2 * iter; ++pass) {
#pragma omp parallel for schedule(static)
        for (int t = 1; t < D.T - 1; ++t)
            for (int k = 1; k < D.Z - 1; ++k)
                for (int j = 1; j < D.Y - 1; ++j)
                    for (int i = 1; i < D.X - 1; ++i) {
                        T val(0);
                        R weight = 0;
                        for (int vt = t - s; vt <= t + s; ++vt)
                            for (int zk = k - s; zk <= k + s; ++zk)
                                for (int yj = j - s; yj <= j + s; ++yj)
                                    for (int xi = i - s; xi <= i + s; ++xi) {
                                        if (!D.inb(xi, yj, zk, vt, 0)) continue;
                                        R d = (R)((xi - i) * (xi - i) +
                                                  (yj - j) * (yj - j));
                                        d += (R)((zk - k) * (zk - k));
                                        d += (R)((vt - t) * (vt - t));
                                        const R w = (R)std::exp(
                                            -(double)d /
                                            (2. * (double)sigma * (double)sigma));
                                        weight += w;
                                        val = val + cur[D.at(xi, yj, zk, vt)] * w;
                                    }
                        const idx_t c = D.at(i, j, k, t);
                        if (weight > Eps<R>::v)
                            tmp[c] = divs(val, weight);
                        else
                            tmp[c] = cur[c];
                    }
        std::swap(cur, tmp);
    }
    // the reference swaps data pointers; the caller's grid ends up holding
    // whichever buffer `cur` is now
    if (cur != a) std::copy(cur, cur + D.n(), a);
}


// ---------------------------------------------------------------------------
// One-level optical flow, optflow4d.cpp:361-553.  The 4N x 4N system is
// assembled as the reference stores it (per row: one off-diagonal weight and
// a 4-wide block row), and solved by the reference's Jacobi-PCG
// (optflow4d.cpp:211-355) with FP64 dot products of FP32 products, FP64 axpy
// scalars stored back to R, and the signed-max stopping rule.
struct CgResult {
    int iters;
    double residual;  // ret_residual (stored as Real by the reference)
    int status;       // 0 converged, 1 not converged, 2 sigma 0/NaN, 3 tiny rhs
    int warm;         // 1 when a warm start was used (not the cold fallback)
};

template <class R>
static void applyMat(const std::vector<R> &offd, const std::vector<R> &blockd,
                     const idx_t off[8], const std::vector<R> &x,
                     std::vector<R> &y)
{
    const idx_t n = (idx_t)x.size();
    for (idx_t i = 0; i < n; ++i) {
        R v = 0;
        const R od = offd[i];
        if (od != 0.)
            for (int m = 0; m < 8; ++m) v += od * x[i + off[m]];
        const idx_t b = (i / 4) * 4;
        for (int m = 0; m < 4; ++m) v += blockd[i * 4 + m] * x[b + m];
        y[i] = v;
    }
}
template <class R>
static double dotProd(const std::vector<R> &a, const std::vector<R> &b)
{
    DSum d;
    for (size_t i = 0; i < a.size(); ++i) d.add(a[i] * b[i]);
    return d.value();
}
template <class R> static R maxSigned(const std::vector<R> &a)
{
    R m = a[0];
    for (size_t i = 1; i < a.size(); ++i)
        if (a[i] > m) m = a[i];
    return m;
}
template <class R>
static void addScaled(std::vector<R> &a, double b, const std::vector<R> &c)
{
    for (size_t i = 0; i < a.size(); ++i) a[i] += b * c[i];
}

template <class R>
static CgResult pcg(const std::vector<R> &offd, const std::vector<R> &blockd,
                    const idx_t off[8], const std::vector<R> &rhs,
                    std::vector<R> &x, R accuracy, int maxIter,
                    const std::vector<R> *x0)
{
    CgResult out{0, 1e10, 1, 0};
    const size_t n = rhs.size();
    std::vector<R> srch(n), res(n), prec(n), tmp(n);
    std::fill(x.begin(), x.end(), R(0));
    res = rhs;
    double residual = maxSigned(res);
    if (residual < Eps<R>::v) {
        out.iters = 0;
        out.residual = 0;
        out.status = 3;
        return out;
    }
    const double acc = accuracy * residual;
    const double resIni = residual;
    if (x0) {
        // Warm start (extension): the guess w = x0 is scaled by the
        // energy-optimal factor gamma = (b.w) / (w.A w) (the multiple of w
        // closest to the solution in the A-norm, 0 if w.A w is not positive),
        // then x = gamma w and r = b - gamma (A w), kept only when
        // ||r||_2 <= ||b||_2, else the cold start x = 0 is used.  The stopping
        // threshold stays relative to the cold-start residual.
        const std::vector<R> &w = *x0;
        applyMat(offd, blockd, off, w, tmp);
        DSum bw, waw, bb;
        for (size_t i = 0; i < n; ++i) {
            bw.add((double)rhs[i] * (double)w[i]);
            waw.add((double)w[i] * (double)tmp[i]);
            bb.add((double)rhs[i] * (double)rhs[i]);
        }
        const double wAw = waw.value();
        double gamma = 0.;
        if (wAw > 0. && wAw < HUGE_VAL) gamma = bw.value() / wAw;
        if (gamma != gamma) gamma = 0.;
        std::vector<R> xw(n), rw(n);
        DSum r2;
        for (size_t i = 0; i < n; ++i) {
            xw[i] = (R)(gamma * (double)w[i]);
            rw[i] = (R)((double)rhs[i] - gamma * (double)tmp[i]);
            r2.add((double)rw[i] * (double)rw[i]);
        }
        if (r2.value() <= bb.value()) {
            out.warm = 1;
            x = xw;
            res = rw;
            residual = maxSigned(res);
            if (residual <= acc) {
                out.iters = 0;
                out.residual = (R)(residual / resIni);
                out.status = 0;
                return out;
            }
        }
    }
    for (size_t i = 0; i < n; ++i) {
        const R diag = blockd[i * 4 + (i & 3)];
        prec[i] = (R)(1. / diag);
    }
    for (size_t i = 0; i < n; ++i) tmp[i] = res[i] * prec[i];
    double sigma = dotProd(tmp, res);
    if (sigma == 0 || sigma != sigma) {
        out.iters = 0;
        out.status = 2;
        return out;
    }
    srch = tmp;
    if (g_pstore)
        for (auto &v : srch) v = round_p(v);
    for (int it = 0; it < maxIter; ++it) {
        applyMat(offd, blockd, off, srch, tmp);
        const double alpha = sigma / dotProd(srch, tmp);
        addScaled(x, alpha, srch);
        addScaled(res, -alpha, tmp);
        residual = maxSigned(res);
        out.residual = (R)(residual / resIni);
        out.iters = it + 1;
        if (residual <= acc) {
            out.status = 0;
            return out;
        }
        for (size_t i = 0; i < n; ++i) tmp[i] = res[i] * prec[i];
        const double sigmaNew = dotProd(tmp, res);
        const double beta = sigmaNew / sigma;
        addScaled(tmp, beta, srch);
        srch = tmp;
        if (g_pstore)
            for (auto &v : srch) v = round_p(v);
        sigma = sigmaNew;
    }
    out.residual = (R)(residual / resIni);
    out.status = 1;
    return out;
}

struct OfParams {
    double wSmooth, wEnergy, postVelBlur, cgAccuracy, cfl, resetBndWidth;
    int level, blurType, multiStep, projSizeThresh, minGridSize,
        doFinalProject, cgMaxIter;
};

static OfParams fromC(const flof_of_params *p)
{
    OfParams o;
    o.wSmooth = p->wSmooth;
    o.wEnergy = p->wEnergy;
    o.postVelBlur = p->postVelBlur;
    o.cgAccuracy = p->cgAccuracy;
    o.cfl = p->cfl;
    o.resetBndWidth = p->resetBndWidth;
    o.level = p->level;
    o.blurType = p->blurType;
    o.multiStep = p->multiStep;
    o.projSizeThresh = p->projSizeThresh;
    o.minGridSize = p->minGridSize;
    o.doFinalProject = p->doFinalProject;
    o.cgMaxIter = p->cgMaxIter;
    return o;
}

// opticalFlowDim<..., 4>: vel is the incoming flow (used by the rhs) and
// receives the solution of this level.
template <class R>
CgResult opticalFlowLevel(V4<R> *vel, const R *i0, const R *i1, R *rhsT,
                          const Dims &D, R wSmooth, R wEnergy, R postVelBlur,
                          R cgAccuracy, R resetBndWidth, R dt, int maxIter,
                          const V4<R> *warm)
{
    const R mDt = dt;
    const R mDx = (R)(1. / D.X);
    const R mDx2Inv = 1;
    const idx_t nc = D.n(), N = nc * 4;
    std::vector<R> offd(N, R(0)), blockd(N * 4, R(0)), rhs(N, R(0)),
        result(N, R(0));
    static const int nbx[8] = {0, 0, 0, -1, +1, 0, 0, 0};
    static const int nby[8] = {0, 0, -1, 0, 0, +1, 0, 0};
    static const int nbz[8] = {0, -1, 0, 0, 0, 0, +1, 0};
    static const int nbt[8] = {-1, 0, 0, 0, 0, 0, 0, +1};
    idx_t off[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    {
        const idx_t base = D.at(1, 1, 1, 1) * 4;
        for (int m = 0; m < 8; ++m) {
            const int ti = 1 + nbx[m], tj = 1 + nby[m], tk = 1 + nbz[m],
                      tt = 1 + nbt[m];
            if (D.inb(ti, tj, tk, tt, 0)) off[m] = D.at(ti, tj, tk, tt) * 4 - base;
        }
    }
    const R dxf = (R)(1. / (2. * (double)mDx));
    for (int t = 0; t < D.T; ++t)
        for (int k = 0; k < D.Z; ++k)
            for (int j = 0; j < D.Y; ++j)
                for (int i = 0; i < D.X; ++i) {
                    const idx_t c = D.at(i, j, k, t), row = c * 4;
                    if (!D.inb(i, j, k, t, 1)) {
                        for (int d = 0; d < 4; ++d) {
                            result[row + d] = 0;
                            blockd[(row + d) * 4 + d] = 1;
                        }
                        continue;
                    }
                    const R tderiv = (i1[c] - i0[c]) / mDt;
                    R g[4];
                    g[0] = grad_round<R>((i1[c + 1] - i1[c - 1]) * dxf);
                    g[1] = grad_round<R>((i1[c + D.X] - i1[c - D.X]) * dxf);
                    g[2] = grad_round<R>((i1[c + D.sz()] - i1[c - D.sz()]) * dxf);
                    g[3] = grad_round<R>((i1[c + D.st()] - i1[c - D.st()]) * dxf);
                    for (int d = 0; d < 4; ++d) {
                        R r = -g[d] * tderiv;
                        for (int m = 0; m < 8; ++m) {
                            const int ti = std::max(0, std::min(D.X - 1, i + nbx[m]));
                            const int tj = std::max(0, std::min(D.Y - 1, j + nby[m]));
                            const int tk = std::max(0, std::min(D.Z - 1, k + nbz[m]));
                            const int tt = std::max(0, std::min(D.T - 1, t + nbt[m]));
                            r -= wSmooth *
                                 (vel[c][d] - vel[D.at(ti, tj, tk, tt)][d]) *
                                 mDx * mDx2Inv;
                        }
                        offd[row + d] = -wSmooth * mDx2Inv;
                        R diag = 0;
                        diag += (R)(8) * wSmooth * mDx2Inv;
                        diag += wEnergy;
                        for (int m = 0; m < 4; ++m)
                            blockd[(row + d) * 4 + m] =
                                (d == m) ? g[d] * g[m] + diag : g[d] * g[m];
                        r -= wEnergy * vel[c][d] * mDx;
                        rhs[row + d] = r;
                    }
                }
    std::vector<R> x0;
    if (warm) {
        x0.assign(N, R(0));
        for (idx_t c = 0; c < nc; ++c)
            for (int d = 0; d < 4; ++d) x0[c * 4 + d] = warm[c][d] * mDx;
        // identity rows keep a zero guess
        for (int t = 0; t < D.T; ++t)
            for (int k = 0; k < D.Z; ++k)
                for (int j = 0; j < D.Y; ++j)
                    for (int i = 0; i < D.X; ++i)
                        if (!D.inb(i, j, k, t, 1))
                            for (int d = 0; d < 4; ++d)
                                x0[D.at(i, j, k, t) * 4 + d] = 0;
    }
    CgResult cg = pcg(offd, blockd, off, rhs, result, cgAccuracy, maxIter,
                      warm ? &x0 : nullptr);
    if (cg.residual != cg.residual)
        std::fill(result.begin(), result.end(), R(0));
    for (idx_t c = 0; c < nc; ++c) {
        for (int d = 0; d < 4; ++d) vel[c][d] = result[c * 4 + d] / mDx;
        if (rhsT) rhsT[c] = rhs[c * 4];
    }
    if (postVelBlur > 0.) gaussianBlur<R, V4<R>>(vel, D, (R)(0.5 * postVelBlur), 1);
    if (resetBndWidth > 0.) {
        const int rb = int(resetBndWidth * (R)D.X) + 1;
        for (int t = 0; t < D.T; ++t)
            for (int k = 0; k < D.Z; ++k)
                for (int j = 0; j < D.Y; ++j)
                    for (int i = 0; i < D.X; ++i)
                        if (!D.inb(i, j, k, t, rb)) vel[D.at(i, j, k, t)] = V4<R>(0);
    }
    return cg;
}

// ---------------------------------------------------------------------------
// level-set projection, optflow4d.cpp:613-626,651-802
template <class R>
inline bool inBoundsPos(const Dims &D, const V4<R> &p)
{
    // Grid4dBase::isInBounds(Vec4) truncates toward zero (vector4d.h:394-397)
    const int i = (int)p.x, j = (int)p.y, k = (int)p.z, t = (int)p.t;
    return i >= 0 && j >= 0 && k >= 0 && t >= 0 && i < D.X && j < D.Y &&
           k < D.Z && t < D.T;
}

template <class R>
inline V4<R> normalInterp(const R *phi, const Dims &D, const V4<R> &p)
{
    const R h = (R)0.5;
    V4<R> n;
    n.x = interp4<R, R>(phi, D, p.x + h, p.y, p.z, p.t) -
          interp4<R, R>(phi, D, p.x - h, p.y, p.z, p.t);
    n.y = interp4<R, R>(phi, D, p.x, p.y + h, p.z, p.t) -
          interp4<R, R>(phi, D, p.x, p.y - h, p.z, p.t);
    n.z = interp4<R, R>(phi, D, p.x, p.y, p.z + h, p.t) -
          interp4<R, R>(phi, D, p.x, p.y, p.z - h, p.t);
    n.t = interp4<R, R>(phi, D, p.x, p.y, p.z, p.t + h) -
          interp4<R, R>(phi, D, p.x, p.y, p.z, p.t - h);
    // normalize(), vector4d.h:331-345
    const R l = normSq(n);
    const R e2 = Eps<R>::v * Eps<R>::v;
    if (std::fabs((double)l - 1.) < e2) {
    } else if (l > e2) {
        const R nrm = (R)std::sqrt(l);
        n = n * (R)(1. / nrm);
    } else {
        n = V4<R>(0);
    }
    return n;
}

template <class R>
inline R projectCell(const V4<R> &normal, const R *phiOrg, const R *phiTarget,
                     const Dims &D, const V4<R> &pos, idx_t cell, int maxIter,
                     R dt)
{
    R step = (R)0.25;
    if (maxIter < 50) step = (R)0.5;
    int lastDir = 0;
    const R target = phiTarget[cell];
    R vn = interp4<R, R>(phiOrg, D, pos.x, pos.y, pos.z, pos.t);
    R d = vn - target;
    for (int s = 0; s < maxIter; ++s) {
        const V4<R> pn = pos + (normal * d) * dt;
        if (!inBoundsPos(D, pn)) {
            s = maxIter;
            d = (R)9999.;
            continue;
        }
        vn = interp4<R, R>(phiOrg, D, pn.x, pn.y, pn.z, pn.t);
        if (vn < target) {
            if (lastDir < 0) step = (R)(step * 0.5);
            d += step;
            lastDir = 1;
        } else {
            if (lastDir > 0) step = (R)(step * 0.5);
            d -= step;
            lastDir = -1;
        }
        if (step < 1e-01) s = maxIter + 1;
    }
    return d;
}

template <class R>
void corrVelsOf(V4<R> *dst, V4<R> *vel, const R *phiOrg, const R *phiTarget,
                const Dims &D, R threshPhi, R postVelBlur, R resetBndWidth,
                int maxIter, R dt)
{
    const R blurThreshold = (R)0.98;
    const int doCvExtrapol = (int)(maxNormVec(vel, D.n()) + 4);
    std::vector<R> marker(D.n(), R(0));
    std::vector<V4<R>> tmp(D.n());
    int singleStep = 0;
    if (postVelBlur < blurThreshold) singleStep = 1;
    while ((postVelBlur >= blurThreshold) || singleStep) {
        std::fill(dst, dst + D.n(), V4<R>(0));
        std::fill(marker.begin(), marker.end(), R(0));
#pragma omp parallel for schedule(dynamic)
        for (int t = 2; t < D.T - 2; ++t)
            for (int k = 2; k < D.Z - 2; ++k)
                for (int j = 2; j < D.Y - 2; ++j)
                    for (int i = 2; i < D.X - 2; ++i) {
                        const idx_t c = D.at(i, j, k, t);
                        const V4<R> pos =
                            V4<R>((R)(i + 0.5f), (R)(j + 0.5f), (R)(k + 0.5f),
                                  (R)(t + 0.5f)) -
                            vel[c] * dt;
                        const V4<R> n = normalInterp(phiOrg, D, pos);
                        const R d = projectCell(n, phiOrg, phiTarget, D, pos, c,
                                                maxIter, dt);
                        if (std::fabs(d) > threshPhi) {
                            dst[c] = V4<R>(0);
                            continue;
                        }
                        dst[c] = n * d;
                        marker[c] = 1;
                    }
        for (int it = 0; it < doCvExtrapol; ++it) {
            std::copy(dst, dst + D.n(), tmp.begin());
#pragma omp parallel for schedule(static)
            for (int t = 1; t < D.T - 1; ++t)
                for (int k = 1; k < D.Z - 1; ++k)
                    for (int j = 1; j < D.Y - 1; ++j)
                        for (int i = 1; i < D.X - 1; ++i) {
                            const idx_t c = D.at(i, j, k, t);
                            if (marker[c] != 0.) continue;
                            V4<R> val(0);
                            for (int vt = t - 1; vt <= t + 1; ++vt)
                                for (int zk = k - 1; zk <= k + 1; ++zk)
                                    for (int yj = j - 1; yj <= j + 1; ++yj)
                                        for (int xi = i - 1; xi <= i + 1; ++xi)
                                            val = val + dst[D.at(xi, yj, zk, vt)];
                            tmp[c] = muld(val, 1. / 81.0);
                        }
            std::copy(tmp.begin(), tmp.end(), dst);
        }
        if (postVelBlur > Eps<R>::v)
            gaussianBlur<R, V4<R>>(dst, D, (R)(0.5 * postVelBlur), 1);
        const int resetBnd = int(resetBndWidth * (R)D.X) + 1;
        if (resetBndWidth > 0.) setBound<V4<R>>(dst, D, V4<R>(0), resetBnd);
        for (idx_t c = 0; c < D.n(); ++c) vel[c] = vel[c] + V4<R>(-1) * dst[c];
        postVelBlur = (R)(postVelBlur * 0.5);
        if (singleStep > 0) singleStep--;
    }
}

// ---------------------------------------------------------------------------
// multiscale driver, optflow4d.cpp:936-1173
template <class R>
R multiscale(V4<R> *vel, const R *i0, const R *i1, R *rhsT, const Dims &D,
             const OfParams &P, int level, R dt, flof_solve_info *info,
             const V4<R> *warm)
{
    const R wSmooth = (R)P.wSmooth, wEnergy = (R)P.wEnergy,
            postVelBlur = (R)P.postVelBlur, cgAccuracy = (R)P.cgAccuracy,
            cfl = (R)P.cfl, resetBndWidth = (R)P.resetBndWidth;
    int multiStep = P.multiStep;
    bool doFinalProject = P.doFinalProject != 0;
    const int resetBnd = resetBndWidth > 0 ? int(resetBndWidth * (R)D.X) + 1 : 0;
    const R projMaxDist = (R)4.;
    const R projMaxIter = (R)40.;
    const R lsDiffFac = (R)(0.1 / 20.);
    const idx_t n = D.n();

    if (info && level < FLOF_MAX_LEVELS) {
        info->dims[level][0] = D.X;
        info->dims[level][1] = D.Y;
        info->dims[level][2] = D.Z;
        info->dims[level][3] = D.T;
        if (level + 1 > info->levels) info->levels = level + 1;
    }
    std::vector<R> i0warped(i0, i0 + n);
    const R errPreOf = lsDiff<R>(i0, i1, nullptr, D, lsDiffFac, resetBnd);
    if (info && level == 0) info->err_pre = errPreOf;

    if (D.X > P.minGridSize) {
        Dims Ds{D.X / 2, D.Y / 2, D.Z / 2, D.T / 2};
        std::vector<R> i0s(Ds.n()), i1s(Ds.n()), dbg(Ds.n());
        std::vector<V4<R>> vs(Ds.n()), ws;
        pyramidResample<R, R>(i0s.data(), Ds, i0, D);
        pyramidResample<R, R>(i1s.data(), Ds, i1, D);
        img_round_all<R>(i0s.data(), Ds.n());
        img_round_all<R>(i1s.data(), Ds.n());
        pyramidResample<R, V4<R>>(vs.data(), Ds, vel, D);
        for (auto &v : vs) v = v * V4<R>((R)0.5);
        if (warm) {
            // the warm flow is a guess of the total flow; each level's CG
            // receives (warm - incoming vel) as the guess of its increment
            ws.resize(Ds.n());
            pyramidResample<R, V4<R>>(ws.data(), Ds, warm, D);
            for (auto &v : ws) v = v * V4<R>((R)0.5);
        }
        multiscale<R>(vs.data(), i0s.data(), i1s.data(), dbg.data(), Ds, P,
                      level + 1, dt, info, warm ? ws.data() : nullptr);
        pyramidResample<R, V4<R>>(vel, D, vs.data(), Ds);
        for (idx_t c = 0; c < n; ++c) vel[c] = vel[c] * V4<R>((R)2.);
        if (rhsT) std::copy(i0warped.begin(), i0warped.end(), rhsT);
    }

    advectCfl4d<R, R>(vel, i0warped.data(), D, cfl, (R)1., dt, true);
    setBoundNeumann<R>(i0warped.data(), D, 0);
    R errCurr = lsDiff<R>(i0warped.data(), i1, nullptr, D, lsDiffFac, resetBnd);
    if (info && level < FLOF_MAX_LEVELS) info->err_level[level] = errCurr;

    bool doProject = false;
    if (D.X > P.projSizeThresh) {
        doProject = true;
        multiStep = 1;
        if (doFinalProject) doFinalProject = false;
    }

    auto record = [&](int step, const CgResult &cg) {
        if (!info || level >= FLOF_MAX_LEVELS || step >= FLOF_MAX_STEPS) return;
        info->iters[level][step] = cg.iters;
        info->residual[level][step] = cg.residual;
        info->status[level][step] = cg.status;
        info->warm_used[level] = cg.warm;
        if (step + 1 > info->steps[level]) info->steps[level] = step + 1;
        info->total_iters += cg.iters;
    };

    if (multiStep > 1) {
        std::vector<std::vector<V4<R>>> vsv(multiStep), vs2(multiStep);
        for (int s = 0; s < multiStep; ++s) vs2[s].assign(n, V4<R>(0));
        std::vector<V4<R>> tmpVel(vel, vel + n);
        R velBlur = postVelBlur;
        R errLast = errCurr;
        int steps = multiStep;
        for (int of = 0; of < steps; ++of) {
            vsv[of].assign(n, V4<R>(0));
            CgResult cg = opticalFlowLevel<R>(
                vsv[of].data(), i0warped.data(), i1, rhsT, D, wSmooth, wEnergy,
                velBlur, cgAccuracy, resetBndWidth, dt, P.cgMaxIter, nullptr);
            record(of, cg);
            velBlur = (R)(velBlur * (3. / 4.));
            if (velBlur < 2.) velBlur = 2.;
            for (int k = of; k >= 0; --k) vs2[k] = vsv[k];
            for (int k = of - 1; k >= 0; --k)
                for (int l = 0; l < k; ++l)
                    advectCfl4d<R, V4<R>>(vs2[k].data(), vs2[l].data(), D, cfl,
                                          (R)1., dt);
            std::copy(vel, vel + n, tmpVel.begin());
            for (int k = of; k >= 0; --k)
                for (idx_t c = 0; c < n; ++c) tmpVel[c] = tmpVel[c] + vs2[k][c];
            std::vector<R> w2(i0, i0 + n);
            advectCfl4d<R, R>(tmpVel.data(), w2.data(), D, cfl, (R)1., dt, true);
            setBoundNeumann<R>(w2.data(), D, 0);
            const R errC = lsDiff<R>(w2.data(), i1, nullptr, D, lsDiffFac, resetBnd);
            i0warped = w2;
            if (of > 0 && (errC / errLast) > 0.95) {
                std::fill(vsv[of].begin(), vsv[of].end(), V4<R>(0));
                steps = of + 1;
            }
            errLast = errC;
        }
        for (int of = steps - 1; of >= 0; --of)
            for (int l = 0; l < of; ++l)
                advectCfl4d<R, V4<R>>(vsv[of].data(), vsv[l].data(), D, cfl,
                                      (R)1., dt);
        for (int of = 0; of < steps; ++of)
            for (idx_t c = 0; c < n; ++c) vel[c] = vel[c] + vsv[of][c];
    } else {
        std::vector<V4<R>> velCurr(n, V4<R>(0));
        if (!doProject) {
            std::vector<V4<R>> guess;
            if (warm) {
                guess.resize(n);
                for (idx_t c = 0; c < n; ++c) guess[c] = warm[c] - vel[c];
            }
            CgResult cg = opticalFlowLevel<R>(
                velCurr.data(), i0warped.data(), i1, rhsT, D, wSmooth, wEnergy,
                postVelBlur, cgAccuracy, resetBndWidth, dt, P.cgMaxIter,
                warm ? guess.data() : nullptr);
            record(0, cg);
            for (idx_t c = 0; c < n; ++c) vel[c] = vel[c] + velCurr[c];
        } else {
            std::vector<V4<R>> velTmp2(n, V4<R>(0));
            corrVelsOf<R>(velCurr.data(), velTmp2.data(), i0warped.data(), i1, D,
                          projMaxDist, postVelBlur, resetBndWidth,
                          (int)projMaxIter, dt);
            for (idx_t c = 0; c < n; ++c) vel[c] = vel[c] + velTmp2[c];
            for (idx_t c = 0; c < n; ++c)
                velCurr[c] = velCurr[c] * V4<R>((R)-1.);
            for (idx_t c = 0; c < n; ++c) vel[c] = vel[c] + velCurr[c];
        }
    }

    if (level == 0 && doFinalProject) {
        std::vector<V4<R>> velCurr(n, V4<R>(0));
        const R finalProjBlur = (R)4.;
        corrVelsOf<R>(velCurr.data(), vel, i0, i1, D, projMaxDist, finalProjBlur,
                      resetBndWidth, (int)projMaxIter, dt);
    }
    if (level == 0) {
        std::copy(i0, i0 + n, i0warped.begin());
        advectCfl4d<R, R>(vel, i0warped.data(), D, cfl, (R)1., dt, true);
        setBoundNeumann<R>(i0warped.data(), D, 0);
        errCurr = lsDiff<R>(i0warped.data(), i1, nullptr, D, lsDiffFac, resetBnd);
        if (info) info->err_final = errCurr;
    }
    return errCurr;
}

// ---------------------------------------------------------------------------
// frame placement, optflow4d.cpp:1445-1595 (frames given in memory; frame f
// holds file index fileIdxStart + f)
template <class R>
void placeFrames(R *phi, const Dims &D, const R *frames, int sx, int sy, int sz,
                 const V4<R> &offset, const V4<R> &scale, int fileIdxStart,
                 int fileIdxEnd, int debugSkipLoad, R spread,
                 const V4<R> &overrideSize, R overrideTimeOff,
                 int overrideGoodRegion, R loadTimeScale, bool rescaleSdfValues,
                 R sdfIsoOff, R repeatStartFrame)
{
    V4<R> defoSize((R)D.X, (R)D.Y, (R)D.Z, (R)D.T);
    if (overrideSize.x > 0.) defoSize = overrideSize;
    V4<R> sf, off2 = offset;
    gridFactor4d<R>(V4<R>((R)sx, (R)sy, (R)sz, (R)1), defoSize, V4<R>((R)-1.),
                    scale, sf, off2);
    R valueScale = 1;
    if (rescaleSdfValues) valueScale = (R)(1. / sf.x);
    std::vector<R> w(D.T, R(0));
    if (overrideGoodRegion <= 0)
        std::fill(phi, phi + D.n(), R(0));
    else
        std::fill(phi + D.st() * overrideGoodRegion, phi + D.n(), R(0));
    if (fileIdxStart < 0) fileIdxStart = 0;
    if (fileIdxEnd < 0) fileIdxEnd = (int)defoSize.t;
    int repeatOff = 0;
    if (repeatStartFrame > 0)
        repeatOff = int((fileIdxEnd - fileIdxStart) * repeatStartFrame + 1.0);
    const idx_t fsz = (idx_t)sx * sy * sz;
    std::vector<R> tmp(fsz);
    for (int fileid = fileIdxStart - repeatOff;
         fileid < std::min(fileIdxEnd, debugSkipLoad); ++fileid) {
        const int fc = fileid > fileIdxStart ? fileid : fileIdxStart;
        bool loaded = false;
        const R duration = (R)(1. * (defoSize.t * scale.t / loadTimeScale) /
                               R(fileIdxEnd - fileIdxStart));
        const R dstt = offset.t + overrideTimeOff + duration * (R)(fileid - fileIdxStart);
        for (int slice = int(dstt - spread * duration);
             slice < int(dstt + spread * duration) + 2; ++slice) {
            if (slice < overrideGoodRegion || slice >= D.T) continue;
            const R wc = (R)(1. - std::fabs((R)slice - dstt) / (duration * spread));
            if (wc < Eps<R>::v) continue;
            w[slice] += wc;
            if (!loaded) {
                std::copy(frames + (idx_t)(fc - fileIdxStart) * fsz,
                          frames + (idx_t)(fc - fileIdxStart + 1) * fsz,
                          tmp.begin());
                if (rescaleSdfValues && sdfIsoOff != 0.) {
                    for (int k = 1; k < sz - 1; ++k)
                        for (int j = 1; j < sy - 1; ++j)
                            for (int i = 1; i < sx - 1; ++i)
                                tmp[i + (idx_t)sx * j + (idx_t)sx * sy * k] += sdfIsoOff;
                    for (int k = 0; k < sz; ++k)
                        for (int j = 0; j < sy; ++j)
                            for (int i = 0; i < sx; ++i)
                                if (i <= 1 || i >= sx - 2 || j <= 1 || j >= sy - 2 ||
                                    k <= 1 || k >= sz - 2)
                                    tmp[i + (idx_t)sx * j + (idx_t)sx * sy * k] = 1;
                }
                loaded = true;
            }
            R *ph = phi + D.st() * slice;
#pragma omp parallel for schedule(static)
            for (int k = 1; k < D.Z - 1; ++k)
                for (int j = 1; j < D.Y - 1; ++j)
                    for (int i = 1; i < D.X - 1; ++i) {
                        const R px = (R)i * sf.x + off2.x, py = (R)j * sf.y + off2.y,
                                pz = (R)k * sf.z + off2.z;
                        ph[i + (idx_t)D.X * j + D.sz() * k] +=
                            wc * interp3<R, R>(tmp.data(), sx, sy, sz,
                                               (idx_t)sx * sy, px, py, pz);
                    }
        }
    }
    for (int s = 0; s < D.T; ++s) {
        if (w[s] < 1e-03)
            w[s] = 0;
        else {
            w[s] = (R)(1. / w[s]);
            w[s] *= valueScale;
        }
    }
#pragma omp parallel for schedule(static)
    for (int t = 1; t < D.T - 1; ++t) {
        if (w[t] <= 0.) continue;
        for (int k = 1; k < D.Z - 1; ++k)
            for (int j = 1; j < D.Y - 1; ++j)
                for (int i = 1; i < D.X - 1; ++i) phi[D.at(i, j, k, t)] *= w[t];
    }
}

// knSemiLagrangeLookupSlice4d(_WithVel), optflow4d.cpp:1622-1665: vt is the
// velocity slice at deformation resolution; bnd 10 (with the border skip) for
// the cached variant, 1 for the other
template <class R>
void lookupSlice(const V4<R> *vt, const Dims &Dv, R *dst, int dx, int dy, int dz,
                 const R *phi, const Dims &Dp, R tl, R alpha, const V4<R> &sf,
                 const V4<R> &off2, const V4<R> &fac, int bordSkip, bool optimized)
{
    const int bnd = optimized ? 10 : 1;
#pragma omp parallel for schedule(static)
    for (int k = bnd; k < dz - bnd; ++k)
        for (int j = bnd; j < dy - bnd; ++j)
            for (int i = bnd; i < dx - bnd; ++i) {
                if (optimized && !(i >= bordSkip && j >= bordSkip && k >= bordSkip &&
                                   i < dx - bordSkip && j < dy - bordSkip &&
                                   k < dz - bordSkip))
                    continue;
                const R px = (R)i * sf.x + off2.x, py = (R)j * sf.y + off2.y,
                        pz = (R)k * sf.z + off2.z;
                const V4<R> v = interp3<R, V4<R>>(vt, Dv.X, Dv.Y, Dv.Z, Dv.sz(), px, py, pz) *
                                fac;
                const V4<R> pos =
                    V4<R>((R)(i + 0.5f), (R)(j + 0.5f), (R)(k + 0.5f), (R)(tl + 0.5f)) -
                    v * alpha;
                dst[i + (idx_t)dx * j + (idx_t)dx * dy * k] =
                    interp4<R, R>(phi, Dp, pos.x, pos.y, pos.z, pos.t);
            }
}

// frame warp, optflow4d.cpp:1951-2105 (one deformation volume, slices taken
// from the in-memory 4D defo), and its non-cached twin :1671-1760
template <class R>
void warpSlice(const V4<R> *defo, const Dims &Dv, R *dst, int dx, int dy, int dz,
               const R *phi, const Dims &Dp, R time, R blendAlpha,
               R loadTimeScale, const V4<R> &defoOffset, const V4<R> &defoScale,
               const V4<R> &defoFactor, const V4<R> &overrideSize,
               R overrideTimeOff, bool zeroVel, int bordSkip, R defoAniFac,
               bool optimized)
{
    V4<R> sf, off2 = defoOffset;
    int ds[4] = {Dp.X, Dp.Y, Dp.Z, Dp.T};
    if (overrideSize.x > 0.) {
        ds[0] = (int)overrideSize.x;
        ds[1] = (int)overrideSize.y;
        ds[2] = (int)overrideSize.z;
        ds[3] = (int)overrideSize.t;
    }
    gridFactor4d<R>(V4<R>((R)Dv.X, (R)Dv.Y, (R)Dv.Z, (R)Dv.T),
                    V4<R>((R)ds[0], (R)ds[1], (R)ds[2], (R)ds[3]),
                    V4<R>((R)-1.), defoScale, sf, off2);
    const R srcTime = (R)(time * sf.t * loadTimeScale + off2.t - 0.5);
    int t = (int)srcTime;
    int tp1 = t + 1;
    const R tw = srcTime - (R)t;
    t = std::min(t, Dv.T - 1);
    tp1 = std::min(tp1, Dv.T - 1);
    const idx_t ns = Dv.st();
    std::vector<V4<R>> vt(ns);
    const V4<R> f0((R)(1. - tw)), f1(tw);
    for (idx_t c = 0; c < ns; ++c) {
        V4<R> v(0);
        if (optimized) {
            v = v + f0 * defo[ns * t + c];
            v = v + f1 * defo[ns * tp1 + c];
        } else {
            v = defo[ns * t + c] * f0 + defo[ns * tp1 + c] * f1;
        }
        vt[c] = zeroVel ? V4<R>(0) : v;
    }
    lookupSlice<R>(vt.data(), Dv, dst, dx, dy, dz, phi, Dp, time + overrideTimeOff,
                   blendAlpha, sf, off2, defoFactor * V4<R>(defoAniFac), bordSkip,
                   optimized);
}

// Deformation volumes of loadAdvectTimeSlice_OptInit(useDefoVols=true) /
// _OptAdd / _OptRun (optflow4d.cpp:1786-1870, 1871-1929, 2013-2085): each
// volume is seen through a window of W = int(dimT * max(0.2, partialLoadFac))
// slices, slice tl holding file slice clamp(t - W/2 + tl, 0, dimT - 1); the
// velocity slice vt is combined from the volumes at window time
// srcTime - (t - W/2), then the lookup runs with alpha 1.
template <class R>
void warpSliceVols(const std::vector<const V4<R> *> &vols, const Dims &Dv, R *dst, int dx,
                   int dy, int dz, const R *phi, const Dims &Dp, R time, R blendAlpha,
                   R loadTimeScale, const V4<R> &defoOffset, const V4<R> &defoScale,
                   const V4<R> &defoFactor, const V4<R> &overrideSize, R overrideTimeOff,
                   bool zeroVel, int bordSkip, R defoAniFac, bool doAligned,
                   R partialLoadFac, R thirdAlpha, R fourthAlpha)
{
    V4<R> sf, off2 = defoOffset;
    int ds[4] = {Dp.X, Dp.Y, Dp.Z, Dp.T};
    if (overrideSize.x > 0.) {
        ds[0] = (int)overrideSize.x;
        ds[1] = (int)overrideSize.y;
        ds[2] = (int)overrideSize.z;
        ds[3] = (int)overrideSize.t;
    }
    gridFactor4d<R>(V4<R>((R)Dv.X, (R)Dv.Y, (R)Dv.Z, (R)Dv.T),
                    V4<R>((R)ds[0], (R)ds[1], (R)ds[2], (R)ds[3]),
                    V4<R>((R)-1.), defoScale, sf, off2);
    const R srcTime = (R)(time * sf.t * loadTimeScale + off2.t - 0.5);
    int t = (int)srcTime;
    t = std::min(t, Dv.T - 1);
    const R width = std::max((R)0.2, partialLoadFac);
    const int W = (int)((R)Dv.T * width);
    const int currt = W / 2;
    const int dvOff = t - currt;
    const Dims Dw{Dv.X, Dv.Y, Dv.Z, W};
    const idx_t ns = Dv.st();
    std::vector<std::vector<V4<R>>> win(vols.size(), std::vector<V4<R>>(ns * W));
    for (size_t v = 0; v < vols.size(); ++v)
        for (int tl = 0; tl < W; ++tl) {
            const int it = std::min(std::max(t - currt + tl, 0), Dv.T - 1);
            std::copy(vols[v] + ns * it, vols[v] + ns * (it + 1), win[v].begin() + ns * tl);
        }
    std::vector<V4<R>> vt(ns);
    const R pt = srcTime - (R)dvOff;
    const V4<R> h((R)0.5);
    auto get = [&](size_t v, const V4<R> &q) {
        return interp4<R, V4<R>>(win[v].data(), Dw, q.x, q.y, q.z, q.t);
    };
    if (vols.size() == 2 && !doAligned) {
        for (int k = 0; k < Dv.Z; ++k)
            for (int j = 0; j < Dv.Y; ++j)
                for (int i = 0; i < Dv.X; ++i) {
                    const V4<R> p((R)i, (R)j, (R)k, pt);
                    const V4<R> v2 = get(1, p + h);
                    vt[Dv.at(i, j, k, 0)] = get(0, p + h - v2) * blendAlpha + v2 * thirdAlpha;
                }
    } else if (vols.size() == 2) {
        std::vector<V4<R>> dvt(ns * W);
        for (int tt = 0; tt < W; ++tt)
            for (int k = 0; k < Dv.Z; ++k)
                for (int j = 0; j < Dv.Y; ++j)
                    for (int i = 0; i < Dv.X; ++i) {
                        const V4<R> p((R)i, (R)j, (R)k, (R)tt);
                        const V4<R> v2 = get(1, p + h);
                        dvt[Dw.at(i, j, k, tt)] =
                            get(0, p + h - v2) * blendAlpha + v2 * thirdAlpha;
                    }
        const R back = (R)(1. - blendAlpha);
        for (int k = 0; k < Dv.Z; ++k)
            for (int j = 0; j < Dv.Y; ++j)
                for (int i = 0; i < Dv.X; ++i) {
                    const V4<R> p((R)i, (R)j, (R)k, pt);
                    const V4<R> v1 = get(0, p + h) * (R)-1;
                    const V4<R> q = p + h - v1 * back;
                    vt[Dv.at(i, j, k, 0)] = interp4<R, V4<R>>(dvt.data(), Dw, q.x, q.y, q.z, q.t);
                }
    } else if (vols.size() == 3) {
        for (int k = 0; k < Dv.Z; ++k)
            for (int j = 0; j < Dv.Y; ++j)
                for (int i = 0; i < Dv.X; ++i) {
                    const V4<R> p((R)i, (R)j, (R)k, pt);
                    const V4<R> v3 = get(2, p + h);
                    const V4<R> v2 = get(1, p + h - v3);
                    const V4<R> v1 = get(0, p + h - v3 - v2);
                    vt[Dv.at(i, j, k, 0)] =
                        v1 * blendAlpha + v2 * thirdAlpha + v3 * fourthAlpha;
                }
    } else {
        throw std::runtime_error("Code currently only supports 2 deformation volumes");
    }
    if (zeroVel) std::fill(vt.begin(), vt.end(), V4<R>(0));
    lookupSlice<R>(vt.data(), Dims{Dv.X, Dv.Y, Dv.Z, 1}, dst, dx, dy, dz, phi, Dp,
                   time + overrideTimeOff, (R)1, sf, off2, defoFactor * V4<R>(defoAniFac),
                   bordSkip, true);
}

// ---------------------------------------------------------------------------
// Seeded falling-drop SDF (no reference counterpart; replaces the FLIP input
// simulation of scenes/dataGen2Drop.py).  splitmix64 stream `seed`; spheres
// draw (cx, y0, cz[, r]); pool plane at 0.2*Y; fall y0 - 0.2*Y*s^2 with
// s = t/(T-1); union = min; clamp to +-clampCells.  Frames at `scale` times
// the spatial resolution use the same drops in fine-cell units.
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

struct Drop {
    double cx, y0, cz, r;
};
static int makeDrops(Drop *d, int maxd, int X, int Y, int Z, uint64_t seed,
                     int nspheres, double rmin, double rmax)
{
    uint64_t s = seed;
    int n = std::min(nspheres, maxd);
    for (int i = 0; i < n; ++i) {
        d[i].cx = X * (0.3 + 0.4 * u01(s));
        d[i].y0 = Y * (0.6 + 0.2 * u01(s));
        d[i].cz = Z * (0.3 + 0.4 * u01(s));
        d[i].r = (rmax > rmin) ? X * (rmin + (rmax - rmin) * u01(s)) : X * rmin;
    }
    return n;
}

template <class R>
void dropSdf(R *out, int X, int Y, int Z, int T, int scale, uint64_t seed,
             int nspheres, double rmin, double rmax, double clampCells)
{
    Drop d[16];
    const int n = makeDrops(d, 16, X, Y, Z, seed, nspheres, rmin, rmax);
    const int fx = X * scale, fy = Y * scale, fz = Z * scale;
    const double F = scale, g = 0.4 * Y, cl = clampCells * F;
#pragma omp parallel for schedule(static)
    for (int t = 0; t < T; ++t) {
        const double s = T > 1 ? (double)t / (double)(T - 1) : 0.0;
        for (int k = 0; k < fz; ++k)
            for (int j = 0; j < fy; ++j)
                for (int i = 0; i < fx; ++i) {
                    const double px = i + 0.5, py = j + 0.5, pz = k + 0.5;
                    double v = py - 0.2 * Y * F;
                    for (int q = 0; q < n; ++q) {
                        const double cy = d[q].y0 - 0.5 * g * s * s;
                        const double ddx = px - F * d[q].cx, ddy = py - F * cy,
                                     ddz = pz - F * d[q].cz;
                        const double sd =
                            std::sqrt(ddx * ddx + ddy * ddy + ddz * ddz) - F * d[q].r;
                        if (sd < v) v = sd;
                    }
                    if (v > cl) v = cl;
                    if (v < -cl) v = -cl;
                    out[(idx_t)i + (idx_t)fx * j + (idx_t)fx * fy * k +
                        (idx_t)fx * fy * fz * t] = (R)v;
                }
    }
}

// ---------------------------------------------------------------------------
// input preprocessing (optflow4d.cpp:1254-1266, 1323-1406)

// repeatFrame4d: slices t in [int(srct - range + 0.5), int(srct)] take the
// quadrilinear sample at time srct of every (bnd-interior) cell; loops run in
// the reference's order (t, k, j, i) and update phi in place.
template <class R>
void repeatFrame4d(R *phi, const Dims &D, R srct, R range, int bnd)
{
    for (int t = int(srct - range + (R)0.5); t <= int(srct); t++)
        for (int k = bnd; k < D.Z - bnd; ++k)
            for (int j = bnd; j < D.Y - bnd; ++j)
                for (int i = bnd; i < D.X - bnd; ++i)
                    phi[D.at(i, j, k, t)] = interp4<R, R>(
                        phi, D, (R)i + (R)0.5, (R)j + (R)0.5, (R)k + (R)0.5,
                        srct + (R)0.5);
}

// extrap4dLsSimple: mark the inside (phi < 0, or > 0 with `inside`), the
// first layer around it, then `distance - 1` layers each taking the mean of
// the previous layer's neighbours plus `direction`; unreached cells get
// direction*(distance+2).  Neighbour order x-, x+, y-, y+, z-, z+, t-, t+.
template <class R>
void extrap4dLsSimple(R *phi, const Dims &D, int distance, bool inside)
{
    std::vector<int> tmp(D.n(), 0);
    const idx_t off[8] = {-1, 1, -(idx_t)D.X, (idx_t)D.X, -D.sz(), D.sz(), -D.st(), D.st()};
    R direction = 1;
    if (inside) direction = -1;
    for (int t = 1; t < D.T - 1; ++t)
        for (int k = 1; k < D.Z - 1; ++k)
            for (int j = 1; j < D.Y - 1; ++j)
                for (int i = 1; i < D.X - 1; ++i) {
                    const idx_t c = D.at(i, j, k, t);
                    if (!inside ? phi[c] < 0. : phi[c] > 0.) tmp[c] = 1;
                }
    for (int t = 1; t < D.T - 1; ++t)
        for (int k = 1; k < D.Z - 1; ++k)
            for (int j = 1; j < D.Y - 1; ++j)
                for (int i = 1; i < D.X - 1; ++i) {
                    const idx_t c = D.at(i, j, k, t);
                    if (tmp[c]) continue;
                    for (int n = 0; n < 8; ++n)
                        if (tmp[c + off[n]] == 1) {
                            tmp[c] = 2;
                            break;
                        }
                }
    for (int d = 2; d < 1 + distance; ++d) {
        // cells of layer d+1 read only cells of layer d: order independent
#pragma omp parallel for schedule(static)
        for (int t = 1; t < D.T - 1; ++t)
            for (int k = 1; k < D.Z - 1; ++k)
                for (int j = 1; j < D.Y - 1; ++j)
                    for (int i = 1; i < D.X - 1; ++i) {
                        const idx_t c = D.at(i, j, k, t);
                        if (tmp[c] != 0) continue;
                        int nbs = 0;
                        R avg = 0;
                        for (int n = 0; n < 8; ++n)
                            if (tmp[c + off[n]] == d) {
                                avg += phi[c + off[n]];
                                nbs++;
                            }
                        if (nbs > 0) {
                            tmp[c] = d + 1;
                            phi[c] = avg / (R)nbs + direction;
                        }
                    }
    }
    const R set = direction * (R)(distance + 2);
    for (int t = 1; t < D.T - 1; ++t)
        for (int k = 1; k < D.Z - 1; ++k)
            for (int j = 1; j < D.Y - 1; ++j)
                for (int i = 1; i < D.X - 1; ++i) {
                    const idx_t c = D.at(i, j, k, t);
                    if (tmp[c] == 0) phi[c] = set;
                }
}

}  // namespace orc

// ===========================================================================
// extern "C" surface used by the Python checker (oracle/oracle.py)
using namespace orc;

#define ORC_DEFINE(R, SUF)                                                     \
    extern "C" void orc_sample4_real##SUF(const R *data, int X, int Y, int Z,  \
                                          int T, const R *pos, int64_t npts,   \
                                          R *out)                              \
    {                                                                          \
        Dims D{X, Y, Z, T};                                                    \
        for (int64_t p = 0; p < npts; ++p)                                     \
            out[p] = interp4<R, R>(data, D, pos[4 * p], pos[4 * p + 1],        \
                                   pos[4 * p + 2], pos[4 * p + 3]);            \
    }                                                                          \
    extern "C" void orc_sample3_real##SUF(const R *data, int X, int Y, int Z,  \
                                          const R *pos, int64_t npts, R *out)  \
    {                                                                          \
        for (int64_t p = 0; p < npts; ++p)                                     \
            out[p] = interp3<R, R>(data, X, Y, Z, (idx_t)X * Y, pos[3 * p],    \
                                   pos[3 * p + 1], pos[3 * p + 2]);            \
    }                                                                          \
    extern "C" void orc_advect4d_real##SUF(const R *vel, R *grid, int X,       \
                                           int Y, int Z, int T, double dt,     \
                                           double dtFac)                       \
    {                                                                          \
        advect4d<R, R>((const V4<R> *)vel, grid, Dims{X, Y, Z, T},             \
                       (R)dt * (R)dtFac);                                      \
    }                                                                          \
    extern "C" void orc_advect4d_vec4##SUF(const R *vel, R *grid, int X,       \
                                           int Y, int Z, int T, double dt,     \
                                           double dtFac)                       \
    {                                                                          \
        advect4d<R, V4<R>>((const V4<R> *)vel, (V4<R> *)grid,                  \
                           Dims{X, Y, Z, T}, (R)dt * (R)dtFac);                \
    }                                                                          \
    extern "C" void orc_advect_cfl4d_real##SUF(const R *vel, R *grid, int X,   \
                                               int Y, int Z, int T,            \
                                               double cfl, double velFactor,   \
                                               double dt)                      \
    {                                                                          \
        advectCfl4d<R, R>((const V4<R> *)vel, grid, Dims{X, Y, Z, T}, (R)cfl,  \
                          (R)velFactor, (R)dt);                                \
    }                                                                          \
    extern "C" void orc_interpolate_grid4d##SUF(                               \
        R *target, int tX, int tY, int tZ, int tT, const R *source, int sX,    \
        int sY, int sZ, int sT, int isVec4, const double *offset,              \
        const double *scale, const double *size)                               \
    {                                                                          \
        V4<R> sf(1), off2((R)offset[0], (R)offset[1], (R)offset[2],            \
                          (R)offset[3]);                                       \
        gridFactor4d<R>(V4<R>((R)sX, (R)sY, (R)sZ, (R)sT),                     \
                        V4<R>((R)tX, (R)tY, (R)tZ, (R)tT),                     \
                        V4<R>((R)size[0], (R)size[1], (R)size[2], (R)size[3]), \
                        V4<R>((R)scale[0], (R)scale[1], (R)scale[2],           \
                              (R)scale[3]),                                    \
                        sf, off2);                                             \
        Dims Dt{tX, tY, tZ, tT}, Ds{sX, sY, sZ, sT};                           \
        if (isVec4)                                                            \
            resample4d<R, V4<R>>((V4<R> *)target, Dt, (const V4<R> *)source,   \
                                 Ds, sf, off2);                                \
        else                                                                   \
            resample4d<R, R>(target, Dt, source, Ds, sf, off2);                \
    }                                                                          \
    extern "C" void orc_pyramid_resample##SUF(R *target, int tX, int tY,       \
                                              int tZ, int tT,                  \
                                              const R *source, int sX, int sY, \
                                              int sZ, int sT, int isVec4)      \
    {                                                                          \
        Dims Dt{tX, tY, tZ, tT}, Ds{sX, sY, sZ, sT};                           \
        if (isVec4)                                                            \
            pyramidResample<R, V4<R>>((V4<R> *)target, Dt,                     \
                                      (const V4<R> *)source, Ds);              \
        else                                                                   \
            pyramidResample<R, R>(target, Dt, source, Ds);                     \
    }                                                                          \
    extern "C" void orc_set_bound_neumann_real##SUF(R *g, int X, int Y, int Z, \
                                                    int T, int w)              \
    {                                                                          \
        setBoundNeumann<R>(g, Dims{X, Y, Z, T}, w);                            \
    }                                                                          \
    extern "C" void orc_set_bound_real##SUF(R *g, int X, int Y, int Z, int T,  \
                                            double value, int w)               \
    {                                                                          \
        setBound<R>(g, Dims{X, Y, Z, T}, (R)value, w);                         \
    }                                                                          \
    extern "C" double orc_ls_diff##SUF(const R *a, const R *b, R *out, int X,  \
                                       int Y, int Z, int T, double corr,       \
                                       int bnd)                                \
    {                                                                          \
        return (double)lsDiff<R>(a, b, out, Dims{X, Y, Z, T}, (R)corr, bnd);   \
    }                                                                          \
    extern "C" void orc_gaussian_blur##SUF(R *g, int X, int Y, int Z, int T,   \
                                           int isVec4, double sigma, int iter) \
    {                                                                          \
        if (isVec4)                                                            \
            gaussianBlur<R, V4<R>>((V4<R> *)g, Dims{X, Y, Z, T}, (R)sigma,     \
                                   iter);                                      \
        else                                                                   \
            gaussianBlur<R, R>(g, Dims{X, Y, Z, T}, (R)sigma, iter);           \
    }                                                                          \
    extern "C" void orc_optical_flow4d##SUF(                                   \
        R *vel, const R *i0, const R *i1, R *rhsT, int X, int Y, int Z, int T, \
        const flof_of_params *p, double dt, const R *warm, int32_t *iters,     \
        double *residual, int32_t *status)                                     \
    {                                                                          \
        g_pstore = p->precision & 1;                                           \
        g_imgmode = (p->precision & 2) != 0;                                   \
        const idx_t nn = (idx_t)X * Y * Z * T;                                 \
        std::vector<R> b0(i0, i0 + nn), b1(i1, i1 + nn);                       \
        img_round_all<R>(b0.data(), nn);                                       \
        img_round_all<R>(b1.data(), nn);                                       \
        CgResult cg = opticalFlowLevel<R>(                                     \
            (V4<R> *)vel, b0.data(), b1.data(), rhsT, Dims{X, Y, Z, T},        \
            (R)p->wSmooth,                                                     \
            (R)p->wEnergy, (R)p->postVelBlur, (R)p->cgAccuracy,                \
            (R)p->resetBndWidth, (R)dt, p->cgMaxIter, (const V4<R> *)warm);    \
        if (iters) *iters = cg.iters;                                          \
        if (residual) *residual = cg.residual;                                 \
        if (status) *status = cg.status;                                       \
    }                                                                          \
    extern "C" double orc_multiscale4d##SUF(                                   \
        R *vel, const R *i0, const R *i1, R *rhsT, int X, int Y, int Z, int T, \
        const flof_of_params *p, double dt, const R *warm,                     \
        flof_solve_info *info)                                                 \
    {                                                                          \
        if (info) std::memset(info, 0, sizeof(*info));                         \
        OfParams P = fromC(p);                                                 \
        g_pstore = p->precision & 1;                                           \
        g_imgmode = (p->precision & 2) != 0;                                   \
        const idx_t nn = (idx_t)X * Y * Z * T;                                 \
        std::vector<R> b0(i0, i0 + nn), b1(i1, i1 + nn);                       \
        img_round_all<R>(b0.data(), nn);                                       \
        img_round_all<R>(b1.data(), nn);                                       \
        return (double)multiscale<R>((V4<R> *)vel, b0.data(), b1.data(), rhsT, \
                                     Dims{X, Y, Z, T}, P, p->level, (R)dt,     \
                                     info, (const V4<R> *)warm);               \
    }                                                                          \
    extern "C" void orc_corr_vels_of4d##SUF(                                   \
        R *dst, R *vel, const R *phiOrg, const R *phiTarget, int X, int Y,     \
        int Z, int T, double threshPhi, double postVelBlur,                    \
        double resetBndWidth, int maxIter, double dt)                          \
    {                                                                          \
        corrVelsOf<R>((V4<R> *)dst, (V4<R> *)vel, phiOrg, phiTarget,           \
                      Dims{X, Y, Z, T}, (R)threshPhi, (R)postVelBlur,          \
                      (R)resetBndWidth, maxIter, (R)dt);                       \
    }                                                                          \
    extern "C" void orc_place_frames##SUF(                                     \
        R *phi, int X, int Y, int Z, int T, const R *frames, int sx, int sy,   \
        int sz, const double *offset, const double *scale, int fileIdxStart,   \
        int fileIdxEnd, int debugSkipLoad, double spread,                      \
        const double *overrideSize, double overrideTimeOff,                    \
        int overrideGoodRegion, double loadTimeScale, int rescaleSdfValues,    \
        double sdfIsoOff, double repeatStartFrame)                             \
    {                                                                          \
        placeFrames<R>(                                                        \
            phi, Dims{X, Y, Z, T}, frames, sx, sy, sz,                         \
            V4<R>((R)offset[0], (R)offset[1], (R)offset[2], (R)offset[3]),     \
            V4<R>((R)scale[0], (R)scale[1], (R)scale[2], (R)scale[3]),         \
            fileIdxStart, fileIdxEnd, debugSkipLoad, (R)spread,                \
            V4<R>((R)overrideSize[0], (R)overrideSize[1], (R)overrideSize[2],  \
                  (R)overrideSize[3]),                                         \
            (R)overrideTimeOff, overrideGoodRegion, (R)loadTimeScale,          \
            rescaleSdfValues != 0, (R)sdfIsoOff, (R)repeatStartFrame);         \
    }                                                                          \
    extern "C" void orc_warp_slice##SUF(                                       \
        const R *defo, int vX, int vY, int vZ, int vT, R *dst, int dx, int dy, \
        int dz, const R *phi, int pX, int pY, int pZ, int pT, double time,     \
        double blendAlpha, double loadTimeScale, const double *defoOffset,     \
        const double *defoScale, const double *defoFactor,                     \
        const double *overrideSize, double overrideTimeOff, int zeroVel,       \
        int bordSkip, double defoAniFac, int optimized)                        \
    {                                                                          \
        auto v4 = [](const double *a) {                                        \
            return V4<R>((R)a[0], (R)a[1], (R)a[2], (R)a[3]);                  \
        };                                                                     \
        warpSlice<R>((const V4<R> *)defo, Dims{vX, vY, vZ, vT}, dst, dx, dy,   \
                     dz, phi, Dims{pX, pY, pZ, pT}, (R)time, (R)blendAlpha,    \
                     (R)loadTimeScale, v4(defoOffset), v4(defoScale),          \
                     v4(defoFactor), v4(overrideSize), (R)overrideTimeOff,     \
                     zeroVel != 0, bordSkip, (R)defoAniFac, optimized != 0);   \
    }                                                                          \
    extern "C" int orc_warp_slice_vols##SUF(                                   \
        const R *const *vols, int nvols, int vX, int vY, int vZ, int vT,       \
        R *dst, int dx, int dy, int dz, const R *phi, int pX, int pY, int pZ,  \
        int pT, double time, double blendAlpha, double loadTimeScale,          \
        const double *defoOffset, const double *defoScale,                     \
        const double *defoFactor, const double *overrideSize,                  \
        double overrideTimeOff, int zeroVel, int bordSkip, double defoAniFac,  \
        int doAligned, double partialLoadFac, double thirdAlpha,               \
        double fourthAlpha)                                                    \
    {                                                                          \
        auto v4 = [](const double *a) {                                        \
            return V4<R>((R)a[0], (R)a[1], (R)a[2], (R)a[3]);                  \
        };                                                                     \
        std::vector<const V4<R> *> vv;                                         \
        for (int q = 0; q < nvols; ++q) vv.push_back((const V4<R> *)vols[q]);  \
        try {                                                                  \
            warpSliceVols<R>(vv, Dims{vX, vY, vZ, vT}, dst, dx, dy, dz, phi,   \
                             Dims{pX, pY, pZ, pT}, (R)time, (R)blendAlpha,     \
                             (R)loadTimeScale, v4(defoOffset), v4(defoScale),  \
                             v4(defoFactor), v4(overrideSize),                 \
                             (R)overrideTimeOff, zeroVel != 0, bordSkip,       \
                             (R)defoAniFac, doAligned != 0, (R)partialLoadFac, \
                             (R)thirdAlpha, (R)fourthAlpha);                   \
        } catch (const std::exception &) {                                     \
            return -1;                                                         \
        }                                                                      \
        return 0;                                                              \
    }                                                                          \
    extern "C" void orc_repeat_frame4d##SUF(R *phi, int X, int Y, int Z, int T,  \
                                            double srct, double range,         \
                                            int bnd)                           \
    {                                                                          \
        repeatFrame4d<R>(phi, Dims{X, Y, Z, T}, (R)srct, (R)range, bnd);       \
    }                                                                          \
    extern "C" void orc_extrap4d_ls_simple##SUF(R *phi, int X, int Y, int Z,   \
                                                int T, int distance,           \
                                                int inside)                    \
    {                                                                          \
        extrap4dLsSimple<R>(phi, Dims{X, Y, Z, T}, distance, inside != 0);     \
    }                                                                          \
    extern "C" void orc_drop_sdf##SUF(R *out, int X, int Y, int Z, int T,      \
                                      int scale, uint64_t seed, int nspheres,  \
                                      double rmin, double rmax,                \
                                      double clampCells)                       \
    {                                                                          \
        dropSdf<R>(out, X, Y, Z, T, scale, seed, nspheres, rmin, rmax,         \
                   clampCells);                                                \
    }

extern "C" void orc_set_sum_mode(int mode) { g_sum_mode = mode; }
extern "C" float orc_round_bf16(float f) { return round_bf16(f); }
extern "C" float orc_round_f16(float f) { return round_f16(f); }
extern "C" int orc_get_sum_mode(void) { return g_sum_mode; }

ORC_DEFINE(float, _f32)
ORC_DEFINE(double, _f64)
