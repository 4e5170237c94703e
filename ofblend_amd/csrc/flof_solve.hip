// flof_solve.hip -- the 4D optical-flow solve of libflof_hip: matrix-free
// Jacobi-PCG of the smoothness-regularised system, Gaussian velocity blur,
// level-set projection correction and the multiscale driver.
//
// Reference: source/plugin/optflow4d.cpp -- opticalFlowDim :361-553,
// GridCGOptflow4d :268-355, gaussianBlurGeneric :127-174, corrVelsOfTempl
// :613-802, opticalFlowMultiscaleTemplate :936-1173.
//
// The system of one level, per interior cell c and component d (unknown
// U = vel*h, h = 1/X):
//     (g g^T + (8w + e) I) U(c) - w * sum_{8 nbrs n} U_d(n) = b_d(c)
// with g = central differences of i1 times X/2 and b = -g (i1 - i0)/dt
// (+ terms of the incoming flow).  Outer-layer cells are identity rows.  The
// reference assembles a 4N x 4N matrix (80 B/cell); here the operator is
// applied from g on the fly with the identical FP32 operation order, so
// q = A p is bit-identical to the reference's applyMat.
//
// PCG per iteration: three stencil kernels (p.q, the r update with z.r and
// max r, the x and direction update; 148 B per cell).  The FP64 scalars:
// alpha = sigma / p.q is reduced from the p.q partials by every workgroup of
// the update itself (cg_alpha); the stopping test and beta by the 1-block
// kernel k_cg_beta between the update and the direction kernel.  p.q and
// the update run as LDS-tiled kernels
// (k_cg_stencil_lds) on levels with long t segments and as t-marching
// kernels (k_cg_pq_tm, k_cg_update_tm) elsewhere; the direction update is
// k_cg_dir_tm.  The streaming variant with a q array (k_cg_apply,
// k_cg_update, k_cg_dir, 208 B per cell) is kept for comparison
// (FLOF_NO_TMARCH=1).
// The host enqueues iterations in chunks and reads the device state once per
// chunk; kernels of iterations past convergence return immediately.

#include <algorithm>
#include <functional>
#include <memory>
#include <array>
#include <atomic>
#include <thread>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "flof_internal.h"

// unroll factor of the t-marching loops
#ifndef FLOF_TM_UNROLL
#define FLOF_TM_UNROLL 1
#endif
// 1: t-march work items dealt XCD-contiguously (blockIdx % 8 = XCD);
// 0: in launch order.  Launch order measured 2.5% faster per PCG iteration
// at 64^4 (its z-neighbour chunks, 16 blocks apart, share an XCD).
#ifndef FLOF_TM_XCD_MAP
#define FLOF_TM_XCD_MAP 0
#endif
// non-temporal access to data a kernel touches once (1: r and x of the
// update kernel; 2: also p's own load and store in the direction kernel).
// Measured at 64^4: 626 -> 565 us per PCG iteration.
#ifndef FLOF_STREAM_NT
#define FLOF_STREAM_NT 2
#endif

namespace flof {

struct CgState {
    double sigma;
    double alpha;
    double alpha_prev;  // alpha of the previous iteration (deferred x update)
    double beta;
    double acc;      // stopping threshold cgAccuracy * resIni
    double resIni;   // signed max of the initial residual
    double resRel;   // (float)(residual / resIni), as ret_residual
    double r2, b2;   // warm-start residual check
    double bb, rr;   // ||b||^2 of the rhs and ||r||^2 of the final residual
    double gamma;    // warm-start scale
    int iters;
    int done;
    int status;  // 0 converged, 1 not converged, 2 sigma 0/NaN, 3 tiny rhs
    int maxIter;
    int warm;    // warm start kept
    int zero;    // result must be zeroed (failed solve)
    int dir_mode;  // k_cg_dir: 0 skip, 1 x and p update, 2 x update only
    int fast_rcp;  // set-up: every interior g_d^2 + diag in rcp_newton's range
};

struct LevelGeom {
    Dims4 D;
    Coord4 C;
    float offd;   // -wSmooth (times mDx2Inv = 1)
    float diag;   // 8*wSmooth + wEnergy, formed as the reference does
    int use_offd; // offd != 0
    float eps;    // VECTOR_EPSILON of FP32 builds
    int fast_rcp; // every interior diagonal entry g_d^2 + diag of the level
                  // lies in the range of rcp_newton (set after the set-up)
};

__device__ __forceinline__ bool interior(const Dims4 &D, int i, int j, int k, int t)
{
    return i >= 1 && j >= 1 && k >= 1 && t >= 1 && i < D.X - 1 && j < D.Y - 1 &&
           k < D.Z - 1 && t < D.T - 1;
}

// q = A p at cell c (interior).  Order: 8 neighbour terms t-1, z-1, y-1, x-1,
// x+1, y+1, z+1, t+1, then the 4-wide block row (optflow4d.cpp:211-232).
__device__ __forceinline__ float4 apply_A(const float4 *__restrict__ p, int64_t c,
                                          const float4 &pc, const float4 &g,
                                          const LevelGeom &G)
{
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (G.use_offd) {
        const int64_t o[8] = {-G.D.st, -G.D.sz, -(int64_t)G.D.X, -1,
                              1,       (int64_t)G.D.X, G.D.sz, G.D.st};
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            const float4 n = p[c + o[m]];
            v.x = v.x + G.offd * n.x;
            v.y = v.y + G.offd * n.y;
            v.z = v.z + G.offd * n.z;
            v.w = v.w + G.offd * n.w;
        }
    }
    const float gg[4] = {g.x, g.y, g.z, g.w};
    const float pp[4] = {pc.x, pc.y, pc.z, pc.w};
    float out[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int d = 0; d < 4; ++d) {
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            float b = gg[d] * gg[m];
            if (d == m) b = b + G.diag;
            out[d] = out[d] + b * pp[m];
        }
    }
    return make_float4(out[0], out[1], out[2], out[3]);
}

// identity row: 0 + 1*p_d + 0*p_m, exactly as the assembled row evaluates
__device__ __forceinline__ float4 apply_I(const float4 &pc)
{
    const float pp[4] = {pc.x, pc.y, pc.z, pc.w};
    float out[4];
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        float v = 0.f;
#pragma unroll
        for (int m = 0; m < 4; ++m) v = v + (d == m ? 1.f : 0.f) * pp[m];
        out[d] = v;
    }
    return make_float4(out[0], out[1], out[2], out[3]);
}

__device__ __forceinline__ void dd_add4(DD &a, const float4 &u, const float4 &v)
{
    dd_add(a, (double)(u.x * v.x));
    dd_add(a, (double)(u.y * v.y));
    dd_add(a, (double)(u.z * v.z));
    dd_add(a, (double)(u.w * v.w));
}

__device__ __forceinline__ float4 prec_of(const float4 &g, float diag)
{
    // precond = 1./blockd[d][d] evaluated in double, stored as float
    return make_float4((float)(1.0 / (double)(g.x * g.x + diag)),
                       (float)(1.0 / (double)(g.y * g.y + diag)),
                       (float)(1.0 / (double)(g.z * g.z + diag)),
                       (float)(1.0 / (double)(g.w * g.w + diag)));
}

// Storage of the images i0 and i1 of a solve: FP32 (the reference's Real)
// or, in the mixed-precision form, bfloat16 with every component of the
// gradient of i1 rounded to binary16 where it is formed (FP32 arithmetic
// otherwise, so a GPU and an oracle that round at the same points agree).
struct ImgF32 {
    typedef float T;
    static __device__ __forceinline__ float ld(const float *p, int64_t c) { return p[c]; }
    static __device__ __forceinline__ float gq(float g) { return g; }
};
struct ImgBF16 {
    typedef bf16_t T;
    static __device__ __forceinline__ float ld(const bf16_t *p, int64_t c)
    {
        return bf16_to_f32(p[c].u);
    }
    static __device__ __forceinline__ float gq(float g) { return round_f16(g); }
};

// --- set-up (optflow4d.cpp:398-493 + the start of solve :284-303) ---------
template <class IM>
__global__ void __launch_bounds__(256)
k_of_setup(const typename IM::T *__restrict__ i0, const typename IM::T *__restrict__ i1,
           const float4 *__restrict__ vin, float *__restrict__ rhsT,
           float4 *__restrict__ g, float4 *__restrict__ prec,
           float4 *__restrict__ x, float4 *__restrict__ r, float4 *__restrict__ p,
           LevelGeom G, float wSmooth, float wEnergy, float mDt, float mDx,
           float dxf, DD *part_dd, float *part_max, float *part_out)
{
    const Dims4 &D = G.D;
    const int64_t n = (int64_t)D.st * D.T;
    DD sig{0., 0.};
    float mx = -INFINITY;
    float out = 0.f;  // 1: a diagonal entry outside rcp_newton's range
    FOR_XCD(c, n)
    {
        int i, j, k, t;
        decompose((uint32_t)c, G.C, i, j, k, t);
        const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
        if (!interior(D, i, j, k, t)) {
            g[c] = zero4;
            prec[c] = make_float4(1.f, 1.f, 1.f, 1.f);
            x[c] = zero4;
            r[c] = zero4;
            p[c] = zero4;
            if (rhsT) rhsT[c] = 0.f;
            mx = fmaxf(mx, 0.f);
            continue;
        }
        const float tderiv = (IM::ld(i1, c) - IM::ld(i0, c)) / mDt;
        float4 gc;
        gc.x = IM::gq((IM::ld(i1, c + 1) - IM::ld(i1, c - 1)) * dxf);
        gc.y = IM::gq((IM::ld(i1, c + D.X) - IM::ld(i1, c - D.X)) * dxf);
        gc.z = IM::gq((IM::ld(i1, c + D.sz) - IM::ld(i1, c - D.sz)) * dxf);
        gc.w = IM::gq((IM::ld(i1, c + D.st) - IM::ld(i1, c - D.st)) * dxf);
        float b[4] = {-gc.x * tderiv, -gc.y * tderiv, -gc.z * tderiv, -gc.w * tderiv};
        if (vin) {
            const int nbx[8] = {0, 0, 0, -1, 1, 0, 0, 0};
            const int nby[8] = {0, 0, -1, 0, 0, 1, 0, 0};
            const int nbz[8] = {0, -1, 0, 0, 0, 0, 1, 0};
            const int nbt[8] = {-1, 0, 0, 0, 0, 0, 0, 1};
            const float4 vc = vin[c];
            const float vcc[4] = {vc.x, vc.y, vc.z, vc.w};
            for (int m = 0; m < 8; ++m) {
                const int ti = max(0, min(D.X - 1, i + nbx[m]));
                const int tj = max(0, min(D.Y - 1, j + nby[m]));
                const int tk = max(0, min(D.Z - 1, k + nbz[m]));
                const int tt = max(0, min(D.T - 1, t + nbt[m]));
                const float4 vn = vin[(int64_t)ti + (int64_t)D.X * tj + D.sz * tk + D.st * tt];
                const float vnn[4] = {vn.x, vn.y, vn.z, vn.w};
                for (int d = 0; d < 4; ++d)
                    b[d] = b[d] - wSmooth * (vcc[d] - vnn[d]) * mDx * 1.f;
            }
            for (int d = 0; d < 4; ++d) b[d] = b[d] - wEnergy * vcc[d] * mDx;
        }
        const float4 rc = make_float4(b[0], b[1], b[2], b[3]);
        const float4 pr = prec_of(gc, G.diag);
        if (!(rcp_exact_range(gc.x * gc.x + G.diag) && rcp_exact_range(gc.y * gc.y + G.diag) &&
              rcp_exact_range(gc.z * gc.z + G.diag) && rcp_exact_range(gc.w * gc.w + G.diag)))
            out = 1.f;
        const float4 z = f4mul(rc, pr);
        g[c] = gc;
        prec[c] = pr;
        x[c] = zero4;
        r[c] = rc;
        p[c] = z;
        if (rhsT) rhsT[c] = rc.x;
        mx = fmaxf(mx, fmaxf(fmaxf(rc.x, rc.y), fmaxf(rc.z, rc.w)));
        dd_add4(sig, z, rc);
    }
    sig = block_dd(sig);
    mx = block_max(mx);
    out = block_max(out);
    if (threadIdx.x == 0) {
        part_dd[blockIdx.x] = sig;
        part_max[blockIdx.x] = mx;
        part_out[blockIdx.x] = out;
    }
}

// reduce partials of nb blocks (one block of 256 threads)
__device__ __forceinline__ void fin_dd_max(const DD *pd, const float *pm, int nb, DD &s,
                                           float &m)
{
    s = DD{0., 0.};
    m = -INFINITY;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        dd_merge(s, pd[b]);
        if (pm) m = fmaxf(m, pm[b]);
    }
    s = block_dd(s);
    m = block_max(m);
}

// ||v||^2 of a Vec4 array (FP64 squares of the FP32 values, correctly
// rounded sum): partials, then a 1-block kernel into *out
__global__ void __launch_bounds__(256)
k_sumsq(const float4 *__restrict__ v, int64_t n, DD *part)
{
    DD s{0., 0.};
    FOR_XCD(c, n)
    {
        const float4 a = v[c];
        dd_add(s, (double)a.x * (double)a.x);
        dd_add(s, (double)a.y * (double)a.y);
        dd_add(s, (double)a.z * (double)a.z);
        dd_add(s, (double)a.w * (double)a.w);
    }
    s = block_dd(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}
__global__ void k_sumsq_fin(const DD *part, int nb, double *out)
{
    DD s;
    float m;
    fin_dd_max(part, nullptr, nb, s, m);
    if (threadIdx.x == 0) *out = s.hi + s.lo;
}

// alpha = sigma / p.q in the update kernels: every workgroup reduces the
// p.q partials of the previous launch itself (the order of k_cg_alpha, so
// all hold the same value and no finalizer launch runs between the two
// kernels); workgroup 0 keeps alpha and the previous one in the state for
// the direction kernel.  Called by all threads of the workgroup.
__device__ __forceinline__ double cg_alpha(CgState *st, const DD *part_pq, int nb,
                                           int bid = -1)
{
    DD s;
    float m;
    fin_dd_max(part_pq, nullptr, nb, s, m);
    const double alpha = st->sigma / (s.hi + s.lo);
    if ((bid < 0 ? (int)blockIdx.x : bid) == 0 && threadIdx.x == 0) {
        st->alpha_prev = st->alpha;
        st->alpha = alpha;
    }
    return alpha;
}

// iterations after the last one: the direction kernel does nothing (what
// k_cg_alpha does for the streaming form)
__device__ __forceinline__ void cg_after_last(CgState *st, int bid = -1)
{
    if ((bid < 0 ? (int)blockIdx.x : bid) == 0 && threadIdx.x == 0) st->dir_mode = 0;
}

__global__ void k_setup_fin(CgState *st, const DD *pd, const float *pm, const float *pout,
                            int nb, float cgAcc, float eps)
{
    DD s;
    float m;
    fin_dd_max(pd, pm, nb, s, m);
    float out = 0.f;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) out = fmaxf(out, pout[b]);
    out = block_max(out);
    if (threadIdx.x == 0) {
        const double residual = (double)m;
        st->fast_rcp = out == 0.f ? 1 : 0;
        st->iters = 0;
        st->warm = 0;
        st->zero = 0;
        st->resRel = 1e10;
        st->resIni = residual;
        if (residual < (double)eps) {
            st->resRel = 0.;
            st->status = 3;
            st->done = 1;
        } else {
            st->acc = (double)cgAcc * residual;
            const double sigma = s.hi + s.lo;
            st->sigma = sigma;
            if (sigma == 0. || sigma != sigma) {
                st->status = 2;
                st->done = 1;
            } else {
                st->status = 1;
                st->done = 0;
            }
        }
    }
}

// --- warm start (extension): guess w, scaled by the energy-optimal gamma;
// kept if ||b - gamma A w||_2 <= ||b||_2 ------------------------------------
__global__ void k_warm_guess(const float4 *__restrict__ warm,
                             const float4 *__restrict__ vin_acc, float4 *__restrict__ x0s,
                             Dims4 D, Coord4 C, float mDx)
{
    const int64_t n = (int64_t)D.st * D.T;
    FOR_XCD(c, n)
    {
        int i, j, k, t;
        decompose((uint32_t)c, C, i, j, k, t);
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (interior(D, i, j, k, t)) {
            float4 w = warm[c];
            if (vin_acc) w = f4sub(w, vin_acc[c]);
            v = f4muls(w, mDx);
        }
        x0s[c] = v;
    }
}

// A w, and the partial sums of b.w, w.A w and b.b (FP64 products of the FP32
// values, correctly rounded sums)
__global__ void __launch_bounds__(256)
k_warm_aw(const float4 *__restrict__ w, const float4 *__restrict__ g,
          const float4 *__restrict__ r, float4 *__restrict__ aw, LevelGeom G, DD *part_bw,
          DD *part_waw, DD *part_bb)
{
    const Dims4 &D = G.D;
    const int64_t n = (int64_t)D.st * D.T;
    DD bw{0., 0.}, waw{0., 0.}, bb{0., 0.};
    FOR_XCD(c, n)
    {
        int i, j, k, t;
        decompose((uint32_t)c, G.C, i, j, k, t);
        const float4 wc = w[c];
        const float4 a = interior(D, i, j, k, t) ? apply_A(w, c, wc, g[c], G) : apply_I(wc);
        const float4 b = r[c];
        aw[c] = a;
        dd_add(bw, (double)b.x * (double)wc.x);
        dd_add(bw, (double)b.y * (double)wc.y);
        dd_add(bw, (double)b.z * (double)wc.z);
        dd_add(bw, (double)b.w * (double)wc.w);
        dd_add(waw, (double)wc.x * (double)a.x);
        dd_add(waw, (double)wc.y * (double)a.y);
        dd_add(waw, (double)wc.z * (double)a.z);
        dd_add(waw, (double)wc.w * (double)a.w);
        dd_add(bb, (double)b.x * (double)b.x);
        dd_add(bb, (double)b.y * (double)b.y);
        dd_add(bb, (double)b.z * (double)b.z);
        dd_add(bb, (double)b.w * (double)b.w);
    }
    bw = block_dd(bw);
    waw = block_dd(waw);
    bb = block_dd(bb);
    if (threadIdx.x == 0) {
        part_bw[blockIdx.x] = bw;
        part_waw[blockIdx.x] = waw;
        part_bb[blockIdx.x] = bb;
    }
}

// gamma = (b.w) / (w.A w), the energy-optimal multiple of the guess (0 if
// w.A w is not positive and finite)
__global__ void k_warm_gamma(CgState *st, const DD *pbw, const DD *pwaw, const DD *pbb, int nb)
{
    DD s1, s2, s3;
    float m;
    fin_dd_max(pbw, nullptr, nb, s1, m);
    fin_dd_max(pwaw, nullptr, nb, s2, m);
    fin_dd_max(pbb, nullptr, nb, s3, m);
    if (threadIdx.x == 0) {
        const double wAw = s2.hi + s2.lo;
        double gamma = 0.;
        if (wAw > 0. && wAw < HUGE_VAL) gamma = (s1.hi + s1.lo) / wAw;
        if (gamma != gamma) gamma = 0.;
        st->gamma = gamma;
        st->b2 = s3.hi + s3.lo;
    }
}

// x0 = gamma w, r_w = b - gamma (A w) (in place), partial ||r_w||^2
__global__ void __launch_bounds__(256)
k_warm_scale(float4 *__restrict__ w, float4 *__restrict__ aw, const float4 *__restrict__ r,
             int64_t n, const CgState *st, DD *part)
{
    const double gm = st->gamma;
    DD r2{0., 0.};
    FOR_XCD(c, n)
    {
        const float4 wc = w[c], a = aw[c], b = r[c];
        float4 x, rr;
        x.x = (float)(gm * (double)wc.x);
        x.y = (float)(gm * (double)wc.y);
        x.z = (float)(gm * (double)wc.z);
        x.w = (float)(gm * (double)wc.w);
        rr.x = (float)((double)b.x - gm * (double)a.x);
        rr.y = (float)((double)b.y - gm * (double)a.y);
        rr.z = (float)((double)b.z - gm * (double)a.z);
        rr.w = (float)((double)b.w - gm * (double)a.w);
        w[c] = x;
        aw[c] = rr;
        dd_add(r2, (double)rr.x * (double)rr.x);
        dd_add(r2, (double)rr.y * (double)rr.y);
        dd_add(r2, (double)rr.z * (double)rr.z);
        dd_add(r2, (double)rr.w * (double)rr.w);
    }
    r2 = block_dd(r2);
    if (threadIdx.x == 0) part[blockIdx.x] = r2;
}

__global__ void k_warm_fin(CgState *st, const DD *pa, int nb)
{
    DD s;
    float m;
    fin_dd_max(pa, nullptr, nb, s, m);
    if (threadIdx.x == 0) {
        st->r2 = s.hi + s.lo;
        st->warm = (st->done == 0 && st->r2 <= st->b2) ? 1 : 0;
    }
}

// adopt the warm start: x = x0s, r = r_w, p = z
__global__ void __launch_bounds__(256)
k_warm_adopt(const float4 *__restrict__ x0s, const float4 *__restrict__ rw,
             const float4 *__restrict__ prec, float4 *__restrict__ x,
             float4 *__restrict__ r, float4 *__restrict__ p, int64_t n, const CgState *st,
             DD *part_dd, float *part_max)
{
    if (!st->warm) return;
    DD sig{0., 0.};
    float mx = -INFINITY;
    FOR_XCD(c, n)
    {
        const float4 rr = rw[c];
        const float4 z = f4mul(rr, prec[c]);
        x[c] = x0s[c];
        r[c] = rr;
        p[c] = z;
        mx = fmaxf(mx, fmaxf(fmaxf(rr.x, rr.y), fmaxf(rr.z, rr.w)));
        dd_add4(sig, z, rr);
    }
    sig = block_dd(sig);
    mx = block_max(mx);
    if (threadIdx.x == 0) {
        part_dd[blockIdx.x] = sig;
        part_max[blockIdx.x] = mx;
    }
}

__global__ void k_warm_adopt_fin(CgState *st, const DD *pd, const float *pm, int nb)
{
    if (!st->warm) return;
    DD s;
    float m;
    fin_dd_max(pd, pm, nb, s, m);
    if (threadIdx.x == 0) {
        const double residual = (double)m;
        st->resRel = (double)(float)(residual / st->resIni);
        if (residual <= st->acc) {
            st->status = 0;
            st->done = 1;
        } else {
            const double sigma = s.hi + s.lo;
            st->sigma = sigma;
            if (sigma == 0. || sigma != sigma) {
                st->status = 2;
                st->done = 1;
                st->zero = 1;
            }
        }
    }
}

// --- PCG iteration kernels -------------------------------------------------
__global__ void __launch_bounds__(256)
k_cg_apply(const float4 *__restrict__ p, const float4 *__restrict__ g,
           float4 *__restrict__ q, LevelGeom G, const CgState *st, DD *part_dd)
{
    if (st->done) return;
    const Dims4 &D = G.D;
    const int64_t n = (int64_t)D.st * D.T;
    DD acc{0., 0.};
    FOR_XCD(c, n)
    {
        int i, j, k, t;
        decompose((uint32_t)c, G.C, i, j, k, t);
        const float4 pc = p[c];
        const float4 qc = interior(D, i, j, k, t) ? apply_A(p, c, pc, g[c], G) : apply_I(pc);
        q[c] = qc;
        dd_add4(acc, pc, qc);
    }
    acc = block_dd(acc);
    if (threadIdx.x == 0) part_dd[blockIdx.x] = acc;
}

// t-marching work items: (3D cell c3, t segment); the thread keeps the
// t-1, t, t+1 values of its stencil operands in registers, so every value is
// fetched once from HBM; the x/y/z neighbours at time t are fetched through
// the caches, which hold the few t-planes the co-resident workgroups are
// working on (every workgroup of a launch is resident at 64^4).  Work items
// follow the launch order (FLOF_TM_XCD_MAP selects the XCD-contiguous order).
struct TMarch {
    int64_t n3;      // X*Y*Z
    int nblk3;       // workgroups per t segment (ceil(n3 / 256))
    int seglen;      // t steps per segment
    int nseg;        // segments
    int total;       // nblk3 * nseg real workgroups
};

// alpha = sigma / p.q
__global__ void k_cg_alpha(CgState *st, const DD *pd, int nb)
{
    if (st->done) {
        // iterations after the last one: the direction kernel does nothing
        if (threadIdx.x == 0) st->dir_mode = 0;
        return;
    }
    DD s;
    float m;
    fin_dd_max(pd, nullptr, nb, s, m);
    if (threadIdx.x == 0) {
        st->alpha_prev = st->alpha;
        st->alpha = st->sigma / (s.hi + s.lo);
    }
}

__global__ void __launch_bounds__(256)
k_cg_update(float4 *__restrict__ r, const float4 *__restrict__ q,
            const float4 *__restrict__ prec, int64_t n, const CgState *st, DD *part_dd,
            float *part_max)
{
    // r -= alpha q (optflow4d.cpp:310); x += alpha p is applied by k_cg_dir,
    // which reads p anyway, with the same operands
    if (st->done) return;
    const double nalpha = -st->alpha;
    DD sig{0., 0.};
    float mx = -INFINITY;
    FOR_XCD(c, n)
    {
        const float4 qc = q[c];
        float4 rc = r[c];
        rc.x = (float)((double)rc.x + nalpha * (double)qc.x);
        rc.y = (float)((double)rc.y + nalpha * (double)qc.y);
        rc.z = (float)((double)rc.z + nalpha * (double)qc.z);
        rc.w = (float)((double)rc.w + nalpha * (double)qc.w);
        r[c] = rc;
        const float4 z = f4mul(rc, prec[c]);
        mx = fmaxf(mx, fmaxf(fmaxf(rc.x, rc.y), fmaxf(rc.z, rc.w)));
        dd_add4(sig, z, rc);
    }
    sig = block_dd(sig);
    mx = block_max(mx);
    if (threadIdx.x == 0) {
        part_dd[blockIdx.x] = sig;
        part_max[blockIdx.x] = mx;
    }
}

// signed-max stopping test, then beta = z.r / sigma (optflow4d.cpp:311-325)
__device__ __forceinline__ void cg_beta_body(CgState *st, const DD *pd, const float *pm,
                                             int nb)
{
    if (st->done) return;
    DD s;
    float m;
    fin_dd_max(pd, pm, nb, s, m);
    if (threadIdx.x == 0) {
        const double residual = (double)m;
        const double sigmaNew = s.hi + s.lo;
        const int it = st->iters;  // iterations completed before this one
        st->iters = it + 1;
        st->resRel = (double)(float)(residual / st->resIni);
        if (residual <= st->acc) {
            st->status = 0;
            st->done = 1;
            st->dir_mode = 2;  // the last x update only
        } else if (sigmaNew != sigmaNew || residual != residual) {
            // a NaN residual ends with a zero result (optflow4d.cpp:509-514)
            st->status = 2;
            st->zero = 1;
            st->done = 1;
            st->dir_mode = 0;
        } else {
            st->beta = sigmaNew / st->sigma;
            st->sigma = sigmaNew;
            st->dir_mode = 1;
            if (it + 1 >= st->maxIter) {
                st->status = 1;
                st->done = 1;
                st->dir_mode = 2;
            }
        }
    }
}
__global__ void k_cg_beta(CgState *st, const DD *pd, const float *pm, int nb)
{
    cg_beta_body(st, pd, pm, nb);
}

__global__ void __launch_bounds__(256)
k_cg_dir(float4 *__restrict__ x, float4 *__restrict__ p, const float4 *__restrict__ r,
         const float4 *__restrict__ prec, int64_t n, const CgState *st)
{
    // x += alpha p (optflow4d.cpp:309), then p = z + beta p (:318-323) unless
    // this was the last iteration
    const int mode = st->dir_mode;
    if (mode == 0) return;
    const double alpha = st->alpha, beta = st->beta;
    FOR_XCD(c, n)
    {
        const float4 pc = p[c];
        float4 xc = x[c];
        xc.x = (float)((double)xc.x + alpha * (double)pc.x);
        xc.y = (float)((double)xc.y + alpha * (double)pc.y);
        xc.z = (float)((double)xc.z + alpha * (double)pc.z);
        xc.w = (float)((double)xc.w + alpha * (double)pc.w);
        x[c] = xc;
        if (mode == 1) {
            const float4 z = f4mul(r[c], prec[c]);
            float4 np;
            np.x = (float)((double)z.x + beta * (double)pc.x);
            np.y = (float)((double)z.y + beta * (double)pc.y);
            np.z = (float)((double)z.z + beta * (double)pc.z);
            np.w = (float)((double)z.w + beta * (double)pc.w);
            p[c] = np;
        }
    }
}

// ---------------------------------------------------------------------------
// Fused t-marching PCG (the default).  No q array and no stored g / diag^-1:
// g is recomputed from i1 (4 B per cell instead of 16 B) exactly as the set-up
// forms it, the Jacobi preconditioner from g (1/d in FP32 is correctly
// rounded, as is (float)(1./(double)d): with 53 >= 2*24+2 bits the double
// rounding is innocuous), and q = A p is formed twice, bit-identically: once
// for p.q and once in the update.  Per cell and iteration:
//   k_cg_pq_tm     : read p, i1                        20 B  -> alpha
//   k_cg_update_tm : read p, i1, r; write r            52 B  -> beta
//   k_cg_dir_tm    : read r, i1, p; write p            52 B  (even iterations)
//                    also read x, previous p; write x 100 B  (odd iterations)
// 148 B on average instead of 208 B of the three streaming kernels.

// the Jacobi preconditioner 1 / (g_d^2 + diag); fast (uniform over the
// launch) when the set-up found every diagonal entry of the level in the
// range where rcp_newton equals the division
__device__ __forceinline__ float4 prec_f32(const float4 &g, float diag, int fast)
{
    const float dx = g.x * g.x + diag, dy = g.y * g.y + diag;
    const float dz = g.z * g.z + diag, dw = g.w * g.w + diag;
    if (fast)
        return make_float4(rcp_newton(dx), rcp_newton(dy), rcp_newton(dz), rcp_newton(dw));
    return make_float4(1.f / dx, 1.f / dy, 1.f / dz, 1.f / dw);
}
__device__ __forceinline__ float4 prec_f32(const float4 &g, const LevelGeom &G)
{
    return prec_f32(g, G.diag, G.fast_rcp);
}

// one thread's t-segment of one 3D cell
struct TmCell {
    int64_t c3;
    int t0, t1;
    bool sp_int;  // x, y, z interior
    bool live;
};

__device__ __forceinline__ TmCell tm_cell(const Dims4 &D, const TMarch &M, int bid, int nblk)
{
    TmCell w;
    const int per = nblk / FLOF_NXCD;
#if FLOF_TM_XCD_MAP
    const int L = (bid % FLOF_NXCD) * per + bid / FLOF_NXCD;
#else
    const int L = bid;
    (void)per;
#endif
    w.live = false;
    w.c3 = 0;
    w.t0 = w.t1 = 0;
    w.sp_int = false;
    if (L >= M.total) return w;
    const int seg = L / M.nblk3;
    const int b3 = L - seg * M.nblk3;
    w.t0 = seg * M.seglen;
    w.t1 = min(D.T, w.t0 + M.seglen);
    w.c3 = (int64_t)b3 * blockDim.x + threadIdx.x;
    if (w.c3 >= M.n3 || w.t0 >= w.t1) return w;
    const int i = (int)(w.c3 % D.X);
    const int j = (int)((w.c3 / D.X) % D.Y);
    const int k = (int)(w.c3 / D.sz);
    w.sp_int = i >= 1 && j >= 1 && k >= 1 && i < D.X - 1 && j < D.Y - 1 && k < D.Z - 1;
    w.live = true;
    return w;
}

// g at interior cell c; ipm / ipp are i1 at t-1 / t+1 (optflow4d.cpp:425-433)
template <class IM>
__device__ __forceinline__ float4 grad_i1(const typename IM::T *__restrict__ i1, int64_t c,
                                          const Dims4 &D, float ipm, float ipp, float dxf)
{
    float4 gc;
    gc.x = IM::gq((IM::ld(i1, c + 1) - IM::ld(i1, c - 1)) * dxf);
    gc.y = IM::gq((IM::ld(i1, c + D.X) - IM::ld(i1, c - D.X)) * dxf);
    gc.z = IM::gq((IM::ld(i1, c + D.sz) - IM::ld(i1, c - D.sz)) * dxf);
    gc.w = IM::gq((ipp - ipm) * dxf);
    return gc;
}

// Storage of the search direction p.  FP32 reproduces the reference; the
// bfloat16 form (extension) keeps p as 4 x 16 bit per cell, rounded to
// nearest even on every store, and every kernel works with the stored
// (rounded) direction, so x, r and the FP64 scalars stay consistent with it.
// (binary16 is not offered: p shrinks with the residual and leaves its
// exponent range long before a solve to 1e-5 ends.)
// loads and stores of data used once per kernel (r, x, and p in the
// direction kernel): non-temporal, so the stencil planes of p and i1 stay in
// L2 instead of being evicted by streamed data
typedef float f4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ld_stream(const float4 *a)
{
#if FLOF_STREAM_NT
    const f4v v = __builtin_nontemporal_load((const f4v *)a);
    return make_float4(v.x, v.y, v.z, v.w);
#else
    return *a;
#endif
}
__device__ __forceinline__ void st_stream(float4 *a, const float4 &v)
{
#if FLOF_STREAM_NT
    f4v w = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(w, (f4v *)a);
#else
    *a = v;
#endif
}

struct PStoreF32 {
    typedef float4 T;
    static __device__ __forceinline__ float4 rnd(const float4 &v) { return v; }
    static __device__ __forceinline__ float4 ld(const T *p, int64_t c) { return p[c]; }
    static __device__ __forceinline__ void st(T *p, int64_t c, const float4 &v) { p[c] = v; }
#if FLOF_STREAM_NT >= 2
    static __device__ __forceinline__ float4 ld_once(const T *p, int64_t c)
    {
        return ld_stream(p + c);
    }
    static __device__ __forceinline__ void st_once(T *p, int64_t c, const float4 &v)
    {
        st_stream(p + c, v);
    }
#else
    static __device__ __forceinline__ float4 ld_once(const T *p, int64_t c) { return p[c]; }
    static __device__ __forceinline__ void st_once(T *p, int64_t c, const float4 &v)
    {
        p[c] = v;
    }
#endif
};
typedef unsigned short bf4v __attribute__((ext_vector_type(4)));
struct PStoreBF16 {
    typedef bf4v T;
    // the value a store followed by a load gives
    static __device__ __forceinline__ float4 rnd(const float4 &v)
    {
        return make_float4(bf16_to_f32(f32_to_bf16(v.x)), bf16_to_f32(f32_to_bf16(v.y)),
                           bf16_to_f32(f32_to_bf16(v.z)), bf16_to_f32(f32_to_bf16(v.w)));
    }
    static __device__ __forceinline__ float4 ld(const T *p, int64_t c)
    {
        const bf4v h = p[c];
        return make_float4(bf16_to_f32(h.x), bf16_to_f32(h.y), bf16_to_f32(h.z),
                           bf16_to_f32(h.w));
    }
    static __device__ __forceinline__ float4 ld_once(const T *p, int64_t c) { return ld(p, c); }
    static __device__ __forceinline__ void st_once(T *p, int64_t c, const float4 &v)
    {
        st(p, c, v);
    }
    static __device__ __forceinline__ void st(T *p, int64_t c, const float4 &v)
    {
        bf4v h;
        h.x = f32_to_bf16(v.x);
        h.y = f32_to_bf16(v.y);
        h.z = f32_to_bf16(v.z);
        h.w = f32_to_bf16(v.w);
        p[c] = h;
    }
};

// A p at an interior cell with p(t-1) = pm, p(t) = pc, p(t+1) = pp in registers
template <class PS>
__device__ __forceinline__ float4 apply_A_tm(const typename PS::T *__restrict__ p, int64_t c,
                                             const float4 &pm, const float4 &pc,
                                             const float4 &pp, const float4 &gc,
                                             const LevelGeom &G)
{
    const Dims4 &D = G.D;
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (G.use_offd) {
        const float4 nb[8] = {pm,
                              PS::ld(p, c - D.sz),
                              PS::ld(p, c - D.X),
                              PS::ld(p, c - 1),
                              PS::ld(p, c + 1),
                              PS::ld(p, c + D.X),
                              PS::ld(p, c + D.sz),
                              pp};
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            v.x = v.x + G.offd * nb[m].x;
            v.y = v.y + G.offd * nb[m].y;
            v.z = v.z + G.offd * nb[m].z;
            v.w = v.w + G.offd * nb[m].w;
        }
    }
    const float gg[4] = {gc.x, gc.y, gc.z, gc.w};
    const float pv[4] = {pc.x, pc.y, pc.z, pc.w};
    float o[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int d = 0; d < 4; ++d) {
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            float b = gg[d] * gg[m];
            if (d == m) b = b + G.diag;
            o[d] = o[d] + b * pv[m];
        }
    }
    return make_float4(o[0], o[1], o[2], o[3]);
}

// p.(A p)
template <class PS, class IM>
__device__ __forceinline__ void
cg_pq_tm_body(int bid, int nblk, const typename PS::T *__restrict__ p,
              const typename IM::T *__restrict__ i1, const LevelGeom &G, const TMarch &M,
              float dxf, const CgState *st, DD *part_dd)
{
    if (st->done) return;
    const Dims4 &D = G.D;
    const TmCell w = tm_cell(D, M, bid, nblk);
    DD acc{0., 0.};
    if (w.live) {
        const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
        const int64_t c3 = w.c3;
        float4 pm = w.t0 > 0 ? PS::ld(p, c3 + D.st * (w.t0 - 1)) : zero4;
        float4 pc = PS::ld(p, c3 + D.st * w.t0);
        float im = w.t0 > 0 ? IM::ld(i1, c3 + D.st * (w.t0 - 1)) : 0.f;
        float ic = IM::ld(i1, c3 + D.st * w.t0);
#pragma unroll FLOF_TM_UNROLL
        for (int t = w.t0; t < w.t1; ++t) {
            const int64_t c = c3 + D.st * t;
            const bool tn = t + 1 < D.T;
            const float4 pp = tn ? PS::ld(p, c + D.st) : zero4;
            const float ip = tn ? IM::ld(i1, c + D.st) : 0.f;
            float4 qc;
            if (w.sp_int && t >= 1 && t < D.T - 1)
                qc = apply_A_tm<PS>(p, c, pm, pc, pp, grad_i1<IM>(i1, c, D, im, ip, dxf), G);
            else
                qc = apply_I(pc);
            dd_add4(acc, pc, qc);
            pm = pc;
            pc = pp;
            im = ic;
            ic = ip;
        }
    }
    acc = block_dd(acc);
    if (threadIdx.x == 0) part_dd[bid] = acc;
}
template <class PS, class IM>
__global__ void __launch_bounds__(256)
k_cg_pq_tm(const typename PS::T *__restrict__ p, const typename IM::T *__restrict__ i1,
           LevelGeom G, TMarch M, float dxf, const CgState *st, DD *part_dd)
{
    cg_pq_tm_body<PS, IM>((int)blockIdx.x, (int)gridDim.x, p, i1, G, M, dxf, st, part_dd);
}

// r -= alpha (A p) (optflow4d.cpp:310), z = r diag^-1, z.r and the signed
// max of r (:311-317); x += alpha p (:309) is applied by k_cg_dir_tm
template <class PS, class IM>
__device__ __forceinline__ void
cg_update_tm_body(int bid, int nblk, float4 *__restrict__ r,
                  const typename PS::T *__restrict__ p, const typename IM::T *__restrict__ i1,
                  const LevelGeom &G, const TMarch &M, float dxf, CgState *st,
                  const DD *part_pq, DD *part_dd, float *part_max)
{
    if (st->done) {
        cg_after_last(st, bid);
        return;
    }
    const Dims4 &D = G.D;
    const TmCell w = tm_cell(D, M, bid, nblk);
    const double nalpha = -cg_alpha(st, part_pq, nblk, bid);
    DD sig{0., 0.};
    float mx = -INFINITY;
    if (w.live) {
        const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
        const int64_t c3 = w.c3;
        float4 pm = w.t0 > 0 ? PS::ld(p, c3 + D.st * (w.t0 - 1)) : zero4;
        float4 pc = PS::ld(p, c3 + D.st * w.t0);
        float im = w.t0 > 0 ? IM::ld(i1, c3 + D.st * (w.t0 - 1)) : 0.f;
        float ic = IM::ld(i1, c3 + D.st * w.t0);
#pragma unroll FLOF_TM_UNROLL
        for (int t = w.t0; t < w.t1; ++t) {
            const int64_t c = c3 + D.st * t;
            const bool tn = t + 1 < D.T;
            const float4 pp = tn ? PS::ld(p, c + D.st) : zero4;
            const float ip = tn ? IM::ld(i1, c + D.st) : 0.f;
            float4 qc, pr;
            if (w.sp_int && t >= 1 && t < D.T - 1) {
                const float4 gc = grad_i1<IM>(i1, c, D, im, ip, dxf);
                qc = apply_A_tm<PS>(p, c, pm, pc, pp, gc, G);
                pr = prec_f32(gc, G);
            } else {
                qc = apply_I(pc);
                pr = make_float4(1.f, 1.f, 1.f, 1.f);
            }
            float4 rc = ld_stream(r + c);
            rc.x = (float)((double)rc.x + nalpha * (double)qc.x);
            rc.y = (float)((double)rc.y + nalpha * (double)qc.y);
            rc.z = (float)((double)rc.z + nalpha * (double)qc.z);
            rc.w = (float)((double)rc.w + nalpha * (double)qc.w);
            st_stream(r + c, rc);
            const float4 z = f4mul(rc, pr);
            mx = fmaxf(mx, fmaxf(fmaxf(rc.x, rc.y), fmaxf(rc.z, rc.w)));
            dd_add4(sig, z, rc);
            pm = pc;
            pc = pp;
            im = ic;
            ic = ip;
        }
    }
    sig = block_dd(sig);
    mx = block_max(mx);
    if (threadIdx.x == 0) {
        part_dd[bid] = sig;
        part_max[bid] = mx;
    }
}
template <class PS, class IM>
__global__ void __launch_bounds__(256)
k_cg_update_tm(float4 *__restrict__ r, const typename PS::T *__restrict__ p,
               const typename IM::T *__restrict__ i1, LevelGeom G, TMarch M, float dxf,
               CgState *st, const DD *part_pq, DD *part_dd, float *part_max)
{
    cg_update_tm_body<PS, IM>((int)blockIdx.x, (int)gridDim.x, r, p, i1, G, M, dxf, st,
                              part_pq, part_dd, part_max);
}

// x += alpha p (optflow4d.cpp:309), then p = z + beta p with z = r diag^-1
// (:318-323) unless this was the last iteration.  The direction is double
// buffered: p_k is read from p and p_(k+1) written to pn, which held
// p_(k-1).  The x update of an even iteration k is deferred to the next
// one: there (ODD) x += alpha_(k-1) p_(k-1), then x += alpha_k p_k, the same
// two roundings in the same order, so x is the same while it is read and
// written every other iteration only.  dir_mode 2 (the last iteration)
// applies the pending updates and leaves p.
__device__ __forceinline__ float4 x_plus(const float4 &x, double a, const float4 &p)
{
    return make_float4((float)((double)x.x + a * (double)p.x),
                       (float)((double)x.y + a * (double)p.y),
                       (float)((double)x.z + a * (double)p.z),
                       (float)((double)x.w + a * (double)p.w));
}

template <class PS, bool ODD, class IM>
__device__ __forceinline__ void
cg_dir_tm_body(int bid, int nblk, const typename PS::T *__restrict__ p, typename PS::T *pn,
               float4 *__restrict__ x, const float4 *__restrict__ r,
               const typename IM::T *__restrict__ i1, const LevelGeom &G, const TMarch &M,
               float dxf, const CgState *st)
{
    const int mode = st->dir_mode;
    if (mode != 1 && mode != 2) return;
    const bool xupd = ODD || mode == 2;
    const Dims4 &D = G.D;
    const TmCell w = tm_cell(D, M, bid, nblk);
    if (!w.live) return;
    const double alpha = st->alpha, alpha_prev = st->alpha_prev, beta = st->beta;
    const int64_t c3 = w.c3;
    float im = w.t0 > 0 ? IM::ld(i1, c3 + D.st * (w.t0 - 1)) : 0.f;
    float ic = IM::ld(i1, c3 + D.st * w.t0);
#pragma unroll FLOF_TM_UNROLL
    for (int t = w.t0; t < w.t1; ++t) {
        const int64_t c = c3 + D.st * t;
        const float ip = t + 1 < D.T ? IM::ld(i1, c + D.st) : 0.f;
        const float4 pc = PS::ld_once(p, c);
        if (xupd) {
            float4 xc = ld_stream(x + c);
            if (ODD) xc = x_plus(xc, alpha_prev, PS::ld_once(pn, c));
            st_stream(x + c, x_plus(xc, alpha, pc));
        }
        if (mode == 1) {
            float4 pr = make_float4(1.f, 1.f, 1.f, 1.f);
            if (w.sp_int && t >= 1 && t < D.T - 1)
                pr = prec_f32(grad_i1<IM>(i1, c, D, im, ip, dxf), G);
            const float4 z = f4mul(ld_stream(r + c), pr);
            float4 np;
            np.x = (float)((double)z.x + beta * (double)pc.x);
            np.y = (float)((double)z.y + beta * (double)pc.y);
            np.z = (float)((double)z.z + beta * (double)pc.z);
            np.w = (float)((double)z.w + beta * (double)pc.w);
            PS::st_once(pn, c, np);
        }
        im = ic;
        ic = ip;
    }
}
template <class PS, bool ODD, class IM>
__global__ void __launch_bounds__(256)
k_cg_dir_tm(const typename PS::T *__restrict__ p, typename PS::T *pn, float4 *__restrict__ x,
            const float4 *__restrict__ r, const typename IM::T *__restrict__ i1, LevelGeom G,
            TMarch M, float dxf, const CgState *st)
{
    cg_dir_tm_body<PS, ODD, IM>((int)blockIdx.x, (int)gridDim.x, p, pn, x, r, i1, G, M, dxf,
                                st);
}

// ---------------------------------------------------------------------------
// LDS-tiled variant of the two stencil kernels (k_cg_pq_tm, k_cg_update_tm).
// A workgroup owns a tile of 64 (x) x 4 (y) x 4 (z) cells and marches it
// through its t segment; thread (lane = x, wave = y) owns the column of 4
// z cells.  Per t step the tile's p(t) and i1(t) go to LDS together with the
// y-halo rows, x-halo columns and the z-halo planes, in one of two buffers
// that alternate by step (so one barrier per step suffices, and the own
// cells' t-1 products are read back from the other buffer); the t+1 values
// of the own cells stay in registers.  The halo of step t+1 and the own values
// of t+2 are loaded while step t computes (the loads are waited for only
// when they are written to LDS at the next step).  Every cell is computed
// with the same operands in the same order as in the untiled kernels, so the
// results are identical; only the neighbour loads change (one halo value per
// own value instead of six neighbours through the caches).
struct TileMarch {
    int ntx, nty, ntz;  // tiles along x, y, z
    int ntile;          // ntx * nty * ntz
    int seglen, nseg;   // t segments
    int total;          // ntile * nseg work items
    int xcd;            // 1: the grid (a multiple of 8) gives each XCD a
                        // contiguous range of work items
};
constexpr int LT_X = 64, LT_Y = 4, LT_Z = 4;
constexpr int LS_X = LT_X + 2, LS_Y = LT_Y + 2;
constexpr int LS_N = LT_Z * LS_Y * LS_X;  // 1584 entries: 4 z x 6 rows x 66

// LDS slot of (z q in [0,4), row ry in [-1,4], column x in [-1,64])
__device__ __forceinline__ int lslot(int q, int ry, int x)
{
    return (q * LS_Y + (ry + 1)) * LS_X + (x + 1);
}

__device__ __forceinline__ float4 f4scale(float s, const float4 &v)
{
    return make_float4(s * v.x, s * v.y, s * v.z, s * v.w);
}

template <class PS, bool UPDATE>
__global__ void __launch_bounds__(256)
k_cg_stencil_lds(const typename PS::T *__restrict__ p, const float *__restrict__ i1,
                 float4 *__restrict__ r, LevelGeom G, TileMarch M, float dxf, CgState *st,
                 const DD *part_pq, DD *part_dd, float *part_max)
{
    if (st->done) {  // uniform over the launch
        if (UPDATE) cg_after_last(st);
        return;
    }
    double nalpha = 0.;
    // The neighbour sum of A p only ever uses offd * p(neighbour), so LDS
    // holds those products (each formed once per cell instead of once per
    // use; the same float products, so the sums are unchanged).  The tile
    // arrays are double-buffered by t step (one barrier per step); the
    // z-halo planes need not be: each thread reads only the slots it wrote.
    __shared__ float4 sp2[2][LS_N];         // offd * p: tile, y-halo, x-halo
    __shared__ float si2[2][LS_N];          // i1 of the same cells
    __shared__ float4 szp[2][LT_Y * LT_X];  // offd * p: z-halo planes z0-1, z0+4
    __shared__ float szi[2][LT_Y * LT_X];
    const Dims4 &D = G.D;
    const float od = G.offd;
    DD acc{0., 0.};
    float mx = -INFINITY;
    int L = blockIdx.x;
    if (M.xcd) {
        const int per = (int)gridDim.x / FLOF_NXCD;
        L = (L % FLOF_NXCD) * per + L / FLOF_NXCD;
    }
    if (L < M.total) {  // uniform over the workgroup
        const int seg = L / M.ntile;
        const int tile = L - seg * M.ntile;
        const int tx = tile % M.ntx, tyz = tile / M.ntx;
        const int ty = tyz % M.nty, tz = tyz / M.nty;
        const int x0 = tx * LT_X, y0 = ty * LT_Y, z0 = tz * LT_Z;
        const int t0 = seg * M.seglen, t1 = min(D.T, t0 + M.seglen);
        const int tid = (int)threadIdx.x;
        const int lane = tid & 63, wy = tid >> 6;
        const int i = x0 + lane, j = y0 + wy;
        const bool in_xy = i < D.X && j < D.Y;
        const bool sp_xy = i >= 1 && j >= 1 && i < D.X - 1 && j < D.Y - 1;
        const int sz = (int)D.sz, stt = (int)D.st;
        const int c3 = i + D.X * j;  // + sz * k (cell indices fit 31 bits)
        bool own[LT_Z];
#pragma unroll
        for (int q = 0; q < LT_Z; ++q) own[q] = in_xy && z0 + q < D.Z;
        const bool has_zm = in_xy && z0 >= 1;
        const bool has_zp = in_xy && z0 + LT_Z < D.Z;
        // Loads are unconditional: where the cell of a lane lies outside the
        // grid (x, y, z, or t beyond the last step) the lane reads cell 0
        // instead.  Such values only reach registers and LDS slots that no
        // computation reads: only interior cells read neighbours, and the 8
        // neighbours of an interior cell lie inside the grid.  The loads then
        // need no exec-mask branches (p.Ap at 64^4: 92 -> 80 us).
        const int czm = c3 + sz * (z0 - 1), czp = c3 + sz * (z0 + LT_Z);
        const int czm_l = has_zm ? czm : 0, czp_l = has_zp ? czp : 0;
        // y-halo rows y0-1 / y0+4 of the 4 z planes: 512 cells, 2 per thread
        const int hy_side = tid & 1, hy_q = (tid >> 1) & 3, hy_lane = tid >> 3;
        const int hy_row = hy_side ? LT_Y : -1;
        bool hy_ok[2];
        int hy_c[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int hi = x0 + hy_lane + 32 * h, hj = y0 + hy_row, hk = z0 + hy_q;
            hy_ok[h] = hi < D.X && hj >= 0 && hj < D.Y && hk < D.Z;
            hy_c[h] = hy_ok[h] ? hi + D.X * hj + sz * hk : 0;
        }
        // x-halo columns x0-1 / x0+64 of the 4 rows and 4 planes: threads 0..31
        const int hx_side = tid & 1, hx_r = (tid >> 1) & 3, hx_q = (tid >> 3) & 3;
        const int hx_col = hx_side ? LT_X : -1;
        bool hx_ok;
        int hx_c;
        {
            const int hi = x0 + hx_col, hj = y0 + hx_r, hk = z0 + hx_q;
            hx_ok = tid < 32 && hi >= 0 && hi < D.X && hj < D.Y && hk < D.Z;
            hx_c = hx_ok ? hi + D.X * hj + sz * hk : 0;
        }
        const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
        // halo of one t step in registers until it is written to LDS
        float4 hy_v[2], hx_v, zm_v, zp_v;
        float hy_i[2], hx_i, zm_i, zp_i;
        auto load_halo = [&](int t) {
            const int ct = stt * t;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                hy_v[h] = PS::ld(p, hy_c[h] + ct);
                hy_i[h] = i1[hy_c[h] + ct];
            }
            hx_v = hx_ok ? PS::ld(p, hx_c + ct) : zero4;
            hx_i = hx_ok ? i1[hx_c + ct] : 0.f;
            zm_v = PS::ld(p, czm_l + ct);
            zm_i = i1[czm_l + ct];
            zp_v = PS::ld(p, czp_l + ct);
            zp_i = i1[czp_l + ct];
        };
        // own cells: offd * p at t-1; p at t, t+1 and (prefetched) t+2
        float4 pc[LT_Z], pp[LT_Z], pn[LT_Z];
        float im[LT_Z], ic[LT_Z], ip[LT_Z], in[LT_Z];
#pragma unroll
        for (int q = 0; q < LT_Z; ++q) {
            const int c = c3 + sz * (z0 + q) + stt * t0;
            const bool hm = own[q] && t0 > 0, hp = own[q] && t0 + 1 < D.T;
            sp2[1][lslot(q, wy, lane)] = f4scale(od, PS::ld(p, hm ? c - stt : 0));
            im[q] = i1[hm ? c - stt : 0];
            pc[q] = PS::ld(p, own[q] ? c : 0);
            ic[q] = i1[own[q] ? c : 0];
            pp[q] = PS::ld(p, hp ? c + stt : 0);
            ip[q] = i1[hp ? c + stt : 0];
        }
        load_halo(t0);
        const int wl = wy * LT_X + lane;
        // the update's streamed r: this step's in ro, the next step's
        // requested while this one computes
        float4 ro[LT_Z];
        if (UPDATE) {
#pragma unroll
            for (int q = 0; q < LT_Z; ++q)
                ro[q] = ld_stream(r + (own[q] ? c3 + sz * (z0 + q) + stt * t0 : 0));
        }
        // alpha while the first step's loads are in flight (workgroup 0,
        // which stores it, always has a work item)
        if (UPDATE) nalpha = -cg_alpha(st, part_pq, (int)gridDim.x);
        for (int t = t0; t < t1; ++t) {
            const int ct = stt * t;
            // step t writes buffer (t - t0) & 1; the other one holds the own
            // cells' offd * p(t-1) (written by this thread at the last step;
            // before the loop for t0).  A buffer is written again two steps
            // later, after every thread has passed the barrier between.
            float4 *sp = sp2[(t - t0) & 1];
            float *si = si2[(t - t0) & 1];
            const float4 *spm = sp2[((t - t0) & 1) ^ 1];
#pragma unroll
            for (int q = 0; q < LT_Z; ++q) {
                sp[lslot(q, wy, lane)] = f4scale(od, pc[q]);
                si[lslot(q, wy, lane)] = ic[q];
            }
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                sp[lslot(hy_q, hy_row, hy_lane + 32 * h)] = f4scale(od, hy_v[h]);
                si[lslot(hy_q, hy_row, hy_lane + 32 * h)] = hy_i[h];
            }
            if (tid < 32) {
                sp[lslot(hx_q, hx_r, hx_col)] = f4scale(od, hx_v);
                si[lslot(hx_q, hx_r, hx_col)] = hx_i;
            }
            szp[0][wl] = f4scale(od, zm_v);
            szi[0][wl] = zm_i;
            szp[1][wl] = f4scale(od, zp_v);
            szi[1][wl] = zp_i;
            // prefetch of the next step: r, the halo, own p at t+2
            float4 rq[LT_Z];
            if (UPDATE && t + 1 < t1) {
#pragma unroll
                for (int q = 0; q < LT_Z; ++q)
                    rq[q] = ld_stream(r + (own[q] ? c3 + sz * (z0 + q) + ct + stt : 0));
            }
            if (t + 1 < t1) {
                load_halo(t + 1);
#pragma unroll
                for (int q = 0; q < LT_Z; ++q) {
                    const bool hn = own[q] && t + 2 < D.T;
                    const int c = hn ? c3 + sz * (z0 + q) + ct + 2 * stt : 0;
                    pn[q] = PS::ld(p, c);
                    in[q] = i1[c];
                }
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < LT_Z; ++q) {
                if (!own[q]) continue;
                const int k = z0 + q;
                float4 qc, pr;
                // Branch-free: every lane evaluates the interior row; on an
                // outer-layer cell the result is replaced by the identity
                // row's (the neighbour values read there lie in the tile's
                // LDS slots and are discarded).
                const bool inner = sp_xy && k >= 1 && k < D.Z - 1 && t >= 1 && t < D.T - 1;
                {
                    float4 v = zero4;
                    if (G.use_offd) {
                        // offd * p of the neighbours in the order t-1, z-1, y-1,
                        // x-1, x+1, y+1, z+1, t+1
                        const float4 nb[8] = {
                            spm[lslot(q, wy, lane)],
                            q > 0 ? sp[lslot(q > 0 ? q - 1 : 0, wy, lane)] : szp[0][wl],
                            sp[lslot(q, wy - 1, lane)],
                            sp[lslot(q, wy, lane - 1)],
                            sp[lslot(q, wy, lane + 1)],
                            sp[lslot(q, wy + 1, lane)],
                            q < LT_Z - 1 ? sp[lslot(q < LT_Z - 1 ? q + 1 : 0, wy, lane)]
                                         : szp[1][wl],
                            f4scale(od, pp[q])};
#pragma unroll
                        for (int m = 0; m < 8; ++m) {
                            v.x = v.x + nb[m].x;
                            v.y = v.y + nb[m].y;
                            v.z = v.z + nb[m].z;
                            v.w = v.w + nb[m].w;
                        }
                    }
                    float4 gc;
                    gc.x = (si[lslot(q, wy, lane + 1)] - si[lslot(q, wy, lane - 1)]) * dxf;
                    gc.y = (si[lslot(q, wy + 1, lane)] - si[lslot(q, wy - 1, lane)]) * dxf;
                    {
                        const float izn = q > 0 ? ic[q > 0 ? q - 1 : 0] : szi[0][wl];
                        const float izq = q < LT_Z - 1 ? ic[q < LT_Z - 1 ? q + 1 : 0] : szi[1][wl];
                        gc.z = (izq - izn) * dxf;
                    }
                    gc.w = (ip[q] - im[q]) * dxf;
                    const float gg[4] = {gc.x, gc.y, gc.z, gc.w};
                    const float pv[4] = {pc[q].x, pc[q].y, pc[q].z, pc[q].w};
                    float o[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                    for (int d = 0; d < 4; ++d) {
#pragma unroll
                        for (int m = 0; m < 4; ++m) {
                            float b = gg[d] * gg[m];
                            if (d == m) b = b + G.diag;
                            o[d] = o[d] + b * pv[m];
                        }
                    }
                    // an outer-layer cell's row is the identity: q = p and
                    // the preconditioner 1 (what the interior row gives with
                    // a neighbour sum of +0, g = 0 and a diagonal of 1, for
                    // the +0 that p always holds there)
                    qc = inner ? make_float4(o[0], o[1], o[2], o[3]) : pc[q];
                    if (UPDATE) {
                        pr = prec_f32(gc, G.diag, G.fast_rcp);
                        if (!inner) pr = make_float4(1.f, 1.f, 1.f, 1.f);
                    }
                }
                if (!UPDATE) {
                    dd_add4(acc, pc[q], qc);
                } else {
                    const int c = c3 + sz * k + ct;
                    float4 rc = ro[q];
                    rc.x = (float)((double)rc.x + nalpha * (double)qc.x);
                    rc.y = (float)((double)rc.y + nalpha * (double)qc.y);
                    rc.z = (float)((double)rc.z + nalpha * (double)qc.z);
                    rc.w = (float)((double)rc.w + nalpha * (double)qc.w);
                    st_stream(r + c, rc);
                    const float4 z = f4mul(rc, pr);
                    mx = fmaxf(mx, fmaxf(fmaxf(rc.x, rc.y), fmaxf(rc.z, rc.w)));
                    dd_add4(acc, z, rc);
                }
            }
            // offd * p(t) of the own cells stays in this step's buffer as the
            // t-1 neighbour of the next step (the thread's own slots)
            if (UPDATE && t + 1 < t1) {
#pragma unroll
                for (int q = 0; q < LT_Z; ++q) ro[q] = rq[q];
            }
#pragma unroll
            for (int q = 0; q < LT_Z; ++q) {
                pc[q] = pp[q];
                pp[q] = pn[q];
                im[q] = ic[q];
                ic[q] = ip[q];
                ip[q] = in[q];
            }
        }
    }
    acc = block_dd(acc);
    if (UPDATE) mx = block_max(mx);
    if (threadIdx.x == 0) {
        part_dd[blockIdx.x] = acc;
        if (UPDATE) part_max[blockIdx.x] = mx;
    }
}

// the FP32 direction of the set-up (or warm start) into reduced storage
template <class PS>
__global__ void __launch_bounds__(256)
k_p_pack(const float4 *__restrict__ src, typename PS::T *__restrict__ dst, int64_t n)
{
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n;
         c += (int64_t)gridDim.x * blockDim.x)
        PS::st(dst, c, src[c]);
}

// vel_out (=) or (+=) x / mDx; a failed solve contributes zero
__global__ void k_of_finish(const float4 *__restrict__ x, float4 *__restrict__ out,
                            int64_t n, float mDx, int accumulate, const CgState *st)
{
    const bool zero = st->zero != 0;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n;
         c += (int64_t)gridDim.x * blockDim.x) {
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (!zero) {
            const float4 xc = x[c];
            v = make_float4(xc.x / mDx, xc.y / mDx, xc.z / mDx, xc.w / mDx);
        }
        out[c] = accumulate ? f4add(out[c], v) : v;
    }
}

// ---------------------------------------------------------------------------
// Passes over 3^4 neighbourhoods with LDS-staged planes.  A workgroup owns a
// TX x TY x 2 column of cells (64 x 2 x 2, or 32 x 4 x 2 on grids whose
// width would leave 64-wide tiles largely empty) for a segment of t and
// marches through it, keeping the haloed planes t-1, t, t+1 ((TX+2) x (TY+2)
// x 4 cells each) in a ring of three LDS slots, so the 81 reads per cell hit
// LDS instead of the vector caches.  Interior cells have all 81 neighbours in
// the grid.
//  MARCH_BOX81: knCvExpolBlur4d (optflow4d.cpp:613-626): unmarked interior
//    cells take the 3^4 box mean, summed in the reference's loop order (t,
//    z, y, x outermost to innermost, FP32, starting from zero); other cells
//    are copied.  `aux` is the marker grid.
//  MARCH_BLUR1: the Gaussian blur of k_blur4d for a window of s = 1 (never
//    clipped for an interior cell), same order and operations; other cells
//    are left unchanged.  `aux` is the weight table by squared distance.
enum { MARCH_BOX81 = 0, MARCH_BLUR1 = 1 };
template <int TX>
struct BoxTile {
    static constexpr int X = TX, Y = TX == 64 ? 2 : 4, Z = 2;
    static constexpr int HX = X + 2, HY = Y + 2, HZ = Z + 2;
    static constexpr int PLANE = HX * HY * HZ;        // 1056 / 816 cells
    static constexpr int LOADS = (PLANE + 255) / 256;  // per thread and plane
    static_assert(X * Y * Z == 256, "one cell per thread");
};

struct BoxMarch {
    int tx;                    // tile width: 64 or 32
    int ntx, nty, ntz, ntile;  // column tiles
    int seglen, nseg;          // t segments
};

template <class T, int OP, int TX>
__global__ void __launch_bounds__(256)
k_march3_lds(const T *__restrict__ a, T *__restrict__ out, const float *__restrict__ aux,
             Dims4 D, BoxMarch B, float eps)
{
    typedef VOps<T> O;
    typedef BoxTile<TX> BT;
    constexpr int BX_X = BT::X, BX_Y = BT::Y, BX_Z = BT::Z;
    constexpr int BX_HX = BT::HX, BX_HY = BT::HY;
    constexpr int BX_PLANE = BT::PLANE, BX_LOADS = BT::LOADS;
    __shared__ T pl[3][BX_PLANE];
    const int tile = blockIdx.x % B.ntile;
    const int seg = blockIdx.x / B.ntile;
    const int bx = tile % B.ntx, by = (tile / B.ntx) % B.nty, bz = tile / (B.ntx * B.nty);
    const int x0 = bx * BX_X, y0 = by * BX_Y, z0 = bz * BX_Z;
    const int t0 = seg * B.seglen;
    const int t1 = min(t0 + B.seglen, D.T);
    const int tx = threadIdx.x % BX_X, ty = (threadIdx.x / BX_X) % BX_Y,
              tz = threadIdx.x / (BX_X * BX_Y);
    const int gx = x0 + tx, gy = y0 + ty, gz = z0 + tz;
    const bool own = gx < D.X && gy < D.Y && gz < D.Z;

    // halo cells this thread stages: e = threadIdx.x + 256 m
    int64_t hoff[BX_LOADS];  // offset inside a t-plane, -1: outside the grid
#pragma unroll
    for (int m = 0; m < BX_LOADS; ++m) {
        const int e = threadIdx.x + 256 * m;
        hoff[m] = -1;
        if (e < BX_PLANE) {
            const int hx = e % BX_HX, hy = (e / BX_HX) % BX_HY, hz = e / (BX_HX * BX_HY);
            const int x = x0 + hx - 1, y = y0 + hy - 1, z = z0 + hz - 1;
            if (x >= 0 && x < D.X && y >= 0 && y < D.Y && z >= 0 && z < D.Z)
                hoff[m] = x + (int64_t)D.X * y + D.sz * z;
        }
    }
    T st[BX_LOADS];
    auto fetch = [&](int p) {
#pragma unroll
        for (int m = 0; m < BX_LOADS; ++m)
            st[m] = (p >= 0 && p < D.T && hoff[m] >= 0) ? a[hoff[m] + D.st * p] : O::zero();
    };
    auto stash = [&](int p) {
        T *s = pl[(p + 3) % 3];
#pragma unroll
        for (int m = 0; m < BX_LOADS; ++m) {
            const int e = threadIdx.x + 256 * m;
            if (e < BX_PLANE) s[e] = st[m];
        }
    };
    // planes t0-1 and t0 (t0-1 may be -1: never read, see below)
    fetch(t0 - 1);
    stash(t0 - 1);
    fetch(t0);
    stash(t0);
    const int cl = (tx + 1) + BX_HX * ((ty + 1) + BX_HY * (tz + 1));
    for (int t = t0; t < t1; ++t) {
        fetch(t + 1);
        __syncthreads();  // every wave is done with plane t-2's slot
        stash(t + 1);
        __syncthreads();
        if (!own) continue;
        const int64_t c = gx + (int64_t)D.X * gy + D.sz * gz + D.st * t;
        const bool in = interior(D, gx, gy, gz, t);
        if (OP == MARCH_BOX81) {
            if (!in || aux[c] != 0.f) {
                out[c] = pl[t % 3][cl];
                continue;
            }
            T val = O::zero();
            for (int vt = -1; vt <= 1; ++vt) {
                const T *s = pl[(t + vt + 3) % 3];
#pragma unroll
                for (int zk = -1; zk <= 1; ++zk)
#pragma unroll
                    for (int yj = -1; yj <= 1; ++yj)
#pragma unroll
                        for (int xi = -1; xi <= 1; ++xi)
                            val = O::add(val, s[cl + xi + BX_HX * (yj + BX_HY * zk)]);
            }
            const double f = 1. / 81.0;
            if constexpr (sizeof(T) == 16) {
                const float4 v = *(const float4 *)&val;
                ((float4 *)out)[c] = make_float4((float)(v.x * f), (float)(v.y * f),
                                                 (float)(v.z * f), (float)(v.w * f));
            } else {
                ((float *)out)[c] = (float)(*(const float *)&val * f);
            }
        } else {
            if (!in) continue;
            T val = O::zero();
            float weight = 0.f;
            for (int vt = -1; vt <= 1; ++vt) {
                const T *s = pl[(t + vt + 3) % 3];
                // z not unrolled: 9 taps in flight keep the Vec4 form at
                // 2 or more waves per SIMD
#pragma unroll 1
                for (int zk = -1; zk <= 1; ++zk)
#pragma unroll
                    for (int yj = -1; yj <= 1; ++yj)
#pragma unroll
                        for (int xi = -1; xi <= 1; ++xi) {
                            const float w = aux[vt * vt + zk * zk + yj * yj + xi * xi];
                            weight = weight + w;
                            val = O::add(val, O::muls(s[cl + xi + BX_HX * (yj + BX_HY * zk)], w));
                        }
            }
            if (weight > eps) {
                if constexpr (sizeof(T) == 16) {
                    const float4 v = *(const float4 *)&val;
                    ((float4 *)out)[c] = make_float4(v.x / weight, v.y / weight, v.z / weight,
                                                     v.w / weight);
                } else {
                    ((float *)out)[c] = *(const float *)&val / weight;
                }
            } else {
                out[c] = pl[t % 3][cl];
            }
        }
    }
}

// per_cu: workgroups per CU that the t segments aim at
static BoxMarch make_boxmarch(flof_ctx *ctx, const Dims4 &D, int per_cu = 6)
{
    BoxMarch B;
    // 32-wide tiles where they cover the x extent with 10% fewer lanes
    const int w64 = (D.X + 63) / 64 * 64, w32 = (D.X + 31) / 32 * 32;
    B.tx = 10 * w32 < 9 * w64 ? 32 : 64;
    const int ty = B.tx == 64 ? BoxTile<64>::Y : BoxTile<32>::Y;
    B.ntx = (D.X + B.tx - 1) / B.tx;
    B.nty = (D.Y + ty - 1) / ty;
    B.ntz = (D.Z + 1) / 2;
    B.ntile = B.ntx * B.nty * B.ntz;
    // the default 6: twice the workgroups per CU that LDS admits at once for
    // 64-wide Vec4 tiles (three of 50 KB)
    const int target = ctx->num_cu * per_cu;
    int nseg = (target + B.ntile - 1) / B.ntile;
    nseg = std::max(1, std::min(nseg, D.T));
    B.seglen = (D.T + nseg - 1) / nseg;
    B.nseg = (D.T + B.seglen - 1) / B.seglen;
    return B;
}

template <class T, int OP>
static void launch_march3(flof_ctx *ctx, const BoxMarch &B, const T *src, T *dst,
                          const float *aux, const Dims4 &D, float eps)
{
    const dim3 grid(B.ntile * B.nseg);
    if (B.tx == 32)
        hipLaunchKernelGGL((k_march3_lds<T, OP, 32>), grid, dim3(256), 0, ctx->stream, src, dst,
                           aux, D, B, eps);
    else
        hipLaunchKernelGGL((k_march3_lds<T, OP, 64>), grid, dim3(256), 0, ctx->stream, src, dst,
                           aux, D, B, eps);
}

// Gaussian blur (optflow4d.cpp:127-174): every interior cell sums the
// (2s+1)^4 window clipped to the grid in the reference's order (t, z, y, x
// from low to high), the weights and the values in FP32, then divides by
// the weight sum.  The weight of a squared distance d is
// (float)exp(-d / (2 sigma^2)) evaluated in double on the host, one table
// entry per d in [0, 4 s^2], so the device uses the host's exp.  This
// gathering form serves windows of s >= 3; s = 1 and s = 2 run as the LDS
// marches k_march3_lds and k_blur5_lds with the same sums.
template <class T>
__global__ void __launch_bounds__(256)
k_blur4d(const T *__restrict__ src, T *__restrict__ dst, Dims4 D, Coord4 C, int s,
         const float *__restrict__ wt, float eps)
{
    const int64_t n = (int64_t)D.st * D.T;
    FOR_XCD(c, n)
    {
        int i, j, k, t;
        decompose((uint32_t)c, C, i, j, k, t);
        if (!interior(D, i, j, k, t)) continue;
        T val = VOps<T>::zero();
        float weight = 0.f;
        for (int vt = max(t - s, 0); vt <= min(t + s, D.T - 1); ++vt)
            for (int zk = max(k - s, 0); zk <= min(k + s, D.Z - 1); ++zk)
                for (int yj = max(j - s, 0); yj <= min(j + s, D.Y - 1); ++yj) {
                    const int dtzy = (vt - t) * (vt - t) + (zk - k) * (zk - k) +
                                     (yj - j) * (yj - j);
                    const T *row = src + (int64_t)D.st * vt + D.sz * zk + (int64_t)D.X * yj;
                    for (int xi = max(i - s, 0); xi <= min(i + s, D.X - 1); ++xi) {
                        const float w = wt[dtzy + (xi - i) * (xi - i)];
                        weight = weight + w;
                        val = VOps<T>::add(val, VOps<T>::muls(row[xi], w));
                    }
                }
        if (weight > eps) {
            if constexpr (sizeof(T) == 16) {
                const float4 v = *(const float4 *)&val;
                ((float4 *)dst)[c] = make_float4(v.x / weight, v.y / weight, v.z / weight,
                                                 v.w / weight);
            } else {
                ((float *)dst)[c] = *(const float *)&val / weight;
            }
        } else {
            dst[c] = src[c];
        }
    }
}

// The same blur for windows of s = 2 (5^4 taps) with LDS-staged planes.  A
// workgroup owns a TX x TY x 2 column of cells of one component (the four
// components of a Vec4 grid are four independent sums, one workgroup each)
// and marches it through a t segment, keeping the planes t-2 .. t+2 with a
// two-cell halo ((TX+4) x (TY+4) x 6 values) in a ring of five LDS slots, so
// the 625 reads per cell hit LDS instead of the vector caches.  Window
// positions outside the grid hold +0: the reference skips them, and adding
// w * (+0) = +0 to a sum that started at +0 changes no bit (such a sum is
// never -0).  The weight sum of a clipped window depends only on how far
// the faces clip it: `wsum` holds it for each of the 4^4 clipping patterns
// (per dimension the window of an interior cell is [-1, 2], [-2, 1], [-2, 2]
// or [-1, 1]), summed on the host in the reference's order.
struct Blur5W {
    float w[17];  // weight by squared distance 0..16
};
__device__ __forceinline__ int blur5_pattern(int i, int n)
{
    return (min(2, i) - 1) * 2 + (min(2, n - 1 - i) - 1);
}
template <int TX>
__global__ void __launch_bounds__(256)
k_blur5_lds(const float *__restrict__ src, float *__restrict__ dst, int cstride, Dims4 D,
            BoxMarch B, Blur5W W, const float *__restrict__ wsum, float eps)
{
    typedef BoxTile<TX> BT;
    constexpr int HX = BT::X + 4, HY = BT::Y + 4, HZ = BT::Z + 4;
    constexpr int PLANE = HX * HY * HZ;  // 2448 / 1728 values
    constexpr int LOADS = (PLANE + 255) / 256;
    __shared__ float pl[5][PLANE];
    const int per = B.ntile * B.nseg;
    const int comp = blockIdx.x / per;
    const int rem = blockIdx.x - comp * per;
    const int tile = rem % B.ntile, seg = rem / B.ntile;
    const int bx = tile % B.ntx, by = (tile / B.ntx) % B.nty, bz = tile / (B.ntx * B.nty);
    const int x0 = bx * BT::X, y0 = by * BT::Y, z0 = bz * BT::Z;
    const int t0 = seg * B.seglen;
    const int t1 = min(t0 + B.seglen, D.T);
    const int tx = threadIdx.x % BT::X, ty = (threadIdx.x / BT::X) % BT::Y,
              tz = threadIdx.x / (BT::X * BT::Y);
    const int gx = x0 + tx, gy = y0 + ty, gz = z0 + tz;
    const bool own = gx < D.X && gy < D.Y && gz < D.Z;
    src += comp;  // value of cell c at [c * cstride]
    dst += comp;
    // window cells this thread stages: e = threadIdx.x + 256 m
    int64_t hoff[LOADS];  // cell offset inside a t-plane, -1: outside the grid
#pragma unroll
    for (int m = 0; m < LOADS; ++m) {
        const int e = threadIdx.x + 256 * m;
        hoff[m] = -1;
        if (e < PLANE) {
            const int hx = e % HX, hy = (e / HX) % HY, hz = e / (HX * HY);
            const int x = x0 + hx - 2, y = y0 + hy - 2, z = z0 + hz - 2;
            if (x >= 0 && x < D.X && y >= 0 && y < D.Y && z >= 0 && z < D.Z)
                hoff[m] = x + (int64_t)D.X * y + D.sz * z;
        }
    }
    float stv[LOADS];
    auto fetch = [&](int p) {
#pragma unroll
        for (int m = 0; m < LOADS; ++m)
            stv[m] = (p >= 0 && p < D.T && hoff[m] >= 0) ? src[(hoff[m] + D.st * p) * cstride]
                                                         : 0.f;
    };
    auto stash = [&](int p) {
        float *sl = pl[(p + 10) % 5];
#pragma unroll
        for (int m = 0; m < LOADS; ++m) {
            const int e = threadIdx.x + 256 * m;
            if (e < PLANE) sl[e] = stv[m];
        }
    };
    // planes t0-2 .. t0+1 (those outside the grid hold +0)
    for (int p = t0 - 2; p <= t0 + 1; ++p) {
        fetch(p);
        stash(p);
    }
    const int cl = (tx + 2) + HX * ((ty + 2) + HY * (tz + 2));
    const bool own_in = own && gx >= 1 && gy >= 1 && gz >= 1 && gx < D.X - 1 &&
                        gy < D.Y - 1 && gz < D.Z - 1;
    const int pxyz = (blur5_pattern(gz, D.Z) * 4 + blur5_pattern(gy, D.Y)) * 4 +
                     blur5_pattern(gx, D.X);
    for (int t = t0; t < t1; ++t) {
        fetch(t + 2);
        __syncthreads();  // every wave is done with plane t-3's slot
        stash(t + 2);
        __syncthreads();
        if (!own_in || t < 1 || t >= D.T - 1) continue;
        // the reference's order: t, z, y, x from low to high
        float val = 0.f;
#pragma unroll
        for (int vt = -2; vt <= 2; ++vt) {
            const float *sl = pl[(t + vt + 10) % 5] + cl;
#pragma unroll
            for (int zk = -2; zk <= 2; ++zk)
#pragma unroll
                for (int yj = -2; yj <= 2; ++yj)
#pragma unroll
                    for (int xi = -2; xi <= 2; ++xi)
                        val = val + sl[xi + HX * (yj + HY * zk)] *
                                        W.w[vt * vt + zk * zk + yj * yj + xi * xi];
        }
        const float weight = wsum[blur5_pattern(t, D.T) * 64 + pxyz];
        const int64_t c = gx + (int64_t)D.X * gy + D.sz * gz + D.st * t;
        dst[c * cstride] = weight > eps ? val / weight : pl[t % 5][cl];
    }
}

void gaussian_blur(flof_ctx *ctx, void *a, const Dims4 &D, bool vec4, float sigma,
                   int iter)
{
    int s = int(1. * sigma + 0.5);
    if (s == 0) s = 1;
    FLOF_REQUIRE(s <= 16, FLOF_ERR_RANGE, "gaussianBlur: sigma too large");
    // weights by squared distance d (the reference forms d as a float sum of
    // squared integer offsets: exact)
    const int nw = 4 * s * s + 1;
    std::vector<float> wh(nw);
    for (int d = 0; d < nw; ++d)
        wh[d] = (float)std::exp(-(double)(float)d / (2. * (double)sigma * (double)sigma));
    const int64_t n = (int64_t)D.st * D.T;
    const size_t bytes = (size_t)n * (vec4 ? 16 : 4);
    Tmp wdev(ctx, (size_t)nw * sizeof(float)), b(ctx, bytes);
    FLOF_HIP(hipMemcpyAsync(wdev.p, wh.data(), (size_t)nw * sizeof(float),
                            hipMemcpyHostToDevice, ctx->stream));
    launch_fill(ctx, b.p, bytes);
    Coord4 C = make_coord4(D);
    const int nb = grid_blocks_xcd(ctx, n);
    const float eps = 1e-6f;  // VECTOR_EPSILON of FP32 builds
    const BoxMarch BM = make_boxmarch(ctx, D);
    // s = 2: the weight of each squared distance and the weight sums of the
    // 4^4 clipping patterns (summed in the reference's order)
    Blur5W w5{};
    Tmp wsum5;
    BoxMarch BM5{};
    if (s == 2) {
        for (int d = 0; d < 17; ++d) w5.w[d] = wh[d];
        std::vector<float> ws(256);
        for (int pt = 0; pt < 4; ++pt)
            for (int pz = 0; pz < 4; ++pz)
                for (int py = 0; py < 4; ++py)
                    for (int px = 0; px < 4; ++px) {
                        float weight = 0.f;
                        const int pat[4] = {pt, pz, py, px};
                        int lo[4], hi[4];
                        for (int q = 0; q < 4; ++q) {
                            lo[q] = -(pat[q] / 2 + 1);
                            hi[q] = pat[q] % 2 + 1;
                        }
                        for (int vt = lo[0]; vt <= hi[0]; ++vt)
                            for (int zk = lo[1]; zk <= hi[1]; ++zk)
                                for (int yj = lo[2]; yj <= hi[2]; ++yj)
                                    for (int xi = lo[3]; xi <= hi[3]; ++xi)
                                        weight = weight +
                                                 wh[vt * vt + zk * zk + yj * yj + xi * xi];
                        ws[((pt * 4 + pz) * 4 + py) * 4 + px] = weight;
                    }
        wsum5 = Tmp(ctx, 256 * sizeof(float));
        FLOF_HIP(hipMemcpyAsync(wsum5.p, ws.data(), 256 * sizeof(float),
                                hipMemcpyHostToDevice, ctx->stream));
        FLOF_HIP(hipStreamSynchronize(ctx->stream));  // ws lives on the stack
        BM5 = make_boxmarch(ctx, D, vec4 ? 2 : 6);
    }
    auto pass = [&](const void *src, void *dst) {
        if (s == 2) {
            const int ncomp = vec4 ? 4 : 1;
            const dim3 grid(BM5.ntile * BM5.nseg * ncomp);
            if (BM5.tx == 32)
                hipLaunchKernelGGL(k_blur5_lds<32>, grid, dim3(256), 0, ctx->stream,
                                   (const float *)src, (float *)dst, ncomp, D, BM5, w5,
                                   (const float *)wsum5.as<float>(), eps);
            else
                hipLaunchKernelGGL(k_blur5_lds<64>, grid, dim3(256), 0, ctx->stream,
                                   (const float *)src, (float *)dst, ncomp, D, BM5, w5,
                                   (const float *)wsum5.as<float>(), eps);
            check_launch("k_blur5_lds");
            return;
        }
        if (s == 1) {
            // the LDS-tiled form of the same sums (identical results)
            if (vec4)
                launch_march3<float4, MARCH_BLUR1>(ctx, BM, (const float4 *)src, (float4 *)dst,
                                                   (const float *)wdev.as<float>(), D, eps);
            else
                launch_march3<float, MARCH_BLUR1>(ctx, BM, (const float *)src, (float *)dst,
                                                  (const float *)wdev.as<float>(), D, eps);
            check_launch("k_march3_lds<blur1>");
            return;
        }
        if (vec4)
            hipLaunchKernelGGL(k_blur4d<float4>, dim3(nb), dim3(256), 0, ctx->stream,
                               (const float4 *)src, (float4 *)dst, D, C, s,
                               (const float *)wdev.as<float>(), eps);
        else
            hipLaunchKernelGGL(k_blur4d<float>, dim3(nb), dim3(256), 0, ctx->stream,
                               (const float *)src, (float *)dst, D, C, s,
                               (const float *)wdev.as<float>(), eps);
        check_launch("k_blur4d");
    };
    for (int it = 0; it < iter; ++it) {
        // pass 1 into the temporary (its outer layer stays zero), pass 2 back
        // into the grid, which keeps its own outer layer: the reference's two
        // swaps with a zeroed temporary
        pass(a, b.p);
        pass(b.p, a);
    }
    // the host weights must outlive the async copy
    FLOF_HIP(hipStreamSynchronize(ctx->stream));
}

// ---------------------------------------------------------------------------
// level-set projection (optflow4d.cpp:651-734)

__device__ __forceinline__ bool in_bounds_pos(const Dims4 &D, float x, float y, float z,
                                              float t)
{
    const int i = (int)x, j = (int)y, k = (int)z, l = (int)t;
    return i >= 0 && j >= 0 && k >= 0 && l >= 0 && i < D.X && j < D.Y && k < D.Z &&
           l < D.T;
}

__global__ void k_project_cells(float4 *__restrict__ dst, const float4 *__restrict__ vel,
                                const float *__restrict__ phiOrg,
                                const float *__restrict__ phiTarget,
                                float *__restrict__ marker, Dims4 D, Coord4 C,
                                float threshPhi, int maxIter, float dt)
{
    const int64_t n = (int64_t)D.st * D.T;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n;
         c += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k, t;
        decompose((uint32_t)c, C, i, j, k, t);
        if (i < 2 || j < 2 || k < 2 || t < 2 || i >= D.X - 2 || j >= D.Y - 2 ||
            k >= D.Z - 2 || t >= D.T - 2)
            continue;
        const float4 v = f4muls(vel[c], dt);
        const float px = ((float)i + 0.5f) - v.x, py = ((float)j + 0.5f) - v.y,
                    pz = ((float)k + 0.5f) - v.z, pt = ((float)t + 0.5f) - v.w;
        // line search along the normal (projectCell, optflow4d.cpp:675-710);
        // a cell whose final |d| exceeds threshPhi gets dst = 0 whatever its
        // normal and path, and every remaining step moves d by at most the
        // current step (which never grows), so a cell stops as soon as no
        // remaining steps can bring |d| back within threshPhi (with a margin
        // far above the FP32 rounding of the steps); the result is the same
        float step = 0.25f;
        if (maxIter < 50) step = 0.5f;
        const float target = phiTarget[c];
        float vn = interp4<float>(phiOrg, D, px, py, pz, pt);
        float d = vn - target;
        auto out_of_reach = [&](int rem) {
            const float ad = fabsf(d);
            return ad > threshPhi + (float)max(rem, 0) * step * 1.001f + 1e-3f * (1.f + ad);
        };
        if (out_of_reach(maxIter)) {
            dst[c] = make_float4(0.f, 0.f, 0.f, 0.f);
            continue;
        }
        const float h = 0.5f;
        float4 nrm;
        nrm.x = interp4<float>(phiOrg, D, px + h, py, pz, pt) -
                interp4<float>(phiOrg, D, px - h, py, pz, pt);
        nrm.y = interp4<float>(phiOrg, D, px, py + h, pz, pt) -
                interp4<float>(phiOrg, D, px, py - h, pz, pt);
        nrm.z = interp4<float>(phiOrg, D, px, py, pz + h, pt) -
                interp4<float>(phiOrg, D, px, py, pz - h, pt);
        nrm.w = interp4<float>(phiOrg, D, px, py, pz, pt + h) -
                interp4<float>(phiOrg, D, px, py, pz, pt - h);
        {
            const float l = nrm.x * nrm.x + nrm.y * nrm.y + nrm.z * nrm.z + nrm.w * nrm.w;
            const float e2 = 1e-6f * 1e-6f;
            if (fabs((double)l - 1.) < (double)e2) {
            } else if (l > e2) {
                const float nr = sqrtf(l);
                nrm = f4muls(nrm, (float)(1. / (double)nr));
            } else {
                nrm = make_float4(0.f, 0.f, 0.f, 0.f);
            }
        }
        int lastDir = 0;
        for (int s = 0; s < maxIter; ++s) {
            const float4 o = f4muls(f4muls(nrm, d), dt);
            const float qx = px + o.x, qy = py + o.y, qz = pz + o.z, qt = pt + o.w;
            if (!in_bounds_pos(D, qx, qy, qz, qt)) {
                s = maxIter;
                d = 9999.f;
                continue;
            }
            vn = interp4<float>(phiOrg, D, qx, qy, qz, qt);
            if (vn < target) {
                if (lastDir < 0) step = (float)(step * 0.5);
                d += step;
                lastDir = 1;
            } else {
                if (lastDir > 0) step = (float)(step * 0.5);
                d -= step;
                lastDir = -1;
            }
            if (step < 1e-01) {
                s = maxIter + 1;
            } else if (out_of_reach(maxIter - 1 - s)) {
                d = 9999.f;
                s = maxIter + 1;
            }
        }
        if (fabsf(d) > threshPhi) {
            dst[c] = make_float4(0.f, 0.f, 0.f, 0.f);
            continue;
        }
        dst[c] = f4muls(nrm, d);
        marker[c] = 1.f;
    }
}

void corr_vels_of(flof_ctx *ctx, float4 *dst, float4 *vel, const float *phiOrg,
                  const float *phiTarget, const Dims4 &D, float threshPhi,
                  float postVelBlur, float resetBndWidth, int maxIter, float dt)
{
    const int64_t n = (int64_t)D.st * D.T;
    const float blurThreshold = 0.98f;
    const int doCvExtrapol = (int)(host_max_norm_vec4(ctx, vel, n) + 4);
    Tmp marker(ctx, (size_t)n * 4), tmp(ctx, (size_t)n * 16);
    Coord4 C = make_coord4(D);
    const int nb = grid_blocks(ctx, n);
    int singleStep = 0;
    if (postVelBlur < blurThreshold) singleStep = 1;
    while ((postVelBlur >= blurThreshold) || singleStep) {
        launch_fill(ctx, dst, (size_t)n * 16);
        launch_fill(ctx, marker.p, (size_t)n * 4);
        hipLaunchKernelGGL(k_project_cells, dim3(nb), dim3(256), 0, ctx->stream, dst, vel,
                           phiOrg, phiTarget, marker.as<float>(), D, C, threshPhi, maxIter,
                           dt);
        check_launch("k_project_cells");
        // the passes alternate between dst and tmp (every pass writes every
        // cell); one copy at the end if the last pass wrote tmp
        float4 *cur = dst, *nxt = tmp.as<float4>();
        const BoxMarch BM = make_boxmarch(ctx, D);
        for (int it = 0; it < doCvExtrapol; ++it) {
            launch_march3<float4, MARCH_BOX81>(ctx, BM, (const float4 *)cur, nxt,
                                               (const float *)marker.as<float>(), D, 0.f);
            check_launch("k_march3_lds<box81>");
            std::swap(cur, nxt);
        }
        if (cur != dst)
            FLOF_HIP(hipMemcpyAsync(dst, cur, (size_t)n * 16, hipMemcpyDeviceToDevice,
                                    ctx->stream));
        if (postVelBlur > 1e-6f)
            gaussian_blur(ctx, dst, D, true, (float)(0.5 * postVelBlur), 1);
        const int resetBnd = int(resetBndWidth * (float)D.X) + 1;
        if (resetBndWidth > 0.f)
            launch_set_bound(ctx, dst, D, true, make_float4(0.f, 0.f, 0.f, 0.f), resetBnd);
        launch_axpy4(ctx, vel, dst, make_float4(-1.f, -1.f, -1.f, -1.f), n);
        postVelBlur = (float)(postVelBlur * 0.5);
        if (singleStep > 0) singleStep--;
    }
}

// ---------------------------------------------------------------------------
// one level (opticalFlowDim)

struct SolveOut {
    int iters = 0, status = 0, warm = 0;
    double resRel = 0.;
    double resL2 = 0.;  // ||r||_2 / ||b||_2 at the end
};

// t-march decomposition: enough workgroups to fill the chip, t split into
// segments when one t-slab has too few cells
static TMarch make_tmarch(flof_ctx *ctx, const Dims4 &D, int blocks_per_cu)
{
    TMarch M;
    M.n3 = D.st;
    M.nblk3 = (int)((M.n3 + 255) / 256);
    const int target = ctx->num_cu * blocks_per_cu;
    int nseg = (target + M.nblk3 - 1) / M.nblk3;
    if (nseg < 1) nseg = 1;
    if (nseg > D.T) nseg = D.T;
    M.seglen = (D.T + nseg - 1) / nseg;
    M.nseg = (D.T + M.seglen - 1) / M.seglen;
    M.total = M.nblk3 * M.nseg;
    return M;
}

static TileMarch make_tilemarch(const flof_ctx *ctx, const Dims4 &D, int blocks_per_cu)
{
    TileMarch M;
    M.ntx = (D.X + LT_X - 1) / LT_X;
    M.nty = (D.Y + LT_Y - 1) / LT_Y;
    M.ntz = (D.Z + LT_Z - 1) / LT_Z;
    M.ntile = M.ntx * M.nty * M.ntz;
    const int target = ctx->num_cu * blocks_per_cu;
    int nseg = (target + M.ntile - 1) / M.ntile;
    if (nseg < 1) nseg = 1;
    if (nseg > D.T) nseg = D.T;
    M.seglen = (D.T + nseg - 1) / nseg;
    M.nseg = (D.T + M.seglen - 1) / M.seglen;
    M.total = M.ntile * M.nseg;
    M.xcd = ctx->lds_xcd ? 1 : 0;
    return M;
}

int lds_seglen(const flof_ctx *ctx, const Dims4 &D, int lds_bpc)
{
    return make_tilemarch(ctx, D, lds_bpc).seglen;
}

// candidates of the autotuner: workgroups per CU of the t-marching kernels
// and work items per CU of the LDS-tiled stencil kernels
static const int TM_CANDIDATES[] = {2, 4, 8, 16};
static const int LDS_CANDIDATES[] = {1, 2, 4, 8};

struct LevelWork {
    Tmp g, prec, x, r, p, q, st, part;
    int nb = 0;     // grid of the streaming kernels
    int nb_tm = 0;  // grid of the t-marching kernels
    int nb_lds = 0; // grid of the LDS-tiled stencil kernels
    int cap = 0;    // partial slots
    TMarch tm;
    TileMarch tl;
    // lds_bpc: work items per CU of the LDS-tiled kernels (<= 0: the
    // context's default)
    void set_tm(flof_ctx *ctx, const Dims4 &D, int bpc, int lds_bpc = 0)
    {
        tm = make_tmarch(ctx, D, bpc);
        nb_tm = ((tm.total + FLOF_NXCD - 1) / FLOF_NXCD) * FLOF_NXCD;
        tl = make_tilemarch(ctx, D, lds_bpc > 0 ? lds_bpc : ctx->lds_blocks_per_cu);
        nb_lds = tl.xcd ? ((tl.total + FLOF_NXCD - 1) / FLOF_NXCD) * FLOF_NXCD : tl.total;
        // the LDS tiles pay off on long t segments of full-width tiles with
        // FP32 direction storage (measured: 3% faster at 64^4, 35% slower
        // at 32^4 with 4-step segments, 26% slower with bfloat16 storage);
        // they index cells (and the t+2 prefetch) in 32 bits
        const bool fits = D.st * (int64_t)(D.T + 2) < (1ll << 31);
        const bool pays = ctx->p_store == FLOF_PSTORE_F32 && tl.seglen >= 16 &&
                          10 * D.X >= 9 * LT_X * tl.ntx;
        lds_allowed = fits && !ctx->img_bf16 && ctx->stencil != FLOF_STENCIL_TMARCH;
        lds = lds_allowed &&
              (ctx->stencil == FLOF_STENCIL_LDS || (ctx->stencil == FLOF_STENCIL_AUTO && pays));
    }
    bool lds = false;          // p.Ap and the update run as the LDS-tiled kernels
    bool lds_allowed = false;  // the LDS-tiled kernels may run this level
    int nb_stencil() const { return lds ? nb_lds : nb_tm; }
    LevelWork(flof_ctx *ctx, const Dims4 &D)
    {
        const int64_t n = (int64_t)D.st * D.T;
        const size_t b = (size_t)n * 16;
        g = Tmp(ctx, b);
        prec = Tmp(ctx, b);
        x = Tmp(ctx, b);
        r = Tmp(ctx, b);
        p = Tmp(ctx, b);
        q = Tmp(ctx, b);
        nb = grid_blocks_xcd(ctx, n);
        // partial slots for the largest grids the autotuner may pick
        set_tm(ctx, D, TM_CANDIDATES[3], LDS_CANDIDATES[3]);
        cap = std::max(nb, std::max(nb_tm, nb_lds));
        set_tm(ctx, D, ctx->tm_blocks_per_cu);
        cap = std::max(cap, std::max(nb_tm, nb_lds));
        st = Tmp(ctx, sizeof(CgState));
        part = Tmp(ctx, (size_t)cap * (3 * sizeof(DD) + sizeof(float)) + 256);
        FLOF_HIP(hipMemsetAsync(st.p, 0, sizeof(CgState), ctx->stream));
    }
    DD *pdd() const { return part.as<DD>(); }
    DD *pdd2() const { return part.as<DD>() + cap; }
    DD *pdd3() const { return part.as<DD>() + 2 * cap; }
    float *pmax() const { return (float *)(part.as<DD>() + 3 * cap); }
};

// The two buffers of the search direction, in its storage: p_k lives in
// buffer k % 2 (FP32: W.p and W.q; bfloat16: the two halves of W.q; W.q is
// free during the fused iterations).
static void direction_buffers(LevelWork &W, int pstore, int64_t n, void *pb[2])
{
    if (pstore == FLOF_PSTORE_BF16) {
        pb[0] = W.q.p;
        pb[1] = (char *)W.q.p + (size_t)n * 8;
    } else {
        pb[0] = W.p.p;
        pb[1] = W.q.p;
    }
}

// one PCG iteration k (the three stencil kernels and the two finalizers):
// p_k in pcur, p_(k+1) into pnext, odd = k % 2; ka/ku/kd: timer kinds or -1
template <class PS, class IM>
static void enqueue_tm_iter(flof_ctx *ctx, LevelWork &W, void *pcur, void *pnext, bool odd,
                            const typename IM::T *i1, const LevelGeom &G, float dxf,
                            CgState *st, int ka, int ku, int kd, bool finalizers)
{
    typedef typename PS::T PT;
    const int ns = W.nb_stencil();
    // p.q partials go to pdd2: the update reads them while its own
    // workgroups write their partials to pdd (the LDS tiles stage FP32
    // images only: set_tm never selects them for bfloat16 images)
    bool lds_done = false;
    if constexpr (std::is_same<IM, ImgF32>::value) {
        if (W.lds) {
            klaunch(ctx, ka, k_cg_stencil_lds<PS, false>, dim3(ns), dim3(256), (const PT *)pcur,
                    i1, (float4 *)nullptr, G, W.tl, dxf, st, (const DD *)nullptr, W.pdd2(),
                    (float *)nullptr);
            klaunch(ctx, ku, k_cg_stencil_lds<PS, true>, dim3(ns), dim3(256), (const PT *)pcur,
                    i1, W.r.as<float4>(), G, W.tl, dxf, st, (const DD *)W.pdd2(), W.pdd(),
                    W.pmax());
            lds_done = true;
        }
    }
    if (!lds_done) {
        klaunch(ctx, ka, k_cg_pq_tm<PS, IM>, dim3(ns), dim3(256), (const PT *)pcur, i1, G, W.tm,
                dxf, (const CgState *)st, W.pdd2());
        klaunch(ctx, ku, k_cg_update_tm<PS, IM>, dim3(ns), dim3(256), W.r.as<float4>(),
                (const PT *)pcur, i1, G, W.tm, dxf, st, (const DD *)W.pdd2(), W.pdd(),
                W.pmax());
    }
    if (finalizers)
        hipLaunchKernelGGL(k_cg_beta, dim3(1), dim3(256), 0, ctx->stream, st, W.pdd(),
                           W.pmax(), ns);
    if (odd)
        klaunch(ctx, kd, k_cg_dir_tm<PS, true, IM>, dim3(W.nb_tm), dim3(256), (const PT *)pcur,
                (PT *)pnext, W.x.as<float4>(), (const float4 *)W.r.as<float4>(), i1, G, W.tm,
                dxf, (const CgState *)st);
    else
        klaunch(ctx, kd, k_cg_dir_tm<PS, false, IM>, dim3(W.nb_tm), dim3(256), (const PT *)pcur,
                (PT *)pnext, W.x.as<float4>(), (const float4 *)W.r.as<float4>(), i1, G, W.tm,
                dxf, (const CgState *)st);
}

static void enqueue_tm_iter_any(flof_ctx *ctx, int pstore, LevelWork &W, void *pcur,
                                void *pnext, bool odd, const void *i1, const LevelGeom &G,
                                float dxf, CgState *st, int ka, int ku, int kd, bool finalizers)
{
    if (ctx->img_bf16) {
        const bf16_t *ib = (const bf16_t *)i1;
        if (pstore == FLOF_PSTORE_BF16)
            enqueue_tm_iter<PStoreBF16, ImgBF16>(ctx, W, pcur, pnext, odd, ib, G, dxf, st, ka,
                                                 ku, kd, finalizers);
        else
            enqueue_tm_iter<PStoreF32, ImgBF16>(ctx, W, pcur, pnext, odd, ib, G, dxf, st, ka,
                                                ku, kd, finalizers);
        return;
    }
    const float *ifl = (const float *)i1;
    if (pstore == FLOF_PSTORE_BF16)
        enqueue_tm_iter<PStoreBF16, ImgF32>(ctx, W, pcur, pnext, odd, ifl, G, dxf, st, ka, ku,
                                            kd, finalizers);
    else
        enqueue_tm_iter<PStoreF32, ImgF32>(ctx, W, pcur, pnext, odd, ifl, G, dxf, st, ka, ku,
                                           kd, finalizers);
}

// Autotuner of the launch shape: on the first solve of a grid shape, times
// one PCG iteration (the stencil and direction kernels, twice) for each
// candidate on the level's own (not yet initialised) workspace, with a
// scratch solver state that lets every kernel run, and keeps the fastest
// per shape for the context's lifetime.  The candidates are the t-marching
// kernels' workgroups per CU and then, where the LDS-tiled stencil kernels
// may run, their work items per CU (which fix the t-segment length of the
// 64x4x4-cell tiles; counts that give the same segments are timed once)
// with the best t-marching direction kernel.  Every candidate computes the
// same values (correctly rounded sums, order-free maxima), and the
// workspace is written by the set-up kernel afterwards, so the result of
// the solve does not depend on the tuning.
static flof_ctx::TunedShape tuned_shape(flof_ctx *ctx, const Dims4 &D, LevelWork &W,
                                        const void *i1, const LevelGeom &G, float dxf,
                                        int pstore, void *const pb[2], bool fused = false)
{
    const std::array<int, 5> key{D.X, D.Y, D.Z, D.T, tune_variant(ctx) + (fused ? 64 : 0)};
    auto it = ctx->tm_tuned.find(key);
    if (it != ctx->tm_tuned.end()) return it->second;
    CgState hs;
    std::memset(&hs, 0, sizeof(hs));
    hs.alpha = 0.;
    hs.beta = 0.5;
    hs.dir_mode = 1;
    Tmp stt(ctx, sizeof(CgState));
    FLOF_HIP(hipMemcpyAsync(stt.p, &hs, sizeof(hs), hipMemcpyHostToDevice, ctx->stream));
    FLOF_HIP(hipStreamSynchronize(ctx->stream));  // hs lives on the stack
    hipEvent_t e0, e1;
    FLOF_HIP(hipEventCreate(&e0));
    FLOF_HIP(hipEventCreate(&e1));
    CgState *st = stt.as<CgState>();
    hipError_t err = hipSuccess;
    // milliseconds of two iterations after a warm-up one (< 0: error)
    auto time_it = [&]() -> float {
        enqueue_tm_iter_any(ctx, pstore, W, pb[0], pb[1], false, i1, G, dxf, st, -1, -1, -1,
                            false);
        err = hipEventRecord(e0, ctx->stream);
        for (int k = 0; k < 2; ++k)
            enqueue_tm_iter_any(ctx, pstore, W, pb[0], pb[1], false, i1, G, dxf, st, -1, -1,
                                -1, false);
        if (err == hipSuccess) err = hipEventRecord(e1, ctx->stream);
        if (err == hipSuccess) err = hipEventSynchronize(e1);
        float ms = 0.f;
        if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
        return err == hipSuccess ? ms : -1.f;
    };
    const bool lds_ok = !fused && W.lds_allowed;
    // with the LDS-tiled kernels forced, the first sweep runs them too
    const bool forced = lds_ok && ctx->stencil == FLOF_STENCIL_LDS;
    flof_ctx::TunedShape best{ctx->tm_blocks_per_cu, forced ? ctx->lds_blocks_per_cu : 0};
    float best_ms = 1e30f;
    for (int bpc : TM_CANDIDATES) {
        W.set_tm(ctx, D, bpc);
        W.lds = forced;
        if (W.nb_tm > W.cap || W.nb_lds > W.cap) continue;
        const float ms = time_it();
        if (ms < 0.f) break;
        if (ms < best_ms) {
            best_ms = ms;
            best = {bpc, forced ? ctx->lds_blocks_per_cu : 0};
        }
    }
    if (lds_ok && err == hipSuccess) {
        const int bpc = best.bpc;
        int last_nseg = -1;
        for (int lb : LDS_CANDIDATES) {
            W.set_tm(ctx, D, bpc, lb);
            W.lds = true;
            if (W.nb_tm > W.cap || W.nb_lds > W.cap || W.tl.nseg == last_nseg) continue;
            last_nseg = W.tl.nseg;
            const float ms = time_it();
            if (ms < 0.f) break;
            if (ms < best_ms) {
                best_ms = ms;
                best = {bpc, lb};
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    FLOF_HIP(err);
    check_launch("autotune");
    ctx->tm_tuned[key] = best;
    return best;
}

// One level's PCG: workspace, geometry, direction buffers and the solver
// state after the set-up (read back to the host).
struct LevelCg {
    LevelWork W;
    LevelGeom G;
    Dims4 D;
    int64_t n = 0;
    float mDx = 0.f, dxf = 0.f;
    int pstore = FLOF_PSTORE_F32;
    void *pb[2] = {nullptr, nullptr};
    const void *i1 = nullptr;
    CgState hs;
    LevelCg(flof_ctx *ctx, const Dims4 &D_) : W(ctx, D_), D(D_)
    {
        n = (int64_t)D_.st * D_.T;
    }
    CgState *st() { return W.st.as<CgState>(); }
};

// The set-up of one level (optflow4d.cpp:398-493, the start of the CG
// :284-303, the optional warm start).  vin: incoming flow for the rhs (may be
// null = zero).  fused: the level joins a batch whose PCG launches span
// several problems (t-marching kernels only).
static void level_setup(flof_ctx *ctx, LevelCg &L, const void *i0, const void *i1,
                        const float4 *vin, float *rhsT, float wSmooth, float wEnergy,
                        float cgAccuracy, float dt, int maxIter, const float4 *warm,
                        const float4 *warm_base, bool fused)
{
    const Dims4 &D = L.D;
    FLOF_REQUIRE(D.X >= 1 && D.Y >= 1 && D.Z >= 1 && D.T >= 1, FLOF_ERR_DIM,
                 "opticalFlow4d: bad size");
    const int64_t n = L.n;
    FLOF_REQUIRE(n < (1ll << 31), FLOF_ERR_DIM, "opticalFlow4d: grid too large");
    FLOF_REQUIRE(!fused || ctx->use_tmarch, FLOF_ERR_ARG,
                 "fused batch: needs the t-marching kernels");
    LevelWork &W = L.W;
    L.i1 = i1;
    CgState *st = W.st.as<CgState>();
    const float mDx = (float)(1. / D.X);
    const float dxf = (float)(1. / (2. * (double)mDx));
    L.mDx = mDx;
    L.dxf = dxf;
    LevelGeom &G = L.G;
    G.D = D;
    G.C = make_coord4(D);
    G.offd = -wSmooth * 1.f;
    float diag = 0.f;
    diag += (float)(8) * wSmooth * 1.f;
    diag += wEnergy;
    G.diag = diag;
    G.use_offd = G.offd != 0.f;
    G.eps = 1e-6f;
    G.fast_rcp = 0;  // set from the set-up's range check below
    {
        int maxIt = maxIter;
        FLOF_HIP(hipMemcpyAsync(&st->maxIter, &maxIt, sizeof(int), hipMemcpyHostToDevice,
                                ctx->stream));
        FLOF_HIP(hipStreamSynchronize(ctx->stream));  // maxIt lives on the stack
    }
    // search direction storage of the fused kernels: two buffers (W.p and
    // W.q, or two bfloat16 halves of W.q; the fused kernels need no q array,
    // and the warm start uses W.q before the iterations)
    const int pstore = ctx->use_tmarch ? ctx->p_store : FLOF_PSTORE_F32;
    L.pstore = pstore;
    void **pb = L.pb;
    direction_buffers(W, pstore, n, pb);
    if (ctx->use_tmarch) {
        if (ctx->autotune && n >= (1 << 16)) {
            const flof_ctx::TunedShape ts = tuned_shape(ctx, D, W, i1, G, dxf, pstore, pb, fused);
            W.set_tm(ctx, D, ts.bpc, ts.lds_bpc);
            W.lds = ts.lds_bpc > 0 && W.lds_allowed;
        } else {
            W.set_tm(ctx, D, ctx->tm_blocks_per_cu);
        }
        if (fused) W.lds = false;
    }
    // with fewer than 3 cells in a dimension every row is an identity row and
    // the neighbour offsets are never used
    const int nb = W.nb;
    if (ctx->img_bf16)
        hipLaunchKernelGGL(k_of_setup<ImgBF16>, dim3(nb), dim3(256), 0, ctx->stream,
                           (const bf16_t *)i0, (const bf16_t *)i1, vin, rhsT, W.g.as<float4>(),
                           W.prec.as<float4>(), W.x.as<float4>(), W.r.as<float4>(),
                           W.p.as<float4>(), G, wSmooth, wEnergy, dt, mDx, dxf, W.pdd(),
                           W.pmax(), (float *)W.pdd2());
    else
        hipLaunchKernelGGL(k_of_setup<ImgF32>, dim3(nb), dim3(256), 0, ctx->stream,
                           (const float *)i0, (const float *)i1, vin, rhsT, W.g.as<float4>(),
                           W.prec.as<float4>(), W.x.as<float4>(), W.r.as<float4>(),
                           W.p.as<float4>(), G, wSmooth, wEnergy, dt, mDx, dxf, W.pdd(),
                           W.pmax(), (float *)W.pdd2());
    check_launch("k_of_setup");
    hipLaunchKernelGGL(k_setup_fin, dim3(1), dim3(256), 0, ctx->stream, st, W.pdd(),
                       W.pmax(), (const float *)W.pdd2(), nb, cgAccuracy, G.eps);
    check_launch("k_setup_fin");
    // ||b||^2 while r still holds the rhs (x = 0)
    hipLaunchKernelGGL(k_sumsq, dim3(nb), dim3(256), 0, ctx->stream,
                       (const float4 *)W.r.as<float4>(), n, W.pdd3());
    hipLaunchKernelGGL(k_sumsq_fin, dim3(1), dim3(256), 0, ctx->stream, (const DD *)W.pdd3(),
                       nb, &st->bb);
    check_launch("k_sumsq");
    if (warm) {
        hipLaunchKernelGGL(k_warm_guess, dim3(nb), dim3(256), 0, ctx->stream, warm,
                           warm_base, W.q.as<float4>(), D, G.C, mDx);
        check_launch("k_warm_guess");
        Tmp rw(ctx, (size_t)n * 16);
        hipLaunchKernelGGL(k_warm_aw, dim3(nb), dim3(256), 0, ctx->stream,
                           (const float4 *)W.q.as<float4>(), (const float4 *)W.g.as<float4>(),
                           (const float4 *)W.r.as<float4>(), rw.as<float4>(), G, W.pdd(),
                           W.pdd2(), W.pdd3());
        check_launch("k_warm_aw");
        hipLaunchKernelGGL(k_warm_gamma, dim3(1), dim3(256), 0, ctx->stream, st,
                           (const DD *)W.pdd(), (const DD *)W.pdd2(), (const DD *)W.pdd3(),
                           nb);
        check_launch("k_warm_gamma");
        hipLaunchKernelGGL(k_warm_scale, dim3(nb), dim3(256), 0, ctx->stream,
                           W.q.as<float4>(), rw.as<float4>(), (const float4 *)W.r.as<float4>(),
                           n, (const CgState *)st, W.pdd());
        check_launch("k_warm_scale");
        hipLaunchKernelGGL(k_warm_fin, dim3(1), dim3(256), 0, ctx->stream, st,
                           (const DD *)W.pdd(), nb);
        check_launch("k_warm_fin");
        hipLaunchKernelGGL(k_warm_adopt, dim3(nb), dim3(256), 0, ctx->stream,
                           W.q.as<float4>(), rw.as<float4>(), W.prec.as<float4>(),
                           W.x.as<float4>(), W.r.as<float4>(), W.p.as<float4>(), n, st,
                           W.pdd(), W.pmax());
        check_launch("k_warm_adopt");
        hipLaunchKernelGGL(k_warm_adopt_fin, dim3(1), dim3(256), 0, ctx->stream, st,
                           W.pdd(), W.pmax(), nb);
        check_launch("k_warm_adopt_fin");
    }
    if (pstore == FLOF_PSTORE_BF16) {
        hipLaunchKernelGGL(k_p_pack<PStoreBF16>, dim3(grid_blocks(ctx, n)), dim3(256), 0,
                           ctx->stream, (const float4 *)W.p.as<float4>(),
                           (PStoreBF16::T *)pb[0], n);
        check_launch("k_p_pack");
    }
    // the state after the set-up (and warm start), which also says whether
    // the fast reciprocal is exact on this level
    FLOF_HIP(hipMemcpyAsync(&L.hs, st, sizeof(CgState), hipMemcpyDeviceToHost, ctx->stream));
    FLOF_HIP(hipStreamSynchronize(ctx->stream));
    G.fast_rcp = L.hs.fast_rcp;
}

// vel (=) or (+=) x / h of a finished level, and ||r||_2 / ||b||_2
static SolveOut level_finish(flof_ctx *ctx, LevelCg &L, float4 *out, bool accumulate)
{
    LevelWork &W = L.W;
    CgState *st = W.st.as<CgState>();
    const CgState &hs = L.hs;
    hipLaunchKernelGGL(k_of_finish, dim3(W.nb), dim3(256), 0, ctx->stream, W.x.as<float4>(),
                       out, L.n, L.mDx, accumulate ? 1 : 0, st);
    check_launch("k_of_finish");
    SolveOut so;
    so.iters = hs.iters;
    so.status = hs.status;
    so.warm = hs.warm;
    so.resRel = hs.resRel;
    // ||r||_2 / ||b||_2 of the final (recursive) residual
    hipLaunchKernelGGL(k_sumsq, dim3(W.nb), dim3(256), 0, ctx->stream,
                       (const float4 *)W.r.as<float4>(), L.n, W.pdd3());
    hipLaunchKernelGGL(k_sumsq_fin, dim3(1), dim3(256), 0, ctx->stream, (const DD *)W.pdd3(),
                       W.nb, &st->rr);
    check_launch("k_sumsq");
    CgState *slots = (CgState *)ctx->host_scratch;
    FLOF_HIP(hipMemcpyAsync(&slots[0], st, sizeof(CgState), hipMemcpyDeviceToHost,
                            ctx->stream));
    FLOF_HIP(hipStreamSynchronize(ctx->stream));
    so.resL2 = slots[0].bb > 0. ? std::sqrt(slots[0].rr / slots[0].bb) : 0.;
    return so;
}

// Solves one level.  vin: incoming flow for the rhs (may be null = zero).
// The result x/h is written to (accumulate ? += : =) out.
// The PCG iterations of one level set up by level_setup (fused = false),
// until it converges or reaches maxIter; the final state goes to L.hs.
static void level_iterate(flof_ctx *ctx, LevelCg &L, int maxIter)
{
    LevelWork &W = L.W;
    const void *i1 = L.i1;
    CgState *st = W.st.as<CgState>();
    const int64_t n = L.n;
    const float dxf = L.dxf;
    const int pstore = L.pstore;
    void **pb = L.pb;
    const LevelGeom G = L.G;  // captured by value in the graph
    const int nb = W.nb;
    CgState hs = L.hs;
    // Iterations are enqueued in chunks of CHUNK (as one captured hipGraph
    // unless per-kernel timing is on); after each chunk the state is copied to
    // pinned host memory behind an event, and the host waits only for the
    // chunk before the last one, so the GPU never idles on the check.
    // Kernels of iterations after convergence (or the cap, counted on the
    // device) return at once.  CHUNK is even, so iteration k of a chunk has
    // the parity of the solve's iteration and reads p from buffer k % 2.
    const int CHUNK = 16;
    static_assert(CHUNK % 2 == 0, "direction buffers alternate per iteration");
    ctx->timer_tag = n;  // timed launches of this level carry its cell count
    bool timed = false;
    for (int k : {FLOF_KT_CG_APPLY, FLOF_KT_CG_UPDATE, FLOF_KT_CG_DIR})
        timed = timed || ctx->timers[k].on;
    auto enqueue_iters = [&](int count, bool with_timers) {
        const int ka = with_timers ? FLOF_KT_CG_APPLY : -1;
        const int ku = with_timers ? FLOF_KT_CG_UPDATE : -1;
        const int kd = with_timers ? FLOF_KT_CG_DIR : -1;
        for (int k = 0; k < count; ++k) {
            if (ctx->use_tmarch) {
                enqueue_tm_iter_any(ctx, pstore, W, pb[k & 1], pb[(k + 1) & 1], (k & 1) != 0, i1,
                                    G, dxf, st, ka, ku, kd, true);
                continue;
            }
            klaunch(ctx, ka, k_cg_apply, dim3(nb), dim3(256),
                    (const float4 *)W.p.as<float4>(), (const float4 *)W.g.as<float4>(),
                    W.q.as<float4>(), G, (const CgState *)st, W.pdd());
            hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(256), 0, ctx->stream, st, W.pdd(),
                               nb);
            klaunch(ctx, ku, k_cg_update, dim3(nb), dim3(256), W.r.as<float4>(),
                    (const float4 *)W.q.as<float4>(), (const float4 *)W.prec.as<float4>(), n,
                    (const CgState *)st, W.pdd(), W.pmax());
            hipLaunchKernelGGL(k_cg_beta, dim3(1), dim3(256), 0, ctx->stream, st, W.pdd(),
                               W.pmax(), nb);
            klaunch(ctx, kd, k_cg_dir, dim3(nb), dim3(256), W.x.as<float4>(), W.p.as<float4>(),
                    (const float4 *)W.r.as<float4>(), (const float4 *)W.prec.as<float4>(), n,
                    (const CgState *)st);
        }
    };
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    if (!timed && ctx->use_graphs) {
        FLOF_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        enqueue_iters(CHUNK, false);
        hipError_t ce = hipStreamEndCapture(ctx->stream, &graph);
        if (ce == hipSuccess && graph)
            ce = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (ce != hipSuccess) {
            (void)hipGetLastError();
            if (graph) (void)hipGraphDestroy(graph);
            graph = nullptr;
            exec = nullptr;
        }
    }
    CgState *slots = (CgState *)ctx->host_scratch;  // pinned, 2 slots
    hipEvent_t ev[2];
    FLOF_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    FLOF_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    int64_t chunks = 0;
    const int64_t max_chunks = (int64_t)maxIter / CHUNK + 3;
    hipError_t loop_err = hipSuccess;
    while (!hs.done && chunks < max_chunks) {
        if (exec) {
            loop_err = hipGraphLaunch(exec, ctx->stream);
            if (loop_err != hipSuccess) break;
        } else {
            enqueue_iters(CHUNK, timed);
        }
        loop_err = hipMemcpyAsync(&slots[chunks & 1], st, sizeof(CgState),
                                  hipMemcpyDeviceToHost, ctx->stream);
        if (loop_err == hipSuccess) loop_err = hipEventRecord(ev[chunks & 1], ctx->stream);
        if (loop_err != hipSuccess) break;
        ++chunks;
        if (chunks >= 2) {
            // state after the previous chunk
            loop_err = hipEventSynchronize(ev[chunks & 1]);
            if (loop_err != hipSuccess) break;
            hs = slots[chunks & 1];
        }
    }
    if (loop_err == hipSuccess) loop_err = hipStreamSynchronize(ctx->stream);
    if (loop_err == hipSuccess && chunks >= 1) hs = slots[(chunks - 1) & 1];
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    FLOF_HIP(loop_err);
    check_launch("k_cg_*");
    L.hs = hs;
}

SolveOut solve_level(flof_ctx *ctx, const void *i0, const void *i1, const float4 *vin,
                     float *rhsT, const Dims4 &D, float wSmooth, float wEnergy,
                     float cgAccuracy, float dt, int maxIter, const float4 *warm,
                     const float4 *warm_base, float4 *out, bool accumulate)
{
    LevelCg L(ctx, D);
    level_setup(ctx, L, i0, i1, vin, rhsT, wSmooth, wEnergy, cgAccuracy, dt, maxIter, warm,
                warm_base, false);
    level_iterate(ctx, L, maxIter);
    return level_finish(ctx, L, out, accumulate);
}

// ---------------------------------------------------------------------------
// Fused batch (no reference counterpart; the reference solves one problem at
// a time): the PCG iterations of several problems' levels share their
// launches, one per kernel kind and iteration.  Workgroup b of a launch
// belongs to the problem whose range [off, off + nb) holds it and runs that
// problem's kernel body as workgroup b - off of nb -- the same work items,
// partial sums and order as a launch of the problem alone, so each result
// is bit-identical to the problem's own solve.  The ranges start at
// multiples of 8, so every problem keeps its workgroups' XCD assignment.
// Problems that have converged return at once in every launch.
struct BItem {
    LevelGeom G;
    TMarch M;
    float dxf;
    int off, nb;  // workgroups of the stencil and direction kernels
    const void *i1;
    float4 *r, *x;
    void *pb[2];  // direction of even / odd iterations
    CgState *st;
    DD *pdd, *pdd2;
    float *pmax;
};

__device__ __forceinline__ int bitem_of(const BItem *it, int P, int b)
{
    int j = 0;
    while (j + 1 < P && b >= it[j + 1].off) ++j;
    return j;
}

template <class PS, class IM>
__global__ void __launch_bounds__(256) kb_cg_pq(const BItem *__restrict__ it, int P, int par)
{
    const BItem &I = it[bitem_of(it, P, (int)blockIdx.x)];
    cg_pq_tm_body<PS, IM>((int)blockIdx.x - I.off, I.nb, (const typename PS::T *)I.pb[par],
                          (const typename IM::T *)I.i1, I.G, I.M, I.dxf, I.st, I.pdd2);
}

template <class PS, class IM>
__global__ void __launch_bounds__(256) kb_cg_update(const BItem *__restrict__ it, int P, int par)
{
    const BItem &I = it[bitem_of(it, P, (int)blockIdx.x)];
    cg_update_tm_body<PS, IM>((int)blockIdx.x - I.off, I.nb, I.r,
                              (const typename PS::T *)I.pb[par], (const typename IM::T *)I.i1,
                              I.G, I.M, I.dxf, I.st, (const DD *)I.pdd2, I.pdd, I.pmax);
}

__global__ void kb_cg_beta(const BItem *__restrict__ it)
{
    const BItem &I = it[blockIdx.x];
    cg_beta_body(I.st, I.pdd, I.pmax, I.nb);
}

template <class PS, bool ODD, class IM>
__global__ void __launch_bounds__(256) kb_cg_dir(const BItem *__restrict__ it, int P, int par)
{
    const BItem &I = it[bitem_of(it, P, (int)blockIdx.x)];
    cg_dir_tm_body<PS, ODD, IM>((int)blockIdx.x - I.off, I.nb, (const typename PS::T *)I.pb[par],
                                (typename PS::T *)I.pb[par ^ 1], I.x, I.r,
                                (const typename IM::T *)I.i1, I.G, I.M, I.dxf, I.st);
}

__global__ void kb_gather_state(const BItem *__restrict__ it, int P, CgState *out)
{
    for (int j = threadIdx.x; j < P; j += blockDim.x) out[j] = *it[j].st;
}

template <class PS, class IM>
static void enqueue_fused_iter(flof_ctx *ctx, const BItem *tab, int P, int total, int k)
{
    const int par = k & 1;
    hipLaunchKernelGGL((kb_cg_pq<PS, IM>), dim3(total), dim3(256), 0, ctx->stream, tab, P, par);
    hipLaunchKernelGGL((kb_cg_update<PS, IM>), dim3(total), dim3(256), 0, ctx->stream, tab, P,
                       par);
    hipLaunchKernelGGL(kb_cg_beta, dim3(P), dim3(256), 0, ctx->stream, tab);
    if (par)
        hipLaunchKernelGGL((kb_cg_dir<PS, true, IM>), dim3(total), dim3(256), 0, ctx->stream,
                           tab, P, par);
    else
        hipLaunchKernelGGL((kb_cg_dir<PS, false, IM>), dim3(total), dim3(256), 0, ctx->stream,
                           tab, P, par);
}

// The PCG iterations of every level in Ls (each set up by level_setup with
// fused = true), until all have converged or reached their cap; the final
// states go to each level's hs.
static void fused_iterations(flof_ctx *ctx, std::vector<LevelCg *> &Ls)
{
    const int P = (int)Ls.size();
    if (P == 0) return;
    std::vector<BItem> h(P);
    int total = 0, max_iter = 0;
    bool all_done = true;
    for (int j = 0; j < P; ++j) {
        LevelCg &L = *Ls[j];
        FLOF_REQUIRE(!L.W.lds && L.W.nb_tm % FLOF_NXCD == 0, FLOF_ERR_STATE,
                     "fused batch: a level is not set up for shared launches");
        BItem &I = h[j];
        I.G = L.G;
        I.M = L.W.tm;
        I.dxf = L.dxf;
        I.off = total;
        I.nb = L.W.nb_tm;
        I.i1 = L.i1;
        I.r = L.W.r.as<float4>();
        I.x = L.W.x.as<float4>();
        I.pb[0] = L.pb[0];
        I.pb[1] = L.pb[1];
        I.st = L.st();
        I.pdd = L.W.pdd();
        I.pdd2 = L.W.pdd2();
        I.pmax = L.W.pmax();
        total += I.nb;
        max_iter = std::max(max_iter, L.hs.maxIter);
        all_done = all_done && L.hs.done;
    }
    if (all_done) return;
    Tmp tab(ctx, (size_t)P * sizeof(BItem));
    Tmp states(ctx, (size_t)P * sizeof(CgState));
    FLOF_HIP(hipMemcpyAsync(tab.p, h.data(), (size_t)P * sizeof(BItem), hipMemcpyHostToDevice,
                            ctx->stream));
    FLOF_HIP(hipStreamSynchronize(ctx->stream));  // h lives on the host stack
    const BItem *dtab = tab.as<BItem>();
    const int pstore = Ls[0]->pstore;
    const bool bf = ctx->img_bf16;
    const int CHUNK = 16;
    auto enqueue_chunk = [&]() {
        for (int k = 0; k < CHUNK; ++k) {
            if (bf) {
                if (pstore == FLOF_PSTORE_BF16)
                    enqueue_fused_iter<PStoreBF16, ImgBF16>(ctx, dtab, P, total, k);
                else
                    enqueue_fused_iter<PStoreF32, ImgBF16>(ctx, dtab, P, total, k);
            } else {
                if (pstore == FLOF_PSTORE_BF16)
                    enqueue_fused_iter<PStoreBF16, ImgF32>(ctx, dtab, P, total, k);
                else
                    enqueue_fused_iter<PStoreF32, ImgF32>(ctx, dtab, P, total, k);
            }
        }
        check_launch("kb_cg_*");
    };
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    if (ctx->use_graphs) {
        FLOF_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        enqueue_chunk();
        hipError_t ce = hipStreamEndCapture(ctx->stream, &graph);
        if (ce == hipSuccess && graph) ce = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (ce != hipSuccess) {
            (void)hipGetLastError();
            if (graph) (void)hipGraphDestroy(graph);
            graph = nullptr;
            exec = nullptr;
        }
    }
    CgState *slots = nullptr;
    FLOF_HIP(hipHostMalloc((void **)&slots, 2 * (size_t)P * sizeof(CgState), hipHostMallocDefault));
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipError_t loop_err = hipEventCreateWithFlags(&ev[0], hipEventDisableTiming);
    if (loop_err == hipSuccess) loop_err = hipEventCreateWithFlags(&ev[1], hipEventDisableTiming);
    int64_t chunks = 0;
    const int64_t max_chunks = (int64_t)max_iter / CHUNK + 3;
    // the host waits for the chunk before the last one, as solve_level does
    while (loop_err == hipSuccess && !all_done && chunks < max_chunks) {
        if (exec) {
            loop_err = hipGraphLaunch(exec, ctx->stream);
            if (loop_err != hipSuccess) break;
        } else {
            enqueue_chunk();
        }
        hipLaunchKernelGGL(kb_gather_state, dim3(1), dim3(64), 0, ctx->stream, dtab, P,
                           states.as<CgState>());
        loop_err = hipMemcpyAsync(slots + (chunks & 1) * P, states.p, (size_t)P * sizeof(CgState),
                                  hipMemcpyDeviceToHost, ctx->stream);
        if (loop_err == hipSuccess) loop_err = hipEventRecord(ev[chunks & 1], ctx->stream);
        if (loop_err != hipSuccess) break;
        ++chunks;
        if (chunks >= 2) {
            loop_err = hipEventSynchronize(ev[chunks & 1]);
            if (loop_err != hipSuccess) break;
            all_done = true;
            for (int j = 0; j < P; ++j) all_done = all_done && slots[(chunks & 1) * P + j].done;
        }
    }
    if (loop_err == hipSuccess) loop_err = hipStreamSynchronize(ctx->stream);
    if (loop_err == hipSuccess && chunks >= 1)
        for (int j = 0; j < P; ++j) Ls[j]->hs = slots[((chunks - 1) & 1) * P + j];
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    (void)hipHostFree(slots);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    FLOF_HIP(loop_err);
    check_launch("kb_cg_*");
}

// ---------------------------------------------------------------------------
// multiscale driver (optflow4d.cpp:936-1173)

struct MsParams {
    float wSmooth, wEnergy, postVelBlur, cgAccuracy, cfl, resetBndWidth;
    int multiStep, projSizeThresh, minGridSize, doFinalProject, cgMaxIter;
};

static void record(flof_solve_info *info, int level, int step, const SolveOut &so)
{
    if (!info || level >= FLOF_MAX_LEVELS || step >= FLOF_MAX_STEPS) return;
    info->iters[level][step] = so.iters;
    info->residual[level][step] = so.resRel;
    info->residual_l2[level][step] = so.resL2;
    info->status[level][step] = so.status;
    info->warm_used[level] = so.warm;
    if (step + 1 > info->steps[level]) info->steps[level] = step + 1;
    info->total_iters += so.iters;
}

// level solve with the post-processing of opticalFlowDim (blur, band reset)
static SolveOut of_level_post(flof_ctx *ctx, float4 *velCurr, const void *i0w,
                              const void *i1, float *rhsT, const Dims4 &D, float wSmooth,
                              float wEnergy, float postVelBlur, float cgAccuracy,
                              float resetBndWidth, float dt, int maxIter,
                              const float4 *warm, const float4 *warm_base)
{
    SolveOut so = solve_level(ctx, i0w, i1, nullptr, rhsT, D, wSmooth, wEnergy, cgAccuracy,
                              dt, maxIter, warm, warm_base, velCurr, false);
    if (postVelBlur > 0.f)
        gaussian_blur(ctx, velCurr, D, true, (float)(0.5 * postVelBlur), 1);
    if (resetBndWidth > 0.f) {
        const int rb = int(resetBndWidth * (float)D.X) + 1;
        launch_set_bound(ctx, velCurr, D, true, make_float4(0.f, 0.f, 0.f, 0.f), rb - 1);
    }
    return so;
}

// Extensions of the multiscale solve (no reference counterpart): a learned
// warm start of a level's CG from that level's pre-warped i0 and i1, and a
// copy of a level's pre-warped i0 and increment (training data of the
// network).  Both act on levels that run one plain solve (multiStep 1, no
// projection).
struct LevelHooks {
    // guess (cells) of the increment of `level`; false: cold start
    std::function<bool(int level, const void *i0w, const void *i1, const Dims4 &D,
                       float4 *guess)>
        predict;
    int capture_level = -1;
    float *cap_i0w = nullptr;
    float4 *cap_inc = nullptr;
};

static float multiscale(flof_ctx *ctx, float4 *vel, const void *i0, const void *i1,
                        float *rhsT, const Dims4 &D, const MsParams &P, int level, float dt,
                        flof_solve_info *info, const float4 *warm,
                        const LevelHooks *hooks = nullptr)
{
    const int64_t n = (int64_t)D.st * D.T;
    // images: FP32, or bfloat16 in the mixed-precision form (projection,
    // capture and the network's warm start take FP32 images only)
    const bool bf = ctx->img_bf16;
    const size_t ib = bf ? 2 : 4;
    const size_t rb = (size_t)n * ib, vb = (size_t)n * 16;
    auto img_lsdiff = [&](const void *a, const void *b, const Dims4 &Dl, float corr, int bnd) {
        return bf ? host_ls_diff_bf16(ctx, a, b, Dl, corr, bnd)
                  : host_ls_diff(ctx, (const float *)a, (const float *)b, nullptr, Dl, corr,
                                 bnd);
    };
    auto img_pyramid = [&](void *dst, const Dims4 &Dd, const void *src, const Dims4 &Dsrc) {
        if (bf)
            launch_pyramid_bf16(ctx, dst, Dd, src, Dsrc);
        else
            launch_pyramid(ctx, dst, Dd, src, Dsrc, false, 1.f);
    };
    auto img_prewarp = [&](const float4 *v, void *g) {
        if (bf) {
            advect_cfl_bf16(ctx, v, g, D, P.cfl, dt);
            launch_neumann_bf16(ctx, g, D, 0);
        } else {
            advect_cfl(ctx, v, g, D, false, P.cfl, 1.f, dt);
            launch_neumann(ctx, g, D, false, 0);
        }
    };
    int multiStep = P.multiStep;
    bool doFinalProject = P.doFinalProject != 0;
    const int resetBnd = P.resetBndWidth > 0 ? int(P.resetBndWidth * (float)D.X) + 1 : 0;
    const float projMaxDist = 4.f;
    const int projMaxIter = 40;
    const float lsDiffFac = (float)(0.1 / 20.);
    if (info && level < FLOF_MAX_LEVELS) {
        info->dims[level][0] = D.X;
        info->dims[level][1] = D.Y;
        info->dims[level][2] = D.Z;
        info->dims[level][3] = D.T;
        if (level + 1 > info->levels) info->levels = level + 1;
    }
    Tmp i0w(ctx, rb);
    FLOF_HIP(hipMemcpyAsync(i0w.p, i0, rb, hipMemcpyDeviceToDevice, ctx->stream));
    if (info && level == 0) info->err_pre = img_lsdiff(i0, i1, D, lsDiffFac, resetBnd);

    if (D.X > P.minGridSize) {
        Dims4 Ds{D.X / 2, D.Y / 2, D.Z / 2, D.T / 2, 0, 0};
        Ds.sz = (int64_t)Ds.X * Ds.Y;
        Ds.st = Ds.sz * Ds.Z;
        FLOF_REQUIRE(Ds.X >= 1 && Ds.Y >= 1 && Ds.Z >= 1 && Ds.T >= 1, FLOF_ERR_DIM,
                     "opticalFlowMultiscale4d: pyramid level too small");
        const int64_t ns = (int64_t)Ds.st * Ds.T;
        Tmp i0s(ctx, (size_t)ns * ib), i1s(ctx, (size_t)ns * ib), vs(ctx, (size_t)ns * 16),
            dbg(ctx, (size_t)ns * 4), ws;
        img_pyramid(i0s.p, Ds, i0, D);
        img_pyramid(i1s.p, Ds, i1, D);
        launch_pyramid(ctx, vs.p, Ds, vel, D, true, 0.5f);
        if (warm) {
            ws = Tmp(ctx, (size_t)ns * 16);
            launch_pyramid(ctx, ws.p, Ds, warm, D, true, 0.5f);
        }
        launch_fill(ctx, dbg.p, (size_t)ns * 4);
        multiscale(ctx, vs.as<float4>(), i0s.p, i1s.p, dbg.as<float>(), Ds, P, level + 1, dt,
                   info, warm ? ws.as<float4>() : nullptr, hooks);
        launch_pyramid(ctx, vel, D, vs.p, Ds, true, 2.f);
        if (rhsT && !bf)
            FLOF_HIP(hipMemcpyAsync(rhsT, i0w.p, rb, hipMemcpyDeviceToDevice, ctx->stream));
    }

    img_prewarp(vel, i0w.p);
    double errCurr = 0.;
    if (info || multiStep > 1) errCurr = img_lsdiff(i0w.p, i1, D, lsDiffFac, resetBnd);
    if (info && level < FLOF_MAX_LEVELS) info->err_level[level] = errCurr;

    bool doProject = false;
    if (D.X > P.projSizeThresh) {
        doProject = true;
        multiStep = 1;
        doFinalProject = false;
    }

    if (multiStep > 1) {
        FLOF_REQUIRE(multiStep < FLOF_MAX_STEPS, FLOF_ERR_RANGE, "Too many of substeps!");
        std::vector<Tmp> vsv, vs2;
        for (int s = 0; s < multiStep; ++s) {
            vs2.emplace_back(ctx, vb);
            launch_fill(ctx, vs2.back().p, vb);
        }
        for (int s = 0; s < multiStep; ++s) vsv.emplace_back();
        Tmp tmpVel(ctx, vb);
        float velBlur = P.postVelBlur;
        double errLast = errCurr;
        int steps = multiStep;
        for (int of = 0; of < steps; ++of) {
            vsv[of] = Tmp(ctx, vb);
            SolveOut so = of_level_post(ctx, vsv[of].as<float4>(), i0w.p, i1, rhsT,
                                        D, P.wSmooth, P.wEnergy, velBlur, P.cgAccuracy,
                                        P.resetBndWidth, dt, P.cgMaxIter, nullptr, nullptr);
            record(info, level, of, so);
            velBlur = (float)(velBlur * (3. / 4.));
            if (velBlur < 2.f) velBlur = 2.f;
            for (int k = of; k >= 0; --k)
                FLOF_HIP(hipMemcpyAsync(vs2[k].p, vsv[k].p, vb, hipMemcpyDeviceToDevice,
                                        ctx->stream));
            for (int k = of - 1; k >= 0; --k)
                for (int l = 0; l < k; ++l)
                    advect_cfl(ctx, vs2[k].as<float4>(), vs2[l].p, D, true, P.cfl, 1.f, dt);
            FLOF_HIP(hipMemcpyAsync(tmpVel.p, vel, vb, hipMemcpyDeviceToDevice, ctx->stream));
            for (int k = of; k >= 0; --k)
                launch_axpy4(ctx, tmpVel.as<float4>(), vs2[k].as<float4>(),
                             make_float4(1.f, 1.f, 1.f, 1.f), n);
            Tmp w2(ctx, rb);
            FLOF_HIP(hipMemcpyAsync(w2.p, i0, rb, hipMemcpyDeviceToDevice, ctx->stream));
            img_prewarp(tmpVel.as<float4>(), w2.p);
            const double errC = img_lsdiff(w2.p, i1, D, lsDiffFac, resetBnd);
            FLOF_HIP(hipMemcpyAsync(i0w.p, w2.p, rb, hipMemcpyDeviceToDevice, ctx->stream));
            if (of > 0 && ((float)errC / (float)errLast) > 0.95) {
                launch_fill(ctx, vsv[of].p, vb);
                steps = of + 1;
            }
            errLast = errC;
        }
        for (int of = steps - 1; of >= 0; --of)
            for (int l = 0; l < of; ++l)
                advect_cfl(ctx, vsv[of].as<float4>(), vsv[l].p, D, true, P.cfl, 1.f, dt);
        for (int of = 0; of < steps; ++of)
            launch_axpy4(ctx, vel, vsv[of].as<float4>(), make_float4(1.f, 1.f, 1.f, 1.f), n);
    } else if (!doProject) {
        const bool post = P.postVelBlur > 0.f || P.resetBndWidth > 0.f;
        const bool capture = hooks && hooks->capture_level == level;
        Tmp guess;
        bool predicted = false;
        if (hooks && hooks->predict) {
            guess = Tmp(ctx, vb);
            predicted = hooks->predict(level, i0w.p, i1, D, guess.as<float4>());
        }
        if (capture)
            FLOF_HIP(hipMemcpyAsync(hooks->cap_i0w, i0w.p, rb, hipMemcpyDeviceToDevice,
                                    ctx->stream));
        if (predicted || capture) {
            // the increment on its own (the network's guess is one of the
            // increment, not of the total flow)
            float4 *inc = capture ? hooks->cap_inc : nullptr;
            Tmp own;
            if (!inc) {
                own = Tmp(ctx, vb);
                inc = own.as<float4>();
            }
            SolveOut so = of_level_post(
                ctx, inc, i0w.p, i1, rhsT, D, P.wSmooth, P.wEnergy, P.postVelBlur,
                P.cgAccuracy, P.resetBndWidth, dt, P.cgMaxIter,
                predicted ? guess.as<float4>() : warm, predicted ? nullptr : vel);
            record(info, level, 0, so);
            launch_axpy4(ctx, vel, inc, make_float4(1.f, 1.f, 1.f, 1.f), n);
        } else if (!post) {
            // vel += x/h fused into the finish kernel
            SolveOut so = solve_level(ctx, i0w.p, i1, nullptr, rhsT, D, P.wSmooth,
                                      P.wEnergy, P.cgAccuracy, dt, P.cgMaxIter, warm, vel,
                                      vel, true);
            record(info, level, 0, so);
        } else {
            Tmp velCurr(ctx, vb);
            SolveOut so = of_level_post(ctx, velCurr.as<float4>(), i0w.p, i1, rhsT,
                                        D, P.wSmooth, P.wEnergy, P.postVelBlur, P.cgAccuracy,
                                        P.resetBndWidth, dt, P.cgMaxIter, warm, vel);
            record(info, level, 0, so);
            launch_axpy4(ctx, vel, velCurr.as<float4>(), make_float4(1.f, 1.f, 1.f, 1.f), n);
        }
    } else {
        Tmp velCurr(ctx, vb), velTmp2(ctx, vb);
        launch_fill(ctx, velTmp2.p, vb);
        FLOF_REQUIRE(!bf, FLOF_ERR_ARG, "mixed precision: no projection (FP32 images only)");
        corr_vels_of(ctx, velCurr.as<float4>(), velTmp2.as<float4>(), i0w.as<float>(),
                     (const float *)i1, D,
                     projMaxDist, P.postVelBlur, P.resetBndWidth, projMaxIter, dt);
        launch_axpy4(ctx, vel, velTmp2.as<float4>(), make_float4(1.f, 1.f, 1.f, 1.f), n);
        launch_scale4(ctx, velCurr.as<float4>(), make_float4(-1.f, -1.f, -1.f, -1.f), n);
        launch_axpy4(ctx, vel, velCurr.as<float4>(), make_float4(1.f, 1.f, 1.f, 1.f), n);
    }

    if (level == 0 && doFinalProject) {
        Tmp velCurr(ctx, vb);
        FLOF_REQUIRE(!bf, FLOF_ERR_ARG, "mixed precision: no projection (FP32 images only)");
        corr_vels_of(ctx, velCurr.as<float4>(), vel, (const float *)i0, (const float *)i1, D,
                     projMaxDist, 4.f,
                     P.resetBndWidth, projMaxIter, dt);
    }
    if (level == 0) {
        FLOF_HIP(hipMemcpyAsync(i0w.p, i0, rb, hipMemcpyDeviceToDevice, ctx->stream));
        img_prewarp(vel, i0w.p);
        errCurr = img_lsdiff(i0w.p, i1, D, lsDiffFac, resetBnd);
        if (info) info->err_final = errCurr;
    }
    return (float)errCurr;
}

// The multiscale solve (optflow4d.cpp:936-1173) of n problems at once, for
// one plain solve per level (multiStep 1, no blur, band, projection or warm
// start): the pyramids are built first, then the problems' levels are
// solved from the coarsest level index down, all problems that have a level
// of that index together, their PCG iterations in shared launches
// (fused_iterations).  Per problem the operations and their order are those
// of multiscale(), so every flow and record is the one a single call gives.
// cells from which a level of a fused batch may run on its own
static const int64_t SOLO_MIN_CELLS = (int64_t)1 << 22;

static void multiscale_fused(flof_ctx *ctx, int np, float4 *const *vel_top,
                             const void *const *i0_top, const void *const *i1_top,
                             const Dims4 *D_top, const MsParams &P, float dt,
                             flof_solve_info *const *info, float *errs)
{
    FLOF_REQUIRE(ctx->use_tmarch, FLOF_ERR_ARG, "fused batch: needs the t-marching kernels");
    const bool bf = ctx->img_bf16;
    const size_t ib = bf ? 2 : 4;
    const int resetBnd = 0;  // the band of resetBndWidth is not part of this form
    const float lsDiffFac = (float)(0.1 / 20.);
    struct Prob {
        std::vector<Dims4> D;
        std::vector<const void *> i0, i1;
        std::vector<float4 *> vel;
        std::vector<Tmp> own;
    };
    std::vector<Prob> pr(np);
    int maxL = 0;
    auto lsdiff = [&](const void *a, const void *b, const Dims4 &Dl) {
        return bf ? host_ls_diff_bf16(ctx, a, b, Dl, lsDiffFac, resetBnd)
                  : host_ls_diff(ctx, (const float *)a, (const float *)b, nullptr, Dl,
                                 lsDiffFac, resetBnd);
    };
    auto prewarp = [&](const float4 *v, void *g, const Dims4 &Dl) {
        if (bf) {
            advect_cfl_bf16(ctx, v, g, Dl, P.cfl, dt);
            launch_neumann_bf16(ctx, g, Dl, 0);
        } else {
            advect_cfl(ctx, v, g, Dl, false, P.cfl, 1.f, dt);
            launch_neumann(ctx, g, Dl, false, 0);
        }
    };
    // pyramids: images and the incoming flow (halved) per level
    for (int j = 0; j < np; ++j) {
        Prob &q = pr[j];
        q.D.push_back(D_top[j]);
        q.i0.push_back(i0_top[j]);
        q.i1.push_back(i1_top[j]);
        q.vel.push_back(vel_top[j]);
        if (info[j]) info[j]->err_pre = lsdiff(i0_top[j], i1_top[j], D_top[j]);
        while (q.D.back().X > P.minGridSize) {
            const Dims4 &Dc = q.D.back();
            Dims4 Ds{Dc.X / 2, Dc.Y / 2, Dc.Z / 2, Dc.T / 2, 0, 0};
            Ds.sz = (int64_t)Ds.X * Ds.Y;
            Ds.st = Ds.sz * Ds.Z;
            FLOF_REQUIRE(Ds.X >= 1 && Ds.Y >= 1 && Ds.Z >= 1 && Ds.T >= 1, FLOF_ERR_DIM,
                         "opticalFlowMultiscale4d: pyramid level too small");
            const int64_t ns = (int64_t)Ds.st * Ds.T;
            q.own.emplace_back(ctx, (size_t)ns * ib);
            void *a = q.own.back().p;
            q.own.emplace_back(ctx, (size_t)ns * ib);
            void *b = q.own.back().p;
            q.own.emplace_back(ctx, (size_t)ns * 16);
            float4 *v = q.own.back().as<float4>();
            if (bf) {
                launch_pyramid_bf16(ctx, a, Ds, q.i0.back(), Dc);
                launch_pyramid_bf16(ctx, b, Ds, q.i1.back(), Dc);
            } else {
                launch_pyramid(ctx, a, Ds, q.i0.back(), Dc, false, 1.f);
                launch_pyramid(ctx, b, Ds, q.i1.back(), Dc, false, 1.f);
            }
            launch_pyramid(ctx, v, Ds, q.vel.back(), Dc, true, 0.5f);
            q.D.push_back(Ds);
            q.i0.push_back(a);
            q.i1.push_back(b);
            q.vel.push_back(v);
        }
        const int L = (int)q.D.size();
        maxL = std::max(maxL, L);
        if (info[j]) {
            info[j]->levels = std::min(L, (int)FLOF_MAX_LEVELS);
            for (int l = 0; l < L && l < FLOF_MAX_LEVELS; ++l) {
                info[j]->dims[l][0] = q.D[l].X;
                info[j]->dims[l][1] = q.D[l].Y;
                info[j]->dims[l][2] = q.D[l].Z;
                info[j]->dims[l][3] = q.D[l].T;
            }
        }
    }
    // levels from the coarsest index down; a problem joins at its coarsest
    for (int l = maxL - 1; l >= 0; --l) {
        std::vector<int> act;
        for (int j = 0; j < np; ++j)
            if ((int)pr[j].D.size() > l) act.push_back(j);
        std::vector<std::unique_ptr<LevelCg>> Ls;
        std::vector<Tmp> warped;
        for (int j : act) {
            Prob &q = pr[j];
            const Dims4 &D = q.D[l];
            const int64_t n = (int64_t)D.st * D.T;
            if (l + 1 < (int)q.D.size())
                launch_pyramid(ctx, q.vel[l], D, q.vel[l + 1], q.D[l + 1], true, 2.f);
            warped.emplace_back(ctx, (size_t)n * ib);
            void *w = warped.back().p;
            FLOF_HIP(hipMemcpyAsync(w, q.i0[l], (size_t)n * ib, hipMemcpyDeviceToDevice,
                                    ctx->stream));
            prewarp(q.vel[l], w, D);
            if (info[j] && l < FLOF_MAX_LEVELS) info[j]->err_level[l] = lsdiff(w, q.i1[l], D);
            // A level large enough to fill the chip on its own whose single
            // form runs the LDS-tiled stencil kernels is solved on its own
            // (those kernels run one problem per launch); the others share
            // the launches of the fused iterations.  Either way the result
            // is the problem's own solve, bit for bit.
            Ls.emplace_back(new LevelCg(ctx, D));
            const bool may_solo = n >= SOLO_MIN_CELLS;
            level_setup(ctx, *Ls.back(), w, q.i1[l], nullptr, nullptr, P.wSmooth, P.wEnergy,
                        P.cgAccuracy, dt, P.cgMaxIter, nullptr, nullptr, !may_solo);
        }
        std::vector<LevelCg *> raw;
        for (auto &L : Ls) {
            if (L->W.lds)
                level_iterate(ctx, *L, P.cgMaxIter);
            else
                raw.push_back(L.get());
        }
        fused_iterations(ctx, raw);
        for (size_t a = 0; a < act.size(); ++a) {
            const int j = act[a];
            SolveOut so = level_finish(ctx, *Ls[a], pr[j].vel[l], true);
            record(info[j], l, 0, so);
        }
    }
    // the error of the finally warped i0 (level 0)
    for (int j = 0; j < np; ++j) {
        Prob &q = pr[j];
        const Dims4 &D = q.D[0];
        const int64_t n = (int64_t)D.st * D.T;
        Tmp w(ctx, (size_t)n * ib);
        FLOF_HIP(hipMemcpyAsync(w.p, q.i0[0], (size_t)n * ib, hipMemcpyDeviceToDevice,
                                ctx->stream));
        prewarp(q.vel[0], w.p, D);
        const double e = lsdiff(w.p, q.i1[0], D);
        if (info[j]) info[j]->err_final = e;
        if (errs) errs[j] = (float)e;
    }
}

// the storage precisions of one call (flof_of_params.precision: bit 0 the
// search direction in bfloat16, bit 1 the images in bfloat16 with binary16
// gradients)
struct PStoreScope {
    flof_ctx *c;
    int old;
    bool old_img;
    PStoreScope(flof_ctx *c_, int v) : c(c_), old(c_->p_store), old_img(c_->img_bf16)
    {
        FLOF_REQUIRE(v >= 0 && v <= (FLOF_PSTORE_BF16 | FLOF_PREC_IMG_BF16), FLOF_ERR_ARG,
                     "precision: a combination of FLOF_PSTORE_BF16 and FLOF_PREC_IMG_BF16");
        FLOF_REQUIRE(v == FLOF_PSTORE_F32 || c->use_tmarch, FLOF_ERR_ARG,
                     "precision: reduced storage needs the t-marching kernels");
        c->p_store = (v & FLOF_PSTORE_BF16) ? FLOF_PSTORE_BF16 : FLOF_PSTORE_F32;
        c->img_bf16 = (v & FLOF_PREC_IMG_BF16) != 0;
    }
    ~PStoreScope()
    {
        c->p_store = old;
        c->img_bf16 = old_img;
    }
};

// the images of a call as the solve stores them: the caller's FP32 grids,
// or bfloat16 copies (rounded to nearest even) in the mixed-precision form
struct SolveImages {
    Tmp b0, b1;
    const void *i0, *i1;
    SolveImages(flof_ctx *ctx, const flof_grid4d *g0, const flof_grid4d *g1)
    {
        i0 = g0->data;
        i1 = g1->data;
        if (!ctx->img_bf16) return;
        const int64_t n = g0->n();
        b0 = Tmp(ctx, (size_t)n * 2);
        b1 = Tmp(ctx, (size_t)n * 2);
        launch_to_bf16(ctx, b0.p, (const float *)g0->data, n);
        launch_to_bf16(ctx, b1.p, (const float *)g1->data, n);
        i0 = b0.p;
        i1 = b1.p;
    }
};

static MsParams ms_params(const flof_of_params *p)
{
    FLOF_REQUIRE(p->blurType == 1 || p->postVelBlur <= 0., FLOF_ERR_ARG, "NYI");
    MsParams P;
    P.wSmooth = (float)p->wSmooth;
    P.wEnergy = (float)p->wEnergy;
    P.postVelBlur = (float)p->postVelBlur;
    P.cgAccuracy = (float)p->cgAccuracy;
    P.cfl = (float)p->cfl;
    P.resetBndWidth = (float)p->resetBndWidth;
    P.multiStep = p->multiStep;
    P.projSizeThresh = p->projSizeThresh;
    P.minGridSize = p->minGridSize;
    P.doFinalProject = p->doFinalProject;
    P.cgMaxIter = p->cgMaxIter;
    FLOF_REQUIRE(P.multiStep < FLOF_MAX_STEPS, FLOF_ERR_RANGE, "Too many of substeps!");
    return P;
}

// Worker contexts of ctx (same device, own stream and pool), created on
// first use with the parent's launch configuration.
static flof_ctx *worker_ctx(flof_ctx *ctx, int w)
{
    while ((int)ctx->workers.size() <= w) {
        flof_ctx *c = nullptr;
        const int rc = flof_ctx_create(&c, ctx->device);
        FLOF_REQUIRE(rc == FLOF_OK && c, rc ? rc : FLOF_ERR_HIP,
                     "opticalFlowMultiscale4dBatch: cannot create a worker stream");
        c->use_graphs = ctx->use_graphs;
        c->blocks_per_cu = ctx->blocks_per_cu;
        c->use_tmarch = ctx->use_tmarch;
        c->tm_blocks_per_cu = ctx->tm_blocks_per_cu;
        c->stencil = ctx->stencil;
        c->lds_blocks_per_cu = ctx->lds_blocks_per_cu;
        c->lds_xcd = ctx->lds_xcd;
        c->autotune = ctx->autotune;
        ctx->workers.push_back(c);
    }
    return ctx->workers[w];
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

static void need4(const flof_grid4d *g, int elem, const char *what)
{
    FLOF_REQUIRE(g, FLOF_ERR_ARG, std::string(what) + ": null grid");
    FLOF_REQUIRE(g->dtype == FLOF_F32, FLOF_ERR_TYPE, std::string(what) + ": FP32 storage");
    FLOF_REQUIRE(g->elem == elem, FLOF_ERR_TYPE, std::string(what) + ": wrong element type");
}
static void need_dims(const flof_grid4d *a, const flof_grid4d *b, const char *what)
{
    FLOF_REQUIRE(a->X == b->X && a->Y == b->Y && a->Z == b->Z && a->T == b->T,
                 FLOF_ERR_DIM, std::string(what) + ": grid sizes differ");
}

extern "C" {

int flof_gaussian_blur4d(flof_grid4d *grid, float sigma, int iter)
{
    flof_ctx *ctx = grid ? grid->ctx : nullptr;
    API_TRY
    FLOF_REQUIRE(grid && grid->dtype == FLOF_F32 && grid->elem != FLOF_INT, FLOF_ERR_TYPE,
                 "gaussianBlur: REAL or VEC4 FP32 grid");
    gaussian_blur(ctx, grid->data, dims_of(grid), grid->elem == FLOF_VEC4, sigma, iter);
    API_CATCH(ctx)
}

int flof_optical_flow4d(flof_grid4d *vel, flof_grid4d *i0, flof_grid4d *i1,
                        flof_grid4d *rhsT, const flof_of_params *p, flof_solve_info *info)
{
    flof_ctx *ctx = vel ? vel->ctx : nullptr;
    API_TRY
    need4(vel, FLOF_VEC4, "opticalFlow4d(vel)");
    need4(i0, FLOF_REAL, "opticalFlow4d(i0)");
    need4(i1, FLOF_REAL, "opticalFlow4d(i1)");
    need_dims(vel, i0, "opticalFlow4d");
    need_dims(vel, i1, "opticalFlow4d");
    if (rhsT) {
        need4(rhsT, FLOF_REAL, "opticalFlow4d(rhsT)");
        need_dims(vel, rhsT, "opticalFlow4d");
    }
    FLOF_REQUIRE(p, FLOF_ERR_ARG, "opticalFlow4d: null params");
    FLOF_REQUIRE(p->blurType == 1 || p->postVelBlur <= 0., FLOF_ERR_ARG, "NYI");
    PStoreScope ps(ctx, p->precision);
    Dims4 D = dims_of(vel);
    const int64_t n = vel->n();
    // the incoming flow feeds the rhs; the solution replaces it
    Tmp vin(ctx, (size_t)n * 16);
    FLOF_HIP(hipMemcpyAsync(vin.p, vel->data, (size_t)n * 16, hipMemcpyDeviceToDevice,
                            ctx->stream));
    SolveImages im(ctx, i0, i1);
    SolveOut so = solve_level(ctx, im.i0, im.i1,
                              vin.as<float4>(), rhsT ? (float *)rhsT->data : nullptr, D,
                              (float)p->wSmooth, (float)p->wEnergy, (float)p->cgAccuracy, 1.f,
                              p->cgMaxIter, nullptr, nullptr, (float4 *)vel->data, false);
    if (p->postVelBlur > 0.)
        gaussian_blur(ctx, vel->data, D, true, (float)(0.5 * (float)p->postVelBlur), 1);
    if (p->resetBndWidth > 0.) {
        const int rb = int((float)p->resetBndWidth * (float)D.X) + 1;
        launch_set_bound(ctx, vel->data, D, true, make_float4(0.f, 0.f, 0.f, 0.f), rb - 1);
    }
    if (info) {
        std::memset(info, 0, sizeof(*info));
        info->levels = 1;
        info->dims[0][0] = D.X;
        info->dims[0][1] = D.Y;
        info->dims[0][2] = D.Z;
        info->dims[0][3] = D.T;
        record(info, 0, 0, so);
    }
    FLOF_HIP(hipStreamSynchronize(ctx->stream));
    API_CATCH(ctx)
}

int flof_optical_flow_multiscale4d(flof_grid4d *vel, flof_grid4d *i0, flof_grid4d *i1,
                                   flof_grid4d *rhsT, const flof_of_params *p,
                                   flof_grid4d *warm, flof_solve_info *info, float *err)
{
    flof_ctx *ctx = vel ? vel->ctx : nullptr;
    API_TRY
    need4(vel, FLOF_VEC4, "opticalFlowMultiscale4d(vel)");
    need4(i0, FLOF_REAL, "opticalFlowMultiscale4d(i0)");
    need4(i1, FLOF_REAL, "opticalFlowMultiscale4d(i1)");
    need_dims(vel, i0, "opticalFlowMultiscale4d");
    need_dims(vel, i1, "opticalFlowMultiscale4d");
    if (rhsT) {
        need4(rhsT, FLOF_REAL, "opticalFlowMultiscale4d(rhsT)");
        need_dims(vel, rhsT, "opticalFlowMultiscale4d");
    }
    FLOF_REQUIRE(p, FLOF_ERR_ARG, "opticalFlowMultiscale4d: null params");
    if (p->warmStart) {
        FLOF_REQUIRE(warm, FLOF_ERR_ARG, "opticalFlowMultiscale4d: warmStart without warm grid");
        need4(warm, FLOF_VEC4, "opticalFlowMultiscale4d(warm)");
        need_dims(vel, warm, "opticalFlowMultiscale4d");
    }
    const MsParams P = ms_params(p);
    PStoreScope ps(ctx, p->precision);
    flof_solve_info local;
    flof_solve_info *inf = info ? info : &local;
    std::memset(inf, 0, sizeof(*inf));
    FLOF_REQUIRE(!(ctx->img_bf16 && rhsT), FLOF_ERR_ARG,
                 "opticalFlowMultiscale4d: rhsT with FP32 images only");
    SolveImages im(ctx, i0, i1);
    const float e = multiscale(ctx, (float4 *)vel->data, im.i0, im.i1,
                               rhsT ? (float *)rhsT->data : nullptr, dims_of(vel), P, p->level,
                               1.f, inf, p->warmStart ? (const float4 *)warm->data : nullptr);
    if (err) *err = e;
    FLOF_HIP(hipStreamSynchronize(ctx->stream));
    API_CATCH(ctx)
}

int flof_optical_flow_multiscale4d_net(flof_grid4d *vel, flof_grid4d *i0, flof_grid4d *i1,
                                       const flof_of_params *p, flof_warmnet *net,
                                       int levels, flof_solve_info *info, float *err)
{
    flof_ctx *ctx = vel ? vel->ctx : nullptr;
    API_TRY
    need4(vel, FLOF_VEC4, "opticalFlowMultiscale4d(vel)");
    need4(i0, FLOF_REAL, "opticalFlowMultiscale4d(i0)");
    need4(i1, FLOF_REAL, "opticalFlowMultiscale4d(i1)");
    need_dims(vel, i0, "opticalFlowMultiscale4d");
    need_dims(vel, i1, "opticalFlowMultiscale4d");
    FLOF_REQUIRE(p && net, FLOF_ERR_ARG, "opticalFlowMultiscale4d (net): null argument");
    FLOF_REQUIRE(!p->warmStart, FLOF_ERR_ARG,
                 "opticalFlowMultiscale4d (net): the network replaces the warm grid");
    const MsParams P = ms_params(p);
    PStoreScope ps(ctx, p->precision);
    flof_solve_info local;
    flof_solve_info *inf = info ? info : &local;
    std::memset(inf, 0, sizeof(*inf));
    LevelHooks hk;
    FLOF_REQUIRE(!ctx->img_bf16, FLOF_ERR_ARG,
                 "opticalFlowMultiscale4d (net): FP32 images only");
    hk.predict = [&](int level, const void *a, const void *b, const Dims4 &D,
                     float4 *guess) -> bool {
        if (levels >= 0 && level >= levels) return false;
        warmnet_infer(ctx, net, (const float *)a, (const float *)b, D, guess);
        return true;
    };
    const float e = multiscale(ctx, (float4 *)vel->data, (const float *)i0->data,
                               (const float *)i1->data, nullptr, dims_of(vel), P, p->level, 1.f,
                               inf, nullptr, &hk);
    if (err) *err = e;
    FLOF_HIP(hipStreamSynchronize(ctx->stream));
    API_CATCH(ctx)
}

int flof_optical_flow_multiscale4d_capture(flof_grid4d *vel, flof_grid4d *i0,
                                           flof_grid4d *i1, const flof_of_params *p,
                                           int level, flof_grid4d *i0w_out,
                                           flof_grid4d *inc_out, flof_solve_info *info,
                                           float *err)
{
    flof_ctx *ctx = vel ? vel->ctx : nullptr;
    API_TRY
    need4(vel, FLOF_VEC4, "opticalFlowMultiscale4d(vel)");
    need4(i0, FLOF_REAL, "opticalFlowMultiscale4d(i0)");
    need4(i1, FLOF_REAL, "opticalFlowMultiscale4d(i1)");
    need_dims(vel, i0, "opticalFlowMultiscale4d");
    need_dims(vel, i1, "opticalFlowMultiscale4d");
    need4(i0w_out, FLOF_REAL, "opticalFlowMultiscale4d (capture i0w)");
    need4(inc_out, FLOF_VEC4, "opticalFlowMultiscale4d (capture increment)");
    need_dims(i0w_out, inc_out, "opticalFlowMultiscale4d (capture)");
    FLOF_REQUIRE(p && level >= 0, FLOF_ERR_ARG, "opticalFlowMultiscale4d (capture): bad argument");
    FLOF_REQUIRE(!p->warmStart, FLOF_ERR_ARG, "opticalFlowMultiscale4d (capture): no warm start");
    FLOF_REQUIRE(p->multiStep <= 1, FLOF_ERR_ARG,
                 "opticalFlowMultiscale4d (capture): one solve per level");
    // the size of the captured level: halved once per level below the top
    Dims4 D = dims_of(vel);
    for (int l = 0; l < level; ++l) {
        FLOF_REQUIRE(D.X > p->minGridSize, FLOF_ERR_RANGE,
                     "opticalFlowMultiscale4d (capture): the pyramid has no such level");
        D = Dims4{D.X / 2, D.Y / 2, D.Z / 2, D.T / 2, 0, 0};
    }
    FLOF_REQUIRE(i0w_out->X == D.X && i0w_out->Y == D.Y && i0w_out->Z == D.Z &&
                     i0w_out->T == D.T,
                 FLOF_ERR_DIM, "opticalFlowMultiscale4d (capture): grids of another size");
    const MsParams P = ms_params(p);
    PStoreScope ps(ctx, p->precision);
    flof_solve_info local;
    flof_solve_info *inf = info ? info : &local;
    std::memset(inf, 0, sizeof(*inf));
    LevelHooks hk;
    FLOF_REQUIRE(!ctx->img_bf16, FLOF_ERR_ARG,
                 "opticalFlowMultiscale4d (capture): FP32 images only");
    hk.capture_level = level;
    hk.cap_i0w = (float *)i0w_out->data;
    hk.cap_inc = (float4 *)inc_out->data;
    const float e = multiscale(ctx, (float4 *)vel->data, (const float *)i0->data,
                               (const float *)i1->data, nullptr, dims_of(vel), P, p->level, 1.f,
                               inf, nullptr, &hk);
    if (err) *err = e;
    FLOF_HIP(hipStreamSynchronize(ctx->stream));
    API_CATCH(ctx)
}

int flof_optical_flow_multiscale4d_batch(flof_ctx *ctx, int n, flof_grid4d *const *vel,
                                         flof_grid4d *const *i0, flof_grid4d *const *i1,
                                         const flof_of_params *p, int concurrency,
                                         flof_solve_info *infos, float *errs)
{
    API_TRY
    FLOF_REQUIRE(ctx && n >= 0 && p && (n == 0 || (vel && i0 && i1)), FLOF_ERR_ARG,
                 "opticalFlowMultiscale4dBatch: bad argument");
    FLOF_REQUIRE(!p->warmStart, FLOF_ERR_ARG,
                 "opticalFlowMultiscale4dBatch: warm starts are per call");
    for (int j = 0; j < n; ++j) {
        need4(vel[j], FLOF_VEC4, "opticalFlowMultiscale4dBatch(vel)");
        need4(i0[j], FLOF_REAL, "opticalFlowMultiscale4dBatch(i0)");
        need4(i1[j], FLOF_REAL, "opticalFlowMultiscale4dBatch(i1)");
        need_dims(vel[j], i0[j], "opticalFlowMultiscale4dBatch");
        need_dims(vel[j], i1[j], "opticalFlowMultiscale4dBatch");
        FLOF_REQUIRE(vel[j]->ctx->device == ctx->device && i0[j]->ctx->device == ctx->device &&
                         i1[j]->ctx->device == ctx->device,
                     FLOF_ERR_ARG, "opticalFlowMultiscale4dBatch: grids on another device");
        for (int m = 0; m < j; ++m)
            FLOF_REQUIRE(vel[m] != vel[j], FLOF_ERR_ARG,
                         "opticalFlowMultiscale4dBatch: one vel grid in two solves");
    }
    const MsParams P = ms_params(p);
    if (n == 0) return FLOF_OK;
    const int W = std::max(1, std::min(std::min(concurrency > 0 ? concurrency : 4, 16), n));
    std::vector<flof_ctx *> wk;
    for (int w = 0; w < W; ++w) wk.push_back(worker_ctx(ctx, w));
    // inputs may still be written by work enqueued on the grids' streams
    FLOF_HIP(hipStreamSynchronize(ctx->stream));
    for (int j = 0; j < n; ++j)
        for (flof_grid4d *g : {vel[j], i0[j], i1[j]})
            if (g->ctx != ctx) FLOF_HIP(hipStreamSynchronize(g->ctx->stream));
    std::atomic<int> next{0};
    // a solve no worker took (every worker failed to select the device)
    // reports that instead of success
    std::vector<int> rc(n, FLOF_ERR_HIP);
    std::vector<std::string> msg(n, "not run: no worker could select the device");
    auto run = [&](flof_ctx *w) {
        if (hipSetDevice(w->device) != hipSuccess) return;
        for (;;) {
            const int j = next.fetch_add(1);
            if (j >= n) break;
            try {
                PStoreScope ps(w, p->precision);
                flof_solve_info local;
                flof_solve_info *inf = infos ? &infos[j] : &local;
                std::memset(inf, 0, sizeof(*inf));
                SolveImages im(w, i0[j], i1[j]);
                const float e = multiscale(w, (float4 *)vel[j]->data, im.i0, im.i1, nullptr,
                                           dims_of(vel[j]), P, p->level, 1.f, inf, nullptr);
                FLOF_HIP(hipStreamSynchronize(w->stream));
                if (errs) errs[j] = e;
                rc[j] = FLOF_OK;
            } catch (const flof::Error &e) {
                rc[j] = e.code;
                msg[j] = e.msg;
                (void)hipStreamSynchronize(w->stream);
            } catch (const std::exception &e) {
                rc[j] = FLOF_ERR_ARG;
                msg[j] = e.what();
                (void)hipStreamSynchronize(w->stream);
            }
        }
    };
    std::vector<std::thread> th;
    for (int w = 1; w < W; ++w) th.emplace_back(run, wk[w]);
    run(wk[0]);
    for (auto &t : th) t.join();
    for (int j = 0; j < n; ++j)
        FLOF_REQUIRE(rc[j] == FLOF_OK, rc[j],
                     "opticalFlowMultiscale4dBatch: solve " + std::to_string(j) + ": " + msg[j]);
    API_CATCH(ctx)
}

int flof_optical_flow_multiscale4d_fused(flof_ctx *ctx, int n, flof_grid4d *const *vel,
                                         flof_grid4d *const *i0, flof_grid4d *const *i1,
                                         const flof_of_params *p, flof_solve_info *infos,
                                         float *errs)
{
    API_TRY
    FLOF_REQUIRE(ctx && n >= 0 && p && (n == 0 || (vel && i0 && i1)), FLOF_ERR_ARG,
                 "opticalFlowMultiscale4dFused: bad argument");
    FLOF_REQUIRE(!p->warmStart && p->multiStep <= 1 && p->postVelBlur <= 0. &&
                     p->resetBndWidth <= 0. && !p->doFinalProject,
                 FLOF_ERR_ARG,
                 "opticalFlowMultiscale4dFused: one plain solve per level (multiStep 1, no "
                 "blur, band, final projection or warm start)");
    for (int j = 0; j < n; ++j) {
        need4(vel[j], FLOF_VEC4, "opticalFlowMultiscale4dFused(vel)");
        need4(i0[j], FLOF_REAL, "opticalFlowMultiscale4dFused(i0)");
        need4(i1[j], FLOF_REAL, "opticalFlowMultiscale4dFused(i1)");
        need_dims(vel[j], i0[j], "opticalFlowMultiscale4dFused");
        need_dims(vel[j], i1[j], "opticalFlowMultiscale4dFused");
        FLOF_REQUIRE(vel[j]->ctx->device == ctx->device && i0[j]->ctx->device == ctx->device &&
                         i1[j]->ctx->device == ctx->device,
                     FLOF_ERR_ARG, "opticalFlowMultiscale4dFused: grids on another device");
        FLOF_REQUIRE(vel[j]->X <= p->projSizeThresh, FLOF_ERR_ARG,
                     "opticalFlowMultiscale4dFused: no projection-only levels");
        for (int m = 0; m < j; ++m)
            FLOF_REQUIRE(vel[m] != vel[j], FLOF_ERR_ARG,
                         "opticalFlowMultiscale4dFused: one vel grid in two solves");
    }
    const MsParams P = ms_params(p);
    if (n == 0) return FLOF_OK;
    for (int j = 0; j < n; ++j)
        for (flof_grid4d *g : {vel[j], i0[j], i1[j]})
            if (g->ctx != ctx) FLOF_HIP(hipStreamSynchronize(g->ctx->stream));
    PStoreScope ps(ctx, p->precision);
    std::vector<SolveImages> ims;
    ims.reserve(n);
    std::vector<float4 *> vp(n);
    std::vector<const void *> a(n), b(n);
    std::vector<Dims4> D(n);
    std::vector<flof_solve_info> local(n);
    std::vector<flof_solve_info *> inf(n);
    for (int j = 0; j < n; ++j) {
        ims.emplace_back(ctx, i0[j], i1[j]);
        vp[j] = (float4 *)vel[j]->data;
        a[j] = ims.back().i0;
        b[j] = ims.back().i1;
        D[j] = dims_of(vel[j]);
        inf[j] = infos ? &infos[j] : &local[j];
        std::memset(inf[j], 0, sizeof(flof_solve_info));
    }
    multiscale_fused(ctx, n, vp.data(), a.data(), b.data(), D.data(), P, 1.f, inf.data(), errs);
    FLOF_HIP(hipStreamSynchronize(ctx->stream));
    API_CATCH(ctx)
}

int flof_corr_vels_of4d(flof_grid4d *dst, flof_grid4d *vel, flof_grid4d *phiOrg,
                        flof_grid4d *phiCurr, flof_grid4d *phiTarget, float threshPhi,
                        float threshNorm, float postVelBlur, float resetBndWidth, int maxIter)
{
    flof_ctx *ctx = vel ? vel->ctx : nullptr;
    (void)phiCurr;
    (void)threshNorm;  // unused by the reference as well (optflow4d.cpp:736,744)
    API_TRY
    need4(dst, FLOF_VEC4, "corrVelsOf4d(dst)");
    need4(vel, FLOF_VEC4, "corrVelsOf4d(vel)");
    need4(phiOrg, FLOF_REAL, "corrVelsOf4d(phiOrg)");
    need4(phiTarget, FLOF_REAL, "corrVelsOf4d(phiTarget)");
    need_dims(vel, dst, "corrVelsOf4d");
    need_dims(vel, phiOrg, "corrVelsOf4d");
    need_dims(vel, phiTarget, "corrVelsOf4d");
    corr_vels_of(ctx, (float4 *)dst->data, (float4 *)vel->data, (const float *)phiOrg->data,
                 (const float *)phiTarget->data, dims_of(vel), threshPhi, postVelBlur,
                 resetBndWidth, maxIter, 1.f);
    FLOF_HIP(hipStreamSynchronize(ctx->stream));
    API_CATCH(ctx)
}

}  // extern "C"
