// flof_warmnet.hip -- the learned warm start of the multiscale solve (no
// reference counterpart; the reference always starts every level's CG from
// zero, source/plugin/optflow4d.cpp:284-303).
//
// A small space-time convolutional network maps the two SDFs of a solve,
// restricted to a coarse pyramid level, to a 4D flow; the flow is prolonged
// to the finest level and handed to the solve as its warm start (the solve
// keeps it only where its residual is not larger than the zero start's).
//
// Network (C = 16 channels, activations FP32 channels-last [cell][16]):
//   h0 = lift(i0/4, i1/4)                       pointwise 2 -> 16
//   4 blocks:  s = Ws * h + bs                  3x3x3 space conv 16 -> 16
//              a = relu(s)
//              h = h + (Wt * a + bt)            3-tap time conv 16 -> 16
//   block b spaces its taps 2^b cells apart (dilation 1, 2, 4, 8), so the
//   network sees 31 cells in x, y, z and t: as far as the flows reach
//   flow = head(h)                              pointwise 16 -> 4
// The convolutions run on the matrix cores: v_mfma_f32_16x16x32_bf16 with
// BF16 weights and BF16-rounded activations, FP32 accumulation.  A wavefront
// computes out^T (16 channels x 16 cells) for two taps per instruction:
//   A (16 x 32) = weights   lane l: out channel l%16, k = 8*(l/16) .. +7
//   B (32 x 16) = inputs    lane l: cell l%16,        k = 8*(l/16) .. +7
//   D (16 x 16) = out^T     lane l: cell l%16, out channels 4*(l/16) .. +3
// with k = 16*(tap - 2*pair) + input channel, so every lane loads 8
// consecutive channels (32 B) of one neighbour and stores 4 consecutive
// output channels (16 B) of one cell.  The backward data pass is the same
// kernel with mirrored taps and transposed weights; the weight gradient uses
// the same instruction with the cells as the reduction dimension.  Training
// is Adam on FP32 master weights.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "flof_internal.h"

namespace flof {

namespace wn {

constexpr int C = 16;       // channels
constexpr int NB = 4;       // residual blocks
constexpr int NS = 27;      // space taps
constexpr int NTT = 3;      // time taps
constexpr int PAIRS_S = 14; // ceil(27 / 2)
constexpr int PAIRS_T = 2;  // ceil(3 / 2)

// parameter layout (FP32, flat)
constexpr int OFF_WL = 0;                   // [2][16]
constexpr int OFF_BL = OFF_WL + 2 * C;      // [16]
constexpr int OFF_BLK = OFF_BL + C;         // blocks
constexpr int BLK_WS = 0;                   // [27][16][16]  (tap, in, out)
constexpr int BLK_BS = BLK_WS + NS * C * C; // [16]
constexpr int BLK_WT = BLK_BS + C;          // [3][16][16]
constexpr int BLK_BT = BLK_WT + NTT * C * C;// [16]
constexpr int BLK_SIZE = BLK_BT + C;        // 7712
constexpr int OFF_WH = OFF_BLK + NB * BLK_SIZE;  // [16][4]
constexpr int OFF_BH = OFF_WH + C * 4;      // [4]
constexpr int NPARAM = OFF_BH + 4;          // 30964

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short to_bf16(float f)
{
    unsigned u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ bf16x8 pack8(const float4 &a, const float4 &b)
{
    u16x8 u;
    u[0] = to_bf16(a.x);
    u[1] = to_bf16(a.y);
    u[2] = to_bf16(a.z);
    u[3] = to_bf16(a.w);
    u[4] = to_bf16(b.x);
    u[5] = to_bf16(b.y);
    u[6] = to_bf16(b.z);
    u[7] = to_bf16(b.w);
    return __builtin_bit_cast(bf16x8, u);
}

struct Geo {
    int X, Y, Z, T;
    int n;    // cells (< 2^31, checked by the host)
    int dil;  // tap spacing of the convolutions
};

__device__ __forceinline__ void cell_coords(const Geo &g, int c, int &i, int &j, int &k, int &t)
{
    i = c % g.X;
    int r = c / g.X;
    j = r % g.Y;
    r /= g.Y;
    k = r % g.Z;
    t = r / g.Z;
}

// tap m of a kind: space taps (dz, dy, dx) in {-1,0,1}^3 with dx fastest;
// time taps dt in {-1, 0, 1}
__device__ __forceinline__ void tap_offset(bool space, int m, int &dx, int &dy, int &dz, int &dt)
{
    if (space) {
        dx = m % 3 - 1;
        dy = (m / 3) % 3 - 1;
        dz = m / 9 - 1;
        dt = 0;
    } else {
        dx = dy = dz = 0;
        dt = m - 1;
    }
}

enum { EP_BIAS = 1, EP_RELU = 2, EP_MASK = 4, EP_RES = 8 };

// out[c][o] = ep( sum_tap sum_i W[tap][i][o] in[c + sgn*off(tap)][i] )
// with zero padding outside the grid; ep: (+ bias[o]) (relu) (* (mask > 0))
// (+ res); out_relu (optional) receives max(out, 0) as well.  wf: packed A
// fragments [pairs][64 lanes] (16 B each).
template <bool SPACE>
__global__ void __launch_bounds__(256)
k_conv(const float *__restrict__ in, float *__restrict__ out, float *__restrict__ out_relu,
       const bf16x8 *__restrict__ wf, const float *__restrict__ bias,
       const float *__restrict__ mask, const float *__restrict__ res, Geo g, int sgn,
       int flags)
{
    constexpr int NT = SPACE ? NS : NTT;
    const int dil = g.dil;
    constexpr int NP = (NT + 1) / 2;
    const int lane = threadIdx.x & 63;
    const int grp = lane >> 4;       // 0..3
    const int sub = grp >> 1;        // tap of the pair
    const int half = grp & 1;        // input channels 8*half .. +7
    const int waves = gridDim.x * (blockDim.x >> 6);
    const int ntiles = (g.n + 15) >> 4;
    for (int tile = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); tile < ntiles;
         tile += waves) {
        const int c = tile * 16 + (lane & 15);
        const bool live = c < g.n;
        int i = 0, j = 0, k = 0, t = 0;
        if (live) cell_coords(g, c, i, j, k, t);
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
        for (int p = 0; p < NP; ++p) {
            const int m = 2 * p + sub;
            float4 lo = make_float4(0.f, 0.f, 0.f, 0.f), hi = lo;
            if (live && m < NT) {
                int dx, dy, dz, dt;
                tap_offset(SPACE, m, dx, dy, dz, dt);
                const int sd = sgn * dil;
                const int ni = i + sd * dx, nj = j + sd * dy, nk = k + sd * dz,
                          nt = t + sd * dt;
                if (ni >= 0 && nj >= 0 && nk >= 0 && nt >= 0 && ni < g.X && nj < g.Y &&
                    nk < g.Z && nt < g.T) {
                    const int nc = ni + g.X * (nj + g.Y * (nk + g.Z * nt));
                    const float4 *src = (const float4 *)(in + (int64_t)nc * C + 8 * half);
                    lo = src[0];
                    hi = src[1];
                }
            }
            const bf16x8 b = pack8(lo, hi);
            const bf16x8 a = wf[p * 64 + lane];
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
        }
        if (!live) continue;
        const int o0 = 4 * grp;
        float v[4] = {acc[0], acc[1], acc[2], acc[3]};
        const int64_t base = (int64_t)c * C + o0;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            if (flags & EP_BIAS) v[r] += bias[o0 + r];
            if (flags & EP_RELU) v[r] = fmaxf(v[r], 0.f);
            if (flags & EP_MASK) v[r] = mask[base + r] > 0.f ? v[r] : 0.f;
            if (flags & EP_RES) v[r] += res[base + r];
        }
        *(float4 *)(out + base) = make_float4(v[0], v[1], v[2], v[3]);
        if (out_relu)
            *(float4 *)(out_relu + base) = make_float4(fmaxf(v[0], 0.f), fmaxf(v[1], 0.f),
                                                       fmaxf(v[2], 0.f), fmaxf(v[3], 0.f));
    }
}

// A fragments of a conv from W[tap][in][out] (FP32).  transpose: the
// backward data pass, whose "out" channel is W's input channel.
__global__ void k_pack_frag(const float *__restrict__ W, bf16x8 *__restrict__ wf, int ntaps,
                            int npairs, int transpose)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= npairs * 64) return;
    const int p = idx >> 6, lane = idx & 63;
    const int o = lane & 15, grp = lane >> 4;
    const int m = 2 * p + (grp >> 1);
    float v[8];
    for (int q = 0; q < 8; ++q) {
        const int ii = 8 * (grp & 1) + q;
        float w = 0.f;
        if (m < ntaps) w = transpose ? W[(m * C + o) * C + ii] : W[(m * C + ii) * C + o];
        v[q] = w;
    }
    wf[idx] = pack8(make_float4(v[0], v[1], v[2], v[3]), make_float4(v[4], v[5], v[6], v[7]));
}

// weight gradient of a forward conv (sgn +1):
//   dW[tap][i][o] = sum_c bf16(in[c + off(tap)][i]) * bf16(dout[c][o])
// D = A B with A (16 x 32) = in^T over 32 cells, B (32 x 16) = dout over the
// same cells; each wavefront accumulates taps M0 .. M0+NT-1 over its cells
// and writes one partial [tap][16][16] (taps of one launch: at most 9, so
// the accumulators stay in registers).
template <bool SPACE, int M0, int NT>
__global__ void __launch_bounds__(256)
k_conv_wgrad(const float *__restrict__ in, const float *__restrict__ dout,
             float *__restrict__ part, Geo g)
{
    const int lane = threadIdx.x & 63;
    const int grp = lane >> 4;
    const int ch = lane & 15;
    const int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int waves = gridDim.x * (blockDim.x >> 6);
    f32x4 acc[NT];
#pragma unroll
    for (int m = 0; m < NT; ++m) acc[m] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int nchunks = (g.n + 31) >> 5;
    for (int chunk = wave; chunk < nchunks; chunk += waves) {
        int ci[8], cj[8], ck[8], ct[8];
        bool live[8];
        float bo[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int c = chunk * 32 + 8 * grp + q;
            live[q] = c < g.n;
            ci[q] = cj[q] = ck[q] = ct[q] = 0;
            if (live[q]) cell_coords(g, c, ci[q], cj[q], ck[q], ct[q]);
            bo[q] = live[q] ? dout[(int64_t)c * C + ch] : 0.f;
        }
        const bf16x8 b = pack8(make_float4(bo[0], bo[1], bo[2], bo[3]),
                               make_float4(bo[4], bo[5], bo[6], bo[7]));
#pragma unroll
        for (int m = 0; m < NT; ++m) {
            int dx, dy, dz, dt;
            tap_offset(SPACE, M0 + m, dx, dy, dz, dt);
            float av[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int ni = ci[q] + g.dil * dx, nj = cj[q] + g.dil * dy,
                          nk = ck[q] + g.dil * dz, nt = ct[q] + g.dil * dt;
                av[q] = (live[q] && ni >= 0 && nj >= 0 && nk >= 0 && nt >= 0 && ni < g.X &&
                         nj < g.Y && nk < g.Z && nt < g.T)
                            ? in[(int64_t)(ni + g.X * (nj + g.Y * (nk + g.Z * nt))) * C + ch]
                            : 0.f;
            }
            const bf16x8 a = pack8(make_float4(av[0], av[1], av[2], av[3]),
                                   make_float4(av[4], av[5], av[6], av[7]));
            acc[m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[m], 0, 0, 0);
        }
    }
    // D[i][o]: lane holds i = 4*grp + r, o = lane % 16
    float *dst = part + (int64_t)wave * NT * C * C;
#pragma unroll
    for (int m = 0; m < NT; ++m)
#pragma unroll
        for (int r = 0; r < 4; ++r) dst[(m * C + 4 * grp + r) * C + ch] = acc[m][r];
}

// rows x cols partial sums -> one row (sum over rows), one thread per column
__global__ void k_sum_rows(const float *__restrict__ part, int rows, int cols,
                           float *__restrict__ dst, int accumulate)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= cols) return;
    double s = 0.;
    for (int r = 0; r < rows; ++r) s += (double)part[(int64_t)r * cols + col];
    dst[col] = accumulate ? dst[col] + (float)s : (float)s;
}

// per-block column sums of a [n][cols] tensor (bias gradients etc.)
__global__ void __launch_bounds__(256)
k_col_partials(const float *__restrict__ a, int n, int cols, float *__restrict__ part)
{
    // blockDim 256 = 16 rows x 16 columns; cols <= 16
    const int col = threadIdx.x & 15;
    const int row0 = threadIdx.x >> 4;
    float s = 0.f;
    if (col < cols)
        for (int r = blockIdx.x * 16 + row0; r < n; r += gridDim.x * 16)
            s += a[(int64_t)r * cols + col];
    __shared__ float sh[256];
    sh[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < 16) {
        float t = 0.f;
        for (int q = 0; q < 16; ++q) t += sh[q * 16 + threadIdx.x];
        if (threadIdx.x < cols) part[(int64_t)blockIdx.x * cols + threadIdx.x] = t;
    }
}

// h0[c][o] = Wl[0][o] * i0/4 + Wl[1][o] * i1/4 + bl[o]
__global__ void k_lift(const float *__restrict__ i0, const float *__restrict__ i1,
                       const float *__restrict__ P, float *__restrict__ h, int n)
{
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        const float x0 = i0[c] * 0.25f, x1 = i1[c] * 0.25f;
        float v[C];
#pragma unroll
        for (int o = 0; o < C; ++o)
            v[o] = P[OFF_WL + o] * x0 + P[OFF_WL + C + o] * x1 + P[OFF_BL + o];
        float4 *d = (float4 *)(h + (int64_t)c * C);
#pragma unroll
        for (int q = 0; q < 4; ++q) d[q] = make_float4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
    }
}

// flow[c][d] = sum_i Wh[i][d] h[c][i] + bh[d]
__global__ void k_head(const float *__restrict__ h, const float *__restrict__ P,
                       float4 *__restrict__ flow, int n)
{
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        const float4 *s = (const float4 *)(h + (int64_t)c * C);
        float hv[C];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float4 t = s[q];
            hv[4 * q] = t.x;
            hv[4 * q + 1] = t.y;
            hv[4 * q + 2] = t.z;
            hv[4 * q + 3] = t.w;
        }
        float o[4];
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            float a = P[OFF_BH + d];
#pragma unroll
            for (int i = 0; i < C; ++i) a += P[OFF_WH + i * 4 + d] * hv[i];
            o[d] = a;
        }
        flow[c] = make_float4(o[0], o[1], o[2], o[3]);
    }
}

// loss = mean((flow - tgt)^2) over 4n values: dy = 2 (flow - tgt) / (4n);
// per-block partial sums of the squared error (FP64)
__global__ void __launch_bounds__(256)
k_loss(const float4 *__restrict__ flow, const float4 *__restrict__ tgt,
       float4 *__restrict__ dy, int n, double *__restrict__ part)
{
    const float sc = 2.f / (4.f * (float)n);
    double s = 0.;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        const float4 f = flow[c], t = tgt[c];
        const float4 e = make_float4(f.x - t.x, f.y - t.y, f.z - t.z, f.w - t.w);
        s += (double)e.x * e.x + (double)e.y * e.y + (double)e.z * e.z + (double)e.w * e.w;
        dy[c] = make_float4(sc * e.x, sc * e.y, sc * e.z, sc * e.w);
    }
    __shared__ double sh[256];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// head backward: dh[c][i] = sum_d Wh[i][d] dy[c][d]; per-block partials of
// dWh[i][d] = sum_c h[c][i] dy[c][d] (64) and dbh[d] = sum_c dy[c][d] (4)
__global__ void __launch_bounds__(256)
k_head_bwd(const float *__restrict__ h, const float4 *__restrict__ dy,
           const float *__restrict__ P, float *__restrict__ dh, int n,
           float *__restrict__ part)
{
    float gw[C * 4], gb[4];
#pragma unroll
    for (int q = 0; q < C * 4; ++q) gw[q] = 0.f;
#pragma unroll
    for (int d = 0; d < 4; ++d) gb[d] = 0.f;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        const float4 e = dy[c];
        const float ev[4] = {e.x, e.y, e.z, e.w};
        const float *hc = h + (int64_t)c * C;
        float *dc = dh + (int64_t)c * C;
#pragma unroll
        for (int i = 0; i < C; ++i) {
            float a = 0.f;
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                a += P[OFF_WH + i * 4 + d] * ev[d];
                gw[i * 4 + d] += hc[i] * ev[d];
            }
            dc[i] = a;
        }
#pragma unroll
        for (int d = 0; d < 4; ++d) gb[d] += ev[d];
    }
    __shared__ float sh[256];
    for (int q = 0; q < C * 4 + 4; ++q) {
        sh[threadIdx.x] = q < C * 4 ? gw[q] : gb[q - C * 4];
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[(int64_t)blockIdx.x * (C * 4 + 4) + q] = sh[0];
        __syncthreads();
    }
}

// lift backward: per-block partials of dWl[k][o] = sum_c x_k[c] dh0[c][o]
// (32) and dbl[o] = sum_c dh0[c][o] (16)
__global__ void __launch_bounds__(256)
k_lift_bwd(const float *__restrict__ i0, const float *__restrict__ i1,
           const float *__restrict__ dh, int n, float *__restrict__ part)
{
    float gw[2 * C], gb[C];
#pragma unroll
    for (int q = 0; q < 2 * C; ++q) gw[q] = 0.f;
#pragma unroll
    for (int q = 0; q < C; ++q) gb[q] = 0.f;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        const float x0 = i0[c] * 0.25f, x1 = i1[c] * 0.25f;
        const float *d = dh + (int64_t)c * C;
#pragma unroll
        for (int o = 0; o < C; ++o) {
            gw[o] += x0 * d[o];
            gw[C + o] += x1 * d[o];
            gb[o] += d[o];
        }
    }
    __shared__ float sh[256];
    for (int q = 0; q < 3 * C; ++q) {
        sh[threadIdx.x] = q < 2 * C ? gw[q] : gb[q - 2 * C];
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[(int64_t)blockIdx.x * (3 * C) + q] = sh[0];
        __syncthreads();
    }
}

__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace wn
}  // namespace flof

using namespace flof;
using namespace flof::wn;

struct flof_warmnet {
    flof_ctx *ctx = nullptr;
    float *P = nullptr;      // FP32 master parameters
    float *G = nullptr;      // gradients
    float *M = nullptr, *V = nullptr;  // Adam moments
    bf16x8 *frag = nullptr;  // packed A fragments, forward and backward
    int step = 0;
    bool frag_ok = false;
};

namespace {

constexpr int FRAG_S = PAIRS_S * 64, FRAG_T = PAIRS_T * 64;
// fragment slots per block: space fwd, space bwd, time fwd, time bwd
constexpr int FRAG_BLK = 2 * FRAG_S + 2 * FRAG_T;

void pack_all(flof_warmnet *net)
{
    flof_ctx *ctx = net->ctx;
    for (int b = 0; b < NB; ++b) {
        const float *blk = net->P + OFF_BLK + b * BLK_SIZE;
        bf16x8 *f = net->frag + (size_t)b * FRAG_BLK;
        hipLaunchKernelGGL(k_pack_frag, dim3((FRAG_S + 255) / 256), dim3(256), 0, ctx->stream,
                           blk + BLK_WS, f, NS, PAIRS_S, 0);
        hipLaunchKernelGGL(k_pack_frag, dim3((FRAG_S + 255) / 256), dim3(256), 0, ctx->stream,
                           blk + BLK_WS, f + FRAG_S, NS, PAIRS_S, 1);
        hipLaunchKernelGGL(k_pack_frag, dim3(1), dim3(256), 0, ctx->stream, blk + BLK_WT,
                           f + 2 * FRAG_S, NTT, PAIRS_T, 0);
        hipLaunchKernelGGL(k_pack_frag, dim3(1), dim3(256), 0, ctx->stream, blk + BLK_WT,
                           f + 2 * FRAG_S + FRAG_T, NTT, PAIRS_T, 1);
    }
    check_launch("k_pack_frag");
    net->frag_ok = true;
}

Geo geo_of(const flof_grid4d *g)
{
    FLOF_REQUIRE(g->n() < (1ll << 31) / C, FLOF_ERR_DIM, "warm-start net: grid too large");
    return Geo{g->X, g->Y, g->Z, g->T, (int)g->n(), 1};
}

int conv_blocks(flof_ctx *ctx, int n)
{
    const int tiles = (n + 15) / 16;
    return std::max(1, std::min((tiles + 3) / 4, ctx->num_cu * 8));
}

template <bool SPACE>
void conv(flof_ctx *ctx, const float *in, float *out, float *out_relu, const bf16x8 *wf,
          const float *bias, const float *mask, const float *res, const Geo &g, int sgn,
          int flags)
{
    hipLaunchKernelGGL(k_conv<SPACE>, dim3(conv_blocks(ctx, g.n)), dim3(256), 0, ctx->stream,
                       in, out, out_relu, wf, bias, mask, res, g, sgn, flags);
    check_launch("k_conv");
}

// activations of one forward pass, kept for the backward pass
struct Acts {
    Tmp h[NB + 1], s[NB], a[NB];
};

void forward(flof_warmnet *net, const float *i0, const float *i1, const Geo &g0, Acts &A,
             float4 *flow)
{
    const Geo &g = g0;
    flof_ctx *ctx = net->ctx;
    if (!net->frag_ok) pack_all(net);
    const size_t ab = (size_t)g.n * C * 4;
    for (int b = 0; b <= NB; ++b) A.h[b] = Tmp(ctx, ab);
    for (int b = 0; b < NB; ++b) {
        A.s[b] = Tmp(ctx, ab);
        A.a[b] = Tmp(ctx, ab);
    }
    hipLaunchKernelGGL(k_lift, dim3(grid_blocks(ctx, g.n)), dim3(256), 0, ctx->stream, i0, i1,
                       (const float *)net->P, A.h[0].as<float>(), g.n);
    check_launch("k_lift");
    for (int b = 0; b < NB; ++b) {
        const float *blk = net->P + OFF_BLK + b * BLK_SIZE;
        const bf16x8 *f = net->frag + (size_t)b * FRAG_BLK;
        Geo g = g0;
        g.dil = 1 << b;
        conv<true>(ctx, A.h[b].as<float>(), A.s[b].as<float>(), A.a[b].as<float>(), f,
                   blk + BLK_BS, nullptr, nullptr, g, 1, EP_BIAS);
        conv<false>(ctx, A.a[b].as<float>(), A.h[b + 1].as<float>(), nullptr, f + 2 * FRAG_S,
                    blk + BLK_BT, nullptr, A.h[b].as<float>(), g, 1, EP_BIAS | EP_RES);
    }
    hipLaunchKernelGGL(k_head, dim3(grid_blocks(ctx, g.n)), dim3(256), 0, ctx->stream,
                       (const float *)A.h[NB].as<float>(), (const float *)net->P, flow, g.n);
    check_launch("k_head");
}

}  // namespace

void flof::warmnet_infer(flof_ctx *ctx, flof_warmnet *net, const float *i0, const float *i1,
                         const Dims4 &D, float4 *flow)
{
    FLOF_REQUIRE(net && net->ctx == ctx, FLOF_ERR_ARG,
                 "warm-start net: the network belongs to another context");
    const int64_t n = D.st * D.T;
    FLOF_REQUIRE(n >= 1 && n < (1ll << 30), FLOF_ERR_DIM,
                 "warm-start net: 1 to 2^30 cells");
    const Geo g0{D.X, D.Y, D.Z, D.T, (int)n, 1};
    if (!net->frag_ok) pack_all(net);
    const size_t ab = (size_t)n * C * 4;
    Tmp h(ctx, ab), h2(ctx, ab), s(ctx, ab), a(ctx, ab);
    hipLaunchKernelGGL(k_lift, dim3(grid_blocks(ctx, n)), dim3(256), 0, ctx->stream, i0, i1,
                       (const float *)net->P, h.as<float>(), (int)n);
    check_launch("k_lift");
    for (int b = 0; b < NB; ++b) {
        const float *blk = net->P + OFF_BLK + b * BLK_SIZE;
        const bf16x8 *f = net->frag + (size_t)b * FRAG_BLK;
        Geo g = g0;
        g.dil = 1 << b;
        conv<true>(ctx, h.as<float>(), s.as<float>(), a.as<float>(), f, blk + BLK_BS, nullptr,
                   nullptr, g, 1, EP_BIAS);
        conv<false>(ctx, a.as<float>(), h2.as<float>(), nullptr, f + 2 * FRAG_S, blk + BLK_BT,
                    nullptr, h.as<float>(), g, 1, EP_BIAS | EP_RES);
        std::swap(h, h2);
    }
    hipLaunchKernelGGL(k_head, dim3(grid_blocks(ctx, n)), dim3(256), 0, ctx->stream,
                       (const float *)h.as<float>(), (const float *)net->P, flow, (int)n);
    check_launch("k_head");
}


namespace {

int wgrad_blocks(flof_ctx *ctx, int n)
{
    const int chunks = (n + 31) / 32;
    return std::max(1, std::min((chunks + 3) / 4, ctx->num_cu * 2));
}

// dW (and db) of one conv: partials per wavefront, then a column sum
template <bool SPACE, int M0, int NT>
void wgrad_taps(flof_ctx *ctx, const float *in, const float *dout, const Geo &g, float *dW)
{
    const int nbk = wgrad_blocks(ctx, g.n);
    const int waves = nbk * 4;
    Tmp part(ctx, (size_t)waves * NT * C * C * 4);
    hipLaunchKernelGGL((k_conv_wgrad<SPACE, M0, NT>), dim3(nbk), dim3(256), 0, ctx->stream, in,
                       dout, part.as<float>(), g);
    check_launch("k_conv_wgrad");
    hipLaunchKernelGGL(k_sum_rows, dim3((NT * C * C + 255) / 256), dim3(256), 0, ctx->stream,
                       (const float *)part.as<float>(), waves, NT * C * C, dW + M0 * C * C, 0);
    check_launch("k_sum_rows");
}

template <bool SPACE>
void conv_wgrad(flof_ctx *ctx, const float *in, const float *dout, const Geo &g, float *dW,
                float *db)
{
    if (SPACE) {
        wgrad_taps<true, 0, 9>(ctx, in, dout, g, dW);
        wgrad_taps<true, 9, 9>(ctx, in, dout, g, dW);
        wgrad_taps<true, 18, 9>(ctx, in, dout, g, dW);
    } else {
        wgrad_taps<false, 0, NTT>(ctx, in, dout, g, dW);
    }
    const int cb = grid_blocks(ctx, (g.n + 15) / 16, 1, 4);
    Tmp bp(ctx, (size_t)cb * C * 4);
    hipLaunchKernelGGL(k_col_partials, dim3(cb), dim3(256), 0, ctx->stream, dout, g.n, C,
                       bp.as<float>());
    hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(256), 0, ctx->stream,
                       (const float *)bp.as<float>(), cb, C, db, 0);
    check_launch("k_col_partials");
}

// forward + loss + backward; gradients into net->G, returns the loss
double loss_grad(flof_warmnet *net, const float *i0, const float *i1, const float4 *tgt,
                 const Geo &g0)
{
    const Geo &g = g0;
    flof_ctx *ctx = net->ctx;
    Acts A;
    Tmp flow(ctx, (size_t)g.n * 16), dy(ctx, (size_t)g.n * 16);
    forward(net, i0, i1, g, A, flow.as<float4>());
    const int lb = grid_blocks(ctx, g.n, 256, 4);
    Tmp lp(ctx, (size_t)lb * 8);
    hipLaunchKernelGGL(k_loss, dim3(lb), dim3(256), 0, ctx->stream,
                       (const float4 *)flow.as<float4>(), tgt, dy.as<float4>(), g.n,
                       lp.as<double>());
    check_launch("k_loss");
    const size_t ab = (size_t)g.n * C * 4;
    Tmp dh(ctx, ab), dh2(ctx, ab), ds(ctx, ab);
    const int hb = grid_blocks(ctx, g.n, 256, 2);
    Tmp hp(ctx, (size_t)hb * (C * 4 + 4) * 4);
    hipLaunchKernelGGL(k_head_bwd, dim3(hb), dim3(256), 0, ctx->stream,
                       (const float *)A.h[NB].as<float>(), (const float4 *)dy.as<float4>(),
                       (const float *)net->P, dh.as<float>(), g.n, hp.as<float>());
    check_launch("k_head_bwd");
    hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(256), 0, ctx->stream,
                       (const float *)hp.as<float>(), hb, C * 4 + 4, net->G + OFF_WH, 0);
    for (int b = NB - 1; b >= 0; --b) {
        Geo g = g0;
        g.dil = 1 << b;
        float *gb = net->G + OFF_BLK + b * BLK_SIZE;
        const float *blk = net->P + OFF_BLK + b * BLK_SIZE;
        const bf16x8 *f = net->frag + (size_t)b * FRAG_BLK;
        (void)blk;
        // h_{b+1} = h_b + Wt * a + bt
        conv_wgrad<false>(ctx, A.a[b].as<float>(), dh.as<float>(), g, gb + BLK_WT,
                          gb + BLK_BT);
        // ds = (Wt^T * dh) masked by s > 0
        conv<false>(ctx, dh.as<float>(), ds.as<float>(), nullptr, f + 2 * FRAG_S + FRAG_T,
                    nullptr, A.s[b].as<float>(), nullptr, g, -1, EP_MASK);
        conv_wgrad<true>(ctx, A.h[b].as<float>(), ds.as<float>(), g, gb + BLK_WS,
                         gb + BLK_BS);
        // dh_b = dh_{b+1} + Ws^T * ds
        conv<true>(ctx, ds.as<float>(), dh2.as<float>(), nullptr, f + FRAG_S, nullptr,
                   nullptr, dh.as<float>(), g, -1, EP_RES);
        std::swap(dh, dh2);
    }
    const int fb = grid_blocks(ctx, g.n, 256, 2);
    Tmp fp(ctx, (size_t)fb * 3 * C * 4);
    hipLaunchKernelGGL(k_lift_bwd, dim3(fb), dim3(256), 0, ctx->stream, i0, i1,
                       (const float *)dh.as<float>(), g.n, fp.as<float>());
    check_launch("k_lift_bwd");
    hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(256), 0, ctx->stream,
                       (const float *)fp.as<float>(), fb, 3 * C, net->G + OFF_WL, 0);
    check_launch("k_sum_rows");
    std::vector<double> h(lb);
    FLOF_HIP(hipMemcpyAsync(h.data(), lp.p, (size_t)lb * 8, hipMemcpyDeviceToHost,
                            ctx->stream));
    FLOF_HIP(hipStreamSynchronize(ctx->stream));
    double s = 0.;
    for (double v : h) s += v;
    return s / (4.0 * (double)g.n);
}

void need_pair(const flof_grid4d *i0, const flof_grid4d *i1, const flof_grid4d *v,
               const char *what)
{
    FLOF_REQUIRE(i0 && i1 && v, FLOF_ERR_ARG, std::string(what) + ": null grid");
    FLOF_REQUIRE(i0->elem == FLOF_REAL && i1->elem == FLOF_REAL && v->elem == FLOF_VEC4 &&
                     i0->dtype == FLOF_F32 && i1->dtype == FLOF_F32 && v->dtype == FLOF_F32,
                 FLOF_ERR_TYPE, std::string(what) + ": REAL inputs and a VEC4 flow (FP32)");
    for (const flof_grid4d *q : {i1, v})
        FLOF_REQUIRE(q->X == i0->X && q->Y == i0->Y && q->Z == i0->Z && q->T == i0->T,
                     FLOF_ERR_DIM, std::string(what) + ": grid sizes differ");
}

}  // namespace

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

int flof_warmnet_num_params(void) { return NPARAM; }

int flof_warmnet_create(flof_ctx *ctx, uint64_t seed, flof_warmnet **out)
{
    API_TRY
    FLOF_REQUIRE(ctx && out, FLOF_ERR_ARG, "warmnet_create: null argument");
    *out = nullptr;
    flof_warmnet *net = new flof_warmnet();
    net->ctx = ctx;
    try {
        const size_t pb = (size_t)NPARAM * 4;
        FLOF_HIP(hipMalloc((void **)&net->P, pb));
        FLOF_HIP(hipMalloc((void **)&net->G, pb));
        FLOF_HIP(hipMalloc((void **)&net->M, pb));
        FLOF_HIP(hipMalloc((void **)&net->V, pb));
        FLOF_HIP(hipMalloc((void **)&net->frag, (size_t)NB * FRAG_BLK * sizeof(bf16x8)));
        FLOF_HIP(hipMemsetAsync(net->G, 0, pb, ctx->stream));
        FLOF_HIP(hipMemsetAsync(net->M, 0, pb, ctx->stream));
        FLOF_HIP(hipMemsetAsync(net->V, 0, pb, ctx->stream));
        // He-style initial weights from a seeded splitmix64 stream; the
        // residual branches and the head start small
        std::vector<float> w(NPARAM, 0.f);
        uint64_t s = seed;
        auto uni = [&]() {
            s += 0x9E3779B97F4A7C15ull;
            uint64_t z = s;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            return (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
        };
        for (int q = 0; q < 2 * C; ++q) w[OFF_WL + q] = (float)(uni() * 1.0);
        for (int b = 0; b < NB; ++b) {
            float *blk = w.data() + OFF_BLK + b * BLK_SIZE;
            const double as = std::sqrt(3.0 / (NS * C)), at = 0.1 * std::sqrt(3.0 / (NTT * C));
            for (int q = 0; q < NS * C * C; ++q) blk[BLK_WS + q] = (float)(uni() * as);
            for (int q = 0; q < NTT * C * C; ++q) blk[BLK_WT + q] = (float)(uni() * at);
        }
        for (int q = 0; q < C * 4; ++q) w[OFF_WH + q] = (float)(uni() * 0.05);
        FLOF_HIP(hipMemcpyAsync(net->P, w.data(), pb, hipMemcpyHostToDevice, ctx->stream));
        FLOF_HIP(hipStreamSynchronize(ctx->stream));
    } catch (...) {
        flof_warmnet_destroy(net);
        throw;
    }
    *out = net;
    API_CATCH(ctx)
}

int flof_warmnet_destroy(flof_warmnet *net)
{
    if (!net) return FLOF_ERR_ARG;
    if (net->ctx) (void)hipStreamSynchronize(net->ctx->stream);
    for (void *p : {(void *)net->P, (void *)net->G, (void *)net->M, (void *)net->V,
                    (void *)net->frag})
        if (p) (void)hipFree(p);
    delete net;
    return FLOF_OK;
}

int flof_warmnet_get_params(flof_warmnet *net, float *host, int n)
{
    flof_ctx *ctx = net ? net->ctx : nullptr;
    API_TRY
    FLOF_REQUIRE(net && host && n == NPARAM, FLOF_ERR_ARG, "warmnet_get_params: bad argument");
    FLOF_HIP(hipMemcpyAsync(host, net->P, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream));
    FLOF_HIP(hipStreamSynchronize(ctx->stream));
    API_CATCH(ctx)
}

int flof_warmnet_set_params(flof_warmnet *net, const float *host, int n)
{
    flof_ctx *ctx = net ? net->ctx : nullptr;
    API_TRY
    FLOF_REQUIRE(net && host && n == NPARAM, FLOF_ERR_ARG, "warmnet_set_params: bad argument");
    FLOF_HIP(hipMemcpyAsync(net->P, host, (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream));
    FLOF_HIP(hipStreamSynchronize(ctx->stream));
    net->frag_ok = false;
    API_CATCH(ctx)
}

int flof_warmnet_get_grads(flof_warmnet *net, float *host, int n)
{
    flof_ctx *ctx = net ? net->ctx : nullptr;
    API_TRY
    FLOF_REQUIRE(net && host && n == NPARAM, FLOF_ERR_ARG, "warmnet_get_grads: bad argument");
    FLOF_HIP(hipMemcpyAsync(host, net->G, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream));
    FLOF_HIP(hipStreamSynchronize(ctx->stream));
    API_CATCH(ctx)
}

int flof_warmnet_forward(flof_warmnet *net, flof_grid4d *i0, flof_grid4d *i1,
                         flof_grid4d *flow)
{
    flof_ctx *ctx = net ? net->ctx : nullptr;
    API_TRY
    FLOF_REQUIRE(net, FLOF_ERR_ARG, "warmnet_forward: null net");
    need_pair(i0, i1, flow, "warmnet_forward");
    const Geo g = geo_of(i0);
    Acts A;
    forward(net, (const float *)i0->data, (const float *)i1->data, g, A, (float4 *)flow->data);
    FLOF_HIP(hipStreamSynchronize(ctx->stream));
    API_CATCH(ctx)
}

int flof_warmnet_loss_grad(flof_warmnet *net, flof_grid4d *i0, flof_grid4d *i1,
                           flof_grid4d *target, double *loss)
{
    flof_ctx *ctx = net ? net->ctx : nullptr;
    API_TRY
    FLOF_REQUIRE(net, FLOF_ERR_ARG, "warmnet_loss_grad: null net");
    need_pair(i0, i1, target, "warmnet_loss_grad");
    const double l = loss_grad(net, (const float *)i0->data, (const float *)i1->data,
                               (const float4 *)target->data, geo_of(i0));
    if (loss) *loss = l;
    API_CATCH(ctx)
}

int flof_warmnet_adam(flof_warmnet *net, float lr, float beta1, float beta2, float eps)
{
    flof_ctx *ctx = net ? net->ctx : nullptr;
    API_TRY
    FLOF_REQUIRE(net && lr > 0.f && beta1 >= 0.f && beta1 < 1.f && beta2 >= 0.f && beta2 < 1.f,
                 FLOF_ERR_ARG, "warmnet_adam: bad argument");
    net->step += 1;
    const float c1 = (float)(1.0 - std::pow((double)beta1, net->step));
    const float c2 = (float)(1.0 - std::pow((double)beta2, net->step));
    hipLaunchKernelGGL(k_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, ctx->stream, net->P,
                       (const float *)net->G, net->M, net->V, NPARAM, lr, beta1, beta2, eps,
                       c1, c2);
    check_launch("k_adam");
    net->frag_ok = false;
    FLOF_HIP(hipStreamSynchronize(ctx->stream));
    API_CATCH(ctx)
}

int flof_warmnet_warm_start(flof_warmnet *net, flof_grid4d *i0, flof_grid4d *i1, int netX,
                            flof_grid4d *warm)
{
    flof_ctx *ctx = net ? net->ctx : nullptr;
    API_TRY
    FLOF_REQUIRE(net && netX >= 2, FLOF_ERR_ARG, "warmnet_warm_start: bad argument");
    need_pair(i0, i1, warm, "warmnet_warm_start");
    // the pyramid of the multiscale solve down to X <= netX
    std::vector<Dims4> dims{dims_of(i0)};
    while (dims.back().X > netX) {
        const Dims4 &D = dims.back();
        Dims4 Ds{D.X / 2, D.Y / 2, D.Z / 2, D.T / 2, 0, 0};
        Ds.sz = (int64_t)Ds.X * Ds.Y;
        Ds.st = Ds.sz * Ds.Z;
        FLOF_REQUIRE(Ds.X >= 2 && Ds.Y >= 2 && Ds.Z >= 2 && Ds.T >= 2, FLOF_ERR_DIM,
                     "warmnet_warm_start: pyramid level too small");
        dims.push_back(Ds);
    }
    const size_t L = dims.size();
    std::vector<Tmp> a(L), b(L), v(L);
    const float *ca = (const float *)i0->data, *cb = (const float *)i1->data;
    for (size_t l = 1; l < L; ++l) {
        const int64_t n = dims[l].st * dims[l].T;
        a[l] = Tmp(ctx, (size_t)n * 4);
        b[l] = Tmp(ctx, (size_t)n * 4);
        launch_pyramid(ctx, a[l].p, dims[l], ca, dims[l - 1], false, 1.f);
        launch_pyramid(ctx, b[l].p, dims[l], cb, dims[l - 1], false, 1.f);
        ca = a[l].as<float>();
        cb = b[l].as<float>();
    }
    const Dims4 &Dn = dims[L - 1];
    FLOF_REQUIRE(Dn.st * Dn.T < (1ll << 31) / C, FLOF_ERR_DIM, "warmnet: grid too large");
    const Geo g{Dn.X, Dn.Y, Dn.Z, Dn.T, (int)(Dn.st * Dn.T), 1};
    float4 *out = L == 1 ? (float4 *)warm->data : nullptr;
    if (L > 1) {
        v[L - 1] = Tmp(ctx, (size_t)g.n * 16);
        out = v[L - 1].as<float4>();
    }
    {
        Acts A;
        forward(net, ca, cb, g, A, out);
    }
    for (size_t l = L - 1; l >= 1; --l) {
        void *dst = l == 1 ? warm->data : nullptr;
        if (l > 1) {
            v[l - 1] = Tmp(ctx, (size_t)dims[l - 1].st * dims[l - 1].T * 16);
            dst = v[l - 1].p;
        }
        launch_pyramid(ctx, dst, dims[l - 1], v[l].p, dims[l], true, 2.f);
    }
    FLOF_HIP(hipStreamSynchronize(ctx->stream));
    API_CATCH(ctx)
}

}  // extern "C"
