// flof_mesh.hip -- the level-set mesh of the output frames and its .bobj.gz
// file (the last step of the pipeline: scenes/ofHelpers.py:116-122 meshes
// every warped frame and saves it).
//
// Reference: LevelsetGrid::createMesh (source/levelset.cpp:266-365), Mesh::
// computeVertexNormals (source/mesh.cpp:460-479), writeBobjFile
// (source/fileio.cpp:121-160).
//
// The reference visits the cubes one after another (x outermost, then y,
// then z) and numbers a vertex when the first cube that uses its grid edge
// meets it.  Here every cube works at once: a cube owns an active edge when
// no earlier cube (in that visiting order) that contains the edge is valid
// (a cube is skipped when a corner is at or below the fast-marching invalid
// time -1000), and counts its owned edges and its triangles; exclusive scans
// over the cubes in visiting order give every cube the numbers of its first
// vertex and first triangle, which are exactly the reference's numbers.  The
// owner computes the vertex from its own corners (the edge's orientation in
// that cube fixes the rounding).  The vertex normals of the file sum the
// triangles' contributions in triangle order: a vertex walks the cubes that
// contain its edge in visiting order and, in each, the cube's triangles in
// table order -- the reference's accumulation order, so all FP32 results
// are identical.

#include <zlib.h>

#include <algorithm>
#include <string>
#include <vector>

#include "flof_internal.h"
#include "flof_mc_tables.h"

namespace flof {
namespace {

__constant__ mc::McTable d_mc = mc::make_mc_table();

struct MeshGeo {
    int X, Y, Z;
    int64_t nc;  // (X-1)(Y-1)(Z-1) cubes
};

__device__ __forceinline__ float phi_at(const float *phi, const MeshGeo &g, int i, int j, int k)
{
    return phi[(int64_t)i + (int64_t)g.X * (j + (int64_t)g.Y * k)];
}

// cube (i, j, k) <-> its number in the reference's visiting order
__device__ __forceinline__ int64_t cube_id(const MeshGeo &g, int i, int j, int k)
{
    return ((int64_t)i * (g.Y - 1) + j) * (g.Z - 1) + k;
}
__device__ __forceinline__ void cube_ijk(const MeshGeo &g, int64_t t, int &i, int &j, int &k)
{
    const int64_t yz = (int64_t)(g.Y - 1) * (g.Z - 1);
    i = (int)(t / yz);
    const int64_t r = t - (int64_t)i * yz;
    j = (int)(r / (g.Z - 1));
    k = (int)(r - (int64_t)j * (g.Z - 1));
}

// the case of a valid cube, -1 for a skipped cube or one with no edge
__device__ int cube_case(const float *phi, const MeshGeo &g, int i, int j, int k)
{
    const float invalidTime = -1000.f, iso = 1e-4f;
    int cs = 0;
    bool skip = false;
#pragma unroll
    for (int l = 0; l < 8; ++l) {
        const float v = -phi_at(phi, g, i + mc::kMcCorner[l][0], j + mc::kMcCorner[l][1],
                                k + mc::kMcCorner[l][2]);
        if (-v <= invalidTime) skip = true;
        if (v < iso) cs |= 1 << l;
    }
    if (skip || mc::mc_edge_mask(cs) == 0) return -1;
    return cs;
}

// grid edge of slot e of cube (i, j, k): direction and its lower end
__device__ __forceinline__ void slot_edge(int e, int i, int j, int k, int &dir, int &gi, int &gj,
                                          int &gk)
{
    const int a = mc::kMcEdge[e][0], b = mc::kMcEdge[e][1];
    dir = mc::kMcCorner[a][0] != mc::kMcCorner[b][0] ? 0
          : mc::kMcCorner[a][1] != mc::kMcCorner[b][1] ? 1
                                                       : 2;
    gi = i + min(mc::kMcCorner[a][0], mc::kMcCorner[b][0]);
    gj = j + min(mc::kMcCorner[a][1], mc::kMcCorner[b][1]);
    gk = k + min(mc::kMcCorner[a][2], mc::kMcCorner[b][2]);
}

// the cubes containing grid edge (dir, gi, gj, gk), in visiting order
// (lexicographic in i, j, k); out of range ones are marked with id -1
__device__ __forceinline__ void edge_cubes(const MeshGeo &g, int dir, int gi, int gj, int gk,
                                           int ci[4], int cj[4], int ck[4], int64_t id[4])
{
    int q = 0;
    for (int a = 1; a >= 0; --a)
        for (int b = 1; b >= 0; --b) {
            int i = gi, j = gj, k = gk;
            if (dir == 0) { j -= a; k -= b; }
            if (dir == 1) { i -= a; k -= b; }
            if (dir == 2) { i -= a; j -= b; }
            ci[q] = i;
            cj[q] = j;
            ck[q] = k;
            const bool in = i >= 0 && j >= 0 && k >= 0 && i < g.X - 1 && j < g.Y - 1 && k < g.Z - 1;
            id[q] = in ? cube_id(g, i, j, k) : -1;
            ++q;
        }
}

// edge slot e of cube t is owned by t: no earlier valid cube contains it
__device__ bool owns(const float *phi, const MeshGeo &g, int i, int j, int k, int e)
{
    int dir, gi, gj, gk;
    slot_edge(e, i, j, k, dir, gi, gj, gk);
    int ci[4], cj[4], ck[4];
    int64_t id[4];
    edge_cubes(g, dir, gi, gj, gk, ci, cj, ck, id);
    const int64_t self = cube_id(g, i, j, k);
    for (int q = 0; q < 4; ++q) {
        if (id[q] < 0) continue;
        if (id[q] == self) return true;
        if (cube_case(phi, g, ci[q], cj[q], ck[q]) >= 0) return false;
    }
    return false;
}

__global__ void k_mc_count(const float *__restrict__ phi, MeshGeo g, short *__restrict__ cases,
                           int *__restrict__ nvert, int *__restrict__ ntri)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < g.nc;
         t += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        cube_ijk(g, t, i, j, k);
        const int cs = cube_case(phi, g, i, j, k);
        int nv = 0, nt = 0;
        if (cs >= 0) {
            const int em = mc::mc_edge_mask(cs);
            for (int e = 0; e < 12; ++e)
                if ((em >> e) & 1) nv += owns(phi, g, i, j, k, e) ? 1 : 0;
            nt = d_mc.ntri[cs];
        }
        cases[t] = (short)cs;
        nvert[t] = nv;
        ntri[t] = nt;
    }
}

__device__ __forceinline__ int64_t edge_index(const MeshGeo &g, int dir, int gi, int gj, int gk)
{
    return (int64_t)dir * ((int64_t)g.X * g.Y * g.Z) + gi +
           (int64_t)g.X * (gj + (int64_t)g.Y * gk);
}

__global__ void k_mc_vertices(const float *__restrict__ phi, MeshGeo g,
                              const short *__restrict__ cases, const int *__restrict__ vbase,
                              float *__restrict__ pos, int *__restrict__ eid,
                              int64_t *__restrict__ vedge)
{
    const float iso = 1e-4f;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < g.nc;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int cs = cases[t];
        if (cs < 0) continue;
        int i, j, k;
        cube_ijk(g, t, i, j, k);
        const int em = mc::mc_edge_mask(cs);
        int id = vbase[t];
        for (int e = 0; e < 12; ++e) {
            if (!((em >> e) & 1) || !owns(phi, g, i, j, k, e)) continue;
            const int a = mc::kMcEdge[e][0], b = mc::kMcEdge[e][1];
            const float v1 = -phi_at(phi, g, i + mc::kMcCorner[a][0], j + mc::kMcCorner[a][1],
                                     k + mc::kMcCorner[a][2]);
            const float v2 = -phi_at(phi, g, i + mc::kMcCorner[b][0], j + mc::kMcCorner[b][1],
                                     k + mc::kMcCorner[b][2]);
            const float mu = (iso - v1) / (v2 - v1);
            const float p1[3] = {(float)(i + mc::kMcCorner[a][0]), (float)(j + mc::kMcCorner[a][1]),
                                 (float)(k + mc::kMcCorner[a][2])};
            const float p2[3] = {(float)(i + mc::kMcCorner[b][0]), (float)(j + mc::kMcCorner[b][1]),
                                 (float)(k + mc::kMcCorner[b][2])};
            for (int d = 0; d < 3; ++d) pos[3 * (int64_t)id + d] = (p1[d] + (p2[d] - p1[d]) * mu) + 0.5f;
            int dir, gi, gj, gk;
            slot_edge(e, i, j, k, dir, gi, gj, gk);
            const int64_t ei = edge_index(g, dir, gi, gj, gk);
            eid[ei] = id;
            vedge[id] = ei;
            ++id;
        }
    }
}

__global__ void k_mc_triangles(MeshGeo g, const short *__restrict__ cases,
                               const int *__restrict__ tbase, const int *__restrict__ eid,
                               int *__restrict__ tris)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < g.nc;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int cs = cases[t];
        if (cs < 0) continue;
        int i, j, k;
        cube_ijk(g, t, i, j, k);
        const int nt = d_mc.ntri[cs];
        const int64_t tb = tbase[t];
        for (int q = 0; q < 3 * nt; ++q) {
            int dir, gi, gj, gk;
            slot_edge(d_mc.tri[cs][q], i, j, k, dir, gi, gj, gk);
            tris[3 * tb + q] = eid[edge_index(g, dir, gi, gj, gk)];
        }
    }
}

// corner c of triangle (p0, p1, p2): cross(p0 - p1, p1 - p2) times
// 1 / (|edge before c|^2 |edge after c|^2) in double, stored as float
__device__ __forceinline__ float3 tri_normal_part(const float *pos, const int *tri, int c)
{
    float p[3][3];
    for (int q = 0; q < 3; ++q)
        for (int d = 0; d < 3; ++d) p[q][d] = pos[3 * (int64_t)tri[q] + d];
    const float n0[3] = {p[0][0] - p[1][0], p[0][1] - p[1][1], p[0][2] - p[1][2]};
    const float n1[3] = {p[1][0] - p[2][0], p[1][1] - p[2][1], p[1][2] - p[2][2]};
    const float n2[3] = {p[2][0] - p[0][0], p[2][1] - p[0][1], p[2][2] - p[0][2]};
    const float l0 = n0[0] * n0[0] + n0[1] * n0[1] + n0[2] * n0[2];
    const float l1 = n1[0] * n1[0] + n1[1] * n1[1] + n1[2] * n1[2];
    const float l2 = n2[0] * n2[0] + n2[1] * n2[1] + n2[2] * n2[2];
    const float nm[3] = {(n0[1] * n1[2]) - (n0[2] * n1[1]), (n0[2] * n1[0]) - (n0[0] * n1[2]),
                         (n0[0] * n1[1]) - (n0[1] * n1[0])};
    const float lp = c == 0 ? l0 * l2 : c == 1 ? l0 * l1 : l1 * l2;
    const double f = 1.0 / (double)lp;
    return make_float3((float)((double)nm[0] * f), (float)((double)nm[1] * f),
                       (float)((double)nm[2] * f));
}

__global__ void k_mc_normals(MeshGeo g, int64_t nv, const short *__restrict__ cases,
                             const int *__restrict__ tbase, const int *__restrict__ tris,
                             const float *__restrict__ pos, const int64_t *__restrict__ vedge,
                             float *__restrict__ nrm)
{
    const int64_t ng = (int64_t)g.X * g.Y * g.Z;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x) {
        const int64_t ei = vedge[v];
        const int dir = (int)(ei / ng);
        int64_t r = ei - dir * ng;
        const int gk = (int)(r / ((int64_t)g.X * g.Y));
        r -= (int64_t)gk * g.X * g.Y;
        const int gj = (int)(r / g.X), gi = (int)(r - (int64_t)gj * g.X);
        int ci[4], cj[4], ck[4];
        int64_t id[4];
        edge_cubes(g, dir, gi, gj, gk, ci, cj, ck, id);
        float a[3] = {0.f, 0.f, 0.f};
        for (int q = 0; q < 4; ++q) {
            if (id[q] < 0) continue;
            const int cs = cases[id[q]];
            if (cs < 0) continue;
            const int64_t tb = tbase[id[q]];
            const int nt = d_mc.ntri[cs];
            for (int m = 0; m < nt; ++m) {
                const int *tri = tris + 3 * (tb + m);
                for (int c = 0; c < 3; ++c) {
                    if (tri[c] != (int)v) continue;
                    const float3 f = tri_normal_part(pos, tri, c);
                    a[0] = a[0] + f.x;
                    a[1] = a[1] + f.y;
                    a[2] = a[2] + f.z;
                }
            }
        }
        // normalize (util/vectorbase.h:401-415)
        const float l = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
        const float e2 = 1e-6f * 1e-6f;
        float o[3] = {a[0], a[1], a[2]};
        if (fabs((double)l - 1.) < (double)e2) {
        } else if (l > e2) {
            const float nrmv = (float)sqrt((double)l);
            const float s = (float)(1. / (double)nrmv);
            for (int d = 0; d < 3; ++d) o[d] = a[d] * s;
        } else {
            o[0] = o[1] = o[2] = 0.f;
        }
        for (int d = 0; d < 3; ++d) nrm[3 * v + d] = o[d];
    }
}

// exclusive scan of n ints in place (1024 per workgroup, then the workgroup
// sums by one workgroup, then the offsets); *total receives the sum
constexpr int SCAN_PER = 1024;

__device__ __forceinline__ int block_incl_scan(int v, int *sh)
{
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const int t = tid >= off ? sh[tid - off] : 0;
        __syncthreads();
        sh[tid] += t;
        __syncthreads();
    }
    const int r = sh[tid];
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(256) k_scan_local(int *a, int64_t n, int *bsum)
{
    __shared__ int sh[256];
    const int64_t base = (int64_t)blockIdx.x * SCAN_PER + 4 * threadIdx.x;
    int v[4], s = 0;
    for (int q = 0; q < 4; ++q) {
        v[q] = base + q < n ? a[base + q] : 0;
        const int x = v[q];
        v[q] = s;
        s += x;
    }
    const int incl = block_incl_scan(s, sh);
    const int excl = incl - s;
    for (int q = 0; q < 4; ++q)
        if (base + q < n) a[base + q] = excl + v[q];
    if (threadIdx.x == 255) bsum[blockIdx.x] = incl;
}

__global__ void __launch_bounds__(256) k_scan_sums(int *bsum, int nb, int *total)
{
    __shared__ int sh[256];
    int carry = 0;
    for (int c0 = 0; c0 < nb; c0 += 256) {
        const int i = c0 + (int)threadIdx.x;
        const int v = i < nb ? bsum[i] : 0;
        const int incl = block_incl_scan(v, sh);
        if (i < nb) bsum[i] = carry + incl - v;
        sh[threadIdx.x] = incl;
        __syncthreads();
        carry += sh[255];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ void k_scan_add(int *a, int64_t n, const int *bsum)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        a[i] += bsum[i / SCAN_PER];
}

int scan_exclusive(flof_ctx *ctx, int *a, int64_t n)
{
    const int nb = (int)((n + SCAN_PER - 1) / SCAN_PER);
    Tmp bs(ctx, (size_t)std::max(nb, 1) * 4 + 64);
    int *bsum = bs.as<int>();
    int *total = bsum + std::max(nb, 1);
    hipLaunchKernelGGL(k_scan_local, dim3(std::max(nb, 1)), dim3(256), 0, ctx->stream, a, n,
                       bsum);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(256), 0, ctx->stream, bsum, nb, total);
    hipLaunchKernelGGL(k_scan_add, dim3(grid_blocks(ctx, n)), dim3(256), 0, ctx->stream, a, n,
                       (const int *)bsum);
    check_launch("mesh scan");
    int h = 0;
    FLOF_HIP(hipMemcpyAsync(&h, total, 4, hipMemcpyDeviceToHost, ctx->stream));
    FLOF_HIP(hipStreamSynchronize(ctx->stream));
    return h;
}

}  // namespace
}  // namespace flof

using namespace flof;

struct flof_mesh {
    flof_ctx *ctx = nullptr;
    int X = 0, Y = 0, Z = 0;   // grid of the level set (the file's scale)
    int64_t nv = 0, nt = 0;
    float *pos = nullptr, *nrm = nullptr;
    int *tris = nullptr;
};

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

int flof_create_mesh(flof_ctx *ctx, const float *phi, int X, int Y, int Z, flof_mesh **out)
{
    API_TRY
    FLOF_REQUIRE(ctx && phi && out, FLOF_ERR_ARG, "createMesh: null argument");
    *out = nullptr;
    FLOF_REQUIRE(X >= 2 && Y >= 2 && Z >= 2, FLOF_ERR_DIM,
                 "createMesh: only 3D grids (at least 2 cells per dimension)");
    // vertex (3 per grid point) and triangle (5 per cube) counts and their
    // scans are 32-bit
    FLOF_REQUIRE((int64_t)X * Y * Z < (1ll << 31) / 5, FLOF_ERR_DIM, "createMesh: grid too large");
    MeshGeo g{X, Y, Z, (int64_t)(X - 1) * (Y - 1) * (Z - 1)};
    const int64_t ng = (int64_t)X * Y * Z;
    Tmp cases(ctx, (size_t)g.nc * 2), vb(ctx, (size_t)g.nc * 4), tb(ctx, (size_t)g.nc * 4),
        eid(ctx, (size_t)ng * 3 * 4);
    hipLaunchKernelGGL(k_mc_count, dim3(grid_blocks(ctx, g.nc)), dim3(256), 0, ctx->stream, phi,
                       g, cases.as<short>(), vb.as<int>(), tb.as<int>());
    check_launch("k_mc_count");
    flof_mesh *m = new flof_mesh();
    m->ctx = ctx;
    m->X = X;
    m->Y = Y;
    m->Z = Z;
    try {
        m->nv = scan_exclusive(ctx, vb.as<int>(), g.nc);
        m->nt = scan_exclusive(ctx, tb.as<int>(), g.nc);
        if (m->nv > 0) {
            FLOF_HIP(hipMalloc((void **)&m->pos, (size_t)m->nv * 12));
            FLOF_HIP(hipMalloc((void **)&m->nrm, (size_t)m->nv * 12));
        }
        if (m->nt > 0) FLOF_HIP(hipMalloc((void **)&m->tris, (size_t)m->nt * 12));
        Tmp vedge(ctx, (size_t)std::max<int64_t>(m->nv, 1) * 8);
        if (m->nv > 0) {
            hipLaunchKernelGGL(k_mc_vertices, dim3(grid_blocks(ctx, g.nc)), dim3(256), 0,
                               ctx->stream, phi, g, (const short *)cases.as<short>(),
                               (const int *)vb.as<int>(), m->pos, eid.as<int>(),
                               vedge.as<int64_t>());
            check_launch("k_mc_vertices");
        }
        if (m->nt > 0) {
            hipLaunchKernelGGL(k_mc_triangles, dim3(grid_blocks(ctx, g.nc)), dim3(256), 0,
                               ctx->stream, g, (const short *)cases.as<short>(),
                               (const int *)tb.as<int>(), (const int *)eid.as<int>(), m->tris);
            check_launch("k_mc_triangles");
        }
        if (m->nv > 0) {
            hipLaunchKernelGGL(k_mc_normals, dim3(grid_blocks(ctx, m->nv)), dim3(256), 0,
                               ctx->stream, g, m->nv, (const short *)cases.as<short>(),
                               (const int *)tb.as<int>(), (const int *)m->tris,
                               (const float *)m->pos, (const int64_t *)vedge.as<int64_t>(),
                               m->nrm);
            check_launch("k_mc_normals");
        }
        FLOF_HIP(hipStreamSynchronize(ctx->stream));
    } catch (...) {
        flof_mesh_destroy(m);
        throw;
    }
    *out = m;
    API_CATCH(ctx)
}

int flof_mesh_size(const flof_mesh *m, int64_t *nodes, int64_t *tris)
{
    if (!m) return FLOF_ERR_ARG;
    if (nodes) *nodes = m->nv;
    if (tris) *tris = m->nt;
    return FLOF_OK;
}

int flof_mesh_download(const flof_mesh *m, float *pos, float *normals, int32_t *tris)
{
    flof_ctx *ctx = m ? m->ctx : nullptr;
    API_TRY
    FLOF_REQUIRE(m, FLOF_ERR_ARG, "mesh download: null mesh");
    if (pos && m->nv)
        FLOF_HIP(hipMemcpyAsync(pos, m->pos, (size_t)m->nv * 12, hipMemcpyDeviceToHost,
                                ctx->stream));
    if (normals && m->nv)
        FLOF_HIP(hipMemcpyAsync(normals, m->nrm, (size_t)m->nv * 12, hipMemcpyDeviceToHost,
                                ctx->stream));
    if (tris && m->nt)
        FLOF_HIP(hipMemcpyAsync(tris, m->tris, (size_t)m->nt * 12, hipMemcpyDeviceToHost,
                                ctx->stream));
    FLOF_HIP(hipStreamSynchronize(ctx->stream));
    API_CATCH(ctx)
}

int flof_mesh_save_bobj(const flof_mesh *m, const char *name)
{
    flof_ctx *ctx = m ? m->ctx : nullptr;
    API_TRY
    FLOF_REQUIRE(m && name, FLOF_ERR_ARG, "writeBobj: null argument");
    FLOF_REQUIRE(m->nv < (1ll << 31) && m->nt < (1ll << 31), FLOF_ERR_RANGE,
                 "writeBobj: more than 2^31 nodes or triangles");
    std::vector<float> pos((size_t)m->nv * 3), nrm((size_t)m->nv * 3);
    std::vector<int32_t> tri((size_t)m->nt * 3);
    const int rc = flof_mesh_download(m, pos.data(), nrm.data(), tri.data());
    if (rc != FLOF_OK) return rc;
    // positions normalised to the unit cube around 0: (p - gs/2) * dx with
    // dx = 1 / max(gs) of the level set's grid
    const int gmax = std::max(m->X, std::max(m->Y, m->Z));
    const float dx = (float)(1.0 / (double)gmax);
    const float half[3] = {(float)((double)(float)m->X * 0.5), (float)((double)(float)m->Y * 0.5),
                           (float)((double)(float)m->Z * 0.5)};
    for (size_t v = 0; v < (size_t)m->nv; ++v)
        for (int d = 0; d < 3; ++d) pos[3 * v + d] = (pos[3 * v + d] - half[d]) * dx;
    gzFile f = gzopen(name, "wb1");
    FLOF_REQUIRE(f, FLOF_ERR_ARG, std::string("writeBobj: unable to open file ") + name);
    const int nv = (int)m->nv, nt = (int)m->nt;
    bool ok = gzwrite(f, &nv, 4) == 4;
    if (ok && nv) ok = gzwrite(f, pos.data(), (unsigned)(pos.size() * 4)) == (int)(pos.size() * 4);
    ok = ok && gzwrite(f, &nv, 4) == 4;
    if (ok && nv) ok = gzwrite(f, nrm.data(), (unsigned)(nrm.size() * 4)) == (int)(nrm.size() * 4);
    ok = ok && gzwrite(f, &nt, 4) == 4;
    if (ok && nt) ok = gzwrite(f, tri.data(), (unsigned)(tri.size() * 4)) == (int)(tri.size() * 4);
    const int cl = gzclose(f);
    FLOF_REQUIRE(ok && cl == Z_OK, FLOF_ERR_ARG, std::string("writeBobj: error writing ") + name);
    API_CATCH(ctx)
}

int flof_mesh_destroy(flof_mesh *m)
{
    if (!m) return FLOF_ERR_ARG;
    if (m->ctx) (void)hipStreamSynchronize(m->ctx->stream);
    for (void *p : {(void *)m->pos, (void *)m->nrm, (void *)m->tris})
        if (p) (void)hipFree(p);
    delete m;
    return FLOF_OK;
}

}  // extern "C"
