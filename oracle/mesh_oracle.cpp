// mesh_oracle.cpp -- CPU restatement of the level-set mesher and the mesh
// file's vertex normals (TEST INFRASTRUCTURE ONLY: used by tests/ as the
// checker of the GPU mesher, never shipped).
//
// LevelsetGrid::createMesh (source/levelset.cpp:266-365): cubes visited
// with x outermost, then y, then z; a cube is skipped if a corner value is
// at or below the fast-marching invalid time (-1000); its case has bit c set
// when -phi(corner c) < 1e-4; each active edge's vertex is created by the
// first cube that visits it (slot order 0..11 within the cube) at
// p1 + (p2 - p1) mu + 0.5 with mu = (iso - v1) / (v2 - v1) on the negated
// values; triangles in table order.  Mesh::computeVertexNormals
// (source/mesh.cpp:460-479), which writeBobjFile runs before it writes the
// normals (source/fileio.cpp:121-160): per triangle, in triangle order,
// corner c gets cross(p0 - p1, p1 - p2) / (|edge before c|^2 |edge after
// c|^2), then every normal is normalised.  FP32 throughout, with the
// double factors where the reference computes in double.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../ofblend_amd/csrc/flof_mc_tables.h"

namespace {

using flof::mc::kMcCorner;
using flof::mc::kMcEdge;
using flof::mc::kMcTris;

struct V3 {
    float x, y, z;
};

inline int hexv(char c) { return c <= '9' ? c - '0' : c - 'a' + 10; }

// normalize (util/vectorbase.h:401-415) of a float vector
inline V3 normalize3(V3 v)
{
    const float l = v.x * v.x + v.y * v.y + v.z * v.z;
    const float e2 = 1e-6f * 1e-6f;
    if (std::fabs((double)l - 1.) < e2) return v;
    if (l > e2) {
        const float nrm = (float)std::sqrt((double)l);
        const float s = (float)(1. / (double)nrm);
        return V3{v.x * s, v.y * s, v.z * s};
    }
    return V3{0.f, 0.f, 0.f};
}

// computeVertexNormals: per triangle, in triangle order, corner c gets
// cross(p0 - p1, p1 - p2) / (|edge before c|^2 |edge after c|^2), then every
// normal is normalised
std::vector<V3> vertex_normals(const std::vector<V3> &pos, const int32_t *tris, int64_t nt)
{
    std::vector<V3> nrm(pos.size(), V3{0.f, 0.f, 0.f});
    for (int64_t t = 0; t < nt; ++t) {
        const int32_t *c = &tris[3 * t];
        const V3 p0 = pos[c[0]], p1 = pos[c[1]], p2 = pos[c[2]];
        const V3 n0{p0.x - p1.x, p0.y - p1.y, p0.z - p1.z};
        const V3 n1{p1.x - p2.x, p1.y - p2.y, p1.z - p2.z};
        const V3 n2{p2.x - p0.x, p2.y - p0.y, p2.z - p0.z};
        const float l0 = n0.x * n0.x + n0.y * n0.y + n0.z * n0.z;
        const float l1 = n1.x * n1.x + n1.y * n1.y + n1.z * n1.z;
        const float l2 = n2.x * n2.x + n2.y * n2.y + n2.z * n2.z;
        const V3 nm{(n0.y * n1.z) - (n0.z * n1.y), (n0.z * n1.x) - (n0.x * n1.z),
                    (n0.x * n1.y) - (n0.y * n1.x)};
        const double f[3] = {1.0 / (double)(l0 * l2), 1.0 / (double)(l0 * l1),
                             1.0 / (double)(l1 * l2)};
        for (int q = 0; q < 3; ++q) {
            V3 &a = nrm[c[q]];
            a.x = a.x + (float)((double)nm.x * f[q]);
            a.y = a.y + (float)((double)nm.y * f[q]);
            a.z = a.z + (float)((double)nm.z * f[q]);
        }
    }
    for (auto &n : nrm) n = normalize3(n);
    return nrm;
}

}  // namespace

extern "C" int orc_create_mesh(const float *phi, int X, int Y, int Z, float *pos_out,
                               float *nrm_out, int32_t *tri_out, int64_t maxv, int64_t maxt,
                               int64_t *nv_out, int64_t *nt_out)
{
    const float invalidTime = -1000.f;
    const float iso = 1e-4f;
    auto at = [&](int i, int j, int k) { return phi[(int64_t)i + (int64_t)X * (j + (int64_t)Y * k)]; };
    const size_t ng = (size_t)X * Y * Z;
    std::vector<int> ev[3];
    for (auto &e : ev) e.assign(ng, 0);
    std::vector<V3> pos;
    std::vector<int32_t> tris;
    for (int i = 0; i < X - 1; i++)
        for (int j = 0; j < Y - 1; j++)
            for (int k = 0; k < Z - 1; k++) {
                float value[8];
                bool skip = false;
                int cs = 0;
                for (int l = 0; l < 8; l++) {
                    value[l] = at(i + kMcCorner[l][0], j + kMcCorner[l][1], k + kMcCorner[l][2]);
                    value[l] *= -1;
                    if (-value[l] <= invalidTime) skip = true;
                    if (value[l] < iso) cs |= 1 << l;
                }
                const int emask = flof::mc::mc_edge_mask(cs);
                if (skip || emask == 0) continue;
                int tri_idx[12];
                for (int e = 0; e < 12; e++) {
                    if (!(emask & (1 << e))) continue;
                    const int ca = kMcEdge[e][0], cb = kMcEdge[e][1];
                    int dir = 0;
                    for (int d = 0; d < 3; ++d)
                        if (kMcCorner[ca][d] != kMcCorner[cb][d]) dir = d;
                    const int gi = i + std::min(kMcCorner[ca][0], kMcCorner[cb][0]);
                    const int gj = j + std::min(kMcCorner[ca][1], kMcCorner[cb][1]);
                    const int gk = k + std::min(kMcCorner[ca][2], kMcCorner[cb][2]);
                    int &slot = ev[dir][(size_t)gi + (size_t)X * (gj + (size_t)Y * gk)];
                    if (slot == 0) {
                        const V3 p1{(float)(i + kMcCorner[ca][0]), (float)(j + kMcCorner[ca][1]),
                                    (float)(k + kMcCorner[ca][2])};
                        const V3 p2{(float)(i + kMcCorner[cb][0]), (float)(j + kMcCorner[cb][1]),
                                    (float)(k + kMcCorner[cb][2])};
                        const float v1 = value[ca], v2 = value[cb];
                        const float mu = (iso - v1) / (v2 - v1);
                        V3 p;
                        p.x = (p1.x + (p2.x - p1.x) * mu) + 0.5f;
                        p.y = (p1.y + (p2.y - p1.y) * mu) + 0.5f;
                        p.z = (p1.z + (p2.z - p1.z) * mu) + 0.5f;
                        pos.push_back(p);
                        slot = (int)pos.size();  // node index + 1
                    }
                    tri_idx[e] = slot;
                }
                const char *tt = kMcTris[cs];
                for (int q = 0; tt[q]; q += 3)
                    for (int c = 0; c < 3; ++c) tris.push_back(tri_idx[hexv(tt[q + c])] - 1);
            }
    const int64_t nv = (int64_t)pos.size(), nt = (int64_t)tris.size() / 3;
    *nv_out = nv;
    *nt_out = nt;
    if (nv > maxv || nt > maxt) return 1;
    const std::vector<V3> nrm = vertex_normals(pos, tris.data(), nt);
    for (int64_t v = 0; v < nv; ++v) {
        pos_out[3 * v + 0] = pos[v].x;
        pos_out[3 * v + 1] = pos[v].y;
        pos_out[3 * v + 2] = pos[v].z;
        nrm_out[3 * v + 0] = nrm[v].x;
        nrm_out[3 * v + 1] = nrm[v].y;
        nrm_out[3 * v + 2] = nrm[v].z;
    }
    std::memcpy(tri_out, tris.data(), tris.size() * sizeof(int32_t));
    return 0;
}

// Mesh::computeVertexNormals on given nodes and triangles (the routine the
// mesher uses, exposed so that its weighting can be checked on its own).
extern "C" void orc_vertex_normals(const float *pos, int64_t nv, const int32_t *tris,
                                   int64_t nt, float *nrm_out)
{
    std::vector<V3> p(nv);
    for (int64_t v = 0; v < nv; ++v) p[v] = V3{pos[3 * v], pos[3 * v + 1], pos[3 * v + 2]};
    const std::vector<V3> n = vertex_normals(p, tris, nt);
    for (int64_t v = 0; v < nv; ++v) {
        nrm_out[3 * v + 0] = n[v].x;
        nrm_out[3 * v + 1] = n[v].y;
        nrm_out[3 * v + 2] = n[v].z;
    }
}
