"""Pins the oracle's level-set mesher (LevelsetGrid::createMesh,
levelset.cpp:266-365) and the mesh file's vertex normals
(Mesh::computeVertexNormals, mesh.cpp:460-479) without a GPU, by properties
the marching-cubes construction fixes:

* a closed surface gives a closed, consistently oriented 2-manifold with the
  Euler characteristic of its topology (2 for a sphere, 0 for a torus) --
  a wrong triangle in any case table entry breaks this on random fields;
* every vertex lies on a grid edge (cell centres at +0.5) where the linear
  interpolation of -phi equals the iso value 1e-4, one vertex per active
  edge;
* vertices are numbered in the order of the first cube (x outermost, then
  y, then z) that contains their edge;
* cubes with a corner at or below -1000 (invalid) produce nothing;
* the file normals of a sphere point outwards, close to radial.
"""
import numpy as np
import pytest

from oracle import oracle as O


def grid(shape):
    Z, Y, X = shape
    return np.meshgrid(*[np.arange(n, dtype=np.float64) + 0.5 for n in (Z, Y, X)],
                       indexing="ij")


def sphere(shape, c, r):
    z, y, x = grid(shape)
    return (np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) - r).astype(np.float32)


def torus(shape, c, R, r):
    z, y, x = grid(shape)
    q = np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2) - R
    return (np.sqrt(q ** 2 + (z - c[2]) ** 2) - r).astype(np.float32)


def edges_of(tri):
    d = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]])
    return d


def check_closed(tri, nv):
    d = edges_of(tri)
    # consistent orientation: every directed edge once, its reverse once
    keys = d[:, 0].astype(np.int64) * nv + d[:, 1]
    assert len(np.unique(keys)) == len(keys)
    rev = d[:, 1].astype(np.int64) * nv + d[:, 0]
    assert np.array_equal(np.sort(keys), np.sort(rev))
    und = np.unique(np.sort(d, axis=1).astype(np.int64) @ np.array([nv, 1]))
    return nv - len(und) + len(tri)  # Euler characteristic


def test_sphere_closed_manifold_and_outward_normals():
    phi = sphere((22, 21, 20), (10.3, 9.7, 11.1), 6.2)
    p, n, t = O.create_mesh(phi)
    assert len(t) > 500
    assert check_closed(t, len(p)) == 2
    c = np.array([10.3, 9.7, 11.1])
    radial = (p - c) / np.linalg.norm(p - c, axis=1)[:, None]
    cos = (n * radial).sum(1)
    assert cos.min() > 0.95
    assert np.abs(np.linalg.norm(n, axis=1) - 1).max() < 1e-5


def test_torus_euler_zero():
    phi = torus((16, 30, 30), (15.2, 14.9, 8.1), 8.0, 3.1)
    p, n, t = O.create_mesh(phi)
    assert check_closed(t, len(p)) == 0


def _edge_vertex_table(phi):
    """Active grid edges (dir, i, j, k) and the position of their vertex
    where the linear interpolation of -phi equals 1e-4."""
    v = -phi.astype(np.float64)
    Z, Y, X = phi.shape
    out = []
    for d, (dz, dy, dx) in enumerate(((0, 0, 1), (0, 1, 0), (1, 0, 0))):
        a = v[:Z - dz, :Y - dy, :X - dx]
        b = v[dz:, dy:, dx:]
        act = (a < 1e-4) != (b < 1e-4)
        for (k, j, i) in zip(*np.nonzero(act)):
            va, vb = a[k, j, i], b[k, j, i]
            mu = (1e-4 - va) / (vb - va)
            out.append((d, i, j, k, np.array([i + dx * mu, j + dy * mu, k + dz * mu]) + 0.5))
    return out


@pytest.mark.parametrize("seed", [3, 4])
def test_vertices_on_active_edges_in_cube_order(seed):
    rng = np.random.default_rng(seed)
    shape = (9, 10, 11)
    a = rng.standard_normal(shape)
    for ax in range(3):
        a = (np.roll(a, 1, ax) + 2 * a + np.roll(a, -1, ax)) / 4
    phi = (a * 3).astype(np.float32)
    p, n, t = O.create_mesh(phi)
    ref = _edge_vertex_table(phi)
    assert len(p) == len(ref)
    # every vertex sits at the interpolated crossing of one active edge
    keyed = {}
    for d, i, j, k, q in ref:
        keyed[(d, i, j, k)] = q
    Z, Y, X = shape
    owners = []
    for v in p:
        best = min(keyed, key=lambda e: np.abs(keyed[e] - v).sum())
        assert np.abs(keyed[best] - v).max() < 1e-5
        d, i, j, k = best
        # the first cube (x outermost) containing the edge
        cand = []
        for di in ((0,) if d == 0 else (0, -1)):
            for dj in ((0,) if d == 1 else (0, -1)):
                for dk in ((0,) if d == 2 else (0, -1)):
                    ci, cj, ck = i + di, j + dj, k + dk
                    if 0 <= ci < X - 1 and 0 <= cj < Y - 1 and 0 <= ck < Z - 1:
                        cand.append((ci * (Y - 1) + cj) * (Z - 1) + ck)
        owners.append(min(cand))
    assert np.all(np.diff(owners) >= 0)
    # triangles away from the grid's faces are shared by two triangles
    d = np.sort(edges_of(t), axis=1)
    u, cnt = np.unique(d, axis=0, return_counts=True)
    assert cnt.max() <= 2


def test_invalid_cubes_are_skipped():
    phi = sphere((16, 16, 16), (8.2, 7.9, 8.1), 5.0)
    cut = phi.copy()
    cut[:, :, 9:] = -2000.0  # invalid: every cube touching x >= 9 is skipped
    p, n, t = O.create_mesh(cut)
    p_full, _, t_full = O.create_mesh(phi)
    assert 0 < len(t) < len(t_full)
    assert p[:, 0].max() <= 9.0 + 1e-6  # no vertex beyond the last valid cube


def test_vertex_normal_weights_are_exact_on_a_sphere():
    """Mesh::computeVertexNormals weights each triangle's normal at a corner by
    sin(angle) / (|e1| |e2|) of the two edges there (mesh.cpp:460-479), the
    weighting for which Max (1999, "Weights for computing vertex normals from
    facet normals") shows the vertex normal is exact when the vertex and its
    neighbours lie on one sphere.  An irregular closed fan on a sphere: the
    oracle's normal is the radial direction to float precision, while the
    area- and angle-weighted normals of the same fan are measurably off (so
    the check tells the weightings apart)."""
    rng = np.random.default_rng(12)
    c = np.array([0.3, -0.2, 0.1])
    r = 1.7
    k = 7
    az = np.sort(rng.uniform(0, 2 * np.pi, k))
    polar = rng.uniform(0.15, 0.6, k)
    # vertex at the sphere's top, neighbours around it on the sphere
    v = c + np.array([0, 0, r])
    nb = c + r * np.stack([np.sin(polar) * np.cos(az), np.sin(polar) * np.sin(az),
                           np.cos(polar)], 1)
    pos = np.vstack([v, nb]).astype(np.float32)
    tris = np.array([[0, 1 + i, 1 + (i + 1) % k] for i in range(k)], np.int32)
    n = O.vertex_normals(pos, tris)[0].astype(np.float64)
    radial = (pos[0] - c) / np.linalg.norm(pos[0] - c)
    assert np.dot(n, radial) > 0  # counter-clockwise fan seen from outside
    assert np.abs(n - radial).max() < 2e-6
    # the other common weightings miss the radial direction on this fan
    p = pos.astype(np.float64)
    area, angle = np.zeros(3), np.zeros(3)
    for a, b, d in tris:
        e1, e2 = p[b] - p[a], p[d] - p[a]
        cr = np.cross(e1, e2)
        area += cr
        ang = np.arccos(np.dot(e1, e2) / np.linalg.norm(e1) / np.linalg.norm(e2))
        angle += ang * cr / np.linalg.norm(cr)
    for w in (area, angle):
        w = w / np.linalg.norm(w)
        assert np.abs(w - radial).max() > 1e-3
