"""The GPU level-set mesher and mesh file against the oracle: nodes,
triangles (the reference's numbering) and the file's vertex normals bit for
bit, on a sphere, a torus, a random smooth field that crosses the grid's
faces, a field with invalid cubes and a warped output frame; the .bobj.gz
file read back by an independent gzip/struct parser of the documented
layout (source/fileio.cpp:121-160)."""
import gzip
import struct

import numpy as np
import pytest

from oracle import oracle as O
from ofblend_amd import manta as M

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _grid(shape):
    Z, Y, X = shape
    return np.meshgrid(*[np.arange(n, dtype=np.float64) + 0.5 for n in (Z, Y, X)],
                       indexing="ij")


def _fields():
    rng = np.random.default_rng(5)
    z, y, x = _grid((22, 21, 20))
    sph = np.sqrt((x - 10.3) ** 2 + (y - 9.7) ** 2 + (z - 11.1) ** 2) - 6.2
    z, y, x = _grid((16, 30, 31))
    q = np.sqrt((x - 15.2) ** 2 + (y - 14.9) ** 2) - 8.0
    tor = np.sqrt(q ** 2 + (z - 8.1) ** 2) - 3.1
    a = rng.standard_normal((13, 17, 70))
    for ax in range(3):
        a = (np.roll(a, 1, ax) + 2 * a + np.roll(a, -1, ax)) / 4
    inv = sph.copy()
    inv[:, :, 9:] = -2000.0
    return {"sphere": sph, "torus": tor, "random": a * 3, "invalid": inv}


@pytest.mark.parametrize("name", ["sphere", "torus", "random", "invalid"])
def test_mesh_bit_exact(name, tmp_path):
    phi = _fields()[name].astype(np.float32)
    say("mesh", name, phi.shape)
    ref = O.create_mesh(phi)
    d = M.DeviceArray.from_numpy(phi)
    mesh = M.Mesh()
    d.createMesh(mesh)
    got = mesh.to_numpy()
    assert mesh.numNodes() == len(ref[0]) and mesh.numTris() == len(ref[2])
    for g, r in zip(got, ref):
        assert np.array_equal(g, r)
    # the file
    path = tmp_path / "m.bobj.gz"
    mesh.save(str(path))
    raw = gzip.open(path, "rb").read()
    nv = struct.unpack_from("<i", raw, 0)[0]
    pos = np.frombuffer(raw, np.float32, 3 * nv, 4).reshape(nv, 3)
    o = 4 + 12 * nv
    assert struct.unpack_from("<i", raw, o)[0] == nv
    nrm = np.frombuffer(raw, np.float32, 3 * nv, o + 4).reshape(nv, 3)
    o += 4 + 12 * nv
    nt = struct.unpack_from("<i", raw, o)[0]
    tri = np.frombuffer(raw, np.int32, 3 * nt, o + 4).reshape(nt, 3)
    assert len(raw) == o + 4 + 12 * nt
    Z, Y, X = phi.shape
    gs = np.array([X, Y, Z], np.float32)
    dx = np.float32(1.0 / max(X, Y, Z))
    want = (ref[0] - (gs.astype(np.float64) * 0.5).astype(np.float32)) * dx
    assert np.array_equal(pos, want.astype(np.float32))
    assert np.array_equal(nrm, ref[1]) and np.array_equal(tri, ref[2])


def test_mesh_of_warped_frame():
    """A frame of the output pipeline: the drop warped by a random
    deformation (loadAdvectTimeSlice_OptRun, which writes only cells 10 or
    more from the border), then meshed."""
    N, T, F = 24, 8, 2
    rng = np.random.default_rng(9)
    frames = O.drop_sdf(N, N, N, T, seed=4, scale=F)
    phi_np = O.place_frames((T, N * F, N * F, N * F), frames)
    s = M.Solver(gridSize=M.vec3(N * F, N * F, N * F), fourthDim=T)
    phi = s.create(M.Grid4Real)
    phi.from_numpy(phi_np)
    sd = M.Solver(gridSize=M.vec3(N, N, N), fourthDim=T)
    dv = (rng.standard_normal((T, N, N, N, 4)) * 0.8).astype(np.float32)
    defo = sd.create(M.Grid4Vec4)
    defo.from_numpy(dv)
    dst = M.DeviceArray.from_numpy(np.zeros((N * F,) * 3, np.float32))
    M.loadAdvectTimeSliceFromDefo(defo, dst, phi, 3.5, 1.0, defoFactor=M.Vec4(2, 2, 2, 1))
    frame = dst.to_numpy()
    mesh = M.Mesh()
    dst.createMesh(mesh)
    got = mesh.to_numpy()
    ref = O.create_mesh(frame)
    say("warped frame mesh", len(ref[0]), "nodes", len(ref[2]), "triangles")
    assert len(ref[2]) > 100
    for g, r in zip(got, ref):
        assert np.array_equal(g, r)


def test_mesh_empty_and_smallest_grid(tmp_path):
    """Edge cases: a level set with no surface gives an empty mesh (and a file
    of three zero counts); the smallest grid (2 x 2 x 2: one cube) with one
    corner inside gives the oracle's single triangle."""
    phi = np.full((9, 8, 7), 3.0, np.float32)
    mesh = M.Mesh()
    M.DeviceArray.from_numpy(phi).createMesh(mesh)
    assert mesh.numNodes() == 0 and mesh.numTris() == 0
    ref = O.create_mesh(phi)
    assert len(ref[0]) == 0 and len(ref[2]) == 0
    path = tmp_path / "e.bobj.gz"
    mesh.save(str(path))
    assert gzip.open(path, "rb").read() == struct.pack("<iii", 0, 0, 0)
    tiny = np.full((2, 2, 2), 1.0, np.float32)
    tiny[0, 0, 0] = -0.6
    mesh2 = M.Mesh()
    M.DeviceArray.from_numpy(tiny).createMesh(mesh2)
    ref2 = O.create_mesh(tiny)
    say("one-cube mesh", len(ref2[0]), "nodes", len(ref2[2]), "triangles")
    assert len(ref2[2]) == 1 and len(ref2[0]) == 3
    for g, r in zip(mesh2.to_numpy(), ref2):
        assert np.array_equal(g, r)
