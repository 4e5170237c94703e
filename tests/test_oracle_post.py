"""The post-processing restatement (oracle/post.py) against closed forms, and
the blend weights of the union blends (host-side glue of the shim) against
the barycentric identities they must satisfy."""
import numpy as np
import pytest

from oracle import post as P
from ofblend_amd import manta as M


def test_blur_constant_field_and_outer_layer():
    a = np.full((5, 6, 7), 3.0, np.float32)
    b1 = P.simple_blur_special(a, 1)
    assert np.all(b1[1:-1, 1:-1, 1:-1] == 3.0)
    # the zeroed temporary's outer layer comes along with the first swap
    outer = np.ones_like(a, bool)
    outer[1:-1, 1:-1, 1:-1] = False
    assert np.all(b1[outer] == 0.0)
    b2 = P.simple_blur_special(a, 2)
    assert np.all(b2[outer] == 3.0)  # the input's outer layer, swapped back
    # a corner-adjacent cell of the second pass sees three zero neighbours
    assert b2[1, 1, 1] == np.float32(9.0 / 6.0)
    assert b2[2, 2, 2] == 3.0


def test_blur_linear_field_threshold_band_and_2d():
    z, y, x = np.meshgrid(np.arange(6), np.arange(7), np.arange(8), indexing="ij")
    f = (x + 2 * y + 3 * z).astype(np.float32)
    b = P.simple_blur_special(f, 1)
    assert np.array_equal(b[1:-1, 1:-1, 1:-1], f[1:-1, 1:-1, 1:-1])
    g = f - 20.0  # cells below the threshold 0 stay
    bg = P.simple_blur_special(g, 1, thresh=0.0)
    inner = g[1:-1, 1:-1, 1:-1]
    assert np.array_equal(bg[1:-1, 1:-1, 1:-1][inner < 0], inner[inner < 0])
    # cells outside the band bord stay as they are
    r = np.random.default_rng(3).standard_normal((7, 8, 9)).astype(np.float32) + 5.0
    br = P.simple_blur_special(r, 1, thresh=-99.0, bord=2)
    assert br[1, 4, 4] == r[1, 4, 4] and br[4, 1, 4] == r[4, 1, 4]
    assert br[3, 3, 3] != r[3, 3, 3]
    # 2D (one z plane): the 4-neighbour mean
    a2 = np.full((1, 5, 6), 2.0, np.float32)
    a2[0, 2, 3] = 6.0
    b2 = P.simple_blur_special(a2, 1, thresh=-1.0)
    assert b2[0, 2, 2] == np.float32((2.0 + 6.0 + 2.0 + 2.0) / 4.0)
    assert b2[0, 2, 3] == 2.0


def test_set_bound3_brute_force():
    a = np.arange(5 * 6 * 7, dtype=np.float32).reshape(5, 6, 7)
    for w in (0, 1, 2):
        b = P.set_bound3(a, -1.0, w)
        m = np.zeros(a.shape, bool)
        m[:w + 1] = m[-(w + 1):] = True
        m[:, :w + 1] = m[:, -(w + 1):] = True
        m[:, :, :w + 1] = m[:, :, -(w + 1):] = True
        assert np.array_equal(b[m], np.full(m.sum(), -1.0, np.float32))
        assert np.array_equal(b[~m], a[~m])
    a2 = np.zeros((1, 5, 5), np.float32)
    b2 = P.set_bound3(a2, 1.0, 0)
    assert b2[0, 0, 2] == 1.0 and b2[0, 2, 2] == 0.0  # z is not a border in 2D


def test_join_is_the_minimum():
    a = np.array([1.0, -2.0, 3.0, np.nan], np.float32)
    b = np.array([0.5, -1.0, np.nan, 1.0], np.float32)
    j = P.frame_join(a, b)
    assert j[0] == 0.5 and j[1] == -2.0 and j[2] == 3.0 and np.isnan(j[3])


def test_blendW_closed_forms():
    assert M.blendW(0.0, 0.1) == (1.0, 0.0, 0.0)
    assert M.blendW(1.0, 0.1) == (0.0, 0.0, 1.0)
    w0, wu, w1 = M.blendW(0.5, 0.1)
    assert abs(w0 - 0.1 / 1.1) < 1e-15 and abs(w1 - 0.1 / 1.1) < 1e-15
    for t in np.linspace(0, 1, 23):
        w0, wu, w1 = M.blendW(float(t), 0.1)
        assert min(w0, wu, w1) >= -1e-15 and abs(w0 + wu + w1 - 1.0) < 1e-15


VERTS = [(1, 0, 0), (0, 1, 0), (0, 0, 1), (.5, .5, 0), (0, .5, .5), (.5, 0, .5),
         (1. / 3, 1. / 3, 1. / 3)]


@pytest.mark.parametrize("seed", range(4))
def test_weights2dim_barycentric(seed):
    """The seven weights are barycentric coordinates of uvw in the
    subdivided triangle: non-negative, summing to 1 and reproducing uvw as
    the weighted sum of their vertices (pure, pairwise and triple unions)."""
    rng = np.random.default_rng(seed)
    for _ in range(200):
        u, v = rng.random(2)
        if u + v > 1:
            u, v = 1 - u, 1 - v
        w = 1 - u - v
        ws = M.weights2dim((u, v, w))
        assert min(ws) > -1e-6 and abs(sum(ws) - 1.0) < 1e-6
        p = np.sum([wk * np.array(vk) for wk, vk in zip(ws, VERTS)], axis=0)
        assert np.abs(p - np.float32([u, v, w])).max() < 1e-5
    for k, vk in enumerate(VERTS):
        ws = M.weights2dim(vk)
        assert abs(ws[k] - 1.0) < 1e-6
