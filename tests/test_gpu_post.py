"""GPU parity of the output post-processing of warped 3D frames (through the
C-ABI) against oracle/post.py: bit-exact."""
import numpy as np
import pytest

from oracle import post as P
from ofblend_amd import manta as M

pytestmark = pytest.mark.gpu


def dev(a):
    return M.DeviceArray.from_numpy(np.asarray(a, np.float32))


def test_frame_ops_bit_exact():
    rng = np.random.default_rng(11)
    a = rng.standard_normal((9, 10, 11)).astype(np.float32)
    b = rng.standard_normal((9, 10, 11)).astype(np.float32)
    print("frame ops", a.shape, flush=True)
    d, e = dev(a), dev(b)
    d.join(e)
    assert np.array_equal(d.to_numpy(), P.frame_join(a, b))
    d.multConst(0.3)
    ref = P.frame_mult_const(P.frame_join(a, b), 0.3)
    assert np.array_equal(d.to_numpy(), ref)
    d.addScaled(e, 0.7)
    ref = P.frame_add_scaled(ref, b, 0.7)
    assert np.array_equal(d.to_numpy(), ref)
    d.copyFrom(e)
    assert np.array_equal(d.to_numpy(), b)
    d.setConst(2.5)
    assert np.all(d.to_numpy() == 2.5)
    with pytest.raises(RuntimeError):
        d.join(dev(np.zeros((3, 3, 3), np.float32)))


@pytest.mark.parametrize("shape,it,thresh,bord", [
    ((6, 7, 8), 1, 0.0, 0), ((6, 7, 8), 2, 0.0, 0), ((9, 8, 7), 3, -999.0, 2),
    ((1, 12, 13), 1, 0.0, 0), ((1, 12, 13), 2, -999.0, -1), ((5, 5, 5), 0, 0.0, 0)])
def test_simple_blur_special_bit_exact(shape, it, thresh, bord):
    print("simpleBlurSpecial", shape, it, thresh, bord, flush=True)
    rng = np.random.default_rng(sum(shape) + it)
    a = (rng.standard_normal(shape) * 3).astype(np.float32)
    d = dev(a)
    M.simpleBlurSpecial(d, it, thresh, bord)
    assert np.array_equal(d.to_numpy(), P.simple_blur_special(a, it, thresh, bord))


@pytest.mark.parametrize("shape,w", [((6, 7, 8), 0), ((6, 7, 8), 2), ((1, 6, 7), 1)])
def test_frame_set_bound_bit_exact(shape, w):
    a = np.random.default_rng(5).standard_normal(shape).astype(np.float32)
    d = dev(a)
    d.setBound(0.5, w)
    assert np.array_equal(d.to_numpy(), P.set_bound3(a, 0.5, w))


def test_union_blends_and_temporal_merge():
    """The compositions of scenes/flof.py:737-748 (two-way union blend),
    :759-788 (three-way union blend) and :811-820 (temporal SDF merge and
    level-set border), each against the same composition of the oracle."""
    rng = np.random.default_rng(21)
    shp = (8, 9, 10)
    f = [(rng.standard_normal(shp) * 2).astype(np.float32) for _ in range(3)]
    print("union blends", shp, flush=True)
    # two-way
    w0, wu, w1 = M.blendW(0.4, 0.1)
    dst, other, tmp = dev(f[0]), dev(f[1]), M.DeviceArray(shp)
    tmp.copyFrom(dst)
    tmp.join(other)
    dst.multConst(w0)
    dst.addScaled(tmp, wu)
    dst.addScaled(other, w1)
    u = P.frame_join(f[0], f[1])
    ref = P.frame_add_scaled(P.frame_add_scaled(P.frame_mult_const(f[0], w0), u, wu), f[1], w1)
    assert np.array_equal(dst.to_numpy(), ref)
    # three-way
    ws = M.weights2dim((0.3, 0.45, 0.25))
    d = [dev(x) for x in f]
    acc, t3 = M.DeviceArray(shp), M.DeviceArray(shp)
    acc.setConst(0.)
    racc = np.zeros(shp, np.float32)
    pairs = [(3, (0, 1)), (4, (1, 2)), (5, (0, 2)), (6, (0, 1, 2))]
    for wi, members in pairs:
        if ws[wi] > 1e-03:
            t3.copyFrom(d[members[0]])
            rt = f[members[0]]
            for m in members[1:]:
                t3.join(d[m])
                rt = P.frame_join(rt, f[m])
            acc.addScaled(t3, ws[wi])
            racc = P.frame_add_scaled(racc, rt, ws[wi])
    for k in range(3):
        if ws[k] > 1e-03:
            acc.addScaled(d[k], ws[k])
            racc = P.frame_add_scaled(racc, f[k], ws[k])
    assert np.array_equal(acc.to_numpy(), racc)
    # temporal merge over three frames, then the level-set border
    prev1, prev2, cur = M.DeviceArray(shp), M.DeviceArray(shp), M.DeviceArray(shp)
    r1 = r2 = None
    for k in range(3):
        cur.copyFrom(d[k])
        rc = f[k]
        if k == 0:
            prev2.copyFrom(cur)
            prev1.copyFrom(cur)
            r1 = r2 = rc
        prev2.copyFrom(prev1)
        prev1.copyFrom(cur)
        r2, r1 = r1, rc
        cur.join(prev2)
        M.simpleBlurSpecial(cur, 1, 0.)
        cur.setBound(0.5, 2)
        rc = P.set_bound3(P.simple_blur_special(P.frame_join(rc, r2), 1, 0.0), 0.5, 2)
        assert np.array_equal(cur.to_numpy(), rc), k
