"""Pins the oracle's sampling, advection, resampling and boundary functions
against library routines and closed forms (no GPU).

interp4 / interp3 (util/vector4d.h:422-516, util/interpol.h:57-116) are
multilinear interpolation of cell-centred data with edge clamping, i.e.
scipy.ndimage.map_coordinates(order=1, mode='nearest') at p - 1/2.
"""
import numpy as np
import pytest
from scipy.ndimage import map_coordinates

from oracle import oracle as O

RNG = np.random.default_rng(7)


def mc4(data, pos):
    """Reference sampler: pos (n,4) as (x,y,z,t), data (T,Z,Y,X)."""
    coords = np.stack([pos[:, 3] - 0.5, pos[:, 2] - 0.5, pos[:, 1] - 0.5,
                       pos[:, 0] - 0.5])
    return map_coordinates(data, coords, order=1, mode="nearest")


def mc3(data, pos):
    coords = np.stack([pos[:, 2] - 0.5, pos[:, 1] - 0.5, pos[:, 0] - 0.5])
    return map_coordinates(data, coords, order=1, mode="nearest")


def test_sample4_matches_map_coordinates():
    T, Z, Y, X = 5, 6, 7, 8
    d = RNG.standard_normal((T, Z, Y, X))
    hi = np.array([X, Y, Z, T], float)
    # positions inside and up to 1.5 cells outside the grid
    pos = RNG.uniform(-1.5, 1.5, (4000, 4)) + RNG.uniform(0, 1, (4000, 4)) * hi
    got = O.sample4(d, pos)
    np.testing.assert_allclose(got, mc4(d, pos), rtol=0, atol=1e-12)
    got32 = O.sample4(d.astype(np.float32), pos.astype(np.float32))
    np.testing.assert_allclose(got32, mc4(d, pos.astype(np.float32).astype(float)),
                               rtol=0, atol=2e-5)


def test_sample4_exact_on_multilinear_field():
    T, Z, Y, X = 4, 5, 6, 7
    t, z, y, x = np.meshgrid(np.arange(T) + .5, np.arange(Z) + .5,
                             np.arange(Y) + .5, np.arange(X) + .5, indexing="ij")

    def f(x, y, z, t):
        return 0.3 + 1.1 * x - 0.7 * y + 0.2 * z + 0.9 * t + 0.05 * x * y * z * t

    d = f(x, y, z, t)
    pos = np.stack(
        [RNG.uniform(0.5, X - 0.5, 500), RNG.uniform(0.5, Y - 0.5, 500),
         RNG.uniform(0.5, Z - 0.5, 500), RNG.uniform(0.5, T - 0.5, 500)], 1)
    np.testing.assert_allclose(O.sample4(d, pos), f(*pos.T), rtol=1e-12, atol=1e-12)


def test_sample3_matches_map_coordinates():
    d = RNG.standard_normal((6, 7, 8))
    pos = RNG.uniform(-1.5, 9.5, (3000, 3))
    np.testing.assert_allclose(O.sample3(d, pos), mc3(d, pos), atol=1e-12)


def _cell_grid(T, Z, Y, X):
    t, z, y, x = np.meshgrid(np.arange(T), np.arange(Z), np.arange(Y),
                             np.arange(X), indexing="ij")
    return x, y, z, t


def test_advect_zero_flow_is_identity_with_zero_layer():
    a = RNG.standard_normal((6, 7, 8, 9))
    v = np.zeros(a.shape + (4,))
    out = O.advect4d(v, a)
    inner = (slice(1, -1),) * 4
    np.testing.assert_array_equal(out[inner], a[inner])
    mask = np.ones(a.shape, bool)
    mask[inner] = False
    assert (out[mask] == 0).all()


def test_advect_matches_backtrace():
    T, Z, Y, X = 6, 7, 8, 9
    a = RNG.standard_normal((T, Z, Y, X))
    v = RNG.standard_normal((T, Z, Y, X, 4)) * 1.7
    out = O.advect4d(v, a, dt=1.0, dtFac=0.8)
    x, y, z, t = _cell_grid(T, Z, Y, X)
    pos = np.stack([x + .5 - 0.8 * v[..., 0], y + .5 - 0.8 * v[..., 1],
                    z + .5 - 0.8 * v[..., 2], t + .5 - 0.8 * v[..., 3]], -1)
    ref = mc4(a, pos.reshape(-1, 4)).reshape(a.shape)
    inner = (slice(1, -1),) * 4
    np.testing.assert_allclose(out[inner], ref[inner], atol=1e-12)
    # Vec4 grids advect componentwise
    av = RNG.standard_normal((T, Z, Y, X, 4))
    outv = O.advect4d(v, av)
    for c in range(4):
        np.testing.assert_allclose(outv[..., c], O.advect4d(v, av[..., c].copy()),
                                   atol=1e-13)


def test_advect_integer_shift():
    T, Z, Y, X = 6, 7, 8, 9
    a = RNG.standard_normal((T, Z, Y, X))
    v = np.zeros(a.shape + (4,))
    v[..., 0] = 1.0   # x shift by one cell
    v[..., 3] = -1.0  # t shift by minus one
    out = O.advect4d(v, a)
    # out(i,..,t) = a(i-1, .., t+1) where inside
    np.testing.assert_array_equal(out[1:-2, 1:-1, 1:-1, 1:-1],
                                  a[2:-1, 1:-1, 1:-1, 0:-2])


def test_advect_cfl_substeps():
    """cfl below the max speed splits into int(|v|max/cfl)+1 substeps of
    dt/steps, each one advect4d (optflow4d.cpp:1309-1321)."""
    a = RNG.standard_normal((6, 6, 6, 6))
    v = RNG.standard_normal((6, 6, 6, 6, 4))
    vmax = np.sqrt((v ** 2).sum(-1)).max()
    steps = int(vmax / 0.5) + 1
    ref = a.copy()
    for _ in range(steps):
        ref = O.advect4d(v, ref, dt=1.0 / steps)
    np.testing.assert_allclose(O.advect_cfl4d(v, a, cfl=0.5), ref, atol=1e-14)


def _resample_ref(src, shape, fac, off):
    T, Z, Y, X = shape
    x, y, z, t = _cell_grid(T, Z, Y, X)
    pos = np.stack([x * fac[0] + off[0], y * fac[1] + off[1],
                    z * fac[2] + off[2], t * fac[3] + off[3]], -1)
    return mc4(src, pos.reshape(-1, 4)).reshape(shape)


def test_interpolate_grid4d_matches_map_coordinates():
    src = RNG.standard_normal((6, 7, 8, 9))
    for shape, offset, scale in [((3, 4, 5, 6), (0, 0, 0, 0), (1, 1, 1, 1)),
                                 ((10, 11, 12, 13), (0.3, -0.5, 0.2, 0.1),
                                  (0.8, 1.2, 1.0, 0.9))]:
        got = O.interpolate_grid4d(shape, src, offset=offset, scale=scale)
        # gridFactor4d: f = (src/dst)/scale, off = -offset*f + f/2
        sdims = np.array([9, 8, 7, 6], float)
        ddims = np.array(shape[::-1], float)
        f = sdims / ddims / np.array(scale, float)
        off = -np.array(offset, float) * f + f * 0.5
        np.testing.assert_allclose(got, _resample_ref(src, shape, f, off), atol=1e-12)


def test_restriction_is_16_cell_mean_and_prolongation_weights():
    src = RNG.standard_normal((4, 6, 8, 10))
    half = O.pyramid_resample((2, 3, 4, 5), src)
    mean = src.reshape(2, 2, 3, 2, 4, 2, 5, 2).mean(axis=(1, 3, 5, 7))
    np.testing.assert_allclose(half, mean, atol=1e-14)
    # prolongation: 1D weights (3/4, 1/4) towards the nearer coarse cell,
    # clamped at the ends
    c = RNG.standard_normal((3, 3, 3, 4))
    up = O.pyramid_resample((6, 6, 6, 8), c)
    x = np.arange(8)
    q = x * 0.5 + 0.25 - 0.5  # coarse coordinate of the fine cell centre
    i0 = np.clip(np.floor(q).astype(int), 0, 2)
    w1 = np.clip(q - i0, 0, 1)
    w1[q < 0] = 0
    ref_x = c[..., i0] * (1 - w1) + c[..., np.minimum(i0 + 1, 3)] * w1
    line = up[0, 0, 0]
    # along x only, at a fine (t,z,y) whose coarse samples coincide (0,0,0)
    np.testing.assert_allclose(line, ref_x[0, 0, 0], atol=1e-13)
    assert set(np.round(w1[1:-1], 12)) <= {0.25, 0.75}


def test_interpol4d_chain_reference_test():
    """tools/tests/test_0042_interpol4d.py at a reduced size (8->16->32->16->8):
    setRegion4d box resampled up and down; each step equals the separable
    linear resampling of the previous grid."""
    n = 8
    a = np.zeros((n, n, n, n))
    lo, hi = n * 0.3, n * 0.7
    idx = np.arange(n)
    m = (idx >= lo) & (idx <= hi)
    a[np.ix_(m, m, m, m)] = 1.0
    cur = a
    for s in (16, 32, 16, 8):
        nxt = O.interpolate_grid4d((s, s, s, s), cur)
        f = cur.shape[0] / s
        ref = _resample_ref(cur, (s, s, s, s), (f,) * 4, (f / 2,) * 4)
        np.testing.assert_allclose(nxt, ref, atol=1e-12)
        cur = nxt
    # the reference test's threshold for the single-precision chain
    cur32 = a.astype(np.float32)
    for s in (16, 32, 16, 8):
        cur32 = O.interpolate_grid4d((s, s, s, s), cur32)
    assert np.abs(cur32 - cur).max() <= 1e-5


def test_set_bound_and_neumann():
    a = RNG.standard_normal((6, 7, 8, 9))
    T, Z, Y, X = a.shape
    x, y, z, t = _cell_grid(T, Z, Y, X)
    for w in (0, 1):
        band = ((x <= w) | (x >= X - 1 - w) | (y <= w) | (y >= Y - 1 - w) |
                (z <= w) | (z >= Z - 1 - w) | (t <= w) | (t >= T - 1 - w))
        ref = a.copy()
        ref[band] = 3.5
        np.testing.assert_array_equal(O.set_bound(a, 3.5, w), ref)
        src = (np.clip(t, w + 1, T - 2 - w), np.clip(z, w + 1, Z - 2 - w),
               np.clip(y, w + 1, Y - 2 - w), np.clip(x, w + 1, X - 2 - w))
        ref = a.copy()
        ref[band] = a[src][band]
        np.testing.assert_array_equal(O.set_bound_neumann(a, w), ref)
