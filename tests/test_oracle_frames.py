"""Pins the oracle's frame placement (loadPlaceGrid4d), frame warp
(loadAdvectTimeSlice[_OptRun]) and the seeded drop generator (no GPU)."""
import numpy as np
import pytest
from scipy.ndimage import map_coordinates

from oracle import oracle as O

RNG = np.random.default_rng(5)


def splitmix64(state):
    state = (state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return state, z ^ (z >> 31)


def drops_numpy(X, Y, Z, T, seed, nspheres, rmin, rmax, clamp, scale=1):
    s = seed
    dr = []
    for _ in range(nspheres):
        vals = []
        for _ in range(3 if rmax <= rmin else 4):
            s, z = splitmix64(s)
            vals.append((z >> 11) * 2.0 ** -53)
        cx = X * (0.3 + 0.4 * vals[0])
        y0 = Y * (0.6 + 0.2 * vals[1])
        cz = Z * (0.3 + 0.4 * vals[2])
        r = X * (rmin + (rmax - rmin) * vals[3]) if rmax > rmin else X * rmin
        dr.append((cx, y0, cz, r))
    F = scale
    k, j, i = np.meshgrid(np.arange(Z * F) + .5, np.arange(Y * F) + .5,
                          np.arange(X * F) + .5, indexing="ij")
    out = np.empty((T, Z * F, Y * F, X * F))
    for t in range(T):
        sv = t / (T - 1) if T > 1 else 0.0
        v = j - 0.2 * Y * F
        for (cx, y0, cz, r) in dr:
            cy = y0 - 0.5 * (0.4 * Y) * sv * sv
            v = np.minimum(v, np.sqrt((i - F * cx) ** 2 + (j - F * cy) ** 2 +
                                      (k - F * cz) ** 2) - F * r)
        out[t] = np.clip(v, -clamp * F, clamp * F)
    return out


def test_drop_generator_closed_form():
    for args in [(16, 16, 16, 8, 1, 1, 0.15, 0.15, 4.0, 1),
                 (12, 12, 12, 6, 9, 2, 0.08, 0.18, 4.0, 1),
                 (8, 8, 8, 5, 1, 1, 0.15, 0.15, 4.0, 2)]:
        X, Y, Z, T, seed, ns, rmin, rmax, cl, sc = args
        got = O.drop_sdf(X, Y, Z, T, seed=seed, nspheres=ns, rmin=rmin, rmax=rmax,
                         clampCells=cl, scale=sc, dtype=np.float64)
        ref = drops_numpy(X, Y, Z, T, seed, ns, rmin, rmax, cl, sc)
        np.testing.assert_allclose(got, ref, atol=1e-12)
    # different seeds give different drops; the pool plane is the same
    a = O.drop_sdf(16, 16, 16, 4, seed=1)
    b = O.drop_sdf(16, 16, 16, 4, seed=2)
    assert np.abs(a - b).max() > 1.0
    assert np.array_equal(a[:, :, 0, :], b[:, :, 0, :])


def test_place_frames_one_to_one_is_copy():
    T, N = 6, 10
    frames = RNG.standard_normal((T, N, N, N))
    phi = O.place_frames((T, N, N, N), frames)
    inner = (slice(None),) + (slice(1, -1),) * 3
    np.testing.assert_array_equal(phi[inner], frames[inner])
    mask = np.ones(phi.shape, bool)
    mask[inner] = False
    assert not phi[mask].any()


def test_place_frames_tent_weights():
    """Half as many frames as slices: duration 2, slice s gets weights
    1 - |s - 2f|/2 from the frames around it, normalised on 4D-interior
    slices only (optflow4d.cpp:1544-1594)."""
    T, N, F = 8, 6, 4
    frames = RNG.standard_normal((F, N, N, N))
    phi = O.place_frames((T, N, N, N), frames, fileIdxStart=0, fileIdxEnd=F)
    inner = (slice(1, -1),) * 3
    ref = np.zeros((T, N, N, N))
    wsum = np.zeros(T)
    for f in range(F):
        for s in range(T):
            w = 1.0 - abs(s - 2.0 * f) / 2.0
            if w >= 1e-10:
                ref[s][inner] += w * frames[f][inner]
                wsum[s] += w
    for s in range(1, T - 1):
        ref[s] /= wsum[s]
    np.testing.assert_allclose(phi, ref, atol=1e-12)


def mc4(data, pos):
    coords = np.stack([pos[..., 3] - 0.5, pos[..., 2] - 0.5, pos[..., 1] - 0.5,
                       pos[..., 0] - 0.5])
    return map_coordinates(data, coords, order=1, mode="nearest")


def test_warp_slice_zero_and_constant_defo():
    N, T, F = 8, 6, 3   # defo 8^3 x 6, output frames 24^3
    phi = RNG.standard_normal((T, N * F, N * F, N * F))
    defo = np.zeros((T, N, N, N, 4))
    dst0 = RNG.standard_normal((N * F,) * 3)
    k, j, i = np.meshgrid(*[np.arange(N * F)] * 3, indexing="ij")
    band = (slice(10, -10),) * 3
    for time in (1.0, 2.5):
        out = O.warp_slice(defo, phi, dst0, time, blendAlpha=1.0, zeroVel=False)
        pos = np.stack([i + .5, j + .5, k + .5, np.full(i.shape, time + .5)], -1)
        np.testing.assert_allclose(out[band], mc4(phi, pos)[band], atol=1e-12)
        out_m = out.copy()
        out_m[band] = dst0[band]
        np.testing.assert_array_equal(out_m, dst0)  # outside the band: untouched
    v = np.array([0.4, -0.3, 0.2, 0.5])
    defo[:] = v
    fac = np.array([F, F, F, 1.0])
    out = O.warp_slice(defo, phi, dst0, 2.0, blendAlpha=0.7,
                       defoFactor=tuple(fac))
    pos = np.stack([i + .5, j + .5, k + .5, np.full(i.shape, 2.5)], -1) - 0.7 * v * fac
    np.testing.assert_allclose(out[band], mc4(phi, pos)[band], atol=1e-12)
    # the non-cached loadAdvectTimeSlice writes the bnd=1 interior
    out1 = O.warp_slice(defo, phi, dst0, 2.0, blendAlpha=0.7, defoFactor=tuple(fac),
                        optimized=False)
    b1 = (slice(1, -1),) * 3
    np.testing.assert_allclose(out1[b1], mc4(phi, pos)[b1], atol=1e-12)


def test_warp_slice_time_blend():
    """The deformation is blended linearly between slices t and t+1 of
    srcTime = time*sf_t + off_t - 1/2 (optflow4d.cpp:1980-2013)."""
    N, T = 12, 6
    phi = RNG.standard_normal((T, N, N, N))
    defo = np.zeros((T, N, N, N, 4))
    for t in range(T):
        defo[t, ..., 0] = 0.1 * t
    dst0 = np.zeros((N, N, N))
    # defo time axis half as long as phi's: sf_t = 0.5, off_t = 0.25
    phi2 = RNG.standard_normal((2 * T, N, N, N))
    time = 5.0
    out = O.warp_slice(defo, phi2, dst0, time, blendAlpha=1.0)
    src = time * 0.5 + 0.25 - 0.5
    t0 = int(src)
    tw = src - t0
    vx = (1 - tw) * 0.1 * t0 + tw * 0.1 * (t0 + 1)
    k, j, i = np.meshgrid(*[np.arange(N)] * 3, indexing="ij")
    pos = np.stack([i + .5 - vx, j + .5, k + .5, np.full(i.shape, time + .5)], -1)
    band = (slice(10, -10),) * 3 if N > 20 else None
    assert band is None  # a 12^3 frame has no cells inside the 10-cell band
    assert np.array_equal(out, dst0)
    out1 = O.warp_slice(defo, phi2, dst0, time, blendAlpha=1.0, optimized=False)
    b1 = (slice(1, -1),) * 3
    np.testing.assert_allclose(out1[b1], mc4(phi2, pos)[b1], atol=1e-12)


def test_repeat_frame_closed_form():
    """repeatFrame4d: slices int(srct-range+0.5)..int(srct) take the linear
    blend of slices floor(srct), floor(srct)+1 at every bnd-interior cell."""
    T, N = 8, 6
    phi = RNG.standard_normal((T, N, N, N))
    out = O.repeat_frame4d(phi, 3.0, 2.0, 0)
    ref = phi.copy()
    ref[1:4] = phi[3]
    np.testing.assert_allclose(out, ref, atol=0)
    out = O.repeat_frame4d(phi, 2.25, 1.0, 1)
    ref = phi.copy()
    blend = 0.75 * phi[2] + 0.25 * phi[3]
    for t in (1, 2):
        ref[t, 1:-1, 1:-1, 1:-1] = blend[1:-1, 1:-1, 1:-1]
    np.testing.assert_allclose(out, ref, atol=1e-14)


def test_extrap_planar_closed_form():
    """extrap4dLsSimple outside a planar level set x < a: the first layer keeps
    its values, layer d (3 <= d <= distance+1) holds first-layer value +
    (d - 2), cells beyond get distance + 2; inside is untouched."""
    T, Z, Y, X = 6, 6, 6, 16
    x = np.arange(X) + 0.5
    phi = np.broadcast_to(x - 4.2, (T, Z, Y, X)).copy()  # inside: i <= 3
    dist = 6
    out = O.extrap4d_ls_simple(phi, distance=dist, inside=False)
    core = (slice(1, -1),) * 3
    first = phi[(slice(1, -1),) * 3 + (4,)]
    for i in range(1, X - 1):
        got = out[core + (i,)]
        if i <= 3:
            np.testing.assert_array_equal(got, phi[core + (i,)])
        elif i == 4:
            np.testing.assert_array_equal(got, first)
        elif i <= dist + 3:  # passes d = 2..dist fill layers 3..dist+1
            np.testing.assert_allclose(got, first + (i - 4), atol=1e-12)
        else:
            assert (got == dist + 2).all()
    # inside=True mirrors: the inside region (phi <= 0 here) is extrapolated
    # with direction -1 from the first layer inside the interface
    out_in = O.extrap4d_ls_simple(phi, distance=dist, inside=True)
    first_in = phi[core + (3,)]
    np.testing.assert_allclose(out_in[core + (2,)], first_in - 1, atol=1e-12)


def _window(v, time, W):
    """The reference's defo window of optflow4d.cpp:1822-1866 for srcTime."""
    T = v.shape[0]
    t = min(int(time), T - 1)
    cur = W // 2
    idx = [min(max(t - cur + tl, 0), T - 1) for tl in range(W)]
    return v[idx], time - (t - cur)


def mc4v(vol, pos):
    """map_coordinates of a (T,Z,Y,X,4) field at positions (..., 4) (x,y,z,t),
    cell centres at c + 1/2, edge clamp."""
    coords = np.stack([pos[..., 3] - 0.5, pos[..., 2] - 0.5, pos[..., 1] - 0.5,
                       pos[..., 0] - 0.5])
    return np.stack([map_coordinates(vol[..., c], coords, order=1, mode="nearest")
                     for c in range(4)], -1)


def test_warp_defo_volumes_against_map_coordinates():
    """Two volumes (not aligned): vt = alpha dv1(p - dv2(p)) + third dv2(p) on
    the clamped window; three volumes chain; aligned: dvt on the whole window,
    then vt = dvt(p + (1 - alpha) dv1(p)); the lookup then runs with alpha 1
    over the bnd-10 interior.  Restated with map_coordinates on windows built
    here."""
    N, T = 24, 10
    time = 4.3
    phi = RNG.standard_normal((T, N, N, N))
    vols = [RNG.standard_normal((T, N, N, N, 4)) * 0.6 for _ in range(3)]
    dst0 = np.zeros((N, N, N))
    k, j, i = np.meshgrid(*[np.arange(N)] * 3, indexing="ij")
    W = int(T * 0.2)
    wins = []
    for v in vols:
        wv, pt = _window(v, time, W)   # srcTime = time (sf_t 1, off_t 1/2)
        wins.append(wv)
    p = np.stack([i + .5, j + .5, k + .5, np.full(i.shape, pt + .5)], -1)
    pos0 = np.stack([i + .5, j + .5, k + .5, np.full(i.shape, time + .5)], -1)
    b = (slice(10, -10),) * 3

    def check(out, vt):
        np.testing.assert_allclose(out[b], mc4(phi, pos0 - vt)[b], atol=1e-9)
        mask = np.ones(out.shape, bool)
        mask[b] = False
        assert not out[mask].any()    # outside the bnd-10 interior: untouched

    out = O.warp_slice_vols(vols[:2], phi, dst0, time, blendAlpha=0.7, thirdAlpha=0.4,
                            bordSkip=0)
    v2 = mc4v(wins[1], p)
    check(out, 0.7 * mc4v(wins[0], p - v2) + 0.4 * v2)
    out = O.warp_slice_vols(vols, phi, dst0, time, blendAlpha=0.5, thirdAlpha=0.3,
                            fourthAlpha=0.2, bordSkip=0)
    v3 = mc4v(wins[2], p)
    v2 = mc4v(wins[1], p - v3)
    v1 = mc4v(wins[0], p - v3 - v2)
    check(out, 0.5 * v1 + 0.3 * v2 + 0.2 * v3)
    out = O.warp_slice_vols(vols[:2], phi, dst0, time, blendAlpha=0.7, thirdAlpha=0.4,
                            bordSkip=0, doAligned=True)
    kk, jj, ii, tt = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), np.arange(W),
                                 indexing="ij")
    q = np.stack([ii + .5, jj + .5, kk + .5, tt + .5], -1).transpose(3, 0, 1, 2, 4)
    dvt = 0.7 * mc4v(wins[0], q - mc4v(wins[1], q)) + 0.4 * mc4v(wins[1], q)
    check(out, mc4v(dvt, p + 0.3 * mc4v(wins[0], p)))
    with pytest.raises(RuntimeError):
        O.warp_slice_vols(vols[:1], phi, dst0, time)
