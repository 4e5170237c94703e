"""Pins the oracle's level-set projection corrVelsOf4d (optflow4d.cpp:613-802)
without a GPU.

* The discrete line search of projectCell (optflow4d.cpp:676-708) on planar
  level sets phi = x - a, phi_target = x - b: the endpoints below were worked
  out by hand from the rule (initial step 0.5 when maxIter < 50, else 0.25;
  move toward the target value, halve the step on a direction change, stop
  once the step is below 0.1 or after maxIter moves).  Two cases let maxIter
  bind, one per initial step, so a wrong initial step or a wrong stopping
  test changes the endpoint.
* The 81-term extrapolation (knCvExpolBlur4d, optflow4d.cpp:613-626): an
  unmarked cell whose 80 neighbours are marked with the value c converges as
  v_k = (80 c + v_(k-1)) / 81, summed in the reference's loop order.
* The whole call on random smooth data against an independent numpy
  restatement (map_coordinates interpolation, vectorised line search, the
  81-term sum in loop order, the separable Gaussian, setBound's band).
"""
import numpy as np
import pytest

from oracle import oracle as O
from test_oracle_solve import mc4, np_blur_pass


def planar(shape, a, b):
    T, Z, Y, X = shape
    x = np.arange(X) + 0.5
    phiOrg = np.broadcast_to(x - a, shape).astype(np.float64).copy()
    phiTgt = np.broadcast_to(x - b, shape).astype(np.float64).copy()
    return phiOrg, phiTgt


# (a - b, maxIter, endpoint of the line search, moves it takes)
#  1.3, 40: start 0.5: -1.3 -> ... -> 1.7, 1.45, 1.2, 1.325, 1.2625 (stop)
#  1.3, 60: start 0.25: -1.3 -> ... -> 1.45, 1.325, 1.2, 1.2625 (stop)
#  5.3, 40: start 0.5: 22 moves to 5.7, then 5.45, 5.2, 5.325, 5.2625
#  9.3, 40: start 0.5: 38 moves to 9.7, then 9.45, 9.2 -- maxIter binds
#  9.3, 60: start 0.25: 60 moves from -9.3 end at 5.7 -- maxIter binds
LINE_SEARCH = [(1.3, 40, 1.2625), (1.3, 60, 1.2625), (5.3, 40, 5.2625),
               (9.3, 40, 9.2), (9.3, 60, 5.7)]


@pytest.mark.parametrize("amb,maxIter,endpoint", LINE_SEARCH)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_line_search_endpoint_planar(amb, maxIter, endpoint, dtype):
    T, Z, Y, X = 6, 6, 6, 40
    b = 12.0
    a = b + amb
    phiOrg, phiTgt = planar((T, Z, Y, X), a, b)
    vel = np.zeros((T, Z, Y, X, 4), dtype=dtype)
    dst, v = O.corr_vels_of4d(vel, phiOrg.astype(dtype), phiTgt.astype(dtype),
                              threshPhi=10.0, postVelBlur=0.0, maxIter=maxIter)
    # cells whose whole search path x + 0.5 + d (d in [-amb, amb + 0.5]) stays
    # inside the grid and away from the clamped outer half cell
    lo = int(np.ceil(amb + 1.0))
    hi = X - int(np.ceil(amb + 2.0))
    core = dst[2:-2, 2:-2, 2:-2, lo:hi]
    tol = 1e-12 if dtype == np.float64 else 2e-5
    assert np.abs(core[..., 0] - endpoint).max() <= tol, (amb, maxIter)
    assert not core[..., 1:].any()
    np.testing.assert_array_equal(v, -dst)


def test_box_fill_isolated_cell_closed_form():
    """One unmarked cell (its search leaves the grid: d = 9999 > threshPhi)
    among marked cells of value c: the 4 extrapolation passes (int(|vel|max
    + 4) with vel = 0) give v_k = (80 c + v_(k-1)) / 81 in the loop order
    t, z, y, x with the cell itself as the centre (41st) term."""
    shape = (10, 10, 10, 16)
    phiOrg, phiTgt = planar(shape, 7.3, 6.0)
    cell = (4, 4, 4, 8)
    phiTgt[cell] = 50.0  # d0 = -48.8: the first probe leaves the grid
    vel = np.zeros(shape + (4,))
    dst, _ = O.corr_vels_of4d(vel, phiOrg, phiTgt, threshPhi=4.0,
                              postVelBlur=0.0, maxIter=40)
    c = dst[3, 3, 3, 7, 0]  # a marked neighbour (the line search's endpoint)
    assert abs(c - 1.2625) < 1e-12
    # every neighbour of the cell is marked with exactly c
    nb = dst[3:6, 3:6, 3:6, 7:10, 0].copy()
    nb[1, 1, 1, 1] = c
    assert (nb == c).all()
    v = 0.0
    for _ in range(4):
        s = 0.0
        for term in range(81):
            s += v if term == 40 else c
        v = s * (1.0 / 81.0)
    assert dst[cell + (0,)] == v
    assert abs(v - c * (1.0 - 81.0 ** -4)) < 1e-14
    assert not dst[cell][1:].any()


# ----------------------------------------------------------------------------
# independent numpy restatement of corrVelsOfTempl (FP64)
def np_normal(phi, pos):
    n = np.empty(pos.shape)
    for d in range(4):
        e = np.zeros(4)
        e[d] = 0.5
        n[..., d] = mc4(phi, pos + e) - mc4(phi, pos - e)
    l = n[..., 0] * n[..., 0] + n[..., 1] * n[..., 1] + n[..., 2] * n[..., 2] \
        + n[..., 3] * n[..., 3]
    eps2 = 1e-10 * 1e-10
    keep = np.abs(l - 1.0) < eps2
    zero = ~keep & (l <= eps2)
    with np.errstate(divide="ignore"):
        fac = 1.0 / np.sqrt(l)
    out = np.where(keep[..., None], n, n * fac[..., None])
    out[zero] = 0.0
    return out


def np_line_search(phi, target, pos, n, maxIter):
    shape = np.array(phi.shape[::-1])  # X, Y, Z, T
    step = np.full(target.shape, 0.5 if maxIter < 50 else 0.25)
    last = np.zeros(target.shape, dtype=int)
    d = mc4(phi, pos) - target
    active = np.ones(target.shape, dtype=bool)
    for _ in range(maxIter):
        pn = pos + n * d[..., None]
        ip = np.trunc(pn)
        inb = ((ip >= 0) & (ip < shape)).all(-1)
        out = active & ~inb
        d[out] = 9999.0
        active &= inb
        vn = mc4(phi, np.where(active[..., None], pn, pos))
        lt = active & (vn < target)
        ge = active & ~(vn < target)
        step[lt & (last < 0)] *= 0.5
        d[lt] += step[lt]
        last[lt] = 1
        step[ge & (last > 0)] *= 0.5
        d[ge] -= step[ge]
        last[ge] = -1
        active &= ~(step < 0.1)
        if not active.any():
            break
    return d


def np_box_pass(dst, marker):
    T, Z, Y, X = marker.shape
    inner = (slice(1, -1),) * 4
    val = np.zeros(dst[inner].shape)
    for dt in (-1, 0, 1):
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    val = val + dst[1 + dt:T - 1 + dt, 1 + dz:Z - 1 + dz,
                                    1 + dy:Y - 1 + dy, 1 + dx:X - 1 + dx]
    out = dst.copy()
    um = marker[inner] == 0
    sub = out[inner]
    sub[um] = (val * (1.0 / 81.0))[um]
    out[inner] = sub
    return out


def np_gauss(a, sigma):
    s = max(1, int(sigma + 0.5))
    inner = (slice(1, -1),) * 4
    b = np.zeros_like(a)
    b[inner] = np_blur_pass(a, sigma, s)[inner]
    out = a.copy()
    out[inner] = np_blur_pass(b, sigma, s)[inner]
    return out


def np_set_bound_zero(a, w):
    """Grid4d::setBound (grid4d.cpp:355-367): cells with an index <= w or
    >= n - 1 - w."""
    out = a.copy()
    T, Z, Y, X = a.shape[:4]
    t, z, y, x = np.meshgrid(*[np.arange(s) for s in (T, Z, Y, X)], indexing="ij")
    band = ((x <= w) | (x >= X - 1 - w) | (y <= w) | (y >= Y - 1 - w) |
            (z <= w) | (z >= Z - 1 - w) | (t <= w) | (t >= T - 1 - w))
    out[band] = 0.0
    return out


def np_corr_vels(vel, phiOrg, phiTgt, threshPhi, postVelBlur, resetBndWidth,
                 maxIter):
    vel = vel.copy()
    T, Z, Y, X = phiOrg.shape
    passes = int(np.sqrt((vel * vel).sum(-1).max()) + 4)
    single = 1 if postVelBlur < 0.98 else 0
    core = (slice(2, -2),) * 4
    t, z, y, x = np.meshgrid(*[np.arange(s) for s in (T, Z, Y, X)], indexing="ij")
    centre = np.stack([x + 0.5, y + 0.5, z + 0.5, t + 0.5], -1)
    dst = None
    while postVelBlur >= 0.98 or single:
        dst = np.zeros_like(vel)
        marker = np.zeros(phiOrg.shape)
        pos = (centre - vel)[core]
        n = np_normal(phiOrg, pos)
        d = np_line_search(phiOrg, phiTgt[core], pos, n, maxIter)
        ok = np.abs(d) <= threshPhi
        sub = np.where(ok[..., None], n * d[..., None], 0.0)
        dst[core] = sub
        marker[core] = ok
        for _ in range(passes):
            dst = np_box_pass(dst, marker)
        if postVelBlur > 1e-10:
            dst = np_gauss(dst, 0.5 * postVelBlur)
        if resetBndWidth > 0:
            dst = np_set_bound_zero(dst, int(resetBndWidth * X) + 1)
        vel = vel + (-1.0) * dst
        postVelBlur *= 0.5
        if single > 0:
            single -= 1
    return dst, vel


def smooth_field(shape, rng, scale):
    a = rng.standard_normal(shape)
    for ax in range(4):
        a = (np.roll(a, 1, ax) + 2 * a + np.roll(a, -1, ax)) / 4
    return a * scale


@pytest.mark.parametrize("postVelBlur,resetBndWidth,maxIter",
                         [(0.0, -1.0, 40), (4.0, 0.1, 40), (0.5, -1.0, 100)])
def test_corr_vels_matches_numpy_restatement(postVelBlur, resetBndWidth, maxIter):
    rng = np.random.default_rng(5)
    shape = (8, 9, 10, 11)
    T, Z, Y, X = shape
    t, z, y, x = np.meshgrid(*[np.arange(s) + 0.5 for s in shape], indexing="ij")
    r = np.sqrt((x - 5.2) ** 2 + (y - 4.7) ** 2 + (z - 4.1) ** 2 + (t - 3.9) ** 2)
    phiOrg = r - 2.6 + smooth_field(shape, rng, 0.3)
    phiTgt = r - 3.4 + smooth_field(shape, rng, 0.3)
    vel = np.stack([smooth_field(shape, rng, 0.8) for _ in range(4)], -1)
    dst, v = O.corr_vels_of4d(vel, phiOrg, phiTgt, threshPhi=4.0,
                              postVelBlur=postVelBlur,
                              resetBndWidth=resetBndWidth, maxIter=maxIter)
    rd, rv = np_corr_vels(vel, phiOrg, phiTgt, 4.0, postVelBlur, resetBndWidth,
                          maxIter)
    assert np.abs(dst).max() > 0.1
    # the unmarked cells hold extrapolated values, the marked ones n * d
    np.testing.assert_allclose(dst, rd, atol=1e-10, rtol=0)
    np.testing.assert_allclose(v, rv, atol=1e-10, rtol=0)
