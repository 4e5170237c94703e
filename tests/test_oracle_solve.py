"""Pins the oracle's optical-flow solve, multiscale driver, blur, error metric
and projection (no GPU).

The one-level system of opticalFlowDim (optflow4d.cpp:361-553) is assembled
here independently with scipy.sparse and solved by spsolve; an independent
numpy Jacobi-PCG with the reference's signed-max stopping rule must need the
same number of iterations as the oracle.  The multiscale driver is checked
against a composition of map_coordinates warps, 16-cell-mean restriction and
spsolve solves.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl
from scipy.ndimage import map_coordinates

from oracle import oracle as O
from ofblend_amd.abi import default_params

RNG = np.random.default_rng(11)
NB = [(-1, 0, 0, 0), (0, -1, 0, 0), (0, 0, -1, 0), (0, 0, 0, -1),
      (0, 0, 0, 1), (0, 0, 1, 0), (0, 1, 0, 0), (1, 0, 0, 0)]  # (t,z,y,x)


def assemble(i0, i1, w, e, vin=None):
    """Interior unknowns U (4 per interior cell) of
    (g g^T + (8w+e) I) U_c - w sum_n U_n = b_c; boundary neighbours are 0.
    Returns (A, b, interior index array)."""
    T, Z, Y, X = i0.shape
    h = 1.0 / X
    ints = [(t, z, y, x) for t in range(1, T - 1) for z in range(1, Z - 1)
            for y in range(1, Y - 1) for x in range(1, X - 1)]
    idx = {c: k for k, c in enumerate(ints)}
    rows, cols, vals = [], [], []
    b = np.zeros(4 * len(ints))
    for k, (t, z, y, x) in enumerate(ints):
        g = np.array([i1[t, z, y, x + 1] - i1[t, z, y, x - 1],
                      i1[t, z, y + 1, x] - i1[t, z, y - 1, x],
                      i1[t, z + 1, y, x] - i1[t, z - 1, y, x],
                      i1[t + 1, z, y, x] - i1[t - 1, z, y, x]]) * (X / 2.0)
        It = i1[t, z, y, x] - i0[t, z, y, x]
        B = np.outer(g, g) + (8 * w + e) * np.eye(4)
        for d in range(4):
            for m in range(4):
                rows.append(4 * k + d); cols.append(4 * k + m); vals.append(B[d, m])
            b[4 * k + d] = -g[d] * It
        for (dt, dz, dy, dx) in NB:
            n = (t + dt, z + dz, y + dy, x + dx)
            if n in idx:
                for d in range(4):
                    rows.append(4 * k + d); cols.append(4 * idx[n] + d); vals.append(-w)
        if vin is not None:
            vc = vin[t, z, y, x]
            for (dt, dz, dy, dx) in NB:
                nn = (min(max(t + dt, 0), T - 1), min(max(z + dz, 0), Z - 1),
                      min(max(y + dy, 0), Y - 1), min(max(x + dx, 0), X - 1))
                b[4 * k:4 * k + 4] -= w * (vc - vin[nn]) * h
            b[4 * k:4 * k + 4] -= e * vc * h
    A = sp.csr_matrix((vals, (rows, cols)), shape=(4 * len(ints),) * 2)
    return A, b, ints


def scatter(U, ints, shape):
    v = np.zeros(shape + (4,))
    for k, c in enumerate(ints):
        v[c] = U[4 * k:4 * k + 4]
    return v


def drop_pair(N, T, seeds=(1, 2), dtype=np.float64):
    return [O.drop_sdf(N, N, N, T, seed=s, dtype=dtype) for s in seeds]


def test_level_solve_matches_spsolve():
    N, T = 7, 6
    i0, i1 = drop_pair(N, T)
    vin = RNG.standard_normal((T, N, N, N, 4)) * 0.3
    w, e = 0.1, 1e-4
    A, b, ints = assemble(i0, i1, w, e, vin)
    U = spl.spsolve(A.tocsc(), b)
    ref = scatter(U, ints, i0.shape) * N  # vel = U / h
    p = default_params(wSmooth=w, wEnergy=e, cgAccuracy=1e-13, cgMaxIter=20000)
    vel, it, res, st, rhs = O.optical_flow4d(vin, i0, i1, p, want_rhs=True)
    assert st == 0
    np.testing.assert_allclose(vel, ref, atol=1e-8 * max(1, np.abs(ref).max()))
    # rhsT is component 0 of the right-hand side
    np.testing.assert_allclose(rhs[tuple(np.array(ints).T)], b[0::4], atol=1e-12)


def numpy_pcg(A, b, acc_rel, max_iter):
    """Jacobi-PCG from x=0 with the signed-max stopping rule; the identity
    rows of the outer layer contribute residual 0 to the max."""
    x = np.zeros_like(b)
    r = b.copy()
    res0 = max(r.max(), 0.0)
    acc = acc_rel * res0
    D = A.diagonal()
    z = r / D
    sigma = z @ r
    p = z.copy()
    for it in range(max_iter):
        q = A @ p
        alpha = sigma / (p @ q)
        x += alpha * p
        r -= alpha * q
        if max(r.max(), 0.0) <= acc:
            return x, it + 1
        z = r / D
        sn = z @ r
        p = z + (sn / sigma) * p
        sigma = sn
    return x, max_iter


def test_pcg_iteration_count_matches_independent_pcg():
    N, T = 8, 8
    i0, i1 = drop_pair(N, T)
    w, e = 0.1, 1e-4
    A, b, ints = assemble(i0, i1, w, e)
    for acc in (1e-3, 1e-5):
        _, it_ref = numpy_pcg(A, b, acc, 1000)
        p = default_params(wSmooth=w, wEnergy=e, cgAccuracy=acc)
        _, it, _, st = O.optical_flow4d(np.zeros(i0.shape + (4,)), i0, i1, p)
        assert it == it_ref and st == 0


def test_level_solve_exits():
    N, T = 6, 6
    i0 = O.drop_sdf(N, N, N, T, seed=3, dtype=np.float64)
    p = default_params(wSmooth=0.1, wEnergy=1e-4, cgAccuracy=1e-5)
    # identical images: rhs 0 -> early exit with zero flow (optflow4d.cpp:286-291)
    vel, it, res, st = O.optical_flow4d(np.zeros(i0.shape + (4,)), i0, i0.copy(), p)
    assert it == 0 and st == 3 and not vel.any()
    # iteration cap: status "not converged", result still used
    p = default_params(wSmooth=0.1, wEnergy=1e-4, cgAccuracy=1e-12, cgMaxIter=3)
    i1 = O.drop_sdf(N, N, N, T, seed=4, dtype=np.float64)
    vel, it, res, st = O.optical_flow4d(np.zeros(i0.shape + (4,)), i0, i1, p)
    assert it == 3 and st == 1 and np.abs(vel).max() > 0


def mc4(data, pos):
    coords = np.stack([pos[..., 3] - 0.5, pos[..., 2] - 0.5, pos[..., 1] - 0.5,
                       pos[..., 0] - 0.5])
    return map_coordinates(data, coords, order=1, mode="nearest")


def np_advect(vel, a):
    T, Z, Y, X = a.shape
    t, z, y, x = np.meshgrid(*[np.arange(s) for s in a.shape], indexing="ij")
    pos = np.stack([x + .5, y + .5, z + .5, t + .5], -1) - vel
    out = np.zeros_like(a)
    inner = (slice(1, -1),) * 4
    out[inner] = mc4(a, pos)[inner]
    return out


def np_neumann0(a):
    T, Z, Y, X = a.shape
    t, z, y, x = np.meshgrid(*[np.arange(s) for s in a.shape], indexing="ij")
    src = (np.clip(t, 1, T - 2), np.clip(z, 1, Z - 2), np.clip(y, 1, Y - 2),
           np.clip(x, 1, X - 2))
    return a[src]


def np_resample(src, shape, f):
    t, z, y, x = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    pos = np.stack([x * f + f / 2, y * f + f / 2, z * f + f / 2, t * f + f / 2], -1)
    if src.ndim == 5:
        return np.stack([mc4(src[..., c], pos) for c in range(4)], -1)
    return mc4(src, pos)


def np_multiscale(vel, i0, i1, w, e, mg):
    """opticalFlowMultiscaleTemplate with multiStep 1, no blur/projection."""
    T, Z, Y, X = i0.shape
    if X > mg:
        small = (T // 2, Z // 2, Y // 2, X // 2)
        vs = np_resample(vel, small, 2.0) * 0.5
        vs = np_multiscale(vs, np_resample(i0, small, 2.0), np_resample(i1, small, 2.0),
                           w, e, mg)
        vel = np_resample(vs, i0.shape, 0.5) * 2.0
    i0w = np_neumann0(np_advect(vel, i0))
    A, b, ints = assemble(i0w, i1, w, e)
    U = spl.spsolve(A.tocsc(), b)
    return vel + scatter(U, ints, i0.shape) * X


def test_multiscale_matches_independent_composition():
    N, T = 8, 8
    i0, i1 = drop_pair(N, T)
    w, e = 0.1, 1e-4
    ref = np_multiscale(np.zeros(i0.shape + (4,)), i0, i1, w, e, 4)
    p = default_params(wSmooth=w, wEnergy=e, cgAccuracy=1e-13, cgMaxIter=20000,
                       minGridSize=4)
    vel, err, info = O.multiscale4d(np.zeros(i0.shape + (4,)), i0, i1, p)
    s = info.summary()
    assert [l["dims"][0] for l in s["levels"]] == [8, 4]
    np.testing.assert_allclose(vel, ref, atol=1e-7 * max(1, np.abs(ref).max()))
    # the returned error is calcLsDiff of the finally warped i0 (corr 0.1/20)
    i0w = np_neumann0(np_advect(vel, i0))
    assert abs(err - O.ls_diff(i0w, i1, correction=0.1 / 20)) <= 1e-9 * max(1, err)


def test_ls_diff_brute_force():
    a = RNG.standard_normal((6, 7, 8, 9))
    b = RNG.standard_normal((6, 7, 8, 9)) * 2
    a[0, 0, 0, 0] = 0.0  # >= 0 counts as positive
    b[0, 0, 0, 0] = -0.0
    for bnd in (0, 1):
        sl = (slice(bnd, a.shape[0] - bnd),) + (slice(bnd, None if bnd == 0 else -bnd),) * 3
        aa, bb = a[sl], b[sl]
        mism = (aa < 0) != (bb < 0)
        d = np.minimum(np.abs(aa - bb) * 0.3, 1.0)
        ref = d[mism].sum() * 1e6 / aa.size
        got, out = O.ls_diff(a, b, correction=0.3, bnd=bnd, want_out=True)
        assert abs(got - ref) <= 1e-12 * ref
        assert np.allclose(out[sl][mism], d[mism]) and not out[sl][~mism].any()


def np_blur_pass(a, sigma, s):
    """One normalised Gaussian pass over the clipped (2s+1)^4 window,
    written for interior cells only (separable: Gaussian weights factor)."""
    out = a.copy()
    w = np.exp(-np.arange(-s, s + 1) ** 2 / (2 * sigma * sigma))
    for ax in range(4):
        n = out.shape[ax]
        num = np.zeros_like(out)
        den = np.zeros(n)
        for o in range(-s, s + 1):
            lo, hi = max(0, -o), min(n, n - o)
            src = [slice(None)] * out.ndim
            dst = [slice(None)] * out.ndim
            src[ax] = slice(lo + o, hi + o)
            dst[ax] = slice(lo, hi)
            num[tuple(dst)] += w[o + s] * out[tuple(src)]
            den[lo:hi] += w[o + s]
        shape = [1] * out.ndim
        shape[ax] = n
        out = num / den.reshape(shape)
    return out


def test_gaussian_blur_separable_and_boundary_rule():
    a = RNG.standard_normal((7, 8, 9, 10, 4))
    sigma = 1.0
    s = 1
    inner = (slice(1, -1),) * 4
    b = np.zeros_like(a)
    b[inner] = np_blur_pass(a, sigma, s)[inner]   # pass 1 into a zeroed temp
    ref = a.copy()
    ref[inner] = np_blur_pass(b, sigma, s)[inner]  # pass 2 back: boundary kept
    np.testing.assert_allclose(O.gaussian_blur(a, sigma, 1), ref, atol=1e-12)
    # a constant is preserved where neither pass reaches the zeroed outer
    # layer of the temporary (distance >= 2s from the border)
    c = np.full((8, 8, 8, 8), 2.5)
    g = O.gaussian_blur(c, 1.0, 1)
    assert abs(g[2:-2, 2:-2, 2:-2, 2:-2] - 2.5).max() < 1e-12
    assert abs(g[1, 3, 3, 3] - 2.5) > 1e-3


def test_projection_planar_level_sets():
    """corrVelsOf4d on planar SDFs phi = x - a and phi_t = x - b: the normal
    is +x and the line search (steps 0.5, halved on a direction flip, stop
    below 0.1) lands within its last step of d = a - b; vel -= n*d."""
    T, Z, Y, X = 8, 8, 8, 16
    x = np.arange(X) + 0.5
    a, b = 7.3, 6.0
    phiOrg = np.broadcast_to(x - a, (T, Z, Y, X)).astype(np.float64).copy()
    phiTgt = np.broadcast_to(x - b, (T, Z, Y, X)).astype(np.float64).copy()
    vel = np.zeros((T, Z, Y, X, 4))
    dst, v = O.corr_vels_of4d(vel, phiOrg, phiTgt, threshPhi=4.0, postVelBlur=0.0,
                              maxIter=40)
    core = dst[2:-2, 2:-2, 2:-2, 4:-4]
    assert np.abs(core[..., 1:]).max() == 0.0
    assert np.abs(core[..., 0] - (a - b)).max() <= 0.125
    np.testing.assert_allclose(v, -dst, atol=0)


def test_warm_start_energy_optimal_scaling():
    """The warm start is scaled by gamma = b.w / w.A w: a multiple of the
    converged flow is rescaled to it (the solve needs no iterations), a flow
    orthogonal in the energy sense is dropped (gamma 0), the zero guess is
    the cold start."""
    N, T = 8, 8
    i0, i1 = drop_pair(N, T)
    w, e = 0.1, 1e-4
    p = default_params(wSmooth=w, wEnergy=e, cgAccuracy=1e-9, cgMaxIter=5000)
    zero = np.zeros(i0.shape + (4,))
    v, it, _, _ = O.optical_flow4d(zero, i0, i1, p)
    p2 = default_params(wSmooth=w, wEnergy=e, cgAccuracy=1e-6, cgMaxIter=5000)
    for s in (1.0, 2.5, -0.7):
        vw, it_w, _, st = O.optical_flow4d(zero, i0, i1, p2, warm=s * v)
        assert st == 0 and it_w <= 1, (s, it_w)
        np.testing.assert_allclose(vw, v, atol=1e-5 * np.abs(v).max())
    # the exact Galerkin factor: gamma = b.w / w.A w on the assembled system
    A, b, ints = assemble(i0, i1, w, e)
    g = np.random.default_rng(0).standard_normal(v.shape)
    wv = np.concatenate([g[c] for c in ints]) / N
    gamma = (b @ wv) / (wv @ (A @ wv))
    p1 = default_params(wSmooth=w, wEnergy=e, cgAccuracy=1e-6, cgMaxIter=0)
    vg, it_g, _, st_g = O.optical_flow4d(zero, i0, i1, p1, warm=g)
    # with no iterations the result is the adopted guess gamma w (this guess
    # passes the ||b - gamma A w|| <= ||b|| test)
    assert np.linalg.norm(b - gamma * (A @ wv)) <= np.linalg.norm(b)
    ref = scatter(gamma * wv, ints, i0.shape) * N
    np.testing.assert_allclose(vg, ref, atol=1e-12 * np.abs(ref).max())
