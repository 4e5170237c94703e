"""Pins the oracle's reduced-precision storage of the CG search direction
(extension): its float -> bfloat16 rounding against the definition (8
significant bits, half to even), and the PCG with rounded directions against
an independent numpy PCG that rounds the same way (no GPU)."""
import numpy as np

from oracle import oracle as O

RNG = np.random.default_rng(11)


def _samples():
    a = np.concatenate([
        RNG.standard_normal(3000) * 10.0 ** RNG.integers(-9, 6, 3000),
        [0.0, -0.0, 1.0, 65504.0, 65519.99, 65520.0, 7e4, 6.1035e-05, 5.96e-08,
         2.98e-08, 2.99e-08, 1e-9, 1.00390625, 1.001953125, 1.005859375,
         3.0e38, np.inf, -np.inf]])
    return a.astype(np.float32)


def _bf16_ref(x):
    """Round to nearest even at 8 significand bits via exact float64
    arithmetic on the value (no bit tricks)."""
    x = float(np.float32(x))
    if not np.isfinite(x) or x == 0.0:
        return np.float32(x)
    m, e = np.frexp(abs(x))          # x = m 2^e, m in [0.5, 1)
    q = np.round(m * 256.0)          # 8 significant bits, half to even
    y = np.ldexp(q, e - 8)
    if y > np.finfo(np.float32).max:
        y = np.inf
    return np.float32(np.copysign(y, x))


def test_round_bf16_matches_definition():
    for x in _samples():
        if abs(x) > 3.3e38:
            continue
        want = _bf16_ref(x)
        got = np.float32(O.round_bf16(x))
        assert got == want, (x, got, want)
    assert np.isnan(O.round_bf16(float("nan")))


def numpy_pcg_rounded(A, b, acc_rel, max_iter, rnd):
    """Jacobi-PCG with the signed-max stopping rule where every stored search
    direction is replaced by rnd(p) and the rounded direction is used in A p,
    the step and the next direction."""
    x = np.zeros_like(b)
    r = b.copy()
    acc = acc_rel * max(r.max(), 0.0)
    D = A.diagonal()
    z = r / D
    sigma = z @ r
    p = rnd(z)
    for it in range(max_iter):
        q = A @ p
        alpha = sigma / (p @ q)
        x += alpha * p
        r -= alpha * q
        if max(r.max(), 0.0) <= acc:
            return x, it + 1
        z = r / D
        sn = z @ r
        p = rnd(z + (sn / sigma) * p)
        sigma = sn
    return x, max_iter


def _tobf16(v):
    return np.array([_bf16_ref(x) for x in v.astype(np.float32)], dtype=np.float64)


def test_pcg_with_rounded_directions_matches_numpy():
    """opticalFlow4d (FP64 oracle) with the direction stored as bfloat16
    follows a numpy PCG of the independently assembled system that rounds p
    the same way.  Rounding is discontinuous, so last-bit differences of the
    two FP64 computations flip single roundings after some iterations and
    the trajectories then part; before that they agree to 1e-11 while the
    unrounded oracle is 1e-6 away, and both converge to the same flow."""
    from test_oracle_solve import assemble, scatter, drop_pair
    from ofblend_amd.abi import default_params
    N, T = 8, 8
    i0, i1 = drop_pair(N, T)
    A, b, ints = assemble(i0, i1, 0.1, 1e-4)

    def oracle(prec, acc, cap):
        pp = default_params(wSmooth=0.1, wEnergy=1e-4, cgAccuracy=acc, precision=prec,
                            cgMaxIter=cap)
        return O.optical_flow4d(np.zeros(i0.shape + (4,)), i0, i1, pp)

    for cap in (3, 8):
        v, it, _, _ = oracle(1, 1e-30, cap)
        v0, _, _, _ = oracle(0, 1e-30, cap)
        x, _ = numpy_pcg_rounded(A, b, 1e-30, cap, _tobf16)
        ref = scatter(x, ints, i0.shape) * N
        scale = np.abs(ref).max()
        assert np.abs(v - ref).max() <= 1e-11 * scale
        assert np.abs(v0 - ref).max() >= 1e-6 * scale
    v, it, _, st = oracle(1, 1e-6, 1000)
    x, it_np = numpy_pcg_rounded(A, b, 1e-6, 1000, _tobf16)
    ref = scatter(x, ints, i0.shape) * N
    assert st == 0 and abs(it - it_np) <= 3
    assert np.abs(v - ref).max() <= 1e-5 * np.abs(ref).max()
    # the plain solve needs at most as many iterations, and the same flow
    v0, it0, _, _ = oracle(0, 1e-6, 1000)
    assert it0 <= it and np.abs(v - v0).max() <= 1e-5 * np.abs(v0).max()


# ---------------------------------------------------------------------------
# mixed-precision images (precision FLOF_PREC_IMG_BF16 = 2)

def test_round_f16_matches_numpy():
    """The oracle's float -> binary16 rounding equals numpy's IEEE
    conversion (nearest even, subnormals, overflow to infinity)."""
    a = np.concatenate([
        RNG.standard_normal(20000) * 10.0 ** RNG.integers(-10, 6, 20000),
        RNG.uniform(-1, 1, 2000) * 6.2e-5,          # around the subnormal range
        np.float32(2.0 ** -24) * np.arange(-40, 41),  # subnormal grid points
        (np.arange(1, 200) + 0.5) * np.float32(2.0 ** -24),  # subnormal ties
        [65504.0, 65519.99, 65520.0, 65535.0, 7e4, -7e4, 6.1035156e-05,
         6.1035153e-05, 2.98023224e-08, 2.9802326e-08, 1.0 + 2 ** -11,
         1.0 + 3 * 2 ** -11, 0.0, -0.0]]).astype(np.float32)
    with np.errstate(over="ignore"):
        want = a.astype(np.float16).astype(np.float32)
    got = np.array([O.round_f16(x) for x in a], dtype=np.float32)
    bad = ~((got == want) & (np.signbit(got) == np.signbit(want)))
    assert not bad.any(), (a[bad][:5], got[bad][:5], want[bad][:5])


def _bf16(a):
    return np.array([O.round_bf16(x) for x in np.asarray(a, np.float32).ravel()],
                    dtype=np.float32).reshape(np.shape(a))


def _assemble_mixed(i0, i1, w, e, round_grad):
    """The level system from FP32 images (already bfloat16 values) with the
    gradient formed in FP32 and (optionally) rounded to binary16, the rest
    in FP64."""
    import scipy.sparse as sp
    T, Z, Y, X = i0.shape
    dxf = np.float32(1.0 / (2.0 * float(np.float32(1.0 / X))))
    ints = [(t, z, y, x) for t in range(1, T - 1) for z in range(1, Z - 1)
            for y in range(1, Y - 1) for x in range(1, X - 1)]
    idx = {c: k for k, c in enumerate(ints)}
    NB = [(-1, 0, 0, 0), (0, -1, 0, 0), (0, 0, -1, 0), (0, 0, 0, -1),
          (0, 0, 0, 1), (0, 0, 1, 0), (0, 1, 0, 0), (1, 0, 0, 0)]
    rows, cols, vals = [], [], []
    b = np.zeros(4 * len(ints))
    f16 = (lambda v: np.float32(np.float16(v))) if round_grad else (lambda v: v)
    for k, (t, z, y, x) in enumerate(ints):
        g = np.array([f16((i1[t, z, y, x + 1] - i1[t, z, y, x - 1]) * dxf),
                      f16((i1[t, z, y + 1, x] - i1[t, z, y - 1, x]) * dxf),
                      f16((i1[t, z + 1, y, x] - i1[t, z - 1, y, x]) * dxf),
                      f16((i1[t + 1, z, y, x] - i1[t - 1, z, y, x]) * dxf)], np.float64)
        It = float(np.float32(i1[t, z, y, x] - i0[t, z, y, x]))
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
    return sp.csr_matrix((vals, (rows, cols)), shape=(4 * len(ints),) * 2), b, ints


def test_mixed_level_solve_is_the_rounded_system():
    """One level in the mixed-precision form: the FP32 oracle solves the
    system of the bfloat16 images with binary16 gradients (spsolve of that
    system agrees to FP32 accuracy, the unrounded system does not)."""
    import scipy.sparse.linalg as spl
    from ofblend_amd.abi import default_params
    from test_oracle_solve import scatter
    N, T = 8, 7
    i0 = O.drop_sdf(N, N, N, T, seed=21, nspheres=2, rmin=0.08, rmax=0.18)
    i1 = O.drop_sdf(N, N, N, T, seed=22, nspheres=2, rmin=0.08, rmax=0.18)
    w, e = 0.1, 1e-4
    p = default_params(wSmooth=w, wEnergy=e, cgAccuracy=1e-9, cgMaxIter=20000,
                       precision=2)
    v, it, _, st = O.optical_flow4d(np.zeros(i0.shape + (4,), np.float32), i0, i1, p)
    assert st == 0
    b0, b1 = _bf16(i0), _bf16(i1)
    A, b, ints = _assemble_mixed(b0, b1, w, e, True)
    ref = scatter(spl.spsolve(A.tocsc(), b), ints, i0.shape) * N
    A2, b2, _ = _assemble_mixed(i0, i1, w, e, False)
    ref_unrounded = scatter(spl.spsolve(A2.tocsc(), b2), ints, i0.shape) * N
    scale = np.abs(ref).max()
    d = np.abs(v - ref).max()
    d_un = np.abs(ref_unrounded - ref).max()
    assert d <= 2e-4 * scale, (d, scale)
    assert d_un > 20 * d, (d_un, d)
