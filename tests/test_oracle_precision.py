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
