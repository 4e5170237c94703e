"""Pins the CPU restatement of the warm-start network (no GPU): its backward
pass against central finite differences of its forward pass (exact
arithmetic), its convolutions against a direct loop, and its bfloat16
rounding against the definition."""
import numpy as np

from oracle import warmnet as W

RNG = np.random.default_rng(21)


def _params(scale=0.3):
    P = RNG.standard_normal(W.NPARAM) * scale
    return P


def test_conv_matches_direct_loop():
    x = RNG.standard_normal((3, 4, 3, 5, 16))
    Wt = RNG.standard_normal((27, 16, 7))
    got = W.conv(x, Wt, True, rnd=False)
    T, Z, Y, X, _ = x.shape
    ref = np.zeros((T, Z, Y, X, 7))
    for t in range(T):
        for z in range(Z):
            for y in range(Y):
                for xx in range(X):
                    for m in range(27):
                        dx, dy, dz = m % 3 - 1, (m // 3) % 3 - 1, m // 9 - 1
                        zz, yy, xs = z + dz, y + dy, xx + dx
                        if 0 <= zz < Z and 0 <= yy < Y and 0 <= xs < X:
                            ref[t, z, y, xx] += x[t, zz, yy, xs] @ Wt[m]
    np.testing.assert_allclose(got, ref, atol=1e-12)
    for dil in (1, 2):
        got_t = W.conv(x, Wt[:3], False, sgn=-1, rnd=False, dil=dil)
        ref_t = np.zeros_like(got_t)
        for m in range(3):
            dt = -(m - 1) * dil
            for t in range(T):
                if 0 <= t + dt < T:
                    ref_t[t] += x[t + dt] @ Wt[m]
        np.testing.assert_allclose(got_t, ref_t, atol=1e-12)
    # dilation 2 in space equals the dense conv of a 5x5x5 kernel with
    # zeros between the taps
    got2 = W.conv(x, Wt, True, rnd=False, dil=2)
    ref2 = np.zeros_like(got2)
    for m in range(27):
        off = (0, 2 * (m // 9 - 1), 2 * ((m // 3) % 3 - 1), 2 * (m % 3 - 1))
        for t in range(T):
            for z in range(Z):
                for y in range(Y):
                    for xx in range(X):
                        zz, yy, xs = z + off[1], y + off[2], xx + off[3]
                        if 0 <= zz < Z and 0 <= yy < Y and 0 <= xs < X:
                            ref2[t, z, y, xx] += x[t, zz, yy, xs] @ Wt[m]
    np.testing.assert_allclose(got2, ref2, atol=1e-12)


def test_backward_matches_finite_differences():
    shape = (5, 5, 4, 9)
    i0 = RNG.standard_normal(shape) * 2
    i1 = RNG.standard_normal(shape) * 2
    tgt = RNG.standard_normal(shape + (4,))
    P = _params()
    loss, g = W.loss_grad(P, i0, i1, tgt, rnd=False)
    idx = np.concatenate([
        RNG.choice(W.NPARAM, 40, replace=False),
        [W.OFF_WL + 3, W.OFF_BL + 5, W.OFF_WH + 7, W.OFF_BH + 2,
         W.OFF_BLK + W.BLK_BS + 1, W.OFF_BLK + 3 * W.BLK_SIZE + W.BLK_BT + 4,
         W.OFF_BLK + 2 * W.BLK_SIZE + W.BLK_WT + 100]])
    for k in idx:
        e = np.zeros(W.NPARAM)
        e[k] = 1e-6
        lp, _ = W.loss_grad(P + e, i0, i1, tgt, rnd=False)
        lm, _ = W.loss_grad(P - e, i0, i1, tgt, rnd=False)
        fd = (lp - lm) / 2e-6
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.001953125, 1.005859375, -3.0e-3, 65504.0, 0.0],
                 np.float32)
    want = [1.0, 1.0, 1.0, 1.0078125]  # 1+2^-8 is a tie: to even (1.0)
    got = W.bf16(x)
    assert list(got[:4]) == want       # 8 significant bits, ties to even
    assert got[-1] == 0.0
    m, e = np.frexp(abs(got[4]))
    assert m * 256 == np.round(m * 256)
