"""CPU restatement of the learned warm-start network (test infrastructure;
only the tests may import it).

The network has no counterpart in the reference: it is this library's
extension that predicts a warm start for the multiscale solve.  This module
states its forward pass, loss and backward pass in float64 numpy from the
definitions (pointwise lift 2->16, four residual blocks of a 3x3x3 space
convolution, ReLU and a 3-tap time convolution with 16 channels, pointwise
head 16->4, mean squared error), with the operands of every convolution
rounded to bfloat16 as the matrix-core kernels do (rnd=True), or exact
(rnd=False: the form the finite-difference check pins).

Parameter layout (flat float32, identical to the library):
  Wl[2][16], bl[16], then per block Ws[27][16][16] (tap, in, out), bs[16],
  Wt[3][16][16], bt[16], then Wh[16][4], bh[4].
Space tap m: (dx, dy, dz) = (m % 3 - 1, (m // 3) % 3 - 1, m // 9 - 1);
time tap m: dt = m - 1; block b multiplies the offsets by 2^b (dilation).
Zero padding outside the grid.
"""
import numpy as np

C = 16
NB = 4
NS = 27
NT = 3
OFF_WL = 0
OFF_BL = 2 * C
OFF_BLK = OFF_BL + C
BLK_WS = 0
BLK_BS = BLK_WS + NS * C * C
BLK_WT = BLK_BS + C
BLK_BT = BLK_WT + NT * C * C
BLK_SIZE = BLK_BT + C
OFF_WH = OFF_BLK + NB * BLK_SIZE
OFF_BH = OFF_WH + C * 4
NPARAM = OFF_BH + 4


def bf16(a):
    """Round to bfloat16 (nearest even) through float32; returns float64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(P):
    P = np.asarray(P, dtype=np.float64)
    d = {"Wl": P[OFF_WL:OFF_WL + 2 * C].reshape(2, C),
         "bl": P[OFF_BL:OFF_BL + C], "blocks": []}
    for b in range(NB):
        o = OFF_BLK + b * BLK_SIZE
        d["blocks"].append({
            "Ws": P[o + BLK_WS:o + BLK_BS].reshape(NS, C, C),
            "bs": P[o + BLK_BS:o + BLK_WT],
            "Wt": P[o + BLK_WT:o + BLK_BT].reshape(NT, C, C),
            "bt": P[o + BLK_BT:o + BLK_SIZE]})
    d["Wh"] = P[OFF_WH:OFF_WH + C * 4].reshape(C, 4)
    d["bh"] = P[OFF_BH:OFF_BH + 4]
    return d


def pack(d):
    P = np.zeros(NPARAM)
    P[OFF_WL:OFF_WL + 2 * C] = d["Wl"].ravel()
    P[OFF_BL:OFF_BL + C] = d["bl"]
    for b, blk in enumerate(d["blocks"]):
        o = OFF_BLK + b * BLK_SIZE
        P[o + BLK_WS:o + BLK_BS] = blk["Ws"].ravel()
        P[o + BLK_BS:o + BLK_WT] = blk["bs"]
        P[o + BLK_WT:o + BLK_BT] = blk["Wt"].ravel()
        P[o + BLK_BT:o + BLK_SIZE] = blk["bt"]
    P[OFF_WH:OFF_WH + C * 4] = d["Wh"].ravel()
    P[OFF_BH:OFF_BH + 4] = d["bh"]
    return P


def taps(space, dil=1):
    """Offsets (dt, dz, dy, dx) of the taps in tap order."""
    if space:
        return [(0, dil * (m // 9 - 1), dil * ((m // 3) % 3 - 1), dil * (m % 3 - 1))
                for m in range(NS)]
    return [(dil * (m - 1), 0, 0, 0) for m in range(NT)]


def shift(a, off):
    """s[c] = a[c + off] with zeros outside; a is (T, Z, Y, X, ch)."""
    out = np.zeros_like(a)
    if any(abs(d) >= n for d, n in zip(off, a.shape[:4])):
        return out
    src = []
    dst = []
    for d, n in zip(off, a.shape[:4]):
        if d >= 0:
            src.append(slice(d, n))
            dst.append(slice(0, n - d))
        else:
            src.append(slice(0, n + d))
            dst.append(slice(-d, n))
    out[tuple(dst)] = a[tuple(src)]
    return out


def conv(x, W, space, sgn=1, rnd=True, dil=1):
    """out[c] = sum_tap x[c + sgn*off(tap)] @ W[tap] (W[tap] is in x out)."""
    xr = bf16(x) if rnd else x
    Wr = bf16(W) if rnd else W
    out = np.zeros(x.shape[:4] + (W.shape[2],))
    for m, off in enumerate(taps(space, dil)):
        out += shift(xr, tuple(sgn * o for o in off)) @ Wr[m]
    return out


def wgrad(x, dout, space, rnd=True, dil=1):
    """dW[tap] = sum_c x[c + off(tap)]^T dout[c]."""
    xr = bf16(x) if rnd else x
    dr = bf16(dout) if rnd else dout
    n = x.shape[4]
    g = np.zeros((NS if space else NT, n, dout.shape[4]))
    d2 = dr.reshape(-1, dr.shape[4])
    for m, off in enumerate(taps(space, dil)):
        g[m] = shift(xr, off).reshape(-1, n).T @ d2
    return g


def forward(P, i0, i1, rnd=True, keep=False):
    """i0, i1: (T, Z, Y, X) SDFs; returns the flow (T, Z, Y, X, 4)."""
    d = unpack(P)
    x = np.stack([np.asarray(i0, np.float64) * 0.25, np.asarray(i1, np.float64) * 0.25], -1)
    h = x @ d["Wl"] + d["bl"]
    hs, ss, as_ = [h], [], []
    for b, blk in enumerate(d["blocks"]):
        s = conv(h, blk["Ws"], True, rnd=rnd, dil=2 ** b) + blk["bs"]
        a = np.maximum(s, 0.0)
        h = h + (conv(a, blk["Wt"], False, rnd=rnd, dil=2 ** b) + blk["bt"])
        ss.append(s)
        as_.append(a)
        hs.append(h)
    y = h @ d["Wh"] + d["bh"]
    if keep:
        return y, (x, hs, ss, as_)
    return y


def loss_grad(P, i0, i1, tgt, rnd=True):
    """Mean squared error over all 4 components and its gradient (flat)."""
    d = unpack(P)
    y, (x, hs, ss, as_) = forward(P, i0, i1, rnd=rnd, keep=True)
    e = y - tgt
    n4 = e.size
    loss = float((e ** 2).sum() / n4)
    dy = 2.0 * e / n4
    g = {"Wh": hs[-1].reshape(-1, C).T @ dy.reshape(-1, 4), "bh": dy.reshape(-1, 4).sum(0),
         "blocks": [None] * NB}
    dh = dy @ d["Wh"].T
    for b in range(NB - 1, -1, -1):
        blk = d["blocks"][b]
        gb = {}
        dil = 2 ** b
        gb["Wt"] = wgrad(as_[b], dh, False, rnd=rnd, dil=dil)
        gb["bt"] = dh.reshape(-1, C).sum(0)
        Wt_T = np.transpose(blk["Wt"], (0, 2, 1))
        ds = conv(dh, Wt_T, False, sgn=-1, rnd=rnd, dil=dil) * (ss[b] > 0)
        gb["Ws"] = wgrad(hs[b], ds, True, rnd=rnd, dil=dil)
        gb["bs"] = ds.reshape(-1, C).sum(0)
        Ws_T = np.transpose(blk["Ws"], (0, 2, 1))
        dh = dh + conv(ds, Ws_T, True, sgn=-1, rnd=rnd, dil=dil)
        g["blocks"][b] = gb
    g["Wl"] = x.reshape(-1, 2).T @ dh.reshape(-1, C)
    g["bl"] = dh.reshape(-1, C).sum(0)
    return loss, pack(g)
