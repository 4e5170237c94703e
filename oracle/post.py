"""CPU restatement (numpy, float32 arithmetic) of the output post-processing
of warped 3D frames.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker; the product path
(ofblend_amd) never imports it.

A frame is a (Z, Y, X) float32 array; Z == 1 is a 2D grid (is3D() false).

  frame_join (source/levelset.cpp:114-118, KnJoin: a = min(a, b) with
      std::min, i.e. b < a ? b : a) -- elementwise closed form
  frame_mult_const / frame_add_scaled (source/grid.h:735-743: a *= s,
      a += s * b) -- elementwise closed forms
  simple_blur_special (source/test.cpp:94-127) -- constant and linear fields
      are fixed points inside, values below the threshold and cells outside
      the band stay, the outer layer alternates with the zeroed temporary
      [test_oracle_post.py]
  set_bound3 (source/grid.cpp:462-473) -- brute force [test_oracle_post.py]
"""
import numpy as np


def frame_join(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return np.where(b < a, b, a).astype(np.float32)


def frame_mult_const(a, s):
    return (np.asarray(a, np.float32) * np.float32(s)).astype(np.float32)


def frame_add_scaled(a, b, s):
    return (np.asarray(a, np.float32) +
            np.float32(s) * np.asarray(b, np.float32)).astype(np.float32)


def _blur_pass(a, dst, thresh, bord):
    """One knSimpleBlurSpecial pass (KERNEL bnd = 1) from a into dst."""
    Z, Y, X = a.shape
    is3d = Z > 1
    out = dst.copy()
    ks = range(1, Z - 1) if is3d else [0]
    f = np.float32
    for k in ks:
        for j in range(1, Y - 1):
            for i in range(1, X - 1):
                inb = i >= bord and j >= bord and i < X - bord and j < Y - bord
                inb = inb and ((k >= bord and k < Z - bord) if is3d else k == 0)
                v = a[k, j, i]
                if not inb or v < f(thresh):
                    out[k, j, i] = v
                    continue
                s = f(f(f(a[k, j, i + 1] + a[k, j, i - 1]) + a[k, j + 1, i]) + a[k, j - 1, i])
                if is3d:
                    s = f(s + f(a[k + 1, j, i] + a[k - 1, j, i]))
                    s = f(np.float64(s) * (1.0 / 6.0))
                else:
                    s = f(np.float64(s) * (1.0 / 4.0))
                out[k, j, i] = s
    return out


def simple_blur_special(a, iter=1, thresh=0.0, bord=0):
    """simpleBlurSpecial: the temporary grid starts zeroed and is swapped with
    the grid after every pass (its untouched outer layer comes along)."""
    cur = np.array(a, np.float32, copy=True)
    tmp = np.zeros_like(cur)
    for _ in range(iter):
        new = _blur_pass(cur, tmp, thresh, bord)
        tmp = cur
        cur = new
    return cur


def set_bound3(a, value, w):
    a = np.array(a, np.float32, copy=True)
    Z, Y, X = a.shape
    for k in range(Z):
        for j in range(Y):
            for i in range(X):
                bnd = (i <= w or i >= X - 1 - w or j <= w or j >= Y - 1 - w or
                       (Z > 1 and (k <= w or k >= Z - 1 - w)))
                if bnd:
                    a[k, j, i] = np.float32(value)
    return a
