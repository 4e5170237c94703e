"""Learned warm start for opticalFlowMultiscale4d (an extension; the reference
always starts each level's CG from zero, source/plugin/optflow4d.cpp:284-303).

The network (libflof_hip's flof_warmnet_*) maps the two SDFs, restricted along
the solve's pyramid to a coarse level, to a 4D flow that is prolonged to the
finest level and passed to the solve as `warm`.  The solve keeps the warm
start per level only if its residual norm is not larger than the zero
start's, so a poor prediction costs at most the check.

Training data are flows that the plain GPU solve computes for seeded drop
pairs; training runs on the GPU with the library's own forward and backward
kernels and Adam.
"""
import ctypes as C
import os

import numpy as np

from . import _lib
from . import manta as M
from ._lib import FlofError, check

NPARAM = None


def num_params():
    global NPARAM
    if NPARAM is None:
        NPARAM = _lib.lib().flof_warmnet_num_params()
    return NPARAM


class WarmNet:
    """Handle of one network on the default context's GPU."""

    def __init__(self, seed=1, ctx=None):
        self.ctx = ctx if ctx is not None else M.context()
        h = C.c_void_p()
        check(_lib.lib().flof_warmnet_create(self.ctx, int(seed), C.byref(h)), self.ctx)
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            try:
                _lib.lib().flof_warmnet_destroy(h)
            except Exception:
                pass
            self.h = None

    def get_params(self):
        a = np.empty(num_params(), np.float32)
        check(_lib.lib().flof_warmnet_get_params(
            self.h, a.ctypes.data_as(C.POINTER(C.c_float)), a.size), self.ctx)
        return a

    def set_params(self, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        if a.size != num_params():
            raise FlofError(-1, "warm-start net: %d parameters expected" % num_params())
        check(_lib.lib().flof_warmnet_set_params(
            self.h, a.ctypes.data_as(C.POINTER(C.c_float)), a.size), self.ctx)

    def get_grads(self):
        a = np.empty(num_params(), np.float32)
        check(_lib.lib().flof_warmnet_get_grads(
            self.h, a.ctypes.data_as(C.POINTER(C.c_float)), a.size), self.ctx)
        return a

    def forward(self, i0, i1, flow):
        check(_lib.lib().flof_warmnet_forward(self.h, i0.h, i1.h, flow.h), self.ctx)

    def loss_grad(self, i0, i1, target):
        loss = C.c_double()
        check(_lib.lib().flof_warmnet_loss_grad(self.h, i0.h, i1.h, target.h,
                                                C.byref(loss)), self.ctx)
        return loss.value

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        check(_lib.lib().flof_warmnet_adam(self.h, lr, beta1, beta2, eps), self.ctx)

    def warm_start(self, i0, i1, warm, netX=32):
        check(_lib.lib().flof_warmnet_warm_start(self.h, i0.h, i1.h, int(netX), warm.h),
              self.ctx)

    def solve(self, vel, i0, i1, levels=1, **kw):
        """opticalFlowMultiscale4d with this network's warm start on the
        `levels` finest levels (-1: all); returns (err, SolveInfo)."""
        p = _params_of(kw)
        info = M.SolveInfo()
        err = C.c_float()
        check(_lib.lib().flof_optical_flow_multiscale4d_net(
            vel.h, i0.h, i1.h, C.byref(p), self.h, int(levels), C.byref(info),
            C.byref(err)), self.ctx)
        return err.value, info

    def save(self, path):
        np.save(path, self.get_params(), allow_pickle=False)

    def load(self, path):
        self.set_params(np.load(path, allow_pickle=False))


def _params_of(kw):
    kw = dict(kw)
    if "doFinalProject" in kw:
        kw["doFinalProject"] = int(kw["doFinalProject"])
    return M.default_params(**kw)


def capture(vel, i0, i1, level=0, **kw):
    """opticalFlowMultiscale4d that also returns the pre-warped i0 of pyramid
    level `level` and the increment its solve adds (new grids of that
    level's size): (err, info, i0w, inc)."""
    X, Y, Z, T = vel.size
    for _ in range(level):
        X, Y, Z, T = X // 2, Y // 2, Z // 2, T // 2
    i0w = M.Grid4Real(vel.parent, (X, Y, Z, T))
    inc = M.Grid4Vec4(vel.parent, (X, Y, Z, T))
    p = _params_of(kw)
    info = M.SolveInfo()
    err = C.c_float()
    check(_lib.lib().flof_optical_flow_multiscale4d_capture(
        vel.h, i0.h, i1.h, C.byref(p), int(level), i0w.h, inc.h, C.byref(info),
        C.byref(err)), vel.ctx)
    return err.value, info, i0w, inc


PLAIN = dict(wSmooth=0.1, wEnergy=1e-4, cgAccuracy=1e-5, multiStep=1, postVelBlur=0.0,
             doFinalProject=False, resetBndWidth=-1.0, cfl=999.0)


def drop_pair(solver, seed, nspheres=1, rmin=0.15, rmax=0.15):
    """Two seeded drop SDF grids (seeds seed, seed+1)."""
    a, b = solver.create(M.Grid4Real), solver.create(M.Grid4Real)
    M.dropSdf4d(a, seed=seed, nspheres=nspheres, rmin=rmin, rmax=rmax)
    M.dropSdf4d(b, seed=seed + 1, nspheres=nspheres, rmin=rmin, rmax=rmax)
    return a, b


def generate(n, N, T, seed0=1000, minGridSize=8, **kw):
    """n training samples (i0, i1, flow) at N^3 x T: drop pairs of 1-3 spheres
    with radii in [0.08N, 0.18N], flows from the plain multiscale solve."""
    s = M.Solver(gridSize=M.vec3(N, N, N), fourthDim=T)
    prm = dict(PLAIN, minGridSize=minGridSize)
    prm.update(kw)
    out = []
    for k in range(n):
        i0, i1 = drop_pair(s, seed0 + 2 * k, nspheres=1 + k % 3, rmin=0.08, rmax=0.18)
        v = s.create(M.Grid4Vec4)
        M.opticalFlowMultiscale4d(vel=v, i0=i0, i1=i1, **prm)
        out.append((i0, i1, v))
    return s, out


def train(net, data, steps, lr=2e-3, seed=0, log=None, lr_end=None):
    """Adam on single samples drawn from data (seeded); cosine decay of the
    learning rate to lr_end (default lr/10).  Returns the loss per step."""
    rng = np.random.default_rng(seed)
    lr_end = lr / 10 if lr_end is None else lr_end
    losses = []
    for st in range(steps):
        i0, i1, v = data[rng.integers(len(data))]
        losses.append(net.loss_grad(i0, i1, v))
        f = 0.5 * (1 + np.cos(np.pi * st / max(1, steps - 1)))
        net.adam(lr=lr_end + (lr - lr_end) * f)
        if log and (st % 50 == 0 or st == steps - 1):
            log("warmnet step %d loss %.5f" % (st, losses[-1]))
    return losses


DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                               "warmnet_32.npy")
# the network trained on the finest-level increments of 2000 plain 64^4
# solves (python -m ofblend_amd.train_warmnet), used by WarmNet.solve
NET64_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                             "warmnet_64.npy")
