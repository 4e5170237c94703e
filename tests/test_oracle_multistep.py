"""Pins the oracle's multiscale driver with several optical-flow solves per
level (multiStep > 1, optflow4d.cpp:1031-1116), its Gaussian blur decay,
border band and the final projection (optflow4d.cpp:1140-1150) without a GPU.

An independent numpy/scipy composition of the driver: every level solve is
the scipy.sparse system of test_oracle_solve.assemble solved by spsolve, the
warps are map_coordinates backtraces, the blur is the separable Gaussian, the
error metric is a brute-force calcLsDiff and the projection is the numpy
restatement of test_oracle_project.  It follows the reference's steps: the
n-th solve of a level gets blur max(2, 3/4 of the previous one), earlier
increments are re-advected by the later ones before they are summed with the
incoming flow, i0 is warped by that sum, and the level stops (discarding the
last increment) once the sign-error ratio exceeds 0.95.  The cases are chosen
so that both a stopped level and a level that runs all its solves occur.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spl

from oracle import oracle as O
from ofblend_amd.abi import default_params
from test_oracle_solve import (assemble, np_advect, np_neumann0, np_resample,
                               scatter)
from test_oracle_project import np_corr_vels, np_gauss


def np_ls_diff(a, b, corr, bnd):
    sl = (slice(bnd, a.shape[0] - bnd),) * 4 if bnd else (slice(None),) * 4
    aa, bb = a[sl], b[sl]
    mism = ((aa < 0) & (bb >= 0)) | ((aa >= 0) & (bb < 0))
    d = np.minimum(np.abs(aa - bb) * corr, 1.0)
    return d[mism].sum() * 1e6 / aa.size


def np_advect_vec(vel, a):
    return np.stack([np_advect(vel, a[..., c]) for c in range(4)], -1)


def np_level(i0w, i1, w, e, blur, rbw):
    """opticalFlowDim (optflow4d.cpp:361-553) from a zero flow: solve, copy
    back as U / h, blur with sigma blur / 2, zero the band of width
    int(rbw X) + 1."""
    X = i0w.shape[3]
    A, b, ints = assemble(i0w, i1, w, e)
    v = scatter(spl.spsolve(A.tocsc(), b), ints, i0w.shape) * X
    if blur > 0:
        v = np_gauss(v, 0.5 * blur)
    if rbw > 0:
        rb = int(rbw * X) + 1
        inside = (slice(rb, -rb),) * 4
        keep = v[inside].copy()
        v[:] = 0.0
        v[inside] = keep
    return v


def np_multiscale_ms(vel, i0, i1, prm, level=0, record=None):
    T, Z, Y, X = i0.shape
    w, e, rbw = prm["wSmooth"], prm["wEnergy"], prm["resetBndWidth"]
    rb = int(rbw * X) + 1 if rbw > 0 else 0
    corr = 0.1 / 20
    if X > prm["minGridSize"]:
        small = (T // 2, Z // 2, Y // 2, X // 2)
        vs = np_resample(vel, small, 2.0) * 0.5
        vs = np_multiscale_ms(vs, np_resample(i0, small, 2.0),
                              np_resample(i1, small, 2.0), prm, level + 1, record)
        vel = np_resample(vs, i0.shape, 0.5) * 2.0
    i0w = np_neumann0(np_advect(vel, i0))
    err_last = np_ls_diff(i0w, i1, corr, rb)
    blur = prm["postVelBlur"]
    steps = prm["multiStep"]
    incs = []
    of = 0
    while of < steps:
        incs.append(np_level(i0w, i1, w, e, blur, rbw))
        blur = max(2.0, blur * 0.75)
        aligned = [v.copy() for v in incs]
        for k in range(of - 1, -1, -1):
            for l in range(k):
                aligned[l] = np_advect_vec(aligned[k], aligned[l])
        tmp = vel.copy()
        for k in range(of, -1, -1):
            tmp = tmp + aligned[k]
        i0w = np_neumann0(np_advect(tmp, i0))
        err = np_ls_diff(i0w, i1, corr, rb)
        with np.errstate(divide="ignore", invalid="ignore"):
            ratio = np.float64(err) / np.float64(err_last)
        if of > 0 and ratio > 0.95:
            incs[of] = np.zeros_like(incs[of])
            steps = of + 1
        err_last = err
        of += 1
    if record is not None:
        record[level] = steps
    for of in range(steps - 1, -1, -1):
        for l in range(of):
            incs[l] = np_advect_vec(incs[of], incs[l])
    for of in range(steps):
        vel = vel + incs[of]
    if level == 0 and prm["doFinalProject"]:
        _, vel = np_corr_vels(vel, i0, i1, 4.0, 4.0, rbw, 40)
    if level == 0:
        i0w = np_neumann0(np_advect(vel, i0))
        err_last = np_ls_diff(i0w, i1, corr, rb)
    return (vel, err_last) if level == 0 else vel


CASES = [
    # seeds, multiStep, postVelBlur, doFinalProject: solves run per level
    ((1, 2), 3, 0.0, 0),  # (2, 3): the fine level stops after its 2nd solve
    ((3, 4), 3, 4.0, 0),  # (3, 2): the coarse level stops
    ((1, 2), 3, 4.0, 1),  # flof.py's settings (multiStep 3, blur 4, band 0.1,
                          # final projection)
    ((1, 2), 5, 1.0, 0),  # (5, 5): from the 4th solve on, i0 is warped by
                          # re-advected earlier increments
]


@pytest.mark.parametrize("seeds,ms,blur,proj", CASES)
def test_multistep_driver_matches_numpy_composition(seeds, ms, blur, proj):
    N, T = 8, 8
    i0 = O.drop_sdf(N, N, N, T, seed=seeds[0], dtype=np.float64)
    i1 = O.drop_sdf(N, N, N, T, seed=seeds[1], dtype=np.float64)
    prm = dict(wSmooth=0.1, wEnergy=1e-4, minGridSize=4, multiStep=ms,
               postVelBlur=blur, resetBndWidth=0.1, doFinalProject=proj)
    steps = {}
    ref, ref_err = np_multiscale_ms(np.zeros(i0.shape + (4,)), i0, i1, prm,
                                    record=steps)
    p = default_params(cgAccuracy=1e-13, cgMaxIter=20000, **prm)
    vel, err, info = O.multiscale4d(np.zeros(i0.shape + (4,)), i0, i1, p)
    s = info.summary()
    got_steps = {k: len(l["iters"]) for k, l in enumerate(s["levels"])}
    assert got_steps == steps, (got_steps, steps)
    assert np.abs(ref).max() > 0.1
    np.testing.assert_allclose(vel, ref, atol=1e-8 * np.abs(ref).max(), rtol=0)
    assert abs(err - ref_err) <= 1e-8 * max(1.0, ref_err)


def test_multistep_cases_cover_both_outcomes():
    """The cases above include a level that stops early and one that runs
    every solve (the early-stop rule is exercised both ways)."""
    outcomes = set()
    for seeds, ms, blur, proj in CASES[:2]:
        N, T = 8, 8
        i0 = O.drop_sdf(N, N, N, T, seed=seeds[0], dtype=np.float64)
        i1 = O.drop_sdf(N, N, N, T, seed=seeds[1], dtype=np.float64)
        p = default_params(wSmooth=0.1, wEnergy=1e-4, cgAccuracy=1e-13,
                           cgMaxIter=20000, minGridSize=4, multiStep=3,
                           postVelBlur=blur, resetBndWidth=0.1)
        _, _, info = O.multiscale4d(np.zeros(i0.shape + (4,)), i0, i1, p)
        for l in info.summary()["levels"]:
            outcomes.add(len(l["iters"]) == 3)
    assert outcomes == {True, False}
