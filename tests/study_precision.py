"""Accuracy study of the GPU solve against the FP64 CPU oracle (a script, not
a pytest test; the oracle is the checker):

    python tests/study_precision.py

A 2-sphere drop pair (seeds 7, 8; radii 0.08N-0.18N) at 32^3 x 32, P_plain,
3 levels, every level to relative residual 1e-5: the GPU flow with FP32 and
with bfloat16 direction storage against the FP64 oracle of the same
algorithm (the reference's DOUBLEPRECISION build, serial sums).  Prints one
JSON line.
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402  (the checker)
from ofblend_amd import manta as M  # noqa: E402
from ofblend_amd.abi import FLOF_PSTORE_BF16, default_params  # noqa: E402

PARAMS = dict(wSmooth=0.1, wEnergy=1e-4, cgAccuracy=1e-5, multiStep=1, postVelBlur=0.0,
              doFinalProject=False, resetBndWidth=-1.0, cfl=999.0)


def main(n=32, t=32):
    sv = M.Solver(gridSize=M.vec3(n, n, n), fourthDim=t)
    a, b = sv.create(M.Grid4Real), sv.create(M.Grid4Real)
    M.dropSdf4d(a, seed=7, nspheres=2, rmin=0.08, rmax=0.18)
    M.dropSdf4d(b, seed=8, nspheres=2, rmin=0.08, rmax=0.18)
    v = sv.create(M.Grid4Vec4)
    prm = dict(PARAMS, minGridSize=n // 4, cgMaxIter=20000)
    res = {"workload": "2-sphere drop pair %d^3 x %d (seeds 7, 8), P_plain, residual 1e-5"
                       % (n, t)}
    flows = {}
    for name, prec in (("fp32", 0), ("bf16", FLOF_PSTORE_BF16)):
        v.clear()
        e = M.opticalFlowMultiscale4d(vel=v, i0=a, i1=b, precision=prec, **prm)
        flows[name] = v.to_numpy()
        res[name + "_final_err"] = e
    ha, hb = a.to_numpy().astype(np.float64), b.to_numpy().astype(np.float64)
    t0 = time.perf_counter()
    ref, e64, _ = O.multiscale4d(np.zeros(flows["fp32"].shape), ha, hb, default_params(**prm))
    res["fp64_oracle_s"] = time.perf_counter() - t0
    res["fp64_final_err"] = e64
    for name in ("fp32", "bf16"):
        d = flows[name] - ref
        res[name + "_vs_fp64"] = {"max": float(np.abs(d).max()),
                                  "rms": float(np.sqrt((d ** 2).mean()))}
    res["flow_abs_max_fp64"] = float(np.abs(ref).max())
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
