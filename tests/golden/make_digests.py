"""Writes tests/golden/digests.json: SHA-256 digests and convergence records
of the CPU oracle's FlOF solves of the seeded drop pairs (seeds 1, 2), for
the GPU parity tests at sizes whose oracle run takes minutes: drop32_l3
(32^4, 3 levels, minGridSize 8: ~45 s per solve) and drop64_l3_cap1000
(64^4, 3 levels, minGridSize 16, the reference's 1000-iteration cap: ~25 min
per solve on 6 cores).

Per case, three oracle solves of the same inputs:
  * FP32 with correctly rounded FP64 sums (sum_mode 1) -- the summation the
    GPU library uses, so a matching digest means a bit-identical flow field;
  * FP32 with the reference's serial FP64 sums (sum_mode 0) -- the
    reference's own operation order;
  * FP64 (the reference's DOUBLEPRECISION build) -- the accuracy yardstick.
and the max / RMS distances between them (in cells).  The frame warp of the
same drops (T frames at twice the spatial resolution placed 1:1 onto t,
alpha 1, defoFactor (2, 2, 2, 1), the loadAdvectTimeSlice_OptRun lookup into
zeroed frames) is recorded for the first, middle and last frame: the digest
of the three FP32 frames warped by the sum_mode-1 flow, and the distances of
the FP32 warps to the FP64 warp (FP64 frames, FP64 flow).

The mixed-precision case mixed32_l3 records the digest of the oracle's
mixed-precision flow (images stored as bfloat16, gradients of i1 rounded to
binary16, sum_mode 1) and its distance, and the FP32 flow's, to the FP64
flow.

    python tests/golden/make_digests.py [drop32_l3] [drop64_l3_cap1000]
        [mixed32_l3]
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402
from ofblend_amd.abi import default_params  # noqa: E402

CASES = {
    "drop32_l3": dict(N=32, minGridSize=8, cgMaxIter=1000),
    "drop64_l3_cap1000": dict(N=64, minGridSize=16, cgMaxIter=1000),
}
# the mixed-precision solve (bfloat16 images, binary16 gradients) on the
# mixed-precision configuration's generator: 2 spheres per grid, radii
# uniform in [0.08N, 0.18N], seeds 7 and 8, at a size the oracle covers
MIXED_CASES = {
    "mixed32_l3": dict(N=32, T=32, minGridSize=8, cgMaxIter=20000, seeds=(7, 8)),
}
WARP_SCALE = 2


def dist(a, b):
    d = np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64))
    return {"max_abs_diff": float(d.max()), "rms_diff": float(np.sqrt((d ** 2).mean()))}


def warp_frames(N, flow, dtype):
    """The first, middle and last frame of the frame warp: T frames at twice
    the resolution, placed 1:1 onto t, warped by the flow."""
    T = N
    frames = O.drop_sdf(N, N, N, T, seed=1, scale=WARP_SCALE, dtype=dtype)
    F = WARP_SCALE * N
    phi = O.place_frames((T, F, F, F), frames)
    out = []
    for t in (0, T // 2 - 1, T - 1):
        out.append(O.warp_slice(flow.astype(dtype), phi, np.zeros((F, F, F), dtype),
                                float(t), blendAlpha=1.0, defoFactor=(2, 2, 2, 1)))
    return np.stack(out)


def run(name):
    c = CASES[name]
    N = c["N"]
    prm = dict(O.P_PLAIN, minGridSize=c["minGridSize"], cgMaxIter=c["cgMaxIter"])
    p = default_params(**prm)
    out = {"config": c, "seeds": [1, 2], "params": dict(O.P_PLAIN)}
    i0 = O.drop_sdf(N, N, N, N, seed=1)
    i1 = O.drop_sdf(N, N, N, N, seed=2)
    t = time.time()
    with O.sum_mode(1):
        v1, err1, info1 = O.multiscale4d(np.zeros(i0.shape + (4,), np.float32),
                                         i0, i1, p)
    out["oracle_seconds"] = time.time() - t
    s = info1.summary()
    out["sha256"] = hashlib.sha256(v1.tobytes()).hexdigest()
    out["err"] = err1
    out["iters"] = [l["iters"][0] for l in s["levels"]]
    out["absmax"] = float(np.abs(v1).max())
    with O.sum_mode(0):
        v0, err0, info0 = O.multiscale4d(np.zeros(i0.shape + (4,), np.float32),
                                         i0, i1, p)
    out["serial_sum_oracle"] = dict(dist(v0, v1), err=err0,
                                    iters=[l["iters"][0] for l in info0.summary()["levels"]])
    d0 = O.drop_sdf(N, N, N, N, seed=1, dtype=np.float64)
    d1 = O.drop_sdf(N, N, N, N, seed=2, dtype=np.float64)
    with O.sum_mode(0):
        v64, err64, info64 = O.multiscale4d(np.zeros(d0.shape + (4,)), d0, d1, p)
    out["fp64_oracle"] = {"err": err64,
                          "iters": [l["iters"][0] for l in info64.summary()["levels"]],
                          "absmax": float(np.abs(v64).max())}
    out["distances"] = {"gpu_vs_fp64": dist(v1, v64), "serial_vs_fp64": dist(v0, v64),
                        "gpu_vs_serial": dist(v1, v0)}
    w1 = warp_frames(N, v1, np.float32)
    w0 = warp_frames(N, v0, np.float32)
    w64 = warp_frames(N, v64, np.float64)
    out["warp"] = {"frames": [0, N // 2 - 1, N - 1], "scale": WARP_SCALE,
                   "sha256": hashlib.sha256(w1.tobytes()).hexdigest(),
                   "gpu_vs_fp64": dist(w1, w64), "serial_vs_fp64": dist(w0, w64),
                   "gpu_vs_serial": dist(w1, w0)}
    return out


def run_mixed(name):
    c = MIXED_CASES[name]
    N, T = c["N"], c["T"]
    prm = dict(O.P_PLAIN, minGridSize=c["minGridSize"], cgMaxIter=c["cgMaxIter"])
    out = {"config": {k: v for k, v in c.items() if k != "seeds"}, "seeds": list(c["seeds"]),
           "params": dict(O.P_PLAIN), "spheres": 2, "radii": [0.08, 0.18]}
    gen = dict(nspheres=2, rmin=0.08, rmax=0.18)
    i0 = O.drop_sdf(N, N, N, T, seed=c["seeds"][0], **gen)
    i1 = O.drop_sdf(N, N, N, T, seed=c["seeds"][1], **gen)
    zero = np.zeros(i0.shape + (4,), np.float32)
    with O.sum_mode(1):
        vm, errm, infm = O.multiscale4d(zero, i0, i1, default_params(precision=2, **prm))
        v32, err32, inf32 = O.multiscale4d(zero, i0, i1, default_params(**prm))
    d0 = O.drop_sdf(N, N, N, T, seed=c["seeds"][0], dtype=np.float64, **gen)
    d1 = O.drop_sdf(N, N, N, T, seed=c["seeds"][1], dtype=np.float64, **gen)
    v64, err64, inf64 = O.multiscale4d(np.zeros(d0.shape + (4,)), d0, d1, default_params(**prm))
    it = lambda inf: [l["iters"][0] for l in inf.summary()["levels"]]
    out.update({"sha256": hashlib.sha256(vm.tobytes()).hexdigest(), "err": errm,
                "iters": it(infm), "iters_fp32": it(inf32), "iters_fp64": it(inf64),
                "err_fp32": err32, "err_fp64": err64, "absmax_fp64": float(np.abs(v64).max()),
                "distances": {"mixed_vs_fp64": dist(vm, v64), "fp32_vs_fp64": dist(v32, v64),
                              "mixed_vs_fp32": dist(vm, v32)}})
    return out


def main():
    names = [a for a in sys.argv[1:] if a in CASES or a in MIXED_CASES] or ["drop32_l3"]
    path = os.path.join(HERE, "digests.json")
    for n in names:
        print("computing", n, flush=True)
        res = run(n) if n in CASES else run_mixed(n)
        data = json.load(open(path)) if os.path.exists(path) else {}
        data[n] = res
        print(json.dumps(res), flush=True)
        json.dump(data, open(path, "w"), indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
