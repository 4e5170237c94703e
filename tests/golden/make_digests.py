"""Writes tests/golden/digests.json: SHA-256 digests and convergence records
of the CPU oracle's FlOF solves of the seeded drop pairs (seeds 1, 2), for
the GPU parity tests at sizes whose oracle run takes minutes: drop32_l3
(32^4, 3 levels, minGridSize 8: ~45 s) and drop64_l3_cap1000 (64^4, 3
levels, minGridSize 16, the reference's 1000-iteration cap: ~25 min on 6
cores).

The oracle runs with correctly rounded FP64 sums (sum_mode 1), the summation
the GPU library uses, so a matching digest means a bit-identical flow field.
The serial-sum (reference order) oracle's distance to it is recorded too.

    python tests/golden/make_digests.py [drop32_l3] [drop64_l3_cap1000]
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


def run(name):
    c = CASES[name]
    N = c["N"]
    i0 = O.drop_sdf(N, N, N, N, seed=1)
    i1 = O.drop_sdf(N, N, N, N, seed=2)
    p = default_params(**O.P_PLAIN, minGridSize=c["minGridSize"],
                       cgMaxIter=c["cgMaxIter"])
    out = {"config": c, "seeds": [1, 2], "params": dict(O.P_PLAIN)}
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
    v0, err0, _ = O.multiscale4d(np.zeros(i0.shape + (4,), np.float32), i0, i1, p)
    d = np.abs(v0.astype(np.float64) - v1)
    out["serial_sum_oracle"] = {"max_abs_diff": float(d.max()),
                                "rms_diff": float(np.sqrt((d ** 2).mean())),
                                "err": err0}
    return out


def main():
    names = [a for a in sys.argv[1:] if a in CASES] or ["drop32_l3"]
    path = os.path.join(HERE, "digests.json")
    data = json.load(open(path)) if os.path.exists(path) else {}
    for n in names:
        print("computing", n, flush=True)
        data[n] = run(n)
        print(json.dumps(data[n]), flush=True)
        json.dump(data, open(path, "w"), indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
