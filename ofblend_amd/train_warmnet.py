"""Trains the warm-start network on flows of the plain GPU solve.

    python -m ofblend_amd.train_warmnet --out weights.npy [--samples 64]
        [--N 32] [--T 32] [--steps 1500] [--lr 2e-3] [--seed 1]

Samples are seeded drop pairs of 1-3 spheres (radii 0.08N-0.18N) at N^3 x T,
their targets the multiscale solve's flows (wSmooth 0.1, wEnergy 1e-4,
relative residual 1e-5).  The weights (a flat float32 .npy) serve
WarmNet.load(); the network is applied at the pyramid level with X <= N.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import warmstart as WS


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--samples", type=int, default=64)
    ap.add_argument("--N", type=int, default=32)
    ap.add_argument("--T", type=int, default=32)
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=1)
    a = ap.parse_args(argv)

    def log(*m):
        print(*m, file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    s, data = WS.generate(a.samples, a.N, a.N if a.T <= 0 else a.T, seed0=1000 * a.seed,
                          minGridSize=8)
    t_gen = time.perf_counter() - t0
    log("generated %d samples in %.1f s" % (len(data), t_gen))
    net = WS.WarmNet(seed=a.seed)
    hold = data[-4:]
    train = data[:-4] if len(data) > 8 else data

    def held_loss():
        return float(np.mean([net.loss_grad(*d) for d in hold]))

    before = held_loss()
    t0 = time.perf_counter()
    losses = WS.train(net, train, a.steps, lr=a.lr, seed=a.seed, log=log)
    t_train = time.perf_counter() - t0
    after = held_loss()
    net.save(a.out)
    print(json.dumps({"samples": len(data), "N": a.N, "T": a.T, "steps": a.steps,
                      "generate_s": t_gen, "train_s": t_train,
                      "heldout_loss_before": before, "heldout_loss_after": after,
                      "train_loss_last50": float(np.mean(losses[-50:]))}), flush=True)


if __name__ == "__main__":
    main()
