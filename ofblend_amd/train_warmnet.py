"""Trains the warm-start network on the finest-level increments of the plain
GPU solve.

    python -m ofblend_amd.train_warmnet --out weights.npy [--flows 2000]
        [--N 64] [--T 64] [--steps-per-flow 1] [--ring 8] [--lr 1e-3]
        [--time-budget SECONDS] [--resume weights.npy --start K] [--seed 1]

Each flow is one plain multiscale solve (wSmooth 0.1, wEnergy 1e-4, relative
residual 1e-5, minGridSize 16) of a seeded drop pair of 1-3 spheres (radii
0.08N-0.18N) at N^3 x T, computed on the GPU; the training pair is the
finest level's pre-warped i0 and i1 (the network's input) and the increment
that level's solve adds to the prolonged coarse flow (its target).  Flows
are generated as training proceeds; the last `ring` of them stay on the GPU
and Adam draws its samples from them.  With a time budget the run stops
early and reports how many flows it used; --resume/--start continue it.
The weights (a flat float32 .npy) serve WarmNet.load() and WarmNet.solve().
"""
import argparse
import json
import sys
import time

import numpy as np

from . import manta as M
from . import warmstart as WS


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--flows", type=int, default=2000)
    ap.add_argument("--N", type=int, default=64)
    ap.add_argument("--T", type=int, default=64)
    ap.add_argument("--min-grid", type=int, default=16)
    ap.add_argument("--steps-per-flow", type=int, default=1)
    ap.add_argument("--ring", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-end", type=float, default=1e-4)
    ap.add_argument("--time-budget", type=float, default=0.0)
    ap.add_argument("--resume", default=None)
    ap.add_argument("--start", type=int, default=0)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--heldout", type=int, default=2)
    a = ap.parse_args(argv)

    def log(*m):
        print(*m, file=sys.stderr, flush=True)

    N, T = a.N, a.T
    s = M.Solver(gridSize=M.vec3(N, N, N), fourthDim=T)
    prm = dict(WS.PLAIN, minGridSize=a.min_grid)
    net = WS.WarmNet(seed=a.seed)
    if a.resume:
        net.load(a.resume)
    rng = np.random.default_rng(a.seed + a.start)

    def sample(seed, k):
        i0, i1 = WS.drop_pair(s, seed, nspheres=1 + k % 3, rmin=0.08, rmax=0.18)
        vel = s.create(M.Grid4Vec4)
        _, info, i0w, inc = WS.capture(vel, i0, i1, level=0, **prm)
        return (i0w, i1, inc), info

    # held-out pairs (seeds outside the training stream)
    held = [sample(10 ** 6 + 2 * k, k)[0] for k in range(a.heldout)]

    def held_loss():
        return float(np.mean([net.loss_grad(*h) for h in held])) if held else 0.0

    before = held_loss()
    ring = []
    t0 = time.perf_counter()
    t_solve = t_train = 0.0
    k = a.start
    losses = []
    while k < a.flows:
        if a.time_budget > 0 and time.perf_counter() - t0 > a.time_budget:
            break
        ts = time.perf_counter()
        smp, info = sample(1000 * a.seed + 2 * k, k)
        t_solve += time.perf_counter() - ts
        ring.append(smp)
        if len(ring) > a.ring:
            ring.pop(0)
        tt = time.perf_counter()
        for _ in range(a.steps_per_flow):
            f = 0.5 * (1 + np.cos(np.pi * min(1.0, k / max(1, a.flows - 1))))
            losses.append(net.loss_grad(*ring[rng.integers(len(ring))]))
            net.adam(lr=a.lr_end + (a.lr - a.lr_end) * f)
        t_train += time.perf_counter() - tt
        k += 1
        if k % 25 == 0 or k == a.flows:
            log("flows %d loss(last 25 steps) %.4f solve %.1f s train %.1f s"
                % (k, float(np.mean(losses[-25:])), t_solve, t_train))
            net.save(a.out)
    after = held_loss()
    net.save(a.out)
    print(json.dumps({"flows_used": k, "flows_target": a.flows, "N": N, "T": T,
                      "steps_per_flow": a.steps_per_flow, "solve_s": t_solve,
                      "train_s": t_train, "heldout_loss_before": before,
                      "heldout_loss_after": after,
                      "train_loss_last50": float(np.mean(losses[-50:])) if losses else None}),
          flush=True)


if __name__ == "__main__":
    main()
