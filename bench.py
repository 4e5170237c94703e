"""Benchmark of the FlOF deformation path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-frames]

A step is one device-resident multiscale FlOF solve at 64^4 (seeded drop
SDFs, wSmooth 0.1, wEnergy 1e-4, 3 levels, every level run to relative
signed-max residual 1e-5 -- the iteration cap is raised above the
reference's 1000 so the residual is the same at every level).  The JSON line
reports solves/s of K solves on the production launch path (PCG iterations
as captured HIP graphs), the HBM roofline of the dominant kernel (the PCG
direction kernel) from the same K solves repeated with a HIP event pair on every PCG
launch on the library's stream (events force plain launches, which costs a
few percent of the solve, so they are kept out of the timed solves), and the
CPU oracle timed on a bounded sample of the same workload on this host.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "64^4 FlOF solves/s at fixed residual; % HBM roofline; warm-start vs plain"
HBM_PEAK = 8.0e12
N = 64
PARAMS = dict(wSmooth=0.1, wEnergy=1e-4, cgAccuracy=1e-5, minGridSize=16,
              multiStep=1, postVelBlur=0.0, doFinalProject=False,
              resetBndWidth=-1.0, cfl=999.0)
FIXED_RESIDUAL_CAP = 20000
# algorithmic bytes per cell of one launch (FP32 Vec4 = 16 B per vector,
# i1 = 4 B; g and diag^-1 are recomputed from i1)
B_UPDATE = 16 * 2 + 4 + 16       # read p,r,i1; write r
B_APPLY = 16 + 4                 # read p,i1 (p.Ap only)
# the direction kernel moves x every other iteration (the x update of an even
# iteration is deferred to the next one): even read r,p,i1, write p; odd also
# read x and the previous direction, write x
B_DIR_EVEN = 16 * 2 + 4 + 16
B_DIR_ODD = 16 * 4 + 4 + 16 * 2
B_DIR = (B_DIR_EVEN + B_DIR_ODD) / 2  # per iteration
B_ITER_MIN = 112                 # x,r,p read+written once, g read once
# kernel names of the PCG launch kinds on the finest level: the library runs
# the LDS-tiled stencil kernels there (64^4, FP32: long t segments of full
# tiles) and the t-marching ones on the coarser levels; FLOF_LDS_STENCIL=0
# selects the t-marching kernels everywhere
LDS_STENCIL = os.environ.get("FLOF_LDS_STENCIL", "auto")[:1] != "0"
KN_UPDATE = "k_cg_stencil_lds<update>" if LDS_STENCIL else "k_cg_update_tm"
KN_APPLY = "k_cg_stencil_lds<pq>" if LDS_STENCIL else "k_cg_pq_tm"
KN_DIR = "k_cg_dir_tm"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-frames", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=12)
    ap.add_argument("--no-graphs", action="store_true",
                    help="enqueue CG iterations as plain launches instead of "
                         "captured graphs (for profilers that trace graphs poorly)")
    ap.add_argument("--batch-workload", action="store_true",
                    help="16 solves (forward and backward of 8 seeded drop pairs, "
                         "N in {48,64,80}, T in {32,48,64}) one after another and "
                         "as one batch on worker streams; prints its own JSON line")
    ap.add_argument("--precision-workload", action="store_true",
                    help="96^3 x 64 and 32^3 x 32 drop pairs (2 spheres, radii "
                         "0.08N-0.18N): FP32 vs BF16 search direction on the GPU; "
                         "prints its own JSON line")
    ap.add_argument("--pmc-workload", action="store_true",
                    help="short workload for rocprofv3 --pmc runs: a 256 MiB "
                         "grid copy (calibration) and 3 level-0 PCG iterations")
    ap.add_argument("--warp-workload", action="store_true",
                    help="short workload for rocprofv3 --pmc runs of the frame "
                         "warp: a 64^4 solve capped at 50 iterations per level, "
                         "then the bench's 64-frame 128^3 warp; prints its JSON")
    ap.add_argument("--ops-workload", action="store_true",
                    help="the path's other operations at 64^4 (grid ops, advect4d, "
                         "interpolateGrid4d, gaussianBlur4d, calcLsDiff4d, "
                         "corrVelsOf4d, preprocessing), each timed over repeated "
                         "calls against its algorithmic bytes; prints its JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU: the FlOF solve is not "
              "sharded across GPUs (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    if args.no_graphs:
        os.environ["FLOF_NO_GRAPH"] = "1"  # read when the context is created
    from ofblend_amd import _lib
    from ofblend_amd import manta as M
    from ofblend_amd.abi import (FLOF_KT_CG_APPLY, FLOF_KT_CG_DIR, FLOF_KT_CG_UPDATE,
                                 FLOF_KT_WARP, FLOF_PSTORE_BF16)

    L = _lib.lib()
    if args.batch_workload:
        batch_workload(M, L)
        return
    if args.precision_workload:
        precision_workload(M)
        return
    s = M.Solver(gridSize=M.vec3(N, N, N), fourthDim=N)
    ctx = M.context()
    i0 = s.create(M.Grid4Real)
    i1 = s.create(M.Grid4Real)
    M.dropSdf4d(i0, seed=1)
    M.dropSdf4d(i1, seed=2)
    vel = s.create(M.Grid4Vec4)
    if args.warp_workload:
        M.opticalFlowMultiscale4d(vel=vel, i0=i0, i1=i1, cgMaxIter=50, **PARAMS)
        M.sync()
        print(json.dumps(frame_warp(M, L, ctx, s, vel, FLOF_KT_WARP)), flush=True)
        return
    if args.ops_workload:
        M.opticalFlowMultiscale4d(vel=vel, i0=i0, i1=i1, **PARAMS)
        M.sync()
        print(json.dumps(ops_workload(M, s, i0, i1, vel)), flush=True)
        return
    if args.pmc_workload:
        # copy of a 64^4 Vec4 grid: 256 MiB read + 256 MiB written
        other = s.create(M.Grid4Vec4)
        other.copyFrom(vel)
        M.sync()
        M.opticalFlow4d(vel, i0, i1, wSmooth=0.1, wEnergy=1e-4, cgAccuracy=1e-30,
                        cgMaxIter=3)
        M.sync()
        log("bench: pmc workload done")
        return

    def solve(cap):
        vel.clear()
        return M.opticalFlowMultiscale4d(vel=vel, i0=i0, i1=i1, cgMaxIter=cap, **PARAMS)

    log("bench: warmup %d solve(s) at %d^4" % (args.warmup, N))
    for _ in range(args.warmup):
        solve(FIXED_RESIDUAL_CAP)
    M.sync()
    # timed region: K solves on the production launch path
    times = []
    infos = []
    for st in range(args.steps):
        M.sync()
        t0 = time.perf_counter()
        err = solve(FIXED_RESIDUAL_CAP)
        M.sync()
        times.append(time.perf_counter() - t0)
        infos.append(M.getLastSolveInfo().summary())
        log("bench: step %d %.3f s iters %s" % (
            st, times[-1], [l["iters"][0] for l in infos[-1]["levels"]]))
    # the same K solves with a HIP event pair on every PCG launch: the
    # per-kernel durations of the roofline
    kinds = (FLOF_KT_CG_APPLY, FLOF_KT_CG_UPDATE, FLOF_KT_CG_DIR)
    for k in kinds:
        L.flof_ctx_timing(ctx, k, None, None)  # drop earlier records
        L.flof_ctx_enable_timing(ctx, k, 1)
    t_events = []
    for st in range(args.steps):
        M.sync()
        t0 = time.perf_counter()
        solve(FIXED_RESIDUAL_CAP)
        M.sync()
        t_events.append(time.perf_counter() - t0)
        it_ev = [l["iters"][0] for l in M.getLastSolveInfo().summary()["levels"]]
        log("bench: step %d with kernel events %.3f s iters %s" % (st, t_events[-1], it_ev))
    cells0 = N ** 4
    ktime, ktime0 = {}, {}
    for k in kinds:
        n, ms = C.c_int64(), C.c_double()
        L.flof_ctx_timing_tag(ctx, k, cells0, C.byref(n), C.byref(ms))
        ktime0[k] = (n.value, ms.value)
        L.flof_ctx_timing(ctx, k, C.byref(n), C.byref(ms))
        L.flof_ctx_enable_timing(ctx, k, 0)
        ktime[k] = (n.value, ms.value)
    info = infos[-1]
    levels = info["levels"]
    iters = [l["iters"][0] for l in levels]
    cells = [int(np.prod(l["dims"])) for l in levels]
    cell_iters = sum(i * c for i, c in zip(iters, cells))
    ms_step = 1000.0 * float(np.median(times))
    solves_per_s = 1000.0 / ms_step

    # all levels: algorithmic bytes of all real launches over the summed
    # event time of every launch of that kernel
    def kstat(kind, bpc):
        n, ms = ktime[kind]
        byts = bpc * cell_iters * args.steps
        return {"launches": n, "ms_total": ms, "avg_us": 1000.0 * ms / max(1, n),
                "GBps": byts / (ms * 1e-3) / 1e9 if ms > 0 else None}

    # finest level only: launches tagged with its cell count; the iterations
    # after convergence that the chunked host loop enqueues return at once
    # and are counted in the time, not in the bytes
    def kstat0(kind, bpc):
        n, ms = ktime0[kind]
        real = iters[0] * args.steps
        byts = bpc * cells0 * real
        return {"launches": n, "iterations": real, "ms_total": ms,
                "avg_us_per_iteration": 1000.0 * ms / max(1, real),
                "bytes_per_launch": bpc * cells0,
                "GBps": byts / (ms * 1e-3) / 1e9 if ms > 0 else None}
    names = {FLOF_KT_CG_UPDATE: (KN_UPDATE, B_UPDATE),
             FLOF_KT_CG_APPLY: (KN_APPLY, B_APPLY), FLOF_KT_CG_DIR: (KN_DIR, B_DIR)}
    kst = {names[k][0]: kstat(k, names[k][1]) for k in kinds}
    kst0 = {names[k][0]: kstat0(k, names[k][1]) for k in kinds}
    dom = max(kst0, key=lambda k: kst0[k]["ms_total"])
    achieved = kst0[dom]["GBps"] * 1e9
    iter_ms0 = sum(kst0[k]["ms_total"] for k in kst0) / (iters[0] * args.steps)
    traffic = committed_traffic(dom)
    roofline = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                "unit": "GB/s", "frac": achieved / HBM_PEAK,
                "traffic": traffic,
                "kernel": dom, "launch": "finest level, %d cells" % cells0,
                "bytes_per_cell": {KN_UPDATE: B_UPDATE, KN_APPLY: B_APPLY,
                                   KN_DIR: B_DIR}[dom],
                "timed_over": "the K solves repeated with per-launch HIP events",
                "avg_launch_us": kst0[dom]["avg_us_per_iteration"],
                "pcg_iteration_us": 1000.0 * iter_ms0,
                "pcg_iteration_frac_112B": (B_ITER_MIN * cells0 / (iter_ms0 * 1e-3)) / HBM_PEAK}

    out = {
        "metric": METRIC, "value": solves_per_s, "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic: seeded falling-drop SDFs (seeds 1, 2)",
        "config": {"workload": "FlOF opticalFlowMultiscale4d 64^4, 3 levels, "
                               "P_plain (wSmooth 0.1, wEnergy 1e-4), every level to "
                               "relative residual 1e-5",
                   "grid": [N, N, N, N], "levels": len(levels),
                   "iters_per_level": iters, "final_err": info["err_final"],
                   "residual_per_level": [l["residual"][0] for l in levels],
                   "residual_l2_per_level": [l["residual_l2"][0] for l in levels]},
        "roofline": roofline,
        "kernels_finest_level": kst0,
        "kernels_all_levels": kst,
        "kernels_all_levels_note": "per launch kind; the coarser levels run the "
                                   "t-marching p.Ap and update kernels",
        "ms_per_step_with_kernel_events": 1000.0 * float(np.median(t_events)),
    }

    # what HBM sustains for the byte mixes of the PCG kernels when a plain
    # streaming kernel moves them (no stencil work), on the finest level's
    # cell count: the practical ceiling the kernels are compared with
    from ofblend_amd.abi import FLOF_MIX_COPY, FLOF_MIX_DIR, FLOF_MIX_DIR_X
    stream = {}
    for nm, mix in (("mix_52B", FLOF_MIX_DIR), ("mix_100B", FLOF_MIX_DIR_X),
                    ("copy_32B", FLOF_MIX_COPY)):
        g = C.c_double()
        rc = L.flof_stream_bandwidth(ctx, mix, cells0, 20, C.byref(g))
        if rc == 0:
            stream[nm] = g.value
    out["stream_bandwidth_GBps"] = stream
    out["stream_bandwidth_note"] = ("52 B: read 2 float4 + 1 float, write 1 float4 (the "
                                    "update and the even direction launches); 100 B: read "
                                    "4 float4 + 1 float, write 2 float4 (the odd ones)")
    ach = None
    if dom == KN_UPDATE and "mix_52B" in stream:
        ach = stream["mix_52B"]
    elif dom == KN_DIR and "mix_52B" in stream and "mix_100B" in stream:
        # an even and an odd launch at the stream rates of their mixes
        ach = (B_DIR_EVEN + B_DIR_ODD) / (B_DIR_EVEN / stream["mix_52B"] +
                                          B_DIR_ODD / stream["mix_100B"])
    if ach:
        roofline["achievable_same_mix"] = {
            "GBps": ach, "frac": achieved / 1e9 / ach,
            "note": "streaming kernels with the same bytes per cell and no stencil work"}

    # host-buffer boundary: the inputs come from and the flow goes to host
    # memory (pageable numpy arrays): copies around one solve
    h0, h1 = i0.to_numpy(), i1.to_numpy()
    M.sync()
    t0 = time.perf_counter()
    i0.from_numpy(h0)
    i1.from_numpy(h1)
    M.sync()
    t_up = time.perf_counter() - t0
    t0 = time.perf_counter()
    hv = vel.to_numpy()
    t_down = time.perf_counter() - t0
    out["pcie_inclusive"] = {"upload_ms": 1000.0 * t_up, "download_ms": 1000.0 * t_down,
                             "ms_per_step": ms_step + 1000.0 * (t_up + t_down),
                             "bytes": h0.nbytes + h1.nbytes + hv.nbytes}

    # the same solve with the reference's iteration cap (1000)
    t0 = time.perf_counter()
    solve(1000)
    M.sync()
    capinfo = M.getLastSolveInfo().summary()
    out["reference_cap_1000"] = {
        "ms": 1000.0 * (time.perf_counter() - t0),
        "iters_per_level": [l["iters"][0] for l in capinfo["levels"]],
        "residual_per_level": [l["residual"][0] for l in capinfo["levels"]]}
    vcap = vel.to_numpy()
    import hashlib
    out["reference_cap_1000"]["sha256"] = hashlib.sha256(vcap.tobytes()).hexdigest()

    # the search direction stored as bfloat16 (extension, not the headline:
    # the reference computes in FP32); second of two solves is timed
    vel.clear()
    solve(FIXED_RESIDUAL_CAP)
    vf = vel.to_numpy()
    for _ in range(2):
        vel.clear()
        M.sync()
        t0 = time.perf_counter()
        eb = M.opticalFlowMultiscale4d(vel=vel, i0=i0, i1=i1, cgMaxIter=FIXED_RESIDUAL_CAP,
                                       precision=FLOF_PSTORE_BF16, **PARAMS)
        M.sync()
        tb = time.perf_counter() - t0
    bi = M.getLastSolveInfo().summary()
    out["bf16_direction"] = {
        "ms": 1000.0 * tb, "iters_per_level": [l["iters"][0] for l in bi["levels"]],
        "final_err": eb, "final_err_fp32": info["err_final"],
        "max_abs_diff_vs_fp32": float(np.abs(vel.to_numpy() - vf).max()),
        "note": "x, r, g and all arithmetic FP32, CG scalars FP64; p stored as bfloat16"}
    out["warm_start"] = warm_start(M, s, vel, i0, i1, vf, iters, ms_step)

    if not args.no_frames:
        out["frame_warp"] = frame_warp(M, L, ctx, s, vel, FLOF_KT_WARP)

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_iters, iters, cells, ms_step)
        out["speedup_vs_cpu"] = out["cpu_baseline"]["s_per_solve"] / (ms_step / 1000.0)
    print(json.dumps(out), flush=True)


def precision_workload(M):
    """Reduced-precision search direction: FP32 and BF16 solves of the same
    2-sphere drop pair at 96^3 x 64 and 32^3 x 32 against each other (the
    comparison with the FP64 CPU oracle is tests/study_precision.py)."""
    from ofblend_amd.abi import FLOF_PSTORE_BF16
    out = {"workload": "2-sphere drop pairs (seeds 7, 8; radii 0.08N-0.18N), P_plain, "
                       "residual 1e-5"}
    for (n, t) in ((96, 64), (32, 32)):
        sv = M.Solver(gridSize=M.vec3(n, n, n), fourthDim=t)
        a, b = sv.create(M.Grid4Real), sv.create(M.Grid4Real)
        M.dropSdf4d(a, seed=7, nspheres=2, rmin=0.08, rmax=0.18)
        M.dropSdf4d(b, seed=8, nspheres=2, rmin=0.08, rmax=0.18)
        v = sv.create(M.Grid4Vec4)
        prm = dict(PARAMS, minGridSize=n // 4, cgMaxIter=FIXED_RESIDUAL_CAP)  # 3 levels
        res = {}
        flows = {}
        for name, prec in (("fp32", 0), ("bf16", FLOF_PSTORE_BF16)):
            v.clear()
            M.sync()
            t0 = time.perf_counter()
            e = M.opticalFlowMultiscale4d(vel=v, i0=a, i1=b, precision=prec, **prm)
            M.sync()
            dt = time.perf_counter() - t0
            info = M.getLastSolveInfo().summary()
            flows[name] = v.to_numpy()
            res[name] = {"ms": 1000 * dt, "iters_per_level": [l["iters"][0] for l in info["levels"]],
                         "final_err": e}
        d = flows["bf16"] - flows["fp32"]
        res["bf16_vs_fp32"] = {"max": float(np.abs(d).max()),
                               "rms": float(np.sqrt((d ** 2).mean()))}
        out["%dx%d" % (n, t)] = res
        log("bench: precision %d^3 x %d done" % (n, t))
    print(json.dumps(out), flush=True)


def batch_workload(M, L):
    """Forward and backward deformations of 8 seeded drop configurations of
    uneven sizes: sequential solves vs one batch on worker streams (same
    results, checked), with the launch-shape autotuner on."""
    rng = np.random.default_rng(4)
    jobs = []
    for k in range(8):
        n = int(rng.choice([48, 64, 80]))
        t = int(rng.choice([32, 48, 64]))
        sv = M.Solver(gridSize=M.vec3(n, n, n), fourthDim=t)
        a, b = sv.create(M.Grid4Real), sv.create(M.Grid4Real)
        M.dropSdf4d(a, seed=100 + 2 * k, nspheres=1 + k % 3, rmin=0.08, rmax=0.18)
        M.dropSdf4d(b, seed=101 + 2 * k, nspheres=1 + k % 3, rmin=0.08, rmax=0.18)
        for (x, y) in ((a, b), (b, a)):
            jobs.append((sv, x, y, sv.create(M.Grid4Vec4), (n, t)))
    prm = dict(PARAMS, cgMaxIter=FIXED_RESIDUAL_CAP)
    L.flof_ctx_set_autotune(M.context(), 1)
    log("bench: batch warm-up")
    for sv, x, y, v, _ in jobs[:2]:
        M.opticalFlowMultiscale4d(vel=v, i0=x, i1=y, **prm)
    M.sync()
    t0 = time.perf_counter()
    seq = []
    for sv, x, y, v, _ in jobs:
        v.clear()
        M.opticalFlowMultiscale4d(vel=v, i0=x, i1=y, **prm)
        seq.append(v.to_numpy())
    M.sync()
    t_seq = time.perf_counter() - t0
    log("bench: sequential %.2f s" % t_seq)
    res = {}
    for conc in (2,):
        for _, _, _, v, _ in jobs:
            v.clear()
        M.sync()
        t0 = time.perf_counter()
        errs, infos = M.opticalFlowMultiscale4dBatch([j[3] for j in jobs], [j[1] for j in jobs],
                                                     [j[2] for j in jobs], concurrency=conc,
                                                     **prm)
        M.sync()
        tb = time.perf_counter() - t0
        same = all(np.array_equal(j[3].to_numpy(), seq[q]) for q, j in enumerate(jobs))
        res["worker_streams_%d" % conc] = {"s": tb, "solves_per_s": len(jobs) / tb,
                                           "identical_to_sequential": same}
        log("bench: batch x%d %.2f s identical %s" % (conc, tb, same))
    # the fused batch: one launch per PCG kernel kind and iteration spans the
    # levels of equal index of all 16 problems (the second run is timed)
    for rep in range(2):
        for _, _, _, v, _ in jobs:
            v.clear()
        M.sync()
        t0 = time.perf_counter()
        errs, infos = M.opticalFlowMultiscale4dFused([j[3] for j in jobs], [j[1] for j in jobs],
                                                     [j[2] for j in jobs], **prm)
        M.sync()
        tf = time.perf_counter() - t0
    same = all(np.array_equal(j[3].to_numpy(), seq[q]) for q, j in enumerate(jobs))
    res["fused"] = {"s": tf, "solves_per_s": len(jobs) / tf, "identical_to_sequential": same,
                    "finest_level_iters": [inf.summary()["levels"][0]["iters"][0]
                                           for inf in infos]}
    log("bench: fused batch %.2f s identical %s" % (tf, same))
    cells = sum(j[4][0] ** 3 * j[4][1] for j in jobs)
    print(json.dumps({"workload": "16 solves: forward+backward of 8 seeded drop pairs, "
                                  "N in {48,64,80}, T in {32,48,64}, P_plain, residual 1e-5",
                      "sizes": [j[4] for j in jobs], "total_cells": cells,
                      "sequential_s": t_seq, "sequential_solves_per_s": len(jobs) / t_seq,
                      **res}), flush=True)


def committed_traffic(kernel):
    """HBM bytes per finest-level launch of `kernel` from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated by a copy kernel),
    or None if the summary is absent."""
    path = os.path.join(ROOT, "profiles", "pcg_kernels_64x4_hbm_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("machine-readable:"):
            d = json.loads(line.split(":", 1)[1]).get(kernel)
            if d:
                return d["fetch_bytes"] + d["write_bytes"]
    return None


def warm_start(M, s, vel, i0, i1, v_plain, iters_plain, ms_plain):
    """The solve warm-started by the trained network (weights committed in
    ofblend_amd/data) against the plain solve: iterations, time and the
    SDF-alignment error of both (calcLsDiff, and the RMS of i0 warped by the
    flow minus i1)."""
    from ofblend_amd import warmstart as WS
    if not os.path.exists(WS.DEFAULT_WEIGHTS):
        return None
    net = WS.WarmNet(seed=1)
    net.load(WS.DEFAULT_WEIGHTS)
    warm = s.create(M.Grid4Vec4)
    res = {}
    for rep in range(2):
        M.sync()
        t0 = time.perf_counter()
        net.warm_start(i0, i1, warm, netX=32)
        vel.clear()
        err = M.opticalFlowMultiscale4d(vel=vel, i0=i0, i1=i1, warm=warm,
                                        cgMaxIter=FIXED_RESIDUAL_CAP, **PARAMS)
        M.sync()
        t = time.perf_counter() - t0
    wi = M.getLastSolveInfo().summary()
    vw = vel.to_numpy()
    h1 = i1.to_numpy()

    def rms_sdf(flow):
        # RMS of i0 warped by the flow (advect4d) minus i1, over all cells
        w = s.create(M.Grid4Real)
        w.copyFrom(i0)
        M.advect4d(flow, w)
        d = w.to_numpy().astype(np.float64) - h1
        return float(np.sqrt(np.mean(d * d)))

    rms_warm = rms_sdf(vel)
    vel.clear()
    e_plain = M.opticalFlowMultiscale4d(vel=vel, i0=i0, i1=i1, cgMaxIter=FIXED_RESIDUAL_CAP,
                                        **PARAMS)
    rms_plain = rms_sdf(vel)
    it_w = [l["iters"][0] for l in wi["levels"]]
    return {"ms_including_net": 1000.0 * t, "ms_plain": ms_plain,
            "iters_per_level": it_w, "iters_per_level_plain": iters_plain,
            "finest_level_iteration_ratio": iters_plain[0] / max(1, it_w[0]),
            "warm_kept_per_level": [l["warm_used"] for l in wi["levels"]],
            "final_err": err, "final_err_plain": e_plain,
            "rms_sdf_diff": rms_warm, "rms_sdf_diff_plain": rms_plain,
            "max_abs_diff_vs_plain": float(np.abs(vw - v_plain).max()),
            "net_level": 32}


def frame_warp(M, L, ctx, s, vel, kind):
    """The frame warp workload: 64 analytic 3D frames at 128^3 placed into a
    128^3 x 64 level set and warped by the solved 64^4 deformation (alpha 1,
    defoFactor (2,2,2,1))."""
    F = 2
    fr = M.DeviceArray((N, F * N, F * N, F * N))
    M.dropFrames3d(fr, N, N, N, N, F, seed=1)
    sp = M.Solver(gridSize=M.vec3(F * N, F * N, F * N), fourthDim=N)
    phi = sp.create(M.Grid4Real)
    t0 = time.perf_counter()
    M.loadPlaceGrid4dFrames(fr, phi, M.Vec4(0.), M.Vec4(1.))
    M.sync()
    t_place = time.perf_counter() - t0
    dst = M.DeviceArray((N, F * N, F * N, F * N))
    args = (vel.h, dst.ptr, F * N, F * N, F * N, N, 0.0, 1.0, phi.h, 1.0, 1.0,
            _f4(0, 0, 0, 0), _f4(1, 1, 1, 1), _f4(2, 2, 2, 1), 1, 1.0)
    _lib_check(L.flof_advect_time_slices_opt(*args), ctx)  # warm-up
    M.sync()
    L.flof_ctx_timing(ctx, kind, None, None)
    L.flof_ctx_enable_timing(ctx, kind, 1)
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        _lib_check(L.flof_advect_time_slices_opt(*args), ctx)
    M.sync()
    wall = (time.perf_counter() - t0) / reps
    n, ms = C.c_int64(), C.c_double()
    L.flof_ctx_timing(ctx, kind, C.byref(n), C.byref(ms))
    L.flof_ctx_enable_timing(ctx, kind, 0)
    kern_s = ms.value / 1000.0 / max(1, n.value)
    M3, Md = (F * N) ** 3, N ** 3
    byts = N * (4 * M3 + 8 * M3 + 2 * 16 * Md)  # per frame: write, 2 phi slices, 2 defo slices
    return {"frames": N, "frame": [F * N] * 3, "ms_per_64_frames": 1000 * wall,
            "frames_per_s": N / wall, "kernel_ms": 1000 * kern_s,
            "GBps_algorithmic": byts / kern_s / 1e9, "frac": byts / kern_s / HBM_PEAK,
            "place_ms": 1000 * t_place}


def ops_workload(M, s, i0, i1, vel, reps=10):
    """The path's operations other than the PCG and the frame warp, on the
    64^4 bench grids (vel: the solved flow).  Each is called once to warm up,
    then `reps` times back to back; the time per call is the wall time over
    the calls, closed by a device synchronisation (calls that return a host
    value synchronise every time, so theirs includes one round trip).
    Algorithmic bytes: each cell's compulsory HBM reads and writes (gathers of
    a neighbourhood count once)."""
    cells = N ** 4
    h = N // 2
    sc = M.Solver(gridSize=M.vec3(h, h, h), fourthDim=h)
    coarse_r = sc.create(M.Grid4Real)
    coarse_v = sc.create(M.Grid4Vec4)
    M.interpolateGrid4d(coarse_v, vel)
    tmp_r = s.create(M.Grid4Real)
    tmp_v = s.create(M.Grid4Vec4)
    dst_v = s.create(M.Grid4Vec4)
    phi = s.create(M.Grid4Real)

    def advect():
        tmp_r.copyFrom(i0)
        M.advect4d(vel, tmp_r)

    def prep():
        phi.copyFrom(i1)
        M.extrap4dLsSimple(phi, distance=4, inside=False)

    # corrVelsOf4d subtracts its correction from vel in place, and its pass
    # count follows max|vel|: every call starts from the solved flow
    vel_c = s.create(M.Grid4Vec4)

    def corr():
        vel_c.copyFrom(vel)
        M.corrVelsOf4d(dst_v, vel_c, i0, i0, i1)

    ops = [
        # name, call, algorithmic bytes per call, what the bytes are
        ("grid4d copyFrom Vec4", lambda: tmp_v.copyFrom(vel), 32 * cells,
         "read and write 16 B per cell"),
        ("grid4d addScaled Real", lambda: tmp_r.addScaled(i1, 0.5), 12 * cells,
         "read 2 x 4 B, write 4 B per cell"),
        ("grid4d getMaxAbs Vec4", lambda: vel.getMaxAbs(), 16 * cells,
         "read 16 B per cell (one host round trip)"),
        ("advect4d Real by Vec4 (after a Real copy)", advect, (8 + 24) * cells,
         "copy 8 B; advection reads vel 16 B and the source 4 B, writes 4 B per cell"),
        ("interpolateGrid4d restrict Real 64^4 -> 32^4",
         lambda: M.interpolateGrid4d(coarse_r, i0), 4 * cells + 4 * h ** 4,
         "read the fine grid once, write the coarse one"),
        ("interpolateGrid4d prolong Vec4 32^4 -> 64^4",
         lambda: M.interpolateGrid4d(tmp_v, coarse_v), 16 * h ** 4 + 16 * cells,
         "read the coarse grid once, write the fine one"),
        ("gaussianBlur4d Vec4 sigma 1 (81 taps)",
         lambda: M.gaussianBlur4d(tmp_v, sigma=1.0), 64 * cells,
         "two passes (into a temporary and back), each reading and writing "
         "16 B per cell"),
        ("calcLsDiff4d", lambda: M.calcLsDiff4d(i0, i1), 8 * cells,
         "read 2 x 4 B per cell (one host round trip)"),
        ("corrVelsOf4d (vel restored before each call)", corr, None,
         "projection with its line searches, then int(max|vel|) + 4 "
         "extrapolation passes; no single byte count"),
        ("extrap4dLsSimple distance 4 (after a Real copy)", prep, None,
         "layer-by-layer extrapolation; no single byte count"),
        ("repeatFrame4d", lambda: M.repeatFrame4d(phi, 10.0), None,
         "copies one t slice into the others"),
    ]
    out = {"grid": [N] * 4, "cells": cells, "reps": reps, "ops": []}
    for name, fn, byts, what in ops:
        log("bench: ops workload", name)
        fn()
        M.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        M.sync()
        t = (time.perf_counter() - t0) / reps
        rec = {"op": name, "ms_per_call": 1000 * t, "bytes": what}
        if byts is not None:
            rec["GBps_algorithmic"] = byts / t / 1e9
            rec["frac"] = byts / t / HBM_PEAK
        out["ops"].append(rec)
    return out


def _lib_check(rc, ctx):
    from ofblend_amd._lib import check
    check(rc, ctx)


def _f4(*v):
    return (C.c_float * 4)(*v)


def cpu_baseline(k_iters, iters, cells, gpu_ms):
    """The CPU oracle (FP32, the reference's serial CG and FP64 dot order)
    timed on a bounded sample: the finest-level set-up plus PCG at 64^4 with
    an iteration cap of k and 2k; the difference gives the time per iteration,
    scaled to a full solve with the GPU's per-level iteration counts."""
    from oracle import oracle as O
    from ofblend_amd.abi import default_params
    threads = min(16, os.cpu_count() or 1)
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    i0 = O.drop_sdf(N, N, N, N, seed=1)
    i1 = O.drop_sdf(N, N, N, N, seed=2)
    v = np.zeros(i0.shape + (4,), np.float32)
    tt = []
    for k in (k_iters, 2 * k_iters):
        p = default_params(wSmooth=0.1, wEnergy=1e-4, cgAccuracy=1e-5, cgMaxIter=k)
        t0 = time.perf_counter()
        O.optical_flow4d(v, i0, i1, p)
        tt.append(time.perf_counter() - t0)
        log("bench: cpu oracle 64^4 level solve, %d iterations: %.2f s" % (k, tt[-1]))
    t_it = (tt[1] - tt[0]) / k_iters
    t_setup = max(0.0, tt[0] - k_iters * t_it)
    per_cell_it = t_it / cells[0]
    per_cell_setup = t_setup / cells[0]
    est = sum(i * c * per_cell_it + c * per_cell_setup for i, c in zip(iters, cells))
    return {"value": 1.0 / est, "unit": "solves/s", "cores": threads, "kind": "port",
            "s_per_solve": est,
            "sample": "CPU oracle (FP32-faithful, serial CG as the reference) on "
                      "the 64^4 finest level: set-up + %d and %d PCG iterations "
                      "(%.1f s); time per cell-iteration %.1f ns scaled to the "
                      "solve's per-level iteration counts %s" % (
                          k_iters, 2 * k_iters, sum(tt), per_cell_it * 1e9, iters)}


if __name__ == "__main__":
    main()
