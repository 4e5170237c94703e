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
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the sections after the frame warp (mixed precision at "
                         "96^3 x 64, the learned warm start at 128^3 x 64, the fused "
                         "batch, the flof.py parameter set)")
    ap.add_argument("--flof-workload", action="store_true",
                    help="flof.py's mode-1 parameters (multiStep 3, blur 4, final "
                         "projection, band 0.1, wSmooth 1e-3, cgAccuracy 1e-2, "
                         "minGridSize 20) at 32^3 x 48: GPU and the CPU oracle; prints "
                         "its own JSON line")
    ap.add_argument("--cpu-calibration", action="store_true",
                    help="CPU only: the oracle's full 32^4 solve timed against the "
                         "bounded-sample estimate the CPU baseline uses; prints JSON")
    ap.add_argument("--mesh-workload", action="store_true",
                    help="createMesh of warped 128^3 frames (a short solve first); "
                         "prints its own JSON line")
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

    if args.cpu_calibration:
        print(json.dumps(cpu_calibration(args.cpu_iters)), flush=True)
        return
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
    if args.flof_workload:
        print(json.dumps(flof_workload(M, cpu=not args.no_cpu)), flush=True)
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
    if args.mesh_workload:
        M.opticalFlowMultiscale4d(vel=vel, i0=i0, i1=i1, cgMaxIter=50, **PARAMS)
        M.sync()
        print(json.dumps(mesh_workload(M, L, ctx, s, vel)), flush=True)
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
    out["parity"] = parity_block(M, vcap)

    if not args.no_frames:
        out["frame_warp"] = frame_warp(M, L, ctx, s, vel, FLOF_KT_WARP)

    if not args.no_extras:
        out["mixed_precision"] = mixed_precision(M)
        out["warm_start"] = warm_start_c5(M)
        out["batch"] = fused_batch(M, L)
        # the same batch at sizes whose levels cannot fill the chip alone,
        # where sharing launches pays
        out["batch_small"] = fused_batch(M, L, ns=(16, 24, 32), ts=(16, 24, 32), min_grid=8)
        out["flof_params_32x48"] = flof_workload(M, cpu=False)

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_iters, iters, cells, ms_step)
        out["speedup_vs_cpu"] = out["cpu_baseline"]["s_per_solve"] / (ms_step / 1000.0)
    print(json.dumps(out), flush=True)


def _golden():
    path = os.path.join(ROOT, "tests", "golden", "digests.json")
    return json.load(open(path)) if os.path.exists(path) else {}


def parity_block(M, vcap):
    """The 64^4 solve with the reference's 1000-iteration cap against the
    oracle's golden digest (checked live: a match means the flow is the
    oracle's bit for bit), and the recorded distances (cells) of that flow,
    and of the 32^4 one, to the FP64 oracle and to the oracle in the
    reference's serial summation order, for the flow and for the frame warp
    of the same drops."""
    import hashlib
    d = _golden()
    out = {}
    if "drop64_l3_cap1000" in d:
        out["digest_64x4_cap1000_matches"] = (
            hashlib.sha256(vcap.tobytes()).hexdigest() == d["drop64_l3_cap1000"]["sha256"])
    for k in ("drop32_l3", "drop64_l3_cap1000"):
        c = d.get(k, {})
        if "distances" not in c:
            continue
        out[k] = {"iters": c["iters"], "iters_fp64": c["fp64_oracle"]["iters"],
                  "flow": c["distances"]}
        if "warp" in c:
            out[k]["warp_frames"] = {kk: c["warp"][kk] for kk in
                                     ("frames", "gpu_vs_fp64", "serial_vs_fp64",
                                      "gpu_vs_serial")}
    out["note"] = ("gpu_vs_*: the GPU flow (= the correctly rounded oracle's, digest) "
                   "against FP64 and against the reference's serial FP64 sums; "
                   "serial_vs_fp64: the reference order's own distance to FP64")
    return out


def _drop_pair(M, n, t, seeds, nspheres, rmin=0.08, rmax=0.18):
    sv = M.Solver(gridSize=M.vec3(n, n, n), fourthDim=t)
    a, b = sv.create(M.Grid4Real), sv.create(M.Grid4Real)
    M.dropSdf4d(a, seed=seeds[0], nspheres=nspheres, rmin=rmin, rmax=rmax)
    M.dropSdf4d(b, seed=seeds[1], nspheres=nspheres, rmin=rmin, rmax=rmax)
    return sv, a, b


def mixed_precision(M, sizes=((96, 64),)):
    """The mixed-precision solve (bfloat16 images, binary16 gradients of i1,
    FP32 stencil arithmetic, FP64 CG scalars) against the FP32 solve of the
    same 2-sphere drop pair (radii 0.08N-0.18N, seeds 7, 8), P_plain, 3
    levels, residual 1e-5; with the 32^3 x 32 distances of both to the FP64
    oracle recorded by tests/golden/make_digests.py (the GPU reproduces that
    oracle's mixed-precision flow bit for bit)."""
    from ofblend_amd.abi import FLOF_PREC_IMG_BF16
    out = {"workload": "2-sphere drop pair (seeds 7, 8; radii 0.08N-0.18N), P_plain, "
                       "3 levels, relative residual 1e-5"}
    for (n, t) in sizes:
        sv, a, b = _drop_pair(M, n, t, (7, 8), 2)
        v = sv.create(M.Grid4Vec4)
        prm = dict(PARAMS, minGridSize=n // 4, cgMaxIter=FIXED_RESIDUAL_CAP)
        res, flows = {}, {}
        for name, prec in (("fp32", 0), ("mixed", FLOF_PREC_IMG_BF16)):
            v.clear()
            M.sync()
            t0 = time.perf_counter()
            e = M.opticalFlowMultiscale4d(vel=v, i0=a, i1=b, precision=prec, **prm)
            M.sync()
            dt = time.perf_counter() - t0
            info = M.getLastSolveInfo().summary()
            flows[name] = v.to_numpy()
            res[name] = {"ms": 1000 * dt,
                         "iters_per_level": [l["iters"][0] for l in info["levels"]],
                         "final_err": e}
        d = flows["mixed"] - flows["fp32"]
        res["mixed_vs_fp32"] = {"max": float(np.abs(d).max()),
                                "rms": float(np.sqrt((d ** 2).mean()))}
        out["%dx%dx%dx%d" % (n, n, n, t)] = res
        del flows, d, v, a, b, sv
        log("bench: mixed precision %d^3 x %d done" % (n, t))
    g = _golden().get("mixed32_l3")
    if g:
        out["32x32x32x32_vs_fp64_oracle"] = {
            "mixed": g["distances"]["mixed_vs_fp64"], "fp32": g["distances"]["fp32_vs_fp64"],
            "iters_mixed": g["iters"], "iters_fp32": g["iters_fp32"],
            "iters_fp64": g["iters_fp64"], "absmax_fp64": g["absmax_fp64"]}
    return out


def precision_workload(M):
    """The mixed-precision section at 96^3 x 64 and 32^3 x 32 on its own."""
    print(json.dumps(mixed_precision(M, sizes=((96, 64), (32, 32)))), flush=True)


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


def frame_warp(M, L, ctx, s, vel, kind, keep=None):
    """The frame warp workload: 64 analytic 3D frames at 128^3 placed into a
    128^3 x 64 level set and warped by the solved 64^4 deformation (alpha 1,
    defoFactor (2,2,2,1)).  keep (a dict) receives the warped frames."""
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
    # one launch warps all N frames: it writes N frames and needs the 4D
    # level set (N time slices of M3 cells) and the deformation (N slices of
    # Md Vec4 cells) once each -- consecutive frames share their slices
    byts = N * 4 * M3 + N * 4 * M3 + N * 16 * Md
    if keep is not None:
        keep["frames"] = dst
    return {"frames": N, "frame": [F * N] * 3, "ms_per_64_frames": 1000 * wall,
            "frames_per_s": N / wall, "kernel_ms": 1000 * kern_s,
            "bytes_per_launch": byts,
            "GBps_algorithmic": byts / kern_s / 1e9, "frac": byts / kern_s / HBM_PEAK,
            "place_ms": 1000 * t_place}


def mesh_workload(M, L, ctx, s, vel, reps=5):
    """createMesh of warped output frames (the warp workload's frames 0, 31
    and 63 at 128^3): wall time per mesh (the mesher synchronises: count
    pass, two scans read back to size the arrays, vertices, triangles,
    normals), nodes and triangles."""
    from ofblend_amd.abi import FLOF_KT_WARP
    keep = {}
    frame_warp(M, L, ctx, s, vel, FLOF_KT_WARP, keep)
    dst = keep["frames"]
    F = 2
    M3 = (F * N) ** 3
    out = {"frame": [F * N] * 3, "frames": {}}
    for t in (0, N // 2 - 1, N - 1):
        ptr = C.c_void_p(dst.ptr.value + t * M3 * 4)
        h = C.c_void_p()
        _lib_check(L.flof_create_mesh(ctx, ptr, F * N, F * N, F * N, C.byref(h)), ctx)
        L.flof_mesh_destroy(h)  # warm-up
        t0 = time.perf_counter()
        for _ in range(reps):
            _lib_check(L.flof_create_mesh(ctx, ptr, F * N, F * N, F * N, C.byref(h)), ctx)
            if _ != reps - 1:
                L.flof_mesh_destroy(h)
        wall = (time.perf_counter() - t0) / reps
        nv, nt = C.c_int64(), C.c_int64()
        L.flof_mesh_size(h, C.byref(nv), C.byref(nt))
        L.flof_mesh_destroy(h)
        out["frames"][str(t)] = {"ms": 1000 * wall, "nodes": nv.value, "triangles": nt.value,
                                 "cells_per_s": M3 / wall}
    return out


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
        ("gaussianBlur4d Vec4 sigma 2 (625 taps)",
         lambda: M.gaussianBlur4d(tmp_v, sigma=2.0), 64 * cells,
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


def _rms_sdf(M, sv, flow, i0, h1):
    """RMS of i0 warped by the flow (advect4d) minus i1, over all cells."""
    w = sv.create(M.Grid4Real)
    w.copyFrom(i0)
    M.advect4d(flow, w)
    d = w.to_numpy().astype(np.float64) - h1
    return float(np.sqrt(np.mean(d * d)))


def warm_start_c5(M, n=128, t=64, seeds=(9001, 9002), nspheres=2):
    """The learned warm start at 128^3 x 64 (a drop pair the network was not
    trained on): the network (trained on the finest-level increments of 2000
    plain 64^4 solves, weights in ofblend_amd/data) predicts each level's
    increment from that level's pre-warped i0 and i1; the CG refines from it
    to the same residual (1e-5), falling back to zero where the guess's
    residual is larger.  Against the cold solve: iterations per level, time,
    and the SDF-alignment error of both (calcLsDiff and the RMS of i0 warped
    by the flow minus i1)."""
    from ofblend_amd import warmstart as WS
    if not os.path.exists(WS.NET64_WEIGHTS):
        return None
    net = WS.WarmNet(seed=1)
    net.load(WS.NET64_WEIGHTS)
    sv, a, b = _drop_pair(M, n, t, seeds, nspheres)
    v = sv.create(M.Grid4Vec4)
    prm = dict(PARAMS, minGridSize=16, cgMaxIter=FIXED_RESIDUAL_CAP)
    h1 = b.to_numpy()
    out = {"grid": [n, n, n, t], "seeds": list(seeds), "spheres": nspheres}
    for name in ("cold", "net_all_levels"):
        v.clear()
        M.sync()
        t0 = time.perf_counter()
        if name == "cold":
            e = M.opticalFlowMultiscale4d(vel=v, i0=a, i1=b, **prm)
            info = M.getLastSolveInfo().summary()
        else:
            e, inf = net.solve(v, a, b, levels=-1, **prm)
            info = inf.summary()
        M.sync()
        dt = time.perf_counter() - t0
        out[name] = {"ms": 1000 * dt, "iters_per_level": [l["iters"][0] for l in info["levels"]],
                     "warm_kept_per_level": [l["warm_used"] for l in info["levels"]],
                     "residual_per_level": [l["residual"][0] for l in info["levels"]],
                     "final_err": e, "rms_sdf_diff": _rms_sdf(M, sv, v, a, h1)}
        log("bench: warm start %s %.2f s iters %s" % (name, dt, out[name]["iters_per_level"]))
    c, w = out["cold"], out["net_all_levels"]
    out["finest_level_iteration_ratio"] = c["iters_per_level"][0] / max(1, w["iters_per_level"][0])
    out["total_iteration_ratio"] = sum(c["iters_per_level"]) / max(1, sum(w["iters_per_level"]))
    return out


def fused_batch(M, L, ns=(48, 64, 80), ts=(32, 48, 64), min_grid=None):
    """The 16-solve batch (forward and backward of 8 seeded drop pairs of
    1-3 spheres, N in ns, T in ts; by default {48,64,80} and {32,48,64}):
    one solve after another against the fused batch (every PCG launch spans
    the levels of equal index of all problems); the results must be
    identical."""
    rng = np.random.default_rng(4)
    jobs = []
    for k in range(8):
        n = int(rng.choice(list(ns)))
        t = int(rng.choice(list(ts)))
        sv, a, b = _drop_pair(M, n, t, (100 + 2 * k, 101 + 2 * k), 1 + k % 3)
        for (x, y) in ((a, b), (b, a)):
            jobs.append((sv, x, y, sv.create(M.Grid4Vec4), (n, t)))
    prm = dict(PARAMS, cgMaxIter=FIXED_RESIDUAL_CAP)
    if min_grid:
        prm["minGridSize"] = min_grid
    # launch shapes per grid shape from the autotuner (both runs)
    L.flof_ctx_set_autotune(M.context(), 1)
    M.sync()
    t0 = time.perf_counter()
    seq = []
    for sv, x, y, v, _ in jobs:
        v.clear()
        M.opticalFlowMultiscale4d(vel=v, i0=x, i1=y, **prm)
        seq.append(v.to_numpy())
    M.sync()
    t_seq = time.perf_counter() - t0
    for _, _, _, v, _ in jobs:
        v.clear()
    M.sync()
    t0 = time.perf_counter()
    errs, infos = M.opticalFlowMultiscale4dFused([j[3] for j in jobs], [j[1] for j in jobs],
                                                 [j[2] for j in jobs], **prm)
    M.sync()
    t_f = time.perf_counter() - t0
    tuned = {}
    for n, t in sorted(set(j[4] for j in jobs)):
        tb, lb, sl = C.c_int(), C.c_int(), C.c_int()
        if L.flof_ctx_tuned_shape(M.context(), n, n, n, t, C.byref(tb), C.byref(lb),
                                  C.byref(sl)) == 0:
            tuned["%dx%d" % (n, t)] = {"tmarch_wg_per_cu": tb.value,
                                       "lds_items_per_cu": lb.value, "lds_t_segment": sl.value}
    L.flof_ctx_set_autotune(M.context(), 0)
    same = all(np.array_equal(j[3].to_numpy(), seq[q]) for q, j in enumerate(jobs))
    log("bench: batch sequential %.2f s fused %.2f s identical %s" % (t_seq, t_f, same))
    cells = sum(j[4][0] ** 3 * j[4][1] for j in jobs)
    return {"workload": "16 solves: forward+backward of 8 seeded drop pairs (1-3 spheres), "
                        "N in %s, T in %s, P_plain, residual 1e-5" % (list(ns), list(ts)),
            "sizes": [j[4] for j in jobs], "total_cells": cells,
            "sequential_s": t_seq, "fused_s": t_f, "fused_solves_per_s": len(jobs) / t_f,
            "speedup_vs_sequential": t_seq / t_f, "identical_to_sequential": same,
            "autotuned_finest_shapes": tuned}


FLOF_PARAMS = dict(wSmooth=0.001, wEnergy=0.0001, cgAccuracy=0.01, postVelBlur=4.0,
                   cfl=999.0, multiStep=3, minGridSize=20, doFinalProject=True,
                   resetBndWidth=0.1)


def flof_workload(M, cpu=True):
    """flof.py's mode-1 solve (scenes/flof.py:913-915: wSmooth 1e-3, wEnergy
    1e-4, cgAccuracy 1e-2, postVelBlur 4, multiStep 3, minGridSize 20, final
    projection, band 0.1) at its 32^3 x 48 grid, on the seeded drop pair
    (seeds 1, 2) instead of the FLIP simulations: the GPU solve (median of 3)
    and, with cpu, one solve of the CPU oracle in the reference's order on
    this host's cores (the reference quotes "circa 2-3 minutes" for the whole
    mode-1 run, README.md:35)."""
    n, t = 32, 48
    sv, a, b = _drop_pair(M, n, t, (1, 2), 1, rmin=0.15, rmax=0.15)
    v = sv.create(M.Grid4Vec4)
    times = []
    for _ in range(4):
        v.clear()
        M.sync()
        t0 = time.perf_counter()
        e = M.opticalFlowMultiscale4d(vel=v, i0=a, i1=b, **FLOF_PARAMS)
        M.sync()
        times.append(time.perf_counter() - t0)
    info = M.getLastSolveInfo().summary()
    out = {"grid": [n, n, n, t], "params": FLOF_PARAMS,
           "gpu_ms": 1000 * float(np.median(times[1:])), "final_err": e,
           "iters_per_level_and_step": [l["iters"] for l in info["levels"]]}
    if cpu:
        from oracle import oracle as O
        from ofblend_amd.abi import default_params
        i0, i1 = a.to_numpy(), b.to_numpy()
        t0 = time.perf_counter()
        ref, rerr, _ = O.multiscale4d(np.zeros(i0.shape + (4,), np.float32), i0, i1,
                                      default_params(**FLOF_PARAMS))
        out["cpu_s"] = time.perf_counter() - t0
        out["cpu_threads"] = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out["nproc"] = os.cpu_count()
        out["cpu_kind"] = "port (the oracle: serial CG, OpenMP over t in the KERNEL loops)"
        out["speedup"] = out["cpu_s"] / (out["gpu_ms"] / 1000)
        d = v.to_numpy() - ref
        out["max_abs_diff_gpu_vs_reference_order"] = float(np.abs(d).max())
        log("bench: flof.py parameters: gpu %.1f ms, cpu %.1f s" % (out["gpu_ms"], out["cpu_s"]))
    return out


def cpu_calibration(k_iters):
    """The CPU baseline's bounded sample checked against a full solve: at
    32^4 (P_plain, 3 levels) the oracle's whole solve is timed and compared
    with the estimate from a finest-level sample of k and 2k iterations
    scaled by the same per-level iteration counts (the method cpu_baseline
    applies at 64^4)."""
    from oracle import oracle as O
    from ofblend_amd.abi import default_params
    n = 32
    i0 = O.drop_sdf(n, n, n, n, seed=1)
    i1 = O.drop_sdf(n, n, n, n, seed=2)
    t0 = time.perf_counter()
    _, _, info = O.multiscale4d(np.zeros(i0.shape + (4,), np.float32), i0, i1,
                                default_params(**PARAMS_C1))
    full = time.perf_counter() - t0
    lv = info.summary()["levels"]
    iters = [l["iters"][0] for l in lv]
    cells = [int(np.prod(l["dims"])) for l in lv]
    est, tt = _cpu_estimate(i0, i1, k_iters, iters, cells)
    return {"grid": [n] * 4, "full_solve_s": full, "estimate_s": est,
            "estimate_over_full": est / full, "iters_per_level": iters,
            "sample_s": sum(tt), "nproc": os.cpu_count(),
            "omp_threads": os.environ.get("OMP_NUM_THREADS")}


PARAMS_C1 = dict(wSmooth=0.1, wEnergy=1e-4, cgAccuracy=1e-5, minGridSize=8, multiStep=1,
                 postVelBlur=0.0, doFinalProject=False, resetBndWidth=-1.0, cfl=999.0)


def _cpu_estimate(i0, i1, k_iters, iters, cells):
    """Finest-level set-up + k and 2k PCG iterations of the oracle; the
    difference gives the time per iteration, scaled to every level."""
    from oracle import oracle as O
    from ofblend_amd.abi import default_params
    v = np.zeros(i0.shape + (4,), np.float32)
    tt = []
    for k in (k_iters, 2 * k_iters):
        p = default_params(wSmooth=0.1, wEnergy=1e-4, cgAccuracy=1e-5, cgMaxIter=k)
        t0 = time.perf_counter()
        O.optical_flow4d(v, i0, i1, p)
        tt.append(time.perf_counter() - t0)
        log("bench: cpu oracle %s level solve, %d iterations: %.2f s" % (i0.shape, k, tt[-1]))
    t_it = (tt[1] - tt[0]) / k_iters
    t_setup = max(0.0, tt[0] - k_iters * t_it)
    per_cell_it = t_it / cells[0]
    per_cell_setup = t_setup / cells[0]
    est = sum(i * c * per_cell_it + c * per_cell_setup for i, c in zip(iters, cells))
    return est, tt


def cpu_baseline(k_iters, iters, cells, gpu_ms):
    """The CPU oracle (FP32, the reference's serial CG and FP64 dot order)
    timed on a bounded sample: the finest-level set-up plus PCG at 64^4 with
    an iteration cap of k and 2k; the difference gives the time per iteration,
    scaled to a full solve with the GPU's per-level iteration counts (the
    same estimate matches a full 32^4 solve: bench.py --cpu-calibration).
    The CG is serial like the reference's; the set-up loops use OpenMP."""
    from oracle import oracle as O
    threads = min(16, os.cpu_count() or 1)
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    i0 = O.drop_sdf(N, N, N, N, seed=1)
    i1 = O.drop_sdf(N, N, N, N, seed=2)
    est, tt = _cpu_estimate(i0, i1, k_iters, iters, cells)
    return {"value": 1.0 / est, "unit": "solves/s", "cores": 1, "kind": "port",
            "s_per_solve": est, "nproc": os.cpu_count(),
            "omp_threads_setup": int(os.environ["OMP_NUM_THREADS"]),
            "sample": "CPU oracle (FP32-faithful, serial CG as the reference: one core "
                      "does the PCG) on the 64^4 finest level: set-up + %d and %d PCG "
                      "iterations (%.1f s), scaled to the solve's per-level iteration "
                      "counts %s" % (k_iters, 2 * k_iters, sum(tt), iters)}


if __name__ == "__main__":
    main()
