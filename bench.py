"""Benchmark of the Wilson-flow hot path on one MI355X.

Default workload: 32^3 x 64 lattice at beta = 6.2, thermalised on the GPU by
1000 Philox checkerboard Metropolis sweeps (seed 62, 10 hits, eps 0.24,
exp(i eps H) candidates) from a Haar-random start, then the mixed-precision Wilson
flow (FP32 links/staples/clover; FP64 exponential, reunitarisation and
compensated reductions) with eps = 0.02, observables every 10 steps.  A step
is one RK3 flow step over the whole lattice (4V link-updates).  The same
config is then flowed in plain FP64 for the speed ratio and the error
statement.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""
import argparse
import ctypes
import json
import math
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "flow link-updates/s (FP32-mixed vs FP64 @ rel ΔE≤1e-6); batched configs/s"
HBM_PEAK_GBS = 8000.0
VALU_PEAK_TFLOPS = {"fp64": 78.6, "fp32": 78.6}  # non-packed FP32 issues at the FP64 rate
# algorithmic bytes per link-update (one link through one RK3 step): links 18
# reals read + 18 written per stage, RK accumulator 8 reals written (stage 1),
# read + written (stage 2), read (stage 3): 140 reals.
REALS_PER_LINK_UPDATE = 140
STAGE_REALS = (44, 52, 44)
# VALU arithmetic per link-update (three RK stages), from the rocprofv3 VALU
# instruction counters of the stage kernels at 32^3x64 (64 lanes x (ADD + MUL
# + 2 FMA), FP32 and FP64 summed; per-stage figures in PROFILE_FILE): the
# mixed path (two-row staples, links in two-row form with row 2 rebuilt where
# read) 3.08 kflop per link per stage, FP64 (all three rows multiplied out)
# 3.17 kflop.
FLOP_PER_LINK_UPDATE = {"mixed": 9245, "fp64": 9513}
# HBM-side bytes per k_flow_stage launch, averaged over the three RK stages,
# from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of this bench (fetch
# scaled by 2.0 and writes by 1.0, the factors a 1 GiB streaming copy reads
# back), per schedule, since the schedule sets how often neighbour links come
# from HBM rather than L2; the runs, their commands and the per-stage figures
# are in the profile files.  The mixed flow keeps its links as two rows (12
# reals), so it moves fewer bytes than the canonical 18-real count of
# REALS_PER_LINK_UPDATE.  A schedule not profiled reports traffic null.
PROFILE_FILE = "profiles/k_flow_stage_32x32x32x64.md"
PROFILE_FILE_48 = "profiles/k_flow_stage_48x48x48x96.md"
PROFILED_TRAFFIC = {
    ("mixed-32x64", "mixed", "column+lds+2row target=4096"): (1.444e9, PROFILE_FILE),
    ("mixed-32x64", "mixed", "linear+lds+2row target=0"): (1.872e9, PROFILE_FILE),
    ("mixed-32x64", "fp64", "column+lds target=4096"): (6.396e9, PROFILE_FILE),
    ("mixed-48x96", "mixed", "column+lds+2row target=4096"): (8.530e9, PROFILE_FILE_48),
    ("mixed-48x96", "mixed", "column+lds+2row target=1024"): (8.736e9, PROFILE_FILE_48),
    ("mixed-48x96", "fp64", "column+lds target=1024"): (25.81e9, PROFILE_FILE_48),
}


def schedule_label(tuner):
    """'column+lds+2row target=4096' from an autotuner summary line."""
    f = dict(kv.split("=", 1) for kv in tuner.split() if "=" in kv)
    return f"{f.get('schedule')} target={f.get('target')}"


WORKLOADS = {
    # 1000 sweeps from a Haar-random start put t0 (t^2 <E> = 0.3) near t = 5.1,
    # inside the flow range; 200 sweeps from a cold start leave the long-distance
    # modes so ordered that t^2 <E> is still 0.15 at t = 6.
    "mixed-32x64": dict(dims=(32, 32, 32, 64), beta=6.2, sweeps=1000, seed=62, eps=0.02, tau=6.0, every=10,
                        precision="mixed", start="hot"),
    "fp64-24": dict(dims=(24, 24, 24, 24), beta=6.0, sweeps=500, seed=7, eps=0.01, tau=4.0, every=10,
                    precision="fp64"),
    "mixed-48x96": dict(dims=(48, 48, 48, 96), beta=6.4, sweeps=200, seed=64, eps=0.02, tau=8.0, every=5,
                        precision="mixed"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed_flow(ctx, eps, tau, every):
    ctx.synchronize()
    t0 = time.perf_counter()
    series = ctx.flow(eps, tau, every)
    ctx.synchronize()
    return series, time.perf_counter() - t0


def _oracle():
    so = os.path.join(ROOT, "oracle", "_build", "liblym_oracle.so")
    if not os.path.isfile(so):
        return None
    orc = ctypes.CDLL(so)
    D = ctypes.POINTER(ctypes.c_double)
    orc.or_flow_step.argtypes = [D, ctypes.c_int * 4, ctypes.c_double]
    orc.or_flow_step_lattice.argtypes = [D, ctypes.c_int * 4, ctypes.c_double]
    orc.or_observables.argtypes = [D, ctypes.c_int * 4, ctypes.c_int, D]
    orc.or_hot_start.argtypes = [D, ctypes.c_int * 4, ctypes.c_uint64]
    return orc


def cpu_baseline(links, dims, eps, beta, budget_s=30.0):
    """The oracle's restatement of the reference's flow step in the
    reference's own structure (FP64, one thread, whole-lattice temporaries
    and shift copies, the reference's operation order; equal bit for bit to
    the per-site oracle) on the benchmark's own lattice and configuration:
    whole flow steps timed with a wall clock, as the reference times its
    runs; as many steps as fit the budget, at least one."""
    orc = _oracle()
    if orc is None:
        return None
    D = ctypes.POINTER(ctypes.c_double)
    work = np.array(links, copy=True)
    V = int(np.prod(dims))
    d = (ctypes.c_int * 4)(*dims)
    steps, t0 = 0, time.perf_counter()
    while True:
        assert orc.or_flow_step_lattice(work.ctypes.data_as(D), d, eps) == 0
        steps += 1
        el = time.perf_counter() - t0
        if el + el / steps > budget_s:
            break
    return {"value": 4 * V * steps / el, "unit": "link-updates/s", "cores": 1, "kind": "port",
            "sample": f"oracle FP64 RK3 flow in the reference's whole-lattice structure, {steps} step(s) of "
                      f"eps={eps} on the benchmark's GPU-thermalised {'x'.join(map(str, dims))} beta={beta} "
                      f"configuration ({el:.1f} s wall)",
            "cpu": platform.processor() or platform.machine(), "nproc": os.cpu_count()}


def _ensemble_cpu_worker(args):
    """One process of the ensemble CPU leg: oracle flow steps and one
    measurement on a Haar-random lattice of the given shape, timed."""
    dims, seed, nsteps, eps = args
    orc = _oracle()
    D = ctypes.POINTER(ctypes.c_double)
    V = int(np.prod(dims))
    links = np.empty((4, V, 18))
    d = (ctypes.c_int * 4)(*dims)
    orc.or_hot_start(links.ctypes.data_as(D), d, seed)
    t0 = time.perf_counter()
    for _ in range(nsteps):
        assert orc.or_flow_step(links.ctypes.data_as(D), d, eps) == 0
    t1 = time.perf_counter()
    out = np.empty(4)
    orc.or_observables(links.ctypes.data_as(D), d, 0, out.ctypes.data_as(D))
    t2 = time.perf_counter()
    return tuple(dims), (t1 - t0) / nsteps, t2 - t1


def cpu_ensemble_baseline(eps=0.01, tau=2.0, every=10, max_procs=16):
    """The ensemble's CPU leg: P single-thread oracle processes at once (P = the
    cores this process may use, at most max_procs), one configuration each, the
    four lattice shapes of the ensemble in turn.  Each process times two flow
    steps and one measurement of its lattice; a whole configuration costs
    (int(tau/eps) - 1) steps and one measurement per `every` steps, so P
    processes flow P / (mean configuration time) configurations per second.
    Run before the GPU is touched (the workers are forked)."""
    import multiprocessing as mp
    if _oracle() is None:
        return None
    procs = max(1, min(len(os.sched_getaffinity(0)), max_procs))
    shapes = [(8, 8, 8, 8), (12, 12, 12, 12), (16, 16, 16, 16), (16, 16, 16, 32)]
    jobs = [(shapes[i % 4], i, 2, eps) for i in range(max(procs, 4))]
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(procs) as pool:
        res = pool.map(_ensemble_cpu_worker, jobs)
    wall = time.perf_counter() - t0
    steps = int(tau / eps) - 1
    nmeas = 1 + steps // every
    per_shape = {}
    for dims, t_step, t_obs in res:
        per_shape.setdefault(dims, []).append(steps * t_step + nmeas * t_obs)
    conf_s = float(np.mean([np.mean(v) for v in per_shape.values()]))
    return {"value": procs / conf_s, "unit": "configs/s", "cores": procs, "kind": "port",
            "sample": f"{len(jobs)} oracle processes ({procs} at once), each 2 flow steps + 1 measurement of one "
                      f"Haar-random lattice (8^4, 12^4, 16^4, 16^3x32 in turn; {wall:.1f} s wall); a configuration "
                      f"= {steps} steps + {nmeas} measurements, mean {conf_s:.1f} s",
            "seconds_per_config": {"x".join(map(str, k)): float(np.mean(v)) for k, v in per_shape.items()}}


BATCH_DIMS = [(8, 8, 8, 8)] * 16 + [(12, 12, 12, 12)] * 16 + [(16, 16, 16, 16)] * 16 + [(16, 16, 16, 32)] * 16


def batched_ensemble(tune_cache, every=10, use_graph=True):
    """64 uneven lattices (16 each of 8^4, 12^4, 16^4, 16^3x32), beta = 6.0,
    Haar hot starts with seeds 0..63, 100 Metropolis sweeps, then the mixed
    flow eps = 0.01 to t = 2.0 (199 steps, observables every `every` steps),
    graph-captured, in the launch form and schedule the autotuner picks (one
    launch per RK stage, or one persistent launch per measurement interval);
    both launch forms are also timed on their own; against the
    64 lattices flowed one after another by the plain FP64 single-lattice path,
    and (to separate the batching gain from the precision gain) by the mixed
    single-lattice path.  use_graph=False launches the batch eagerly (for a
    kernel trace of its launches one by one)."""
    from latticeyangmills_amd import LatticeContext
    from latticeyangmills_amd.batch import BatchContext
    eps, tau = 0.01, 2.0
    seeds = list(range(len(BATCH_DIMS)))
    with BatchContext(BATCH_DIMS, "fp64") as th:
        th.set_hot(seeds)
        th.metropolis(6.0, 0.24, 10, 100, seeds)
        confs = [th.download(c) for c in range(len(BATCH_DIMS))]
    with BatchContext(BATCH_DIMS, "mixed") as b:
        for c, L in enumerate(confs):
            b.upload(c, L)
        tuner = b.autotune(tune_cache)
        b.flow(eps, tau, every, use_graph=use_graph)  # captures the graph
        for c, L in enumerate(confs):
            b.upload(c, L)
        t0 = time.perf_counter()
        series_b = b.flow(eps, tau, every, use_graph=use_graph)
        t_batch = time.perf_counter() - t0
        # both launch forms of the tuned schedule, best of three graph replays
        # each: one launch per RK stage, and one persistent launch per
        # measurement interval (grid barriers between stages)
        tuned = b.info()
        forms = {}
        for whole in (False, True):
            b.set_schedule(0 if whole else tuned["grid"], tuned["kind"], tuned["target"], tuned["lds"],
                           tuned["two_row"], whole_flow=whole)
            best = None
            for rep in range(4):
                for c, L in enumerate(confs):
                    b.upload(c, L)
                t0 = time.perf_counter()
                b.flow(eps, tau, every, use_graph=use_graph)
                el = time.perf_counter() - t0
                if rep:  # the first run captures the graph
                    best = el if best is None else min(best, el)
            forms["whole_flow" if whole else "per_stage"] = {
                "configs_per_s": len(BATCH_DIMS) / best, "seconds": best, "grid": b.info()["grid"]}
        forms["note"] = ("per_stage: three launches per flow step (597 for the flow), one workgroup per schedule slot "
                         "unless the tuner chose a persistent grid; whole_flow: one persistent launch per 10-step "
                         "measurement interval (20 launches), grid-wide barriers between RK stages, as many "
                         "workgroups as the device holds at once")
    def sequential(precision):
        ctxs = []
        for c, L in enumerate(confs):
            ctx = LatticeContext(BATCH_DIMS[c], precision)
            ctx.upload(L)
            ctxs.append(ctx)
        ctxs[0].flow(eps, 0.05, every)  # warm the code objects
        ctxs[0].upload(confs[0])
        t0 = time.perf_counter()
        series = [ctx.flow(eps, tau, every) for ctx in ctxs]
        el = time.perf_counter() - t0
        for ctx in ctxs:
            ctx.close()
        return series, el

    series_s, t_seq = sequential("fp64")
    _, t_seq_mixed = sequential("mixed")
    e_err = max(float(np.nanmax(np.abs(series_b[c][:, 3] / series_s[c][:, 3] - 1))) for c in range(len(confs)))
    n = len(BATCH_DIMS)
    return {"configs_per_s": n / t_batch, "seconds": t_batch, "sequential_plain_fp64_configs_per_s": n / t_seq,
            "sequential_seconds": t_seq, "speedup": t_seq / t_batch, "tuner": tuner,
            "sequential_mixed_configs_per_s": n / t_seq_mixed, "speedup_vs_sequential_mixed": t_seq_mixed / t_batch,
            "max_rel_err_E_vs_fp64": e_err, "graph": use_graph, "launch_forms": forms,
            "workload": "64 lattices: 16 x 8^4, 16 x 12^4, 16 x 16^4, 16 x 16^3x32; beta 6.0; Haar seeds 0..63; "
                        f"100 sweeps; mixed flow eps {eps} to t {tau} (199 steps), observables every {every}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed flow steps (default: the whole flow)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="mixed-32x64", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--sweeps", type=int, default=None, help="override the Metropolis sweep count")
    ap.add_argument("--calibrate", action="store_true", help="also time a 1 GiB streaming copy (HBM probe)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-ensemble measurement")
    ap.add_argument("--batch-eager", action="store_true", help="launch the batched flow eagerly, not from its graph")
    ap.add_argument("--tune-cache", default=None, help="autotuner cache file (default: tune afresh)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design note only",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    # the ensemble's CPU leg forks its worker processes, so it runs before
    # anything touches the GPU
    cpu_ens = None if (args.no_cpu or args.no_batch) else cpu_ensemble_baseline()
    if cpu_ens:
        log(f"[bench] CPU ensemble leg: {cpu_ens['value']:.4f} configs/s on {cpu_ens['cores']} processes")

    from latticeyangmills_amd import LatticeContext
    w = WORKLOADS[args.workload]
    dims, eps, every = w["dims"], w["eps"], w["every"]
    V = int(np.prod(dims))
    full_steps = int(w["tau"] / eps) - 1  # reference semantics: int(tau/eps) - 1 steps
    steps = args.steps if args.steps else full_steps
    # int(tau / eps) = steps + 1 rows, whatever the rounding of (steps + 1) * eps
    tau = (steps + 1.5) * eps
    log(f"[bench] workload {args.workload} dims={dims} V={V} steps={steps} warmup={args.warmup}")

    # thermalise in FP64 on the GPU
    t = time.perf_counter()
    with LatticeContext(dims, "fp64") as th:
        sweeps = args.sweeps if args.sweeps is not None else w["sweeps"]
        if w.get("start") == "hot":
            th.set_hot(w["seed"])
        acc = th.metropolis(w["beta"], 0.24, 10, sweeps, 0, "exph", w["seed"])
        plaq0 = th.observables()[0]
        links = th.download()
    log(f"[bench] thermalised: {sweeps} sweeps in {time.perf_counter() - t:.1f} s, plaq {plaq0:.6f}, "
        f"acceptance {acc / max(sweeps * 4 * V * 10, 1):.3f}")

    results = {}
    for prec in ([w["precision"], "fp64"] if w["precision"] != "fp64" else ["fp64"]):
        with LatticeContext(dims, prec) as ctx:
            # host <-> device transfer of the configuration at the boundary
            # (GluonField layout in host memory), for the transfer-inclusive rate
            ctx.synchronize()
            t = time.perf_counter()
            ctx.upload(links)
            ctx.synchronize()
            up_s = time.perf_counter() - t
            t = time.perf_counter()
            ctx.download()
            down_s = time.perf_counter() - t
            tuner = ctx.autotune(args.tune_cache)
            if args.warmup:
                # the warm-up runs every kernel of the timed flow (stages, form
                # conversions, observables, reduction) and sizes its measurement
                # buffer (one slot per step), so the timed call allocates nothing
                ctx.flow(eps, (args.warmup + 1.5) * eps, 1)
                ctx.synchronize()
                ctx.upload(links)
            series, el = timed_flow(ctx, eps, tau, every)
            assert series.shape[0] == steps + 1
            if steps != full_steps:
                # the error statement always covers the whole flow to t = (full_steps) eps,
                # whatever the timed step count (untimed)
                ctx.upload(links)
                series = ctx.flow(eps, w["tau"], every)
            ctx.upload(links)
            total_ms, st = ctx.profile_flow(eps, args.profile_steps, every)
        results[prec] = dict(series=series, seconds=el, stage_ms=st, ms_per_step=1e3 * el / steps, tuner=tuner,
                             upload_s=up_s, download_s=down_s)
        log(f"[bench] {prec}: {steps} steps in {el:.3f} s ({1e3 * el / steps:.3f} ms/step), stage ms {st}, {tuner}")

    main_prec = w["precision"]
    r = results[main_prec]
    lups = 4 * V * steps / r["seconds"]
    real = 4 if main_prec == "mixed" else 8
    st = r["stage_ms"]
    step_kernel_ms = float(st[0] + st[1] + st[2])
    bytes_step = 4 * V * REALS_PER_LINK_UPDATE * real
    # per launch, averaged over the three stage launches of a step
    achieved_gbs = bytes_step / (step_kernel_ms * 1e-3) / 1e9
    traffic, traffic_source = PROFILED_TRAFFIC.get((args.workload, main_prec, schedule_label(r["tuner"])),
                                                   (None, None))
    out = {
        "metric": METRIC, "value": lups, "unit": "link-updates/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64" if main_prec == "mixed" else "f64",
        "data": "synthetic: Philox-seeded checkerboard Metropolis thermalised on the GPU",
        "transfers": {"upload_ms": 1e3 * r["upload_s"], "download_ms": 1e3 * r["download_s"],
                      "link_updates_per_s_with_transfers": 4 * V * steps / (r["seconds"] + r["upload_s"] + r["download_s"]),
                      "note": "one upload and one download of the configuration (GluonField layout in host memory) "
                              "around the timed flow; value excludes them"},
        "config": {"workload": args.workload, "lattice": "x".join(map(str, dims)), "beta": w["beta"],
                   "start": w.get("start", "cold"), "metropolis_sweeps": sweeps, "seed": w["seed"], "plaquette_start": plaq0, "flow_eps": eps, "timed_flow_t_max": steps * eps,
                   "measure_every": every, "precision": main_prec},
        "roofline": {"bound": "hbm", "achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved_gbs / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_source,
                     "traffic_gbs": traffic / (step_kernel_ms / 3 * 1e-3) / 1e9 if traffic else None,
                     "algorithmic_bytes_per_launch": bytes_step / 3,
                     "kernel": "k_flow_stage (RK stages 1+2+3 per step)",
                     "stage_ms": [float(x) for x in st[:3]], "observables_ms": float(st[3]),
                     "valu_tflops": 4 * V * FLOP_PER_LINK_UPDATE[main_prec] / (step_kernel_ms * 1e-3) / 1e12,
                     "valu_peak_tflops": VALU_PEAK_TFLOPS["fp32" if main_prec == "mixed" else "fp64"],
                     "valu_frac": 4 * V * FLOP_PER_LINK_UPDATE[main_prec] / (step_kernel_ms * 1e-3) / 1e12
                     / VALU_PEAK_TFLOPS["fp32" if main_prec == "mixed" else "fp64"],
                     "schedule": r["tuner"]},
    }
    if "fp64" in results and main_prec != "fp64":
        f = results["fp64"]
        lups64 = 4 * V * steps / f["seconds"]
        sm, s64 = r["series"], f["series"]
        rows = [i for i in range(len(s64)) if math.isfinite(s64[i, 1])]
        e_err = max(abs(sm[i, 3] / s64[i, 3] - 1.0) for i in rows)
        q_err = max(abs(sm[i, 2] - s64[i, 2]) for i in rows)
        p_err = max(abs(sm[i, 1] / s64[i, 1] - 1.0) for i in rows)
        from latticeyangmills_amd import scale as _scale
        t2e = [(s64[i, 0] ** 2 * _scale.e_mean(sm[i, 3], V), s64[i, 0] ** 2 * _scale.e_mean(s64[i, 3], V)) for i in rows]
        t2e_err = max((abs(a / b - 1.0) for a, b in t2e if b != 0), default=None)
        from latticeyangmills_amd import scale
        t0m, t064 = scale.t0(sm, V), scale.t0(s64, V)
        w0m, w064 = scale.w0(sm, V), scale.w0(s64, V)
        a_fm = scale.lattice_spacing_fm(w["beta"])
        last = rows[-1]
        out["fp64"] = {"value": lups64, "ms_per_step": f["ms_per_step"], "speedup_mixed": lups / lups64,
                       "stage_ms": [float(x) for x in f["stage_ms"][:3]], "schedule": f["tuner"]}
        out["accuracy_vs_fp64"] = {
            "flow": f"whole flow, eps {eps} to t = {s64[-1, 0]:.2f} ({len(s64) - 1} steps), measured every {every} "
                    f"steps (last measured row t = {s64[last, 0]:.2f})",
            "max_rel_err_E": e_err, "max_rel_err_t2E": t2e_err, "max_rel_err_plaq": p_err, "max_abs_err_Q": q_err,
            "t0_a2_mixed": t0m, "t0_a2_fp64": t064,
            "t0_rel_err": (abs(t0m / t064 - 1) if (t0m and t064) else None),
            "w0_a_mixed": w0m, "w0_a_fp64": w064,
            "w0_rel_err": (abs(w0m / w064 - 1) if (w0m and w064) else None),
            "a_fm": a_fm, "t0_fm2": (t064 * a_fm * a_fm if t064 else None),
            "t2E_at_tmax": s64[last, 0] ** 2 * scale.e_mean(s64[last, 3], V), "Q_at_tmax": s64[last, 2],
        }
    if not args.no_batch:
        out["batched"] = batched_ensemble(args.tune_cache, use_graph=not args.batch_eager)
        if cpu_ens:
            out["batched"]["cpu_baseline"] = cpu_ens
        log(f"[bench] batched: {out['batched']}")
    if args.calibrate:
        from latticeyangmills_amd.capi import copy_bandwidth
        out["hbm_copy_gbs"] = copy_bandwidth(1 << 30, 10)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(links, dims, eps, w["beta"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
