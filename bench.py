#!/usr/bin/env python3
"""Benchmark of the device pipeline of World.internalStep on one MI355X.

Default workload: the batched box-stack / pile workload -- 4096 independent worlds, body count
per world log-uniform in [8, 512] (box stacks or piles of boxes and spheres over a plane),
GSSolver 10 iterations, dt = 1/60, FP32 -- advanced one launch per stage over every world.
A "step" is one World.step(dt) of every world of the batch.

Prints one JSON line: body-steps/s of the batched path (value), the plain one-world-per-launch
path on a sample of the same worlds, the roofline of the dominant stage kernel measured with
HIP events inside the timed region, and the CPU baseline: the C++ restatement of the
reference semantics (double arithmetic over FP32 storage) on the host cores.
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
# algorithmic bytes (per body, pair and row of one step): 13-float state read + written and
# 6 constants read per dynamic body; two int32 indices per pair; a 64-B row record written by
# the narrowphase and read by the solver plus a 4-B lambda written back per row
B_BODY, B_PAIR, B_ROW = 128, 16, 136
# the GS sweep reads each row record once and writes its lambda once (iterations are meant to
# stay on chip)
B_ROW_GS = 68


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_worlds(workload, nworlds, seed0=0):
    from cannon_physics_amd import scenes
    recs, dyn, desc = [], [], ""
    if workload == "c3":
        for i in range(nworlds):
            w = scenes.c3_world(seed0 + i)
            recs.append(w.recipe())
            dyn.append(len(w.bodies) - 1)
        desc = ("batched box stacks / piles: %d worlds, log-uniform [8,512] bodies, boxes "
                "(half-extents U[0.25,0.5]) and spheres (r U[0.2,0.4]) on a plane, NaiveBroadphase, "
                "GSSolver 10 iterations, dt 1/60" % nworlds)
        dt = 1 / 60
    elif workload == "c4":
        for i in range(nworlds):
            w = scenes.c4_chain_world(seed0 + i)
            recs.append(w.recipe())
            dyn.append(len(w.bodies) - 1)
        desc = ("batched jointed chains: %d worlds, U{4..64} box links (0.1x0.3x0.1) joined by "
                "PointToPoint / Hinge+motor / Distance / ConeTwist / Lock, on a plane, GSSolver 20 "
                "iterations, dt 1/120" % nworlds)
        dt = 1 / 120
    elif workload == "c5":
        # the served request mix as one batch (no serving loop): worlds of log-uniform [16,2048]
        # bodies over a heightfield, the serving leg's templates
        for i in range(nworlds):
            w = scenes.c5_terrain_world(5 * 1000 + i)
            recs.append(w.recipe())
            dyn.append(len(w.bodies) - 1)
        desc = ("batched heightfield worlds (the served request mix): %d worlds, log-uniform [16,2048] "
                "bodies (spheres, boxes, 8-16 vertex hulls) over a 64x64 heightfield, NaiveBroadphase, "
                "GSSolver 10 iterations, dt 1/60" % nworlds)
        dt = 1 / 60
    elif workload == "c1":
        w = scenes.c1_spheres()
        recs.append(w.recipe())
        dyn.append(216)
        desc = "216 spheres r=0.5 on a 6x6x6 grid over a plane, NaiveBroadphase, GSSolver 10, dt 1/60"
        dt = 1 / 60
    elif workload == "c2":
        w = scenes.c2_large_world()
        recs.append(w.recipe())
        dyn.append(len(w.bodies) - 1)
        desc = ("one world of 32,768 spheres / rotated boxes in 64x32x64 m over a plane, "
                "SAPBroadphase (auto axis), GSSolver 10, dt 1/60")
        dt = 1 / 60
    else:
        raise SystemExit("unknown workload " + workload)
    return recs, dyn, desc, dt


def copy_probe(nbytes):
    """One device-to-device copy of nbytes through the HIP runtime (its copy kernel reads and
    writes nbytes): the calibration point for the HBM byte counters."""
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(src, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    log("copy probe: %d bytes device to device" % nbytes)


def measured_traffic(workload, precision, kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    profiles/<kernel>_<workload>_traffic.md (FETCH_SIZE + WRITE_SIZE per dispatch, scaled by the
    same counters on a copy kernel of known size); None when there is none for this workload."""
    path = os.path.join(ROOT, "profiles", "%s_%s_%s_traffic.md" % (kernel, workload, precision))
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def stats_totals(ctx, wids):
    tot = {"sum_pairs": 0, "sum_rows": 0, "sum_contacts": 0, "steps": 0}
    for w in wids:
        s = ctx.stats(w)
        for k in ("sum_pairs", "sum_rows", "sum_contacts"):
            tot[k] += s[k]
        tot["steps"] = max(tot["steps"], s["steps"])
    return tot


def serve_bench(ntemplates, nrequests, budget_gib, tick_steps, seed=5, max_submit_s=60.0):
    """Serving: a Poisson stream of requests (worlds of log-uniform [16, 2048] bodies over a
    64x64 heightfield, K ~ U{60..600} steps at dt = 1/60) through chp_serve_*.  The arrival
    rate is set from Little's law so that the context's HBM budget is about 70% occupied.
    Request worlds are drawn from ntemplates pre-built recipes."""
    import numpy as np
    from cannon_physics_amd import scenes
    from cannon_physics_amd.world import Context
    t0 = time.perf_counter()
    temps = []
    for i in range(ntemplates):
        w = scenes.c5_terrain_world(seed * 1000 + i)
        temps.append((w.recipe(), len(w.bodies) - 1))
    log("serving: %d request templates built in %.1fs" % (ntemplates, time.perf_counter() - t0))
    budget = int(budget_gib * (1 << 30))
    ctx = Context(device=0, hbm_budget=budget)
    # bytes per template, for the arrival rate
    sizes = []
    for r, _ in temps:
        wid = ctx.create_world(r)
        sizes.append(ctx.bytes_in_use())
        ctx.destroy_world(wid)
    rng = np.random.default_rng(seed)
    K = rng.integers(60, 601, size=nrequests)
    pick = rng.integers(0, ntemplates, size=nrequests)
    mean_ticks = float(np.mean(K)) / tick_steps
    rate = 0.7 * budget / (float(np.mean(sizes)) * mean_ticks)      # arrivals per tick
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=nrequests))
    tickets, submitted, tick, occ = [], 0, 0, []
    t0 = time.perf_counter()
    last = t0
    while True:
        now = time.perf_counter()
        if now - last > 2.0:
            log("serving: tick %d, %d submitted, %.0f%% of the budget in use" % (tick, submitted, 100 * occ[-1]))
            last = now
        if now - t0 > max_submit_s:
            nrequests = submitted        # stop the arrival stream; drain what is in
        while submitted < nrequests and arrivals[submitted] <= tick:
            r, nb = temps[pick[submitted]]
            tickets.append((ctx.serve_submit(r, 1 / 60, int(K[submitted])), nb, int(K[submitted])))
            submitted += 1
        running, queued = ctx.serve_tick(tick_steps)
        occ.append(ctx.bytes_in_use() / budget)
        tick += 1
        if submitted == nrequests and running == 0 and queued == 0:
            break
    wall = time.perf_counter() - t0
    lat, waits, bsteps = [], [], 0
    for t, nb, k in tickets:
        st = ctx.serve_status(t)
        if st["state"] == 2:
            lat.append(st["latency_ms"])
            waits.append(st["wait_ms"])
            bsteps += nb * k
    ctx.close()
    lat = np.array(lat)
    # the steady state: from one mean residence time after the first arrival to the last arrival
    # (the ramp-up and the drain of the stream are left out)
    last_arrival = int(arrivals[min(submitted, len(arrivals)) - 1]) if submitted else 0
    steady = occ[int(mean_ticks): last_arrival + 1] or occ
    return {"requests": submitted, "completed": int(len(lat)), "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "mean_wait_ms": float(np.mean(waits)),
            "body_steps_per_s": bsteps / wall, "wall_s": wall, "mean_occupancy": float(np.mean(occ)),
            "steady_occupancy": float(np.mean(steady)), "steady_ticks": [int(mean_ticks), last_arrival],
            "hbm_budget_gib": budget_gib, "steps_per_tick": tick_steps,
            "config": "Poisson arrivals at %.2f requests per tick, %d request templates (log-uniform [16,2048] "
                      "bodies: spheres, boxes, 8-16 vertex hulls over a 64x64 heightfield), K ~ U{60..600} "
                      "steps, dt 1/60, FP32" % (rate, ntemplates)}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(recs, dyn, dt, warm, stride, target_s=6.0, runs=3):
    """The oracle (reference semantics, FP64 arithmetic) on every stride-th world, one
    std::thread per world slice on every host core: one warm-up run, then the median of `runs`
    timed runs, each of about target_s seconds (the steps per run are sized from the warm-up)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    L = oracle_lib.lib("f64")
    sample = list(range(0, len(recs), stride))
    ws = [L.orc_create(recs[i].ptr) for i in sample]
    arr = (C.c_void_p * len(ws))(*ws)
    threads = max(1, os.cpu_count() or 1)
    t0 = time.perf_counter()
    L.orc_step_many(arr, len(ws), dt, warm, threads)
    per_step = (time.perf_counter() - t0) / max(1, warm)
    steps = int(max(3, min(600, target_s / max(per_step, 1e-6))))
    rates = []
    nb = sum(dyn[i] for i in sample)
    for _ in range(runs):
        t0 = time.perf_counter()
        L.orc_step_many(arr, len(ws), dt, steps, threads)
        rates.append(nb * steps / (time.perf_counter() - t0))
    for w in ws:
        L.orc_destroy(w)
    rates.sort()
    return {"value": rates[len(rates) // 2], "unit": "body-steps/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "runs": rates,
            "sample": "%d of the %d worlds (every %d-th), %d dynamic bodies, %d warm-up steps, then the "
                      "median of %d runs of %d steps; C++ restatement of the reference semantics (FP64 "
                      "arithmetic over FP32 storage), one std::thread per world slice on %d host threads"
                      % (len(sample), len(recs), stride, nb, warm, runs, steps, threads)}


def precision_report(ctx, recs, wids, dyn, dt, preroll, steps_total, stride):
    """FP32 device path against the reference's own arithmetic (FP64 expressions over FP32
    storage: the device's F64 mode, which the parity tests hold bit-identical to the oracle's
    FP64 mode) on every stride-th world of the batch:
      one step from identical state -- max |dx|, max |dv| over the worlds where both make the
      same narrowphase decisions (the same contact count); the worlds whose contact count
      differs (a decision that flipped) and their max |dx|, reported apart;
      drift -- both paths from the initial state over the config's step count, max |dx|."""
    import numpy as np
    from cannon_physics_amd import _abi as A
    from cannon_physics_amd.world import Context, state_to_numpy
    sample = list(range(0, len(recs), stride))
    c64 = Context(device=0, precision=A.PRECISION_F64)
    w64 = [c64.create_world(recs[i]) for i in sample]
    w32 = [wids[i] for i in sample]
    # identical state: the FP32 worlds as the timed region left them
    for a, b in zip(w32, w64):
        arr, n = ctx.download(a)
        c64.upload(b, arr, n)
    ctx.step(w32, dt, 1)
    c64.step(w64, dt, 1)
    dx = dv = dx_flip = 0.0
    flips = 0
    per_world = []
    for a, b in zip(w32, w64):
        g, o = state_to_numpy(*ctx.download(a)), state_to_numpy(*c64.download(b))
        ex = float(np.max(np.abs(g["position"] - o["position"])))
        ev = float(np.max(np.abs(g["velocity"] - o["velocity"])))
        if ctx.stats(a)["contacts"] != c64.stats(b)["contacts"]:
            flips += 1
            dx_flip = max(dx_flip, ex)
        else:
            dx, dv = max(dx, ex), max(dv, ev)
            per_world.append(ex)
    for b in w64:
        c64.destroy_world(b)
    # drift over the whole run from the initial state, both precisions on the device
    from cannon_physics_amd.world import Context as Ctx
    c32 = Ctx(device=0, precision=A.PRECISION_F32)
    d32 = [c32.create_world(recs[i]) for i in sample]
    d64 = [c64.create_world(recs[i]) for i in sample]
    c32.step(d32, dt, steps_total)
    c64.step(d64, dt, steps_total)
    drift = 0.0
    for a, b in zip(d32, d64):
        g, o = state_to_numpy(*c32.download(a)), state_to_numpy(*c64.download(b))
        drift = max(drift, float(np.max(np.abs(g["position"] - o["position"]))))
    c32.close()
    c64.close()
    return {"worlds": len(sample), "after_steps": preroll,
            "one_step_max_abs_dx_m": dx, "one_step_max_abs_dv_m_per_s": dv,
            "one_step_dx_median_m": float(np.median(per_world)) if per_world else 0.0,
            "one_step_dx_p90_m": float(np.percentile(per_world, 90)) if per_world else 0.0,
            "worlds_within_tolerance": int(sum(1 for e in per_world if e <= 1e-4)),
            "worlds_with_contact_count_flip": flips, "flip_worlds_max_abs_dx_m": dx_flip,
            "tolerance_m": 1e-4, "within_tolerance": dx <= 1e-4,
            "drift_steps": steps_total, "drift_max_abs_dx_m": drift,
            "reference_model": "FP64 arithmetic over FP32-stored vectors (the device's F64 mode)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--preroll", type=int, default=None,
                    help="untimed steps before the warm-up, so that the timed window sits mid-run "
                         "(default: about half of the workload's step count)")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--workload", default="c3", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--worlds", type=int, default=None)
    ap.add_argument("--precision", default="f32", choices=["f32", "f64"])
    ap.add_argument("--solve-block", type=int, default=0,
                    help="GS workgroup size (64, 128, 256); 0: the autotuner picks it during warm-up")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serve-templates", type=int, default=12)
    ap.add_argument("--serve-requests", type=int, default=600)
    ap.add_argument("--serve-budget-gib", type=float, default=8.0)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="after the run, copy a buffer of this size once on the device (the "
                         "calibration kernel for HBM byte counters)")
    args = ap.parse_args()
    world_size = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world_size != 1:
        print("bench.py measures one GPU: worlds shard across GPUs without communication, and "
              "this build runs and reports the single-GPU figure only (--gpus 1)", file=sys.stderr)
        sys.exit(2)

    from cannon_physics_amd import _abi as A
    from cannon_physics_amd.world import Context

    nworlds = args.worlds or {"c3": 4096, "c4": 1024, "c5": 48}.get(args.workload, 1)
    t0 = time.perf_counter()
    recs, dyn, desc, dt = build_worlds(args.workload, nworlds)
    log("built %d worlds (%d dynamic bodies) in %.1fs" % (len(recs), sum(dyn), time.perf_counter() - t0))

    prec = A.PRECISION_F64 if args.precision == "f64" else A.PRECISION_F32
    ctx = Context(device=0, precision=prec)
    if args.solve_block:
        ctx.set_tuning(solve_block=args.solve_block)
    wids = [ctx.create_world(r) for r in recs]
    log("uploaded: %.2f GB on the device" % (ctx.bytes_in_use() / 1e9))

    steps_total = {"c3": 600, "c4": 480, "c5": 300, "c1": 600, "c2": 300}[args.workload]
    preroll = args.preroll if args.preroll is not None else max(0, steps_total // 2 - args.warmup - args.steps // 2)
    if preroll:
        t1 = time.perf_counter()
        ctx.step(wids, dt, preroll)
        log("pre-roll: %d steps in %.1fs" % (preroll, time.perf_counter() - t1))
    if args.warmup:
        # the autotuner times the level / GS workgroup sizes on the last warm-up step, once
        # contacts have formed (the choice does not change results)
        if args.warmup > 1:
            ctx.step(wids, dt, args.warmup - 1)
        if not args.solve_block:
            ctx.autotune()
        ctx.step(wids, dt, 1)
    log("launch shape: " + json.dumps(ctx.tuning()))
    before = stats_totals(ctx, wids)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(wids, dt, args.steps)          # synchronous: returns after the device finished
    wall = time.perf_counter() - t0
    dev_ms = ctx.last_step_ms()
    after = stats_totals(ctx, wids)
    # the stage breakdown: the same number of steps again with per-stage events (a profiled call
    # launches the whole batch on one stream; unprofiled batches of 16+ worlds run as two halves
    # on two streams)
    ctx.set_profiling(True)
    t1 = time.perf_counter()
    ctx.step(wids, dt, args.steps)
    profiled_wall = time.perf_counter() - t1
    stages = ctx.stage_times()
    ctx.set_profiling(False)
    K = args.steps
    n_dyn = sum(dyn)
    value = n_dyn * K / wall
    P = (after["sum_pairs"] - before["sum_pairs"]) / K
    E = (after["sum_rows"] - before["sum_rows"]) / K
    Cn = (after["sum_contacts"] - before["sum_contacts"]) / K
    log("timed %d steps: wall %.3fs device %.1f ms; per step %.0f pairs %.0f contacts %.0f rows"
        % (K, wall, dev_ms, P, Cn, E))
    log("stage ms over the profiled steps (one stream, %.3f ms per step): " % (profiled_wall / K * 1e3)
        + json.dumps({k: round(v, 2) for k, v in stages.items()}))

    # the same number of steps again without per-stage events, from a captured graph of the
    # step (captured, with the tuned launch shape, by one untimed call first; graphs are opt-in)
    ctx.set_graphs(True)
    ctx.step(wids, dt, 2)
    t0 = time.perf_counter()
    ctx.step(wids, dt, K)
    graph_wall = time.perf_counter() - t0
    ctx.set_graphs(False)
    log("graph replay: %d steps in %.3fs" % (K, graph_wall))
    dom = max(stages, key=stages.get)
    gs_ms = stages["k_gs"] / K
    achieved_gs = B_ROW_GS * E / (gs_ms / 1e3) / 1e9
    b_step = B_BODY * n_dyn + B_PAIR * P + B_ROW * E
    step_gbs = b_step / (dev_ms / K / 1e3) / 1e9

    plain = None
    if not args.no_plain and len(wids) > 1:
        stride = max(1, len(wids) // 32)
        sample = list(range(0, len(wids), stride))
        tot_t, tot_bs = 0.0, 0
        for i in sample:
            ctx.step_plain(wids[i], dt, 1)
            t1 = time.perf_counter()
            ctx.step_plain(wids[i], dt, 5)
            tot_t += time.perf_counter() - t1
            tot_bs += dyn[i] * 5
        plain = {"value": tot_bs / tot_t, "unit": "body-steps/s",
                 "sample": "%d worlds (every %d-th), 5 steps each, one launch per stage per step "
                           "and a host synchronisation per step" % (len(sample), stride)}

    launch_shape = ctx.tuning()
    precision = None
    if not args.no_precision and args.precision == "f32":
        t1 = time.perf_counter()
        precision = precision_report(ctx, recs, wids, dyn, dt, preroll + args.warmup + 3 * K + 2, steps_total,
                                     max(1, len(recs) // 64))
        log("fp32 vs fp64 (%.1fs): %s" % (time.perf_counter() - t1, json.dumps(precision)))
    serving = None
    if not args.no_serving:
        ctx.close()
        serving = serve_bench(args.serve_templates, args.serve_requests, args.serve_budget_gib, 8)
        log("serving: " + json.dumps(serving))

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(recs, dyn, dt, 5, 1)

    out = {
        "metric": "body-steps/s",
        "value": value,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": wall / K * 1e3,
        "profiled_one_stream_ms_per_step": profiled_wall / K * 1e3,
        "graph_replay_ms_per_step": graph_wall / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic",
        "config": {"workload": desc, "worlds": len(recs), "dynamic_bodies": n_dyn, "dt": dt,
                   "run_steps": steps_total,
                   "timed_steps": "%d..%d of the %d-step run (after %d pre-roll and %d warm-up steps)"
                                  % (preroll + args.warmup + 1, preroll + args.warmup + K, steps_total,
                                     preroll, args.warmup),
                   "launch_shape": launch_shape,
                   "pairs_per_step": P, "contacts_per_step": Cn, "rows_per_step": E},
        "roofline": {"bound": "hbm", "kernel": "k_gs", "achieved": achieved_gs, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved_gs / HBM_PEAK_GBS,
                     "traffic": measured_traffic(args.workload, args.precision, "k_gs"),
                     "bytes_per_launch": B_ROW_GS * E, "avg_launch_ms": gs_ms},
        "step_roofline": {"achieved": step_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": step_gbs / HBM_PEAK_GBS, "bytes_per_step": b_step},
        # the stage of most time; the roofline above is k_gs's, the longest single kernel (the
        # narrowphase stage is a dozen kernels, one per pair class)
        "dominant_stage": dom,
        "stage_ms_per_step": {k: v / K for k, v in stages.items()},   # the profiled steps
        "plain_gpu": plain,
        "serving": serving,
        "batched_vs_plain": (value / plain["value"]) if plain else None,
        "cpu_baseline": cpu,
        "fp32_vs_fp64": precision,
    }
    print(json.dumps(out), flush=True)
    if args.copy_probe_mib:
        copy_probe(args.copy_probe_mib << 20)


if __name__ == "__main__":
    main()
