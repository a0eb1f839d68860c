"""Throughput benchmark of the MI355X constant-Q PSTD engine.

Workload (default, "batch128"): the batched multi-shot forward-modeling job --
128 shots on a 600 x 2400 model with a 40-point CPML, dx = dz = 5 m,
dt = 0.4 ms, 20 Hz Ricker, shots every 18 cells, windows cropped to +-600
cells around each source (uneven sizes), half acoustic and half
viscoacoustic.  One "step" advances every shot of the batch by one time step
(2 fused passes per acoustic shot, 6 per viscoacoustic shot).  All inputs are
resident in HBM before the timed region.

Reported: batched Mcell-steps/s (the metric; the timed steps run the batch's
four launch groups on four streams), the plain one-shot-per-launch path on
one shot of the same workload, the dominant pass's HBM roofline fraction
(HIP-event timing on the engine's stream over extra steps that launch each
pass once over every shot) with its measured traffic,
the FP32 reference-faithful CPU path (the oracle's FP32 build: 9 full 2-D C2C
FFTs per step, OpenMP) timed on a bounded sample of the same workload, the
serving phase (continuous batching of Poisson arrivals: throughput, p50/p99
latency; shortest remaining work first within a run set of 20 M work cells,
and beside it every resident job advancing at every step) and the Q-RTM
phase (32 revolve-checkpointed shots).  Cell-steps
count the windows the shots ask for (CPML included); the engine computes on
windows widened to FFT-friendly sizes (about 1% more cells in the batched
workload), reported beside as computed_mcell_steps_per_s.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload batch128|config1]
                       [--no-serve] [--no-qrtm] [--no-plain] [--no-cpu] [--no-accuracy]
                       [--serve-run-cells C] [--no-serve-all] [--serve-nx-max NX] [--serve-nz-max NZ]
                       [--graph-steps G] [--copy-calibration GB]
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "constant-q-pstd_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(jobs):
    """Minimal-traffic bytes per launch of each pass, summed over the shots it
    serves.  A cell is a padded-grid point (nxx*nzz); a half-spectrum word is
    one complex kz sample per column (nxx*(nzz/2+1) complex = 8 B each).
    Acoustic row pass: reads div, Lz, p, c1, c2, vx, dt*bx; writes p, vx,
      dvx (40 B/cell) + x-CPML strips (phi, psi read+write, 16 B per strip cell).
    Acoustic column pass: reads p, vz, dt*bz, dvx; writes vz, div, Lz
      (28 B/cell) + z-CPML strips (16 B per strip cell).
    Viscoacoustic (one half spectrum = nxx*(nzz/2+1) complex words):
      C1 reads div (4 B/cell), writes 1 half spectrum;
      R1 reads 2, writes 4 half spectra;
      C2 reads 3 half spectra, div, gam, y2, c, writes total (20 B/cell) and
        1 half spectrum;
      R2 reads 1, writes 1 half spectrum;
      C3 reads 2 half spectra, p, total, 2 coefficients (rho v^2 dt lam and
        dt tao / 2; dt tao is twice the latter), vz, dt*bz, writes p, vz, dvz
        (36 B/cell), 1 half spectrum + z-CPML strips;
      R3 reads p, vx, dt*bx, dvz, writes vx, div (24 B/cell) + x-CPML strips."""
    out = {}

    def add(k, v):
        out[k] = out.get(k, 0) + v

    for j in jobs:
        nxx, nzz, pml = j["nxx"], j["nzz"], j["pml"]
        n = nxx * nzz
        hs = nxx * (nzz // 2 + 1) * 8
        sx, sz = 16 * 2 * pml * nzz, 16 * 2 * pml * nxx
        if j["kind"] == 0:
            add("k_row_ac", 40 * n + sx)
            add("k_col_ac", 28 * n + sz)
        else:
            add("k_q_c1", 4 * n + hs)
            add("k_q_r1", 6 * hs)
            add("k_q_c2", 20 * n + 4 * hs)
            add("k_q_r2", 2 * hs)
            add("k_q_c3", 36 * n + 3 * hs + sz)
            add("k_q_r3", 24 * n + sx)
    return out


def shortened(job, steps):
    j = dict(job)
    j["nt"] = steps
    j["rik"] = job["rik"][:steps]
    return j


def with_k2(job):
    from cqpy import host
    j = dict(job)
    if "k2d1" not in j:
        j.update(host.k2d(j["dx"], j["dz"], j["nxx"], j["nzz"], q=int(j["kind"]) in (1, 3)))
    return j


def cells_of(job):
    """Cells a shot's step advances, counted on the window the shot asked for
    (CPML included; synth.useful_cells), not on the FFT-friendly widening it
    is computed on: every rate below is in these useful cell-steps."""
    from cqpy import synth
    return synth.useful_cells(job)


def cpu_rate(job, steps, threads):
    """FP32 reference-faithful CPU path (the oracle's FP32 build: 9 / 14 full
    2-D C2C FFTs per acoustic / viscoacoustic step, OpenMP) on `steps` steps
    of one shot; returns (Mcell-steps/s, seconds)."""
    import orc
    j = with_k2(shortened(job, steps))
    orc.set_threads(threads, 32)
    t0 = time.perf_counter()
    orc.forward(j, bits=32)
    dt = time.perf_counter() - t0
    return cells_of(j) * steps / dt / 1e6, dt


def host_threads():
    """(threads used, nproc): OMP_NUM_THREADS where set (the GPU box sets
    it to its CPU share), else the cores this process may run on."""
    nproc = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS")
    return (int(env) if env and env.isdigit() else nproc), nproc


def cpu_bench(pair, steps_all=(48, 32), steps_one=(5, 4)):
    """The CPU baseline on one acoustic and one viscoacoustic shot of the
    workload, on all the process's cores and on one core.  The mix rate
    weights the two kinds by equal cell-steps (the batch is half of each):
    cells x steps summed over both, over the summed time."""
    threads, nproc = host_threads()
    res = {}
    for label, n, steps in (("all", threads, steps_all), ("one", 1, steps_one)):
        cells = secs = 0.0
        for job, st in zip(pair, steps):
            v, sec = cpu_rate(job, st, n)
            res["%s_%s" % (("acoustic", "visco")[int(job["kind"])], label)] = round(v, 3)
            cells += cells_of(job) * st
            secs += sec
        res["mix_" + label] = round(cells / secs / 1e6, 3)
        res["seconds_" + label] = round(secs, 1)
    return res, threads, nproc


def plain_bench(ctx, pair, steps):
    """The plain path (cq_plain_forward: one shot, one launch sequence per
    step, the gather row copied to the host and the stream synchronised
    after every step) on one acoustic and one viscoacoustic shot of the
    workload, each a whole shot of `steps` steps including its allocation
    and upload as the reference's acforward does; the mix rate weights them
    like the batch (equal cell-steps)."""
    res, cells, secs = {}, 0.0, 0.0
    for job in pair:
        pj = shortened(job, steps)
        t0 = time.perf_counter()
        ctx.plain_forward(pj)
        el = time.perf_counter() - t0
        c = cells_of(pj) * steps
        res[("acoustic", "visco")[int(job["kind"])]] = round(c / el / 1e6, 1)
        cells += c
        secs += el
    res["mix"] = round(cells / secs / 1e6, 1)
    return res


def accuracy_check(ctx, pair, steps):
    """rel-L2 of the batched engine's gathers of one shot of each kind of the
    workload against the FP64 oracle over the first `steps` steps."""
    import orc
    from cqpy import engine
    jobs = [shortened(j, steps) for j in pair]
    b = ctx.batch(jobs)
    b.run()
    got = [b.rec(i) for i in range(len(jobs))]
    b.close()
    threads, _ = host_threads()
    orc.set_threads(threads, 64)
    out = {}
    for j, g in zip(jobs, got):
        ref = orc.forward(with_k2(j))
        out[("acoustic", "visco")[int(j["kind"])]] = float("%.3g" % engine.rel_l2(g, ref))
    out["steps"] = steps
    out["window"] = "%dx%d" % (jobs[0]["nxx"], jobs[0]["nzz"])
    return out


PROFILE = os.path.join(ROOT, "profiles", "batch128_mixed_passes.md")


def profiled_traffic(kernel):
    """HBM bytes per dispatch of `kernel` (FETCH_SIZE + WRITE_SIZE, each
    calibrated with a streaming copy) from the committed rocprofv3 summary of
    the default workload, or None."""
    try:
        text = open(PROFILE).read().split("## HBM bytes per dispatch", 1)
    except OSError:
        return None
    if len(text) < 2:
        return None
    for line in text[1].splitlines():
        # | kernel<tile> | FETCH KiB | WRITE KiB | HBM bytes | algorithmic bytes | ratio |
        m = re.match(r"\| (k_\w+)<\d+> \| [\d.]+ \| [\d.]+ \| ([0-9.e+]+) \|", line)
        if m and m.group(1) == kernel:
            return float(m.group(2))
    return None


def serve_bench(rate, window_s, seed, chunk, run_cells=0, nx_max=4096, nz_max=1024):
    """Serving workload (Poisson arrivals of mixed acoustic / viscoacoustic /
    Q-RTM requests, each with its own window, nt, dx and wavelet) through the
    engine's continuous-batching server.  Requests are generated before the
    clock starts and submitted at their arrival times; the server admits them
    into the running launches.  Latency = completion - arrival, wall clock."""
    from cqpy import engine, host, synth
    reqs = synth.serve_jobs(rate=rate, window_s=window_s, seed=seed, nx_range=(512, nx_max), nz_range=(256, nz_max))
    mx = max(j["nxx"] for _, j, _ in reqs)
    mz = max(j["nzz"] for _, j, _ in reqs)
    useful = executed = 0
    for _, j, kind in reqs:
        c, cp = cells_of(j), j["nxx"] * j["nzz"]
        if kind == 3:
            useful += c * 2 * j["nt"]
            executed += cp * (host.numforw(j["nt"], j["snaps"]) + j["nt"])
        else:
            useful += c * j["nt"]
            executed += cp * j["nt"]
    ctx = engine.Context(0)
    srv = ctx.server(mx, mz, run_cells=run_cells)
    pending = list(reqs)
    info, lat = {}, {}
    t0 = time.perf_counter()
    while pending or len(lat) < len(info):
        now = time.perf_counter() - t0
        while pending and pending[0][0] <= now:
            arr, job, kind = pending.pop(0)
            info[srv.submit(job)] = (arr, kind)
        running, queued = srv.advance(chunk)
        now = time.perf_counter() - t0
        for i, (arr, kind) in info.items():
            if i not in lat and srv.state(i) == engine.DONE:
                lat[i] = now - arr
        if running == 0 and queued == 0 and pending:
            time.sleep(max(0.0, pending[0][0] - (time.perf_counter() - t0)))
    wall = time.perf_counter() - t0
    srv.close()
    ctx.close()
    first = reqs[0][0]

    def pct(v, p):
        v = sorted(v)
        return round(v[min(len(v) - 1, int(round(p / 100.0 * (len(v) - 1))))], 3)

    names = {0: "acoustic", 1: "viscoacoustic", 3: "qrtm"}
    per = {}
    for k, nm in names.items():
        v = [lat[i] for i, (a, kk) in info.items() if kk == k]
        if v:
            per[nm] = {"jobs": len(v), "p50_s": pct(v, 50), "p99_s": pct(v, 99)}
    allv = list(lat.values())
    return {"workload": "serve: Poisson %.0f jobs/s for %.1f s (%d jobs; 50%% acoustic, 30%% "
                        "viscoacoustic, 20%% Q-RTM with 40 checkpoints), windows %d-%d x %d-%d "
                        "cells + 40-pt CPML (padded to 7-smooth sizes, at most %dx%d), nt 2000-8000, "
                        "dx 5-12.5 m, Ricker 10-30 Hz"
                        % (rate, window_s, len(reqs), 512, nx_max, 256, nz_max, mx, mz),
            "jobs": len(reqs), "wall_s": round(wall, 3),
            "useful_mcell_steps_per_s": round(useful / (wall - first) / 1e6, 1),
            "executed_mcell_steps_per_s": round(executed / (wall - first) / 1e6, 1),
            "p50_s": pct(allv, 50), "p99_s": pct(allv, 99), "per_class": per,
            "steps_per_advance": chunk, "run_cells": run_cells}


def qrtm_bench(nshots, nt, snaps, graph_steps):
    """Q-RTM with revolve checkpointing on the 800 x 3000 workload (2100 x 880
    padded windows, `snaps` checkpoints per shot), all shots in one batch, run
    to completion with a shortened time axis (nt steps; spans of graph_steps
    steps between checkpoint copies replayed as HIP graphs).  Executed
    cell-steps count every forward and adjoint step the schedules take
    (numforw + nt per shot); useful ones are 2 nt per shot."""
    from cqpy import engine, host, synth
    jobs = synth.qrtm_jobs(nshots=nshots, nt=nt, snaps=snaps)
    cells = sum(j["nxx"] * j["nzz"] for j in jobs)
    useful_cells = sum(cells_of(j) for j in jobs)
    nf = host.numforw(nt, snaps)
    ctx = engine.Context(0)
    b = ctx.batch(jobs)
    b.use_graphs(graph_steps)
    gb = b.nbytes() / 1e9
    t0 = time.perf_counter()
    b.run()
    el = time.perf_counter() - t0
    b.close()
    # per-pass device times over the first 300 global steps of a second run
    b = ctx.batch(jobs)
    b.enable_timing(True)
    b.step(300)
    kt = {k: round(v[0], 4) for k, v in b.kernel_times().items()}
    b.close()
    ctx.close()
    executed = cells * (nf + nt) / el / 1e6
    useful = useful_cells * 2 * nt / el / 1e6
    return {"workload": "qrtm: %d Q-RTM shots, 800x3000 model, 2100x880 padded windows, 40-pt CPML, "
                        "dx=6.25 m, dt=0.5 ms, %d checkpoints, nt=%d (time axis shortened from 6000)"
                        % (nshots, snaps, nt),
            "shots": nshots, "device_GB": round(gb, 2), "wall_s": round(el, 3),
            "executed_steps_per_shot": nf + nt, "useful_steps_per_shot": 2 * nt,
            "executed_mcell_steps_per_s": round(executed, 1),
            "useful_mcell_steps_per_s": round(useful, 1),
            "useful_mcell_steps_per_s_at_nt6000": round(
                useful * (nf + nt) / (2 * nt) * 2 * 6000 / (host.numforw(6000, snaps) + 6000), 1),
            "pass_ms_first_300_steps": kt}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="batch128", choices=["batch128", "config1"])
    ap.add_argument("--shots", type=int, default=128)
    ap.add_argument("--kinds", default="mixed", choices=["mixed", "acoustic", "visco"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="the engine's own tiles instead of the autotuner's")
    ap.add_argument("--graph-steps", type=int, default=16,
                    help="global steps per captured HIP graph (0: direct launches)")
    ap.add_argument("--accuracy-steps", type=int, default=24)
    ap.add_argument("--plain-steps", type=int, default=0,
                    help="steps of the plain path per shot (0: the workload's whole time axis, so that its "
                         "per-shot allocation and upload weigh as in a real shot)")
    ap.add_argument("--copy-calibration", type=float, default=0.0,
                    help="only run a streaming copy of this many GB (counter calibration)")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving (latency) phase")
    ap.add_argument("--no-qrtm", action="store_true", help="skip the Q-RTM phase")
    ap.add_argument("--qrtm-shots", type=int, default=32)
    ap.add_argument("--qrtm-nt", type=int, default=600)
    ap.add_argument("--serve-rate", type=float, default=40.0)
    ap.add_argument("--serve-window", type=float, default=1.0)
    ap.add_argument("--serve-chunk", type=int, default=50)
    ap.add_argument("--serve-nx-max", type=int, default=4096,
                    help="widest serving window (interior columns; 1216 with --serve-nz-max 600 gives the "
                         "smaller-window mix of earlier measurements)")
    ap.add_argument("--serve-nz-max", type=int, default=1024, help="deepest serving window (interior rows)")
    ap.add_argument("--serve-run-cells", type=int, default=20000000,
                    help="server run set in work cells: the jobs with the least remaining work "
                         "advance (0: every resident job advances every step)")
    ap.add_argument("--no-serve-all", action="store_true",
                    help="skip the second serving run with every resident job advancing")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (every GPU box here has one "
              "GPU); --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from cqpy import engine, synth

    if args.copy_calibration > 0:
        with engine.Context(0) as c:
            nb = int(args.copy_calibration * 1e9)
            ms = c.copy_calibrate(nb, reps=5)
        print(json.dumps({"copy_bytes_read": nb, "copy_bytes_written": nb, "ms_per_copy": ms,
                          "GBs_read_plus_write": 2 * nb / (ms * 1e-3) / 1e9}))
        return

    nt_total = args.warmup + args.steps
    if args.workload == "batch128":
        kinds = [i % 2 if args.kinds == "mixed" else (0 if args.kinds == "acoustic" else 1)
                 for i in range(args.shots)]
        jobs = synth.batch128_jobs(nt=max(nt_total, 5000), nshots=args.shots, kinds=kinds)
        wl = {"workload": "batch128: %d shots, 600x2400 model, 40-pt CPML, dx=dz=5 m, "
                          "dt=0.4 ms, windows cropped to +-600 cells (uneven; computed on windows "
                          "widened to FFT-friendly sizes, counted on the cropped ones)" % args.shots,
              "shots": args.shots,
              "kinds": {"mixed": "half acoustic, half viscoacoustic (the reference's rank-4 "
                                 "Taylor-expanded fractional-Laplacian operator, sum_total)",
                        "acoustic": "acoustic",
                        "visco": "viscoacoustic (Taylor-expanded operator, sum_total)"}[args.kinds],
              "cells_per_step": int(sum(cells_of(j) for j in jobs)),
              "computed_cells_per_step": int(sum(j["nxx"] * j["nzz"] for j in jobs))}
    else:
        jobs = [synth.config1_job(nt=max(nt_total, 3000), want_k2=False)]
        wl = {"workload": "config1: one acoustic shot, 459x220 padded, dt=0.5 ms",
              "shots": 1, "kinds": "acoustic", "cells_per_step": cells_of(jobs[0]),
              "computed_cells_per_step": jobs[0]["nxx"] * jobs[0]["nzz"]}
    cells = wl["cells_per_step"]
    default_workload = args.workload == "batch128" and args.kinds == "mixed" and args.shots == 128
    ctx = engine.Context(0)
    tiles, tune = None, None
    if not args.no_tune:
        # tile shapes and FFT plan from the autotuner (timed trials, cached on
        # disk by device and window set); outside the timed region
        cache = os.environ.get("CQ_TUNE_CACHE") or os.path.join(os.path.expanduser("~"), ".cache", "cq_pstd",
                                                                 "tiles.txt")
        os.makedirs(os.path.dirname(cache), exist_ok=True)
        t0 = time.perf_counter()
        tiles, tms, hit = ctx.autotune(jobs, cache)
        tune = {"tiles": tiles, "trial_ms_per_step": round(tms, 4), "from_cache": hit,
                "seconds": round(time.perf_counter() - t0, 1)}
        print("autotune:", json.dumps(tune), file=sys.stderr, flush=True)
    b = ctx.batch(jobs, tiles=tiles)
    b.use_graphs(args.graph_steps)
    print("batch built: %d shots, %.2f GB on device, tiles %s" % (len(jobs), b.nbytes() / 1e9, b.tiles()),
          file=sys.stderr, flush=True)
    b.step(args.warmup)
    b.sync()
    t0 = time.perf_counter()
    b.step(args.steps)
    b.sync()
    el = time.perf_counter() - t0
    value = cells * args.steps / el / 1e6
    print("timed %d steps in %.3f s" % (args.steps, el), file=sys.stderr, flush=True)

    # per-pass device time with HIP events on the engine's stream
    b.enable_timing(True)
    b.step(min(10, jobs[0]["nt"] - b.time()))
    kt = b.kernel_times()
    b.enable_timing(False)
    b.close()
    algo = algorithmic_bytes(jobs)
    passes = {}
    for name, (ms, calls) in kt.items():
        gbs = algo[name] / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        passes[name] = {"ms": round(ms, 4), "bytes": algo[name], "GBs": round(gbs, 1),
                        "frac": round(gbs / HBM_PEAK_GBS, 4)}
    dom = max(kt, key=lambda k: kt[k][0])
    achieved = passes[dom]["GBs"]
    step_bytes = sum(algo.values())
    print("passes:", json.dumps(passes), file=sys.stderr, flush=True)

    # one interior shot of each kind of the workload: plain path, accuracy, CPU
    mid = len(jobs) // 2
    pair = []
    for k in (0, 1):
        cand = [j for j in jobs[mid:] + jobs[:mid] if int(j["kind"]) == k]
        if cand:
            pair.append(cand[0])
    plain = None
    if not args.no_plain:
        plain_steps = args.plain_steps or int(pair[0]["nt"])
        plain = plain_bench(ctx, pair, plain_steps)
        plain["steps_per_shot"] = plain_steps
        print("plain:", json.dumps(plain), file=sys.stderr, flush=True)
    accuracy = None
    if not args.no_accuracy:
        accuracy = accuracy_check(ctx, pair, args.accuracy_steps)
        print("accuracy:", json.dumps(accuracy), file=sys.stderr, flush=True)
    ctx.close()

    serve = None
    if not args.no_serve:
        print("serving phase", file=sys.stderr, flush=True)
        serve = serve_bench(args.serve_rate, args.serve_window, 501, args.serve_chunk, args.serve_run_cells,
                            args.serve_nx_max, args.serve_nz_max)
        print("serve:", json.dumps(serve), file=sys.stderr, flush=True)
        if args.serve_run_cells > 0 and not args.no_serve_all:
            # the same arrivals with every resident job advancing at every
            # global step (throughput-first policy), for comparison
            print("serving phase, all resident jobs advancing", file=sys.stderr, flush=True)
            serve["all_resident"] = serve_bench(args.serve_rate, args.serve_window, 501, args.serve_chunk, 0,
                                                args.serve_nx_max, args.serve_nz_max)
            print("serve all resident:", json.dumps(serve["all_resident"]), file=sys.stderr, flush=True)

    qrtm = None
    if not args.no_qrtm:
        print("Q-RTM phase", file=sys.stderr, flush=True)
        qrtm = qrtm_bench(args.qrtm_shots, args.qrtm_nt, 40, args.graph_steps)
        print("qrtm:", json.dumps(qrtm), file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        res, threads, nproc = cpu_bench(pair)
        cpu = {"value": res["mix_all"], "unit": "Mcell-steps/s", "cores": threads, "kind": "port",
               "nproc": nproc, "one_core": res["mix_one"], "per_kind": res,
               "sample": "FP32 reference-faithful CPU path (the oracle's FP32 build: 9 / 14 full 2-D "
                         "C2C FFTs per acoustic / viscoacoustic step, OpenMP) on one %dx%d shot of each "
                         "kind of the workload: 48 + 32 steps on %d threads (%.1f s), 5 + 4 steps on one "
                         "core (%.1f s); value weights the kinds by equal cell-steps like the batch"
                         % (pair[0]["nxx"], pair[0]["nzz"], threads, res["seconds_all"], res["seconds_one"])}

    out = {
        "metric": "Mcell-steps/s batched vs plain at rel-L2<=1e-4 vs FP64; p50/p99 shot latency",
        "value": round(value, 1),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 4),
        "computed_mcell_steps_per_s": round(wl["computed_cells_per_step"] * args.steps / el / 1e6, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded layered model with salt body; no field data)",
        "config": wl,
        "plain_mcell_steps_per_s": None if plain is None else plain["mix"],
        "plain": plain,
        "batched_over_plain": None if plain is None else round(value / plain["mix"], 2),
        "accuracy_rel_l2_vs_fp64": accuracy,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": profiled_traffic(dom) if default_workload else None,
                     "kernel": dom, "passes": passes,
                     "step_bytes": step_bytes,
                     "step_GBs": round(step_bytes * args.steps / el / 1e9, 1)},
        "cpu_baseline": cpu,
        "autotune": tune,
        "graph_steps": args.graph_steps,
        "serve": serve,
        "qrtm": qrtm,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
