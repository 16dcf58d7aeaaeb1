#!/usr/bin/env python3
"""Benchmark of the filament BD step on one MI355X.

Metric: segment-steps/s = bonds x steps / wall-seconds, one step being one
reference i_step (two per Delta t), with the mixed path's force error
against the FP64 path reported beside it.  Default workload: the
uneven-length system, 50,000 filaments with log-normal site counts (median
24, sigma_ln 0.8, clipped to [8, 256]; 1.61M bonds), persistence length
U[10, 1000] sigma, periodic cube at volume fraction 0.08, WCA eps=1 cutoff
2^(1/6) sigma, dt=1e-5 (the explicit midstep scheme is unstable at 1e-4 for
lp = 1000 sigma); FP32 pair forces, FP64 tension solve and positions, the
step replayed from a hipGraph, launch shapes autotuned.  Synthetic initial
state (reference-style random insertion), state resident in HBM before the
timed region.

Printed on one JSON line: the metric; the roofline of the step's dominant
kernel by measured time (the integrator on both large systems), with the
pair kernel's and the integrator's beside it, from HIP-event timing on the
system's stream; the per-bond force error of the mixed path against the
FP64 path from identical state; the plain FP64 GPU path on the same
workload (eager and graph-replayed); and the CPU oracle (OpenMP, the
reference's decomposition) on a bounded sample with the host CPU named.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(name, sc):
    """(params, filaments, description dict, crosslinks or None, crosslink
    parameters or None) of a named configuration (simcore_amd.workloads)."""
    from simcore_amd import workloads as W
    if name == "c1":
        p, fil, desc = W.dense()
        return p, fil, desc, None, None
    if name == "c2":
        p, fil, desc = W.uneven()
        return p, fil, desc, None, None
    if name == "c3":
        return W.crosslinked()
    if name == "c0":
        p, spec = sc.params_from_yaml(os.path.join(ROOT, "simcore_amd", "configs",
                                                   "c0_plain_fp64.yaml"))
        nb, _ = sc.filament_layout(spec["length"], spec["min_bond_length"],
                                   spec["min_length"], spec["max_length"])
        fil = sc.generate_filaments(p, [nb + 1] * int(spec["num"]), spec["length"],
                                    spec["persistence_length"], spec["diameter"],
                                    spec["friction_ratio"], seed=7)
        desc = {"workload": "parity system: 64 filaments x 20 sites, lp 100 sigma, box 40 "
                            "sigma, WCA, dt 1e-4", "n_filaments": 64, "sites_per_filament": 20,
                "bonds": int(fil.nb)}
        return p, fil, desc, None, None
    raise ValueError(name)


def serving_requests(sc, n_requests, rate, seed=20240004):
    """Synthetic serving trace: Poisson arrivals at `rate`/s of independent
    systems with 1k-100k bonds (log-uniform) of 32-site filaments at volume
    fraction 0.05, 2000 steps each, persistence length and noise seed drawn
    per request (lp in [10, 1000] sigma, hence dt 1e-5 as for the
    uneven-length system)."""
    from simcore_amd.workloads import box_for_fraction
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    out = []
    for k in range(n_requests):
        nb = float(np.exp(rng.uniform(np.log(1e3), np.log(1e5))))
        nf = max(1, int(round(nb / 31)))
        lp = float(rng.uniform(10.0, 1000.0))
        R = box_for_fraction([32] * nf, 1.0, 1.0, 0.05)
        p = sc.params_init(R, delta=1e-5, min_cell_length=1.0,
                           seed=int(rng.integers(1, 2 ** 62)))
        fil = sc.generate_filaments(p, [32] * nf, 31.0, lp, 1.0, 2.0,
                                    seed=int(rng.integers(1, 2 ** 31)))
        out.append((float(t[k]), p, fil))
    return out


def serve_trace(args, sc, reqs):
    """Replays one arrival trace through a fresh ensemble runner; returns its
    stats."""
    ens = sc.Ensemble(hbm_budget=int(args.hbm_budget_gb * (1 << 30)),
                      max_active=args.max_active, steps_per_launch=args.steps_per_launch,
                      batch_size=args.batch_size)
    t0 = ens.clock()
    k = 0
    last_log = time.perf_counter()
    host = 0.0  # seconds spent inside the runner's scheduling rounds
    while True:
        now = ens.clock() - t0
        while k < len(reqs) and reqs[k][0] <= now:
            ens.submit(reqs[k][1], reqs[k][2], args.request_steps, precision=args.precision,
                       t_submit=t0 + reqs[k][0],
                       spatial_order=args.precision == "mixed" and not args.caller_order)
            k += 1
        th = time.perf_counter()
        nr, nq = ens.progress()
        host += time.perf_counter() - th
        if k == len(reqs) and nr == 0 and nq == 0:
            break
        if time.perf_counter() - last_log > 5.0:
            last_log = time.perf_counter()
            log(f"[bench] t={now:.1f}s submitted {k} running {nr} queued {nq}")
        time.sleep(50e-6)
    st = ens.stats()
    st["host_s"] = host
    ens.close()
    return st


def run_sweep(args, sc):
    """Saturation throughput: the same request mix replayed at doubling
    arrival rates from --rate, --requests requests per rate; stops at the
    first rate with a failed request, or whose p99 latency exceeds 10x the
    lowest rate's, or whose throughput gained less than 5% over the
    previous rate (the runner no longer keeps up).  Prints one JSON line
    with every rate's throughput and latencies and the largest throughput
    reached."""
    base = serving_requests(sc, args.requests, 1.0)
    nbonds = np.array([f.nb for _, _, f in base])
    warm = sc.Ensemble(hbm_budget=int(args.hbm_budget_gb * (1 << 30)),
                       max_active=args.max_active, steps_per_launch=args.steps_per_launch,
                      batch_size=args.batch_size)
    warm.submit(base[0][1], base[0][2], 20, precision=args.precision)
    warm.drain(120)
    warm.close()
    points, rate, stop = [], args.rate, ""
    while not stop and rate <= 65536:
        # the same systems; arrival times scaled to the rate
        reqs = [(t / rate, p, f) for t, p, f in base]
        st = serve_trace(args, sc, reqs)
        value = st["segment_steps"] / st["span"] if st["span"] > 0 else 0.0
        pt = {"rate_per_s": rate, "value": value, "latency_p50_s": st["latency_p50"],
              "latency_p99_s": st["latency_p99"], "completed": st["completed"],
              "failed": st["failed"], "peak_active": st["peak_active"],
              "batched_steps": st["batched_steps"], "solo_steps": st["solo_steps"],
              "host_busy": st["host_s"] / st["span"] if st["span"] > 0 else None,
              "host_admit_s": st["host_admit_s"], "host_chunk_s": st["host_chunk_s"],
              "host_retire_s": st["host_retire_s"], "host_capture_s": st["host_capture_s"],
              "graph_captures": st["graph_captures"],
              "offered": float(nbonds.mean()) * args.request_steps * rate}
        points.append(pt)
        log(f"[bench] sweep {json.dumps(pt)}")
        if st["failed"]:
            stop = "failed request"
        elif st["latency_p99"] > 10 * points[0]["latency_p99_s"]:
            stop = "p99 latency above 10x the lowest rate's"
        elif len(points) > 1 and value < 1.05 * points[-2]["value"]:
            stop = "throughput gained less than 5%"
        rate *= 2
    best = max(points, key=lambda q: q["value"])
    out = {
        "metric": "segment-steps/s", "value": best["value"], "unit": "segment-steps/s",
        "n_gpus": 1, "steps": args.request_steps, "warmup": 0, "ms_per_step": None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if args.precision == "mixed" else "f64", "data": "synthetic",
        "config": {"workload": "ensemble serving, rate sweep: %d systems of 1k-100k bonds "
                               "(log-uniform, 32-site filaments, volume fraction 0.05), %d "
                               "steps each, lp U[10,1000] sigma, dt 1e-5, Poisson arrivals at "
                               "doubling rates from %g/s" % (args.requests, args.request_steps,
                                                             args.rate),
                   "max_active": args.max_active, "steps_per_launch": args.steps_per_launch,
                   "batch_size": args.batch_size,
                   "hbm_budget_gb": args.hbm_budget_gb, "precision": args.precision},
        "saturation": {"value": best["value"], "rate_per_s": best["rate_per_s"],
                       "stopped_at_rate_per_s": points[-1]["rate_per_s"], "reason": stop},
        "sweep": points,
    }
    print(json.dumps(out), flush=True)


def run_serving(args, sc):
    if args.sweep:
        run_sweep(args, sc)
        return
    log(f"[bench] generating {args.requests} requests")
    reqs = serving_requests(sc, args.requests, args.rate)
    nbonds = np.array([f.nb for _, _, f in reqs])
    log(f"[bench] bonds per request: mean {nbonds.mean():.0f}, min {nbonds.min()}, "
        f"max {nbonds.max()}")
    ens = sc.Ensemble(hbm_budget=int(args.hbm_budget_gb * (1 << 30)),
                      max_active=args.max_active, steps_per_launch=args.steps_per_launch,
                      batch_size=args.batch_size)
    # warm the code paths (graph capture, first launches) outside the trace
    w = ens.submit(reqs[0][1], reqs[0][2], 20, precision=args.precision)
    ens.drain(120)
    ens.close()
    ens = sc.Ensemble(hbm_budget=int(args.hbm_budget_gb * (1 << 30)),
                      max_active=args.max_active, steps_per_launch=args.steps_per_launch,
                      batch_size=args.batch_size)
    t0 = ens.clock()
    k = 0
    last_log = time.perf_counter()
    while True:
        now = ens.clock() - t0
        while k < len(reqs) and reqs[k][0] <= now:
            ens.submit(reqs[k][1], reqs[k][2], args.request_steps, precision=args.precision,
                       t_submit=t0 + reqs[k][0],
                       spatial_order=args.precision == "mixed" and not args.caller_order)
            k += 1
        nr, nq = ens.progress()
        if k == len(reqs) and nr == 0 and nq == 0:
            break
        if time.perf_counter() - last_log > 5.0:
            last_log = time.perf_counter()
            log(f"[bench] t={now:.1f}s submitted {k} running {nr} queued {nq}")
        time.sleep(50e-6)
    st = ens.stats()
    ens.close()
    del w
    value = st["segment_steps"] / st["span"] if st["span"] > 0 else 0.0
    offered = float(nbonds.mean()) * args.request_steps * args.rate
    out = {
        "metric": "segment-steps/s", "value": value, "unit": "segment-steps/s",
        "n_gpus": 1, "steps": args.request_steps, "warmup": 0,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64" if args.precision == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": "ensemble serving: Poisson arrivals at %g/s of systems with "
                               "1k-100k bonds (log-uniform, 32-site filaments, volume fraction "
                               "0.05), %d steps each, lp U[10,1000] sigma, dt 1e-5, per-request "
                               "seeds" % (args.rate, args.request_steps),
                   "requests": args.requests, "max_active": args.max_active,
                   "batch_size": args.batch_size,
                   "steps_per_launch": args.steps_per_launch,
                   "hbm_budget_gb": args.hbm_budget_gb, "precision": args.precision},
        "latency_p50_s": st["latency_p50"], "latency_p99_s": st["latency_p99"],
        "latency_mean_s": st["latency_mean"], "latency_max_s": st["latency_max"],
        "completed": st["completed"], "failed": st["failed"], "span_s": st["span"],
        "offered_segment_steps_per_s": offered, "peak_active": st["peak_active"],
        "peak_bytes": st["peak_bytes"],
    }
    print(json.dumps(out), flush=True)


def copy_probe(nbytes, repeats=20):
    import ctypes
    import simcore_amd as sc
    ms = ctypes.c_double()
    rc = sc.lib.sc_copy_probe(int(nbytes), int(repeats), ctypes.byref(ms))
    if rc != 0:
        raise RuntimeError(f"sc_copy_probe failed: {rc}")
    out = {"bytes": int(nbytes), "avg_ms": ms.value,
           "GBps_read_plus_write": 2 * nbytes / (ms.value * 1e-3) / 1e9}
    print(json.dumps(out), flush=True)


def pmc_summary(root, tag, copy_bytes):
    """Per-kernel FETCH_SIZE / WRITE_SIZE averages from rocprofv3 --pmc
    runs under root/{FETCH_SIZE,WRITE_SIZE} and of the copy probe under
    root/copy_{FETCH_SIZE,WRITE_SIZE}; each counter is scaled by
    (copy bytes / what it reads on the copy) and written to
    profiles/traffic_<tag>.json."""
    import glob
    import sqlite3

    def per_kernel(d, counter):
        dbs = glob.glob(os.path.join(root, d, "**", "*.db"), recursive=True)
        acc = {}
        for db in dbs:
            con = sqlite3.connect(db)
            q = ("select kernel_name, dispatch_id, sum(value) from counters_collection "
                 "where counter_name = ? group by kernel_name, dispatch_id")
            for name, _, v in con.execute(q, (counter,)):
                acc.setdefault(name.split("(")[0].replace("void ", ""), []).append(v)
        return {k: (float(np.mean(v)), len(v)) for k, v in acc.items()}

    out = {"tag": tag, "copy_bytes": copy_bytes, "unit": "bytes per launch", "kernels": {}}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        cp = per_kernel("copy_" + counter, counter)
        copy_kib = max(v[0] for k, v in cp.items() if "copy_probe" in k)
        scale = copy_bytes / (copy_kib * 1024.0)
        out[counter + "_scale"] = scale
        for k, (v, n) in per_kernel(counter, counter).items():
            e = out["kernels"].setdefault(k, {})
            e[counter + "_KiB_raw"] = v
            e[counter.lower() + "_bytes"] = v * 1024.0 * scale
            e["launches"] = n
    for k, e in out["kernels"].items():
        if "fetch_size_bytes" in e and "write_size_bytes" in e:
            e["traffic_bytes"] = e["fetch_size_bytes"] + e["write_size_bytes"]
    path = os.path.join(ROOT, "profiles", f"traffic_{tag}.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
    print(json.dumps({"written": path, "fetch_scale": out["FETCH_SIZE_scale"],
                      "write_scale": out["WRITE_SIZE_scale"]}), flush=True)


def measured_traffic(tag, kernel):
    path = os.path.join(ROOT, "profiles", f"traffic_{tag}.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        d = json.load(fh)
    for k, e in d.get("kernels", {}).items():
        if k.endswith(kernel) and "traffic_bytes" in e:
            return e["traffic_bytes"]
    return None


def measured_traffic_sum(tag, part, suffix):
    """Measured bytes of all kernels whose name contains `part` and ends with
    `suffix` (e.g. every length bin's integrator launch of one step kind),
    summed; None without a measurement."""
    path = os.path.join(ROOT, "profiles", f"traffic_{tag}.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        d = json.load(fh)
    vals = [e["traffic_bytes"] for k, e in d.get("kernels", {}).items()
            if part in k and k.endswith(suffix) and "traffic_bytes" in e]
    return float(sum(vals)) if vals else None


# workloads with a committed PMC traffic measurement
TRAFFIC_TAG = {"c1": "dense620k", "c2": "uneven1600k"}


def pair_bytes(precision, nb, nc):
    """Algorithmic HBM bytes of one launch of the pair/scatter kernel (either
    launch shape): per bond its state record read (32 bytes on the mixed
    path, 64 on the FP64 path), its force and torque sums written (6 FP32 on
    the mixed path, 6 FP64 on the FP64 path) and its incidence offset read;
    per candidate the sticky force slot and its two incidence entries
    {candidate, partner} read.  Re-reads of a partner's record and the
    conditional slot write of in-range pairs are not counted."""
    rec = 32 if precision == "mixed" else 64
    slot = 16 if precision == "mixed" else 32
    sums = 24 if precision == "mixed" else 48
    return nb * (rec + sums + 4) + nc * (slot + 2 * 8)


def integrator_bytes(ns, nb):
    """Algorithmic HBM bytes of the integrator per step (all its length-bin
    launches), the per-unit figure of the step's layout: per site 148 bytes
    (position read and written 48, previous position and random force 24
    each -- written on a midstep, read on a full step --, the 12-byte noise
    draw, displacement origin read 24 and accumulator read and written 16),
    per bond 52 (force and torque sums read 24, the new bond state written
    28)."""
    return ns * 148 + nb * 52


def pair_bytes_fp32_state(nb, nc):
    """The same launch priced for a 28-byte FP32 bond state (s, u, l/2) and
    12-byte FP32 slots: per bond the state read and the FP32 sums written,
    per candidate its index pair read and its slot read and written.  This
    build keeps FP64 records (the mixed path's geometry is FP64), so the
    figure is reported beside the roofline, not as it."""
    return nb * (28 + 24) + nc * (8 + 12 + 12)


def timed_steps(sysobj, steps):
    t0 = time.perf_counter()
    sysobj.step(steps)  # returns after the stream has drained
    return time.perf_counter() - t0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="c2", choices=["c0", "c1", "c2", "c3", "c4"])
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--profile-steps", type=int, default=100)
    ap.add_argument("--fp64-steps", type=int, default=200)
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tune-steps", type=int, default=300)
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the mixed-vs-FP64 force error measurement")
    ap.add_argument("--launch-shape", default="",
                    help="pair_block,pair_xcd,integrate_waves,integrate_layout"
                         "[,integrate_streams] instead of autotuning "
                         "(pair_block 0: load-balanced pair kernel)")
    ap.add_argument("--caller-order", action="store_true",
                    help="keep filaments in generation order (no spatial storage order)")
    # serving workload (--workload c4)
    ap.add_argument("--requests", type=int, default=200)
    ap.add_argument("--rate", type=float, default=20.0)
    ap.add_argument("--sweep", action="store_true",
                    help="c4: arrival rates doubling from --rate until saturation")
    ap.add_argument("--request-steps", type=int, default=2000)
    ap.add_argument("--max-active", type=int, default=64)
    ap.add_argument("--batch-size", type=int, default=16,
                    help="c4: systems per batch of grouped launches (0: the runner's "
                         "default, 1: no batches)")
    ap.add_argument("--steps-per-launch", type=int, default=64)
    ap.add_argument("--hbm-budget-gb", type=float, default=64.0)
    # measurement helpers
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run the streaming-copy probe of this many bytes")
    ap.add_argument("--pmc-summary", default="",
                    help="only summarise rocprofv3 --pmc runs under this directory")
    ap.add_argument("--pmc-tag", default="dense620k_mixed")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on exactly one GPU: there is no multi-GPU path here "
              "(independent systems would shard by ensemble member, one large system "
              "by bricks of cells with a halo exchange)", file=sys.stderr)
        sys.exit(2)
    if args.pmc_summary:
        pmc_summary(args.pmc_summary, args.pmc_tag, args.copy_probe or (256 << 20))
        return
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return

    import simcore_amd as sc
    if args.workload == "c4":
        run_serving(args, sc)
        return

    log(f"[bench] building workload {args.workload}")
    p, fil, desc, xl, xl_params = workload(args.workload, sc)
    log(f"[bench] {fil.nf} filaments, {fil.nb} bonds")
    spatial = args.precision == "mixed" and not args.caller_order
    dev = sc.FilamentSystem(p, fil, precision=args.precision, use_graph=True,
                            spatial_order=spatial)
    if xl is not None:
        dev.load_crosslinks(xl[0], xl[1], **xl_params)
    dev.step(args.warmup)
    log(f"[bench] warmup {args.warmup} steps done, counters {dev.counters()}")
    tune = None
    if args.launch_shape:
        keys = ["pair_block", "pair_xcd", "integrate_waves", "integrate_layout",
                "integrate_streams", "pair_parts", "early_long"]
        shape = dict(zip(keys, (int(v) for v in args.launch_shape.split(","))))
        dev.set_launch_shape(**shape)
        tune = dict(dev.launch_shape(), set="fixed")
    elif not args.no_autotune:
        # launch shapes that leave results unchanged, chosen from a
        # checkpoint of the warmed-up state (the state is restored)
        tune = dev.autotune(steps=args.tune_steps)
        log(f"[bench] autotune: {tune}")
    c0 = dev.counters()
    wall = timed_steps(dev, args.steps)
    c1 = dev.counters()
    value = fil.nb * args.steps / wall
    log(f"[bench] timed {args.steps} steps in {wall:.4f} s -> {value:.4e} segment-steps/s, "
        f"rebuilds {c1['n_rebuilds'] - c0['n_rebuilds']}")

    # per-kernel HIP-event timing: the same steps launched eagerly with
    # events right before and after each kernel family on the system's stream
    dev.profile(True)
    dev.step(args.profile_steps)
    prof = dev.profile_read()
    dev.profile(False)
    nc = dev.counters()["n_candidates"]
    pair_ms = prof["pair"][0]
    pair_warm_ms = dev.time_pair(50)
    log(f"[bench] profile: {json.dumps(prof)} nc={nc} pair back-to-back {pair_warm_ms:.5f} ms")
    balanced = int(dev.launch_shape()["pair_block"]) == 0
    tune_layout = int(dev.launch_shape()["integrate_layout"])
    kname = ("k_pair_bal" if balanced else "k_pair_fused") + (
        "<true, false>" if args.precision == "mixed" else "<false, false>")
    ttag = f"{TRAFFIC_TAG[args.workload]}_{args.precision}" if args.workload in TRAFFIC_TAG else None
    # the pair pass: the step's largest single kernel on the uneven-length
    # system and the kernel the HBM target names
    bytes_pair = pair_bytes(args.precision, fil.nb, nc)
    achieved = bytes_pair / (pair_ms * 1e-3) / 1e9 if pair_ms > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(ttag, kname) if ttag else None,
                "kernel": kname,
                "kernel_avg_ms": pair_ms, "kernel_avg_ms_back_to_back": pair_warm_ms,
                "algorithmic_bytes": bytes_pair,
                "algorithmic_bytes_fp32_state": pair_bytes_fp32_state(fil.nb, nc),
                "candidates": nc,
                "per_step_ms": {k: v[0] for k, v in prof.items()}}
    # the integrator: all its length-bin launches of one step
    int_ms = prof["integrate"][0]
    bytes_int = integrator_bytes(fil.ns, fil.nb)
    achieved_int = bytes_int / (int_ms * 1e-3) / 1e9 if int_ms > 0 else 0.0
    int_traffic = None
    if ttag and args.precision == "mixed":
        # the kernels this launch shape runs: layout 2 one launch for every
        # filament (k_integrate_all, or k_integrate_packed_any without
        # filaments above 128 sites); layout 3 one launch per lane-group
        # width plus the longer filaments' length bin
        if tune_layout == 2:
            tf = measured_traffic_sum(ttag, "k_integrate_all<", "false>")
            tm = measured_traffic_sum(ttag, "k_integrate_all<", "true>")
            if tf is None or tm is None:
                tf = measured_traffic_sum(ttag, "k_integrate_packed_any<", "false>")
                tm = measured_traffic_sum(ttag, "k_integrate_packed_any<", "true>")
        else:
            tf = measured_traffic_sum(ttag, "k_integrate_packed<", "false>")
            tm = measured_traffic_sum(ttag, "k_integrate_packed<", "true>")
            if tf is not None and tm is not None:
                tf += measured_traffic_sum(ttag, "k_integrate_long<", "false>") or 0.0
                tm += measured_traffic_sum(ttag, "k_integrate_long<", "true>") or 0.0
        if tf is not None and tm is not None:
            int_traffic = 0.5 * (tf + tm)
    roofline_int = {"bound": "hbm", "achieved": round(achieved_int, 2), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved_int / HBM_PEAK_GBS, 4),
                    "traffic": int_traffic,
                    "kernel": "integrator (every filament of a step: k_integrate_all / "
                              "k_integrate_packed_any, or per-width k_integrate_packed "
                              "and k_integrate_long)",
                    "kernel_avg_ms": int_ms, "algorithmic_bytes": bytes_int}
    r_end = dev.sites()[0]
    dev.close()

    extra = {}
    if not args.no_accuracy and args.precision == "mixed" and xl is None:
        # the metric's accuracy half: FP64 and mixed systems loaded from the
        # same state (the positions the timed run ended at), one step each
        from simcore_amd.accuracy import mixed_vs_fp64
        fil_end = sc.Filaments(fil.site_off, r_end, fil.fil)
        acc = mixed_vs_fp64(p, fil_end, steps=1, spatial_order=spatial)
        extra["accuracy_vs_fp64"] = acc
        log(f"[bench] accuracy: {json.dumps(acc)}")
    if not args.no_fp64_baseline and args.precision != "fp64":
        legs = {}
        for graph in (False, True):
            d64 = sc.FilamentSystem(p, fil, precision="fp64", use_graph=graph)
            if xl is not None:
                d64.load_crosslinks(xl[0], xl[1], **xl_params)
            d64.step(10)
            w64 = timed_steps(d64, args.fp64_steps)
            legs[graph] = fil.nb * args.fp64_steps / w64
            if not graph:
                d64.profile(True)
                d64.step(args.profile_steps)
                p64 = d64.profile_read()
                d64.profile(False)
            d64.close()
        extra["plain_fp64_gpu"] = {"value": legs[False], "unit": "segment-steps/s",
                                   "steps": args.fp64_steps,
                                   "note": "one system at a time, FP64 everywhere "
                                           "(bit-identical to the CPU path), caller order, "
                                           "eager launches",
                                   "per_step_ms": {k: v[0] for k, v in p64.items()}}
        extra["fp64_gpu_graph"] = {"value": legs[True], "unit": "segment-steps/s",
                                   "steps": args.fp64_steps,
                                   "note": "the same FP64 path replayed from hipGraphs"}
        extra["vs_plain_fp64"] = value / legs[False]
        extra["vs_fp64_graph"] = value / legs[True]
        log(f"[bench] fp64 eager {legs[False]:.4e}, graph {legs[True]:.4e}; "
            f"mixed/plain {value / legs[False]:.2f}x")

    cpu = None
    if not args.no_cpu_baseline:
        from oracle.oracle_py import OracleSystem
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
        if threads <= 0:
            threads = len(os.sched_getaffinity(0))
        ora = OracleSystem(p, fil, n_threads=threads)
        if xl is not None:
            ora.load_crosslinks(xl[0], xl[1], **xl_params)
        ora.step(10)
        t0 = time.perf_counter()
        rc = ora.step(args.cpu_steps)
        tcpu = time.perf_counter() - t0
        if rc != 0:
            raise RuntimeError(f"oracle step failed: {rc}")
        host = host_cpu()
        cpu = {"value": fil.nb * args.cpu_steps / tcpu, "unit": "segment-steps/s",
               "cores": threads, "kind": "port",
               "sample": f"{args.cpu_steps} steps of the same workload after 10 warm-up "
                         f"steps (incl. the initial cell-list build), {tcpu:.2f} s wall, "
                         f"{threads} OpenMP threads, static chunks as in the reference",
               "cpu_model": host["model"], "logical_cpus_available": host["logical"],
               "physical_cores_available": host["physical"]}
        log(f"[bench] cpu oracle: {cpu['value']:.4e} on {threads} threads ({host})")
        ora.close()

    out = {
        "metric": "segment-steps/s", "value": value, "unit": "segment-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if args.precision == "mixed" else "f64",
        "data": "synthetic",
        "config": dict(desc, precision=args.precision, graph=True, spatial_order=spatial,
                       launch_shape=tune),
        # `roofline`: the step's dominant kernel by measured time (on both
        # systems now the integrator's one launch); the pair pass beside it
        "roofline": dict(roofline_int if roofline_int["kernel_avg_ms"] >= roofline["kernel_avg_ms"]
                         else roofline),
        "roofline_pair": roofline, "roofline_integrator": roofline_int, "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


def host_cpu():
    """CPU model and the logical / physical cores this process may run on."""
    model, phys = "unknown", set()
    allowed = os.sched_getaffinity(0)
    try:
        cur = {}
        with open("/proc/cpuinfo") as fh:
            for line in fh.read().splitlines() + [""]:
                if not line.strip():
                    if "processor" in cur and int(cur["processor"]) in allowed:
                        phys.add((cur.get("physical id", "0"), cur.get("core id",
                                                                        cur["processor"])))
                    cur = {}
                    continue
                k, _, v = line.partition(":")
                cur[k.strip()] = v.strip()
                if k.strip() == "model name":
                    model = v.strip()
    except OSError:
        pass
    return {"model": model, "logical": len(allowed), "physical": len(phys) or None}


if __name__ == "__main__":
    main()
