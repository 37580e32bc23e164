#!/usr/bin/env python3
"""Benchmark: spin-operator updates/s of the gfx950 loop-cluster update.

Workload (default): S=1/2 Heisenberg antiferromagnet on a 128x128 periodic
square lattice (32768 bonds), J=1, beta=128, one Markov chain on the plain
one-chain-per-launch GPU path, synthetic start (all-up spins, empty operator
string, SEED=23125).  A "step" is one Monte Carlo sweep (diagonal update,
world-line linking, cluster identification, flip, improved estimators).
`value` = sum over the timed sweeps of the operator count after each
diagonal update, divided by wall seconds, with the whole state resident in
HBM.  Prints one JSON line.

--workload scan runs the batched temperature scan instead (512 chains of the
32x32 square lattice, T_i = 0.05 * 40^(i/511)), plain (one chain per launch)
and batched (all chains per launch, HIP graph), and reports both.
"""

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "alps-looper_amd"))

METRIC = "spin-operator updates/s per GPU, batched vs plain; p50/p99 scan-job latency"
HBM_PEAK_GBS = 8000.0

# Algorithmic HBM bytes per operator and sweep of each kernel (compulsory
# traffic of the SoA layout: 12-byte operators, two segment nodes per
# operator).  Stage totals: diagonal update 32, linking 64, union-find 96,
# flip 24, estimators 64 = 280 B per operator-sweep.
BYTES_PER_OP = {
    "diag_count": 12.0, "diag_fill": 20.0,       # diagonal update      32
    "link": 64.0,                                 # world-line linking   64
    "uf_union": 64.0, "uf_compress": 32.0,        # union-find           96
    "estimate": 64.0,                             # estimators           64
    "flip_collect": 24.0,                         # flip                 24
    "scan_reduce": 0.0, "scan_top": 0.0, "scan_down": 0.0, "end_sweep": 0.0,
}
SWEEP_BYTES_PER_OP = 280.0


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def run_chunked(fn, n, chunk=250):
    """run n sweeps through fn(k) in chunks, printing progress; returns records"""
    import numpy as np
    out = []
    done = 0
    while done < n:
        k = min(chunk, n - done)
        out.append(fn(k))
        done += k
        log("%d/%d sweeps" % (done, n))
    return np.concatenate(out, axis=0)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(kind, L, beta, seconds):
    """Oracle looper-mode port (serial restatement of path_integral.C),
    1 core, bounded sample of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    ch = O.LooperChain(kind, L, beta, 1.0, 1.0, seed=23125)
    t_start = time.perf_counter()
    warm = 0.25 * seconds
    nops = 0.0
    tsum = 0.0
    n_timed = 0
    n_warm = 0
    while True:
        t0 = time.perf_counter()
        rec = ch.sweep(1)
        dt = time.perf_counter() - t0
        el = time.perf_counter() - t_start
        if el <= warm:
            n_warm += 1
        else:
            nops += rec[0, 0]
            tsum += dt
            n_timed += 1
        if el > seconds and n_timed > 0:
            break
    return {
        "value": nops / tsum,
        "unit": "spin-operator updates/s",
        "cores": 1,
        "kind": "port",
        "sample": ("oracle looper-mode restatement of path_integral.C (g++ -O2), same lattice/beta, "
                   "from the empty operator string: %d sweeps untimed, then %d timed sweeps "
                   "(mean n_op %.0f) in %.1f s; CPU: %s" % (n_warm, n_timed, nops / n_timed, tsum, cpu_model())),
    }


def measured_traffic(name, kernel, nop_mean):
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE/WRITE_SIZE
    summary (profiles/<name>.json, copy-kernel calibrated), scaled to nop_mean"""
    path = os.path.join(ROOT, "profiles", name + ".json")
    try:
        with open(path) as f:
            d = json.load(f)
        return d["kernels"][kernel]["bytes_per_op"] * nop_mean
    except (OSError, KeyError, TypeError):
        return None


def roofline(prof, nop_mean, traffic_profile=None):
    dom = max(prof, key=lambda k: prof[k])
    ms = prof[dom]
    bytes_launch = BYTES_PER_OP[dom] * nop_mean
    achieved = bytes_launch / (ms * 1e-3) / 1e9
    return {
        "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": measured_traffic(traffic_profile, dom, nop_mean) if traffic_profile else None,
        "kernel": dom, "kernel_ms": ms, "bytes_per_op": BYTES_PER_OP[dom],
        "bytes_per_launch": bytes_launch,
    }


def run_c2(args, lg):
    kind, L, beta = lg.SQUARE, 128, 128.0
    ch = lg.Chain(lg.params(kind, L, beta, 1.0, 1.0, seed=23125))
    log("warmup %d sweeps" % args.warmup)
    run_chunked(lambda k: ch.sweep(k), args.warmup)
    log("timed %d sweeps" % args.steps)
    t0 = time.perf_counter()
    rec = run_chunked(lambda k: ch.sweep(k), args.steps)
    t1 = time.perf_counter()
    dt = t1 - t0
    nop_total = float(rec[:, 0].sum())
    nop_mean = nop_total / args.steps
    value = nop_total / dt
    log("profile %d sweeps" % args.profile_sweeps)
    prof = ch.profile(args.profile_sweeps)
    rl = roofline(prof, nop_mean, "c2_square128_hbm_traffic")
    out = {
        "metric": METRIC, "value": value, "unit": "spin-operator updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "square 128x128 periodic S=1/2 Heisenberg AF, J=1, beta=128, "
                               "one chain, plain GPU path", "lattice": "square lattice", "L": L,
                   "beta": beta, "J": 1.0, "chains": 1, "step": "1 sweep", "seed": 23125},
        "sweeps_per_s": args.steps / dt,
        "n_op_mean": nop_mean,
        "energy_density": float(rec[:, 2].mean() / (L * L)),
        "sweep_hbm_frac": value * SWEEP_BYTES_PER_OP / (HBM_PEAK_GBS * 1e9),
        "kernel_ms": prof,
        "roofline": rl,
    }
    return out, (kind, L, beta)


def scan_params(lg, n=512, L=32):
    temps = [0.05 * 40.0 ** (i / (n - 1)) for i in range(n)]
    return [lg.params(lg.SQUARE, L, 1.0 / T, 1.0, 1.0, seed=23125, chain_id=i) for i, T in enumerate(temps)]


def run_scan(args, lg):
    plist = scan_params(lg)
    b = lg.Batch(plist)
    log("batched warmup %d sweeps" % args.warmup)
    run_chunked(lambda k: b.run(k), args.warmup)
    log("tune")
    b.tune(8)
    b.set_graph(True)
    log("batched timed %d sweeps" % args.steps)
    t0 = time.perf_counter()
    rec = run_chunked(lambda k: b.run(k), args.steps)
    dt = time.perf_counter() - t0
    nop_total = float(rec[:, :, 0].sum())
    batched = nop_total / dt
    # plain path: the same chains, one chain (and one launch per stage) at a
    # time; every 8th chain of the scan, a few sweeps each
    plain_sweeps = max(2, min(20, args.steps // 10))
    sub = plist[:: max(1, len(plist) // 64)]
    log("plain path: %d chains" % len(sub))
    chains = [lg.Chain(p) for p in sub]
    for c in chains:
        c.sweep(args.warmup, records=False)
    t0 = time.perf_counter()
    nops = 0.0
    for c in chains:
        nops += float(c.sweep(plain_sweeps)[:, 0].sum())
    plain = nops / (time.perf_counter() - t0)
    log("profile")
    prof = b.profile(args.profile_sweeps)
    rl = roofline(prof, nop_total / args.steps)
    out = {
        "metric": METRIC, "value": batched, "unit": "spin-operator updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "512-chain temperature scan, square 32x32, T_i=0.05*40^(i/511), batched + HIP graph",
                   "chains": len(plist), "L": 32},
        "plain_value": plain, "batched_over_plain": batched / plain,
        "kernel_ms": prof, "roofline": rl,
    }
    return out, None


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=400)
    ap.add_argument("--profile-sweeps", type=int, default=100)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=["c2", "scan"], default="c2")
    ap.add_argument("--copy", type=int, default=0, metavar="MIB",
                    help="only run the plain copy kernel on MIB MiB (counter calibration) and print GB/s")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on exactly one GPU; --gpus %d is not supported "
              "(independent chains shard trivially across GPUs, see the design notes)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import loopgpu as lg
    if lg.device_count() < 1:
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)
    if args.copy:
        gbs = lg.copy_bandwidth(args.copy << 20, 20)
        print(json.dumps({"copy_mib": args.copy, "iters": 21, "gbs": gbs}), flush=True)
        return
    if args.workload == "c2":
        out, cpu_cfg = run_c2(args, lg)
    else:
        out, cpu_cfg = run_scan(args, lg)
    if cpu_cfg and not args.no_cpu:
        log("cpu baseline %.0f s" % args.cpu_seconds)
        out["cpu_baseline"] = cpu_baseline(*cpu_cfg, args.cpu_seconds)
        out["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
