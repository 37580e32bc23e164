#!/usr/bin/env python3
"""Benchmark: spin-operator updates/s of the gfx950 loop-cluster update.

Default workload (the metric's headline: batched vs plain, p50/p99 job
latency), one JSON line:

* batched temperature scan: 512 independent chains of the S=1/2 Heisenberg
  antiferromagnet on the 32x32 periodic square lattice, J=1,
  T_i = 0.05 * 40^(i/511), chain i keyed (23125, i), all chains packed into
  one launch per sweep stage, the sweep replayed from a captured HIP graph
  with launch settings chosen by the autotuner.  A "step" is one batched
  sweep of all 512 chains.  Before the --warmup steps the batch is
  equilibrated internally (--equilibrate sweeps from the all-up, empty
  operator string start, whatever --warmup is passed; stated in the output).
  `value` = sum over the timed sweeps and chains of the operator count after
  each diagonal update, divided by wall seconds, state resident in HBM;
* `plain_value`: the same chains on the plain one-chain-per-launch path
  (every 8th chain of the scan, equilibrated the same way), and
  `batched_over_plain`;
* `serving`: the serving mix at its full job lengths through the
  continuous-batching job server -- Poisson arrivals, chain L in
  {64, 256, 1024} / square L in {16, 32, 48, 64, 96}, beta ~ U[1, 4L],
  Jz/Jxy ~ U[0.5, 1], 1000-20000 measurement sweeps (+1/8 thermalization)
  per job -- p50/p99 job latency (submit -> result) and delivered updates/s;
* `roofline` of the dominant batched kernel (HIP events on the batch's
  stream), `cpu_baseline` (the oracle's looper-mode port on a thread pool of
  the host cores, the same chains).

--workload scan / serve run one of the two parts; c2 (square 128^2,
beta=128, one chain), c5 (square 512^2, beta=512, one chain), chain128 /
chain1024 (the reference's published chain benchmark parameters) run single
chains on the plain path.
"""

import argparse
import json
import os
import platform
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "alps-looper_amd"))

METRIC = "spin-operator updates/s per GPU, batched vs plain; p50/p99 scan-job latency"
HBM_PEAK_GBS = 8000.0

# Algorithmic HBM bytes per operator and sweep of each kernel (the stage
# budget of the SoA layout: 12-byte operators, two segment nodes per
# operator).  Stage totals: diagonal update 32, linking 64, union-find 96
# (hooking 64, compression 32, done by uf_compress or inside estimate),
# flip 24, estimators 64 = 280 B per operator-sweep.
BYTES_PER_OP = {
    "diag_count": 12.0, "diag_fill": 20.0,       # diagonal update      32
    "link": 64.0,                                 # world-line linking   64
    "uf_union": 64.0, "uf_compress": 32.0,        # union-find           96
    "estimate": 64.0,                             # estimators           64
    "flip_collect": 24.0,                         # flip                 24
    "scan": 0.0, "end_sweep": 0.0,
}
SWEEP_BYTES_PER_OP = 280.0
SEED = 23125


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def run_chunked(fn, n, chunk=250, what="sweeps"):
    """run n sweeps through fn(k) in chunks, printing progress; returns records"""
    import numpy as np
    out = []
    done = 0
    while done < n:
        k = min(chunk, n - done)
        r = fn(k)
        if r is not None:
            out.append(r)
        done += k
        log("%d/%d %s" % (done, n, what))
    return np.concatenate(out, axis=0) if out else None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def _pool_size():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


PORT = ("the oracle's looper-mode restatement of path_integral.C (g++ -O3 -march=x86-64-v3; "
        "0.88x the reference's own standalone/loop.C on its published chain L=128, T=0.01 shape)")
PORT_SSE = "the oracle's looper-mode restatement of sse.C (g++ -O3 -march=x86-64-v3)"


def cpu_baseline(kind, L, beta, seconds):
    """Oracle looper-mode port, 1 core, bounded sample of the same workload:
    the first sweeps (from the empty operator string) and the first quarter
    of the time are untimed."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    ch = O.LooperChain(kind, L, beta, 1.0, 1.0, seed=SEED)
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
        if dt > 5:
            log("cpu sweep %d: %.1f s, n_op %.0f" % (n_warm + n_timed + 1, dt, rec[0, 0]))
        if el <= warm or n_warm == 0:
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
        "sample": ("%s, same lattice and beta, one chain from the empty operator string: %d sweeps untimed, "
                   "then %d timed sweeps (mean n_op %.0f) in %.1f s; CPU: %s"
                   % (PORT, n_warm, n_timed, nops / n_timed, tsum, cpu_model())),
    }


def cpu_pool_baseline(chains, seconds, what, untimed=8, max_sweeps=64, sse=False):
    """Oracle looper-mode chains on a thread pool of the host cores (the
    oracle releases the GIL inside each sweep): each worker takes the next
    chain of the list, runs `untimed` sweeps from the empty operator string,
    then timed sweeps until `max_sweeps` or the shared deadline.  The rate is
    the sum over workers of their timed operator updates / timed seconds.
    sse: the oracle's restatement of the SSE worker (sse.C) instead."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    cores = _pool_size()
    lock = threading.Lock()
    state = {"next": 0, "rate": 0.0, "sweeps": 0, "chains": 0}
    t_end = time.perf_counter() + seconds

    def worker():
        ops, secs, sweeps = 0.0, 0.0, 0
        while time.perf_counter() < t_end:
            with lock:
                i = state["next"]
                state["next"] += 1
            kind, L, beta, jz, seed = chains[i % len(chains)]
            ch = O.LooperChain(kind, L, beta, 1.0, jz, seed=seed)
            step = ch.sweep_sse if sse else ch.sweep
            n = 0
            while time.perf_counter() < t_end and n < untimed + max_sweeps:
                t0 = time.perf_counter()
                r = float(step(1)[0, 0])
                dt = time.perf_counter() - t0
                if n >= untimed:
                    ops += r
                    secs += dt
                    sweeps += 1
                n += 1
            with lock:
                state["chains"] += 1
        with lock:
            if secs > 0:
                state["rate"] += ops / secs
            state["sweeps"] += sweeps

    t0 = time.perf_counter()
    th = [threading.Thread(target=worker) for _ in range(cores)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    wall = time.perf_counter() - t0
    return {
        "value": state["rate"], "unit": "spin-operator updates/s", "cores": cores, "kind": "port",
        "sample": ("%s on a %d-thread pool for %.1f s: %s; %d chains, each %d untimed sweeps from the empty "
                   "operator string then up to %d timed sweeps, %d timed sweeps in total; CPU: %s"
                   % (PORT_SSE if sse else PORT, cores, wall, what, state["chains"], untimed, max_sweeps, state["sweeps"], cpu_model())),
    }


def device_kernels(kernel, link):
    """device kernels behind a profiled sweep stage under linking setting
    `link` (0: k_node_init + k_link_ops, which also unites; 1: k_link_site;
    else the LDS-staged k_link), by their short names"""
    if kernel == "link":
        return ("node_init", "link_ops") if link == 0 else ("link_site",) if link == 1 else ("link",)
    if kernel == "uf_union":
        return () if link == 0 else ("uf_union",)
    return (kernel,)


def measured_traffic(name, kernel, nop_mean, link=0):
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE/WRITE_SIZE
    summary (profiles/<name>.json, copy-kernel calibrated), scaled to nop_mean"""
    if not name:
        return None
    path = os.path.join(ROOT, "profiles", name + ".json")
    try:
        with open(path) as f:
            d = json.load(f)
        ks = device_kernels(kernel, link)
        return sum(d["kernels"][k]["bytes_per_op"] for k in ks) * nop_mean if ks else None
    except (OSError, KeyError, TypeError):
        return None


def kernel_bytes(compress, link=0):
    """algorithmic bytes per operator of each kernel: without the compress
    pass the estimator's own walk does the compression; the per-operator
    linking kernel (link 0) also does the unions"""
    b = dict(BYTES_PER_OP)
    if not compress:
        b["estimate"] += b["uf_compress"]
        b["uf_compress"] = 0.0
    if link == 0:
        b["link"] += b["uf_union"]
        b["uf_union"] = 0.0
    return b


def roofline(prof, nop_mean, traffic_profile=None, compress=1, link=0):
    kb = kernel_bytes(compress, link)
    dom = max(prof, key=lambda k: prof[k])
    ms = prof[dom]
    bytes_launch = kb[dom] * nop_mean
    achieved = bytes_launch / (ms * 1e-3) / 1e9
    return {
        "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": measured_traffic(traffic_profile, dom, nop_mean, link),
        "kernel": dom, "device_kernels": list(device_kernels(dom, link)), "kernel_ms": ms,
        "bytes_per_op": kb[dom], "bytes_per_launch": bytes_launch,
    }


STAGES = {"diagonal update": ("diag_count", "scan", "diag_fill"),
          "linking": ("link",), "union-find": ("uf_union", "uf_compress"),
          "estimators": ("estimate",), "flip + collect": ("flip_collect", "end_sweep")}


def stage_gbs(prof, nop_mean, compress=1, link=0):
    """algorithmic GB/s of each stage (per-stage bytes per operator-sweep)"""
    kb = kernel_bytes(compress, link)
    out = {}
    for st, ks in STAGES.items():
        ms = sum(prof[k] for k in ks)
        b = sum(kb[k] for k in ks) * nop_mean
        out[st] = {"ms": ms, "bytes_per_op": sum(kb[k] for k in ks), "gbs": b / (ms * 1e-3) / 1e9 if ms else None}
    return out


def base_line(args, value, ms_per_step, config):
    return {
        "metric": METRIC, "value": value, "unit": "spin-operator updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (all-up spins, empty operator string, Philox streams keyed by seed and chain)",
        "config": config,
    }


# ---------------------------------------------------------------------------
# single chains, plain path
# ---------------------------------------------------------------------------
def run_single(args, lg, kind, L, beta, traffic_profile=None, workload=None):
    log("allocating %s L=%d beta=%g" % ("chain" if kind == lg.CHAIN else "square", L, beta))
    ch = lg.Chain(lg.params(kind, L, beta, 1.0, 1.0, seed=SEED))
    chunk = 25 if (kind == lg.SQUARE and L > 128) else 250
    equil = args.equilibrate if args.equilibrate is not None else (200 if chunk == 25 else 400)
    log("equilibrate %d sweeps, then warmup %d" % (equil, args.warmup))
    run_chunked(lambda k: ch.sweep(k, records=False), equil, chunk)
    # the warmup takes records like the timed sweeps: the first copy of
    # records to the host costs a few milliseconds once (runtime staging)
    if args.warmup:
        run_chunked(lambda k: ch.sweep(k), args.warmup, chunk)
    log("timed %d sweeps" % args.steps)
    t0 = time.perf_counter()
    rec = run_chunked(lambda k: ch.sweep(k), args.steps, chunk)
    dt = time.perf_counter() - t0
    nop_total = float(rec[:, 0].sum())
    nop_mean = nop_total / args.steps
    value = nop_total / dt
    log("profile %d sweeps" % args.profile_sweeps)
    prof = ch.profile(args.profile_sweeps)
    comp, lk = ch.launch()["compress"], ch.launch()["link"]
    N = L if kind == lg.CHAIN else L * L
    cfg = {"workload": workload or ("%s periodic S=1/2 Heisenberg AF, J=1, beta=%g, one chain, plain GPU path"
                                    % ("chain L=%d" % L if kind == lg.CHAIN else "square %dx%d" % (L, L), beta)),
           "lattice": "chain lattice" if kind == lg.CHAIN else "square lattice", "L": L, "beta": beta, "J": 1.0,
           "chains": 1, "step": "1 sweep", "seed": SEED}
    out = base_line(args, value, 1e3 * dt / args.steps, cfg)
    out.update({
        "equilibration_sweeps": equil,
        "sweeps_per_s": args.steps / dt,
        "n_op_mean": nop_mean,
        "energy_density": float(rec[:, 2].mean() / N),
        "sweep_hbm_frac": value * SWEEP_BYTES_PER_OP / (HBM_PEAK_GBS * 1e9),
        "kernel_ms": prof,
        "stages": stage_gbs(prof, nop_mean, comp, lk),
        "roofline": roofline(prof, nop_mean, traffic_profile, comp, lk),
    })
    if not args.no_cpu:
        log("cpu baseline %.0f s" % args.cpu_seconds)
        out["cpu_baseline"] = cpu_baseline(kind, L, beta, args.cpu_seconds)
        out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    else:
        out["cpu_baseline"] = None
    return out


# ---------------------------------------------------------------------------
# batched temperature scan (headline)
# ---------------------------------------------------------------------------
def scan_temps(n=512):
    return [0.05 * 40.0 ** (i / (n - 1)) for i in range(n)]


def scan_params(lg, n=512, L=32):
    return [lg.params(lg.SQUARE, L, 1.0 / T, 1.0, 1.0, seed=SEED, chain_id=i) for i, T in enumerate(scan_temps(n))]


def run_scan(args, lg):
    plist = scan_params(lg)
    equil = args.equilibrate if args.equilibrate is not None else 400
    # plain path first: the same chains one chain (and one launch per stage)
    # at a time -- every 8th chain of the scan, equilibrated like the batch --
    # so that the batch's profiled sweeps are the workload's last launches
    sub = plist[::8]
    plain_sweeps = 20
    log("scan: plain path, %d chains, %d equilibration + %d timed sweeps each" % (len(sub), equil, plain_sweeps))
    chains = []
    for c in sub:
        ch = lg.Chain(c)
        ch.sweep(equil, records=False)
        ch.sweep(2)  # untimed, with records like the timed sweeps
        chains.append(ch)
    t0 = time.perf_counter()
    nops = 0.0
    for ch in chains:
        nops += float(ch.sweep(plain_sweeps)[:, 0].sum())
    tp = time.perf_counter() - t0
    plain = nops / tp
    for ch in chains:
        ch.close()
    b = lg.Batch(plist)
    log("scan: equilibrate %d batched sweeps" % equil)
    run_chunked(lambda k: b.run(k, records=False), equil, 200)
    log("scan: tune")
    b.tune(8)
    b.set_graph(True)
    log("scan: warmup %d" % args.warmup)
    if args.warmup:
        b.run(args.warmup)  # with records, as the timed sweeps
    log("scan: timed %d batched sweeps" % args.steps)
    t0 = time.perf_counter()
    rec = b.run(args.steps)
    dt = time.perf_counter() - t0
    nop_total = float(rec[:, :, 0].sum())
    batched = nop_total / dt
    nop_sweep = nop_total / args.steps
    # sustained rate over a longer stretch (the same graph replay)
    sustained = None
    if args.sustain > 0:
        log("scan: sustained %d batched sweeps" % args.sustain)
        t0 = time.perf_counter()
        rs = run_chunked(lambda k: b.run(k), args.sustain, 500)
        sustained = float(rs[:, :, 0].sum()) / (time.perf_counter() - t0)
    log("scan: profile %d sweeps" % args.profile_sweeps)
    prof = b.profile(args.profile_sweeps)
    launch = b.launch()
    cfg = {"workload": "512-chain temperature scan: S=1/2 Heisenberg AF on the 32x32 periodic square lattice, "
                       "J=1, T_i=0.05*40^(i/511), chain i keyed (23125, i); batched (one launch per stage for all "
                       "chains), HIP-graph sweep, autotuned launch settings",
           "lattice": "square lattice", "L": 32, "chains": len(plist), "T_min": 0.05, "T_max": 2.0,
           "step": "1 batched sweep of all 512 chains", "seed": SEED}
    out = base_line(args, batched, 1e3 * dt / args.steps, cfg)
    E = rec[:, :, 2].mean(axis=0) / 1024.0
    out.update({
        "equilibration_sweeps": equil,
        "sweeps_per_s": args.steps / dt,
        "n_op_per_batched_sweep": nop_sweep,
        "plain_value": plain, "batched_over_plain": batched / plain,
        "plain_sample": "%d chains (every 8th of the scan), %d sweeps each after %d equilibration sweeps, "
                        "one chain per launch, %.3f s" % (len(sub), plain_sweeps, equil, tp),
        "sustained_value": sustained, "sustained_sweeps": args.sustain,
        "energy_density_coldest_hottest": [float(E[0]), float(E[-1])],
        "launch": launch,
        "kernel_ms": prof,
        "stages": stage_gbs(prof, nop_sweep, launch["compress"], launch["link"]),
        "sweep_hbm_frac": batched * SWEEP_BYTES_PER_OP / (HBM_PEAK_GBS * 1e9),
        "roofline": roofline(prof, nop_sweep, "c3_scan512_hbm_traffic", launch["compress"], launch["link"]),
    })
    b.close()
    if not args.no_cpu:
        log("scan: cpu baseline %.0f s" % args.cpu_seconds)
        chains = [(1, 32, 1.0 / T, 1.0, SEED) for T in scan_temps()[::8]]
        out["cpu_baseline"] = cpu_pool_baseline(chains, args.cpu_seconds,
                                                "every 8th chain of the scan, in order of temperature")
        out["gpu_over_cpu"] = batched / out["cpu_baseline"]["value"]
    else:
        out["cpu_baseline"] = None
    return out


def run_sse_scan(args, lg):
    """the 512-chain scan with the SSE worker ("loop; sse"): the same chains,
    batched, HIP-graph replay, autotuned; updates/s over the timed sweeps"""
    plist = [lg.params(lg.SQUARE, 32, 1.0 / T, 1.0, 1.0, seed=SEED, chain_id=i, algorithm=lg.SSE)
             for i, T in enumerate(scan_temps())]
    equil = args.equilibrate if args.equilibrate is not None else 400
    b = lg.Batch(plist)
    log("sse scan: equilibrate %d batched sweeps" % equil)
    run_chunked(lambda k: b.run(k, records=False), equil, 200)
    log("sse scan: tune")
    b.tune(8)
    b.set_graph(True)
    if args.warmup:
        b.run(args.warmup)
    log("sse scan: timed %d batched sweeps" % args.steps)
    t0 = time.perf_counter()
    rec = b.run(args.steps)
    dt = time.perf_counter() - t0
    nop_total = float(rec[:, :, 0].sum())
    prof = b.profile(args.profile_sweeps)
    cfg = {"workload": "512-chain temperature scan with the SSE worker (loop; sse): S=1/2 Heisenberg AF on "
                       "the 32x32 periodic square lattice, T_i=0.05*40^(i/511), chain i keyed (23125, i); batched, "
                       "HIP-graph sweep, autotuned launch settings",
           "lattice": "square lattice", "L": 32, "chains": len(plist), "algorithm": "loop; sse", "seed": SEED}
    out = base_line(args, nop_total / dt, 1e3 * dt / args.steps, cfg)
    E = rec[:, :, 2].mean(axis=0) / 1024.0
    out.update({"equilibration_sweeps": equil, "n_op_per_batched_sweep": nop_total / args.steps,
                "energy_density_coldest_hottest": [float(E[0]), float(E[-1])], "launch": b.launch(),
                "kernel_ms": prof, "roofline": None, "cpu_baseline": None})
    b.close()
    if not args.no_cpu:
        log("sse scan: cpu baseline %.0f s" % args.cpu_seconds)
        chains = [(1, 32, 1.0 / T, 1.0, SEED) for T in scan_temps()[::8]]
        out["cpu_baseline"] = cpu_pool_baseline(chains, args.cpu_seconds,
                                                "every 8th chain of the scan, in order of temperature, SSE worker",
                                                sse=True)
        out["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    return out


# ---------------------------------------------------------------------------
# serving mix
# ---------------------------------------------------------------------------
SERVE_KINDS = [("chain", 64), ("chain", 256), ("chain", 1024),
               ("square", 16), ("square", 32), ("square", 48), ("square", 64), ("square", 96)]


def serve_trace(n, sweep_scale=1.0, seed=SEED):
    """Job mix of the serving workload: kind uniform over SERVE_KINDS,
    beta ~ U[1, 4L], Jz ~ U[0.5, 1] with Jxy = 1, sweeps ~ U{1000..20000}
    (times sweep_scale), thermalization = sweeps / 8; job i is keyed
    (23125, i)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n):
        kind, L = SERVE_KINDS[int(rng.integers(len(SERVE_KINDS)))]
        beta = float(rng.uniform(1.0, 4.0 * L))
        jz = float(rng.uniform(0.5, 1.0))
        sweeps = max(8, int(round(int(rng.integers(1000, 20001)) * sweep_scale)))
        jobs.append({"kind": kind, "L": L, "beta": beta, "jz": jz, "sweeps": sweeps,
                     "thermalization": sweeps // 8, "chain_id": i})
    return jobs


def job_params(lg, j, seed=SEED):
    return lg.params(lg.CHAIN if j["kind"] == "chain" else lg.SQUARE, j["L"], j["beta"], 1.0, j["jz"],
                     seed=seed, chain_id=j["chain_id"])


def job_updates_estimate(j):
    """expected operator updates of a job: sweeps x beta x bonds x (offset +
    half the largest diagonal energy per bond), the estimate the library
    sizes its stages with"""
    nb = j["L"] if j["kind"] == "chain" else 2 * j["L"] ** 2
    per_bond = 0.25 + 0.5 * (j["jz"] / 4 + 0.5)
    return (j["sweeps"] + j["thermalization"]) * j["beta"] * nb * per_bond


def _percentile(xs, q):
    import numpy as np
    return float(np.percentile(np.asarray(xs), q)) if xs else None


SERVER_CFG = dict(max_jobs=256, max_ops=96 << 20, max_cand=96 << 20, max_sites=1 << 21, max_bonds=1 << 22)


def run_serve(args, lg):
    """Continuous batching of a Poisson job stream at the mix's full job
    lengths.  A calibration burst (all jobs submitted at once, stopped after
    --serve-calib seconds) measures the saturated rate; the arrival rate is
    then set for --load times it, arrivals run for --serve-seconds, and the
    stream drains."""
    import numpy as np
    trace = serve_trace(4096 + args.serve_burst, args.sweep_scale)
    cfg = dict(SERVER_CFG, max_chunk=args.serve_chunk)
    log("serve: calibration burst of %d jobs for %.0f s" % (args.serve_burst, args.serve_calib))
    srv = lg.Server(**cfg)
    for j in trace[:args.serve_burst]:
        srv.submit(job_params(lg, j, seed=SEED + 1), j["sweeps"], j["thermalization"])
    time.sleep(args.serve_calib)
    st = srv.stats()
    srv.stop()
    sat_rate = st["total_updates"] / max(st["busy_s"], 1e-9)
    stream = trace[args.serve_burst:]
    upd_per_job = float(np.mean([job_updates_estimate(j) for j in stream[:512]]))
    lam = args.load * sat_rate / upd_per_job
    log("serve: saturated %.3g updates/s, ~%.3g updates/job -> %.3g jobs/s" % (sat_rate, upd_per_job, lam))
    rng = np.random.default_rng(4242)
    srv = lg.Server(**cfg)
    ids = []
    t_start = time.perf_counter()
    t_next = t_start + rng.exponential(1.0 / lam)
    k = 0
    last_log = t_start
    while t_next - t_start <= args.serve_seconds:
        now = time.perf_counter()
        if now < t_next:
            time.sleep(min(t_next - now, 0.05))
            continue
        j = stream[k]
        k += 1
        ids.append(srv.submit(job_params(lg, j), j["sweeps"], j["thermalization"]))
        t_next += rng.exponential(1.0 / lam)
        if now - last_log > 10:
            s = srv.stats()
            log("serve: %.0f s, %d submitted, %d done, %d running, %d queued"
                % (now - t_start, s["submitted"], s["done"], s["running"], s["queued"]))
            last_log = now
    log("serve: arrivals closed after %d jobs; draining" % len(ids))
    res = []
    for n, i in enumerate(ids):
        while True:
            r = srv.wait(i, 10)
            if r["status"] in (lg.JOB_DONE, lg.JOB_CANCELLED):
                break
            s = srv.stats()
            log("serve: drain %d/%d done, %d running" % (s["done"], len(ids), s["running"]))
        res.append(r)
    wall = time.perf_counter() - t_start
    st = srv.stats()
    srv.stop()
    lat = [r["latency_s"] for r in res if r["status"] == lg.JOB_DONE]
    qs = [r["queue_s"] for r in res if r["status"] == lg.JOB_DONE]
    sweeps = [stream[n]["sweeps"] + stream[n]["thermalization"] for n in range(len(ids))]
    out = {
        "workload": "serving mix: Poisson job stream, chain L in {64,256,1024} / square L in {16,32,48,64,96}, "
                    "beta~U[1,4L], Jz/Jxy~U[0.5,1], Jxy=1, sweeps~U{1000..20000}x%g, thermalization sweeps/8, "
                    "continuous batching" % args.sweep_scale,
        "value": st["total_updates"] / wall, "unit": "spin-operator updates/s",
        "arrival_rate_jobs_per_s": lam, "target_load": args.load, "arrival_seconds": args.serve_seconds,
        "sweep_scale": args.sweep_scale, "max_chunk": args.serve_chunk,
        "jobs": len(ids), "jobs_done": len(lat), "job_sweeps_mean": float(np.mean(sweeps)) if sweeps else None,
        "p50_latency_s": _percentile(lat, 50), "p99_latency_s": _percentile(lat, 99),
        "mean_latency_s": float(np.mean(lat)) if lat else None, "max_latency_s": max(lat) if lat else None,
        "p50_queue_s": _percentile(qs, 50), "p99_queue_s": _percentile(qs, 99),
        "busy_fraction": st["busy_s"] / wall, "updates_per_busy_s": st["updates_per_s"],
        "batched_sweeps": int(st["sweep_launches"]), "ms_per_batched_sweep": 1e3 * st["busy_s"] / max(1, st["sweep_launches"]),
        "graph_captures": int(st["graph_captures"]), "graph_capture_s": st["graph_capture_s"],
        "ms_per_graph_capture": 1e3 * st["graph_capture_s"] / max(1, st["graph_captures"]),
        "saturated_value": sat_rate, "calibration_seconds": args.serve_calib, "wall_s": wall,
    }
    if not args.no_cpu:
        log("serve: cpu baseline %.0f s" % args.serve_cpu_seconds)
        jobs = [(0 if j["kind"] == "chain" else 1, j["L"], j["beta"], j["jz"], SEED) for j in stream[:max(1, k)]]
        out["cpu_baseline"] = cpu_pool_baseline(jobs, args.serve_cpu_seconds,
                                                "the jobs of the same arrival trace in arrival order")
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--equilibrate", type=int, default=None,
                    help="untimed sweeps from the empty operator string before the warmup (default 400; 200 for c5)")
    ap.add_argument("--sustain", type=int, default=2000, help="scan: batched sweeps of the sustained-rate stretch")
    ap.add_argument("--profile-sweeps", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true", help="default workload: skip the serving part")
    ap.add_argument("--workload", choices=["c3", "scan", "serve", "c2", "c5", "chain128", "chain1024", "sse"],
                    default="c3")
    ap.add_argument("--serve-seconds", type=float, default=90.0, help="serve: seconds of Poisson arrivals")
    ap.add_argument("--serve-calib", type=float, default=15.0, help="serve: seconds of the calibration burst")
    ap.add_argument("--serve-burst", type=int, default=64, help="serve: jobs of the calibration burst")
    ap.add_argument("--sweep-scale", type=float, default=1.0, help="serve: factor on every job's sweep count")
    ap.add_argument("--load", type=float, default=0.8, help="serve: offered load / saturated rate")
    ap.add_argument("--serve-chunk", type=int, default=64, help="serve: sweeps between admission points")
    ap.add_argument("--serve-cpu-seconds", type=float, default=15.0)
    ap.add_argument("--copy", type=int, default=0, metavar="MIB",
                    help="only run the plain copy kernel on MIB MiB (counter calibration) and print GB/s")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on exactly one GPU; --gpus %d is not supported "
              "(independent chains and jobs shard trivially across GPUs, one large chain by imaginary-time "
              "slabs; neither is built here)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import loopgpu as lg
    if lg.device_count() < 1:
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)
    if args.copy:
        gbs = lg.copy_bandwidth(args.copy << 20, 20)
        print(json.dumps({"copy_mib": args.copy, "iters": 21, "gbs": gbs}), flush=True)
        return
    if args.workload in ("c3", "scan"):
        out = run_scan(args, lg)
        if args.workload == "c3" and not args.no_serve:
            out["serving"] = run_serve(args, lg)
    elif args.workload == "serve":
        s = run_serve(args, lg)
        out = base_line(args, s.pop("value"), s["ms_per_batched_sweep"], {"workload": s.pop("workload")})
        out.update(s)
        out["roofline"] = None
        out.setdefault("cpu_baseline", None)
        if out["cpu_baseline"]:
            out["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    elif args.workload == "c2":
        out = run_single(args, lg, lg.SQUARE, 128, 128.0, "c2_square128_hbm_traffic")
    elif args.workload == "c5":
        out = run_single(args, lg, lg.SQUARE, 512, 512.0, "c5_square512_hbm_traffic")
    elif args.workload == "sse":
        out = run_sse_scan(args, lg)
    elif args.workload == "chain128":
        out = run_single(args, lg, lg.CHAIN, 128, 100.0, None,
                         "chain L=128 periodic S=1/2 Heisenberg AF, J=1, T=0.01, one chain, plain GPU path "
                         "(the reference's published benchmark parameters)")
        out["published_reference_sweeps_per_s"] = 412.0
    else:
        out = run_single(args, lg, lg.CHAIN, 1024, 1000.0, None,
                         "chain L=1024 periodic S=1/2 Heisenberg AF, J=1, T=0.001, one chain, plain GPU path "
                         "(the reference's published benchmark parameters)")
        out["published_reference_sweeps_per_s"] = 3.30
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
