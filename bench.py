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

--workload serve runs the serving mix through the continuous-batching job
server: a burst of jobs measures the saturated rate, then Poisson arrivals at
--load times that rate for --serve-seconds; reports the delivered updates/s
and p50/p99 job latency (submit -> result).  Job sweep counts are multiplied
by --sweep-scale (0.1 by default, so that the run takes minutes).

--workload c5 runs the 512x512 square lattice at beta=512.  --workload
chain128 / chain1024 run the reference's own published benchmark parameters
of the chain (L=128, T=0.01 and L=1024, T=0.001).

CPU baselines run the oracle's looper-mode restatement: one core for the
one-chain workloads, a thread pool over the host cores for scan and serve;
for the chain workloads the reference's own serial program (built from its
sources into oracle/_ref) on one core.
"""

import argparse
import json
import os
import platform
import re
import subprocess
import sys
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
    1 core, bounded sample of the same workload: the first sweep (from the
    empty operator string) and the first quarter of the time are untimed."""
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
        "sample": ("oracle looper-mode restatement of path_integral.C (g++ -O2), same lattice/beta, "
                   "from the empty operator string: %d sweeps untimed, then %d timed sweeps "
                   "(mean n_op %.0f) in %.1f s; CPU: %s" % (n_warm, n_timed, nops / n_timed, tsum, cpu_model())),
    }


def reference_program_baseline(L, T, nop_mean, seconds):
    """The reference's own serial kernel (standalone/loop.C, compiled from its
    sources into oracle/_ref) on 1 host core for the same chain: it prints its
    own sweep rate ("Speed = ... MCS/sec", standalone/loop.C); updates/s is
    that rate times the GPU run's mean operator count of the same model and
    temperature.  The sweep count is sized from a short first run so that the
    sample takes about `seconds`."""
    exe = os.path.join(ROOT, "oracle", "_ref", "standalone_loop")
    if not os.path.exists(exe):
        return None
    def speed(n):
        r = subprocess.run([exe, "-l", str(L), "-t", repr(T), "-n", str(n)], capture_output=True,
                           text=True, timeout=600)
        m = re.search(r"Speed = ([0-9.eE+-]+) MCS/sec", r.stdout + r.stderr)  # printed on stderr
        return float(m.group(1)) if m else None
    probe = speed(64)
    if not probe:
        return None
    n = max(64, int(probe * seconds / 1.125))  # the program adds n/8 thermalization sweeps
    v = speed(n)
    return {
        "value": v * nop_mean,
        "unit": "spin-operator updates/s",
        "cores": 1,
        "kind": "reference",
        "sample": ("the reference's standalone/loop.C (c++ -O3, oracle/_ref/standalone_loop) -l %d -t %g -n %d: "
                   "%.1f sweeps/s on its own clock, times the GPU run's mean n_op %.0f; CPU: %s"
                   % (L, T, n, v, nop_mean, cpu_model())),
        "sweeps_per_s": v,
    }


def run_chain(args, lg, L, T):
    """The reference's published benchmark configurations of the chain
    (standalone/README.en: "-l 128 -t 0.01" and "-l 1024 -t 0.001"), one chain,
    plain GPU path"""
    beta = 1.0 / T
    ch = lg.Chain(lg.params(lg.CHAIN, L, beta, 1.0, 1.0, seed=23125))
    chunk = 1000
    log("chain L=%d beta=%g: warmup %d sweeps" % (L, beta, args.warmup))
    run_chunked(lambda k: ch.sweep(k), args.warmup, chunk)
    log("timed %d sweeps" % args.steps)
    t0 = time.perf_counter()
    rec = run_chunked(lambda k: ch.sweep(k), args.steps, chunk)
    dt = time.perf_counter() - t0
    nop_mean = float(rec[:, 0].mean())
    prof = ch.profile(args.profile_sweeps)
    out = {
        "metric": METRIC, "value": float(rec[:, 0].sum()) / dt, "unit": "spin-operator updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * dt / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "chain L=%d periodic S=1/2 Heisenberg AF, J=1, T=%g (beta=%g), one chain, "
                               "plain GPU path (the reference's published benchmark parameters)" % (L, T, beta),
                   "lattice": "chain lattice", "L": L, "beta": beta, "J": 1.0, "chains": 1, "seed": 23125},
        "sweeps_per_s": args.steps / dt,
        "n_op_mean": nop_mean,
        "energy_density": float(rec[:, 2].mean() / L),
        "kernel_ms": prof,
        "roofline": roofline(prof, nop_mean, compress=ch.launch()["compress"]),
        "published_reference_sweeps_per_s": {128: 412.0, 1024: 3.30}.get(L),
    }
    return out, ("reference", (L, T, nop_mean))


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


def kernel_bytes(compress):
    """algorithmic bytes per operator of each kernel: without the compress
    pass the estimator's own walk does the compression"""
    b = dict(BYTES_PER_OP)
    if not compress:
        b["estimate"] += b["uf_compress"]
        b["uf_compress"] = 0.0
    return b


def roofline(prof, nop_mean, traffic_profile=None, compress=1):
    kb = kernel_bytes(compress)
    dom = max(prof, key=lambda k: prof[k])
    ms = prof[dom]
    bytes_launch = kb[dom] * nop_mean
    achieved = bytes_launch / (ms * 1e-3) / 1e9
    return {
        "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": measured_traffic(traffic_profile, dom, nop_mean) if traffic_profile else None,
        "kernel": dom, "kernel_ms": ms, "bytes_per_op": kb[dom],
        "bytes_per_launch": bytes_launch,
    }


STAGES = {"diagonal update": ("diag_count", "scan", "diag_fill"),
          "linking": ("link",), "union-find": ("uf_union", "uf_compress"),
          "estimators": ("estimate",), "flip + collect": ("flip_collect", "end_sweep")}


def stage_gbs(prof, nop_mean, compress=1):
    """algorithmic GB/s of each stage (per-stage bytes per operator-sweep)"""
    kb = kernel_bytes(compress)
    out = {}
    for st, ks in STAGES.items():
        ms = sum(prof[k] for k in ks)
        b = sum(kb[k] for k in ks) * nop_mean
        out[st] = {"ms": ms, "bytes_per_op": sum(kb[k] for k in ks), "gbs": b / (ms * 1e-3) / 1e9}
    return out


def run_c2(args, lg, L=128, beta=128.0, traffic_profile="c2_square128_hbm_traffic"):
    kind = lg.SQUARE
    log("allocating square %dx%d beta=%g" % (L, L, beta))
    ch = lg.Chain(lg.params(kind, L, beta, 1.0, 1.0, seed=23125))
    chunk = 250 if L <= 128 else 25  # progress lines every few seconds
    log("warmup %d sweeps" % args.warmup)
    run_chunked(lambda k: ch.sweep(k), args.warmup, chunk)
    log("timed %d sweeps" % args.steps)
    t0 = time.perf_counter()
    rec = run_chunked(lambda k: ch.sweep(k), args.steps, chunk)
    t1 = time.perf_counter()
    dt = t1 - t0
    nop_total = float(rec[:, 0].sum())
    nop_mean = nop_total / args.steps
    value = nop_total / dt
    log("profile %d sweeps" % args.profile_sweeps)
    prof = ch.profile(args.profile_sweeps)
    comp = ch.launch()["compress"]
    rl = roofline(prof, nop_mean, traffic_profile, comp)
    out = {
        "metric": METRIC, "value": value, "unit": "spin-operator updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "square %dx%d periodic S=1/2 Heisenberg AF, J=1, beta=%g, "
                               "one chain, plain GPU path" % (L, L, beta), "lattice": "square lattice", "L": L,
                   "beta": beta, "J": 1.0, "chains": 1, "step": "1 sweep", "seed": 23125},
        "sweeps_per_s": args.steps / dt,
        "n_op_mean": nop_mean,
        "energy_density": float(rec[:, 2].mean() / (L * L)),
        "sweep_hbm_frac": value * SWEEP_BYTES_PER_OP / (HBM_PEAK_GBS * 1e9),
        "kernel_ms": prof,
        "stages": stage_gbs(prof, nop_mean, comp),
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
    rl = roofline(prof, nop_total / args.steps, compress=b.launch()["compress"])
    out = {
        "metric": METRIC, "value": batched, "unit": "spin-operator updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "512-chain temperature scan, square 32x32, T_i=0.05*40^(i/511), batched + HIP graph",
                   "chains": len(plist), "L": 32},
        "plain_value": plain, "batched_over_plain": batched / plain,
        "launch": b.launch(),
        "kernel_ms": prof, "roofline": rl,
    }
    return out, ("scan", [(1, 32, 1.0 / T, 1.0, 23125) for T in (0.05 * 40.0 ** (i / 511) for i in range(512))])


SERVE_KINDS = [("chain", 64), ("chain", 256), ("chain", 1024),
               ("square", 16), ("square", 32), ("square", 48), ("square", 64), ("square", 96)]


def serve_trace(n, sweep_scale, seed=23125):
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


def _job_params(lg, j):
    return lg.params(lg.CHAIN if j["kind"] == "chain" else lg.SQUARE, j["L"], j["beta"], 1.0, j["jz"],
                     seed=23125, chain_id=j["chain_id"])


def _percentile(xs, q):
    import numpy as np
    return float(np.percentile(np.asarray(xs), q)) if xs else None


def run_serve(args, lg):
    """Continuous batching of a Poisson job stream.  A burst of jobs measures
    the saturated rate; the arrival rate is then set for ~80 % of it and the
    stream runs for --serve-seconds of arrivals, after which it drains."""
    import numpy as np
    burst_n = args.serve_burst
    trace = serve_trace(burst_n + 4096, args.sweep_scale)
    cfg = dict(max_jobs=256, max_ops=96 << 20, max_cand=96 << 20, max_sites=1 << 21, max_bonds=1 << 22,
               max_chunk=args.serve_chunk)
    # saturation: every burst job submitted at once
    srv = lg.Server(**cfg)
    t0 = time.perf_counter()
    ids = [srv.submit(_job_params(lg, j), j["sweeps"], j["thermalization"]) for j in trace[:burst_n]]
    for k, i in enumerate(ids):
        srv.wait(i, 3600)
        if (k + 1) % 8 == 0:
            log("burst: %d/%d jobs done" % (k + 1, burst_n))
    burst_wall = time.perf_counter() - t0
    st = srv.stats()
    srv.stop()
    upd_per_job = st["total_updates"] / burst_n
    sat_rate = st["total_updates"] / burst_wall
    lam = args.load * sat_rate / upd_per_job
    log("saturated %.3g updates/s, %.3g updates/job -> arrival rate %.3g jobs/s" % (sat_rate, upd_per_job, lam))
    # Poisson stream
    rng = np.random.default_rng(4242)
    srv = lg.Server(**cfg)
    ids = []
    t_start = time.perf_counter()
    t_next = t_start + rng.exponential(1.0 / lam)
    k = burst_n
    last_log = t_start
    while True:
        now = time.perf_counter()
        if t_next - t_start > args.serve_seconds:
            break
        if now < t_next:
            time.sleep(min(t_next - now, 0.05))
            continue
        j = trace[k]
        k += 1
        ids.append(srv.submit(_job_params(lg, j), j["sweeps"], j["thermalization"]))
        t_next += rng.exponential(1.0 / lam)
        if now - last_log > 10:
            s = srv.stats()
            log("stream %.0f s: %d submitted, %d done, %d running, %d queued" %
                (now - t_start, s["submitted"], s["done"], s["running"], s["queued"]))
            last_log = now
    log("arrivals closed after %d jobs; draining" % len(ids))
    res = []
    for n, i in enumerate(ids):
        res.append(srv.wait(i, 3600))
        if (n + 1) % 32 == 0:
            log("drain: %d/%d" % (n + 1, len(ids)))
    wall = time.perf_counter() - t_start
    st = srv.stats()
    srv.stop()
    lat = [r["latency_s"] for r in res if r["status"] == lg.JOB_DONE]
    qs = [r["queue_s"] for r in res if r["status"] == lg.JOB_DONE]
    value = st["total_updates"] / wall
    out = {
        "metric": METRIC, "value": value, "unit": "spin-operator updates/s",
        "n_gpus": 1, "steps": int(st["sweep_launches"]), "warmup": 0,
        "ms_per_step": 1e3 * st["busy_s"] / max(1, st["sweep_launches"]), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "serving mix: Poisson job stream, chain L in {64,256,1024} / square L in "
                               "{16,32,48,64,96}, beta~U[1,4L], Jz~U[0.5,1], Jxy=1, sweeps~U{1000..20000}"
                               "x%g, thermalization sweeps/8, continuous batching" % args.sweep_scale,
                   "arrival_rate_jobs_per_s": lam, "target_load": args.load, "arrival_seconds": args.serve_seconds,
                   "sweep_scale": args.sweep_scale, "max_chunk": args.serve_chunk, "burst_jobs": burst_n},
        "jobs": len(ids), "jobs_done": len(lat),
        "p50_latency_s": _percentile(lat, 50), "p99_latency_s": _percentile(lat, 99),
        "mean_latency_s": float(np.mean(lat)) if lat else None,
        "p50_queue_s": _percentile(qs, 50), "p99_queue_s": _percentile(qs, 99),
        "busy_fraction": st["busy_s"] / wall, "updates_per_busy_s": st["updates_per_s"],
        "saturated_value": sat_rate, "wall_s": wall, "roofline": None,
    }
    return out, ("serve", trace[burst_n:k] or trace[:burst_n])


def _pool_size():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_pool_baseline(chains, seconds, what):
    """Oracle looper-mode chains run by a thread pool on the host cores (the
    oracle releases the GIL inside each sweep); each worker takes the next
    chain of the list and sweeps it until the shared deadline.  Returns the
    aggregate operator updates/s."""
    import threading
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    cores = _pool_size()
    lock = threading.Lock()
    state = {"next": 0, "ops": 0.0, "sweeps": 0, "chains": 0}
    t_end = time.perf_counter() + seconds

    def worker():
        while time.perf_counter() < t_end:
            with lock:
                i = state["next"]
                state["next"] += 1
            kind, L, beta, jz, seed = chains[i % len(chains)]
            ch = O.LooperChain(kind, L, beta, 1.0, jz, seed=seed)
            n_ops = 0.0
            n = 0
            while time.perf_counter() < t_end and n < 64:
                n_ops += float(ch.sweep(1)[0, 0])
                n += 1
            with lock:
                state["ops"] += n_ops
                state["sweeps"] += n
                state["chains"] += 1

    t0 = time.perf_counter()
    th = [threading.Thread(target=worker) for _ in range(cores)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    wall = time.perf_counter() - t0
    return {
        "value": state["ops"] / wall, "unit": "spin-operator updates/s", "cores": cores, "kind": "port",
        "sample": ("oracle looper-mode restatement of path_integral.C (g++ -O2) on a %d-thread pool for %.1f s: "
                   "%s; %d chains started from the empty operator string, up to 64 sweeps each, %d sweeps in "
                   "total; CPU: %s" % (cores, wall, what, state["chains"], state["sweeps"], cpu_model())),
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=400)
    ap.add_argument("--profile-sweeps", type=int, default=100)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=["c2", "c5", "scan", "serve", "chain128", "chain1024"], default="c2")
    ap.add_argument("--serve-seconds", type=float, default=60.0, help="serve: seconds of Poisson arrivals")
    ap.add_argument("--serve-burst", type=int, default=48, help="serve: jobs of the saturation burst")
    ap.add_argument("--sweep-scale", type=float, default=0.1, help="serve: factor on every job's sweep count")
    ap.add_argument("--load", type=float, default=0.8, help="serve: offered load / saturated rate")
    ap.add_argument("--serve-chunk", type=int, default=64, help="serve: sweeps between admission points")
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
    elif args.workload == "c5":
        out, cpu_cfg = run_c2(args, lg, 512, 512.0, None)
    elif args.workload == "scan":
        out, cpu_cfg = run_scan(args, lg)
    elif args.workload == "chain128":
        out, cpu_cfg = run_chain(args, lg, 128, 0.01)
    elif args.workload == "chain1024":
        out, cpu_cfg = run_chain(args, lg, 1024, 0.001)
    else:
        out, cpu_cfg = run_serve(args, lg)
    if cpu_cfg and not args.no_cpu:
        log("cpu baseline %.0f s" % args.cpu_seconds)
        if cpu_cfg[0] == "reference":
            out["cpu_baseline"] = reference_program_baseline(*cpu_cfg[1], args.cpu_seconds)
            if out["cpu_baseline"] is None:  # reference program not built here: the oracle port
                L, T, _ = cpu_cfg[1]
                out["cpu_baseline"] = cpu_baseline(0, L, 1.0 / T, args.cpu_seconds)
        elif cpu_cfg[0] == "scan":
            out["cpu_baseline"] = cpu_pool_baseline(cpu_cfg[1], args.cpu_seconds,
                                                    "the 512 chains of the scan in order")
        elif cpu_cfg[0] == "serve":
            jobs = [(0 if j["kind"] == "chain" else 1, j["L"], j["beta"], j["jz"], 23125) for j in cpu_cfg[1]]
            out["cpu_baseline"] = cpu_pool_baseline(jobs, args.cpu_seconds, "the jobs of the same arrival "
                                                    "trace in arrival order")
        else:
            out["cpu_baseline"] = cpu_baseline(*cpu_cfg, args.cpu_seconds)
        out["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
