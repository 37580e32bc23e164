"""Summaries of rocprofv3 output (rocpd SQLite databases) for the sweep kernels.

    python alps-looper_amd/profiling.py trace  RESULTS.db [--last N] [--only k1,k2,...]
    python alps-looper_amd/profiling.py pmc    FETCH.db WRITE.db COPY_FETCH.db COPY_WRITE.db \
                                               --copy-mib 1024 [--last N] [--nop NOP] [--json OUT] [--only ...]
    python alps-looper_amd/profiling.py counters RESULTS.db [RESULTS.db ...] [--last N]

`trace` prints per-kernel call counts and average durations (the top_kernels
view); `pmc` prints HBM bytes per dispatch from FETCH_SIZE and WRITE_SIZE,
each scaled by the factor that makes the plain copy kernel (k_copy_stream,
reads and writes COPY_MIB MiB per dispatch) read exactly its algorithmic
bytes.  `counters` prints the mean per dispatch of every counter in the
given databases (one counter-collection run each), kernel by kernel, and the
wait and LDS ratios when their counters are present.  --last N keeps the last
N dispatches of each kernel (the equilibrated sweeps at the end of a run).
"""

import argparse
import json
import re
import sqlite3
import sys


def short(name):
    m = re.search(r"k_([a-z_]+)", name)
    return m.group(1) if m else name


def trace(db, last=None):
    con = sqlite3.connect(db)
    rows = con.execute("select name, duration from kernels order by start").fetchall()
    per = {}
    for n, d in rows:
        per.setdefault(short(n), []).append(d)
    out = {}
    for k, v in per.items():
        v = v[-last:] if last else v
        out[k] = {"calls": len(v), "avg_us": sum(v) / len(v) / 1e3, "total_us": sum(v) / 1e3}
    return out


def pmc(db, counter, last=None):
    con = sqlite3.connect(db)
    rows = con.execute("select kernel_name, value from counters_collection where counter_name=? order by dispatch_id",
                       (counter,)).fetchall()
    per = {}
    for n, v in rows:
        per.setdefault(short(n), []).append(v)
    return {k: sum(v[-last:] if last else v) / len(v[-last:] if last else v) for k, v in per.items()}


def counters(dbs, last=None):
    """kernel -> counter -> mean value per dispatch, over the given databases"""
    out = {}
    for db in dbs:
        con = sqlite3.connect(db)
        names = [r[0] for r in con.execute("select distinct counter_name from counters_collection")]
        for c in names:
            for k, v in pmc(db, c, last).items():
                out.setdefault(k, {})[c] = v
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["trace", "pmc", "counters"])
    ap.add_argument("dbs", nargs="+")
    ap.add_argument("--last", type=int, default=None)
    ap.add_argument("--copy-mib", type=float, default=1024)
    ap.add_argument("--nop", type=float, default=None, help="operators per sweep (bytes per operator column)")
    ap.add_argument("--json", default=None)
    ap.add_argument("--only", default=None, help="comma-separated kernels (short names) to list")
    a = ap.parse_args()
    only = set(a.only.split(",")) if a.only else None
    if a.mode == "trace":
        t = trace(a.dbs[0], a.last)
        if only:
            t = {k: v for k, v in t.items() if k in only}
        tot = sum(x["avg_us"] for k, x in t.items() if k not in ("copy_stream",) and not k.startswith("_"))
        print("| kernel | calls | avg us | share of sweep |")
        print("|---|---|---|---|")
        for k, x in sorted(t.items(), key=lambda kv: -kv[1]["avg_us"]):
            if k.startswith("__amd"):
                continue
            print("| %s | %d | %.2f | %.1f%% |" % (k, x["calls"], x["avg_us"], 100 * x["avg_us"] / tot))
        print("| sum of sweep kernels | | %.2f | |" % tot)
        return
    if a.mode == "counters":
        C = counters(a.dbs, a.last)
        names = sorted({c for k in C for c in C[k]})
        ratios = [("wait/cycles", "SQ_WAIT_ANY", "SQ_WAVE_CYCLES"),
                  ("LDS bank conflict/LDS inst", "SQ_LDS_BANK_CONFLICT", "SQ_INSTS_LDS"),
                  ("L2 hit", "TCC_HIT_sum", None)]
        extra = [r for r in ratios if r[1] in names and (r[2] is None or r[2] in names)]
        print("| kernel | " + " | ".join(names + [r[0] for r in extra]) + " |")
        print("|---|" + "---|" * (len(names) + len(extra)))
        for k in sorted(C, key=lambda k: k):
            if k.startswith("__amd") or k == "copy_stream":
                continue
            row = ["%.4g" % C[k].get(c, float("nan")) for c in names]
            for _, num, den in extra:
                if den is None:  # hit rate
                    h, m = C[k].get("TCC_HIT_sum", 0), C[k].get("TCC_MISS_sum", 0)
                    row.append("%.2f" % (h / (h + m)) if h + m else "-")
                else:
                    d = C[k].get(den, 0)
                    row.append("%.2f" % (C[k].get(num, 0) / d) if d else "-")
            print("| %s | %s |" % (k, " | ".join(row)))
        return
    fetch, write, cfetch, cwrite = a.dbs[:4]
    F = pmc(fetch, "FETCH_SIZE", a.last)
    W = pmc(write, "WRITE_SIZE", a.last)
    cF = pmc(cfetch, "FETCH_SIZE")["copy_stream"]
    cW = pmc(cwrite, "WRITE_SIZE")["copy_stream"]
    copy_kib = a.copy_mib * 1024
    sf, sw = copy_kib / cF, copy_kib / cW
    print("copy kernel: %.0f MiB read + %.0f MiB written per dispatch; FETCH_SIZE %.0f KiB, WRITE_SIZE %.0f KiB"
          % (a.copy_mib, a.copy_mib, cF, cW))
    print("scale factors: fetch %.4f, write %.4f" % (sf, sw))
    print()
    hdr = "| kernel | fetch MB | write MB | HBM MB/launch |" + (" B/op |" if a.nop else "")
    print(hdr)
    print("|---|---|---|---|" + ("---|" if a.nop else ""))
    res = {}
    for k in sorted(F, key=lambda k: -(F[k] * sf + W.get(k, 0) * sw)):
        if k.startswith("__amd") or (only and k not in only):
            continue
        fb = F[k] * sf * 1024
        wb = W.get(k, 0) * sw * 1024
        res[k] = {"fetch_bytes": fb, "write_bytes": wb, "traffic_bytes": fb + wb}
        line = "| %s | %.2f | %.2f | %.2f |" % (k, fb / 1e6, wb / 1e6, (fb + wb) / 1e6)
        if a.nop:
            line += " %.1f |" % ((fb + wb) / a.nop)
            res[k]["bytes_per_op"] = (fb + wb) / a.nop
        print(line)
    if a.json:
        with open(a.json, "w") as f:
            json.dump({"scale_fetch": sf, "scale_write": sw, "kernels": res, "nop": a.nop}, f, indent=1)


if __name__ == "__main__":
    sys.exit(main())
