"""Extract golden vectors from the reference's own test outputs into small JSON
fixtures (run once where the reference checkout is present; the tests read
only the fixtures, never the reference).

Sources (data files of the reference's tests):
  test/union_find.op            100 random unions and the resulting forest
  test/compatibility.op         cut-rule truth table
  test/weight.op                XXZ graph weights (standard and FORCE_SCATTER=0.1)
  test/spinmodel_helper.op      chain lattice bond list (ALPS conventions)
  test/virtual_lattice.op       square lattice L=2 vertices, edges, gauge
  test/temperature.ip/.op       annealing schedules: parameter sets and T(step)
  test/permutation.op           random_shuffle and partitioned_random_shuffle output
                                (mt19937 seed 23094)
  test/virtual_lattice.op       the virtual lattice of a square L=2 lattice with S=3/2 on
                                site 0 and S=1 elsewhere (virtual site and bond order)
  standalone/loop.op            standalone kernel output (L=8, T=0.2)
  loop.op                       end-to-end run: ED and path-integral QMC, chain L=4, T=0.25;
                                and the path-integral run with MEASURE[Correlations] and
                                MEASURE[Structure Factor] (chain L=4, S=1/2, T=0.5): its
                                scalar results and vector entries ("Name[label]"); the ED and
                                path-integral blocks of its other models (S=1 chain L=4, one S=1
                                site in a transverse field Gamma=1 or a field h=0.3, Ising chain,
                                S=1 Ising chain with a single-ion term D=-0.2); its "loop; sse"
                                blocks of the same models
"""

import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))


def read(rel):
    with open(os.path.join(REF, rel)) as f:
        return f.read().splitlines()


def union_find():
    lines = read("test/union_find.op")
    pairs = []
    for l in lines:
        m = re.match(r"connecting node (\d+) to node (\d+)", l)
        if m:
            pairs.append([int(m.group(1)), int(m.group(2))])
    nodes = {}
    start = lines.index("[results]") + 1
    for l in lines[start:start + 100]:
        m = re.match(r"node (\d+) is root and tree size is (\d+)", l)
        if m:
            nodes[int(m.group(1))] = {"parent": -1, "root": int(m.group(1)), "size": int(m.group(2))}
            continue
        m = re.match(r"node (\d+)'s parent is (\d+) and its root is (\d+)", l)
        nodes[int(m.group(1))] = {"parent": int(m.group(2)), "root": int(m.group(3)), "size": 0}
    return {"n": 100, "pairs": pairs,
            "parent": [nodes[i]["parent"] for i in range(100)],
            "root": [nodes[i]["root"] for i in range(100)],
            "size": [nodes[i]["size"] for i in range(100)]}


def compatibility():
    out = []
    for l in read("test/compatibility.op"):
        if l.startswith("[bond graph (general location + general graph)]"):
            collecting = True
            continue
        m = re.match(r"Delta\(g=(\d), c0=([+-]), c1=([+-])\) = (\d)", l)
        if m and len(out) < 16:
            out.append([int(m.group(1)), 0 if m.group(2) == "+" else 1, 0 if m.group(3) == "+" else 1,
                        int(m.group(4))])
    return out


def weights():
    rows = []
    for l in read("test/weight.op"):
        m = re.match(r"bond weight \((standard|ergodic)\): C = (\S+), Jxy = (\S+), Jz = (\S+) : "
                     r"v\[0\] = (\S+), v\[1\] = (\S+), v\[2\] = (\S+), v\[3\] = (\S+), offset = (\S+), sign = (\S+)", l)
        if m:
            g = m.groups()
            rows.append({"force_scatter": 0.0 if g[0] == "standard" else 0.1,
                         "C": float(g[1].rstrip(",")), "jxy": float(g[2].rstrip(",")), "jz": float(g[3]),
                         "v": [float(g[4].rstrip(",")), float(g[5].rstrip(",")), float(g[6].rstrip(",")),
                               float(g[7].rstrip(","))],
                         "offset": float(g[8].rstrip(",")), "sign": int(g[9])})
    return rows


def lattices():
    sm = read("test/spinmodel_helper.op")
    chain = []
    for l in sm:
        m = re.match(r'\s*<EDGE source="(\d+)" target="(\d+)" id="(\d+)"', l)
        if m:
            chain.append([int(m.group(1)) - 1, int(m.group(2)) - 1])
        if "</GRAPH>" in l:
            break
    vl = read("test/virtual_lattice.op")
    verts, edges = [], []
    gauge = None
    for l in vl:
        m = re.match(r'\s*<VERTEX id="(\d+)" type="\d+"><COORDINATE>(\d+) (\d+)</COORDINATE>', l)
        if m:
            verts.append([int(m.group(2)), int(m.group(3))])
        m = re.match(r'\s*<EDGE source="(\d+)" target="(\d+)" id="(\d+)" type="\d+" vector="(\d+) (\d+)"', l)
        if m:
            edges.append([int(m.group(1)) - 1, int(m.group(2)) - 1, int(m.group(4)), int(m.group(5))])
        if "</GRAPH>" in l and gauge is None:
            idx = vl.index(l)
            gauge = [int(x) for x in vl[idx + 1].split()]
            break
    return {"chain4_bonds": chain, "square2_coords": verts, "square2_edges": edges, "square2_gauge": gauge}


def standalone():
    res = {}
    for l in read("standalone/loop.op"):
        m = re.match(r"(.+?)\s+= (\S+)(?: \+- (\S+))?$", l)
        if m:
            res[m.group(1).strip()] = [float(m.group(2))] + ([float(m.group(3))] if m.group(3) else [])
    return res


def loop_op_blocks():
    """every block of loop.op: its input parameters and its results"""
    lines = read("loop.op")
    blocks = []
    cur = None
    for l in lines:
        if l.startswith("[input parameters]"):
            cur = {"params": {}, "results": {}}
            blocks.append(cur)
            continue
        m = re.match(r"(\S[^=]*?) = (.*);$", l)
        if cur is not None and m and not cur["results"]:
            cur["params"][m.group(1).strip()] = m.group(2).strip().strip('"')
            continue
        m = re.match(r"(.+?): (\S+) \+/- (\S+)", l)
        if cur is not None and m:
            err = m.group(3).rstrip(";")
            cur["results"][m.group(1)] = [float(m.group(2)), float("inf") if err == "inf" else float(err)]
    return blocks


def keep_blocks(blocks):
    return [{"algorithm": b["params"].get("ALGORITHM"), "params": b["params"],
             "results": {k: v for k, v in b["results"].items() if not k.startswith("Local")}} for b in blocks]


def loop_op():
    return keep_blocks(loop_op_blocks()[:2])  # L=4 chain, T=0.25: diagonalization and "loop; path integral"


def loop_op_models():
    """the diagonalization and "loop; path integral" blocks of the other
    models of loop.ip, keyed by model: S=1 chain L=4 (T=1/L), one S=1 site
    in a transverse field Gamma=1 (T=1/L), one S=1 site in a longitudinal
    field h=0.3 (T=1/L), S=1/2 Ising chain L=4 (Jxy=0, Jz=1, T=0.5), S=1
    Ising chain L=4 with a single-ion term (Jxy=0, Jz=-1, D=-0.2, T=2)"""
    out = {}
    for b in loop_op_blocks():
        p = b["params"]
        if p.get("ALGORITHM") not in ("diagonalization", "loop; path integral") or "MEASURE[Correlations]" in p:
            continue
        if p.get("LATTICE") == "site":
            key = "site_spin1_gamma1" if p.get("Gamma") == "1" else "site_spin1_h03"
        elif p.get("local_S") == "1" and "D" not in p:
            key = "chain4_spin1"
        elif p.get("local_S") == "1" and p.get("D") == "-0.2":
            key = "chain4_spin1_d"
        elif p.get("Jxy") == "0" and p.get("local_S") == "1/2":
            key = "chain4_ising"
        else:
            continue
        out.setdefault(key, []).extend(keep_blocks([b]))
    assert sorted(out) == ["chain4_ising", "chain4_spin1", "chain4_spin1_d", "site_spin1_gamma1", "site_spin1_h03"]
    assert all(len(v) == 2 for v in out.values())
    return out


def loop_op_sse():
    """the "loop; sse" blocks of loop.op, keyed by model: S=1/2 chain L=4
    (T=0.25), S=1 chain L=4 (T=1/L), S=1/2 Ising chain L=4 (Jxy=0, Jz=1,
    T=0.5), one S=1 site in Gamma=1 (T=1/L), S=1 Ising chain with D=-0.2
    (Jz=-1, T=2)"""
    out = {}
    for b in loop_op_blocks():
        p = b["params"]
        if p.get("ALGORITHM") != "loop; sse":
            continue
        if p.get("LATTICE") == "site":
            key = "site_spin1_gamma1"
        elif "D" in p:
            key = "chain4_spin1_d"
        else:
            key = ("chain4_spin1" if p.get("local_S") == "1" else
                   "chain4_ising" if p.get("Jxy") == "0" else "chain4")
        out[key] = keep_blocks([b])[0]
    assert sorted(out) == ["chain4", "chain4_ising", "chain4_spin1", "chain4_spin1_d", "site_spin1_gamma1"]
    return out


def loop_op_correlations():
    """the "loop; path integral" block of loop.op with MEASURE[Correlations]
    and local_S = 1/2: scalar results and vector entries keyed Name[label]"""
    lines = read("loop.op")
    blocks = []
    cur = None
    vec = None
    for l in lines:
        if l.startswith("[input parameters]"):
            cur = {"params": {}, "results": {}}
            blocks.append(cur)
            vec = None
            continue
        if cur is None:
            continue
        m = re.match(r"(\S[^=]*?) = (.*);$", l)
        if m and not cur["results"] and vec is None:
            cur["params"][m.group(1).strip()] = m.group(2).strip().strip('"')
            continue
        m = re.match(r"Entry\[(.+)\]: (\S+) \+/- (\S+?);", l)
        if m and vec is not None:
            cur["results"]["%s[%s]" % (vec, m.group(1))] = [float(m.group(2)), float(m.group(3))]
            continue
        if l and not l.startswith(" ") and ":" not in l and "=" not in l:
            vec = l.strip()  # name line of a vector observable
            continue
        m = re.match(r"(.+?): (\S+) \+/- (\S+)", l)
        if m and not l.startswith(" "):
            vec = None
            err = m.group(3).rstrip(";")
            cur["results"][m.group(1)] = [float(m.group(2)), float("inf") if err == "inf" else float(err)]
    for b in blocks:
        p = b["params"]
        if (p.get("ALGORITHM") == "loop; path integral" and p.get("MEASURE[Correlations]") == "true"
                and p.get("local_S") == "1/2"):
            return {"params": p, "results": b["results"]}
    raise RuntimeError("correlation block not found in loop.op")


def temperature():
    """parameter sets of test/temperature.ip (globals apply to the following
    {...} blocks, "#clear;" drops them) with the initial / final temperature,
    annealing steps and T(step) lines test/temperature.op prints for each"""
    sets, glob, cur = [], {}, None
    for raw in read("test/temperature.ip"):
        line = raw.strip()
        if line.startswith("#clear"):
            glob = {}
            continue
        if not line:
            continue
        for tok in re.split(r"(\{|\})", line):
            tok = tok.strip()
            if tok == "{":
                cur = dict(glob)
            elif tok == "}":
                sets.append(cur)
                cur = None
            else:
                for stmt in tok.split(";"):
                    if "=" in stmt:
                        k, v = (x.strip() for x in stmt.split("=", 1))
                        (cur if cur is not None else glob)[k] = v
    out, blk = [], None
    for l in read("test/temperature.op"):
        if l.startswith("initial temperature"):
            blk = {"initial": float(l.split("=")[1]), "T": []}
            out.append(blk)
        elif l.startswith("final temperature"):
            blk["final"] = float(l.split("=")[1])
        elif l.startswith("annealing steps"):
            blk["annealing_steps"] = int(l.split("=")[1])
        elif l.strip():
            a, b = l.split()
            blk["T"].append([int(a), float(b)])
    assert len(sets) == len(out)
    return [dict(o, params=p) for p, o in zip(sets, out)]


def permutation():
    """test/permutation.op: 20 shuffles of 0..9, then 5 partitioned shuffles
    with their guides (conf0 at tau = 0, conf1 at tau = 1)"""
    lines = read("test/permutation.op")
    shuffles, parts, cur = [], [], {}
    for l in lines:
        f = l.split()
        if not f:
            continue
        if f[0] in ("conf0", "conf1", "perm"):
            cur[f[0]] = [int(x) for x in f[1:]]
            if f[0] == "perm":
                parts.append(cur)
                cur = {}
        elif all(x.isdigit() for x in f) and len(f) == 10:
            shuffles.append([int(x) for x in f])
    assert len(shuffles) == 20 and len(parts) == 5
    return {"seed": 23094, "n": 10, "shuffles": shuffles, "partitioned": parts}


def virtual_lattice():
    """the first virtual graph of test/virtual_lattice.op (square L=2, site 0
    of type 1 with S=3/2, the other sites S=1): its real edges, the virtual
    edges in order, the virtual sites of each real site and the gauge"""
    lines = read("test/virtual_lattice.op")
    real, virt, gauges, site_map = [], [], [], []
    block = 0
    for l in lines:
        m = re.match(r'\s*<EDGE source="(\d+)" target="(\d+)"', l)
        if l.startswith("<GRAPH"):
            block += 1
        elif m and block == 1:
            real.append([int(m.group(1)) - 1, int(m.group(2)) - 1])
        elif m and block == 2:
            virt.append([int(m.group(1)) - 1, int(m.group(2)) - 1])
        elif block == 2 and re.match(r"^-?1( -?1)+\s*$", l) and not gauges:
            gauges = [int(x) for x in l.split()]
        mm = re.match(r"\s*(\d+) -> \[(\d+),(\d+)\]", l)
        if mm and block == 2 and len(site_map) < 4:
            site_map.append([int(mm.group(2)), int(mm.group(3))])
    assert len(real) == 8 and len(virt) == 40 and len(gauges) == 9
    return {"real_edges": real, "two_s": [3, 2, 2, 2], "virtual_edges": virt, "virtual_gauge": gauges,
            "site_map": site_map}


def main():
    fx = {
        "permutation": permutation(),
        "virtual_lattice": virtual_lattice(),
        "temperature": temperature(),
        "union_find": union_find(),
        "compatibility": compatibility(),
        "weights": weights(),
        "lattices": lattices(),
        "standalone_loop_op": standalone(),
        "loop_op_chain4": loop_op(),
        "loop_op_chain4_correlations": loop_op_correlations(),
        "loop_op_models": loop_op_models(),
        "loop_op_sse": loop_op_sse(),
    }
    with open(os.path.join(OUT, "reference_fixtures.json"), "w") as f:
        json.dump(fx, f, indent=1, sort_keys=True)
    print("wrote", os.path.join(OUT, "reference_fixtures.json"))


if __name__ == "__main__":
    main()
