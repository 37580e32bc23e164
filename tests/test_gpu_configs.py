"""GPU parity at the benchmark configurations, through the C-ABI.

* square 128^2, beta=128 (the single-chain benchmark) at equilibrium: the
  device's configuration after 300 sweeps is moved into the oracle's
  counter-based (philox) mode, both run the next sweeps, and operator
  strings, spins, records and (on the equilibrated configuration) the
  canonical cluster labels of every leg are compared with looper's
  union-find restatement;
* the 512-chain temperature scan on 32^2 (tuned launch settings, HIP-graph
  replay): eight chains spread over the temperatures, bit-exact with the
  oracle's trajectories;
* the serving mix: jobs drawn from the serving trace (chain L=1024 and
  square L=96 among them) submitted with Poisson arrivals; every result
  equals the same job run alone on the plain path, and the plain path equals
  the oracle for the first sweeps;
* square 512^2, beta=512 (the large single-GPU share) after 200 sweeps: one
  sweep and the cluster labels against the oracle.

Reference: path_integral.C:455-864 (the sweep), looper/union_find.h:242-350
(labels), looper/graph_impl.h:277-295 (cut rules).
"""

import time

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

EXACT = [0, 1]                                # n_op, n_clusters
FLOAT = list(range(2, 18)) + [20, 21, 22]     # estimator sums, FP64 atomics in another order
RTOL = 1e-9


def _cmp(gpu, cpu, where):
    np.testing.assert_array_equal(gpu[:, EXACT], cpu[:, EXACT], err_msg="n_op/n_clusters " + where)
    np.testing.assert_allclose(gpu[:, FLOAT], cpu[:, FLOAT], rtol=RTOL, atol=1e-9, err_msg=where)
    np.testing.assert_array_equal(gpu[:, 19], cpu[:, 19], err_msg="sweep index " + where)


def _cmp_config(a, b, where):
    for x, y, name in zip(a, b, ("spins", "bond offsets", "times", "types")):
        np.testing.assert_array_equal(x, y, err_msg="%s %s" % (name, where))


def test_c2_equilibrium_matches_oracle(loopgpu):
    """square 128^2 at beta=128 exactly, equilibrated: both root modes"""
    L, beta, eq, n = 128, 128.0, 300, 2
    p = loopgpu.params(O.SQUARE, L, beta, 1.0, 1.0, seed=23125, chain_id=0)
    print("c2: equilibrate %d sweeps" % eq, flush=True)
    g = loopgpu.Chain(p)
    r = g.sweep(eq)
    assert r[-1, 0] > 2.3e6                          # equilibrium density (about 2.45M operators)
    cfg = g.get_config()
    labels, _, _, _ = loopgpu.cluster_config(p, *cfg, eq)
    print("c2: looper union-find labels of %d elements" % len(labels), flush=True)
    np.testing.assert_array_equal(labels, O.looper_labels(O.SQUARE, L, cfg[1], cfg[2], cfg[3]))
    c = O.PhiloxChain(O.SQUARE, L, beta, 1.0, 1.0, seed=23125, chain_id=0)
    c.load(*cfg, eq)
    rc = c.sweep(n)
    for comp in (0, 1):
        print("c2: %d sweeps from the equilibrated configuration, compress %d" % (n, comp), flush=True)
        h = loopgpu.Chain(p)
        h.set_launch(compress=comp)
        h.set_config(*cfg, sweep=eq)
        _cmp(h.sweep(n), rc, "c2 compress %d" % comp)
        _cmp_config(h.get_config(), c.state(), "c2 compress %d" % comp)
        h.close()


def _scan_params(lg):
    import bench
    return bench.scan_params(lg), bench.scan_temps()


def test_c3_scan_tuned_graph_matches_oracle(loopgpu):
    """512 chains on 32^2, T_i = 0.05 * 40^(i/511): autotuned settings and the
    captured sweep; eight chains across the temperatures against the oracle"""
    plist, temps = _scan_params(loopgpu)
    print("c3: 512-chain batch, tune + graph", flush=True)
    b = loopgpu.Batch(plist)
    b.tune(4)
    b.set_graph(True)
    _, s0 = b.chain_info(0)
    n = 30
    rb = b.run(n)
    assert all(b.chain_info(i)[1] == s0 + n for i in (0, 511))
    for i in (0, 73, 146, 219, 292, 365, 438, 511):
        print("c3: chain %d (T=%.4f) against the oracle, %d + %d sweeps" % (i, temps[i], s0, n), flush=True)
        c = O.PhiloxChain(O.SQUARE, 32, 1.0 / temps[i], 1.0, 1.0, seed=23125, chain_id=i)
        c.sweep(s0)
        _cmp(rb[:, i, :], c.sweep(n), "c3 chain %d" % i)
    # cold chains carry tens of thousands of operators, hot ones a few hundred
    assert rb[-1, 0, 0] > 30 * rb[-1, 511, 0]


def _pick_serving_jobs():
    import bench
    trace = bench.serve_trace(400, sweep_scale=0.05)
    want = [("chain", 1024), ("square", 96), ("chain", 64), ("square", 16), ("chain", 256), ("square", 32),
            ("square", 48), ("square", 64), ("chain", 1024), ("square", 96), ("chain", 64), ("square", 16)]
    jobs, used = [], set()
    for kind in want:
        for j in trace:
            if (j["kind"], j["L"]) == kind and j["chain_id"] not in used:
                jobs.append(j)
                used.add(j["chain_id"])
                break
    return jobs


def _binned_mean(xs):
    s = 0.0
    for x in xs:
        s += x
    return s / len(xs)


def test_c4_serving_mix_poisson_equals_plain_and_oracle(loopgpu):
    """jobs of the serving trace (sweep counts scaled by 0.05) arriving as a
    Poisson stream; each result equals its plain-chain run"""
    import bench
    lg = loopgpu
    jobs = _pick_serving_jobs()
    assert len(jobs) == 12
    assert {(j["kind"], j["L"]) for j in jobs} >= {("chain", 1024), ("square", 96)}
    rng = np.random.default_rng(99)
    srv = lg.Server(**dict(bench.SERVER_CFG, max_chunk=16))
    ids = []
    print("c4: %d jobs with Poisson arrivals" % len(jobs), flush=True)
    for j in jobs:
        ids.append(srv.submit(bench.job_params(lg, j), j["sweeps"], j["thermalization"]))
        time.sleep(float(rng.exponential(0.05)))
    res = [srv.wait(i, 300) for i in ids]
    st = srv.stats()
    srv.stop()
    assert all(r["status"] == lg.JOB_DONE for r in res)
    assert st["done"] == 12 and st["p99_latency_s"] >= st["p50_latency_s"] > 0
    for j, r in zip(jobs, res):
        print("c4: job %d (%s L=%d beta=%.1f) alone on the plain path" % (j["chain_id"], j["kind"], j["L"], j["beta"]),
              flush=True)
        ch = lg.Chain(bench.job_params(lg, j))
        rec = ch.sweep(j["thermalization"] + j["sweeps"])[j["thermalization"]:]
        ch.close()
        V = j["L"] if j["kind"] == "chain" else j["L"] ** 2
        assert r["sweeps_done"] == j["thermalization"] + j["sweeps"]
        assert r["mean"]["Energy Density"] == _binned_mean(list(rec[:, 2] / V))
        assert r["mean"]["Number of Clusters"] == _binned_mean(list(rec[:, 1]))
        assert r["mean"]["Operators"] == _binned_mean(list(rec[:, 0]))
        beta = j["beta"]
        for name, col in (("Susceptibility", beta * rec[:, 9] / V), ("Staggered Structure Factor", rec[:, 12] / V),
                          ("Staggered Susceptibility", beta * rec[:, 16] / V)):
            assert r["mean"][name] == pytest.approx(_binned_mean(list(col)), rel=1e-9, abs=1e-12), name
    for j in jobs:
        if (j["kind"], j["L"]) in (("chain", 64), ("square", 16)):
            print("c4: job %d against the oracle" % j["chain_id"], flush=True)
            kind = O.CHAIN if j["kind"] == "chain" else O.SQUARE
            g = lg.Chain(bench.job_params(lg, j))
            c = O.PhiloxChain(kind, j["L"], j["beta"], 1.0, j["jz"], seed=bench.SEED, chain_id=j["chain_id"])
            _cmp(g.sweep(25), c.sweep(25), "c4 job %d" % j["chain_id"])
            _cmp_config(g.get_config(), c.state(), "c4 job %d" % j["chain_id"])


@pytest.mark.timeout(1200)
def test_c5_equilibrated_sweep_and_labels_match_oracle(loopgpu):
    """square 512^2, beta=512 after 200 sweeps: about 157M operators and 320M
    segments (the 1536-event linking kernel, 16-item estimator chunks, the
    largest grids); one sweep against the oracle, then the labels"""
    L, beta, eq = 512, 512.0, 200
    p = loopgpu.params(O.SQUARE, L, beta, 1.0, 1.0, seed=23125, chain_id=0)
    print("c5: equilibrate %d sweeps" % eq, flush=True)
    g = loopgpu.Chain(p)
    assert g.launch()["link"] == 1536
    for _ in range(eq // 50):
        r = g.sweep(50)
        print("c5: n_op %.0f" % r[-1, 0], flush=True)
    assert r[-1, 0] > 1.45e8
    cfg = g.get_config()
    rg = g.sweep(1)
    after = g.get_config()
    g.close()
    print("c5: oracle sweep from the device's configuration (%d operators)" % len(cfg[2]), flush=True)
    c = O.PhiloxChain(O.SQUARE, L, beta, 1.0, 1.0, seed=23125, chain_id=0)
    c.load(*cfg, eq)
    rc = c.sweep(1)
    _cmp(rg, rc, "c5")
    _cmp_config(after, c.state(), "c5")
    del c, after
    print("c5: cluster labels on the device", flush=True)
    labels, _, _, _ = loopgpu.cluster_config(p, *cfg, eq)
    print("c5: looper union-find labels of %d elements" % len(labels), flush=True)
    np.testing.assert_array_equal(labels, O.looper_labels(O.SQUARE, L, cfg[1], cfg[2], cfg[3]))
