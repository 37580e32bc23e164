"""The "loop; sse" worker (sse.C:174-400): the same loop clusters on the
stochastic series expansion.

CPU: the oracle's looper mode restating sse.C (the gap-by-gap diagonal
update, choose_diagonal / choose_offdiagonal, positions as times) against
exact diagonalisation and against the reference's own "loop; sse" blocks of
loop.op; the oracle's philox mode (the device's formulation: one Philox word
set per step, times p / n) against ED; parameter checks.  GPU: SSE chains
bit-exact with the philox mode (records, operator strings as bond-major
configurations, spins) under every linking kernel and root mode, batched,
resumed from a checkpoint, and the worker with ALGORITHM = "loop; sse"
against loop.op and ED."""

import json
import os

import numpy as np
import pytest

import ed
import oracle_lib as O
import stats

HERE = os.path.dirname(os.path.abspath(__file__))
FXALL = json.load(open(os.path.join(HERE, "golden", "reference_fixtures.json")))
FX = FXALL["loop_op_sse"]
NAMES = ["Energy", "Susceptibility", "Magnetization^2", "Staggered Magnetization^2", "Staggered Susceptibility"]

# kind, L, 2S, T, Jxy, Jz
MODELS = [(O.CHAIN, 8, 1, 0.5, 1.0, 1.0), (O.CHAIN, 4, 1, 0.3, 0.6, 1.2), (O.CHAIN, 4, 1, 0.5, 0.0, 1.0),
          (O.SQUARE, 2, 1, 0.7, 1.0, 1.0), (O.CHAIN, 4, 2, 0.5, 1.0, 1.0), (O.CHAIN, 4, 3, 0.8, 1.0, 1.0),
          (O.SQUARE, 2, 3, 1.5, 1.0, 1.0), (O.CHAIN, 4, 6, 2.5, 1.0, 0.8)]


def _ed(kind, L, two_s, T, jxy, jz):
    return ed.diagonalize_transverse("chain" if kind == O.CHAIN else "square", L, T, two_s, jxy, jz)


@pytest.mark.parametrize("kind,L,two_s,T,jxy,jz", MODELS)
def test_oracle_sse_modes_match_ed(kind, L, two_s, T, jxy, jz):
    ref = _ed(kind, L, two_s, T, jxy, jz)
    beta = 1.0 / T
    n = L if kind == O.CHAIN else L * L
    for name, ch in (("looper", O.LooperChain(kind, L, beta, jxy, jz, seed=3, two_s=two_s)),
                     ("philox", O.PhiloxChain(kind, L, beta, jxy, jz, seed=3, chain_id=1, two_s=two_s))):
        ch.sweep_sse(2000)
        obs = stats.observables_sse(ch.sweep_sse(40000), beta, n, 1 if kind == O.CHAIN else 2)
        for k in NAMES:
            m, e = stats.mean_err(obs[k])
            assert stats.sigmas(m, e, ref[k]) < 4.0, (name, k, m, e, ref[k])


@pytest.mark.parametrize("key,two_s,T,jxy,jz", [("chain4", 1, 0.25, 1.0, 1.0), ("chain4_spin1", 2, 0.25, 1.0, 1.0),
                                                ("chain4_ising", 1, 0.5, 0.0, 1.0)])
def test_oracle_sse_matches_reference_sse_blocks(key, two_s, T, jxy, jz):
    """loop.op's "loop; sse" runs (4096 sweeps each) against the looper-mode
    restatement of the same worker"""
    qmc = FX[key]["results"]
    beta = 1.0 / T
    ch = O.LooperChain(O.CHAIN, 4, beta, jxy, jz, seed=23125, two_s=two_s)
    ch.sweep_sse(2000)
    obs = stats.observables_sse(ch.sweep_sse(40000), beta, 4, 1)
    for k in NAMES:
        if k not in qmc:
            continue
        m, e = stats.mean_err(obs[k])
        assert stats.sigmas(m, e, qmc[k][0], qmc[k][1]) < 4.0, (k, m, e, qmc[k])


# L, 2S, T, Jxy, Jz, hx, D, FORCE_SCATTER
EXTRAS = [(4, 1, 0.5, 0.0, 1.0, 0.8, 0.0, 0.0), (4, 1, 0.5, -1.0, 1.0, 0.8, 0.0, 0.0),
          (4, 2, 2.0, 0.0, -1.0, 0.0, -0.2, 0.0), (4, 2, 0.6, -1.0, 1.0, 0.7, 0.5, 0.1),
          (4, 3, 0.8, 1.0, 1.0, 0.0, -0.4, 0.0)]


@pytest.mark.parametrize("L,two_s,T,jxy,jz,hx,d,fs", EXTRAS)
def test_oracle_sse_site_and_insite_operators_match_ed(L, two_s, T, jxy, jz, hx, d, fs):
    """the SSE walk with the transverse field's site operators and the
    single-ion term's in-site bonds"""
    ref = ed.diagonalize_transverse("chain", L, T, two_s, jxy, jz, hx, 0.0, d)
    beta = 1.0 / T
    for name, ch in (("looper", O.LooperChain(O.CHAIN, L, beta, jxy, jz, seed=5, fs=fs, two_s=two_s, hx=hx, d=d)),
                     ("philox", O.PhiloxChain(O.CHAIN, L, beta, jxy, jz, seed=5, chain_id=2, fs=fs, two_s=two_s,
                                              hx=hx, d=d))):
        ch.sweep_sse(2000)
        obs = stats.observables_sse(ch.sweep_sse(40000), beta, L, 1)
        for k in ("Energy", "Susceptibility", "Magnetization^2", "Staggered Magnetization^2"):
            m, e = stats.mean_err(obs[k])
            assert stats.sigmas(m, e, ref[k]) < 4.0, (name, k, m, e, ref[k])


def test_oracle_sse_matches_reference_gamma_and_d_blocks():
    """loop.op's "loop; sse" runs of one S=1 site in Gamma=1 (two uncoupled
    sites here: per-site values, half the magnetization squares) and of the
    S=1 Ising chain with D=-0.2"""
    site = FX["site_spin1_gamma1"]["results"]
    ch = O.LooperChain(O.CHAIN, 2, 4.0, 0.0, 0.0, seed=23125, two_s=2, hx=1.0)
    ch.sweep_sse(2000)
    obs = stats.observables_sse(ch.sweep_sse(40000), 4.0, 2, 1)
    for k, scale in (("Energy Density", 1.0), ("Susceptibility", 1.0), ("Magnetization^2", 0.5)):
        m, e = stats.mean_err(obs[k] * scale)
        assert stats.sigmas(m, e, site[k][0], site[k][1]) < 4.0, (k, m, e, site[k])
    dch = FX["chain4_spin1_d"]["results"]
    ch = O.LooperChain(O.CHAIN, 4, 0.5, 0.0, -1.0, seed=23125, two_s=2, d=-0.2)
    ch.sweep_sse(2000)
    obs = stats.observables_sse(ch.sweep_sse(40000), 0.5, 4, 1)
    for k in ("Energy", "Susceptibility", "Magnetization^2", "Staggered Magnetization^2"):
        m, e = stats.mean_err(obs[k])
        assert stats.sigmas(m, e, dch[k][0], dch[k][1]) < 4.0, (k, m, e, dch[k])


def test_oracle_sse_normal_estimator_matches_ed():
    """the normal estimator's SSE form beta (n M_a^2 + M_0^2) / (n + 1) / V
    (susceptibility.h:326-333) from both oracle modes"""
    ref = ed.diagonalize_transverse("chain", 6, 0.5, 1)
    for ch in (O.LooperChain(O.CHAIN, 6, 2.0, 1.0, 1.0, seed=4), O.PhiloxChain(O.CHAIN, 6, 2.0, 1.0, 1.0, seed=4,
                                                                               chain_id=1)):
        ch.sweep_sse(2000)
        r = ch.sweep_sse(40000)
        n = r[:, 0]
        m, e = stats.mean_err(2.0 * (n * r[:, 20] ** 2 + r[:, 3] ** 2) / (n + 1) / 6)
        assert stats.sigmas(m, e, ref["Susceptibility"]) < 4.0, (m, e, ref["Susceptibility"])


def test_sse_parameter_checks():
    import loopgpu
    with pytest.raises(loopgpu.LoopGPUError) as e:
        loopgpu.Chain(loopgpu.params(loopgpu.CHAIN, 8, 1.0, h=0.2, algorithm=loopgpu.SSE))
    assert e.value.code == -1 and "longitudinal field" in str(e.value)
    for p in (loopgpu.params(loopgpu.SQUARE, 64, 1.0, algorithm=loopgpu.SSE),
              loopgpu.params(loopgpu.CHAIN, 2050, 1.0, -1.0, 1.0, hx=0.5, algorithm=loopgpu.SSE)):
        with pytest.raises(loopgpu.LoopGPUError) as e:   # (site bonds count against the bond limit)
            loopgpu.Chain(p)
        assert "at most 4096" in str(e.value)
    with pytest.raises(loopgpu.LoopGPUError) as e:
        loopgpu.Batch([loopgpu.params(loopgpu.CHAIN, 8, 1.0), loopgpu.params(loopgpu.CHAIN, 8, 1.0, algorithm=1)])
    assert "one algorithm" in str(e.value)
    with pytest.raises(loopgpu.LoopGPUError) as e:
        loopgpu.Chain(loopgpu.params(loopgpu.CHAIN, 8, 1.0, algorithm=2))
    assert e.value.code == -1
    for txt, frag in (('ALGORITHM = "loop; sse"\nL = 8\nT = 1\nh = 0.1', "longitudinal field"),
                      ('ALGORITHM = "loop; worm"\nL = 8\nT = 1', "unsupported ALGORITHM"),
                      ('ALGORITHM = "loop; sse"\nL = 8\nT = 1\nMEASURE[Correlations] = true', "path-integral worker")):
        with pytest.raises(loopgpu.LoopGPUError) as e:
            loopgpu.Worker(txt)
        assert frag in str(e.value)


EXACT = [0, 1]
FLOAT = list(range(2, 18)) + [20, 21, 22]


def _cmp(g, c, where):
    np.testing.assert_array_equal(g[:, EXACT], c[:, EXACT], err_msg=where)
    np.testing.assert_allclose(g[:, FLOAT], c[:, FLOAT], rtol=1e-9, atol=1e-9, err_msg=where)


# kind, L, 2S, beta, Jxy, Jz (, hx, D, FORCE_SCATTER)
TRAJ = [(O.CHAIN, 4, 1, 2.0, 1.0, 1.0), (O.CHAIN, 64, 1, 10.0, 1.0, 1.0), (O.CHAIN, 16, 1, 4.0, 0.0, 1.0),
        (O.CHAIN, 32, 1, 6.0, -1.0, 0.5), (O.SQUARE, 4, 1, 3.0, 1.0, 1.0), (O.SQUARE, 16, 1, 4.0, 1.0, 1.0),
        (O.CHAIN, 8, 2, 4.0, 1.0, 1.0), (O.CHAIN, 8, 3, 3.0, 1.0, 0.8), (O.SQUARE, 4, 2, 2.0, 1.0, 1.0),
        (O.CHAIN, 8, 7, 1.5, 1.0, 1.0), (O.SQUARE, 4, 4, 1.0, 1.0, 1.0)]
TRAJ_EXTRA = [(O.CHAIN, 16, 1, 4.0, 0.0, 1.0, 0.9, 0.0, 0.0), (O.SQUARE, 4, 1, 3.0, -1.0, 1.0, 0.6, 0.0, 0.0),
              (O.CHAIN, 8, 2, 3.0, 1.0, 1.0, 0.0, -0.4, 0.0), (O.CHAIN, 8, 3, 2.0, -1.0, 0.8, 0.7, 0.5, 0.1),
              (O.CHAIN, 2, 2, 4.0, 0.0, 0.0, 1.0, 0.0, 0.0)]
LAUNCHES = [dict(), dict(link=0), dict(link=0, compress=1), dict(link=1), dict(link=128), dict(link=512),
            dict(link=1536, compress=1)]


def _launches(kind, two_s):
    lists = two_s * (2 if kind == O.CHAIN else 4)
    out = []
    for lc in LAUNCHES:
        if lc.get("link", 0) > 1 and two_s > 1:
            continue
        if lc.get("link", 0) == 1 and lists > 4:
            continue
        out.append(lc)
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("kind,L,two_s,beta,jxy,jz", TRAJ)
def test_sse_bit_exact_with_oracle(loopgpu, kind, L, two_s, beta, jxy, jz):
    c = O.PhiloxChain(kind, L, beta, jxy, jz, seed=5, chain_id=3, two_s=two_s)
    rc = c.sweep_sse(40)
    cs = c.state()
    for lc in _launches(kind, two_s):
        print("sse", kind, L, two_s, beta, jxy, jz, lc, flush=True)
        g = loopgpu.Chain(loopgpu.params(kind, L, beta, jxy, jz, seed=5, chain_id=3, two_s=two_s,
                                         algorithm=loopgpu.SSE))
        if lc:
            g.set_launch(**lc)
        _cmp(g.sweep(40), rc, "%s %s" % ((kind, L, two_s, beta, jxy, jz), lc))
        for a, b in zip(g.get_config(), cs):
            np.testing.assert_array_equal(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("kind,L,two_s,beta,jxy,jz,hx,d,fs", TRAJ_EXTRA)
def test_sse_site_and_insite_bit_exact_with_oracle(loopgpu, kind, L, two_s, beta, jxy, jz, hx, d, fs):
    c = O.PhiloxChain(kind, L, beta, jxy, jz, seed=6, chain_id=2, fs=fs, two_s=two_s, hx=hx, d=d)
    rc = c.sweep_sse(40)
    cs = c.state()
    for lc in (dict(), dict(compress=1)):
        print("sse extra", kind, L, two_s, beta, jxy, jz, hx, d, fs, lc, flush=True)
        g = loopgpu.Chain(loopgpu.params(kind, L, beta, jxy, jz, seed=6, chain_id=2, force_scatter=fs, two_s=two_s,
                                         hx=hx, d=d, algorithm=loopgpu.SSE))
        if lc:
            g.set_launch(**lc)
        assert g.launch()["link"] == 0
        _cmp(g.sweep(40), rc, "%s %s" % ((kind, L, two_s, beta, jxy, jz, hx, d), lc))
        for a, b in zip(g.get_config(), cs):
            np.testing.assert_array_equal(a, b)


@pytest.mark.gpu
def test_worker_sse_gamma_and_d_match_reference_runs(loopgpu):
    site = FX["site_spin1_gamma1"]["results"]
    w = loopgpu.Worker('ALGORITHM = "loop; sse"\nL = 2\nJ = 0\nlocal_S = 1\nGamma = 1\nT = 0.25\n'
                       'SWEEPS = 65536\nSEED = 23125')
    print("worker sse gamma", flush=True)
    w.run()
    for name, scale in (("Energy Density", 1.0), ("Susceptibility", 1.0), ("Magnetization^2", 0.5)):
        m, e, _ = w.result(name)
        assert stats.sigmas(m * scale, e * scale, site[name][0], site[name][1]) < 4.0, (name, m, e, site[name])
    dch = FX["chain4_spin1_d"]["results"]
    ref = ed.diagonalize_transverse("chain", 4, 2.0, 2, 0.0, -1.0, 0.0, 0.0, -0.2)
    w = loopgpu.Worker('ALGORITHM = "loop; sse"\nL = 4\nlocal_S = 1\nJxy = 0\nJz = -1\nD = -0.2\nT = 2\n'
                       'SWEEPS = 65536\nSEED = 23125')
    print("worker sse D", flush=True)
    w.run()
    for name in ("Energy", "Susceptibility", "Magnetization^2", "Staggered Magnetization^2"):
        m, e, _ = w.result(name)
        assert stats.sigmas(m, e, ref[name]) < 4.0, (name, m, e, ref[name])
        assert stats.sigmas(m, e, dch[name][0], dch[name][1]) < 4.0, (name, m, e, dch[name])


@pytest.mark.gpu
def test_sse_batch_and_checkpoint(loopgpu):
    """a batch of SSE chains equals each chain alone; a chain saved after 25
    sweeps and loaded into a fresh handle continues bit for bit"""
    ps = [(O.CHAIN, 16, 4.0, 1.0, 1.0, 1), (O.SQUARE, 8, 3.0, 1.0, 1.0, 1), (O.CHAIN, 8, 2.0, 1.0, 1.0, 2),
          (O.CHAIN, 32, 8.0, 0.5, 1.0, 1)]
    plist = [loopgpu.params(k, L, b, jxy, jz, seed=9, chain_id=i, two_s=t, algorithm=loopgpu.SSE)
             for i, (k, L, b, jxy, jz, t) in enumerate(ps)]
    bt = loopgpu.Batch(plist)
    print("sse batch", flush=True)
    rec = bt.run(30)
    for i, (k, L, b, jxy, jz, t) in enumerate(ps):
        c = O.PhiloxChain(k, L, b, jxy, jz, seed=9, chain_id=i, two_s=t)
        _cmp(rec[:, i, :], c.sweep_sse(30), "chain %d" % i)
    p = loopgpu.params(O.CHAIN, 16, 4.0, seed=4, chain_id=2, algorithm=loopgpu.SSE)
    a = loopgpu.Chain(p)
    ra = a.sweep(50)
    b = loopgpu.Chain(p)
    b.sweep(25)
    blob = b.save()
    c = loopgpu.Chain(p)
    print("sse checkpoint", flush=True)
    c.load(blob)
    _cmp(c.sweep(25), ra[25:], "resumed")


@pytest.mark.gpu
@pytest.mark.parametrize("key,local_s,T,jxy", [("chain4", "1/2", "0.25", "1"), ("chain4_spin1", "1", "0.25", "1"),
                                               ("chain4_ising", "1/2", "0.5", "0")])
def test_worker_sse_matches_reference_run(loopgpu, key, local_s, T, jxy):
    qmc = FX[key]["results"]
    ref = ed.diagonalize_transverse("chain", 4, float(T), 2 if local_s == "1" else 1, float(jxy), 1.0)
    w = loopgpu.Worker('ALGORITHM = "loop; sse"\nLATTICE = "chain lattice"\nL = 4\nlocal_S = %s\nJxy = %s\n'
                       'Jz = 1\nT = %s\nSWEEPS = 65536\nSEED = 23125' % (local_s, jxy, T))
    print("worker sse", key, flush=True)
    w.run()
    for name in ["Energy", "Specific Heat", "Magnetization^2", "Susceptibility", "Staggered Magnetization^2",
                 "Staggered Susceptibility"]:
        m, e, _ = w.result(name)
        assert stats.sigmas(m, e, ref[name]) < 4.0, (name, m, e, ref[name])
        if name in qmc:
            assert stats.sigmas(m, e, qmc[name][0], qmc[name][1]) < 4.0, (name, m, e, qmc[name])


@pytest.mark.gpu
def test_worker_sse_normal_estimator_matches_ed(loopgpu):
    ref = ed.diagonalize_transverse("chain", 6, 0.5, 1)
    w = loopgpu.Worker('ALGORITHM = "loop; sse"\nL = 6\nT = 0.5\nDISABLE_IMPROVED_ESTIMATOR = 1\n'
                       'SWEEPS = 65536\nSEED = 5')
    print("worker sse normal estimator", flush=True)
    w.run()
    for name in ("Energy", "Magnetization^2", "Susceptibility", "Staggered Magnetization^2",
                 "Staggered Susceptibility"):
        m, e, _ = w.result(name)
        assert stats.sigmas(m, e, ref[name]) < 4.0, (name, m, e, ref[name])
