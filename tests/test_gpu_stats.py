"""GPU statistics: observables of the HIP path against exact diagonalisation
and the oracle's looper-mode restatement of the reference worker, within 3
combined standard errors (tolerance stated per test)."""

import numpy as np
import pytest

import oracle_lib as O
import stats

pytestmark = pytest.mark.gpu

OBS = ["Energy Density", "Susceptibility", "Staggered Structure Factor", "Staggered Susceptibility"]


def _ed(kind, L, T, jz):
    import ed
    r = ed.diagonalize("chain" if kind == O.CHAIN else "square", L, T, 1.0, jz)
    V = L if kind == O.CHAIN else L * L
    return {"Energy Density": r["Energy"] / V, "Susceptibility": r["Susceptibility"],
            "Staggered Structure Factor": r["Staggered Magnetization^2"] / V,
            "Staggered Susceptibility": r["Staggered Susceptibility"]}


@pytest.mark.parametrize("kind,L,T,jz", [
    (O.CHAIN, 12, 0.1, 1.0),    # C1' (exact-diagonalisation size of the C1 workload)
    (O.CHAIN, 10, 0.2, 0.6),
    (O.SQUARE, 2, 0.3, 1.0),
    (O.CHAIN, 8, 0.25, 1.5),
])
def test_gpu_matches_exact_diagonalisation(loopgpu, kind, L, T, jz):
    beta = 1 / T
    V = L if kind == O.CHAIN else L * L
    ch = loopgpu.Chain(loopgpu.params(kind, L, beta, 1.0, jz, seed=23125))
    print("ED", kind, L, T, jz, flush=True)
    ch.sweep(4096, records=False)
    rec = ch.sweep(65536)
    obs = stats.observables(rec, beta, V, 1 if kind == O.CHAIN else 2)
    ref = _ed(kind, L, T, jz)
    for k in OBS:
        m, e = stats.mean_err(obs[k])
        assert stats.sigmas(m, e, ref[k]) < 3.0, (k, m, e, ref[k])


def test_c1_chain_matches_reference_worker_restatement(loopgpu):
    # C1: chain L=64, J=1, beta=10, SEED=23125 (thermalisation shortened; the
    # observables agree with the CPU looper-mode run within 3 combined errors)
    L, beta = 64, 10.0
    ch = loopgpu.Chain(loopgpu.params(O.CHAIN, L, beta, seed=23125))
    ch.sweep(2048, records=False)
    g = stats.observables(ch.sweep(65536), beta, L, 1)
    cpu = O.LooperChain(O.CHAIN, L, beta, seed=23125)
    cpu.sweep(2048)
    c = stats.observables(cpu.sweep(32768), beta, L, 1)
    for k in OBS + ["Number of Clusters", "Stiffness"]:
        mg, eg = stats.mean_err(g[k])
        mc, ec = stats.mean_err(c[k])
        assert stats.sigmas(mg, eg, mc, ec) < 3.0, (k, mg, eg, mc, ec)


def test_batched_scan_chains_match_exact_diagonalisation(loopgpu):
    temps = [0.1, 0.2, 0.4, 0.8]
    plist = [loopgpu.params(O.CHAIN, 8, 1 / T, seed=1, chain_id=i) for i, T in enumerate(temps)]
    b = loopgpu.Batch(plist)
    b.run(2048, records=False)
    rec = b.run(32768)
    for i, T in enumerate(temps):
        obs = stats.observables(rec[:, i, :], 1 / T, 8, 1)
        ref = _ed(O.CHAIN, 8, T, 1.0)
        for k in OBS:
            m, e = stats.mean_err(obs[k])
            assert stats.sigmas(m, e, ref[k]) < 3.5, (T, k, m, e, ref[k])
