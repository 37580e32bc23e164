"""S=1 on the chain through virtual sites (path_integral.C:589-615,
looper/permutation.h, the virtual lattice of looper/lattice.h:104-175).

CPU: the oracle's restatement of random_shuffle / partitioned_random_shuffle
against the reference's own test/permutation.op (mt19937 seed 23094, digit
for digit); the virtual site and bond order against test/virtual_lattice.op;
the S=1 exact diagonalisation against closed forms; both oracle modes against
it.  GPU: S=1 chains bit-exact with the oracle's philox mode (operator strings,
spins, records) under both linking kernels that handle virtual sites, in a
batch beside S=1/2 chains, and the worker (local_S = 1) against ED."""

import json
import math
import os

import numpy as np
import pytest

import ed
import oracle_lib as O
import stats

HERE = os.path.dirname(os.path.abspath(__file__))
FX = json.load(open(os.path.join(HERE, "golden", "reference_fixtures.json")))


def test_shuffles_match_reference_permutation_op():
    f = FX["permutation"]
    sh, c0, c1, pm = O.permutation_test(f["seed"], f["n"], len(f["shuffles"]), len(f["partitioned"]))
    np.testing.assert_array_equal(sh, f["shuffles"])
    for t, case in enumerate(f["partitioned"]):
        np.testing.assert_array_equal(c0[t], case["conf0"])
        np.testing.assert_array_equal(c1[t], case["conf1"])
        np.testing.assert_array_equal(pm[t], case["perm"])
        # every tau = 0 spin meets an equal tau = 1 spin
        assert all(case["conf0"][i] == case["conf1"][case["perm"][i]] for i in range(f["n"]))


def _virtual_order(real_edges, two_s):
    """virtual sites of each real site contiguous in real-site order; each
    real edge's virtual edges with the source's virtual site outer and the
    target's inner"""
    first = np.concatenate([[0], np.cumsum(two_s)])
    out = []
    for s0, s1 in real_edges:
        for a in range(two_s[s0]):
            for c in range(two_s[s1]):
                out.append([int(first[s0] + a), int(first[s1] + c)])
    return out, first


def test_virtual_lattice_order_matches_reference():
    f = FX["virtual_lattice"]
    edges, first = _virtual_order(f["real_edges"], f["two_s"])
    assert edges == f["virtual_edges"]
    assert [[int(first[s]), int(first[s + 1] - 1)] for s in range(4)] == f["site_map"]
    # the oracle's S=1 chain follows the same order, and every virtual site
    # takes its real site's gauge
    for L in (2, 4, 6):
        N, Nb, src, tgt, gauge = O.lattice(O.CHAIN, L, 2)
        want, _ = _virtual_order([[i, (i + 1) % L] for i in range(L)], [2] * L)
        assert (N, Nb) == (2 * L, 4 * L)
        assert [[int(a), int(b)] for a, b in zip(src, tgt)] == want
        assert list(gauge) == [1 if (v // 2) % 2 == 0 else -1 for v in range(2 * L)]


def test_spin1_ed_closed_form_two_sites():
    """L=2 has two bonds between the sites: H = 2 S0.S1 = S(S+1) - 4 for total
    spin S = 0, 1, 2 (1, 3, 5 states)"""
    for T in (0.5, 1.0, 3.0):
        b = 1.0 / T
        lev = [(-4.0, 1, 0.0), (-2.0, 3, 2.0), (2.0, 5, 10.0)]   # energy, degeneracy, sum of M^2 over the multiplet
        Z = sum(g * math.exp(-b * e) for e, g, _ in lev)
        E = sum(g * e * math.exp(-b * e) for e, g, _ in lev) / Z
        E2 = sum(g * e * e * math.exp(-b * e) for e, g, _ in lev) / Z
        M2 = sum(m2 * math.exp(-b * e) for e, _, m2 in lev) / Z
        r = ed.diagonalize_spin1_chain(2, T)
        assert r["Energy"] == pytest.approx(E, rel=1e-12)
        assert r["Specific Heat"] == pytest.approx(b * b * (E2 - E * E) / 2, rel=1e-10)
        assert r["Magnetization^2"] == pytest.approx(M2, rel=1e-12)
        assert r["Susceptibility"] == pytest.approx(b * M2 / 2, rel=1e-10)
    # ground-state energies of the S=1 Heisenberg ring (L = 4: -6)
    assert ed.diagonalize_spin1_chain(4, 0.01)["Ground State Energy"] == pytest.approx(-6.0, abs=1e-10)


CASES = [(4, 0.5, 1.0), (6, 0.3, 1.0), (4, 0.6, 0.6)]
NAMES = ["Energy", "Susceptibility", "Staggered Magnetization^2", "Staggered Susceptibility", "Magnetization^2"]


@pytest.mark.parametrize("L,T,jz", CASES)
def test_oracle_modes_match_spin1_ed(L, T, jz):
    ref = ed.diagonalize_spin1_chain(L, T, 1.0, jz)
    beta = 1.0 / T
    for name, ch in (("looper", O.LooperChain(O.CHAIN, L, beta, 1.0, jz, seed=11, two_s=2)),
                     ("philox", O.PhiloxChain(O.CHAIN, L, beta, 1.0, jz, seed=11, chain_id=1, two_s=2))):
        ch.sweep(2000)
        obs = stats.observables(ch.sweep(60000), beta, L, 1)
        for k in NAMES:
            m, e = stats.mean_err(obs[k])
            assert stats.sigmas(m, e, ref[k]) < 4.0, (name, k, m, e, ref[k])


def test_spin1_rejected_beyond_the_chain():
    import loopgpu
    with pytest.raises(loopgpu.LoopGPUError) as e:
        loopgpu.Chain(loopgpu.params(loopgpu.SQUARE, 4, 1.0, two_s=2))
    assert e.value.code == -1
    with pytest.raises(loopgpu.LoopGPUError) as e:
        loopgpu.Worker('LATTICE = "square lattice"; L = 4; T = 1; local_S = 1')
    assert e.value.code == -1
    with pytest.raises(loopgpu.LoopGPUError) as e:
        loopgpu.Worker('L = 4; T = 1; local_S = 3/2')
    assert e.value.code == -1


EXACT = [0, 1]
FLOAT = list(range(2, 18)) + [20, 21, 22]


def _cmp(g, c, where):
    np.testing.assert_array_equal(g[:, EXACT], c[:, EXACT], err_msg=where)
    np.testing.assert_allclose(g[:, FLOAT], c[:, FLOAT], rtol=1e-9, atol=1e-9, err_msg=where)


TRAJ = [(4, 2.0, 1.0), (8, 6.0, 0.7), (32, 12.0, 1.0), (2, 3.0, 1.0)]


@pytest.mark.gpu
@pytest.mark.parametrize("L,beta,jz", TRAJ)
def test_spin1_chain_bit_exact_with_oracle(loopgpu, L, beta, jz):
    c = O.PhiloxChain(O.CHAIN, L, beta, 1.0, jz, seed=3, chain_id=7, two_s=2)
    rc = c.sweep(40)
    cs = c.state()
    for lc in (dict(), dict(link=0), dict(link=1), dict(link=1, compress=1), dict(link=0, compress=1, group=4)):
        print("spin-1 chain", L, beta, jz, lc, flush=True)
        g = loopgpu.Chain(loopgpu.params(O.CHAIN, L, beta, 1.0, jz, seed=3, chain_id=7, two_s=2))
        if lc:
            g.set_launch(**lc)
        _cmp(g.sweep(40), rc, "%s %s" % ((L, beta, jz), lc))
        for a, b in zip(g.get_config(), cs):
            np.testing.assert_array_equal(a, b)
    g = loopgpu.Chain(loopgpu.params(O.CHAIN, L, beta, 1.0, jz, seed=3, chain_id=7, two_s=2))
    with pytest.raises(loopgpu.LoopGPUError) as e:
        g.set_launch(link=512)               # the LDS linking kernels assume wrapping world lines
    assert e.value.code == -1


@pytest.mark.gpu
def test_spin1_chains_in_a_mixed_batch(loopgpu):
    plist = [loopgpu.params(O.CHAIN, 8, 4.0, 1.0, 1.0, seed=9, chain_id=0, two_s=2),
             loopgpu.params(O.SQUARE, 6, 2.0, 1.0, 0.8, seed=9, chain_id=1),
             loopgpu.params(O.CHAIN, 16, 8.0, 1.0, 0.9, seed=9, chain_id=2, two_s=2),
             loopgpu.params(O.CHAIN, 10, 3.0, 1.0, 1.0, seed=9, chain_id=3)]
    n = 30
    b = loopgpu.Batch(plist)
    assert b.launch()["link"] in (0, 1)
    rb = b.run(n)
    for i, p in enumerate(plist):
        rs = loopgpu.Chain(p).sweep(n)
        np.testing.assert_array_equal(rb[:, i, EXACT], rs[:, EXACT])
        np.testing.assert_allclose(rb[:, i, FLOAT], rs[:, FLOAT], rtol=1e-9, atol=1e-9)


@pytest.mark.gpu
@pytest.mark.parametrize("L,T,jz", [(6, 0.3, 1.0), (4, 0.6, 0.6)])
def test_worker_spin1_matches_ed(loopgpu, L, T, jz):
    ref = ed.diagonalize_spin1_chain(L, T, 1.0, jz)
    w = loopgpu.Worker('LATTICE = "chain lattice"; L = %d; T = %r; Jxy = 1; Jz = %r; local_S = 1; '
                       'SWEEPS = 100000; THERMALIZATION = 2000; SEED = 5' % (L, T, jz))
    print("worker spin-1", L, T, jz, flush=True)
    w.run()
    for k in NAMES + ["Specific Heat"]:
        m, e, _ = w.result(k)
        assert stats.sigmas(m, e, ref[k]) < 4.0, (k, m, e, ref[k])
    assert w.result("Volume")[0] == L
