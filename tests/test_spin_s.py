"""Spin S > 1/2 through virtual sites (path_integral.C:589-615,
looper/permutation.h, the virtual lattice of looper/lattice.h:104-175):
S = 1, 3/2, ... 4 on the chain and the square lattice.

CPU: the oracle's restatement of random_shuffle / partitioned_random_shuffle
against the reference's own test/permutation.op (mt19937 seed 23094, digit
for digit); the virtual site and bond order against test/virtual_lattice.op
(a mixed-spin lattice, so the order rule is pinned there and the uniform
lattices are checked against the rule); the spin-S exact diagonalisation
against closed forms; both oracle modes against it.  GPU: virtual-site
chains bit-exact with the oracle's philox mode (operator strings, spins,
records) under the linking kernels that handle them, in a batch beside
S=1/2 chains, and the worker (local_S) against ED."""

import json
import math
import os
from fractions import Fraction

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


def _real_edges(kind, L):
    if kind == O.CHAIN:
        return [[i, (i + 1) % L] for i in range(L)]
    out = []
    for s in range(L * L):
        x, y = divmod(s, L)
        out += [[s, x * L + (y + 1) % L], [s, ((x + 1) % L) * L + y]]
    return out


def test_virtual_lattice_order_matches_reference():
    f = FX["virtual_lattice"]
    edges, first = _virtual_order(f["real_edges"], f["two_s"])
    assert edges == f["virtual_edges"]
    assert [[int(first[s]), int(first[s + 1] - 1)] for s in range(4)] == f["site_map"]
    # the oracle's uniform virtual lattices follow the same order, and every
    # virtual site takes its real site's gauge
    for kind, L, nv in [(O.CHAIN, 2, 2), (O.CHAIN, 4, 2), (O.CHAIN, 6, 2), (O.CHAIN, 4, 3), (O.CHAIN, 6, 4),
                        (O.SQUARE, 2, 2), (O.SQUARE, 4, 2)]:
        N, Nb, src, tgt, gauge = O.lattice(kind, L, nv)
        real = _real_edges(kind, L)
        nreal = L if kind == O.CHAIN else L * L
        want, _ = _virtual_order(real, [nv] * nreal)
        assert (N, Nb) == (nv * nreal, nv * nv * len(real))
        assert [[int(a), int(b)] for a, b in zip(src, tgt)] == want
        _, _, _, _, g1 = O.lattice(kind, L, 1)
        assert list(gauge) == [int(g1[v // nv]) for v in range(N)]


def _two_site_levels(S):
    """L=2 has two bonds between the sites: H = 2 S0.S1 = St(St+1) - 2S(S+1)
    for total spin St = 0..2S; (energy, degeneracy, sum of M^2 over the multiplet)"""
    out = []
    for k in range(int(round(2 * S)) + 1):
        St = float(k)
        out.append((St * (St + 1) - 2 * S * (S + 1), 2 * k + 1, St * (St + 1) * (2 * St + 1) / 3))
    return out


@pytest.mark.parametrize("two_s", [2, 3, 4, 5, 6, 8])
def test_spin_s_ed_closed_forms(two_s):
    S = 0.5 * two_s
    for T in (0.5, 1.0, 3.0):
        b = 1.0 / T
        lev = _two_site_levels(S)
        Z = sum(g * math.exp(-b * e) for e, g, _ in lev)
        E = sum(g * e * math.exp(-b * e) for e, g, _ in lev) / Z
        E2 = sum(g * e * e * math.exp(-b * e) for e, g, _ in lev) / Z
        M2 = sum(m2 * math.exp(-b * e) for e, _, m2 in lev) / Z
        r = ed.diagonalize_spin("chain", 2, T, two_s)
        assert r["Energy"] == pytest.approx(E, rel=1e-12)
        assert r["Specific Heat"] == pytest.approx(b * b * (E2 - E * E) / 2, rel=1e-10)
        assert r["Magnetization^2"] == pytest.approx(M2, rel=1e-12)
        assert r["Susceptibility"] == pytest.approx(b * M2 / 2, rel=1e-10)
    # four-site ring: H = (S_A + S_B)^2 / 2 - ... with the sublattice spins
    # S_A = S_0 + S_2, S_B = S_1 + S_3: E0 = -2S(2S+1)
    assert ed.diagonalize_spin("chain", 4, 0.01, two_s)["Ground State Energy"] == pytest.approx(
        -two_s * (two_s + 1), abs=1e-9)
    # S=1/2 reduces to the S=1/2 routine
    a, c = ed.diagonalize_spin("chain", 6, 0.7, 1, 0.8, 1.1), ed.diagonalize("chain", 6, 0.7, 0.8, 1.1)
    for k in c:
        if k in a:
            assert a[k] == pytest.approx(c[k], rel=1e-10, abs=1e-12)


@pytest.mark.parametrize("two_s", [2, 3, 4])
def test_spin_s_square_ed_closed_form(two_s):
    """2x2 square: every neighbour pair carries two bonds, H = 2 S_A.S_B with the
    sublattice spins S_A = S_0 + S_3, S_B = S_1 + S_2, so E0 = -2 * 2S(2S+1)
    (-12 for S=1, -24 for S=3/2, -40 for S=2)"""
    assert ed.diagonalize_spin("square", 2, 0.01, two_s)["Ground State Energy"] == pytest.approx(
        -2.0 * two_s * (two_s + 1), abs=1e-9)


# kind, L, 2S, T, Jz
CASES = [(O.CHAIN, 4, 2, 0.5, 1.0), (O.CHAIN, 6, 2, 0.3, 1.0), (O.CHAIN, 4, 2, 0.6, 0.6),
         (O.CHAIN, 4, 3, 0.8, 1.0), (O.CHAIN, 4, 4, 1.2, 0.8), (O.SQUARE, 2, 2, 0.9, 1.0),
         (O.CHAIN, 4, 5, 2.0, 1.0), (O.CHAIN, 4, 8, 4.0, 1.0), (O.SQUARE, 2, 3, 1.5, 1.0), (O.SQUARE, 2, 4, 2.0, 0.9)]
NAMES = ["Energy", "Susceptibility", "Staggered Magnetization^2", "Staggered Susceptibility", "Magnetization^2"]


def _ed(kind, L, two_s, T, jz):
    return ed.diagonalize_spin("chain" if kind == O.CHAIN else "square", L, T, two_s, 1.0, jz)


@pytest.mark.parametrize("kind,L,two_s,T,jz", CASES)
def test_oracle_modes_match_spin_s_ed(kind, L, two_s, T, jz):
    ref = _ed(kind, L, two_s, T, jz)
    beta = 1.0 / T
    n = L if kind == O.CHAIN else L * L
    for name, ch in (("looper", O.LooperChain(kind, L, beta, 1.0, jz, seed=11, two_s=two_s)),
                     ("philox", O.PhiloxChain(kind, L, beta, 1.0, jz, seed=11, chain_id=1, two_s=two_s))):
        ch.sweep(2000)
        obs = stats.observables(ch.sweep(40000), beta, n, 1 if kind == O.CHAIN else 2)
        for k in NAMES:
            m, e = stats.mean_err(obs[k])
            assert stats.sigmas(m, e, ref[k]) < 4.0, (name, k, m, e, ref[k])


def test_spin_s_rejected_beyond_the_built_range():
    import loopgpu
    for kind, two_s in ((loopgpu.SQUARE, 9), (loopgpu.CHAIN, 9), (loopgpu.CHAIN, -1)):
        with pytest.raises(loopgpu.LoopGPUError) as e:
            loopgpu.Chain(loopgpu.params(kind, 4, 1.0, two_s=two_s))
        assert e.value.code == -1
    # more than 2^27 virtual bonds: S=4 on a 1026^2 square lattice
    with pytest.raises(loopgpu.LoopGPUError) as e:
        loopgpu.Chain(loopgpu.params(loopgpu.SQUARE, 1026, 1.0, two_s=8))
    assert e.value.code == -1 and "too large" in str(e.value)
    for par in ('LATTICE = "square lattice"; L = 4; T = 1; local_S = 9/2', 'L = 4; T = 1; local_S = 5',
                'L = 4; T = 1; local_S = 0.7', 'L = 4; T = 1; local_S = 0'):
        with pytest.raises(loopgpu.LoopGPUError) as e:
            loopgpu.Worker(par)
        assert e.value.code == -1


EXACT = [0, 1]
FLOAT = list(range(2, 18)) + [20, 21, 22]


def _cmp(g, c, where):
    np.testing.assert_array_equal(g[:, EXACT], c[:, EXACT], err_msg=where)
    np.testing.assert_allclose(g[:, FLOAT], c[:, FLOAT], rtol=1e-9, atol=1e-9, err_msg=where)


# kind, L, 2S, beta, Jz
TRAJ = [(O.CHAIN, 4, 2, 2.0, 1.0), (O.CHAIN, 8, 2, 6.0, 0.7), (O.CHAIN, 32, 2, 12.0, 1.0), (O.CHAIN, 2, 2, 3.0, 1.0),
        (O.CHAIN, 4, 3, 2.0, 1.0), (O.CHAIN, 16, 3, 8.0, 0.8), (O.CHAIN, 8, 4, 4.0, 1.0),
        (O.SQUARE, 2, 2, 1.5, 1.0), (O.SQUARE, 4, 2, 3.0, 1.0), (O.SQUARE, 8, 2, 5.0, 0.9),
        (O.CHAIN, 4, 5, 2.0, 1.0), (O.CHAIN, 8, 6, 3.0, 0.9), (O.CHAIN, 6, 7, 2.0, 1.0), (O.CHAIN, 4, 8, 2.0, 1.0),
        (O.SQUARE, 2, 3, 1.5, 1.0), (O.SQUARE, 4, 4, 2.0, 1.0), (O.SQUARE, 4, 8, 0.7, 1.0)]


def _launches(kind, two_s):
    lists = two_s * (2 if kind == O.CHAIN else 4)
    out = [dict(), dict(link=0), dict(link=0, compress=1, group=4)]
    if lists <= 4:
        out += [dict(link=1), dict(link=1, compress=1)]
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("kind,L,two_s,beta,jz", TRAJ)
def test_spin_s_bit_exact_with_oracle(loopgpu, kind, L, two_s, beta, jz):
    c = O.PhiloxChain(kind, L, beta, 1.0, jz, seed=3, chain_id=7, two_s=two_s)
    rc = c.sweep(40)
    cs = c.state()
    for lc in _launches(kind, two_s):
        print("spin-S", kind, L, two_s, beta, jz, lc, flush=True)
        g = loopgpu.Chain(loopgpu.params(kind, L, beta, 1.0, jz, seed=3, chain_id=7, two_s=two_s))
        if lc:
            g.set_launch(**lc)
        _cmp(g.sweep(40), rc, "%s %s" % ((kind, L, two_s, beta, jz), lc))
        for a, b in zip(g.get_config(), cs):
            np.testing.assert_array_equal(a, b)
    g = loopgpu.Chain(loopgpu.params(kind, L, beta, 1.0, jz, seed=3, chain_id=7, two_s=two_s))
    with pytest.raises(loopgpu.LoopGPUError) as e:
        g.set_launch(link=512)               # the LDS linking kernels assume wrapping world lines
    assert e.value.code == -1
    if two_s * (2 if kind == O.CHAIN else 4) > 4:
        with pytest.raises(loopgpu.LoopGPUError) as e:
            g.set_launch(link=1)             # the per-site kernel merges four lists at most
        assert e.value.code == -1


@pytest.mark.gpu
def test_spin_s_chains_in_a_mixed_batch(loopgpu):
    plist = [loopgpu.params(O.CHAIN, 8, 4.0, 1.0, 1.0, seed=9, chain_id=0, two_s=2),
             loopgpu.params(O.SQUARE, 6, 2.0, 1.0, 0.8, seed=9, chain_id=1),
             loopgpu.params(O.CHAIN, 16, 8.0, 1.0, 0.9, seed=9, chain_id=2, two_s=2),
             loopgpu.params(O.CHAIN, 10, 3.0, 1.0, 1.0, seed=9, chain_id=3),
             loopgpu.params(O.CHAIN, 6, 3.0, 1.0, 1.0, seed=9, chain_id=4, two_s=3),
             loopgpu.params(O.SQUARE, 4, 2.5, 1.0, 1.0, seed=9, chain_id=5, two_s=2),
             loopgpu.params(O.CHAIN, 4, 2.0, 1.0, 0.7, seed=9, chain_id=6, two_s=4)]
    n = 30
    b = loopgpu.Batch(plist)
    assert b.launch()["link"] == 0          # sites with more than four lists
    rb = b.run(n)
    for i, p in enumerate(plist):
        rs = loopgpu.Chain(p).sweep(n)
        np.testing.assert_array_equal(rb[:, i, EXACT], rs[:, EXACT])
        np.testing.assert_allclose(rb[:, i, FLOAT], rs[:, FLOAT], rtol=1e-9, atol=1e-9)
    # the same chains against the oracle, two of them
    for i in (4, 5):
        p = plist[i]
        c = O.PhiloxChain(p.lattice, p.L, p.beta, 1.0, p.jz, seed=9, chain_id=i, two_s=p.two_s)
        _cmp(rb[:, i, :], c.sweep(n), "batch chain %d" % i)


WORKER = [("chain lattice", 6, "1", 0.3, 1.0), ("chain lattice", 4, "1", 0.6, 0.6),
          ("chain lattice", 4, "3/2", 0.8, 1.0), ("chain lattice", 4, "2", 1.2, 1.0),
          ("square lattice", 2, "1", 0.9, 1.0), ("chain lattice", 4, "5/2", 2.0, 1.0), ("chain lattice", 4, "4", 4.0, 1.0),
          ("square lattice", 2, "3/2", 1.5, 1.0), ("square lattice", 2, "2", 2.0, 0.9)]


@pytest.mark.gpu
@pytest.mark.parametrize("lattice,L,S,T,jz", WORKER)
def test_worker_spin_s_matches_ed(loopgpu, lattice, L, S, T, jz):
    two_s = int(2 * Fraction(S))
    ref = ed.diagonalize_spin(lattice.split()[0], L, T, two_s, 1.0, jz)
    w = loopgpu.Worker('LATTICE = "%s"; L = %d; T = %r; Jxy = 1; Jz = %r; local_S = %s; '
                       'SWEEPS = 100000; THERMALIZATION = 2000; SEED = 5' % (lattice, L, T, jz, S))
    print("worker spin-S", lattice, L, S, T, jz, flush=True)
    w.run()
    for k in NAMES + ["Specific Heat"]:
        m, e, _ = w.result(k)
        assert stats.sigmas(m, e, ref[k]) < 4.0, (k, m, e, ref[k])
    assert w.result("Volume")[0] == (L if lattice == "chain lattice" else L * L)
