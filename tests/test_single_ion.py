"""Single-ion term D (S^z)^2 (looper/model_parameter.h:69) for S > 1/2:
looper's in-real-site bonds between the virtual sites of each real site
(lattice.h:632-646), each an Ising coupling 2D (weight_impl.h:194-197), and
the constant D nv/4 per real site (model_parameter.h:592-595,
weight_impl.h:400-404); looper scatters (FORCE_SCATTER 0.1) by default where
D > 0 makes it call the model frustrated (model_parameter.h:646-661,
model_impl.h:51).

CPU: the exact diagonalisation with D against the reference's own loop.op
block (S=1 Ising chain, Jz=-1, D=-0.2, T=2) and against closed forms of one
site; both oracle modes against it (S = 1, 3/2, 2, 5/2; with a transverse field;
D > 0 scattered) and against the reference's path-integral block.  GPU:
bit-exact with the oracle's philox mode under the launches that take in-site
bonds, batched beside plain chains, and the worker's D against loop.op and
ED."""

import json
import math
import os

import numpy as np
import pytest

import ed
import oracle_lib as O
import stats

HERE = os.path.dirname(os.path.abspath(__file__))
FX = json.load(open(os.path.join(HERE, "golden", "reference_fixtures.json")))["loop_op_models"]["chain4_spin1_d"]
SIX = dict(rel=5e-6, abs=1e-9)  # half a unit in the sixth significant digit
NAMES = ["Energy", "Specific Heat", "Magnetization^2", "Susceptibility", "Staggered Magnetization^2",
         "Staggered Susceptibility"]


def test_ed_matches_reference_single_ion_block():
    ref = FX[0]["results"]
    r = ed.diagonalize_transverse("chain", 4, 2.0, 2, 0.0, -1.0, 0.0, 0.0, -0.2)
    for k in NAMES:
        assert r[k] == pytest.approx(ref[k][0], **SIX), k


def test_ed_single_ion_closed_forms():
    # one spin S alone (J = 0): levels D m^2, m = -S..S, per site
    for two_s, d, T in ((2, -0.2, 2.0), (2, 0.7, 0.5), (3, 0.4, 1.0), (4, -0.3, 0.8)):
        S = 0.5 * two_s
        b = 1 / T
        ms = [S - q for q in range(two_s + 1)]
        w = [math.exp(-b * d * m * m) for m in ms]
        Z = sum(w)
        E = sum(d * m * m * x for m, x in zip(ms, w)) / Z
        M2 = sum(m * m * x for m, x in zip(ms, w)) / Z
        r = ed.diagonalize_transverse("chain", 2, T, two_s, 0.0, 0.0, 0.0, 0.0, d)
        assert r["Energy Density"] == pytest.approx(E, rel=1e-12, abs=1e-14)
        assert r["Magnetization^2"] == pytest.approx(2 * M2, rel=1e-12)
        assert r["Susceptibility"] == pytest.approx(b * M2, rel=1e-10)   # S^z conserved: chi = beta <m^2>


# L, 2S, T, Jxy, Jz, D, hx, FORCE_SCATTER
MODELS = [(4, 2, 2.0, 0.0, -1.0, -0.2, 0.0, 0.0),   # loop.op's block
          (4, 2, 0.5, 1.0, 1.0, 0.3, 0.0, 0.1),     # D > 0, scattered as looper does
          (4, 3, 0.8, 1.0, 1.0, -0.4, 0.0, 0.0),
          (4, 2, 0.6, -1.0, 1.0, 0.5, 0.7, 0.1),    # with a transverse field
          (2, 4, 1.0, 1.0, 0.5, -0.3, 0.0, 0.0),
          (4, 5, 2.0, 1.0, 1.0, -0.4, 0.0, 0.0)]


@pytest.mark.parametrize("L,two_s,T,jxy,jz,d,hx,fs", MODELS)
def test_oracle_modes_match_single_ion_ed(L, two_s, T, jxy, jz, d, hx, fs):
    ref = ed.diagonalize_transverse("chain", L, T, two_s, jxy, jz, hx, 0.0, d)
    beta = 1.0 / T
    for name, ch in (("looper", O.LooperChain(O.CHAIN, L, beta, jxy, jz, seed=5, fs=fs, two_s=two_s, hx=hx, d=d)),
                     ("philox", O.PhiloxChain(O.CHAIN, L, beta, jxy, jz, seed=5, chain_id=2, fs=fs, two_s=two_s,
                                              hx=hx, d=d))):
        ch.sweep(2000)
        obs = stats.observables(ch.sweep(40000), beta, L, 1)
        for k in ("Energy", "Susceptibility", "Magnetization^2", "Staggered Magnetization^2"):
            m, e = stats.mean_err(obs[k])
            assert stats.sigmas(m, e, ref[k]) < 4.0, (name, k, m, e, ref[k])


def test_oracle_matches_reference_single_ion_qmc():
    qmc = FX[1]["results"]
    beta = 0.5
    ch = O.LooperChain(O.CHAIN, 4, beta, 0.0, -1.0, seed=7, two_s=2, d=-0.2)
    ch.sweep(2000)
    obs = stats.observables(ch.sweep(40000), beta, 4, 1)
    for k in ("Energy", "Susceptibility", "Magnetization^2", "Staggered Magnetization^2"):
        m, e = stats.mean_err(obs[k])
        assert stats.sigmas(m, e, qmc[k][0], qmc[k][1]) < 4.0, (k, m, e, qmc[k])


def test_single_ion_sizes_and_checks():
    import loopgpu
    assert loopgpu.lattice_size(loopgpu.CHAIN, 4, 2, 0.0, -0.2) == (8, 16 + 4)
    assert loopgpu.lattice_size(loopgpu.CHAIN, 4, 4, 0.5, 0.3) == (16, 64 + 24 + 16)
    assert loopgpu.lattice_size(loopgpu.CHAIN, 4, 1, 0.0, 0.3) == (4, 4)   # S=1/2: a constant only
    with pytest.raises(loopgpu.LoopGPUError) as e:
        loopgpu.Exchange(loopgpu.params(loopgpu.CHAIN, 4, 1.0, 1.0, 1.0, two_s=2, d=0.3), [1.0, 2.0])
    assert "single-ion" in str(e.value)


EXACT = [0, 1]
FLOAT = list(range(2, 18)) + [20, 21, 22]


def _cmp(g, c, where):
    np.testing.assert_array_equal(g[:, EXACT], c[:, EXACT], err_msg=where)
    np.testing.assert_allclose(g[:, FLOAT], c[:, FLOAT], rtol=1e-9, atol=1e-9, err_msg=where)


# kind, L, 2S, beta, Jxy, Jz, D, hx, FORCE_SCATTER
TRAJ = [(O.CHAIN, 4, 2, 0.5, 0.0, -1.0, -0.2, 0.0, 0.0), (O.CHAIN, 16, 2, 4.0, 1.0, 1.0, 0.4, 0.0, 0.1),
        (O.CHAIN, 8, 3, 3.0, 1.0, 0.8, -0.5, 0.0, 0.0), (O.CHAIN, 8, 4, 2.0, 1.0, 1.0, 0.3, 0.0, 0.1),
        (O.CHAIN, 8, 2, 3.0, -1.0, 1.0, 0.5, 0.6, 0.1), (O.SQUARE, 4, 2, 2.0, 1.0, 1.0, -0.6, 0.0, 0.0),
        (O.CHAIN, 8, 1, 2.0, 1.0, 1.0, 0.7, 0.0, 0.0), (O.SQUARE, 2, 4, 1.5, 1.0, 1.0, 0.4, 0.0, 0.1),
        (O.CHAIN, 4, 6, 1.0, 1.0, 1.0, -0.3, 0.0, 0.0)]
LAUNCHES = [dict(), dict(group=4), dict(group=64, compress=1), dict(group=16, blocks=64)]


@pytest.mark.gpu
@pytest.mark.parametrize("kind,L,two_s,beta,jxy,jz,d,hx,fs", TRAJ)
def test_single_ion_bit_exact_with_oracle(loopgpu, kind, L, two_s, beta, jxy, jz, d, hx, fs):
    c = O.PhiloxChain(kind, L, beta, jxy, jz, seed=3, chain_id=7, fs=fs, two_s=two_s, hx=hx, d=d)
    rc = c.sweep(40)
    cs = c.state()
    for lc in LAUNCHES:
        print("single-ion", kind, L, two_s, beta, jxy, jz, d, hx, fs, lc, flush=True)
        g = loopgpu.Chain(loopgpu.params(kind, L, beta, jxy, jz, seed=3, chain_id=7, force_scatter=fs,
                                         two_s=two_s, hx=hx, d=d))
        if lc:
            g.set_launch(**lc)
        _cmp(g.sweep(40), rc, "%s %s" % ((kind, L, two_s, beta, jxy, jz, d, hx, fs), lc))
        for a, b in zip(g.get_config(), cs):
            np.testing.assert_array_equal(a, b)


@pytest.mark.gpu
def test_single_ion_chains_batched_with_plain_ones(loopgpu):
    ps = [dict(L=8, beta=3.0, jxy=1.0, jz=1.0, d=-0.4, hx=0.0, two_s=2, fs=0.0),
          dict(L=8, beta=3.0, jxy=1.0, jz=1.0, d=0.0, hx=0.0, two_s=2, fs=0.0),
          dict(L=16, beta=4.0, jxy=1.0, jz=1.0, d=0.0, hx=0.0, two_s=1, fs=0.0),
          dict(L=8, beta=2.0, jxy=-1.0, jz=0.5, d=0.6, hx=0.8, two_s=3, fs=0.1),
          dict(L=8, beta=2.0, jxy=0.0, jz=1.0, d=0.0, hx=1.0, two_s=1, fs=0.0)]
    plist = [loopgpu.params(loopgpu.CHAIN, q["L"], q["beta"], q["jxy"], q["jz"], seed=9, chain_id=i,
                            force_scatter=q["fs"], two_s=q["two_s"], hx=q["hx"], d=q["d"]) for i, q in enumerate(ps)]
    b = loopgpu.Batch(plist)
    print("batched single-ion", flush=True)
    rec = b.run(30)
    for i, q in enumerate(ps):
        c = O.PhiloxChain(O.CHAIN, q["L"], q["beta"], q["jxy"], q["jz"], seed=9, chain_id=i, fs=q["fs"],
                          two_s=q["two_s"], hx=q["hx"], d=q["d"])
        _cmp(rec[:, i, :], c.sweep(30), "chain %d" % i)


@pytest.mark.gpu
def test_worker_single_ion_matches_reference_run(loopgpu):
    """loop.ip's block: S=1 chain L=4, Jxy=0, Jz=-1, D=-0.2, T=2 against
    loop.op's diagonalization and path-integral results"""
    ed_ref, qmc = FX[0]["results"], FX[1]["results"]
    w = loopgpu.Worker('LATTICE = "chain lattice"\nL = 4\nlocal_S = 1\nJxy = 0\nJz = -1\nD = -0.2\nT = 2\n'
                       'SWEEPS = 65536\nSEED = 23125')
    print("worker D", flush=True)
    w.run()
    for name in NAMES:
        m, e, _ = w.result(name)
        assert stats.sigmas(m, e, ed_ref[name][0]) < 4.0, (name, m, e, ed_ref[name][0])
        if name in qmc:
            assert stats.sigmas(m, e, qmc[name][0], qmc[name][1]) < 4.0, (name, m, e, qmc[name])


@pytest.mark.gpu
def test_worker_positive_single_ion_scatters_and_matches_ed(loopgpu):
    ref = ed.diagonalize_transverse("chain", 4, 0.5, 2, 1.0, 1.0, 0.0, 0.0, 0.3)
    w = loopgpu.Worker('LATTICE = "chain lattice"\nL = 4\nlocal_S = 1\nJ = 1\nD = 0.3\nT = 0.5\n'
                       'SWEEPS = 65536\nSEED = 31')
    print("worker D > 0", flush=True)
    w.run()
    for name in NAMES:
        m, e, _ = w.result(name)
        assert stats.sigmas(m, e, ref[name]) < 4.0, (name, m, e, ref[name])
