"""Transverse field -hx S^x (looper/model_parameter.h:67, the spin model's
Gamma): site operators on one bond per site after the lattice's bonds, the
site graph of weight and offset |hx|/2 (weight_impl.h:63-88), a cut of the
world line at every site operator (graph_impl.h:66-85, path_integral.C:
555-563 and 714-723).

CPU: the dense exact diagonalisation against the reference's own loop.op
(the diagonalization blocks of one S=1 site in Gamma=1 and in h=0.3, the
S=1 chain and the Ising chain), against the S^z-sector routines at hx=0 and
against the closed form of free spins; both oracle modes against it, and the
looper mode against the reference's path-integral block for Gamma=1.  GPU:
bit-exact with the oracle's philox mode (operator strings, spins, records)
under every launch setting that takes site operators, transverse-field chains
batched beside plain ones, and the worker's Gamma against loop.op and ED."""

import json
import math
import os

import numpy as np
import pytest

import ed
import oracle_lib as O
import stats

HERE = os.path.dirname(os.path.abspath(__file__))
FX = json.load(open(os.path.join(HERE, "golden", "reference_fixtures.json")))["loop_op_models"]

# loop.op prints six significant digits
SIX = dict(rel=5e-6, abs=1e-9)  # half a unit in the sixth significant digit


def _fx(model, which):
    return FX[model][0 if which == "ed" else 1]["results"]


def test_ed_matches_reference_site_in_transverse_field():
    """one S=1 site, Gamma=1, T=1/4 (loop.ip's LATTICE = "site" block)"""
    ref = _fx("site_spin1_gamma1", "ed")
    r = ed.diagonalize_transverse("site", 1, 0.25, 2, hx=1.0)
    for k in ("Energy", "Specific Heat", "Magnetization^2", "Magnetization^4", "Susceptibility",
              "Staggered Susceptibility", "Staggered Magnetization^2"):
        assert r[k] == pytest.approx(ref[k][0], **SIX), k
    assert r["Ground State Energy"] == pytest.approx(ref["Ground State Energy"][0], **SIX)


def test_ed_matches_reference_site_in_field_and_chains():
    ref = _fx("site_spin1_h03", "ed")
    r = ed.diagonalize_transverse("site", 1, 0.25, 2, hx=0.0, h=0.3)
    for k in ("Energy", "Specific Heat", "Magnetization^2", "Susceptibility"):
        assert r[k] == pytest.approx(ref[k][0], **SIX), k
    ref = _fx("chain4_spin1", "ed")
    for r in (ed.diagonalize_transverse("chain", 4, 0.25, 2), ed.diagonalize_spin("chain", 4, 0.25, 2)):
        for k in ("Energy", "Specific Heat", "Magnetization^2", "Susceptibility", "Staggered Magnetization^2",
                  "Staggered Susceptibility"):
            assert r[k] == pytest.approx(ref[k][0], **SIX), k
    ref = _fx("chain4_ising", "ed")
    for r in (ed.diagonalize_transverse("chain", 4, 0.5, 1, 0.0, 1.0), ed.diagonalize("chain", 4, 0.5, 0.0, 1.0)):
        for k in ("Energy", "Specific Heat", "Magnetization^2", "Susceptibility", "Staggered Magnetization^2",
                  "Staggered Susceptibility"):
            assert r[k] == pytest.approx(ref[k][0], **SIX), k


def test_ed_transverse_reduces_and_closed_form():
    # hx = 0: the S^z-sector routines
    for kind, L, two_s, T, jxy, jz, h in (("chain", 6, 1, 0.7, 0.8, 1.1, 0.0), ("chain", 4, 1, 0.5, -1.0, 0.4, 0.3),
                                          ("square", 2, 1, 0.9, 1.0, 1.0, 0.0), ("chain", 4, 3, 1.1, 1.0, 0.7, 0.0)):
        a = ed.diagonalize_transverse(kind, L, T, two_s, jxy, jz, 0.0, h)
        b = ed.diagonalize(kind, L, T, jxy, jz, h) if two_s == 1 else ed.diagonalize_spin(kind, L, T, two_s, jxy, jz)
        for k in ("Energy", "Specific Heat", "Magnetization^2", "Susceptibility", "Staggered Magnetization^2"):
            assert a[k] == pytest.approx(b[k], rel=1e-10, abs=1e-12), (kind, L, two_s, k)
    # free spins S=1/2 in -hx S^x: levels -+hx/2 per site; <(S^z)^2> = 1/4,
    # chi = tanh(beta hx/2) / (2 hx) per site (Kubo, the two levels connected by S^z)
    for hx, T in ((0.8, 0.5), (2.0, 0.3), (-1.3, 1.7)):
        b = 1 / T
        r = ed.diagonalize_transverse("chain", 4, T, 1, 0.0, 0.0, hx)
        a = abs(hx)
        assert r["Energy Density"] == pytest.approx(-a / 2 * math.tanh(b * a / 2), rel=1e-12)
        assert r["Magnetization^2"] == pytest.approx(1.0, rel=1e-12)
        assert r["Susceptibility"] == pytest.approx(math.tanh(b * a / 2) / (2 * a), rel=1e-10)
        assert r["Specific Heat"] == pytest.approx((b * a / 2) ** 2 / math.cosh(b * a / 2) ** 2, rel=1e-10)


# kind, L, 2S, T, Jxy, Jz, hx, h
MODELS = [(O.CHAIN, 4, 1, 0.5, 0.0, 1.0, 0.8, 0.0),    # transverse-field Ising chain
          (O.CHAIN, 4, 1, 0.5, -1.0, 1.0, 0.8, 0.0),   # ferromagnetic XY coupling
          (O.CHAIN, 4, 1, 0.8, -0.5, -1.0, -0.6, 0.0),
          (O.CHAIN, 4, 1, 0.6, -1.0, 0.5, 0.7, 0.3),   # with a longitudinal field
          (O.SQUARE, 2, 1, 0.7, 0.0, 1.0, 1.2, 0.0),
          (O.CHAIN, 4, 2, 0.7, -1.0, 1.0, 0.6, 0.0)]   # S = 1
NAMES = ["Energy", "Susceptibility", "Magnetization^2", "Staggered Magnetization^2", "Staggered Susceptibility"]


def _ed(kind, L, two_s, T, jxy, jz, hx, h):
    return ed.diagonalize_transverse("chain" if kind == O.CHAIN else "square", L, T, two_s, jxy, jz, hx, h)


@pytest.mark.parametrize("kind,L,two_s,T,jxy,jz,hx,h", MODELS)
def test_oracle_modes_match_transverse_ed(kind, L, two_s, T, jxy, jz, hx, h):
    ref = _ed(kind, L, two_s, T, jxy, jz, hx, h)
    beta = 1.0 / T
    n = L if kind == O.CHAIN else L * L
    names = NAMES if h == 0 else ["Energy"]   # the improved estimators need h = 0
    for name, ch in (("looper", O.LooperChain(kind, L, beta, jxy, jz, seed=5, field=h, two_s=two_s, hx=hx)),
                     ("philox", O.PhiloxChain(kind, L, beta, jxy, jz, seed=5, chain_id=2, field=h, two_s=two_s,
                                              hx=hx))):
        ch.sweep(2000)
        obs = stats.observables(ch.sweep(40000), beta, n, 1 if kind == O.CHAIN else 2)
        for k in names:
            m, e = stats.mean_err(obs[k])
            assert stats.sigmas(m, e, ref[k]) < 4.0, (name, k, m, e, ref[k])


def test_oracle_matches_reference_site_qmc():
    """loop.op's path-integral run of one S=1 site in Gamma=1 (T=1/4): two
    such sites, uncoupled (a chain of two with J=0), have its energy,
    susceptibility and specific heat per site and twice its <M^2>"""
    ed_ref, qmc = _fx("site_spin1_gamma1", "ed"), _fx("site_spin1_gamma1", "qmc")
    beta = 4.0
    for name, ch in (("looper", O.LooperChain(O.CHAIN, 2, beta, 0.0, 0.0, seed=11, two_s=2, hx=1.0)),
                     ("philox", O.PhiloxChain(O.CHAIN, 2, beta, 0.0, 0.0, seed=11, chain_id=1, two_s=2, hx=1.0))):
        ch.sweep(2000)
        obs = stats.observables(ch.sweep(40000), beta, 2, 1)
        for k, scale in (("Energy Density", 1.0), ("Susceptibility", 1.0), ("Magnetization^2", 0.5)):
            m, e = stats.mean_err(obs[k] * scale)
            assert stats.sigmas(m, e, ed_ref[k][0]) < 4.0, (name, k, m, e)
            assert stats.sigmas(m, e, qmc[k][0], qmc[k][1]) < 4.0, (name, k, m, e, qmc[k])


def test_transverse_field_parameter_checks():
    import loopgpu
    with pytest.raises(loopgpu.LoopGPUError) as e:
        loopgpu.Exchange(loopgpu.params(loopgpu.CHAIN, 4, 1.0, -1.0, 1.0, hx=0.5), [1.0, 2.0])
    assert "transverse" in str(e.value)
    with pytest.raises(loopgpu.LoopGPUError) as e:
        loopgpu.cluster_config(loopgpu.params(loopgpu.CHAIN, 4, 1.0, 0.0, 1.0, hx=0.5), np.zeros(4, np.int32),
                               np.zeros(9, np.int32), np.zeros(0), np.zeros(0, np.int32), 0)
    assert "hx = 0" in str(e.value)
    assert loopgpu.lattice_size(loopgpu.CHAIN, 4, 1, 0.5) == (4, 8)
    assert loopgpu.lattice_size(loopgpu.SQUARE, 2, 2, 0.5) == (8, 32 + 8)


EXACT = [0, 1]
FLOAT = list(range(2, 18)) + [20, 21, 22]


def _cmp(g, c, where):
    np.testing.assert_array_equal(g[:, EXACT], c[:, EXACT], err_msg=where)
    np.testing.assert_allclose(g[:, FLOAT], c[:, FLOAT], rtol=1e-9, atol=1e-9, err_msg=where)


# kind, L, 2S, beta, Jxy, Jz, hx, h
TRAJ = [(O.CHAIN, 4, 1, 2.0, 0.0, 1.0, 0.8, 0.0), (O.CHAIN, 32, 1, 8.0, -1.0, 0.5, 1.5, 0.0),
        (O.CHAIN, 16, 1, 6.0, 0.0, -1.0, 3.0, 0.0), (O.CHAIN, 8, 1, 3.0, -0.5, 1.0, 0.6, 0.3),
        (O.SQUARE, 4, 1, 3.0, 0.0, 1.0, 1.0, 0.0), (O.SQUARE, 8, 1, 5.0, -1.0, 1.0, 0.5, 0.0),
        (O.CHAIN, 2, 2, 4.0, 0.0, 0.0, 1.0, 0.0), (O.CHAIN, 8, 2, 4.0, -1.0, 1.0, 0.7, 0.0),
        (O.CHAIN, 8, 3, 3.0, 0.0, 1.0, 0.9, 0.0), (O.SQUARE, 4, 2, 2.0, -1.0, 0.5, 0.8, 0.0)]
LAUNCHES = [dict(), dict(group=4), dict(group=64, compress=1), dict(group=16, blocks=64)]


@pytest.mark.gpu
@pytest.mark.parametrize("kind,L,two_s,beta,jxy,jz,hx,h", TRAJ)
def test_transverse_bit_exact_with_oracle(loopgpu, kind, L, two_s, beta, jxy, jz, hx, h):
    c = O.PhiloxChain(kind, L, beta, jxy, jz, seed=3, chain_id=7, field=h, two_s=two_s, hx=hx)
    rc = c.sweep(40)
    cs = c.state()
    for lc in LAUNCHES:
        print("transverse", kind, L, two_s, beta, jxy, jz, hx, h, lc, flush=True)
        g = loopgpu.Chain(loopgpu.params(kind, L, beta, jxy, jz, seed=3, chain_id=7, h=h, two_s=two_s, hx=hx))
        if lc:
            g.set_launch(**lc)
        assert g.launch()["link"] == 0
        _cmp(g.sweep(40), rc, "%s %s" % ((kind, L, two_s, beta, jxy, jz, hx, h), lc))
        for a, b in zip(g.get_config(), cs):
            np.testing.assert_array_equal(a, b)


@pytest.mark.gpu
def test_transverse_chains_batched_with_plain_ones(loopgpu):
    """a batch of transverse-field and plain chains: each chain's records
    those of the oracle's philox mode for that chain alone"""
    ps = [dict(kind=O.CHAIN, L=16, beta=4.0, jxy=0.0, jz=1.0, hx=0.9, two_s=1),
          dict(kind=O.CHAIN, L=16, beta=4.0, jxy=1.0, jz=1.0, hx=0.0, two_s=1),
          dict(kind=O.SQUARE, L=4, beta=2.0, jxy=-1.0, jz=1.0, hx=0.7, two_s=1),
          dict(kind=O.CHAIN, L=8, beta=3.0, jxy=1.0, jz=1.0, hx=0.0, two_s=2),
          dict(kind=O.CHAIN, L=8, beta=3.0, jxy=0.0, jz=-1.0, hx=1.1, two_s=2)]
    plist = [loopgpu.params(q["kind"], q["L"], q["beta"], q["jxy"], q["jz"], seed=9, chain_id=i, two_s=q["two_s"],
                            hx=q["hx"]) for i, q in enumerate(ps)]
    b = loopgpu.Batch(plist)
    print("batched transverse", flush=True)
    rec = b.run(30)
    for i, q in enumerate(ps):
        c = O.PhiloxChain(q["kind"], q["L"], q["beta"], q["jxy"], q["jz"], seed=9, chain_id=i, two_s=q["two_s"],
                          hx=q["hx"])
        _cmp(rec[:, i, :], c.sweep(30), "chain %d" % i)
    with pytest.raises(loopgpu.LoopGPUError):
        b.set_launch(link=1)


@pytest.mark.gpu
def test_worker_gamma_matches_reference_site_run(loopgpu):
    """the worker's Gamma on two uncoupled S=1 sites against loop.op's
    diagonalization and path-integral blocks of one S=1 site in Gamma=1"""
    ed_ref, qmc = _fx("site_spin1_gamma1", "ed"), _fx("site_spin1_gamma1", "qmc")
    w = loopgpu.Worker('LATTICE = "chain lattice"\nL = 2\nJ = 0\nlocal_S = 1\nGamma = 1\nT = 0.25\n'
                       'SWEEPS = 65536\nSEED = 23125')
    print("worker Gamma", flush=True)
    w.run()
    for name, scale in (("Energy Density", 1.0), ("Susceptibility", 1.0), ("Specific Heat", 1.0),
                        ("Magnetization^2", 0.5), ("Staggered Magnetization^2", 0.5)):
        m, e, _ = w.result(name)
        m, e = m * scale, e * scale
        assert stats.sigmas(m, e, ed_ref[name][0]) < 4.0, (name, m, e, ed_ref[name][0])
        if name in qmc:
            assert stats.sigmas(m, e, qmc[name][0], qmc[name][1]) < 4.0, (name, m, e, qmc[name])


@pytest.mark.gpu
def test_worker_transverse_ising_matches_ed(loopgpu):
    ref = ed.diagonalize_transverse("chain", 8, 0.5, 1, 0.0, 1.0, 1.0)
    w = loopgpu.Worker('LATTICE = "chain lattice"\nL = 8\nJxy = 0\nJz = 1\nGamma = 1\nT = 0.5\n'
                       'SWEEPS = 65536\nSEED = 77')
    print("worker TFIM", flush=True)
    w.run()
    for name in ("Energy", "Specific Heat", "Magnetization^2", "Susceptibility", "Staggered Magnetization^2",
                 "Staggered Susceptibility"):
        m, e, _ = w.result(name)
        assert stats.sigmas(m, e, ref[name]) < 4.0, (name, m, e, ref[name])


@pytest.mark.gpu
def test_transverse_checkpoint_locations_and_resume(loopgpu):
    """a chain with site operators and in-site bonds saved mid-run: site
    operators sit at their site in the blob (looper's location, bit 0
    clear), bonds at bond << 1 | 1, and the loaded chain continues bit for
    bit"""
    p = loopgpu.params(O.CHAIN, 8, 3.0, -1.0, 1.0, seed=12, chain_id=3, two_s=2, hx=0.6, d=0.4,
                       force_scatter=0.1)
    a = loopgpu.Chain(p)
    a.sweep(20)
    spins, bo, times, types = a.get_config()
    n_sites, n_bonds = loopgpu.lattice_size(O.CHAIN, 8, 2, 0.6, 0.4)
    nb_real = n_bonds - n_sites
    n_site_ops = int(bo[-1] - bo[nb_real])
    assert n_site_ops > 0 and bo[nb_real] > 0
    blob = a.save()
    print("transverse checkpoint", flush=True)
    off = 8 + 8 + 8 + 4 * n_sites
    nops = int(np.frombuffer(blob[off:off + 8], np.uint64)[0])
    rec = np.frombuffer(blob[off + 8:off + 8 + 16 * nops], dtype=[("type", "<i4"), ("loc", "<i4"), ("time", "<f8")])
    assert nops == int(bo[-1])
    site = (rec["loc"] & 1) == 0
    assert int(site.sum()) == n_site_ops
    assert (rec["loc"][site] >> 1).max() < n_sites and (rec["loc"][~site] >> 1).max() < nb_real
    assert (np.diff(rec["time"]) >= 0).all()
    ra = a.sweep(15)
    b = loopgpu.Chain(p)
    b.load(blob)
    _cmp(b.sweep(15), ra, "resumed")
