"""Square-lattice physics at 4x4 against exact diagonalisation
(tests/golden/ed_square4.json, written by tests/golden/make_ed_square4.py from
oracle/ed.py).  On 2x2 every neighbour pair is joined by two bonds, so a
mistake in the order or direction of a site's four distinct bonds could hide
there; 4x4 has them all distinct.  CPU: both oracle modes (looper mode, the
restatement of path_integral.C; philox mode, the device algorithm).  GPU: the
worker's observables."""

import json
import os

import numpy as np
import pytest

import oracle_lib as O
import stats

HERE = os.path.dirname(os.path.abspath(__file__))
ED = json.load(open(os.path.join(HERE, "golden", "ed_square4.json")))
NAMES = ["Energy", "Susceptibility", "Staggered Magnetization^2", "Staggered Susceptibility", "Magnetization^2"]


def _check(rec, case, where, tol=4.0):
    beta = 1.0 / case["T"]
    obs = stats.observables(rec, beta, 16, 2)
    for name in NAMES:
        m, e = stats.mean_err(obs[name])
        ref = case["results"][name]
        assert stats.sigmas(m, e, ref) < tol, (where, name, m, e, ref)


@pytest.mark.parametrize("case", ED, ids=lambda c: "T%g_jz%g" % (c["T"], c["Jz"]))
def test_oracle_modes_match_ed_square4(case):
    beta = 1.0 / case["T"]
    lc = O.LooperChain(O.SQUARE, 4, beta, case["Jxy"], case["Jz"], seed=4711)
    lc.sweep(2000)
    _check(lc.sweep(150000), case, "looper mode")
    pc = O.PhiloxChain(O.SQUARE, 4, beta, case["Jxy"], case["Jz"], seed=4711, chain_id=3)
    pc.sweep(2000)
    _check(pc.sweep(150000), case, "philox mode")


def test_ed_square4_sum_rules():
    """the fixture itself: Susceptibility = beta <M^2> / V for the conserved
    magnetization, structure factor = <M_s^2> / V"""
    for case in ED:
        r = case["results"]
        assert r["Susceptibility"] == pytest.approx(r["Magnetization^2"] / case["T"] / 16, rel=1e-9)
        assert r["Staggered Structure Factor"] == pytest.approx(r["Staggered Magnetization^2"] / 16, rel=1e-12)
        assert -0.7018 * 16 < r["Energy"] < 0   # above the ground-state energy of the 4x4 antiferromagnet


@pytest.mark.gpu
@pytest.mark.parametrize("case", ED, ids=lambda c: "T%g_jz%g" % (c["T"], c["Jz"]))
def test_worker_matches_ed_square4(loopgpu, case):
    w = loopgpu.Worker('LATTICE = "square lattice"; L = 4; T = %r; Jxy = %r; Jz = %r; SWEEPS = 200000; '
                       'THERMALIZATION = 2000; SEED = 99' % (case["T"], case["Jxy"], case["Jz"]))
    print("worker square 4x4", case["T"], case["Jz"], flush=True)
    w.run()
    for name in NAMES + ["Specific Heat"]:
        if name not in case["results"]:
            continue
        m, e, _ = w.result(name)
        assert stats.sigmas(m, e, case["results"][name]) < 4.0, (name, m, e, case["results"][name])
