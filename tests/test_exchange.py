"""Replica exchange (include/loopgpu.h lg_exchange_*): the device's exchange
decisions and every replica's trajectory against the oracle's philox mode
with the exchange step restated here, and per-temperature statistics against
exact diagonalisation."""

import numpy as np
import pytest

import oracle_lib as O
import stats

R_NOP, R_ENERGY, R_UMAG, R_BETA = 0, 2, 9, 18  # lg_sweep_record fields
EXACT = [0, 1]
FLOAT = list(range(2, 18)) + [20, 21, 22]


def test_exchange_argument_errors():
    import loopgpu as lg
    p = lg.params(lg.CHAIN, 8, 1.0)
    with pytest.raises(lg.LoopGPUError) as e:
        lg.Exchange(p, [1.0])                       # fewer than two temperatures
    assert e.value.code == -1
    with pytest.raises(lg.LoopGPUError) as e:
        lg.Exchange(p, [1.0, -2.0])                 # negative temperature
    assert e.value.code == -1
    if lg.device_count() == 0:
        with pytest.raises(lg.LoopGPUError) as e:
            lg.Exchange(p, [1.0, 2.0])
        assert e.value.code == -2


def oracle_exchange(kind, L, jxy, jz, betas, seed, chain0, xseed, n):
    """The exchange worker on the oracle: philox-mode replicas, then after
    each sweep the neighbour swaps t, t+1 (t even on even steps) accepted
    when log u < (n_a - n_b) log(beta_t+1 / beta_t) (W = beta^n,
    path_integral.C:99-109).  Returns per-sweep records of every chain, the
    temperature of every chain in every sweep, the final chain per
    temperature and the accepted swaps per pair."""
    R = len(betas)
    chains = [O.PhiloxChain(kind, L, b, jxy, jz, seed=seed, chain_id=chain0 + t) for t, b in enumerate(betas)]
    lnr = O.det_log(np.array([betas[t + 1] / betas[t] for t in range(R - 1)]))
    slot = list(range(R))
    beta_of = list(betas)
    acc = [0] * (R - 1)
    recs = np.zeros((n, R, O.R))
    temps = np.zeros((n, R))
    for step in range(n):
        for c in range(R):
            recs[step, c] = chains[c].sweep(1)[0]
            temps[step, c] = beta_of[c]
        for t in range(step & 1, R - 1, 2):
            a, b = slot[t], slot[t + 1]
            x = float(recs[step, a, R_NOP] - recs[step, b, R_NOP]) * lnr[t]
            accept = x >= 0.0
            if not accept:
                r = O.philox((step, 4, t, 0), (xseed, 0xFFFFFFFF))
                u53 = (int(r[0]) << 21) | (int(r[1]) >> 11)
                accept = O.det_log(np.array([(u53 + 1) * 2.0 ** -53]))[0] < x
            if accept:
                slot[t], slot[t + 1] = b, a
                beta_of[a], beta_of[b] = betas[t + 1], betas[t]
                chains[a].set_beta(betas[t + 1])
                chains[b].set_beta(betas[t])
                acc[t] += 1
    return recs, temps, slot, acc


CASES = [
    # kind, L, jxy, jz, betas, sweeps
    (O.CHAIN, 8, 1.0, 1.0, [0.5, 1.0, 2.0, 4.0], 60),
    (O.SQUARE, 4, 1.0, 0.8, [0.3, 0.6, 1.0, 1.6, 2.5], 40),
]


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=lambda c: "k%d_L%d_R%d" % (c[0], c[1], len(c[4])))
def test_exchange_matches_oracle_bit_exact(loopgpu, case):
    kind, L, jxy, jz, betas, n = case
    print("exchange parity", case, flush=True)
    seed, chain0, xseed = 77, 10, 4321
    g = loopgpu.Exchange(loopgpu.params(kind, L, betas[0], jxy, jz, seed=seed, chain_id=chain0), betas, seed=xseed)
    rg = g.run(n)
    at, acc, steps = g.state()
    recs, temps, slot, oacc = oracle_exchange(kind, L, jxy, jz, betas, seed, chain0, xseed, n)
    for c in range(len(betas)):
        np.testing.assert_array_equal(rg[:, c, EXACT], recs[:, c, EXACT], err_msg="chain %d" % c)
        np.testing.assert_allclose(rg[:, c, FLOAT], recs[:, c, FLOAT], rtol=1e-9, atol=1e-9)
        np.testing.assert_array_equal(rg[:, c, R_BETA], temps[:, c])
    assert list(at) == slot
    assert [int(a) for a in acc] == oacc
    assert steps == n
    assert sum(oacc) > 0  # the case exercises accepted swaps


@pytest.mark.gpu
def test_exchange_temperatures_match_exact_diagonalisation(loopgpu):
    import ed
    L = 8
    Ts = [0.2, 0.3, 0.45, 0.7, 1.0]
    betas = [1.0 / T for T in reversed(Ts)]  # ascending beta
    g = loopgpu.Exchange(loopgpu.params(loopgpu.CHAIN, L, betas[0], 1.0, 1.0, seed=5, chain_id=0), betas, seed=99)
    g.run(2000, records=False)
    n = 40000
    rec = g.run(n)
    at, acc, steps = g.state()
    assert min(acc) > 0
    for t, beta in enumerate(betas):
        sel = rec[:, :, R_BETA] == beta          # the chain at this temperature in each sweep
        assert np.all(sel.sum(axis=1) == 1)
        r = rec[sel]                             # (n, fields), sweep order
        ref = ed.diagonalize("chain", L, 1.0 / beta, 1.0, 1.0)
        e = r[:, R_ENERGY] / L
        chi = beta * r[:, R_UMAG] / L
        for name, x, rk in [("Energy Density", e, ref["Energy"] / L), ("Susceptibility", chi, ref["Susceptibility"])]:
            m, err = x.mean(), stats.binning_error(x)
            assert stats.sigmas(m, err, rk) < 4.0, (beta, name, m, err, rk)
